/*
 * apocharmm_amd -- the reference's existing C ABI (capi/include/apocharmm_c/)
 * over the MI355X force path and integrator of apocharmm_amd.h.
 *
 * Entry names, argument order and meaning follow the reference's headers so
 * that its ctypes binding (python/apocharmm/) and C callers bind them
 * unchanged; each declaration cites the header line it replaces.  Status
 * codes and apo_last_error() are those of apocharmm_amd.h (same values as
 * capi/include/apocharmm_c/Status.h).  Handles own shared state: destroying
 * a handle that another object still uses (a PSF held by a force manager, a
 * force manager held by a context) keeps the object alive for its users, as
 * the reference's shared_ptr handles do.
 *
 * Not exported: the Nose-Hoover integrator entries
 * (capi/include/apocharmm_c/CudaNoseHooverIntegrator.h); the Langevin
 * thermostat integrator, which the reference's C ABI does not expose, is
 * reachable through apo_cuda_langevin_thermostat_integrator_* below.
 */
#ifndef APOCHARMM_AMD_CAPI_H
#define APOCHARMM_AMD_CAPI_H

#include "apocharmm_amd.h"

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- CharmmPsf (capi/include/apocharmm_c/CharmmPsf.h:22-77) ---------------
 * String getters write 8 characters per atom, space padded, into a buffer of
 * at least 8 * natom bytes. */
typedef struct apo_charmm_psf apo_charmm_psf;
APO_API apo_status apo_charmm_psf_create(apo_charmm_psf **out, const char *path);
APO_API void apo_charmm_psf_destroy(apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_num_atoms(size_t *num_atoms, const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_num_bonds(size_t *num_bonds, const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_num_angles(size_t *num_angles, const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_num_dihedrals(size_t *num_dihedrals,
                                                    const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_num_impropers(size_t *num_impropers,
                                                    const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_num_cross_terms(size_t *num_cross_terms,
                                                      const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_segment_identifiers(char *segis, const size_t len,
                                                          const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_residue_identifiers(int *resis, const size_t len,
                                                          const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_residue_names(char *resns, const size_t len,
                                                    const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_atom_names(char *anams, const size_t len,
                                                 const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_atom_types(char *atyps, const size_t len,
                                                 const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_charges(double *charges, const size_t len,
                                              const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_masses(double *masses, const size_t len,
                                             const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_net_charge(double *net_charge, const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_total_mass(double *total_mass, const apo_charmm_psf *psf);
APO_API apo_status apo_charmm_psf_get_file_name(char *file_name, const size_t len,
                                                const apo_charmm_psf *psf);

/* ---- CharmmCrd (capi/include/apocharmm_c/CharmmCrd.h:22-34) ---------------- */
typedef struct apo_charmm_crd apo_charmm_crd;
APO_API apo_status apo_charmm_crd_create(apo_charmm_crd **out, const char *path);
APO_API void apo_charmm_crd_destroy(apo_charmm_crd *crd);
APO_API apo_status apo_charmm_crd_get_num_atoms(size_t *num_atoms, const apo_charmm_crd *crd);
/* xyz: 3 * natom doubles, atom-major (x0 y0 z0 x1 ...) */
APO_API apo_status apo_charmm_crd_get_coordinates(double *xyz, const size_t xyz_len,
                                                  const apo_charmm_crd *crd);

/* ---- CharmmParameters (capi/include/apocharmm_c/CharmmParameters.h:22-31) - */
typedef struct apo_charmm_parameters apo_charmm_parameters;
APO_API apo_status apo_charmm_parameters_create(apo_charmm_parameters **out, const char *path);
APO_API apo_status apo_charmm_parameters_create_from_files(apo_charmm_parameters **out,
                                                           const char *const *paths,
                                                           const size_t num_paths);
APO_API void apo_charmm_parameters_destroy(apo_charmm_parameters *parameters);

/* ---- ForceManager (capi/include/apocharmm_c/ForceManager.h:22-35) ----------
 * The reference's defaults (src/ForceManager.cu:35-49): kappa 0.34, cutoff
 * 14, ctonnb 12, ctofnb 10, PME order 4, FFT grid 2*floor(L/2) per axis
 * (computeFFTGridSize, :923-943).  The GPU side is created when a context is
 * created on the manager (CharmmContext's constructor initialises it). */
typedef struct apo_force_manager apo_force_manager;
APO_API apo_status apo_force_manager_create(apo_force_manager **out, const apo_charmm_psf *psf,
                                            const apo_charmm_parameters *parameters);
APO_API void apo_force_manager_destroy(apo_force_manager *force_manager);
APO_API apo_status apo_force_manager_set_box_dimensions(apo_force_manager *force_manager,
                                                        const double x, const double y,
                                                        const double z);

/* The remaining ForceManager setters (include/ForceManager.h:200-231), which
 * the reference's C ABI does not expose; they take effect when the manager is
 * initialised (a context is created on it).  fft grid <= 0 restores the
 * 2*floor(L/2) default. */
APO_API apo_status apo_force_manager_set_fft_grid(apo_force_manager *force_manager, const int nx,
                                                  const int ny, const int nz);
APO_API apo_status apo_force_manager_set_kappa(apo_force_manager *force_manager,
                                               const double kappa);
APO_API apo_status apo_force_manager_set_cutoff(apo_force_manager *force_manager,
                                                const double cutoff);
APO_API apo_status apo_force_manager_set_ctonnb(apo_force_manager *force_manager,
                                                const double ctonnb);
APO_API apo_status apo_force_manager_set_ctofnb(apo_force_manager *force_manager,
                                                const double ctofnb);
/* The manager's force path as an apo_nb_system handle (borrowed: valid while
 * the manager lives; NOT_INITIALIZED before a context was created on it), so
 * apo_nb_calc_force evaluates the system read from the files. */
APO_API apo_status apo_force_manager_get_nb_system(apo_force_manager *force_manager,
                                                   apo_nb_system **out);

/* ---- CharmmContext (capi/include/apocharmm_c/CharmmContext.h:28-48) --------
 * Holonomic constraints are on by default (CharmmContext's constructor):
 * SETTLE for (OT, HT, HT) waters and SHAKE for the other bonds to hydrogen;
 * the degrees of freedom are 3N - 3 minus one per constraint. */
typedef struct apo_charmm_context apo_charmm_context;
APO_API apo_status apo_charmm_context_create(apo_charmm_context **out,
                                             const apo_force_manager *force_manager);
APO_API void apo_charmm_context_destroy(apo_charmm_context *context);
APO_API apo_status apo_charmm_context_set_coordinates(apo_charmm_context *context,
                                                      const apo_charmm_crd *crd);
APO_API apo_status apo_charmm_context_set_random_seed_for_velocities(apo_charmm_context *context,
                                                                     const uint64_t seed);
APO_API apo_status apo_charmm_context_use_holonomic_constraints(apo_charmm_context *context,
                                                                const bool useHolonomicConstraints);
APO_API apo_status apo_charmm_context_assign_velocities_at_temperature(
    apo_charmm_context *context, const double temperature);
APO_API apo_status apo_charmm_context_compute_temperature(double *temperature,
                                                          apo_charmm_context *context);

/* ---- Subscribers (capi/include/apocharmm_c/Subscriber.h:22-28,
 * RestartSubscriber.h:22-34, DcdSubscriber.h:22-34) ----------------------------
 * A subscriber's update runs after every report_frequency-th step of a
 * propagate call (CudaIntegrator::reportIfNeeded, src/CudaIntegrator.cu:347+). */
typedef struct apo_subscriber apo_subscriber;
APO_API apo_status apo_subscriber_set_report_frequency(apo_subscriber *subscriber,
                                                       const int report_frequency);
APO_API apo_status apo_subscriber_get_report_frequency(int *report_frequency,
                                                       const apo_subscriber *subscriber);
typedef struct apo_restart_subscriber apo_restart_subscriber;
APO_API apo_status apo_restart_subscriber_create(apo_restart_subscriber **out, const char *path);
APO_API apo_status apo_restart_subscriber_create_with_report_frequency(
    apo_restart_subscriber **out, const char *path, const int report_frequency);
APO_API void apo_restart_subscriber_destroy(apo_restart_subscriber *subscriber);
APO_API apo_status apo_restart_subscriber_as_subscriber(apo_subscriber **out,
                                                        apo_restart_subscriber *subscriber);
typedef struct apo_dcd_subscriber apo_dcd_subscriber;
APO_API apo_status apo_dcd_subscriber_create(apo_dcd_subscriber **out, const char *path);
APO_API apo_status apo_dcd_subscriber_create_with_report_frequency(apo_dcd_subscriber **out,
                                                                   const char *path,
                                                                   const int report_frequency);
APO_API void apo_dcd_subscriber_destroy(apo_dcd_subscriber *subscriber);
APO_API apo_status apo_dcd_subscriber_as_subscriber(apo_subscriber **out,
                                                    apo_dcd_subscriber *subscriber);

/* ---- CudaIntegrator (capi/include/apocharmm_c/CudaIntegrator.h:22-38) ------ */
typedef struct apo_cuda_integrator apo_cuda_integrator;
APO_API apo_status apo_cuda_integrator_set_time_step(apo_cuda_integrator *integrator,
                                                     const double time_step);
APO_API apo_status apo_cuda_integrator_set_charmm_context(apo_cuda_integrator *integrator,
                                                          apo_charmm_context *context);
APO_API apo_status apo_cuda_integrator_subscribe(apo_cuda_integrator *integrator,
                                                 apo_subscriber *subscriber);
APO_API apo_status apo_cuda_integrator_propagate(apo_cuda_integrator *integrator,
                                                 const int num_steps);
/* CudaLangevinThermostatIntegrator::initializeFromRestartFile
 * (src/CudaLangevinThermostatIntegrator.cu:264-520): box, step counters,
 * seed and noise position, coordinates, velocities and the previous
 * half-step displacements. */
APO_API apo_status apo_cuda_integrator_initialize_from_restart_file(
    apo_cuda_integrator *integrator, const char *path);

/* ---- CudaLangevinThermostatIntegrator (include/CudaLangevinThermostatIntegrator.h):
 * no C entry in the reference; named after the class as the Nose-Hoover
 * entries are (capi/include/apocharmm_c/CudaNoseHooverIntegrator.h:28-66).
 * time_step in ps; friction 1/ps; the fused MD step of apo_lt_* runs it. */
typedef struct apo_cuda_langevin_thermostat_integrator apo_cuda_langevin_thermostat_integrator;
APO_API apo_status apo_cuda_langevin_thermostat_integrator_create(
    apo_cuda_langevin_thermostat_integrator **out, const double time_step);
APO_API void apo_cuda_langevin_thermostat_integrator_destroy(
    apo_cuda_langevin_thermostat_integrator *integrator);
APO_API apo_status apo_cuda_langevin_thermostat_integrator_set_friction(
    apo_cuda_langevin_thermostat_integrator *integrator, const double friction);
APO_API apo_status apo_cuda_langevin_thermostat_integrator_set_reference_temperature(
    apo_cuda_langevin_thermostat_integrator *integrator, const double temperature);
APO_API apo_status apo_cuda_langevin_thermostat_integrator_set_rng_seed(
    apo_cuda_langevin_thermostat_integrator *integrator, const uint64_t seed);
APO_API apo_status apo_cuda_langevin_thermostat_integrator_as_cuda_integrator(
    apo_cuda_integrator **out, apo_cuda_langevin_thermostat_integrator *integrator);
/* Copies the state out: xyz, vel, dprev as 3 * natom doubles (atom-major,
 * NULL skips), the total step count (NPRIV). */
APO_API apo_status apo_cuda_integrator_get_state(apo_cuda_integrator *integrator, double *xyz,
                                                 double *vel, double *dprev,
                                                 unsigned long long *total_steps);

#ifdef __cplusplus
}
#endif

#endif
