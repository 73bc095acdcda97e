/*
 * apocharmm_amd -- C ABI of the MI355X-native CHARMM force path.
 *
 * One handle (apo_nb_system) owns everything the reference's ForceManager
 * owns for one system on the hot path: CudaPMEDirectForce (neighbor list +
 * pair kernel + 1-4/exclusion kernel), CudaPMEReciprocalForce, CudaBondedForce
 * and the energy/virial reduction.  Coordinates come in as the reference's
 * xyzq float4 array; the result is the fixed-point (2^-40) gradient in the
 * reference's strided SoA layout, per-term energies and the 3x3 virial.
 *
 * Conventions follow the reference's C ABI (capi/include/apocharmm_c/):
 *  - every entry returns apo_status (Status.h); on failure the message is
 *    kept in thread-local storage and read with apo_last_error() (Error.h,
 *    capi/src/apocharmm_c/Error.cpp:18-22);
 *  - handles are opaque and released with *_destroy.
 * All device pointers are HBM addresses on the current HIP device; `stream`
 * is a hipStream_t passed as void* (NULL = default stream).  Entry points
 * that only enqueue work do not synchronise the host.
 */
#ifndef APOCHARMM_AMD_H
#define APOCHARMM_AMD_H

#include <stddef.h>
#include <stdint.h>

#if defined(_WIN32)
#define APO_API __declspec(dllexport)
#else
#define APO_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: capi/include/apocharmm_c/Status.h (CUDA_ERROR -> GPU_ERROR). */
typedef enum apo_status {
  APO_OK = 0,
  APO_INVALID_ARGUMENT = 1,
  APO_RUNTIME_ERROR = 2,
  APO_GPU_ERROR = 3,
  APO_NOT_INITIALIZED = 4,
  APO_NOT_IMPLEMENTED = 5
} apo_status;

/* Energy terms, in the order of the keys of ForceManager::getEnergyComponents
 * (src/ForceManager.cu:278-295); index APO_E_EPOT holds their sum
 * (UpdatePotentialEnergyKernel, src/ForceManager.cu:641-653). */
enum {
  APO_E_BOND = 0,
  APO_E_ANGLE = 1,
  APO_E_UREYB = 2,
  APO_E_DIHE = 3,
  APO_E_IMDIHE = 4,
  APO_E_EWKS = 5,
  APO_E_EWSE = 6,
  APO_E_EWEX = 7,
  APO_E_ELEC = 8,
  APO_E_VDW = 9,
  APO_E_EPOT = 10,
  APO_NUM_ENERGY = 11
};

/* calc_force flags.  APO_CALC_ENERGY / APO_CALC_VIRIAL are the calcEnergy /
 * calcVirial arguments of ForceManager::calcForce (include/ForceManager.h:424);
 * APO_RESET_LIST is its `reset` argument (rebuild the neighbor list first).
 * The APO_TERM_* bits select force groups (0 = all groups), which the
 * reference does through its three force objects. */
enum {
  APO_CALC_ENERGY = 1 << 0,
  APO_CALC_VIRIAL = 1 << 1,
  APO_RESET_LIST = 1 << 2,
  /* add into d_force, d_energy and d_virial instead of overwriting them, so
   * a caller's own forces (plugins, restraints) computed into the same
   * buffers beforehand are kept (no counterpart flag in the reference, whose
   * manager sums plugin buffers afterwards: apo_nb_add_force) */
  APO_ACCUMULATE = 1 << 3,
  APO_TERM_BONDED = 1 << 8,
  APO_TERM_DIRECT = 1 << 9,
  APO_TERM_PAIRS14 = 1 << 10,
  APO_TERM_RECIP = 1 << 11,
  APO_TERM_SELF = 1 << 12,
  APO_TERM_ALL = 0x1F00
};

/* Arithmetic of the pair and reciprocal kernels: the mixed path (FP32 math,
 * int64 fixed-point accumulation) or the all-FP64 instantiation of the same
 * kernels. */
enum { APO_PRECISION_MIXED = 0, APO_PRECISION_FP64 = 1 };

/* Simulation parameters.  Field meaning follows the ForceManager setters
 * (include/ForceManager.h:200-231): cutnb = setCutoff (list cutoff),
 * ctofnb = setCtofnb (r_off), ctonnb = setCtonnb (r_on), kappa = setKappa,
 * fft_grid = setFFTGrid, pme_order = setPmeSplineOrder (4 only), box =
 * setBoxDimensions (orthorhombic).  e14fac scales 1-4 electrostatics
 * (CudaPMEDirectForce::setup, src/CudaPMEDirectForce.cu:275). */
typedef struct apo_nb_params {
  double box[3];
  double kappa;
  double cutnb;
  double ctofnb;
  double ctonnb;
  double e14fac;
  int32_t fft_grid[3];
  int32_t pme_order;
  int32_t precision;
  int32_t reserved;
} apo_nb_params;

/* Topology and resolved force-field parameters (what ForceManager::initialize
 * extracts from CharmmPSF + CharmmParameters, src/ForceManager.cu:463-527).
 *  vdwtype[natom], vdwtab[ntypes*(ntypes+1)]: (c6, c12) per lower-triangular
 *    type pair a*(a+1)/2+b (a >= b), U = c12/r^12 - c6/r^6 before switching
 *    (CharmmParameters::getVdwParamsAndTypes, src/CharmmParameters.cu:718-844);
 *    vdwtab14 the same for 1-4 pairs (may equal vdwtab);
 *  bonds[2*nbond] + bond_par[2*nbond] (b0, kb): harmonic bonds; the bond list
 *    is also the connectivity from which the 1-2/1-3/1-4 exclusion sets are
 *    derived (CharmmPSF::buildTopologicalExclusions, src/CharmmPSF.cu:331-385);
 *  ureyb[2*n] + ub_par[2*n] (s0, kub);  angles[3*n] + angle_par[2*n]
 *    (theta0 [rad], ktheta);  dihe[4*n] + dihe_par[3*n] (multiplicity, k,
 *    delta [rad]), one row per cosine term;  imdihe[4*n] + imdihe_par[3*n]
 *    (multiplicity 0 = harmonic k (phi - delta)^2).
 * All indices are 0-based.  Arrays are host memory, copied at creation. */
typedef struct apo_nb_topology {
  int32_t natom;
  int32_t ntypes;
  const int32_t *vdwtype;
  const float *vdwtab;
  const float *vdwtab14;
  int32_t nbond;
  const int32_t *bonds;
  const float *bond_par;
  int32_t nureyb;
  const int32_t *ureyb;
  const float *ub_par;
  int32_t nangle;
  const int32_t *angles;
  const float *angle_par;
  int32_t ndihe;
  const int32_t *dihe;
  const float *dihe_par;
  int32_t nimdihe;
  const int32_t *imdihe;
  const float *imdihe_par;
} apo_nb_topology;

typedef struct apo_nb_system apo_nb_system;

/* Thread-local message of the last failed call on this thread
 * (apo_last_error, capi/src/apocharmm_c/Error.cpp:18-22). */
APO_API const char *apo_last_error(void);

/* Library version string. */
APO_API const char *apo_version(void);

/* Creates the force path for one system on the current HIP device
 * (ForceManager(psf, prm) + ForceManager::initialize,
 * src/ForceManager.cu:436-550).  Validates cutoffs against the box as the
 * reference does (cutoff <= box/2, src/ForceManager.cu:443-447). */
APO_API apo_status apo_nb_create(const apo_nb_params *params,
                                 const apo_nb_topology *topology,
                                 apo_nb_system **out);

/* Releases the handle and its device memory (ForceManager::dealloc,
 * src/ForceManager.cu:957-975). */
APO_API apo_status apo_nb_destroy(apo_nb_system *sys);

/* ForceManager::setBoxDimensions (include/ForceManager.h:200).  Takes effect
 * on the next calc_force; the neighbor list is rebuilt on the next reset. */
APO_API apo_status apo_nb_set_box(apo_nb_system *sys, const double *box3);

/* ForceManager::resetNeighborList(xyzq) (src/ForceManager.cu,
 * CudaPMEDirectForce::resetNeighborList src/CudaPMEDirectForce.cu:1347-1366):
 * cell sort + tile list build from device xyzq[4*natom]. */
APO_API apo_status apo_nb_reset_list(apo_nb_system *sys, const float *d_xyzq,
                                     void *stream);

/* ForceManager::calcForce(xyzq, reset, calcEnergy, calcVirial)
 * (include/ForceManager.h:424, src/ForceManager.cu:821-827).
 *  d_xyzq    device float[4*natom] (x, y, z, q), whole molecules as produced
 *            by CharmmContext::imageCentering;
 *  d_force   device int64[3*stride]: OVERWRITTEN with the total gradient
 *            dU/dx in units of 2^-40 kcal/mol/A, x at [0,natom), y at
 *            [stride, ...), z at [2*stride, ...) (Force<long long>,
 *            include/Force.h:32-154);
 *  d_energy  device double[APO_NUM_ENERGY] or NULL; written when
 *            APO_CALC_ENERGY is set;
 *  d_virial  device double[9] or NULL; written when APO_CALC_VIRIAL is set
 *            (W = -sum x (x) g, row-major).
 * Work is enqueued on `stream`; the host is not synchronised. */
APO_API apo_status apo_nb_calc_force(apo_nb_system *sys, const float *d_xyzq,
                                     int32_t flags, int64_t *d_force,
                                     int32_t stride, double *d_energy,
                                     double *d_virial, void *stream);

/* apo_nb_calc_force for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops); no counterpart in the
 * reference.  Same arguments and the same results, bit for bit.  The
 * energy/virial partial sums go to a static device array of 1 MiB per GPU
 * (131,072 doubles, held until the process ends) instead of the system's own
 * buffer, and a list rebuild never grows the list buffers (a rebuild that
 * would need more room is reported at the next call, as by
 * apo_nb_calc_force).  APO_INVALID_ARGUMENT when the system's partials need
 * more than the array, when its list was never built (build it once with
 * apo_nb_reset_list or apo_nb_calc_force) or with kernel timers on.  Calls
 * that overlap on one GPU share the array and race. */
APO_API apo_status apo_nb_calc_force_noalloc(apo_nb_system *sys, const float *d_xyzq,
                                             int32_t flags, int64_t *d_force, int32_t stride,
                                             double *d_energy, double *d_virial, void *stream);

/* Plugin forces (ForceManager::subscribe / ForceView, include/ForceManager.h:
 * 37-105, 427-444): the manager adds each subscribed force's int64 gradient
 * into the total, d_dst[k] += d_src[k] for the three strided rows of natom
 * entries (Force<long long>::add / add_force, src/reduce.cu:177-189, as
 * calcForcePart3 does at src/ForceManager.cu:678-682).  Enqueued on stream. */
APO_API apo_status apo_nb_add_force(const int64_t *d_src, int32_t natom,
                                    int32_t stride, int64_t *d_dst,
                                    void *stream);

/* ... and a plugin's energy (its term 0) into the total potential:
 * d_energy[APO_E_EPOT] += d_plugin_energy[0] (UpdatePotentialEnergyKernel2,
 * src/ForceManager.cu:774-778).  Enqueued on stream. */
APO_API apo_status apo_nb_add_energy(const double *d_plugin_energy,
                                     double *d_energy, void *stream);

/* ForceManager::getForceStride: the smallest valid stride for this system,
 * ceil((natom+31)*8/256)*256/8 (src/Force.cu:344-352). */
APO_API apo_status apo_nb_force_stride(const apo_nb_system *sys,
                                       int32_t *stride);

APO_API apo_status apo_nb_num_atoms(const apo_nb_system *sys, int32_t *natom);

/* Force<long long>::convert / reduce_force<long long,double>
 * (src/reduce.cu:132-143): d_out[k] = (double)d_force[k] * 2^-40 for
 * k in the three strided rows of natom entries. */
APO_API apo_status apo_nb_force_to_double(const int64_t *d_force,
                                          int32_t natom, int32_t stride,
                                          double *d_out, void *stream);

/* Neighbor-list statistics of the last build (synchronises the stream):
 * out[0] cells, out[1] j entries of the pair list (8-slot i-groups times the
 * j slots with a live pair, padded to 64-lane steps), out[2] work items,
 * out[3] atom pairs inside the list cutoff covered by the list, out[4]
 * columns, out[5] largest column population.  The pair count
 * is the reference's debug cross-check quantity
 * (src/CudaNeighborListBuild.cu:2215-2264). */
APO_API apo_status apo_nb_list_stats(apo_nb_system *sys, int64_t *out6,
                                     void *stream);

/* Kernel timing with HIP events recorded on the calc_force stream (no
 * counterpart in the reference, which times with NVTX ranges,
 * src/ForceManager.cu:606-616).  enable != 0 starts recording and clears the
 * accumulators.  apo_nb_profile_read synchronises the recorded events and
 * returns, per slot, the summed milliseconds and the number of launches:
 * slot 0 pair kernel, 1 PME spread with the forward x FFT, 2 PME FFT +
 * convolution (the other 4 passes),
 * 3 PME gather, 4 neighbor-list rebuild, 5 whole calc_force, 6 bonded +
 * 1-4/exclusions, 7 sort/finalize of the direct-space coordinates/forces,
 * 8 integrator update (apo_lt_propagate: leapfrog + SETTLE + kinetic energy),
 * 9 image centering + centre-of-mass removal.  While recording, integrator
 * steps run eagerly (no graph replay). */
enum { APO_PROFILE_SLOTS = 10 };
APO_API apo_status apo_nb_profile(apo_nb_system *sys, int32_t enable);
APO_API apo_status apo_nb_profile_read(apo_nb_system *sys, double *ms,
                                       int64_t *count);

/* Host-side topology helper (no device work): 1-2/1-3/1-4 exclusion CSR and
 * the in14/ex14 pair lists derived from a bond list, exactly as
 * apo_nb_create derives them (CharmmPSF::buildTopologicalExclusions +
 * getInclusionExclusionLists, src/CharmmPSF.cu:235-385).  Pass NULL outputs
 * to query sizes: counts[0] = excl entries, counts[1] = in14 pairs,
 * counts[2] = ex14 pairs. */
APO_API apo_status apo_topology_exclusions(int32_t natom, int32_t nbond,
                                           const int32_t *bonds,
                                           int32_t *counts3,
                                           int32_t *excl_ptr,
                                           int32_t *excl_idx, int32_t *in14,
                                           int32_t *ex14);

/* ---- Langevin thermostat integrator with SETTLE --------------------------
 * CudaLangevinThermostatIntegrator (include/CudaLangevinThermostatIntegrator.h,
 * src/CudaLangevinThermostatIntegrator.cu) with CudaHolonomicConstraint's
 * SETTLE (src/CudaHolonomicConstraint.cu:164-303), driving an apo_nb_system
 * as CharmmContext drives ForceManager.  FP64 coordinates, velocities and
 * half-step displacements stay on the device; each step calls the force path
 * on the float xyzq it keeps. */
typedef struct apo_lt_integrator apo_lt_integrator;

typedef struct apo_lt_params {
  double timestep;    /* ps: CudaIntegrator(timeStep), src/CudaIntegrator.cu:38-40 */
  double temperature; /* K: setReferenceTemperature */
  double friction;    /* 1/ps: setThermostatFriction; 0 gives NVE leapfrog */
  uint64_t seed;      /* setThermostatRngSeed; the noise of (step, atom) is a
                         pure function of (seed, step, atom) */
  int32_t nbfreq;     /* setNonbondedListUpdateFrequency; <= 0 means 20 */
  int32_t comfreq;    /* setRemoveCenterOfMassFrequency; 0 means 1000, < 0 never */
  int32_t nwater;     /* rigid waters constrained by SETTLE */
  int32_t use_graph;  /* nonzero: replay steps without a list rebuild as a
                         hipGraph */
  const int32_t *waters; /* 3*nwater atom indices (O, H1, H2)
                            (CharmmPSF::getWaterMolecules) */
  double water_geom[4];  /* mO, mH, rOH, rHH of the rigid water; all zero
                            selects 15.9994, 1.008, 0.9572, 1.5139
                            (src/CudaHolonomicConstraint.cu:18-33) */
  const double *mass;    /* natom, amu */
  const float *charge;   /* natom, e; the q of the xyzq handed to calc_force */
  int32_t list_policy;   /* 0: rebuild every nbfreq steps (CudaIntegrator's
                            fixed cadence, src/CudaIntegrator.cu:29); 1: also
                            rebuild when an atom may have moved more than half
                            the list buffer (cutnb - ctofnb)/2 since the last
                            build: the step kernel measures every atom's
                            displacement from the build's coordinates, and
                            the host, which reads it at most 4 steps late,
                            adds the largest one-step displacement for each
                            step it has not seen.  Batches need 0. */
  int32_t prune_every;   /* no counterpart in the reference: with list_policy
                            0, every prune_every steps between builds the pair
                            kernel's list is pruned on the current coordinates
                            at r_off + (cutnb - r_off) * prune_every / nbfreq
                            (the build's own buffer per step, so a pair the
                            build's list would reach is kept); 0 = off.
                            Fused batch steps use the build's list. */
  int32_t launch_setting; /* no counterpart in the reference: launch order of
                            the force branches inside captured steps (it
                            changes no result bit).  0: the first list cycles
                            time the four settings (1 untimed + 3 timed cycles
                            each, waiting for the GPU on the host) and keep
                            the fastest; 1..4: use setting (launch_setting - 1)
                            without a trial, bit 0 = PME kernels at raised
                            wave issue priority, bit 1 = pair kernel captured
                            first. */
  int32_t nshake;        /* SHAKE groups of bonds to hydrogen (outside the
                            SETTLE waters), each a heavy atom and up to three
                            hydrogens bonded to it (ForceManager::
                            initializeHolonomicConstraintsVariables,
                            src/ForceManager.cu:844-918; solved as
                            Shake2/3/4Kernel, src/CudaHolonomicConstraint.cu:
                            324-520); 0 = none */
  const int32_t *shake_groups; /* 4*nshake: heavy, h1, h2 or -1, h3 or -1 */
  const double *shake_r2;      /* nshake: squared bond length of the group */
} apo_lt_params;

/* Binds an integrator to `sys` (setCharmmContext) with host coordinates
 * xyz[3*natom] (x0 y0 z0 x1 ...; A) and velocities vel[3*natom] (A per AKMA
 * time unit, CharmmContext::assignVelocitiesAtTemperature; NULL = zero).  The
 * force system must outlive the integrator. */
APO_API apo_status apo_lt_create(apo_nb_system *sys, const apo_lt_params *params,
                                 const double *xyz, const double *vel,
                                 apo_lt_integrator **out);
APO_API apo_status apo_lt_destroy(apo_lt_integrator *lt);

/* CudaLangevinThermostatIntegrator::initialize
 * (src/CudaLangevinThermostatIntegrator.cu:206-258): constrain, force,
 * leapfrog back-step.  apo_lt_propagate calls it when it has not run. */
APO_API apo_status apo_lt_initialize(apo_lt_integrator *lt, void *stream);

/* CudaIntegrator::propagate(nsteps) (src/CudaIntegrator.cu:140-230): image
 * centering and list rebuild every nbfreq steps, centre-of-mass removal every
 * comfreq steps, then per step force, Langevin leapfrog update, SETTLE,
 * velocities, kinetic energy and the float xyzq.  Enqueued after the work
 * already on `stream`.  The host does not wait: list capacities come from the
 * first build, and a later build's counts are checked before the next rebuild
 * (growing the buffers after draining the device, or failing with the
 * build's error). */
APO_API apo_status apo_lt_propagate(apo_lt_integrator *lt, int32_t nsteps,
                                    void *stream);

/* Advances n independent integrators (each bound to its own force system) by
 * nsteps together through a one-off apo_lt_batch (below): one graph launch per
 * batch step runs every system's step in fused launches, with no host
 * synchronisation.  Each system's trajectory is bitwise the one
 * apo_lt_propagate gives it alone.  No
 * counterpart in the reference, which runs one CharmmContext per process; it
 * serves many-system workloads (replica exchange, FEP windows, separate
 * jobs). */
APO_API apo_status apo_lt_propagate_batch(apo_lt_integrator *const *lts, int32_t n,
                                          int32_t nsteps, void *stream);

/* A persistent batch over n integrators (no counterpart in the reference):
 * apo_lt_batch_propagate advances all of them by nsteps with one graph launch
 * per batch step, the systems' steps chained into `lanes` parallel branches
 * (0 selects 4, the default number of hardware queues).  Each trajectory is
 * bitwise the one apo_lt_propagate gives the system alone.  While the batch
 * exists its systems run their force terms on one stream each; destroy the
 * batch before the integrators.  apo_lt_propagate_batch is a one-off batch. */
typedef struct apo_lt_batch apo_lt_batch;
APO_API apo_status apo_lt_batch_create(apo_lt_integrator *const *lts, int32_t n, int32_t lanes,
                                       apo_lt_batch **out);
APO_API apo_status apo_lt_batch_propagate(apo_lt_batch *b, int32_t nsteps, void *stream);
APO_API apo_status apo_lt_batch_destroy(apo_lt_batch *b);

/* Autotuner of the batch's launch shapes (no counterpart in the reference):
 * times the fused spread + FFT launches over seven FFT shapes (threads, row
 * pairs, columns per block), then the whole force step (PME, bonded and pair
 * branches running together) over eight pair-kernel block budgets, on the
 * systems' current coordinates and lists, keeps the fastest, and reports
 * report9 = blocks per CU, FFT threads, FFT rows, FFT columns, the chosen
 * force-step and PME-chain times (ms), the default shapes' force-step and
 * PME-chain times (ms), launches timed.  Trajectories are unchanged: the timed
 * launches rewrite per-step scratch and the gradient buffers, which every step
 * recomputes before it uses them (until the next step, apo_lt_get_state's
 * grad holds those partial sums instead of the last step's gradient).
 * Synchronises. */
APO_API apo_status apo_lt_batch_autotune(apo_lt_batch *b, void *stream, double *report9);

/* Copies the state to the host (synchronises): any output may be NULL.
 *  xyz, vel, dprev, grad: double[3*natom] AoS; vel is the on-step velocity of
 *    the last step, dprev the last half-step displacement (getCoordsDelta),
 *    grad the gradient the last step used;
 *  thermo: double[6] = kinetic energy (mean of the old half-step, on-step and
 *    new half-step estimates), on-step kinetic energy, the two temperatures,
 *    and their running means (getKineticEnergy, getAverageTemperature);
 *  step: steps propagated since creation. */
APO_API apo_status apo_lt_get_state(apo_lt_integrator *lt, double *xyz, double *vel,
                                    double *dprev, double *grad, double *thermo6,
                                    int64_t *step);

/* Neighbor-list health (synchronises): out4 = largest displacement of any
 * atom since the last list build (A), largest one-step displacement since
 * that build (A), rebuilds the displacement check added (list_policy 1), and
 * half the list buffer (cutnb - ctofnb)/2 (A).  No counterpart in the
 * reference, which rebuilds at a fixed cadence without a check. */
APO_API apo_status apo_lt_list_info(apo_lt_integrator *lt, double *out4);

/* Launch settings of the integrator's captured steps (no counterpart in the
 * reference): out6 = chosen setting (0..3, bit 0 PME priority, bit 1 pair
 * kernel first; -1 while the timing trial is still running), trial cycles
 * run so far, and the fastest timed cycle of each of the four settings (ms,
 * 0 if not timed). */
APO_API apo_status apo_lt_launch_info(apo_lt_integrator *lt, double *out6);

/* Device float xyzq[4*natom] the integrator keeps for the force path
 * (CharmmContext::getXYZQ); valid until apo_lt_destroy. */
APO_API apo_status apo_lt_xyzq(apo_lt_integrator *lt, const float **d_xyzq);

#ifdef __cplusplus
}
#endif
#endif
