/*
 * FP64 CPU oracle for the CHARMM nonbonded + bonded force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against (and the timed CPU baseline).  Nothing in the
 * product library links, loads or calls it.
 *
 * Every routine restates the reference algorithm in double precision; each
 * function cites the reference file:line it follows.  Conventions:
 *   - gradients (dU/dx, kcal/mol/A) are ACCUMULATED into grad[3*n] laid out
 *     as x0 y0 z0 x1 y1 z1 ... ;
 *   - virials are accumulated into virial[9] (row-major, W = -sum x (x) g);
 *   - coordinates/charges come as xyzq[4*n] (x, y, z, q).
 */
#ifndef APO_ORACLE_H
#define APO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_CCELEC 332.0716

/* Direct-space PME pair sum (erfc Coulomb + CHARMM VFSW Lennard-Jones) over
 * all minimum-image pairs with r < roff that are not in the exclusion CSR.
 * vdwtab holds (c6, c12) per lower-triangular type pair, pair index
 * p = a*(a+1)/2 + b for a >= b, with U_LJ = c12/r^12 - c6/r^6 unswitched.
 * energy[0] += elec, energy[1] += vdw.  Returns 0 on success. */
int orc_direct(int n, const double *xyzq, const int *vdwtype, int ntypes,
               const double *vdwtab, const double *box, double kappa,
               double roff, double ron, const int *excl_ptr,
               const int *excl_idx, double *grad, double *energy,
               double *virial, int nthreads);

/* 1-4 inclusion pairs (elec qq(erfc+e14fac-1)/r and VFSW LJ with vdwtab14)
 * and 1-2/1-3 exclusion corrections qq(erfc-1)/r.  No cutoff, no image.
 * energy[0] += elec(1-4), energy[1] += vdw(1-4), energy[2] += ewex. */
int orc_pairs14(int nin14, const int *in14, int nex14, const int *ex14,
                const double *xyzq, const int *vdwtype, int ntypes,
                const double *vdwtab14, double kappa, double roff, double ron,
                double e14fac, double *grad, double *energy, double *virial);

/* Smooth PME reciprocal sum, orthorhombic box, B-spline order `order`.
 * energy[0] += ewks.  Grid layout and fractional-coordinate origin follow
 * the reference spread/gather exactly. */
int orc_recip(int n, const double *xyzq, const double *box, const int *K,
              int order, double kappa, double *grad, double *energy,
              double *virial, int nthreads);

/* Ewald self energy -kappa*ccelec/sqrt(pi)*sum q^2. */
double orc_self_energy(int n, const double *xyzq, double kappa);

/* Bonded terms.  Index arrays are 0-based.
 *  bonds    : nb   x (i, j)          bond_par : nb x (b0, kb)   E = kb (r-b0)^2
 *  ureyb    : nub  x (i, k)          ub_par   : nub x (s0, kub)
 *  angles   : na   x (i, j, k)       ang_par  : na x (theta0[rad], ktheta)
 *  dihedrals: nd   x (i, j, k, l)    dihe_par : nd x (n, k, delta[rad])
 *             (several terms on one quadruplet = several entries)
 *  impropers: ni   x (i, j, k, l)    imp_par  : ni x (n, k, delta[rad])
 *             n == 0 -> harmonic k (phi - delta)^2
 * energy[0..4] += bond, angle, ureyb, dihe, imdihe. */
int orc_bonded(const double *xyzq, int nb, const int *bonds,
               const double *bond_par, int nub, const int *ureyb,
               const double *ub_par, int na, const int *angles,
               const double *ang_par, int nd, const int *dihe,
               const double *dihe_par, int ni, const int *imp,
               const double *imp_par, double *grad, double *energy,
               double *virial);

/* Unnormalised in-place 3-D complex DFT on a [K2][K1][K0] (x fastest) grid
 * of interleaved (re, im) doubles; sign = -1 forward, +1 backward. */
int orc_fft3d(const int *K, double *data, int sign, int nthreads);

/* Reciprocal building blocks exposed for closed-form checks. */
void orc_bspline(int order, double w, double *theta, double *dtheta);
void orc_bsp_mod(int order, int nfft, double *bsp_mod);


/* ---- Langevin thermostat leapfrog + SETTLE (md.c) ---------------------------
 * AoS [n][3] doubles; dt in AKMA time units, gamma = dt[ps] * friction[1/ps],
 * kbt in kcal/mol; geom = (mO, mH, rOH, rHH) of the rigid water; waters are
 * (O, H1, H2) index triplets. */
void orc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
void orc_langevin_normals(uint64_t seed, uint64_t step, int atom, double *n3);
void orc_settle_geometry(const double *geom, double *abc);
void orc_settle(int nw, const int *w, const double *geom, const double *xref, double *x);
int orc_shake_water(int nw, const int *w, const double *geom, const double *xref, double *x,
                    double tol, int maxit);
int orc_shake_groups(int ng, const int *g, const double *r2, const double *invmass,
                     const double *xref, double *x, double tol, int maxit);
void orc_lt_init_constrain(int n, int nw, const int *w, const double *geom, int ns, const int *sg,
                           const double *sr2, const double *invmass, double dt, double *x,
                           double *v, double *scratch);
void orc_lt_init_backstep(int n, const double *invmass, int nw, const int *w, const double *geom,
                          int ns, const int *sg, const double *sr2, double dt, double gamma,
                          const double *grad, const double *x, const double *v, double *dprev,
                          double *scratch);
void orc_lt_remove_com(int n, const double *invmass, double *dprev);
void orc_lt_step(int n, const double *invmass, int nw, const int *w, const double *geom, int ns,
                 const int *sg, const double *sr2, double dt, double gamma, double kbt,
                 uint64_t seed, uint64_t step, const double *grad, double *x, double *dprev,
                 double *v, double *ke, double *scratch);
void orc_center(int ng, const int *gptr, const int *gatom, const double *box, double *x);

#ifdef __cplusplus
}
#endif
#endif
