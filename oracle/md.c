/*
 * FP64 CPU oracle for the Langevin thermostat leapfrog integrator with
 * SETTLE water constraints.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP integrator step, never
 * linked or called by the product library.
 *
 * Arrays are AoS doubles [n][3] (x0 y0 z0 x1 ...).  The time step `dt` is in
 * AKMA units (ps / 0.0488882129), `gamma` = dt[ps] * friction[1/ps]
 * (dimensionless), `kbt` = kBoltz * T in kcal/mol, invmass = 1/m.
 *
 * Random numbers: this build's own counter-based stream.  Philox4x32-10
 * (Salmon, Moraes, Dror & Shaw, SC'11) with counter (atom, step_lo, step_hi,
 * 0x4C414E47) and key (seed_lo, seed_hi); the four 32-bit words become
 * uniforms a_k = (u_k + 0.5) 2^-32 and three normals by Box-Muller:
 *   n0 = sqrt(-2 ln a0) cos(2 pi a1), n1 = sqrt(-2 ln a0) sin(2 pi a1),
 *   n2 = sqrt(-2 ln a2) cos(2 pi a3).
 * The reference draws from cuRAND's Philox with its own mapping
 * (src/CudaLangevinThermostatIntegrator.cu:585-587); only the distribution,
 * not the bits, is shared.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define ORC_PHILOX_M0 0xD2511F53u
#define ORC_PHILOX_M1 0xCD9E8D57u
#define ORC_PHILOX_W0 0x9E3779B9u
#define ORC_PHILOX_W1 0xBB67AE85u
#define ORC_LANGEVIN_TAG 0x4C414E47u

void orc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += ORC_PHILOX_W0;
      k1 += ORC_PHILOX_W1;
    }
    const uint64_t p0 = (uint64_t)ORC_PHILOX_M0 * c0;
    const uint64_t p1 = (uint64_t)ORC_PHILOX_M1 * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void orc_langevin_normals(uint64_t seed, uint64_t step, int atom, double *n3) {
  const uint32_t ctr[4] = {(uint32_t)atom, (uint32_t)step, (uint32_t)(step >> 32),
                           ORC_LANGEVIN_TAG};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t u[4];
  orc_philox4x32_10(ctr, key, u);
  double a[4];
  for (int k = 0; k < 4; k++) a[k] = ((double)u[k] + 0.5) * 2.3283064365386963e-10;
  const double two_pi = 6.283185307179586;
  const double r0 = sqrt(-2.0 * log(a[0]));
  const double r1 = sqrt(-2.0 * log(a[2]));
  n3[0] = r0 * cos(two_pi * a[1]);
  n3[1] = r0 * sin(two_pi * a[1]);
  n3[2] = r1 * cos(two_pi * a[3]);
}

/* ---- SETTLE (Miyamoto & Kollman, J. Comput. Chem. 13, 952 (1992)) ---------
 *
 * Rigid water (O, H1, H2) with masses mO, mH, bond rOH and H-H distance rHH.
 * Given the constrained reference positions (xref) and the unconstrained new
 * positions (x), the constraint displacements are along the reference bonds.
 * Those all lie in the reference plane, so in a frame whose z axis is the
 * reference normal every atom keeps its z.  The frame's x axis is
 * perpendicular to z and to the new oxygen's offset from the centre of mass,
 * so the oxygen lies in the y-z plane.  The canonical triangle (O at
 * (0, ra, 0), H at (-/+rc, -rb, 0), centre of mass at the origin) is tilted
 * by phi about x and psi about y to match those z values, then turned by
 * theta about z such that the mass-weighted moment of the displacements about
 * the reference positions vanishes, which is the Lagrange-multiplier form:
 *   alpha sin(theta) + beta cos(theta) = gamma.
 * Follows src/CudaHolonomicConstraint.cu:164-303 (same equations and root).
 */
static void v3sub(const double *a, const double *b, double *o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
static void v3cross(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static double v3dot(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static void v3unit(double *a) {
  const double s = 1.0 / sqrt(v3dot(a, a));
  a[0] *= s;
  a[1] *= s;
  a[2] *= s;
}

void orc_settle_geometry(const double *geom, double *abc) {
  const double mO = geom[0], mH = geom[1], rOH = geom[2], rHH = geom[3];
  const double M = mO + 2.0 * mH;
  const double rc = 0.5 * rHH;
  const double h = sqrt(rOH * rOH - rc * rc); /* O to H-H midpoint */
  abc[0] = 2.0 * mH / M * h;                  /* ra: centre of mass to O */
  abc[1] = h - abc[0];                        /* rb: centre of mass to H-H midpoint */
  abc[2] = rc;
}

void orc_settle(int nw, const int *w, const double *geom, const double *xref, double *x) {
  double abc[3];
  orc_settle_geometry(geom, abc);
  const double ra = abc[0], rb = abc[1], rc = abc[2];
  const double M = geom[0] + 2.0 * geom[1];
  const double wO = geom[0] / M, wH = geom[1] / M;
  for (int m = 0; m < nw; m++) {
    const int ia = w[3 * m], ib = w[3 * m + 1], ic = w[3 * m + 2];
    double *A1 = x + 3 * ia, *B1 = x + 3 * ib, *C1 = x + 3 * ic;
    double com[3], a1[3], b1[3], c1[3], b0[3], c0[3];
    for (int d = 0; d < 3; d++) com[d] = wO * A1[d] + wH * (B1[d] + C1[d]);
    v3sub(A1, com, a1);
    v3sub(B1, com, b1);
    v3sub(C1, com, c1);
    v3sub(xref + 3 * ib, xref + 3 * ia, b0);
    v3sub(xref + 3 * ic, xref + 3 * ia, c0);
    double ez[3], ex[3], ey[3];
    v3cross(b0, c0, ez);
    v3cross(a1, ez, ex);
    v3cross(ez, ex, ey);
    v3unit(ex);
    v3unit(ey);
    v3unit(ez);
    /* coordinates in the frame (e_x, e_y, e_z) */
    const double Xb0 = v3dot(ex, b0), Yb0 = v3dot(ey, b0);
    const double Xc0 = v3dot(ex, c0), Yc0 = v3dot(ey, c0);
    const double Za1 = v3dot(ez, a1);
    const double Xb1 = v3dot(ex, b1), Yb1 = v3dot(ey, b1), Zb1 = v3dot(ez, b1);
    const double Xc1 = v3dot(ex, c1), Yc1 = v3dot(ey, c1), Zc1 = v3dot(ez, c1);
    /* tilt about x (phi) and y (psi) from the z values */
    const double sphi = Za1 / ra;
    const double cphi = sqrt(1.0 - sphi * sphi);
    const double spsi = (Zb1 - Zc1) / (2.0 * rc * cphi);
    const double cpsi = sqrt(1.0 - spsi * spsi);
    const double Ya2 = ra * cphi;
    const double Xb2 = -rc * cpsi;
    const double Yb2 = -rb * cphi - rc * spsi * sphi;
    const double Yc2 = -rb * cphi + rc * spsi * sphi;
    /* turn about z: zero moment of the displacements about the reference */
    const double al = Xb2 * (Xb0 - Xc0) + Yb0 * Yb2 + Yc0 * Yc2;
    const double be = Xb2 * (Yc0 - Yb0) + Xb0 * Yb2 + Xc0 * Yc2;
    const double ga = Xb0 * Yb1 - Xb1 * Yb0 + Xc0 * Yc1 - Xc1 * Yc0;
    const double s2 = al * al + be * be;
    const double sth = (al * ga - be * sqrt(s2 - ga * ga)) / s2;
    const double cth = sqrt(1.0 - sth * sth);
    const double A3[3] = {-Ya2 * sth, Ya2 * cth, Za1};
    const double B3[3] = {Xb2 * cth - Yb2 * sth, Xb2 * sth + Yb2 * cth, Zb1};
    const double C3[3] = {-Xb2 * cth - Yc2 * sth, -Xb2 * sth + Yc2 * cth, Zc1};
    for (int d = 0; d < 3; d++) {
      A1[d] = com[d] + ex[d] * A3[0] + ey[d] * A3[1] + ez[d] * A3[2];
      B1[d] = com[d] + ex[d] * B3[0] + ey[d] * B3[1] + ez[d] * B3[2];
      C1[d] = com[d] + ex[d] * C3[0] + ey[d] * C3[1] + ez[d] * C3[2];
    }
  }
}

/* Iterative SHAKE on the three water distances, to tolerance `tol` on
 * |r^2 - d^2| / d^2.  An independent solver of the same constraint problem,
 * used only to check orc_settle.  Returns the iterations used, -1 if not
 * converged. */
int orc_shake_water(int nw, const int *w, const double *geom, const double *xref, double *x,
                    double tol, int maxit) {
  const double m[3] = {geom[0], geom[1], geom[1]};
  const double d2[3] = {geom[2] * geom[2], geom[2] * geom[2], geom[3] * geom[3]};
  const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
  int worst = 0;
  for (int k = 0; k < nw; k++) {
    int it = 0;
    for (; it < maxit; it++) {
      int done = 1;
      for (int c = 0; c < 3; c++) {
        const int i = w[3 * k + pa[c]], j = w[3 * k + pb[c]];
        double r[3], r0[3];
        v3sub(x + 3 * i, x + 3 * j, r);
        v3sub(xref + 3 * i, xref + 3 * j, r0);
        const double diff = d2[c] - v3dot(r, r);
        if (fabs(diff) > tol * d2[c]) done = 0;
        const double mi = 1.0 / m[pa[c]], mj = 1.0 / m[pb[c]];
        const double g = diff / (2.0 * (mi + mj) * v3dot(r, r0));
        for (int d = 0; d < 3; d++) {
          x[3 * i + d] += g * mi * r0[d];
          x[3 * j + d] -= g * mj * r0[d];
        }
      }
      if (done) break;
    }
    if (it == maxit) return -1;
    if (it > worst) worst = it;
  }
  return worst;
}

/* SHAKE of bonds to hydrogen (src/CudaHolonomicConstraint.cu:324-520 solve
 * the same constraints): groups of a heavy atom and up to three hydrogens
 * (g[4k] heavy, g[4k+1..3] hydrogens or -1), every bond of group k held at
 * squared length r2[k], displacements along the reference bonds weighted by
 * the inverse masses.  Classic iterative SHAKE (Ryckaert, Ciccotti and
 * Berendsen 1977): bonds corrected one after another until every
 * |r^2 - d^2| <= tol d^2 (FP64 rounding of box-scale coordinates limits tol to
 * ~1e-14).  Every group is processed; returns the most iterations a group
 * used, -1 if any did not converge in maxit sweeps. */
int orc_shake_groups(int ng, const int *g, const double *r2, const double *invmass,
                     const double *xref, double *x, double tol, int maxit) {
  int worst = 0, failed = 0;
  for (int k = 0; k < ng; k++) {
    const int h0 = g[4 * k];
    int it = 0;
    for (; it < maxit; it++) {
      int done = 1;
      for (int c = 1; c < 4; c++) {
        const int j = g[4 * k + c];
        if (j < 0) continue;
        double r[3], r0[3];
        v3sub(x + 3 * h0, x + 3 * j, r);
        v3sub(xref + 3 * h0, xref + 3 * j, r0);
        const double diff = r2[k] - v3dot(r, r);
        if (fabs(diff) > tol * r2[k]) done = 0;
        const double mi = invmass[h0], mj = invmass[j];
        const double gg = diff / (2.0 * (mi + mj) * v3dot(r, r0));
        for (int d = 0; d < 3; d++) {
          x[3 * h0 + d] += gg * mi * r0[d];
          x[3 * j + d] -= gg * mj * r0[d];
        }
      }
      if (done) break;
    }
    if (it == maxit) failed = 1;
    if (it > worst) worst = it;
  }
  return failed ? -1 : worst;
}

/* SETTLE of the waters and SHAKE of the hydrogen groups, to FP64 rounding. */
static void constrain_all(int nw, const int *w, const double *geom, int ns, const int *sg,
                          const double *sr2, const double *invmass, const double *xref,
                          double *x) {
  orc_settle(nw, w, geom, xref, x);
  if (ns > 0) orc_shake_groups(ns, sg, sr2, invmass, xref, x, 1e-13, 1000);
}

/* First half of CudaLangevinThermostatIntegrator::initialize
 * (src/CudaLangevinThermostatIntegrator.cu:217-232 with
 * src/CudaHolonomicConstraint.cu:94-108, 638-676): constrain the coordinates
 * against themselves and fold the correction into the velocities. */
void orc_lt_init_constrain(int n, int nw, const int *w, const double *geom, int ns, const int *sg,
                           const double *sr2, const double *invmass, double dt, double *x,
                           double *v, double *scratch) {
  memcpy(scratch, x, sizeof(double) * 3 * (size_t)n);
  constrain_all(nw, w, geom, ns, sg, sr2, invmass, scratch, x);
  for (int k = 0; k < 3 * n; k++) v[k] += (x[k] - scratch[k]) / dt;
}

/* Second half (src/CudaLangevinThermostatIntegrator.cu:138-167, 169-203,
 * 234-258): leapfrog back-step from the gradient at x, constrained against
 * x.  Writes the previous half-step displacement dprev; x is unchanged. */
void orc_lt_init_backstep(int n, const double *invmass, int nw, const int *w, const double *geom,
                          int ns, const int *sg, const double *sr2, double dt, double gamma,
                          const double *grad, const double *x, const double *v, double *dprev,
                          double *scratch) {
  const double hg = 0.5 * gamma;
  const double a1 = dt * (1.0 - hg) / sqrt(1.0 + hg);
  for (int i = 0; i < n; i++) {
    const double fact = 0.5 * dt * dt * invmass[i];
    for (int d = 0; d < 3; d++) {
      const int k = 3 * i + d;
      dprev[k] = a1 * v[k] + fact * grad[k];
      scratch[k] = x[k] - dprev[k];
    }
  }
  constrain_all(nw, w, geom, ns, sg, sr2, invmass, x, scratch);
  for (int k = 0; k < 3 * n; k++) dprev[k] = x[k] - scratch[k];
}

/* Mass-weighted mean of dprev removed
 * (src/CudaLangevinThermostatIntegrator.cu:832-878, P1 branch). */
void orc_lt_remove_com(int n, const double *invmass, double *dprev) {
  double c[3] = {0, 0, 0}, mt = 0;
  for (int i = 0; i < n; i++) {
    const double m = 1.0 / invmass[i];
    for (int d = 0; d < 3; d++) c[d] += dprev[3 * i + d] * m;
    mt += m;
  }
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) dprev[3 * i + d] -= c[d] / mt;
}

/* One propagateOneStep after the force call
 * (src/CudaLangevinThermostatIntegrator.cu:576-614, 616-637, 639-682,
 * 742-796): leapfrog Langevin position update with noise, SETTLE against the
 * start-of-step coordinates, velocity from the two half-step displacements,
 * kinetic energies ke[0] (mean of the old half-step, on-step and new
 * half-step estimates) and ke[1] (on-step), then dprev <- new displacement. */
void orc_lt_step(int n, const double *invmass, int nw, const int *w, const double *geom, int ns,
                 const int *sg, const double *sr2, double dt, double gamma, double kbt,
                 uint64_t seed, uint64_t step,
                 const double *grad, double *x, double *dprev, double *v, double *ke,
                 double *scratch) {
  const double hg = 0.5 * gamma;
  const double den = 1.0 / (1.0 + hg);
  const double alpha = (1.0 - hg) * den;
  double *xref = scratch;
  memcpy(xref, x, sizeof(double) * 3 * (size_t)n);
  for (int i = 0; i < n; i++) {
    double rn[3];
    orc_langevin_normals(seed, step, i, rn);
    const double rdf = sqrt(2.0 * gamma * kbt / invmass[i]) / dt;
    const double fact = dt * dt * invmass[i] * den;
    for (int d = 0; d < 3; d++) {
      const int k = 3 * i + d;
      const double g = grad[k] + rdf * rn[d];
      x[k] += alpha * dprev[k] - fact * g;
    }
  }
  constrain_all(nw, w, geom, ns, sg, sr2, invmass, xref, x);
  const double fv = 0.5 * sqrt(1.0 + hg) / dt;
  const double idt = 1.0 / dt;
  ke[0] = ke[1] = 0.0;
  for (int i = 0; i < n; i++) {
    double kold = 0, kon = 0, knew = 0;
    for (int d = 0; d < 3; d++) {
      const int k = 3 * i + d;
      const double dnew = x[k] - xref[k];
      v[k] = fv * (dnew + dprev[k]);
      kold += dprev[k] * idt * dprev[k] * idt;
      kon += v[k] * v[k];
      knew += dnew * idt * dnew * idt;
      dprev[k] = dnew;
    }
    kold = 0.5 * kold / invmass[i];
    kon = 0.5 * kon / invmass[i];
    knew = 0.5 * knew / invmass[i];
    ke[0] += (1.0 / 3.0) * (kold + kon + knew);
    ke[1] += kon;
  }
}

/* Group image centering (src/CharmmContext.cu:276-372, P1): each group whose
 * unweighted centroid lies outside [-L/2, L/2] on an axis moves by one box
 * edge on that axis.  Groups are CSR (gptr[ng+1], gatom). */
void orc_center(int ng, const int *gptr, const int *gatom, const double *box, double *x) {
  for (int g = 0; g < ng; g++) {
    double c[3] = {0, 0, 0};
    const int n = gptr[g + 1] - gptr[g];
    for (int k = gptr[g]; k < gptr[g + 1]; k++)
      for (int d = 0; d < 3; d++) c[d] += x[3 * gatom[k] + d];
    for (int d = 0; d < 3; d++) {
      c[d] /= n;
      double s = 0.0;
      if (c[d] < -0.5 * box[d]) s = box[d];
      if (c[d] > 0.5 * box[d]) s = -box[d];
      if (s != 0.0)
        for (int k = gptr[g]; k < gptr[g + 1]; k++) x[3 * gatom[k] + d] += s;
    }
  }
}
