// Langevin thermostat leapfrog integrator with SETTLE water constraints on
// top of the force path (CudaLangevinThermostatIntegrator,
// src/CudaLangevinThermostatIntegrator.cu:206-258, 576-800, and
// CudaHolonomicConstraint, src/CudaHolonomicConstraint.cu:164-303).
//
// State lives in HBM in FP64 SoA (x, previous half-step displacement, on-step
// velocity), plus the float4 xyzq the force path reads and the int64 2^-40
// gradient it writes.  One step is:
//   [rebuild steps: image centering + list rebuild]
//   calc_force(xyzq) -> int64 gradient (the pair part left in sorted order)
//   k_lt_step: three waves per 64 waters (one per atom; wave 0 alone for
//              free atoms): the sorted pair gradient added to the rest,
//              leapfrog Langevin update with Philox noise, SETTLE against the
//              start-of-step coordinates (one lane per water, from LDS),
//              velocity, kinetic energy partials, the float xyzq for the next
//              force call
//   k_lt_finalize: fixed-order kinetic energy sum, temperatures, step + 1
// so the integrator costs one pass over the state instead of the reference's
// five kernels and two device copies.  Steps replay as hipGraphs: a whole
// list cycle per graph where the schedule allows, else one step per graph.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "apocharmm_amd.h"
#include "system.hpp"
#include "thermo.hpp"

namespace apo {

void calc_force(NbSystem &S, const float *xyzq, int flags, int64_t *force, int stride,
                double *energy, double *virial, hipStream_t s);
void list_export_check(NbSystem &S, hipStream_t s, bool allow_growth);

// fused (multi-system) launches of the step kernels, one per kernel type
void batch_add_sort(BatchLaunch<SortArgs> &b, NbSystem &S, const float4 *x, int64_t *force,
                    size_t nforce);
void batch_sort(BatchLaunch<SortArgs> &b, hipStream_t s);
void batch_add_pair(BatchLaunch<PairArgs> &b, NbSystem &S, int blocks);
void batch_pair(BatchLaunch<PairArgs> &b, int tab_mode, bool iso, hipStream_t s);
int pair_tab_lds(const NbSystem &S);
bool pair_tab_iso(const NbSystem &S);
struct PmeBatch;
PmeBatch *pme_batch_create();
void pme_batch_destroy(PmeBatch *b);
void pme_batch_clear(PmeBatch &b);
void pme_batch_add(PmeBatch &b, NbSystem &S, int64_t *force, int stride, int fft_rows,
                   int fft_cols, int fft_threads);
void pme_batch_upload(PmeBatch &b, hipStream_t s);
void pme_batch_launch(PmeBatch &b, hipStream_t s, bool gather);
struct BondedBatch;
BondedBatch *bonded_batch_create();
void bonded_batch_destroy(BondedBatch *b);
void bonded_batch_clear(BondedBatch &b);
void bonded_batch_add(BondedBatch &b, NbSystem &S, const float4 *x, int64_t *force, int stride);
void bonded_batch_upload(BondedBatch &b, hipStream_t s);
void bonded_batch_launch(BondedBatch &b, hipStream_t s);

namespace {

constexpr double kTimfac = 0.0488882129;   // AKMA time unit in ps (src/CudaIntegrator.cu:22)
constexpr unsigned kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr unsigned kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;
constexpr unsigned kLangevinTag = 0x4C414E47u;
constexpr int kStepBlock = 64;             // one wave per block in the step kernels

// ---- Philox4x32-10 (Salmon et al., SC'11) and Box-Muller normals ----------
// Counter (atom, step_lo, step_hi, tag), key (seed_lo, seed_hi): stateless, so
// a graph replay or a restart at any step draws the same numbers.
__device__ inline void philox4x32_10(unsigned c[4], unsigned k0, unsigned k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    const unsigned h0 = __umulhi(kPhiloxM0, c[0]), l0 = kPhiloxM0 * c[0];
    const unsigned h1 = __umulhi(kPhiloxM1, c[2]), l1 = kPhiloxM1 * c[2];
    const unsigned n0 = h1 ^ c[1] ^ k0, n2 = h0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = l1;
    c[2] = n2;
    c[3] = l0;
  }
}

__device__ inline double3 langevin_normals(unsigned long long seed, unsigned long long step,
                                           int atom) {
  unsigned c[4] = {(unsigned)atom, (unsigned)step, (unsigned)(step >> 32), kLangevinTag};
  philox4x32_10(c, (unsigned)seed, (unsigned)(seed >> 32));
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double a0 = ((double)c[0] + 0.5) * s, a1 = ((double)c[1] + 0.5) * s;
  const double a2 = ((double)c[2] + 0.5) * s, a3 = ((double)c[3] + 0.5) * s;
  const double two_pi = 6.283185307179586;
  const double r0 = sqrt(-2.0 * log(a0)), r1 = sqrt(-2.0 * log(a2));
  return make_double3(r0 * cos(two_pi * a1), r0 * sin(two_pi * a1), r1 * cos(two_pi * a3));
}

// ---- small double3 algebra -----------------------------------------------------
__device__ inline double3 operator-(double3 a, double3 b) {
  return make_double3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ inline double dot3(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double3 cross3(double3 a, double3 b) {
  return make_double3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline double3 unit3(double3 a) {
  const double s = 1.0 / sqrt(dot3(a, a));
  return make_double3(a.x * s, a.y * s, a.z * s);
}
__device__ inline double3 frame_to_lab(double3 o, double3 ex, double3 ey, double3 ez, double u,
                                       double v, double w) {
  return make_double3(o.x + ex.x * u + ey.x * v + ez.x * w, o.y + ex.y * u + ey.y * v + ez.y * w,
                      o.z + ex.z * u + ey.z * v + ez.z * w);
}

struct SettleGeom {
  double ra, rb, rc;  // centre of mass to O, to the H-H midpoint, half H-H
  double wO, wH;      // mass fractions
};

// SETTLE (Miyamoto & Kollman 1992): the analytic solution of the three
// distance constraints with displacements along the reference bonds.  The
// constraint displacements lie in the reference plane, so in the frame with
// z along the reference normal every atom keeps its z; the canonical triangle
// is tilted (phi about x, psi about y) to those heights, then turned about z
// by theta with alpha sin(theta) + beta cos(theta) = gamma, the condition that
// the mass-weighted moment of the displacements about the reference vanishes.
// Same equations and root as src/CudaHolonomicConstraint.cu:164-303.
__device__ inline void settle(const SettleGeom &G, double3 A0, double3 B0, double3 C0, double3 &A1,
                              double3 &B1, double3 &C1) {
  const double3 com = make_double3(G.wO * A1.x + G.wH * (B1.x + C1.x),
                                   G.wO * A1.y + G.wH * (B1.y + C1.y),
                                   G.wO * A1.z + G.wH * (B1.z + C1.z));
  const double3 a1 = A1 - com, b1 = B1 - com, c1 = C1 - com;
  const double3 b0 = B0 - A0, c0 = C0 - A0;
  double3 ez = cross3(b0, c0);
  double3 ex = cross3(a1, ez);
  double3 ey = cross3(ez, ex);
  ex = unit3(ex);
  ey = unit3(ey);
  ez = unit3(ez);
  const double Xb0 = dot3(ex, b0), Yb0 = dot3(ey, b0);
  const double Xc0 = dot3(ex, c0), Yc0 = dot3(ey, c0);
  const double Za1 = dot3(ez, a1);
  const double Xb1 = dot3(ex, b1), Yb1 = dot3(ey, b1), Zb1 = dot3(ez, b1);
  const double Xc1 = dot3(ex, c1), Yc1 = dot3(ey, c1), Zc1 = dot3(ez, c1);
  const double sphi = Za1 / G.ra;
  const double cphi = sqrt(1.0 - sphi * sphi);
  const double spsi = (Zb1 - Zc1) / (2.0 * G.rc * cphi);
  const double cpsi = sqrt(1.0 - spsi * spsi);
  const double Ya2 = G.ra * cphi;
  const double Xb2 = -G.rc * cpsi;
  const double Yb2 = -G.rb * cphi - G.rc * spsi * sphi;
  const double Yc2 = -G.rb * cphi + G.rc * spsi * sphi;
  const double al = Xb2 * (Xb0 - Xc0) + Yb0 * Yb2 + Yc0 * Yc2;
  const double be = Xb2 * (Yc0 - Yb0) + Xb0 * Yb2 + Xc0 * Yc2;
  const double ga = Xb0 * Yb1 - Xb1 * Yb0 + Xc0 * Yc1 - Xc1 * Yc0;
  const double s2 = al * al + be * be;
  const double sth = (al * ga - be * sqrt(s2 - ga * ga)) / s2;
  const double cth = sqrt(1.0 - sth * sth);
  A1 = frame_to_lab(com, ex, ey, ez, -Ya2 * sth, Ya2 * cth, Za1);
  B1 = frame_to_lab(com, ex, ey, ez, Xb2 * cth - Yb2 * sth, Xb2 * sth + Yb2 * cth, Zb1);
  C1 = frame_to_lab(com, ex, ey, ez, -Xb2 * cth - Yc2 * sth, -Xb2 * sth + Yc2 * cth, Zc1);
}

// SHAKE for one group: a heavy atom (slot 0) and M = 1..3 hydrogens bonded to
// it, every bond held at squared length r2 (the reference's Shake2/3/4Kernel,
// src/CudaHolonomicConstraint.cu:324-520, solve the same constraints).  The
// constraint displacements lie along the reference bonds d_k = ref0 - refk:
// x0 += im0 sum_k l_k d_k, xk -= imk l_k d_k.  The M equations
// |x0 - xk|^2 = r2 are solved for l by Newton's method in FP64 (Jacobian
// J_kj = 2 p_k . (im0 d_j + [k == j] imk d_k), p_k = x0 - xk), to a relative
// residual of 1e-14 or 30 iterations; each Newton system by Gaussian
// elimination with partial pivoting (row swaps).  M is a template argument so
// that every array index is known at compile time and the arrays stay in
// registers.  Momentum is conserved by construction.
template <int M>
__device__ inline void shake_group(double r2, const double (&im)[M + 1],
                                   const double3 (&ref)[M + 1], double3 (&x)[M + 1]) {
  double3 d[M];
#pragma unroll
  for (int k = 0; k < M; k++) d[k] = ref[0] - ref[k + 1];
  double lam[M];
#pragma unroll
  for (int k = 0; k < M; k++) lam[k] = 0.0;
  const double3 x0 = x[0];
  double3 xs[M + 1];
#pragma unroll
  for (int k = 0; k <= M; k++) xs[k] = x[k];
  for (int it = 0; it < 30; it++) {
    // current positions from the multipliers
    double3 c0 = x0;
#pragma unroll
    for (int k = 0; k < M; k++) {
      c0.x += im[0] * lam[k] * d[k].x;
      c0.y += im[0] * lam[k] * d[k].y;
      c0.z += im[0] * lam[k] * d[k].z;
    }
    double3 p[M];
    double f[M], worst = 0.0;
#pragma unroll
    for (int k = 0; k < M; k++) {
      const double3 ck = make_double3(x[k + 1].x - im[k + 1] * lam[k] * d[k].x,
                                      x[k + 1].y - im[k + 1] * lam[k] * d[k].y,
                                      x[k + 1].z - im[k + 1] * lam[k] * d[k].z);
      xs[k + 1] = ck;
      p[k] = c0 - ck;
      f[k] = dot3(p[k], p[k]) - r2;
      worst = fmax(worst, fabs(f[k]));
    }
    xs[0] = c0;
    if (worst <= 1e-14 * r2) break;
    double J[M][M], b[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
      b[k] = -f[k];
#pragma unroll
      for (int j = 0; j < M; j++) {
        double3 t = make_double3(im[0] * d[j].x, im[0] * d[j].y, im[0] * d[j].z);
        if (k == j) {
          t.x += im[k + 1] * d[k].x;
          t.y += im[k + 1] * d[k].y;
          t.z += im[k + 1] * d[k].z;
        }
        J[k][j] = 2.0 * dot3(p[k], t);
      }
    }
    // J dl = -f: for each column the row with the largest magnitude is swapped
    // into the pivot position, then the rows below are eliminated
#pragma unroll
    for (int c = 0; c < M; c++) {
#pragma unroll
      for (int r = c + 1; r < M; r++)
        if (fabs(J[r][c]) > fabs(J[c][c])) {
#pragma unroll
          for (int j = 0; j < M; j++) {
            const double t = J[r][j];
            J[r][j] = J[c][j];
            J[c][j] = t;
          }
          const double t = b[r];
          b[r] = b[c];
          b[c] = t;
        }
#pragma unroll
      for (int r = c + 1; r < M; r++) {
        const double g = J[r][c] / J[c][c];
#pragma unroll
        for (int j = c; j < M; j++) J[r][j] -= g * J[c][j];
        b[r] -= g * b[c];
      }
    }
    double dl[M];
#pragma unroll
    for (int c = M - 1; c >= 0; c--) {
      double sum = b[c];
#pragma unroll
      for (int j = c + 1; j < M; j++) sum -= J[c][j] * dl[j];
      dl[c] = sum / J[c][c];
    }
#pragma unroll
    for (int k = 0; k < M; k++) lam[k] += dl[k];
  }
#pragma unroll
  for (int k = 0; k <= M; k++) x[k] = xs[k];
}

// Calls f(std::integral_constant<int, M>{}) with M the hydrogen count of
// SHAKE group g (heavy, h1, h2 or -1, h3 or -1).
template <typename F> __device__ __forceinline__ void with_group_size(const int4 &g, F &&f) {
  if (g.w >= 0) f(std::integral_constant<int, 3>{});
  else if (g.z >= 0) f(std::integral_constant<int, 2>{});
  else f(std::integral_constant<int, 1>{});
}

struct LtState {
  double *x, *dp, *v;        // SoA, stride n
  const double *invmass;
  const float *charge;
  float4 *xyzq;
  int64_t *grad;             // SoA, stride gstride, 2^-40
  int gstride;
  // when set, the pair part of the gradient is still in the force system's
  // sorted buffer (stride ss, slot glo2loc[i]): the step kernel adds it and
  // stores the total back into grad
  const int64_t *fsort;
  const int *glo2loc;
  int ss;
  int n;
  const int *waters;         // 3 * nwater
  int nwater;
  const int4 *shake;         // nshake SHAKE groups (heavy, h1, h2 / -1, h3 / -1)
  const double *shake_r2;    // squared bond length per group
  int nshake;
  const int *singles;        // atoms in no water and no SHAKE group
  int nsingle;
  const float4 *xref;        // coordinates of the last list build
};

struct LtCoef {
  double alpha;  // (1 - gamma/2) / (1 + gamma/2)
  double den;    // 1 / (1 + gamma/2)
  double dt, dt2, gamma, kbt;
  double fv;     // 0.5 sqrt(1 + gamma/2) / dt
  double idt;
  unsigned long long seed;
  SettleGeom G;
};

__device__ inline double3 ld3(const double *a, int n, int i) {
  return make_double3(a[i], a[n + i], a[2 * n + i]);
}
__device__ inline void st3(double *a, int n, int i, double3 v) {
  a[i] = v.x;
  a[n + i] = v.y;
  a[2 * n + i] = v.z;
}
__device__ inline double3 ldgrad(const LtState &S, int i) {
  const double s = 9.094947017729282e-13;  // 2^-40
  int64_t gx = S.grad[i], gy = S.grad[S.gstride + i], gz = S.grad[2 * (size_t)S.gstride + i];
  if (S.fsort) {  // sorted -> global order (the calc_force finalize, fused here)
    const int l = S.glo2loc[i];
    gx += S.fsort[l];
    gy += S.fsort[S.ss + l];
    gz += S.fsort[2 * (size_t)S.ss + l];
    S.grad[i] = gx;
    S.grad[S.gstride + i] = gy;
    S.grad[2 * (size_t)S.gstride + i] = gz;
  }
  return make_double3((double)gx * s, (double)gy * s, (double)gz * s);
}
__device__ inline void store_xyzq(const LtState &S, int i, double3 x) {
  S.xyzq[i] = make_float4((float)x.x, (float)x.y, (float)x.z, S.charge[i]);
}

// Unconstrained leapfrog Langevin position for one atom
// (src/CudaLangevinThermostatIntegrator.cu:576-614).
__device__ inline double3 lt_position(const LtState &S, const LtCoef &C, unsigned long long step,
                                      int i, double3 x, double3 dp) {
  const double im = S.invmass[i];
  const double3 rn = langevin_normals(C.seed, step, i);
  const double rdf = sqrt(2.0 * C.gamma * C.kbt / im) / C.dt;
  const double fact = C.dt2 * im * C.den;
  const double3 g = ldgrad(S, i);
  const double gx = g.x + rdf * rn.x, gy = g.y + rdf * rn.y, gz = g.z + rdf * rn.z;
  return make_double3(x.x + (C.alpha * dp.x - fact * gx), x.y + (C.alpha * dp.y - fact * gy),
                      x.z + (C.alpha * dp.z - fact * gz));
}

// Velocity, kinetic energies, state store for one atom
// (src/CudaLangevinThermostatIntegrator.cu:616-682, 123-136).
__device__ inline void lt_finish(const LtState &S, const LtCoef &C, int i, double3 xr, double3 xn,
                                 double3 dp, double &ke0, double &ke1) {
  const double im = S.invmass[i];
  const double3 dn = xn - xr;
  const double3 v = make_double3(C.fv * (dn.x + dp.x), C.fv * (dn.y + dp.y), C.fv * (dn.z + dp.z));
  const double3 u1 = make_double3(dp.x * C.idt, dp.y * C.idt, dp.z * C.idt);
  const double3 u2 = make_double3(dn.x * C.idt, dn.y * C.idt, dn.z * C.idt);
  const double kold = 0.5 * dot3(u1, u1) / im;
  const double kon = 0.5 * dot3(v, v) / im;
  const double knew = 0.5 * dot3(u2, u2) / im;
  ke0 += (1.0 / 3.0) * (kold + kon + knew);
  ke1 += kon;
  st3(S.x, S.n, i, xn);
  st3(S.dp, S.n, i, dn);
  st3(S.v, S.n, i, v);
  store_xyzq(S, i, xn);
}

// Squared displacement since the last list build and over this step.
__device__ inline void disp_track(const LtState &S, int i, double3 xr, double3 xn, float &d2,
                                  float &s2) {
  const float4 r = S.xref[i];
  const float dx = (float)xn.x - r.x, dy = (float)xn.y - r.y, dz = (float)xn.z - r.z;
  d2 = fmaxf(d2, dx * dx + dy * dy + dz * dz);
  const double3 st = xn - xr;
  s2 = fmaxf(s2, (float)dot3(st, st));
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Three waves per block of kStepBlock units (a unit is a water, or an atom in
// no water): wave k draws the random force and takes the unconstrained step
// for atom k of each water (wave 0 alone for free atoms), wave 0 then applies
// SETTLE to each water from LDS, and every wave finishes its atoms.  The
// per-atom arithmetic does not depend on the split.
constexpr int kStepThreads = 3 * kStepBlock;
__device__ __forceinline__ void lt_step_body(int bid, const LtState &S, const LtCoef &C,
                                             const unsigned long long *__restrict__ stepc,
                                             double2 *__restrict__ part,
                                             float2 *__restrict__ dpart) {
  __shared__ double old_x[3][3][kStepBlock], new_x[3][3][kStepBlock];
  __shared__ double red_ke[2][3];
  __shared__ float red_d[2][3];
  const int lane = threadIdx.x & (kStepBlock - 1), k = threadIdx.x / kStepBlock;
  const int u = bid * kStepBlock + lane;
  const unsigned long long step = stepc[0];
  const bool water = u < S.nwater;
  // units: the waters, then the atoms in no water and no SHAKE group (the
  // SHAKE groups are stepped by k_lt_shake)
  const bool single = !water && k == 0 && u < S.nwater + S.nsingle;
  const int i = water ? S.waters[3 * u + k] : (single ? S.singles[u - S.nwater] : -1);
  double3 r = make_double3(0, 0, 0), p = r, xn = r;
  if (i >= 0) {
    r = ld3(S.x, S.n, i);
    p = ld3(S.dp, S.n, i);
    xn = lt_position(S, C, step, i, r, p);
  }
  double ke0 = 0.0, ke1 = 0.0;
  float d2 = 0.f, s2 = 0.f;
  if (water) {
    old_x[k][0][lane] = r.x;
    old_x[k][1][lane] = r.y;
    old_x[k][2][lane] = r.z;
    new_x[k][0][lane] = xn.x;
    new_x[k][1][lane] = xn.y;
    new_x[k][2][lane] = xn.z;
  }
  __syncthreads();
  if (water && k == 0) {
    const double3 ra = make_double3(old_x[0][0][lane], old_x[0][1][lane], old_x[0][2][lane]);
    const double3 rb = make_double3(old_x[1][0][lane], old_x[1][1][lane], old_x[1][2][lane]);
    const double3 rc = make_double3(old_x[2][0][lane], old_x[2][1][lane], old_x[2][2][lane]);
    double3 na = make_double3(new_x[0][0][lane], new_x[0][1][lane], new_x[0][2][lane]);
    double3 nb = make_double3(new_x[1][0][lane], new_x[1][1][lane], new_x[1][2][lane]);
    double3 nc = make_double3(new_x[2][0][lane], new_x[2][1][lane], new_x[2][2][lane]);
    settle(C.G, ra, rb, rc, na, nb, nc);
    new_x[0][0][lane] = na.x;
    new_x[0][1][lane] = na.y;
    new_x[0][2][lane] = na.z;
    new_x[1][0][lane] = nb.x;
    new_x[1][1][lane] = nb.y;
    new_x[1][2][lane] = nb.z;
    new_x[2][0][lane] = nc.x;
    new_x[2][1][lane] = nc.y;
    new_x[2][2][lane] = nc.z;
  }
  __syncthreads();
  if (i >= 0) {
    if (water) xn = make_double3(new_x[k][0][lane], new_x[k][1][lane], new_x[k][2][lane]);
    lt_finish(S, C, i, r, xn, p, ke0, ke1);
    disp_track(S, i, r, xn, d2, s2);
  }
  ke0 = wave_sum(ke0);
  ke1 = wave_sum(ke1);
  d2 = wave_max(d2);
  s2 = wave_max(s2);
  if (lane == 0) {
    red_ke[0][k] = ke0;
    red_ke[1][k] = ke1;
    red_d[0][k] = d2;
    red_d[1][k] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[bid] = make_double2(red_ke[0][0] + red_ke[0][1] + red_ke[0][2],
                             red_ke[1][0] + red_ke[1][1] + red_ke[1][2]);
    dpart[bid] = make_float2(fmaxf(fmaxf(red_d[0][0], red_d[0][1]), red_d[0][2]),
                             fmaxf(fmaxf(red_d[1][0], red_d[1][1]), red_d[1][2]));
  }
}
// pend: raised for the kinetic-energy reduction of this step, which the next
// coordinate sort runs (thermo.hpp)
__global__ __launch_bounds__(kStepThreads) void k_lt_step(LtState S, LtCoef C,
                                                          const unsigned long long *__restrict__ stepc,
                                                          double2 *__restrict__ part,
                                                          float2 *__restrict__ dpart,
                                                          unsigned long long *__restrict__ pend) {
  lt_step_body(blockIdx.x, S, C, stepc, part, dpart);
  if (blockIdx.x == 0 && threadIdx.x == 0) *pend = 1ull;
}
struct LtStepArgs {
  LtState S;
  LtCoef C;
  const unsigned long long *stepc;
  double2 *part;
  float2 *dpart;
};
__global__ __launch_bounds__(kStepThreads) void kb_lt_step(const LtStepArgs *__restrict__ tab,
                                                         const int *__restrict__ bstart, int nsys) {
  const int s = batch_system(bstart, nsys);
  const LtStepArgs &a = tab[s];
  lt_step_body(blockIdx.x - bstart[s], a.S, a.C, a.stepc, a.part, a.dpart);
}


// The SHAKE groups of a step: one lane steps, constrains and finishes each
// group's atoms (heavy atom and 1-3 hydrogens).  Its own launch (one wave per
// block), so that the registers of the group solve do not limit the
// occupancy of the water kernel; the blocks' kinetic-energy and displacement
// partials follow the water kernel's (k_lt_finalize sums both).
constexpr int kShakeBlock = 64;
__device__ __forceinline__ void lt_shake_body(int bid, const LtState &S, const LtCoef &C,
                                              const unsigned long long *__restrict__ stepc,
                                              double2 *__restrict__ part,
                                              float2 *__restrict__ dpart) {
  const int u = bid * kShakeBlock + threadIdx.x;
  const unsigned long long step = stepc[0];
  double ke0 = 0.0, ke1 = 0.0;
  float d2 = 0.f, s2 = 0.f;
  if (u < S.nshake) {
    const int4 g = S.shake[u];
    const double r2 = S.shake_r2[u];
    with_group_size(g, [&](auto mc) {
      constexpr int M = decltype(mc)::value;
      const int at[4] = {g.x, g.y, g.z, g.w};
      double im[M + 1];
      double3 ref[M + 1], dp4[M + 1], nw[M + 1];
#pragma unroll
      for (int a = 0; a <= M; a++) {
        im[a] = S.invmass[at[a]];
        ref[a] = ld3(S.x, S.n, at[a]);
        dp4[a] = ld3(S.dp, S.n, at[a]);
        nw[a] = lt_position(S, C, step, at[a], ref[a], dp4[a]);
      }
      shake_group<M>(r2, im, ref, nw);
#pragma unroll
      for (int a = 0; a <= M; a++) {
        lt_finish(S, C, at[a], ref[a], nw[a], dp4[a], ke0, ke1);
        disp_track(S, at[a], ref[a], nw[a], d2, s2);
      }
    });
  }
  ke0 = wave_sum(ke0);
  ke1 = wave_sum(ke1);
  d2 = wave_max(d2);
  s2 = wave_max(s2);
  if (threadIdx.x == 0) {
    part[bid] = make_double2(ke0, ke1);
    dpart[bid] = make_float2(d2, s2);
  }
}
__global__ __launch_bounds__(kShakeBlock) void k_lt_shake(LtState S, LtCoef C,
                                                          const unsigned long long *__restrict__ stepc,
                                                          double2 *__restrict__ part,
                                                          float2 *__restrict__ dpart) {
  lt_shake_body(blockIdx.x, S, C, stepc, part, dpart);
}
__global__ __launch_bounds__(kShakeBlock) void kb_lt_shake(const LtStepArgs *__restrict__ tab,
                                                           const int *__restrict__ bstart,
                                                           int nsys) {
  const int s = batch_system(bstart, nsys);
  const LtStepArgs &a = tab[s];
  lt_shake_body(blockIdx.x - bstart[s], a.S, a.C, a.stepc, a.part, a.dpart);
}

__global__ __launch_bounds__(256) void k_lt_finalize(const double2 *__restrict__ part, int npart,
                                                     double ndof, double *__restrict__ thermo,
                                                     unsigned long long *__restrict__ counters,
                                                     const float2 *__restrict__ dpart,
                                                     const int *__restrict__ meta,
                                                     double *__restrict__ hdisp,
                                                     unsigned long long *__restrict__ pend) {
  lt_finalize_body(part, npart, ndof, thermo, counters, dpart, meta, hdisp, pend);
}
// one block per system
__global__ __launch_bounds__(256) void kb_lt_finalize(const LtFinArgs *__restrict__ tab) {
  const LtFinArgs &a = tab[blockIdx.x];
  lt_finalize_body(a.part, a.npart, a.ndof, a.thermo, a.counters, a.dpart, a.meta, a.hdisp,
                   a.pend);
}


// ---- initialisation (src/CudaLangevinThermostatIntegrator.cu:206-258) ------
// Constrain x against itself, fold the correction into v, refresh xyzq.
__global__ __launch_bounds__(kStepBlock) void k_lt_init_constrain(LtState S, LtCoef C) {
  const int u = blockIdx.x * kStepBlock + threadIdx.x;
  if (u < S.nwater) {
    const int ia = S.waters[3 * u], ib = S.waters[3 * u + 1], ic = S.waters[3 * u + 2];
    const double3 ra = ld3(S.x, S.n, ia), rb = ld3(S.x, S.n, ib), rc = ld3(S.x, S.n, ic);
    double3 na = ra, nb = rb, nc = rc;
    settle(C.G, ra, rb, rc, na, nb, nc);
    const int idx[3] = {ia, ib, ic};
    const double3 o[3] = {ra, rb, rc}, nn[3] = {na, nb, nc};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double3 v = ld3(S.v, S.n, idx[k]);
      v.x += (nn[k].x - o[k].x) / C.dt;
      v.y += (nn[k].y - o[k].y) / C.dt;
      v.z += (nn[k].z - o[k].z) / C.dt;
      st3(S.v, S.n, idx[k], v);
      st3(S.x, S.n, idx[k], nn[k]);
      store_xyzq(S, idx[k], nn[k]);
    }
  } else if (u < S.nwater + S.nshake) {
    const int4 g = S.shake[u - S.nwater];
    const double r2 = S.shake_r2[u - S.nwater];
    with_group_size(g, [&](auto mc) {
      constexpr int M = decltype(mc)::value;
      const int at[4] = {g.x, g.y, g.z, g.w};
      double im[M + 1];
      double3 o[M + 1], nn[M + 1];
#pragma unroll
      for (int a = 0; a <= M; a++) {
        im[a] = S.invmass[at[a]];
        o[a] = nn[a] = ld3(S.x, S.n, at[a]);
      }
      shake_group<M>(r2, im, o, nn);
#pragma unroll
      for (int a = 0; a <= M; a++) {
        double3 v = ld3(S.v, S.n, at[a]);
        v.x += (nn[a].x - o[a].x) / C.dt;
        v.y += (nn[a].y - o[a].y) / C.dt;
        v.z += (nn[a].z - o[a].z) / C.dt;
        st3(S.v, S.n, at[a], v);
        st3(S.x, S.n, at[a], nn[a]);
        store_xyzq(S, at[a], nn[a]);
      }
    });
  } else if (u < S.nwater + S.nshake + S.nsingle) {
    const int i = S.singles[u - S.nwater - S.nshake];
    store_xyzq(S, i, ld3(S.x, S.n, i));
  }
}

// Leapfrog back-step from the gradient at x, constrained against x; stores
// the previous half-step displacement (InitializationKernel and the two
// BackStepInitialization kernels).
__global__ __launch_bounds__(kStepBlock) void k_lt_init_backstep(LtState S, LtCoef C) {
  const int u = blockIdx.x * kStepBlock + threadIdx.x;
  const double hg = 0.5 * C.gamma;
  const double a1 = C.dt * (1.0 - hg) / sqrt(1.0 + hg);
  auto back = [&](int i, double3 x) {
    const double fact = 0.5 * C.dt2 * S.invmass[i];
    const double3 v = ld3(S.v, S.n, i), g = ldgrad(S, i);
    const double3 d = make_double3(a1 * v.x + fact * g.x, a1 * v.y + fact * g.y, a1 * v.z + fact * g.z);
    return make_double3(x.x - d.x, x.y - d.y, x.z - d.z);
  };
  if (u < S.nwater) {
    const int ia = S.waters[3 * u], ib = S.waters[3 * u + 1], ic = S.waters[3 * u + 2];
    const double3 ra = ld3(S.x, S.n, ia), rb = ld3(S.x, S.n, ib), rc = ld3(S.x, S.n, ic);
    double3 pa = back(ia, ra), pb = back(ib, rb), pc = back(ic, rc);
    settle(C.G, ra, rb, rc, pa, pb, pc);
    st3(S.dp, S.n, ia, ra - pa);
    st3(S.dp, S.n, ib, rb - pb);
    st3(S.dp, S.n, ic, rc - pc);
  } else if (u < S.nwater + S.nshake) {
    const int4 g = S.shake[u - S.nwater];
    const double r2 = S.shake_r2[u - S.nwater];
    with_group_size(g, [&](auto mc) {
      constexpr int M = decltype(mc)::value;
      const int at[4] = {g.x, g.y, g.z, g.w};
      double im[M + 1];
      double3 r[M + 1], pb[M + 1];
#pragma unroll
      for (int a = 0; a <= M; a++) {
        im[a] = S.invmass[at[a]];
        r[a] = ld3(S.x, S.n, at[a]);
        pb[a] = back(at[a], r[a]);
      }
      shake_group<M>(r2, im, r, pb);
#pragma unroll
      for (int a = 0; a <= M; a++) st3(S.dp, S.n, at[a], r[a] - pb[a]);
    });
  } else if (u < S.nwater + S.nshake + S.nsingle) {
    const int i = S.singles[u - S.nwater - S.nshake];
    const double3 r = ld3(S.x, S.n, i);
    st3(S.dp, S.n, i, r - back(i, r));
  }
}

// ---- centre-of-mass motion removal (src/CudaLangevinThermostatIntegrator.cu:832-878)
__global__ __launch_bounds__(kStepBlock) void k_com_partial(const double *__restrict__ dp,
                                                            const double *__restrict__ invmass,
                                                            int n, double4 *__restrict__ part) {
  const int i = blockIdx.x * kStepBlock + threadIdx.x;
  double m = 0, a = 0, b = 0, c = 0;
  if (i < n) {
    m = 1.0 / invmass[i];
    a = dp[i] * m;
    b = dp[n + i] * m;
    c = dp[2 * n + i] * m;
  }
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  m = wave_sum(m);
  if (threadIdx.x == 0) part[blockIdx.x] = make_double4(a, b, c, m);
}

__global__ __launch_bounds__(256) void k_com_reduce(double4 *__restrict__ part, int npart) {
  __shared__ double4 s[256];
  double4 t = make_double4(0, 0, 0, 0);
  for (int k = threadIdx.x; k < npart; k += 256) {
    t.x += part[k].x;
    t.y += part[k].y;
    t.z += part[k].z;
    t.w += part[k].w;
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[threadIdx.x].x += s[threadIdx.x + w].x;
      s[threadIdx.x].y += s[threadIdx.x + w].y;
      s[threadIdx.x].z += s[threadIdx.x + w].z;
      s[threadIdx.x].w += s[threadIdx.x + w].w;
    }
    __syncthreads();
  }
  // the result goes one past the partials, read by k_com_apply
  if (threadIdx.x == 0)
    part[npart] = make_double4(s[0].x / s[0].w, s[0].y / s[0].w, s[0].z / s[0].w, 0.0);
}

__global__ void k_com_apply(double *__restrict__ dp, int n, const double4 *__restrict__ com) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 c = com[0];
  dp[i] -= c.x;
  dp[n + i] -= c.y;
  dp[2 * n + i] -= c.z;
}

// ---- image centering (src/CharmmContext.cu:276-372, P1) ---------------------
// One wave per group (whole molecule): unweighted centroid by a fixed-order
// wave sum, then every atom of a group whose centroid lies outside
// [-L/2, L/2] moves by one box edge on that axis.
__global__ __launch_bounds__(256) void k_center(const int *__restrict__ gptr,
                                                const int *__restrict__ gatom, int ngroup,
                                                double3 box, double *__restrict__ x, int n,
                                                const float *__restrict__ charge,
                                                float4 *__restrict__ xyzq) {
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (g >= ngroup) return;  // whole wave leaves together
  const int b = gptr[g], e = gptr[g + 1];
  double cx = 0, cy = 0, cz = 0;
  for (int k = b + lane; k < e; k += 64) {
    const int i = gatom[k];
    cx += x[i];
    cy += x[n + i];
    cz += x[2 * n + i];
  }
  cx = wave_sum(cx);
  cy = wave_sum(cy);
  cz = wave_sum(cz);
  const double inv = 1.0 / (double)(e - b);
  cx *= inv;
  cy *= inv;
  cz *= inv;
  auto shift = [](double c, double L) { return c < -0.5 * L ? L : (c > 0.5 * L ? -L : 0.0); };
  const double sx = shift(cx, box.x), sy = shift(cy, box.y), sz = shift(cz, box.z);
  if (sx == 0.0 && sy == 0.0 && sz == 0.0) return;
  for (int k = b + lane; k < e; k += 64) {
    const int i = gatom[k];
    const double px = x[i] + sx, py = x[n + i] + sy, pz = x[2 * n + i] + sz;
    x[i] = px;
    x[n + i] = py;
    x[2 * n + i] = pz;
    xyzq[i] = make_float4((float)px, (float)py, (float)pz, charge[i]);
  }
}

int nblk(long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// --------------------------------------------------------------------------
struct LtSystem {
  NbSystem *nb = nullptr;
  int n = 0;
  LtCoef C{};
  int nbfreq = 20, comfreq = 1000;
  bool use_graph = true;
  double ndof = 0;
  int nwater = 0, nshake = 0, nsingle = 0, nunit_blocks = 0;
  int nstep_blocks = 0;   // k_lt_step blocks (waters and free atoms)
  int nshake_blocks = 0;  // k_lt_shake blocks; their partials follow k_lt_step's
  DevBuf<int4> shake;
  DevBuf<double> shake_r2;
  DevBuf<double> x, dp, v, invmass;
  DevBuf<float> charge;
  DevBuf<float4> xyzq;
  DevBuf<int64_t> grad;
  int gstride = 0;
  DevBuf<int> waters, singles, gptr, gatom;
  int ngroup = 0;
  DevBuf<double2> ke_part;
  DevBuf<double4> com_part;
  int com_blocks = 0;
  DevBuf<double> thermo;                 // 8 values, see k_lt_finalize
  // [0] step [1] averaging window [2] build number [3] a step's kinetic-energy
  // reduction is pending (thermo.hpp)
  DevBuf<unsigned long long> counters;
  DevBuf<float2> dpart;                  // per-block displacement maxima (squared)
  double *h_disp = nullptr;              // pinned: see k_lt_finalize
  // list rebuild policy: 0 = every nbfreq steps (the reference's cadence);
  // 1 = also whenever an atom may have moved more than half the list buffer
  // since the last build (at most nbfreq steps apart)
  int list_policy = 0;
  double half_buffer = 0.0;              // (cutnb - ctofnb) / 2
  // list pruning (policy 0 only): every prune_every steps between builds the
  // pair kernel's list is pruned at r_off + (cutnb - r_off) prune_every /
  // nbfreq; 0 = the build's list every step
  int prune_every = 0;
  unsigned long long builds_triggered = 0;
  // step completion events of policy 1 (the host looks at most kLag steps
  // behind the GPU)
  static constexpr int kLag = 4;
  hipEvent_t step_ev[kLag] = {};
  unsigned long long step = 0;           // host mirror of counters[0]
  long steps_since_rebuild = 0;
  bool initialized = false;
  // live batches that hold this integrator (it must not be destroyed before them)
  int n_batches = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // captured step bodies: [0] plain step, [1] step with image centering and
  // list rebuild, [2] step with a list prune; each is recaptured when the
  // force system's generation moves
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  unsigned long long graph_gen[3] = {~0ull, ~0ull, ~0ull};

  ~LtSystem() {
    // kernels of this integrator write the pinned h_disp and run on st and on
    // the force system's side streams: drain the device before releasing any
    // of them (destruction is rare; the reference's dealloc synchronises too)
    (void)hipDeviceSynchronize();
    if (nb) nb->n_integrators--;
    for (hipGraphExec_t g : gexec)
      if (g) (void)hipGraphExecDestroy(g);
    if (gcycle) (void)hipGraphExecDestroy(gcycle);
    for (hipEvent_t e : step_ev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : trial_ev)
      if (e) (void)hipEventDestroy(e);
    if (h_disp) (void)hipHostFree(h_disp);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (st) (void)hipStreamDestroy(st);
  }

  // fused: the last force call left the pair gradient in the sorted buffer
  // (kCalcNoFinalize) and the kernel adds it (the step kernels only)
  LtState state(bool fused = false) {
    LtState s;
    s.x = x.get();
    s.dp = dp.get();
    s.v = v.get();
    s.invmass = invmass.get();
    s.charge = charge.get();
    s.xyzq = xyzq.get();
    s.grad = grad.get();
    s.gstride = gstride;
    s.n = n;
    s.waters = waters.get();
    s.nwater = nwater;
    s.shake = shake.get();
    s.shake_r2 = shake_r2.get();
    s.nshake = nshake;
    s.singles = singles.get();
    s.nsingle = nsingle;
    s.xref = nb->xref.get();
    s.fsort = fused ? nb->fsort.get() : nullptr;
    s.glo2loc = nb->glo2loc.get();
    s.ss = nb->ssort;
    return s;
  }

  // fused: the step kernel that follows adds the sorted pair gradient
  void force(hipStream_t s, bool reset, bool prune = false, bool fused = false) {
    int flags = (reset ? APO_RESET_LIST : 0) | (fused ? kCalcNoFinalize : 0);
    if (prune_every > 0) flags |= kCalcUsePruned | (prune ? kCalcPrune : 0);
    // fixed cadence: the coordinate sort also runs the previous step's
    // kinetic-energy reduction (one launch less per step)
    LtFinArgs fin = fin_args();
    struct Reset {
      NbSystem *nb;
      ~Reset() { nb->sort_fin = nullptr; }
    } reset_fin{nb};
    nb->sort_fin = list_policy == 0 ? &fin : nullptr;
    calc_force(*nb, reinterpret_cast<const float *>(xyzq.get()), flags, grad.get(), gstride,
               nullptr, nullptr, s);
  }

  void center(hipStream_t s) {
    hipEvent_t t0 = nullptr;
    nb->tic(9, s, t0);
    const double3 box = make_double3(nb->box[0], nb->box[1], nb->box[2]);
    k_center<<<nblk(ngroup, 4), 256, 0, s>>>(gptr.get(), gatom.get(), ngroup, box, x.get(), n,
                                             charge.get(), xyzq.get());
    APO_LAUNCH_CHECK();
    nb->toc(9, s, t0);
  }

  void remove_com(hipStream_t s) {
    hipEvent_t t0 = nullptr;
    nb->tic(9, s, t0);
    k_com_partial<<<com_blocks, kStepBlock, 0, s>>>(dp.get(), invmass.get(), n, com_part.get());
    k_com_reduce<<<1, 256, 0, s>>>(com_part.get(), com_blocks);
    k_com_apply<<<nblk(n, 256), 256, 0, s>>>(dp.get(), n, com_part.get() + com_blocks);
    APO_LAUNCH_CHECK();
    nb->toc(9, s, t0);
  }

  void update(hipStream_t s) {
    hipEvent_t t0 = nullptr;
    nb->tic(8, s, t0);
    k_lt_step<<<nstep_blocks, kStepThreads, 0, s>>>(state(true), C, counters.get(), ke_part.get(),
                                                  dpart.get(), counters.get() + 3);
    if (nshake_blocks)
      k_lt_shake<<<nshake_blocks, kShakeBlock, 0, s>>>(state(true), C, counters.get(),
                                                      ke_part.get() + nstep_blocks,
                                                      dpart.get() + nstep_blocks);
    // with the fixed cadence the next force call's coordinate sort reduces
    // this step's kinetic energy (fin_args), and propagate() closes with a
    // reduction of the last step; the displacement check reads every step's
    if (list_policy != 0) finalize(s);
    APO_LAUNCH_CHECK();
    nb->toc(8, s, t0);
  }
  // the step's kinetic-energy reduction if one is pending (thermo.hpp)
  void finalize(hipStream_t s) {
    k_lt_finalize<<<1, 256, 0, s>>>(ke_part.get(), nstep_blocks + nshake_blocks, ndof, thermo.get(),
                                    counters.get(), dpart.get(), nb->meta.get(), h_disp,
                                    counters.get() + 3);
    APO_LAUNCH_CHECK();
  }
  LtFinArgs fin_args() {
    return LtFinArgs{ke_part.get(), nstep_blocks + nshake_blocks, ndof, thermo.get(), counters.get(),
                     dpart.get(), nb->meta.get(), h_disp, counters.get() + 3};
  }

  // force + integrator update + kinetic energy (after a list prune)
  void step_body(hipStream_t s, bool prune = false) {
    force(s, false, prune, true);
    update(s);
  }
  // the same after image centering and a list rebuild
  void rebuild_body(hipStream_t s) {
    center(s);
    force(s, true, false, true);
    update(s);
  }

  void initialize(hipStream_t s) {
    k_lt_init_constrain<<<nunit_blocks, kStepBlock, 0, s>>>(state(), C);
    APO_LAUNCH_CHECK();
    force(s, true);
    if (prune_every > 0) nb->pruned_build = nb->builds_queued;
    k_lt_init_backstep<<<nunit_blocks, kStepBlock, 0, s>>>(state(), C);
    APO_LAUNCH_CHECK();
    initialized = true;
  }

  bool graph_usable() const {
    return use_graph && !nb->prof_on && nb->list_valid && nb->cand_cap > 0 && nb->ent_cap > 0 &&
           nb->item_cap > 0;
  }

  void replay(int which, hipStream_t s) {
    hipGraphExec_t &ex = gexec[which];
    if (!ex || graph_gen[which] != nb->generation) {
      if (ex) {
        APO_HIP(hipGraphExecDestroy(ex));
        ex = nullptr;
      }
      hipGraph_t g = nullptr;
      // a build inside the graph counts when the graph is launched
      const unsigned long long bq = nb->builds_queued;
      APO_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        if (which == 1) rebuild_body(s);
        else step_body(s, which == 2);
      } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        nb->builds_queued = bq;
        throw;
      }
      nb->builds_queued = bq;
      APO_HIP(hipStreamEndCapture(s, &g));
      hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      APO_HIP(e);
      graph_gen[which] = nb->generation;
    }
    APO_HIP(hipGraphLaunch(ex, s));
    if (which == 1) nb->builds_queued++;
  }

  // A whole list cycle (the rebuild step and the nbfreq - 1 steps after it,
  // with their prunes) as one graph: one launch per nbfreq steps.  It needs
  // the fixed cadence and no centre-of-mass removal inside the cycle.
  hipGraphExec_t gcycle = nullptr;
  unsigned long long gcycle_gen = ~0ull;
  bool cycle_ok() const {
    if (!graph_usable() || list_policy != 0 || nbfreq < 2) return false;
    if (comfreq <= 0) return true;
    const unsigned long long cf = (unsigned long long)comfreq;
    return (step / cf + 1) * cf >= step + (unsigned long long)nbfreq;
  }
  void run_cycle(hipStream_t s) {
    if (comfreq > 0 && step % (unsigned long long)comfreq == 0) remove_com(s);
    list_export_check(*nb, s, true);
    if (!gcycle || gcycle_gen != nb->generation) {
      if (gcycle) {
        APO_HIP(hipGraphExecDestroy(gcycle));
        gcycle = nullptr;
      }
      hipGraph_t g = nullptr;
      const unsigned long long bq = nb->builds_queued;
      APO_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        rebuild_body(s);
        for (int i = 1; i < nbfreq; i++) step_body(s, prune_every > 0 && i % prune_every == 0);
      } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        nb->builds_queued = bq;
        throw;
      }
      nb->builds_queued = bq;
      APO_HIP(hipStreamEndCapture(s, &g));
      hipError_t e = hipGraphInstantiate(&gcycle, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      APO_HIP(e);
      gcycle_gen = nb->generation;
    }
    APO_HIP(hipGraphLaunch(gcycle, s));
    nb->builds_queued++;
    if (prune_every > 0) nb->pruned_build = nb->builds_queued;
    steps_since_rebuild += nbfreq;
    step += (unsigned long long)nbfreq;
  }

  // Launch-settings trial (NbSystem::pme_prio, NbSystem::pair_first): the
  // first whole list cycles run under each of the four settings (bit 0: PME
  // issue priority, bit 1: pair kernel first), one untimed cycle and
  // kTrialTimed timed cycles each, and the setting with the fastest timed
  // cycle stays.  The settings order instruction issue and dispatch only, so
  // these are ordinary cycles of the trajectory.  The timed cycles wait for
  // the GPU on the host; launch_fixed (apo_lt_params::launch_setting) skips
  // the trial for callers that share the GPU or must not block.
  static constexpr int kTrialSettings = 4;
  static constexpr int kTrialTimed = 3;
  static constexpr int kTrialPer = 1 + kTrialTimed;
  static constexpr int kTrialCycles = kTrialPer * kTrialSettings;
  int trial = 0;  // trial cycles run so far
  int launch_fixed = -1;  // >= 0: the setting, no trial
  float trial_ms[kTrialSettings] = {};
  hipEvent_t trial_ev[2] = {nullptr, nullptr};
  bool trial_pending() const { return launch_fixed < 0 && trial < kTrialCycles; }
  void set_launch(int setting) {
    const int prio = setting & 1, first = (setting >> 1) & 1;
    if (nb->pme_prio == prio && nb->pair_first == first) return;
    nb->pme_prio = prio;
    nb->pair_first = first;
    nb->generation++;  // captured graphs hold both as launch constants
  }
  void trial_cycle(hipStream_t s) {
    const int setting = trial / kTrialPer;
    const bool timed = trial % kTrialPer != 0;
    set_launch(setting);
    if (timed) APO_HIP(hipEventRecord(trial_ev[0], s));
    run_cycle(s);
    if (timed) {
      APO_HIP(hipEventRecord(trial_ev[1], s));
      APO_HIP(hipEventSynchronize(trial_ev[1]));
      float ms = 0.f;
      APO_HIP(hipEventElapsedTime(&ms, trial_ev[0], trial_ev[1]));
      if (trial % kTrialPer == 1 || ms < trial_ms[setting]) trial_ms[setting] = ms;
    }
    if (++trial == kTrialCycles) {
      int best = 0;
      for (int k = 1; k < kTrialSettings; k++)
        if (trial_ms[k] < trial_ms[best]) best = k;
      set_launch(best);
    }
  }

  // CudaIntegrator::propagate (src/CudaIntegrator.cu:140-230) with the
  // thermostat's propagateOneStep (src/CudaLangevinThermostatIntegrator.cu:703-800):
  // whole list cycles as one graph where the schedule allows, else step by step.
  // fresh = false continues the previous call's list cadence (no list reset
  // or centering at the start), for callers that split one propagate into
  // pieces (subscriber reports between them)
  void propagate(int nsteps, hipStream_t s, bool fresh = true) {
    if (fresh) begin(s);
    else if (!initialized) initialize(s);
    for (int k = 0; k < nsteps;) {
      if (nsteps - k >= nbfreq && steps_since_rebuild % nbfreq == 0 && cycle_ok()) {
        if (trial_pending()) trial_cycle(s);
        else run_cycle(s);
        k += nbfreq;
      } else {
        one_step(s);
        k++;
      }
    }
    // the last step's kinetic energy (its reduction would otherwise wait for
    // the next force call)
    if (list_policy == 0 && nsteps > 0) finalize(s);
  }

  void begin(hipStream_t s) {
    if (!initialized) initialize(s);
    // propagate() opens with resetNeighborList(), whose image centering the
    // first step then repeats
    center(s);
    steps_since_rebuild = 0;
  }

  // Policy 1: whether the next step must rebuild the list.  The latest
  // displacements the GPU exported (at most kLag steps old: the host waits
  // for the step kLag back) plus the largest one-step displacement times the
  // steps not yet seen must stay below half the list buffer, with a 10 %
  // margin for the float coordinates of the check.
  bool displacement_trigger(hipStream_t s) {
    (void)s;
    const unsigned long long k = step;  // the step about to be queued
    if (k >= (unsigned long long)kLag) APO_HIP(hipEventSynchronize(step_ev[(k - kLag) % kLag]));
    const volatile double *h = h_disp;
    const double done = h[2];
    const double dmax = h[0], smax = h[1], seq = h[3];
    const double smax_est = std::max(smax, last_smax);
    last_smax = smax_est;
    const double unseen = (double)k - done + 1.0;
    double pred;
    if (seq == (double)nb->builds_queued)
      pred = dmax + std::max(unseen, 1.0) * smax_est;
    else  // the exported values belong to an earlier build
      pred = (double)(steps_since_rebuild + 1) * smax_est;
    last_pred = pred;
    return pred > 0.9 * half_buffer;
  }
  double last_smax = 0.0, last_pred = 0.0;

  void one_step(hipStream_t s) {
    {
      bool rebuild = steps_since_rebuild % nbfreq == 0;
      if (list_policy == 1 && !rebuild && displacement_trigger(s)) {
        rebuild = true;
        builds_triggered++;
        steps_since_rebuild = 0;
      }
      // prune on the schedule, and whenever the pruned copy belongs to an
      // earlier build (a rebuild outside this integrator, a fused batch)
      const bool prune = !rebuild && prune_every > 0 &&
                         (steps_since_rebuild % prune_every == 0 ||
                          nb->pruned_build != nb->builds_queued);
      const bool com = comfreq > 0 && step % (unsigned long long)comfreq == 0;
      // centre-of-mass removal only touches the half-step displacements, so
      // it may precede the force call
      if (com) remove_com(s);
      if (!graph_usable()) {
        if (rebuild) rebuild_body(s);
        else step_body(s, prune);
      } else {
        // errors and capacity growth of earlier (device-side) rebuilds are
        // handled before the next one is queued
        if (rebuild) list_export_check(*nb, s, true);
        replay(rebuild ? 1 : (prune ? 2 : 0), s);
      }
      if (prune_every > 0 && (rebuild || prune)) nb->pruned_build = nb->builds_queued;
      if (list_policy == 1) APO_HIP(hipEventRecord(step_ev[step % kLag], s));
      steps_since_rebuild++;
      step++;
    }
  }
};

// Molecules (connected components over bonds and waters) as centering groups.
static void molecule_groups(int n, const std::vector<int> &pairs, std::vector<int> &gptr,
                            std::vector<int> &gatom) {
  std::vector<int> parent(n);
  for (int i = 0; i < n; i++) parent[i] = i;
  auto find = [&](int a) {
    while (parent[a] != a) {
      parent[a] = parent[parent[a]];
      a = parent[a];
    }
    return a;
  };
  for (size_t k = 0; k + 1 < pairs.size(); k += 2) {
    int a = find(pairs[k]), b = find(pairs[k + 1]);
    if (a != b) parent[a > b ? a : b] = a > b ? b : a;
  }
  // groups ordered by their lowest atom, atoms ascending inside a group
  std::vector<int> gid(n, -1), count;
  int ng = 0;
  for (int i = 0; i < n; i++) {
    int r = find(i);
    if (gid[r] < 0) {
      gid[r] = ng++;
      count.push_back(0);
    }
    count[gid[r]]++;
  }
  gptr.assign(ng + 1, 0);
  for (int g = 0; g < ng; g++) gptr[g + 1] = gptr[g] + count[g];
  std::vector<int> fill(gptr.begin(), gptr.end() - 1);
  gatom.assign(n, 0);
  for (int i = 0; i < n; i++) gatom[fill[gid[find(i)]]++] = i;
}

LtSystem *create_lt(NbSystem *nb, const apo_lt_params &p, const double *xyz, const double *vel) {
  if (!nb) throw InvalidArgument("force system is null");
  if (!xyz) throw InvalidArgument("coordinates are required");
  if (!p.mass || !p.charge) throw InvalidArgument("mass and charge arrays are required");
  if (!(p.timestep > 0)) throw InvalidArgument("timestep must be positive (ps)");
  if (p.friction < 0 || p.temperature < 0) throw InvalidArgument("negative friction or temperature");
  if (p.nwater < 0 || (p.nwater > 0 && !p.waters)) throw InvalidArgument("bad water list");
  const int n = nb->natom;
  auto *L = new LtSystem();
  try {
    L->nb = nb;
    nb->n_integrators++;
    L->n = n;
    L->nbfreq = p.nbfreq > 0 ? p.nbfreq : 20;
    L->comfreq = p.comfreq == 0 ? 1000 : p.comfreq;
    L->use_graph = p.use_graph != 0;
    double geom[4] = {15.99940, 1.00800, 0.9572, 1.5139};
    if (p.water_geom[0] > 0 || p.water_geom[1] > 0 || p.water_geom[2] > 0 || p.water_geom[3] > 0) {
      for (int k = 0; k < 4; k++) {
        if (!(p.water_geom[k] > 0)) throw InvalidArgument("water_geom entries must be positive");
        geom[k] = p.water_geom[k];
      }
      if (!(geom[3] < 2.0 * geom[2])) throw InvalidArgument("water_geom: rHH must be < 2 rOH");
    }
    const double dt = p.timestep / kTimfac;
    LtCoef &C = L->C;
    C.gamma = p.timestep * p.friction;  // m_TimeStep * m_Timfac * friction
    C.dt = dt;
    C.dt2 = dt * dt;
    C.kbt = kBoltz * p.temperature;
    C.den = 1.0 / (1.0 + 0.5 * C.gamma);
    C.alpha = (1.0 - 0.5 * C.gamma) * C.den;
    C.fv = 0.5 * std::sqrt(1.0 + 0.5 * C.gamma) / dt;
    C.idt = 1.0 / dt;
    C.seed = p.seed;
    const double M = geom[0] + 2.0 * geom[1];
    const double rc = 0.5 * geom[3];
    const double h = std::sqrt(geom[2] * geom[2] - rc * rc);
    C.G.ra = 2.0 * geom[1] / M * h;
    C.G.rb = h - C.G.ra;
    C.G.rc = rc;
    C.G.wO = geom[0] / M;
    C.G.wH = geom[1] / M;

    std::vector<double> invm(n);
    for (int i = 0; i < n; i++) {
      if (!(p.mass[i] > 0)) throw InvalidArgument("masses must be positive");
      invm[i] = 1.0 / p.mass[i];
    }
    std::vector<char> in_water(n, 0);
    std::vector<int> wl(p.waters, p.waters + 3 * (size_t)p.nwater);
    for (int k = 0; k < 3 * p.nwater; k++) {
      const int a = wl[k];
      if (a < 0 || a >= n) throw InvalidArgument("water atom index out of range");
      if (in_water[a]) throw InvalidArgument("atom listed in two waters");
      in_water[a] = 1;
    }
    // SHAKE groups: a heavy atom and 1-3 hydrogens, none of them in a water
    // or in another group
    if (p.nshake < 0 || (p.nshake > 0 && (!p.shake_groups || !p.shake_r2)))
      throw InvalidArgument("bad SHAKE group list");
    std::vector<int4> sg(p.nshake);
    std::vector<double> sr2(p.shake_r2, p.shake_r2 + (p.nshake > 0 ? p.nshake : 0));
    int nshake_bonds = 0;
    for (int g = 0; g < p.nshake; g++) {
      const int32_t *a = p.shake_groups + 4 * (size_t)g;
      if (a[1] < 0 || (a[2] < 0 && a[3] >= 0))
        throw InvalidArgument("a SHAKE group needs its hydrogens in the first slots");
      for (int k = 0; k < 4; k++) {
        if (a[k] < 0) continue;
        if (a[k] >= n) throw InvalidArgument("SHAKE atom index out of range");
        if (in_water[a[k]]) throw InvalidArgument("atom listed in two constraint groups");
        in_water[a[k]] = 1;
        if (k > 0) nshake_bonds++;
      }
      if (!(sr2[g] > 0)) throw InvalidArgument("SHAKE bond lengths must be positive");
      sg[g] = make_int4(a[0], a[1], a[2], a[3]);
    }
    // groups ordered by hydrogen count (stable), so that the lanes of a wave
    // of k_lt_shake mostly take the same solver instantiation
    {
      auto nh = [](const int4 &q) { return q.w >= 0 ? 3 : (q.z >= 0 ? 2 : 1); };
      std::vector<int> order(p.nshake);
      for (int g = 0; g < p.nshake; g++) order[g] = g;
      std::stable_sort(order.begin(), order.end(),
                       [&](int x, int y) { return nh(sg[x]) < nh(sg[y]); });
      std::vector<int4> sg2(p.nshake);
      std::vector<double> sr22(p.nshake);
      for (int g = 0; g < p.nshake; g++) {
        sg2[g] = sg[order[g]];
        sr22[g] = sr2[order[g]];
      }
      sg.swap(sg2);
      sr2.swap(sr22);
    }
    std::vector<int> singles;
    for (int i = 0; i < n; i++)
      if (!in_water[i]) singles.push_back(i);
    L->nwater = p.nwater;
    L->nshake = p.nshake;
    L->nsingle = (int)singles.size();
    L->nunit_blocks = nblk(L->nwater + L->nshake + L->nsingle, kStepBlock);
    L->nstep_blocks = std::max(1, nblk(L->nwater + L->nsingle, kStepBlock));
    L->nshake_blocks = nblk(L->nshake, kShakeBlock);
    L->com_blocks = nblk(n, kStepBlock);
    // P1: 3N - 3, less 3 per SETTLE water and 1 per SHAKE bond
    // (CharmmContext::useHolonomicConstraints, src/CharmmContext.cu:919-945)
    L->ndof = 3.0 * n - 3.0 - 3.0 * p.nwater - nshake_bonds;
    if (L->ndof <= 0) throw InvalidArgument("no degrees of freedom left");

    // centering groups: molecules over bonds and waters
    std::vector<int> pairs = nb->h_bonds;
    for (int k = 0; k < p.nwater; k++) {
      pairs.push_back(wl[3 * k]);
      pairs.push_back(wl[3 * k + 1]);
      pairs.push_back(wl[3 * k]);
      pairs.push_back(wl[3 * k + 2]);
    }
    std::vector<int> gptr, gatom;
    molecule_groups(n, pairs, gptr, gatom);
    L->ngroup = (int)gptr.size() - 1;

    APO_HIP(hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking));
    APO_HIP(hipEventCreateWithFlags(&L->ev_in, hipEventDisableTiming));
    APO_HIP(hipEventCreateWithFlags(&L->ev_out, hipEventDisableTiming));
    hipStream_t s = L->st;
    // host AoS -> device SoA
    std::vector<double> soa(3 * (size_t)n);
    auto to_soa = [&](const double *a) {
      for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) soa[(size_t)d * n + i] = a ? a[3 * (size_t)i + d] : 0.0;
    };
    to_soa(xyz);
    L->x.upload(soa, s);
    to_soa(vel);
    L->v.upload(soa, s);
    L->dp.reserve(3 * (size_t)n);
    launch_fill_f64(L->dp.get(), 3 * (size_t)n, s);
    L->invmass.upload(invm, s);
    L->charge.upload(p.charge, n, s);
    std::vector<float4> q4(n);
    for (int i = 0; i < n; i++)
      q4[i] = make_float4((float)xyz[3 * (size_t)i], (float)xyz[3 * (size_t)i + 1],
                          (float)xyz[3 * (size_t)i + 2], p.charge[i]);
    L->xyzq.upload(q4, s);
    L->gstride = ((n + 31) * 8 + 255) / 256 * 256 / 8;
    L->grad.reserve(3 * (size_t)L->gstride);
    launch_fill_i64(L->grad.get(), 3 * (size_t)L->gstride, s);
    L->waters.upload(wl, s);
    L->singles.upload(singles, s);
    L->shake.upload(sg, s);
    L->shake_r2.upload(sr2, s);
    L->gptr.upload(gptr, s);
    L->gatom.upload(gatom, s);
    L->ke_part.reserve(L->nstep_blocks + L->nshake_blocks);
    L->com_part.reserve(L->com_blocks + 1);
    L->thermo.reserve(8);
    launch_fill_f64(L->thermo.get(), 8, s);
    std::vector<unsigned long long> c0 = {0ull, 0ull, ~0ull, 0ull};
    L->counters.upload(c0, s);
    L->dpart.reserve(L->nstep_blocks + L->nshake_blocks);
    APO_HIP(hipHostMalloc(reinterpret_cast<void **>(&L->h_disp), 4 * sizeof(double)));
    for (int k = 0; k < 4; k++) L->h_disp[k] = 0.0;
    L->list_policy = p.list_policy;
    if (L->list_policy != 0 && L->list_policy != 1)
      throw InvalidArgument("list_policy must be 0 (fixed cadence) or 1 (displacement check)");
    L->half_buffer = 0.5 * (nb->cutnb - nb->ctofnb);
    if (p.launch_setting < 0 || p.launch_setting > LtSystem::kTrialSettings)
      throw InvalidArgument("launch_setting must be 0 (timing trial) or 1..4");
    if (p.launch_setting > 0) {
      L->launch_fixed = p.launch_setting - 1;
      L->set_launch(L->launch_fixed);
    }
    L->prune_every = p.prune_every;
    if (L->prune_every < 0) throw InvalidArgument("prune_every must be >= 0");
    if (L->prune_every > 0 && L->list_policy != 0)
      throw InvalidArgument("list pruning needs the fixed rebuild cadence (list_policy 0)");
    if (L->prune_every > 0) {
      const double rp = std::min(nb->cutnb, nb->ctofnb + (nb->cutnb - nb->ctofnb) *
                                                             (double)L->prune_every / L->nbfreq);
      nb->prune_r2 = (float)(rp * rp);
      nb->generation++;  // a launch constant of captured steps
    }
    for (hipEvent_t &e : L->step_ev)
      APO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (hipEvent_t &e : L->trial_ev) APO_HIP(hipEventCreate(&e));
    APO_HIP(hipStreamSynchronize(s));
  } catch (...) {
    delete L;
    throw;
  }
  return L;
}

void destroy_lt(LtSystem *L) {
  if (L && L->n_batches > 0)
    throw InvalidArgument("the integrator is still a member of a batch: destroy the batch first");
  delete L;
}

// Work runs on the integrator's own stream, ordered after and before the
// caller's stream by events, so the non-rebuild steps can be captured.
template <typename F> static void on_stream(LtSystem &L, hipStream_t caller, F &&f) {
  APO_HIP(hipEventRecord(L.ev_in, caller));
  APO_HIP(hipStreamWaitEvent(L.st, L.ev_in, 0));
  f(L.st);
  APO_HIP(hipEventRecord(L.ev_out, L.st));
  APO_HIP(hipStreamWaitEvent(caller, L.ev_out, 0));
}

void lt_initialize(LtSystem &L, hipStream_t caller) {
  on_stream(L, caller, [&](hipStream_t s) { L.initialize(s); });
}

void lt_propagate(LtSystem &L, int nsteps, hipStream_t caller, bool fresh) {
  if (nsteps < 0) throw InvalidArgument("nsteps must be >= 0");
  on_stream(L, caller, [&](hipStream_t s) { L.propagate(nsteps, s, fresh); });
}
void lt_propagate(LtSystem &L, int nsteps, hipStream_t caller) {
  lt_propagate(L, nsteps, caller, true);
}

// Many independent systems advanced together (no counterpart in the
// reference, which runs one CharmmContext per process).  Each batch step is
// one graph launch: the systems' steps are `lanes` independent chains of
// branches (systems i, i + lanes, ... back to back in chain i mod lanes), so
// the chains overlap on the GPU's hardware queues and no host synchronisation
// or per-step event happens between list rebuilds.  Per system the kernels
// and their order are those of lt_propagate, so every trajectory is bitwise
// the one the system gets alone.
struct LtBatch {
  std::vector<LtSystem *> sys;
  std::vector<bool> was_single;
  int lanes = 4;
  hipStream_t m = nullptr;
  std::vector<hipStream_t> lane;
  hipEvent_t fork_ev = nullptr, m_ev = nullptr;
  std::vector<hipEvent_t> lane_ev;
  // force branches inside a step: PME and bonded on side streams, the pair
  // kernel on m (as calc_force forks for one system)
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t sort_ev = nullptr, side_ev[2] = {nullptr, nullptr};
  hipGraphExec_t ex[2] = {nullptr, nullptr};
  std::vector<unsigned long long> gens[2];
  // fused step: one launch per kernel type for all systems
  BatchLaunch<SortArgs> sort;
  BatchLaunch<PairArgs> pair;
  BatchLaunch<LtStepArgs> lstep, lshake;
  BatchLaunch<LtFinArgs> lfin;
  PmeBatch *pme = nullptr;
  BondedBatch *bond = nullptr;
  int tab_mode = 1;  // vdW table layout of the fused pair launch (pair.hip TAB_*)
  bool tab_iso = false;  // every system's box is cubic (fixed-point-unit distances)
  std::vector<unsigned long long> tab_gens;
  // launch shapes (the autotuner's choices): pair-kernel blocks per CU
  // shared out by atom count, FFT threads, row pairs and columns per block
  double pair_bpc = kPairBlocksPerCu;
  int fft_threads = kFftThreads, fft_rows = kFftRowsPerBlock, fft_cols = kFftColsPerBlock;

  ~LtBatch() {
    (void)hipDeviceSynchronize();
    for (LtSystem *L : sys) L->n_batches--;
    for (hipGraphExec_t g : ex)
      if (g) (void)hipGraphExecDestroy(g);
    for (hipEvent_t e : lane_ev) (void)hipEventDestroy(e);
    for (hipStream_t st : lane) (void)hipStreamDestroy(st);
    for (int i = 0; i < 2; i++) {
      if (side_ev[i]) (void)hipEventDestroy(side_ev[i]);
      if (side[i]) (void)hipStreamDestroy(side[i]);
    }
    if (sort_ev) (void)hipEventDestroy(sort_ev);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (m_ev) (void)hipEventDestroy(m_ev);
    if (m) (void)hipStreamDestroy(m);
    if (pme) pme_batch_destroy(pme);
    if (bond) bonded_batch_destroy(bond);
    for (size_t i = 0; i < was_single.size() && i < sys.size(); i++)
      if (!was_single[i]) {
        sys[i]->nb->single_stream = false;
        sys[i]->nb->generation++;
      }
  }

  // per-system argument tables of the fused launches (pointers and launch
  // constants move only with a system's generation)
  void build_tables() {
    sort.clear();
    pair.clear();
    lstep.clear();
    lshake.clear();
    lfin.clear();
    pme_batch_clear(*pme);
    bonded_batch_clear(*bond);
    tab_mode = 1;
    int i_sys = 0;
    double ntot = 0;
    for (LtSystem *L : sys) ntot += L->n;
    const int ncu = sys[0]->nb->ncu;
    for (LtSystem *L : sys) {
      NbSystem &S = *L->nb;
      const float4 *x = L->xyzq.get();
      int64_t *g = L->grad.get();
      batch_add_sort(sort, S, x, g, 3 * (size_t)L->gstride);
      const int pb = std::max(8, (int)std::lround(pair_bpc * ncu * (L->n / ntot)));
      batch_add_pair(pair, S, pb);
      // one layout for all: a layout every system allows (pair.hip TAB_*:
      // 4 switched-coefficient LDS table, 1 square LDS, 2 triangular LDS,
      // 0 global), the global table if any system needs it
      const int m = pair_tab_lds(S);
      if (i_sys == 0) tab_mode = m;
      else if (m != tab_mode) {
        // mixed layouts: the triangular table fits every LDS-capable system;
        // any system beyond it sends the launch to the global table
        const bool lds_ok = (m == 1 || m == 2 || m == 4) && (tab_mode == 1 || tab_mode == 2 ||
                                                             tab_mode == 4);
        tab_mode = lds_ok ? 2 : 0;
      }
      tab_iso = (i_sys == 0 || tab_iso) && pair_tab_iso(S);
      i_sys++;
      pme_batch_add(*pme, S, g, L->gstride, fft_rows, fft_cols, fft_threads);
      bonded_batch_add(*bond, S, x, g, L->gstride);
      lstep.add(LtStepArgs{L->state(true), L->C, L->counters.get(), L->ke_part.get(),
                           L->dpart.get()},
                L->nstep_blocks);
      if (L->nshake_blocks)
        lshake.add(LtStepArgs{L->state(true), L->C, L->counters.get(),
                              L->ke_part.get() + L->nstep_blocks, L->dpart.get() + L->nstep_blocks},
                   L->nshake_blocks);
      lfin.add(LtFinArgs{L->ke_part.get(), L->nstep_blocks + L->nshake_blocks, L->ndof, L->thermo.get(),
                         L->counters.get(), L->dpart.get(), S.meta.get(), L->h_disp, nullptr},
               1);
    }
    sort.upload(m);
    pair.upload(m);
    lstep.upload(m);
    lshake.upload(m);
    lfin.upload(m);
    pme_batch_upload(*pme, m);
    bonded_batch_upload(*bond, m);
    APO_HIP(hipStreamSynchronize(m));
    tab_gens.resize(sys.size());
    for (size_t i = 0; i < sys.size(); i++) tab_gens[i] = sys[i]->nb->generation;
  }

  // force path of every system + integrator update: 12 launches in all (13
  // with SHAKE groups).
  // After the sort, the PME chain and the bonded terms fork onto side
  // streams and the pair kernel runs on s; all three write int64 sums, so
  // their overlap changes no bit.
  void fused_step(hipStream_t s) {
    force_step(s);
    kb_lt_step<<<lstep.blocks(), kStepThreads, 0, s>>>(lstep.d_args.get(), lstep.d_start.get(),
                                                     lstep.nsys());
    if (lshake.nsys())
      kb_lt_shake<<<lshake.blocks(), kShakeBlock, 0, s>>>(lshake.d_args.get(),
                                                         lshake.d_start.get(), lshake.nsys());
    kb_lt_finalize<<<lfin.nsys(), 256, 0, s>>>(lfin.d_args.get());
    APO_LAUNCH_CHECK();
  }
  // the force part: every system's gradient from its current coordinates,
  // with the pair part left in the sorted buffer for kb_lt_step to add
  // (rewrites per-step scratch and the gradient buffers only)
  void force_step(hipStream_t s) {
    batch_sort(sort, s);  // also zeroes every system's gradient
    APO_HIP(hipEventRecord(sort_ev, s));
    APO_HIP(hipStreamWaitEvent(side[0], sort_ev, 0));
    pme_batch_launch(*pme, side[0], true);
    APO_HIP(hipEventRecord(side_ev[0], side[0]));
    APO_HIP(hipStreamWaitEvent(side[1], sort_ev, 0));
    bonded_batch_launch(*bond, side[1]);
    APO_HIP(hipEventRecord(side_ev[1], side[1]));
    batch_pair(pair, tab_mode, tab_iso, s);
    APO_HIP(hipStreamWaitEvent(s, side_ev[0], 0));
    APO_HIP(hipStreamWaitEvent(s, side_ev[1], 0));
    // the sorted pair gradient is added by the step kernel (kb_lt_step)
  }

  // Autotuner of the launch shapes.  It times the fused spread + FFT
  // launches for seven FFT block shapes, then the whole force step (sort,
  // the PME, bonded and pair branches) for eight pair-kernel block budgets,
  // on the systems' current coordinates and lists, and keeps the fastest.
  // Those launches rewrite per-step scratch and the gradient buffers, which
  // each step recomputes before using them, and each row, column and item is
  // computed the same way in any blocking, so trajectories do not change.
  // report[9]:
  // chosen blocks per CU, FFT threads, rows and columns, the force-step and
  // PME-chain times (ms) of the choice and of the default shapes, candidates
  // timed.
  void autotune(hipStream_t caller, double *report) {
    APO_HIP(hipEventRecord(m_ev, caller));
    APO_HIP(hipStreamWaitEvent(m, m_ev, 0));
    for (LtSystem *L : sys)
      if (!L->initialized) {
        APO_HIP(hipStreamWaitEvent(L->st, m_ev, 0));
        L->initialize(L->st);
        APO_HIP(hipEventRecord(L->ev_out, L->st));
        APO_HIP(hipStreamWaitEvent(m, L->ev_out, 0));
      }
    APO_HIP(hipStreamSynchronize(m));
    for (LtSystem *L : sys) APO_HIP(hipStreamSynchronize(L->st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    APO_HIP(hipEventCreate(&e0));
    APO_HIP(hipEventCreate(&e1));
    int ntimed = 0;
    auto timed = [&](auto &&launch) {
      launch();  // warm
      APO_HIP(hipEventRecord(e0, m));
      for (int r = 0; r < 3; r++) launch();
      APO_HIP(hipEventRecord(e1, m));
      APO_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      APO_HIP(hipEventElapsedTime(&ms, e0, e1));
      ntimed++;
      return ms / 3.0;
    };
    try {
      // FFT shape first, on the PME chain alone: (threads, row pairs, columns)
      // per block of the y, z and x-inverse passes
      const int fft_cand[][3] = {{256, 16, 16}, {256, 8, 8}, {128, 8, 8}, {64, 4, 4},
                                 {64, 8, 8},    {64, 2, 2},  {64, 4, 8}};
      int best_t = kFftThreads, best_r = kFftRowsPerBlock, best_c = kFftColsPerBlock;
      double best_pme = 1e30, def_pme = 0;
      for (const auto &c : fft_cand) {
        fft_threads = c[0];
        fft_rows = c[1];
        fft_cols = c[2];
        build_tables();
        const double t = timed([&] { pme_batch_launch(*pme, m, false); });
        if (c[0] == kFftThreads && c[1] == kFftRowsPerBlock && c[2] == kFftColsPerBlock)
          def_pme = t;
        if (t < best_pme) best_pme = t, best_t = c[0], best_r = c[1], best_c = c[2];
      }
      fft_threads = best_t;
      fft_rows = best_r;
      fft_cols = best_c;
      // then the pair kernel's block budget, on the whole force step: the pair
      // kernel shares the GPU with the PME and bonded branches, and a budget
      // that is fastest alone can starve the PME chain
      const double bpc_cand[] = {8, 16, 24, 32, 48, 64, 96, 128};
      double best_bpc = kPairBlocksPerCu, best_force = 1e30, def_force = 0;
      for (double bpc : bpc_cand) {
        pair_bpc = bpc;
        build_tables();
        const double t = timed([&] { force_step(m); });
        if (bpc == kPairBlocksPerCu) def_force = t;
        if (t < best_force) best_force = t, best_bpc = bpc;
      }
      pair_bpc = best_bpc;
      build_tables();
      if (report) {
        report[0] = pair_bpc;
        report[1] = fft_threads;
        report[2] = fft_rows;
        report[3] = fft_cols;
        report[4] = best_force;
        report[5] = best_pme;
        report[6] = def_force;
        report[7] = def_pme;
        report[8] = ntimed;
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // graphs captured with the old shapes are recaptured at the next step
    for (hipGraphExec_t &g : ex)
      if (g) {
        APO_HIP(hipGraphExecDestroy(g));
        g = nullptr;
      }
  }

  void capture(int rb) {
    bool stale_tab = tab_gens.size() != sys.size();
    for (size_t i = 0; i < sys.size() && !stale_tab; i++)
      stale_tab = tab_gens[i] != sys[i]->nb->generation;
    if (stale_tab) build_tables();
    if (ex[rb]) {
      APO_HIP(hipGraphExecDestroy(ex[rb]));
      ex[rb] = nullptr;
    }
    hipGraph_t g = nullptr;
    // builds inside the graph count when it is launched
    std::vector<unsigned long long> bq(sys.size());
    for (size_t i = 0; i < sys.size(); i++) bq[i] = sys[i]->nb->builds_queued;
    auto restore_bq = [&] {
      for (size_t i = 0; i < sys.size(); i++) sys[i]->nb->builds_queued = bq[i];
    };
    APO_HIP(hipStreamBeginCapture(m, hipStreamCaptureModeThreadLocal));
    try {
      if (rb) {
        // image centering and list rebuild of every system, in parallel lanes
        APO_HIP(hipEventRecord(fork_ev, m));
        for (int l = 0; l < lanes; l++) {
          APO_HIP(hipStreamWaitEvent(lane[l], fork_ev, 0));
          for (size_t i = l; i < sys.size(); i += lanes) {
            LtSystem &L = *sys[i];
            L.center(lane[l]);
            nblist_build(*L.nb, L.xyzq.get(), lane[l], false);
          }
          APO_HIP(hipEventRecord(lane_ev[l], lane[l]));
          APO_HIP(hipStreamWaitEvent(m, lane_ev[l], 0));
        }
      }
      fused_step(m);
    } catch (...) {
      (void)hipStreamEndCapture(m, &g);
      if (g) (void)hipGraphDestroy(g);
      restore_bq();
      throw;
    }
    restore_bq();
    APO_HIP(hipStreamEndCapture(m, &g));
    hipError_t e = hipGraphInstantiate(&ex[rb], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    APO_HIP(e);
    gens[rb].resize(sys.size());
    for (size_t i = 0; i < sys.size(); i++) gens[rb][i] = sys[i]->nb->generation;
  }
  void propagate(int nsteps, hipStream_t caller) {
    const int n = (int)sys.size();
    APO_HIP(hipEventRecord(m_ev, caller));
    APO_HIP(hipStreamWaitEvent(m, m_ev, 0));
    for (LtSystem *L : sys) {
      APO_HIP(hipStreamWaitEvent(L->st, m_ev, 0));
      L->begin(L->st);
      APO_HIP(hipEventRecord(L->ev_out, L->st));
      APO_HIP(hipStreamWaitEvent(m, L->ev_out, 0));
    }
    bool whole = true;
    for (LtSystem *L : sys)
      whole = whole && L->graph_usable() && L->nbfreq == sys[0]->nbfreq &&
              L->steps_since_rebuild == sys[0]->steps_since_rebuild;
    for (int k = 0; k < nsteps; k++) {
      if (!whole) {  // systems out of phase: per-system steps on the batch stream
        for (LtSystem *L : sys) L->one_step(m);
        continue;
      }
      const int rb = sys[0]->steps_since_rebuild % sys[0]->nbfreq == 0 ? 1 : 0;
      for (LtSystem *L : sys)
        if (L->comfreq > 0 && L->step % (unsigned long long)L->comfreq == 0) L->remove_com(m);
      if (rb)
        for (LtSystem *L : sys) list_export_check(*L->nb, m, true);
      bool stale = !ex[rb];
      for (int i = 0; i < n && !stale; i++) stale = gens[rb][i] != sys[i]->nb->generation;
      if (stale) capture(rb);
      APO_HIP(hipGraphLaunch(ex[rb], m));
      if (rb)
        for (LtSystem *L : sys) L->nb->builds_queued++;
      for (LtSystem *L : sys) {
        L->steps_since_rebuild++;
        L->step++;
      }
    }
    APO_HIP(hipEventRecord(m_ev, m));
    for (LtSystem *L : sys) APO_HIP(hipStreamWaitEvent(L->st, m_ev, 0));
    APO_HIP(hipStreamWaitEvent(caller, m_ev, 0));
  }
};

LtBatch *create_batch(LtSystem **Ls, int n, int lanes) {
  if (n <= 0) throw InvalidArgument("a batch needs at least one system");
  for (int i = 0; i < n; i++)
    if (!Ls[i]) throw InvalidArgument("null integrator in batch");
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++)
      if (Ls[i] == Ls[j] || Ls[i]->nb == Ls[j]->nb)
        throw InvalidArgument("a batch must not hold one system twice");
  auto *B = new LtBatch();
  try {
    B->sys.assign(Ls, Ls + n);
    for (LtSystem *L : B->sys) L->n_batches++;  // ~LtBatch gives them back
    B->lanes = std::max(1, std::min(lanes > 0 ? lanes : 4, n));
    // one stream per system inside a step: no fork/join within a system
    for (LtSystem *L : B->sys) {
      B->was_single.push_back(L->nb->single_stream);
      if (!L->nb->single_stream) {
        L->nb->single_stream = true;
        L->nb->generation++;
      }
    }
    APO_HIP(hipStreamCreateWithFlags(&B->m, hipStreamNonBlocking));
    APO_HIP(hipEventCreateWithFlags(&B->fork_ev, hipEventDisableTiming));
    APO_HIP(hipEventCreateWithFlags(&B->m_ev, hipEventDisableTiming));
    B->lane.resize(B->lanes, nullptr);
    B->lane_ev.resize(B->lanes, nullptr);
    for (int l = 0; l < B->lanes; l++) {
      APO_HIP(hipStreamCreateWithFlags(&B->lane[l], hipStreamNonBlocking));
      APO_HIP(hipEventCreateWithFlags(&B->lane_ev[l], hipEventDisableTiming));
    }
    APO_HIP(hipEventCreateWithFlags(&B->sort_ev, hipEventDisableTiming));
    for (int i = 0; i < 2; i++) {
      APO_HIP(hipStreamCreateWithFlags(&B->side[i], hipStreamNonBlocking));
      APO_HIP(hipEventCreateWithFlags(&B->side_ev[i], hipEventDisableTiming));
    }
    for (LtSystem *L : B->sys) {
      if (L->nb->precision != 0)
        throw InvalidArgument("a batch runs the mixed-precision force path only");
      if (L->list_policy != 0)
        throw InvalidArgument("a batch rebuilds lists at the fixed cadence (list_policy 0)");
    }
    B->pme = pme_batch_create();
    B->bond = bonded_batch_create();
  } catch (...) {
    delete B;
    throw;
  }
  return B;
}

void destroy_batch(LtBatch *B) { delete B; }

void batch_autotune(LtBatch &B, hipStream_t caller, double *report9) {
  B.autotune(caller, report9);
}

void batch_propagate(LtBatch &B, int nsteps, hipStream_t caller) {
  if (nsteps < 0) throw InvalidArgument("nsteps must be >= 0");
  B.propagate(nsteps, caller);
}

void lt_propagate_batch(LtSystem **Ls, int n, int nsteps, hipStream_t caller) {
  if (n == 0) return;
  LtBatch *B = create_batch(Ls, n, 0);
  try {
    batch_propagate(*B, nsteps, caller);
  } catch (...) {
    delete B;
    throw;
  }
  // the batch's streams may still hold queued work: release after it drains
  APO_HIP(hipStreamSynchronize(B->m));
  delete B;
}

void lt_get_state(LtSystem &L, double *xyz, double *vel, double *dprev, double *grad,
                  double *thermo, int64_t *step) {
  APO_HIP(hipStreamSynchronize(L.st));
  // every rebuild so far has finished: report its errors (capacities grow
  // only right before a rebuild, when the current lists are discarded)
  list_export_check(*L.nb, L.st, false);
  const int n = L.n;
  std::vector<double> soa(3 * (size_t)n);
  auto from_soa = [&](const double *dev, double *out) {
    APO_HIP(hipMemcpy(soa.data(), dev, soa.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) out[3 * (size_t)i + d] = soa[(size_t)d * n + i];
  };
  if (xyz) from_soa(L.x.get(), xyz);
  if (vel) from_soa(L.v.get(), vel);
  if (dprev) from_soa(L.dp.get(), dprev);
  if (grad) {
    std::vector<int64_t> g(3 * (size_t)L.gstride);
    APO_HIP(hipMemcpy(g.data(), L.grad.get(), g.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++)
        grad[3 * (size_t)i + d] = (double)g[(size_t)d * L.gstride + i] * 9.094947017729282e-13;
  }
  if (thermo)
    APO_HIP(hipMemcpy(thermo, L.thermo.get(), 6 * sizeof(double), hipMemcpyDeviceToHost));
  if (step) *step = (int64_t)L.step;
}

// Restart: coordinates, velocities and previous half-step displacements
// (atom-major, 3 * n each) replace the state, and the step counter -- the
// position of the noise stream -- continues from `step`
// (CudaLangevinThermostatIntegrator::initializeFromRestartFile,
// src/CudaLangevinThermostatIntegrator.cu:264-540).  The integrator counts as
// initialised: the next step uses these displacements as they are.
void lt_set_state(LtSystem &L, const double *xyz, const double *vel, const double *dprev,
                  unsigned long long step) {
  if (!xyz || !vel || !dprev) throw InvalidArgument("restart state needs x, v and dprev");
  if (L.n_batches > 0) throw InvalidArgument("the integrator is a member of a batch");
  APO_HIP(hipStreamSynchronize(L.st));
  const int n = L.n;
  std::vector<double> soa(3 * (size_t)n);
  auto up = [&](const double *a, DevBuf<double> &dst) {
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) soa[(size_t)d * n + i] = a[3 * (size_t)i + d];
    APO_HIP(hipMemcpy(dst.get(), soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
  };
  up(xyz, L.x);
  up(vel, L.v);
  up(dprev, L.dp);
  std::vector<float> q(n);
  APO_HIP(hipMemcpy(q.data(), L.charge.get(), n * sizeof(float), hipMemcpyDeviceToHost));
  std::vector<float4> q4(n);
  for (int i = 0; i < n; i++)
    q4[i] = make_float4((float)xyz[3 * (size_t)i], (float)xyz[3 * (size_t)i + 1],
                        (float)xyz[3 * (size_t)i + 2], q[i]);
  APO_HIP(hipMemcpy(L.xyzq.get(), q4.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  unsigned long long c[3];
  APO_HIP(hipMemcpy(c, L.counters.get(), sizeof(c), hipMemcpyDeviceToHost));
  c[0] = step;
  APO_HIP(hipMemcpy(L.counters.get(), c, sizeof(c), hipMemcpyHostToDevice));
  L.step = step;
  L.steps_since_rebuild = 0;
  L.initialized = true;
}

void lt_launch_info(LtSystem &L, double *out6) {
  out6[0] = L.trial_pending() ? -1.0 : (double)(L.nb->pme_prio | (L.nb->pair_first << 1));
  out6[1] = (double)L.trial;
  for (int k = 0; k < LtSystem::kTrialSettings; k++) out6[2 + k] = L.trial_ms[k];
}

const float *lt_xyzq(LtSystem &L) { return reinterpret_cast<const float *>(L.xyzq.get()); }

void lt_list_info(LtSystem &L, double *out4) {
  APO_HIP(hipStreamSynchronize(L.st));
  double t[8];
  APO_HIP(hipMemcpy(t, L.thermo.get(), sizeof(t), hipMemcpyDeviceToHost));
  out4[0] = t[6];
  out4[1] = t[7];
  out4[2] = (double)L.builds_triggered;
  out4[3] = L.half_buffer;
}

}  // namespace apo
