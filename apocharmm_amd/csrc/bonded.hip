// Bonded terms and the 1-4 / exclusion-correction pairs.
//
// Physics follows CudaBondedForce (bond :84-146, angle :222-300 with cos
// clamped to +-0.999, dihedral/improper :386-684 of src/CudaBondedForce.cu)
// and calc_14_force_kernel (src/CudaDirectForceKernels.cu:640-870): no cutoff
// and no periodic image for these lists (ishift 13 in the reference).
// One thread per term, FP32 (or FP64) math, int64 2^-40 atomics, per-block
// FP64 energy/virial partials reduced later in fixed order.
#include <cmath>

#include "system.hpp"

namespace apo {

__device__ __forceinline__ unsigned long long fxr(float v) {
  return (unsigned long long)(long long)__builtin_rintf(v * 1099511627776.0f);
}
__device__ __forceinline__ unsigned long long fxr(double v) {
  return (unsigned long long)(long long)__builtin_rint(v * 1099511627776.0);
}

template <typename CT>
__device__ __forceinline__ void add_grad(unsigned long long *f, int stride, int i, CT gx, CT gy,
                                         CT gz) {
  atomicAdd(&f[i], fxr(gx));
  atomicAdd(&f[(size_t)stride + i], fxr(gy));
  atomicAdd(&f[2 * (size_t)stride + i], fxr(gz));
}

template <int NT>
__device__ __forceinline__ void block_partials(double *vals, int nv, double *sh, double *out) {
  for (int k = 0; k < nv; k++) {
    double v = vals[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) sh[(threadIdx.x >> 6) * 16 + k] = v;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double s = 0;
    for (int w = 0; w < NT / 64; w++) s += sh[w * 16 + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// W -= d (x) g, symmetric part kept as xx xy xz yy yz zz
template <typename CT>
__device__ __forceinline__ void vir_add(double *v, CT dx, CT dy, CT dz, CT gx, CT gy, CT gz) {
  v[0] -= (double)(dx * gx);
  v[1] -= (double)(dx * gy);
  v[2] -= (double)(dx * gz);
  v[3] -= (double)(dy * gy);
  v[4] -= (double)(dy * gz);
  v[5] -= (double)(dz * gz);
}

template <typename CT> __device__ __forceinline__ CT sqrt_ct(CT x);
template <> __device__ __forceinline__ float sqrt_ct<float>(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double sqrt_ct<double>(double x) { return sqrt(x); }
template <typename CT> __device__ __forceinline__ CT acos_ct(CT x);
template <> __device__ __forceinline__ float acos_ct<float>(float x) { return acosf(x); }
template <> __device__ __forceinline__ double acos_ct<double>(double x) { return acos(x); }
template <typename CT> __device__ __forceinline__ CT asin_ct(CT x);
template <> __device__ __forceinline__ float asin_ct<float>(float x) { return asinf(x); }
template <> __device__ __forceinline__ double asin_ct<double>(double x) { return asin(x); }

// Stiff terms (bonds k ~ 450, angles k ~ 55) turn the FP32 rounding of r or
// theta (~3e-8) into ~3e-5 kcal/mol/A of force, larger than the whole
// direct-space error budget; their geometric core (r - b0, theta - theta0)
// is therefore evaluated in FP64 in both modes (a few ops per term).  The
// force components stay in CT.
__device__ __forceinline__ double bond_stretch(const float4 &xi, const float4 &xj, double b0,
                                               double &dx, double &dy, double &dz, double &r) {
  dx = (double)xi.x - (double)xj.x;
  dy = (double)xi.y - (double)xj.y;
  dz = (double)xi.z - (double)xj.z;
  r = sqrt(dx * dx + dy * dy + dz * dz);
  return r - b0;
}

struct BondedArgs {
  const float4 *xyzq;
  const BondT *bonds, *ureyb;
  const AngleT *angles;
  const DiheT *dihe, *imdihe;
  int nb, nub, na, nd, ni;
  int stride;
};

// Dihedral gradient from dE/dphi (F, G, H, A = F x G, B = H x G convention of
// src/CudaBondedForce.cu:515-580): returns cos/sin of phi via (ct, st).
template <typename CT>
__device__ __forceinline__ void dihedral_geom(const float4 *x, int i, int j, int k, int l, CT &ct,
                                              CT &st, CT (&F)[3], CT (&G)[3], CT (&H)[3],
                                              CT (&A)[3], CT (&B)[3], CT &ra2, CT &rb2, CT &rg) {
  const float4 xi = x[i], xj = x[j], xk = x[k], xl = x[l];
  F[0] = (CT)xi.x - (CT)xj.x; F[1] = (CT)xi.y - (CT)xj.y; F[2] = (CT)xi.z - (CT)xj.z;
  G[0] = (CT)xj.x - (CT)xk.x; G[1] = (CT)xj.y - (CT)xk.y; G[2] = (CT)xj.z - (CT)xk.z;
  H[0] = (CT)xl.x - (CT)xk.x; H[1] = (CT)xl.y - (CT)xk.y; H[2] = (CT)xl.z - (CT)xk.z;
  A[0] = F[1] * G[2] - F[2] * G[1]; A[1] = F[2] * G[0] - F[0] * G[2]; A[2] = F[0] * G[1] - F[1] * G[0];
  B[0] = H[1] * G[2] - H[2] * G[1]; B[1] = H[2] * G[0] - H[0] * G[2]; B[2] = H[0] * G[1] - H[1] * G[0];
  ra2 = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  rb2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
  rg = sqrt_ct<CT>(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
  const CT rab = (CT)1 / sqrt_ct<CT>(ra2 * rb2);
  ct = (A[0] * B[0] + A[1] * B[1] + A[2] * B[2]) * rab;
  st = rg * rab * (A[0] * H[0] + A[1] * H[1] + A[2] * H[2]);
}

struct BondedBatchArgs {
  BondedArgs a;
  unsigned long long *f;
};

template <typename CT, bool CALC_E, bool CALC_V>
__device__ __forceinline__ void bonded_body(int bid, const BondedArgs &a,
                                            unsigned long long *__restrict__ f,
                                            double *__restrict__ part) {
  __shared__ double sh[4 * 16];
  double acc[kBondPart];
#pragma unroll
  for (int k = 0; k < kBondPart; k++) acc[k] = 0.0;
  int idx = bid * 256 + threadIdx.x;
  const float4 *x = a.xyzq;
  if (idx < a.nb + a.nub) {
    const bool ub = idx >= a.nb;
    const BondT b = ub ? a.ureyb[idx - a.nb] : a.bonds[idx];
    const float4 xi = x[b.i], xj = x[b.j];
    double ddx, ddy, ddz, r;
    const double db = bond_stretch(xi, xj, (double)b.b0, ddx, ddy, ddz, r);
    if (CALC_E) acc[ub ? 2 : 0] += (double)b.kb * db * db;
    const CT fij = (CT)(2.0 * (double)b.kb * db / r);
    const CT dx = (CT)ddx, dy = (CT)ddy, dz = (CT)ddz;
    add_grad<CT>(f, a.stride, b.i, fij * dx, fij * dy, fij * dz);
    add_grad<CT>(f, a.stride, b.j, -fij * dx, -fij * dy, -fij * dz);
    if (CALC_V) vir_add<CT>(acc + 5, dx, dy, dz, fij * dx, fij * dy, fij * dz);
  } else if ((idx -= a.nb + a.nub) < a.na) {
    const AngleT g = a.angles[idx];
    const float4 xi = x[g.i], xj = x[g.j], xk = x[g.k];
    // FP64 geometric core (see bond_stretch)
    const double ddxij = (double)xi.x - (double)xj.x, ddyij = (double)xi.y - (double)xj.y,
                 ddzij = (double)xi.z - (double)xj.z;
    const double ddxkj = (double)xk.x - (double)xj.x, ddykj = (double)xk.y - (double)xj.y,
                 ddzkj = (double)xk.z - (double)xj.z;
    const double drij = sqrt(ddxij * ddxij + ddyij * ddyij + ddzij * ddzij);
    const double drkj = sqrt(ddxkj * ddxkj + ddykj * ddykj + ddzkj * ddzkj);
    double dcst = (ddxij * ddxkj + ddyij * ddykj + ddzij * ddzkj) / (drij * drkj);
    dcst = fmin(fmax(dcst, -0.999), 0.999);
    const double dda = acos(dcst) - (double)g.t0;
    if (CALC_E) acc[1] += (double)g.kt * dda * dda;
    const CT df = (CT)(-2.0 * (double)g.kt * dda / sqrt(1.0 - dcst * dcst));
    const CT cst = (CT)dcst;
    const CT dxij = (CT)ddxij, dyij = (CT)ddyij, dzij = (CT)ddzij;
    const CT dxkj = (CT)ddxkj, dykj = (CT)ddykj, dzkj = (CT)ddzkj;
    const CT rij_inv = (CT)(1.0 / drij), rkj_inv = (CT)(1.0 / drkj);
    const CT uxi = dxij * rij_inv, uyi = dyij * rij_inv, uzi = dzij * rij_inv;
    const CT uxk = dxkj * rkj_inv, uyk = dykj * rkj_inv, uzk = dzkj * rkj_inv;
    // gradient on i and k: -dE/dtheta / sin * d cos / dx
    const CT gix = df * rij_inv * (uxk - cst * uxi), giy = df * rij_inv * (uyk - cst * uyi),
             giz = df * rij_inv * (uzk - cst * uzi);
    const CT gkx = df * rkj_inv * (uxi - cst * uxk), gky = df * rkj_inv * (uyi - cst * uyk),
             gkz = df * rkj_inv * (uzi - cst * uzk);
    add_grad<CT>(f, a.stride, g.i, gix, giy, giz);
    add_grad<CT>(f, a.stride, g.k, gkx, gky, gkz);
    add_grad<CT>(f, a.stride, g.j, -gix - gkx, -giy - gky, -giz - gkz);
    if (CALC_V) {
      vir_add<CT>(acc + 5, dxij, dyij, dzij, gix, giy, giz);
      vir_add<CT>(acc + 5, dxkj, dykj, dzkj, gkx, gky, gkz);
    }
  } else if ((idx -= a.na) < a.nd + a.ni) {
    const bool imp = idx >= a.nd;
    const DiheT d = imp ? a.imdihe[idx - a.nd] : a.dihe[idx];
    CT ct, st, F[3], G[3], H[3], A[3], B[3], ra2, rb2, rg;
    dihedral_geom<CT>(x, d.i, d.j, d.k, d.l, ct, st, F, G, H, A, B, ra2, rb2, rg);
    const int nper = (int)d.n;
    CT dEdphi, e;
    if (!imp && nper == 0) {
      // multiplicity 0: constant k, no force (src/CudaBondedForce.cu:419-421)
      e = (CT)d.kc;
      dEdphi = 0;
    } else if (nper != 0) {
      // k (1 + cos(n phi - delta)) via the cos/sin recurrence
      CT c1 = 1, s1 = 0;
      for (int p = 0; p < (nper < 0 ? -nper : nper); p++) {
        CT c2 = c1 * ct - s1 * st;
        s1 = c1 * st + s1 * ct;
        c1 = c2;
      }
      const int na_ = nper < 0 ? -nper : nper;
      e = (CT)d.kc * ((CT)1 + c1 * (CT)d.cd + s1 * (CT)d.sd);
      dEdphi = -(CT)d.kc * (CT)na_ * (s1 * (CT)d.cd - c1 * (CT)d.sd);
    } else {
      // harmonic improper on phi - phi0 wrapped to (-pi, pi]
      const CT ca = ct * (CT)d.cd + st * (CT)d.sd;
      const CT sa = st * (CT)d.cd - ct * (CT)d.sd;
      CT ap;
      if (ca > (CT)0.1)
        ap = asin_ct<CT>(sa);
      else {
        ap = acos_ct<CT>(fmax(ca, (CT)-1));
        ap = sa > (CT)0 ? ap : -ap;
      }
      e = (CT)d.kc * ap * ap;
      dEdphi = (CT)2 * (CT)d.kc * ap;
    }
    if (CALC_E) acc[imp ? 4 : 3] += (double)e;
    // dphi/dri = -|G|/A^2 A, dphi/drl = |G|/B^2 B, cross terms on j and k
    const CT fg = F[0] * G[0] + F[1] * G[1] + F[2] * G[2];
    const CT hg = H[0] * G[0] + H[1] * G[1] + H[2] * G[2];
    const CT ia2 = dEdphi / ra2, ib2 = dEdphi / rb2;
    CT gi[3], gj[3], gk[3], gl[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const CT di = -rg * ia2 * A[c];
      const CT dl = rg * ib2 * B[c];
      const CT cr = fg * ia2 / rg * A[c] - hg * ib2 / rg * B[c];
      gi[c] = di;
      gj[c] = -di + cr;
      gk[c] = -dl - cr;
      gl[c] = dl;
    }
    add_grad<CT>(f, a.stride, d.i, gi[0], gi[1], gi[2]);
    add_grad<CT>(f, a.stride, d.j, gj[0], gj[1], gj[2]);
    add_grad<CT>(f, a.stride, d.k, gk[0], gk[1], gk[2]);
    add_grad<CT>(f, a.stride, d.l, gl[0], gl[1], gl[2]);
    if (CALC_V) {
      vir_add<CT>(acc + 5, F[0], F[1], F[2], gi[0], gi[1], gi[2]);
      vir_add<CT>(acc + 5, -G[0], -G[1], -G[2], gk[0], gk[1], gk[2]);
      vir_add<CT>(acc + 5, H[0] - G[0], H[1] - G[1], H[2] - G[2], gl[0], gl[1], gl[2]);
    }
  }
  if (CALC_E || CALC_V) block_partials<256>(acc, kBondPart, sh, part + bid * kBondPart);
}
template <typename CT, bool CALC_E, bool CALC_V>
__global__ __launch_bounds__(256) void k_bonded(BondedArgs a, unsigned long long *__restrict__ f,
                                                double *__restrict__ part) {
  bonded_body<CT, CALC_E, CALC_V>(blockIdx.x, a, f, part);
}
__global__ __launch_bounds__(256) void kb_bonded(const BondedBatchArgs *__restrict__ tab,
                                                 const int *__restrict__ bstart, int nsys) {
  const int s = batch_system(bstart, nsys);
  const BondedBatchArgs &a = tab[s];
  bonded_body<float, false, false>(blockIdx.x - bstart[s], a.a, a.f, nullptr);
}


// 1-4 pairs and 1-2/1-3 Ewald exclusion corrections.
struct P14Args {
  const float4 *xyzq;
  const int *vdwtype;
  const int2 *in14, *ex14;
  int nin, nex, stride;
  int ntypes;
  double e14fac;
};

struct BondedBatchArgs;
struct P14BatchArgs;

template <typename CT> struct Tab2;
template <> struct Tab2<float> { using T = float2; };
template <> struct Tab2<double> { using T = double2; };

template <typename CT> __device__ __forceinline__ CT erfc_ct(CT x);
template <> __device__ __forceinline__ float erfc_ct<float>(float x) { return erfcf(x); }
template <> __device__ __forceinline__ double erfc_ct<double>(double x) { return erfc(x); }
template <typename CT> __device__ __forceinline__ CT erf_ct(CT x);
template <> __device__ __forceinline__ float erf_ct<float>(float x) { return erff(x); }
template <> __device__ __forceinline__ double erf_ct<double>(double x) { return erf(x); }
template <typename CT> __device__ __forceinline__ CT exp_ct(CT x);
template <> __device__ __forceinline__ float exp_ct<float>(float x) { return expf(x); }
template <> __device__ __forceinline__ double exp_ct<double>(double x) { return exp(x); }

struct P14BatchArgs {
  P14Args a;
  PairConsts<float> P;
  const float2 *tab14;
  unsigned long long *f;
};

// Energies of energy steps are evaluated in FP64 from the same float xyzq (a
// 1-4 or exclusion term is ~1e2 larger than the sums it cancels into):
// PE / tab14_e are the FP64 constants and (6 c6, 12 c12) table.
template <typename CT, bool CALC_E, bool CALC_V>
__device__ __forceinline__ void pairs14_body(int bid, const P14Args &a, const PairConsts<CT> &P,
                                             const typename Tab2<CT>::T *__restrict__ tab14,
                                             unsigned long long *__restrict__ f,
                                             double *__restrict__ part,
                                             const PairConsts<double> &PE,
                                             const double2 *__restrict__ tab14_e) {
  __shared__ double sh[4 * 16];
  double acc[kP14Part];
#pragma unroll
  for (int k = 0; k < kP14Part; k++) acc[k] = 0.0;
  const int idx = bid * 256 + threadIdx.x;
  if (idx < a.nin + a.nex) {
    const bool ex = idx >= a.nin;
    const int2 pr = ex ? a.ex14[idx - a.nin] : a.in14[idx];
    const float4 xi = a.xyzq[pr.x], xj = a.xyzq[pr.y];
    const CT dx = (CT)xi.x - (CT)xj.x, dy = (CT)xi.y - (CT)xj.y, dz = (CT)xi.z - (CT)xj.z;
    const CT r2 = dx * dx + dy * dy + dz * dz;
    const CT r = sqrt_ct<CT>(r2), rinv = (CT)1 / r;
    const CT qq = (CT)332.0716 * (CT)xi.w * (CT)xj.w;
    // Ewald with e14fac: qq (erfc + e14fac - 1)/r; exclusions use e14fac = 0,
    // evaluated as -qq erf/r (no cancellation of erfc against 1 at small kappa r)
    const CT efac = ex ? (CT)0 : (CT)a.e14fac;
    const CT u_el = ex ? -qq * erf_ct<CT>(P.kappa * r) * rinv
                       : qq * (erfc_ct<CT>(P.kappa * r) + efac - (CT)1) * rinv;
    CT fr = -qq * P.two_kappa_sqrtpi * exp_ct<CT>(-P.kappa2 * r2) - u_el;
    double ed = 0.0, evd = 0.0;  // FP64 energies
    if (CALC_E) {
      const double ex_ = (double)xi.x - (double)xj.x, ey_ = (double)xi.y - (double)xj.y,
                   ez_ = (double)xi.z - (double)xj.z;
      const double er2 = ex_ * ex_ + ey_ * ey_ + ez_ * ez_, er = sqrt(er2), erinv = 1.0 / er;
      const double qqd = 332.0716 * (double)xi.w * (double)xj.w;
      ed = ex ? -qqd * erf(PE.kappa * er) * erinv
              : qqd * (erfc(PE.kappa * er) + a.e14fac - 1.0) * erinv;
      if (!ex) {
        const int ti = a.vdwtype[pr.x], tj = a.vdwtype[pr.y];
        const int aa = ti > tj ? ti : tj, bb = ti > tj ? tj : ti;
        const double2 c = tab14_e[aa * (aa + 1) / 2 + bb];
        const double er3 = erinv * erinv * erinv, er6 = er3 * er3;
        const bool eo = er2 > PE.ron2;
        const double e6 = eo ? er3 - PE.roffinv3 : er3, e12 = eo ? er6 - PE.roffinv6 : er6;
        const double A6 = eo ? PE.k6 : 1.0, A12 = eo ? PE.k12 : 1.0;
        const double C6 = eo ? 0.0 : PE.dv6, C12 = eo ? 0.0 : PE.dv12;
        evd = c.y * (1.0 / 12.0) * (A12 * e12 * e12 + C12) - c.x * (1.0 / 6.0) * (A6 * e6 * e6 + C6);
      }
    }
    if (!ex) {
      const int ti = a.vdwtype[pr.x], tj = a.vdwtype[pr.y];
      const int aa = ti > tj ? ti : tj, bb = ti > tj ? tj : ti;
      const typename Tab2<CT>::T c = tab14[aa * (aa + 1) / 2 + bb];
      const CT rinv2 = rinv * rinv, rinv3 = rinv * rinv2, rinv6 = rinv3 * rinv3;
      const bool outer = r2 > P.ron2;
      const CT A6 = outer ? P.k6 : (CT)1, B6 = outer ? P.roffinv3 : (CT)0;
      const CT A12 = outer ? P.k12 : (CT)1, B12 = outer ? P.roffinv6 : (CT)0;
      const CT t6 = rinv3 - B6, t12 = rinv6 - B12;
      fr += c.x * A6 * t6 * rinv3 - c.y * A12 * t12 * rinv6;
    }
    if (CALC_E) {
      acc[1] += evd;
      acc[ex ? 2 : 0] += ed;
    }
    const CT fij = fr * rinv * rinv;
    add_grad<CT>(f, a.stride, pr.x, fij * dx, fij * dy, fij * dz);
    add_grad<CT>(f, a.stride, pr.y, -fij * dx, -fij * dy, -fij * dz);
    if (CALC_V) vir_add<CT>(acc + 3, dx, dy, dz, fij * dx, fij * dy, fij * dz);
  }
  if (CALC_E || CALC_V) block_partials<256>(acc, kP14Part, sh, part + bid * kP14Part);
}
template <typename CT, bool CALC_E, bool CALC_V>
__global__ __launch_bounds__(256) void k_pairs14(P14Args a, const PairConsts<CT> P,
                                                 const typename Tab2<CT>::T *__restrict__ tab14,
                                                 unsigned long long *__restrict__ f,
                                                 double *__restrict__ part,
                                                 const PairConsts<double> PE,
                                                 const double2 *__restrict__ tab14_e) {
  pairs14_body<CT, CALC_E, CALC_V>(blockIdx.x, a, P, tab14, f, part, PE, tab14_e);
}
__global__ __launch_bounds__(256) void kb_pairs14(const P14BatchArgs *__restrict__ tab,
                                                  const int *__restrict__ bstart, int nsys) {
  const int s = batch_system(bstart, nsys);
  const P14BatchArgs &a = tab[s];
  const PairConsts<double> no_energy{};
  pairs14_body<float, false, false>(blockIdx.x - bstart[s], a.a, a.P, a.tab14, a.f, nullptr,
                                    no_energy, nullptr);
}


template <typename CT, typename F>
static void dispatch_ev(bool E, bool V, F &&fn) {
  if (E && V) fn(std::integral_constant<int, 3>());
  else if (E) fn(std::integral_constant<int, 1>());
  else if (V) fn(std::integral_constant<int, 2>());
  else fn(std::integral_constant<int, 0>());
}

void bonded_launch(NbSystem &S, const float4 *x, int64_t *force, int stride, bool E, bool V,
                   hipStream_t s) {
  BondedArgs a{x, S.bonds.get(), S.ureyb.get(), S.angles.get(), S.dihe.get(), S.imdihe.get(),
               S.nbond, S.nub, S.nangle, S.ndihe, S.nimdihe, stride};
  const int nt = S.nbond + S.nub + S.nangle + S.ndihe + S.nimdihe;
  auto *f = reinterpret_cast<unsigned long long *>(force);
  double *part = S.parts_ptr() + S.part.bond_off;
  const int G = S.part.bond_n;
  if (nt == 0) {
    if (E || V) launch_fill_f64(part, (size_t)G * kBondPart, s);
    return;
  }
  auto go = [&](auto ev) {
    constexpr int m = decltype(ev)::value;
    if (S.precision == 1)
      k_bonded<double, (m & 1) != 0, (m & 2) != 0><<<G, 256, 0, s>>>(a, f, part);
    else
      k_bonded<float, (m & 1) != 0, (m & 2) != 0><<<G, 256, 0, s>>>(a, f, part);
  };
  dispatch_ev<float>(E, V, go);
  APO_LAUNCH_CHECK();
}

void pairs14_launch(NbSystem &S, const float4 *x, int64_t *force, int stride, bool E, bool V,
                    hipStream_t s) {
  P14Args a{x, S.vdwtype.get(), S.in14.get(), S.ex14.get(), S.nin14, S.nex14, stride,
            S.ntypes, S.e14fac};
  auto *f = reinterpret_cast<unsigned long long *>(force);
  double *part = S.parts_ptr() + S.part.p14_off;
  const int G = S.part.p14_n;
  if (S.nin14 + S.nex14 == 0) {
    if (E || V) launch_fill_f64(part, (size_t)G * kP14Part, s);
    return;
  }
  auto go = [&](auto ev) {
    constexpr int m = decltype(ev)::value;
    if (S.precision == 1)
      k_pairs14<double, (m & 1) != 0, (m & 2) != 0>
          <<<G, 256, 0, s>>>(a, S.pc_d, S.vdwtab14_d.get(), f, part, S.pc_d,
                             S.vdwtab14_d.get());
    else
      k_pairs14<float, (m & 1) != 0, (m & 2) != 0>
          <<<G, 256, 0, s>>>(a, S.pc_f, S.vdwtab14.get(), f, part, S.pc_d, S.vdwtab14_d.get());
  };
  dispatch_ev<float>(E, V, go);
  APO_LAUNCH_CHECK();
}

// ---- fused launches over several systems (mixed precision, no energy) ------
struct BondedBatch {
  BatchLaunch<BondedBatchArgs> bonded;
  BatchLaunch<P14BatchArgs> p14;
};
BondedBatch *bonded_batch_create() { return new BondedBatch(); }
void bonded_batch_destroy(BondedBatch *b) { delete b; }
void bonded_batch_clear(BondedBatch &b) {
  b.bonded.clear();
  b.p14.clear();
}
void bonded_batch_add(BondedBatch &b, NbSystem &S, const float4 *x, int64_t *force, int stride) {
  if (S.precision != 0) throw InvalidArgument("fused batches run the mixed-precision path");
  auto *f = reinterpret_cast<unsigned long long *>(force);
  const int nt = S.nbond + S.nub + S.nangle + S.ndihe + S.nimdihe;
  if (nt > 0)
    b.bonded.add(BondedBatchArgs{BondedArgs{x, S.bonds.get(), S.ureyb.get(), S.angles.get(),
                                            S.dihe.get(), S.imdihe.get(), S.nbond, S.nub,
                                            S.nangle, S.ndihe, S.nimdihe, stride},
                                 f},
                 S.part.bond_n);
  if (S.nin14 + S.nex14 > 0)
    b.p14.add(P14BatchArgs{P14Args{x, S.vdwtype.get(), S.in14.get(), S.ex14.get(), S.nin14,
                                   S.nex14, stride, S.ntypes, S.e14fac},
                           S.pc_f, S.vdwtab14.get(), f},
              S.part.p14_n);
}
void bonded_batch_upload(BondedBatch &b, hipStream_t s) {
  b.bonded.upload(s);
  b.p14.upload(s);
}
void bonded_batch_launch(BondedBatch &b, hipStream_t s) {
  if (b.bonded.nsys())
    kb_bonded<<<b.bonded.blocks(), 256, 0, s>>>(b.bonded.d_args.get(), b.bonded.d_start.get(),
                                                b.bonded.nsys());
  if (b.p14.nsys())
    kb_pairs14<<<b.p14.blocks(), 256, 0, s>>>(b.p14.d_args.get(), b.p14.d_start.get(),
                                              b.p14.nsys());
  APO_LAUNCH_CHECK();
}

}  // namespace apo
