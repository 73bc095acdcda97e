// Direct-space PME pair kernel: erfc Coulomb + CHARMM force-switched LJ over
// 32x32 cell tiles (physics of calcForceKernel, include/CudaDirectForce_util.h
// :16-333, with pair_elec_force EWALD src/CudaDirectForceKernels.cu:476-490
// and pair_vdw_force VFSW :263-286).
//
// gfx950 design (not the reference's warp-rotation scheme):
//  * one wave64 = one work item: an i-group of 16 consecutive cell-sorted
//    slots and its packed j list (slots with a live pair at list build, 64
//    per step), so few lanes idle on pairs far from the group;
//  * the 16 i-atoms are wave-uniform (SGPR operands); i-forces live in 48
//    registers per lane and are reduced once per item; j-forces are flushed
//    per step, skipped when zero;
//  * coordinates are 32-bit fixed point per box edge: the int difference
//    wraps to the minimum image exactly, then converts to float;
//  * forces accumulate in FP32 within a step/item and are rounded to int64
//    2^-40 fixed point per flush, so results are bitwise reproducible.
#include <cmath>

#include "system.hpp"

namespace apo {

template <typename CT> struct Vec2;
template <> struct Vec2<float> { using T = float2; };
template <> struct Vec2<double> { using T = double2; };

__device__ __forceinline__ unsigned long long fx_round(float v) {
  return (unsigned long long)(long long)__builtin_rintf(v * 1099511627776.0f);
}
__device__ __forceinline__ unsigned long long fx_round(double v) {
  return (unsigned long long)(long long)__builtin_rint(v * 1099511627776.0);
}

// Coulomb part: returns r dU/dr; u = U.
__device__ __forceinline__ float ewald_pair(float r2, float r, float rinv, float qq,
                                            const PairConsts<float> &P, float &u) {
  const float e2 = __builtin_amdgcn_exp2f(-P.k2log2e * r2);  // exp(-k^2 r^2)
  const float t = __builtin_amdgcn_rcpf(fmaf(P.erfc_p, r, 1.0f));
  float p = P.erfc_c[8];
#pragma unroll
  for (int k = 7; k >= 0; k--) p = fmaf(p, t, P.erfc_c[k]);
  u = qq * (e2 * p) * rinv;
  return -qq * P.two_kappa_sqrtpi * e2 - u;
}
__device__ __forceinline__ double ewald_pair(double r2, double r, double rinv, double qq,
                                             const PairConsts<double> &P, double &u) {
  const double e2 = exp(-P.kappa2 * r2);
  u = qq * erfc(P.kappa * r) * rinv;
  return -qq * P.two_kappa_sqrtpi * e2 - u;
}

__device__ __forceinline__ float rsqrt_ct(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_ct(double x) { return 1.0 / sqrt(x); }

template <typename CT> __device__ __forceinline__ CT charge_of(int bits);
template <> __device__ __forceinline__ float charge_of<float>(int bits) { return __int_as_float(bits); }
template <> __device__ __forceinline__ double charge_of<double>(int bits) {
  return (double)__int_as_float(bits);
}

// Sums v[3k + c] (atom k < kGroup, component c) over the 64 lanes.  Each
// exchange halves the set a lane keeps (xor 32, 16, 8, 4 select atoms 8, 4,
// 2, 1), then xor 2 and 1 finish the sums: lane L ends with atom (L >> 2) in
// v[0..2].
template <typename CT> __device__ __forceinline__ void group_reduce(CT (&v)[3 * kGroup], int lane) {
  static_assert(kGroup == 16, "group_reduce is written for 16-atom groups");
#pragma unroll
  for (int w = 24, bit = 32; w >= 3; w >>= 1, bit >>= 1) {
#pragma unroll
    for (int k = 0; k < w; k++) {
      const bool up = lane & bit;
      const CT send = up ? v[k] : v[k + w];
      const CT keep = up ? v[k + w] : v[k];
      v[k] = keep + __shfl_xor(send, bit);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++)
    for (int off = 2; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
}

constexpr int kMaxTypesLds = 32;  // square vdW table in LDS up to 32 types
// beyond that the triangular table goes to LDS while it fits this many bytes
constexpr size_t kMaxTriLdsBytes = 32768;

// Where the pair kernel reads the (6 c6, 12 c12) table: TAB_GLOBAL the
// triangular table in global memory, TAB_SQUARE an ntypes x ntypes copy in
// LDS (row = type * ntypes, no index arithmetic per pair), TAB_TRI the
// triangular table copied to LDS.
// TAB_COMBINE (mixed precision, more types than the triangular LDS table
// holds, every entry following the combination rule): 6 c6 = 12 e R^6 and
// 12 c12 = 12 e R^12 from per-type (sqrt(eps), Rmin/2), e = se_i se_j,
// R = rh_i + rh_j.
// TAB_PQA (mixed precision, at most kMaxTypesPqa types): the switched force
// coefficients (P, Q, A12) of the inner and outer regions per type pair
// (NbSystem::vdwpqa) in LDS, so one table read and three multiply-adds give
// the vdW force in either region.
enum { TAB_GLOBAL = 0, TAB_SQUARE = 1, TAB_TRI = 2, TAB_COMBINE = 3, TAB_PQA = 4 };
template <typename T2> static int pair_tab_mode(int ntypes, bool combine) {
  if (sizeof(T2) == sizeof(float2) && ntypes <= kMaxTypesPqa) return TAB_PQA;
  if (ntypes <= kMaxTypesLds) return TAB_SQUARE;
  if ((size_t)ntypes * (ntypes + 1) / 2 * sizeof(T2) <= kMaxTriLdsBytes) return TAB_TRI;
  return combine && sizeof(T2) == sizeof(float2) ? TAB_COMBINE : TAB_GLOBAL;
}
template <typename T2> static size_t pair_tab_lds_bytes(int mode, int ntypes) {
  if (mode == TAB_PQA) return 2 * (size_t)ntypes * ntypes * sizeof(float4);
  return mode == TAB_SQUARE ? (size_t)ntypes * ntypes * sizeof(T2)
                            : (mode == TAB_TRI ? (size_t)ntypes * (ntypes + 1) / 2 * sizeof(T2) : 0);
}

// One wave64 = one work item: an i-group of kGroup consecutive sorted slots
// and up to kItemSteps steps of 64 packed j entries (candidate tile of the
// i-cell, slot in that j-cell, skip mask).  Each lane holds one j
// atom per step; the kGroup i-atoms are wave-uniform (SGPR operands) and
// visited in an unrolled loop, skipped with a scalar branch when no lane has
// a live pair with them.  i-forces accumulate in 3*kGroup registers and are
// reduced once per item; j-forces are flushed per step.
// TAB: where the vdW table is read (TAB_GLOBAL / TAB_SQUARE / TAB_TRI /
// TAB_COMBINE / TAB_PQA).
// ISO (cubic box, mixed precision): the coordinate differences stay in
// fixed-point units, r^2 is compared with the cutoff in those units and
// converted once per pair, and the forces accumulate per unit and are scaled
// to A when they are flushed.
// PRUNE (list pruning fused into an MD step, mixed precision): the kernel
// runs on the build's list and also writes the pruned copy the next steps
// read: each entry gains the skip bits of its pairs at or beyond the prune
// radius (rp2t, in the units r^2 is compared in), entries left with no live
// pair are dropped, and each item's surviving entries are packed to the front
// of its own region (null entries after them, so every entry of the copy
// stays valid) with the item's end shortened to whole steps.  One wave owns an
// item, so the packing needs no coordination.  (The idea of pruning the list
// between builds is GROMACS's dynamic pruning; the reference keeps its list.)
template <typename CT, bool CALC_E, bool CALC_V, int TAB, bool ISO, bool PRUNE>
__device__ __forceinline__ void pair_body(
    int bid, int nb, const int4 *__restrict__ posq, const int *__restrict__ types,
    const int4 *__restrict__ items, const int *__restrict__ iorder, const uint32_t *__restrict__ ent,
    const int *__restrict__ cand_j, const int *__restrict__ meta,
    const typename Vec2<CT>::T *__restrict__ vdwtab, const typename Vec2<CT>::T *__restrict__ vdwsq,
    const float2 *__restrict__ ljtype, const float4 *__restrict__ vdwpqa, int ntypes,
    const PairConsts<CT> &P, CT qfac, unsigned long long *__restrict__ fsort, int ss,
    double *__restrict__ part, const PairConsts<double> &PE,
    const double2 *__restrict__ vdwtab_e, uint32_t *__restrict__ ent_out,
    int4 *__restrict__ items_out, CT rp2t) {
  static_assert(!ISO || sizeof(CT) == sizeof(float), "fixed-point-unit distances are FP32 only");
  static_assert(!PRUNE || (sizeof(CT) == sizeof(float) && !CALC_E && !CALC_V),
                "fused pruning is an MD-step variant");
  static_assert(TAB != TAB_PQA || sizeof(CT) == sizeof(float), "the PQA table is FP32 only");
  using T2 = typename Vec2<CT>::T;
  extern __shared__ char smem_pair[];
  T2 *sh_tab = reinterpret_cast<T2 *>(smem_pair);
  float4 *sh_pqa = reinterpret_cast<float4 *>(smem_pair);
  const int nt2 = ntypes * ntypes;
  if (TAB == TAB_PQA) {
    for (int k = threadIdx.x; k < 2 * nt2; k += 64) sh_pqa[k] = vdwpqa[k];
    __syncthreads();
  } else if (TAB == TAB_SQUARE) {
    for (int k = threadIdx.x; k < ntypes * ntypes; k += 64) sh_tab[k] = vdwsq[k];
    __syncthreads();
  } else if (TAB == TAB_TRI) {
    for (int k = threadIdx.x; k < ntypes * (ntypes + 1) / 2; k += 64) sh_tab[k] = vdwtab[k];
    __syncthreads();
  }
  const int lane = threadIdx.x;
  const int nitem = meta[3];
  double acc[kPairPart];
#pragma unroll
  for (int k = 0; k < kPairPart; k++) acc[k] = 0.0;

  // items in the list's processing order (longest first)
  for (int p = bid; p < nitem; p += nb) {
    const int item = iorder[p];
    const int4 it = items[item];
    const int ibase = it.x * kGroup;
    // i-data to SGPRs: lane k < kGroup loads slot ibase + k
    const int4 pil = posq[ibase + (lane & (kGroup - 1))];
    const int til = types[ibase + (lane & (kGroup - 1))];
    int pix[kGroup], piy[kGroup], piz[kGroup], rowi[kGroup];
    CT qi[kGroup];
    float sei[kGroup], rhi[kGroup];  // TAB_COMBINE only
    float2 ljl = make_float2(0.f, 0.f);
    if (TAB == TAB_COMBINE) ljl = ljtype[til];
#pragma unroll
    for (int k = 0; k < kGroup; k++) {
      pix[k] = __builtin_amdgcn_readlane(pil.x, k);
      piy[k] = __builtin_amdgcn_readlane(pil.y, k);
      piz[k] = __builtin_amdgcn_readlane(pil.z, k);
      qi[k] = charge_of<CT>(__builtin_amdgcn_readlane(pil.w, k));
      const int t = __builtin_amdgcn_readlane(til, k);
      // TAB_PQA: byte offset of the type's row of (inner, outer) float4 pairs
      rowi[k] = TAB == TAB_SQUARE ? t * ntypes : (TAB == TAB_PQA ? t * ntypes * 32 : t);
      if (TAB == TAB_COMBINE) {
        sei[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ljl.x), k));
        rhi[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ljl.y), k));
      }
    }
    CT fi[3 * kGroup];
#pragma unroll
    for (int k = 0; k < 3 * kGroup; k++) fi[k] = (CT)0;

    // the next step's j data is loaded while the current one computes
    const int *tiles = cand_j + it.w;
    auto slot_of = [&](uint32_t e) { return tiles[(e & 0xffffu) >> 5] * kCell + (int)(e & 31u); };
    // two-deep pipeline: the entry two steps ahead, the coordinates one step ahead
    uint32_t e_n = ent[it.y + lane];
    int jl_n = slot_of(e_n);
    int4 pj_n = posq[jl_n];
    int tj_n = types[jl_n];
    uint32_t e_nn = it.y + 64 < it.z ? ent[it.y + 64 + lane] : 0u;
    int wpos = it.y;  // PRUNE: next free slot of the item's pruned entries
    for (int e0 = it.y; e0 < it.z; e0 += 64) {
      const uint32_t e = e_n;
      const int jl = jl_n;
      const int4 pj = pj_n;
      const int tj = tj_n;
      if (e0 + 64 < it.z) {
        e_n = e_nn;
        jl_n = slot_of(e_n);
        pj_n = posq[jl_n];
        tj_n = types[jl_n];
        if (e0 + 128 < it.z) e_nn = ent[e0 + 128 + lane];
      }
      const uint32_t m = e >> 16;
      uint32_t mp = m;  // PRUNE: m plus the pairs at or beyond the prune radius
      // i-atoms with no live pair in this step are skipped with a scalar branch
      uint32_t act = ~m & ((1u << kGroup) - 1u);
      for (int off = 32; off > 0; off >>= 1) act |= __shfl_xor(act, off);
      act = __builtin_amdgcn_readfirstlane(act);
      const CT qj = charge_of<CT>(pj.w) * qfac;
      const int tjb = tj * 32;  // TAB_PQA: byte offset of column tj
      float2 ljj = make_float2(0.f, 0.f);
      if (TAB == TAB_COMBINE) ljj = ljtype[tj];
      CT fjx = 0, fjy = 0, fjz = 0;
      // energies accumulate per pair in FP64 (large cancelling sums); the
      // virial per step in CT
      double el = 0, ev = 0;
      CT v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
#pragma unroll
      for (int k = 0; k < kGroup; k++) {
        if (!(act & (1u << k))) continue;
        // ISO: differences in fixed-point units (dx = unit * dxu)
        const CT dxu = (CT)(int)(pix[k] - pj.x);
        const CT dyu = (CT)(int)(piy[k] - pj.y);
        const CT dzu = (CT)(int)(piz[k] - pj.z);
        const CT dx = ISO ? dxu : dxu * P.sx;
        const CT dy = ISO ? dyu : dyu * P.sy;
        const CT dz = ISO ? dzu : dzu * P.sz;
        const CT r2t = dx * dx + dy * dy + dz * dz;
        if (PRUNE) mp |= (r2t < rp2t ? 0u : 1u) << k;
        if (!(m & (1u << k)) && r2t < (ISO ? P.roff2u : P.roff2)) {
          const CT r2 = ISO ? r2t * P.s2 : r2t;
          const CT rinv = rsqrt_ct(r2);
          const CT rinv2 = rinv * rinv;
          const CT qq = qi[k] * qj;
          CT fij;
          if constexpr (TAB == TAB_PQA) {
            // Coulomb: r dU/dr = -qq exp(-k^2 r^2) (2 k / sqrt(pi) + P(t) / r);
            // vdW: P r^-6 - Q r^-3 - A12 r^-12 with the region's coefficients
            const CT r = r2 * rinv;
            const CT e2 = __builtin_amdgcn_exp2f(-P.k2log2e * r2);
            const CT t = __builtin_amdgcn_rcpf(fmaf(P.erfc_p, r, 1.0f));
            CT p = P.erfc_c[8];
#pragma unroll
            for (int c = 7; c >= 0; c--) p = fmaf(p, t, P.erfc_c[c]);
            const CT w = fmaf(p, rinv, P.two_kappa_sqrtpi);
            const CT rinv3 = rinv2 * rinv;
            const CT rinv6 = rinv3 * rinv3;
            const CT rinv12 = rinv6 * rinv6;
            const float4 c = *reinterpret_cast<const float4 *>(
                reinterpret_cast<const char *>(sh_pqa) + rowi[k] + tjb + (r2 > P.ron2 ? 16 : 0));
            CT fv = c.x * rinv6;
            fv = fmaf(-c.y, rinv3, fv);
            fv = fmaf(-c.z, rinv12, fv);
            fij = fmaf(-(qq * e2), w, fv) * rinv2;
          } else {
            const CT r = r2 * rinv;
            CT u_el;
            const CT f_el = ewald_pair(r2, r, rinv, qq, P, u_el);
            (void)u_el;
            T2 c;
            if (TAB == TAB_SQUARE) {
              c = sh_tab[rowi[k] + tj];
            } else if (TAB == TAB_COMBINE) {
              const float R = rhi[k] + ljj.y, R2 = R * R, R6 = R2 * R2 * R2;
              const float c6 = 12.0f * sei[k] * ljj.x * R6;
              c.x = (CT)c6;
              c.y = (CT)(c6 * R6);
            } else {
              const int ti = rowi[k];
              const int aa = ti > tj ? ti : tj, bb = ti > tj ? tj : ti;
              c = TAB == TAB_TRI ? sh_tab[aa * (aa + 1) / 2 + bb] : vdwtab[aa * (aa + 1) / 2 + bb];
            }
            const CT rinv3 = rinv * rinv2;
            const CT rinv6 = rinv3 * rinv3;
            // force switch: in the switched region (r > r_on) the c6/c12
            // terms are k6 c6 r^-3 (r^-3 - r_off^-3) and k12 c12 r^-6 (r^-6 -
            // r_off^-6); selects between per-lane values and an fma with a -1/0
            // factor avoid copying the switch constants into registers
            const bool outer = r2 > P.ron2;
            const CT mo = outer ? (CT)-1 : (CT)0;
            const CT t6 = fma(mo, P.roffinv3, rinv3), t12 = fma(mo, P.roffinv6, rinv6);
            const CT c6k = c.x * P.k6, c12k = c.y * P.k12;
            const CT c6 = outer ? c6k : c.x, c12 = outer ? c12k : c.y;
            const CT f_vdw = c6 * t6 * rinv3 - c12 * t12 * rinv6;
            fij = (f_vdw + f_el) * rinv2;
          }
          // ISO: gradients in A^-1 x unit, scaled when flushed
          fjx = fma(-fij, dx, fjx);
          fjy = fma(-fij, dy, fjy);
          fjz = fma(-fij, dz, fjz);
          fi[3 * k] = fma(fij, dx, fi[3 * k]);
          fi[3 * k + 1] = fma(fij, dy, fi[3 * k + 1]);
          fi[3 * k + 2] = fma(fij, dz, fi[3 * k + 2]);
          if (CALC_E) {
            // energies of energy steps in FP64 from the same fixed-point
            // geometry: a pair energy is ~1e2 larger than the sums it
            // cancels into, and FP32 pair arithmetic (~2e-7 relative) would
            // dominate the energy error
            const double ex = (double)(int)(pix[k] - pj.x) * PE.sx;
            const double ey = (double)(int)(piy[k] - pj.y) * PE.sy;
            const double ez = (double)(int)(piz[k] - pj.z) * PE.sz;
            const double er2 = ex * ex + ey * ey + ez * ez;
            const double er = sqrt(er2), erinv = 1.0 / er;
            el += (double)qi[k] * (double)charge_of<CT>(pj.w) * kCcelec * erfc(PE.kappa * er) *
                  erinv;
            const int ti = TAB == TAB_SQUARE ? rowi[k] / ntypes
                                             : (TAB == TAB_PQA ? rowi[k] / (ntypes * 32) : rowi[k]);
            const int aa = ti > tj ? ti : tj, bb = ti > tj ? tj : ti;
            const double2 ce = vdwtab_e[aa * (aa + 1) / 2 + bb];
            const double er3 = erinv * erinv * erinv, er6 = er3 * er3;
            const bool eo = er2 > PE.ron2;
            const double e6 = eo ? er3 - PE.roffinv3 : er3, e12 = eo ? er6 - PE.roffinv6 : er6;
            const double A6 = eo ? PE.k6 : 1.0, A12 = eo ? PE.k12 : 1.0;
            const double C6 = eo ? 0.0 : PE.dv6, C12 = eo ? 0.0 : PE.dv12;
            ev += ce.y * (1.0 / 12.0) * (A12 * e12 * e12 + C12) -
                  ce.x * (1.0 / 6.0) * (A6 * e6 * e6 + C6);
          }
          if (CALC_V) {
            const CT gx = fij * dx, gy = fij * dy, gz = fij * dz;
            v0 -= gx * dx;
            v1 -= gx * dy;
            v2 -= gx * dz;
            v3 -= gy * dy;
            v4 -= gy * dz;
            v5 -= gz * dz;
          }
        }
      }
      if (PRUNE) {
        const bool live = e0 + lane < it.z && mp != (1u << kGroup) - 1u;
        const unsigned long long b = __ballot(live);
        if (live)
          ent_out[wpos + __popcll(b & ((1ull << lane) - 1ull))] = (e & 0xffffu) | (mp << 16);
        wpos += __popcll(b);
      }
      // j-forces that stayed zero (no pair inside r_off) cost no memory traffic
      if (fjx != (CT)0 || fjy != (CT)0 || fjz != (CT)0) {
        atomicAdd(&fsort[jl], fx_round(ISO ? fjx * P.sx : fjx));
        atomicAdd(&fsort[ss + jl], fx_round(ISO ? fjy * P.sx : fjy));
        atomicAdd(&fsort[2 * (size_t)ss + jl], fx_round(ISO ? fjz * P.sx : fjz));
      }
      if (CALC_E) {
        acc[0] += el;
        acc[1] += ev;
      }
      if (CALC_V) {
        // ISO: the products are in unit^2 x (energy / A^2)
        const double vs = ISO ? (double)P.s2 : 1.0;
        acc[2] += (double)v0 * vs;
        acc[3] += (double)v1 * vs;
        acc[4] += (double)v2 * vs;
        acc[5] += (double)v3 * vs;
        acc[6] += (double)v4 * vs;
        acc[7] += (double)v5 * vs;
      }
    }
    if (PRUNE) {
      // null entries (every pair skipped) after the survivors, whole steps kept
      const int end = it.y + (wpos - it.y + 63) / 64 * 64;
      for (int p = wpos + lane; p < it.z; p += 64) ent_out[p] = ((1u << kGroup) - 1u) << 16;
      if (lane == 0) items_out[item] = make_int4(it.x, it.y, end, it.w);
    }
    group_reduce<CT>(fi, lane);
    if ((lane & 3) == 0) {
      const int il = ibase + (lane >> 2);
      if (fi[0] != (CT)0 || fi[1] != (CT)0 || fi[2] != (CT)0) {
        atomicAdd(&fsort[il], fx_round(ISO ? fi[0] * P.sx : fi[0]));
        atomicAdd(&fsort[ss + il], fx_round(ISO ? fi[1] * P.sx : fi[1]));
        atomicAdd(&fsort[2 * (size_t)ss + il], fx_round(ISO ? fi[2] * P.sx : fi[2]));
      }
    }
  }
  if (CALC_E || CALC_V) {
#pragma unroll
    for (int k = 0; k < kPairPart; k++) {
      double v = acc[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      acc[k] = v;
    }
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < kPairPart; k++) part[bid * kPairPart + k] = acc[k];
  }
}

template <typename CT, bool CALC_E, bool CALC_V, int TAB, bool ISO, bool PRUNE>
__global__ __launch_bounds__(64) void k_pair(
    const int4 *__restrict__ posq, const int *__restrict__ types, const int4 *__restrict__ items,
    const int *__restrict__ iorder, const uint32_t *__restrict__ ent, const int *__restrict__ cand_j, const int *__restrict__ meta,
    const typename Vec2<CT>::T *__restrict__ vdwtab, const typename Vec2<CT>::T *__restrict__ vdwsq,
    const float2 *__restrict__ ljtype, const float4 *__restrict__ vdwpqa, int ntypes,
    const PairConsts<CT> P, CT qfac, unsigned long long *__restrict__ fsort, int ss,
    double *__restrict__ part, const PairConsts<double> PE, const double2 *__restrict__ vdwtab_e,
    uint32_t *__restrict__ ent_out, int4 *__restrict__ items_out, CT rp2t) {
  // Register floor of 112 VGPRs (an empty asm statement that names v111): a
  // SIMD then holds at most four pair waves (448 of its 512 VGPRs) and keeps
  // room for one wave of the PME passes (44-54 VGPRs), which otherwise waited
  // for a pair wave to end (measured on the 98,304-atom box: z pass with the
  // convolution 102 -> 86 us inside the step, step 705 -> 714 ns/day).  The
  // kernel's own allocation is 92-107 VGPRs, so no spill follows.
  asm volatile("" ::: "v111");
  pair_body<CT, CALC_E, CALC_V, TAB, ISO, PRUNE>(blockIdx.x, gridDim.x, posq, types, items, iorder, ent,
                                                 cand_j, meta, vdwtab, vdwsq, ljtype, vdwpqa,
                                                 ntypes, P, qfac, fsort, ss, part, PE, vdwtab_e,
                                                 ent_out, items_out, rp2t);
}

// Fused launch over several systems: mixed precision, no energy/virial (the
// MD step).
template <int TAB, bool ISO>
__global__ __launch_bounds__(64) void kb_pair(const PairArgs *__restrict__ tab,
                                              const int *__restrict__ bstart, int nsys) {
  const int s = batch_system(bstart, nsys);
  const PairArgs &a = tab[s];
  const PairConsts<double> no_energy{};
  pair_body<float, false, false, TAB, ISO, false>(
      blockIdx.x - bstart[s], bstart[s + 1] - bstart[s], a.posq, a.types, a.items, a.iorder, a.ent,
      a.cand_j, a.meta, a.vdwtab, a.vdwsq, a.ljtype, a.vdwpqa, a.ntypes, a.P, a.qfac, a.fsort,
      a.ss, nullptr, no_energy, nullptr, nullptr, nullptr, 0.0f);
}

struct PairLaunch {
  const int4 *posq;
  const int *types;
  const int4 *items;
  const int *iorder;
  const uint32_t *ent;
  const int *cand_j, *meta;
  const float2 *ljtype;
  const float4 *vdwpqa;
  int ntypes;
  unsigned long long *fsort;
  int ss;
  double *part;
  const double2 *vdwtab_e;
  int blocks;
  size_t lds;
  hipStream_t s;
  // fused pruning: the pruned copy written, and the prune radius squared in
  // the units the kernel compares r^2 in
  uint32_t *ent_out;
  int4 *items_out;
  float rp2t;
};

template <typename CT, bool E, bool V, int TAB, bool ISO>
static void go_pair(const PairLaunch &L, const typename Vec2<CT>::T *tab,
                    const typename Vec2<CT>::T *sq, const PairConsts<CT> &P, CT qfac,
                    const PairConsts<double> &PE) {
  if constexpr (sizeof(CT) == sizeof(float) && !E && !V) {
    if (L.ent_out) {
      k_pair<CT, false, false, TAB, ISO, true><<<L.blocks, 64, L.lds, L.s>>>(
          L.posq, L.types, L.items, L.iorder, L.ent, L.cand_j, L.meta, tab, sq, L.ljtype, L.vdwpqa,
          L.ntypes, P, qfac, L.fsort, L.ss, L.part, PE, L.vdwtab_e, L.ent_out, L.items_out,
          L.rp2t);
      return;
    }
  }
  k_pair<CT, E, V, TAB, ISO, false><<<L.blocks, 64, L.lds, L.s>>>(
      L.posq, L.types, L.items, L.iorder, L.ent, L.cand_j, L.meta, tab, sq, L.ljtype, L.vdwpqa,
      L.ntypes, P,
      qfac, L.fsort, L.ss, L.part, PE, L.vdwtab_e, nullptr, nullptr, (CT)0);
}
template <typename CT, bool E, bool V, bool ISO>
static void go_pair_tab(int mode, const PairLaunch &L, const typename Vec2<CT>::T *tab,
                        const typename Vec2<CT>::T *sq, const PairConsts<CT> &P, CT qfac,
                        const PairConsts<double> &PE) {
  if constexpr (sizeof(CT) == sizeof(float)) {
    if (mode == TAB_PQA) return go_pair<CT, E, V, TAB_PQA, ISO>(L, tab, sq, P, qfac, PE);
    if (mode == TAB_COMBINE) return go_pair<CT, E, V, TAB_COMBINE, ISO>(L, tab, sq, P, qfac, PE);
  }
  if (mode == TAB_SQUARE) return go_pair<CT, E, V, TAB_SQUARE, ISO>(L, tab, sq, P, qfac, PE);
  if (mode == TAB_TRI) return go_pair<CT, E, V, TAB_TRI, ISO>(L, tab, sq, P, qfac, PE);
  go_pair<CT, E, V, TAB_GLOBAL, ISO>(L, tab, sq, P, qfac, PE);
}
template <typename CT, bool E, bool V>
static void go_pair_iso(bool iso, int mode, const PairLaunch &L, const typename Vec2<CT>::T *tab,
                        const typename Vec2<CT>::T *sq, const PairConsts<CT> &P, CT qfac,
                        const PairConsts<double> &PE) {
  if constexpr (sizeof(CT) == sizeof(float))
    if (iso) return go_pair_tab<CT, E, V, true>(mode, L, tab, sq, P, qfac, PE);
  go_pair_tab<CT, E, V, false>(mode, L, tab, sq, P, qfac, PE);
}

// Whether the pair kernel may form distances in fixed-point units: a cubic
// box (one unit for x, y and z), mixed precision.
static bool pair_iso(const NbSystem &S) {
  return S.precision == 0 && S.box[0] == S.box[1] && S.box[1] == S.box[2];
}

// pruned: read the pruned copy; prune_out: read the build's list and write
// the pruned copy (mixed-precision MD steps only)
template <typename CT>
static void launch_pair_t(NbSystem &S, const typename Vec2<CT>::T *tab,
                          const typename Vec2<CT>::T *sq, const PairConsts<CT> &P, CT qfac, bool E,
                          bool V, hipStream_t s, bool pruned, bool prune_out) {
  const int mode = pair_tab_mode<typename Vec2<CT>::T>(S.ntypes, S.lj_combine);
  pruned = pruned && !prune_out;
  const bool iso = pair_iso(S);
  PairLaunch L{S.posq.get(),
               S.type_s.get(),
               pruned ? S.items_p.get() : S.items.get(),
               S.item_order.get(),
               pruned ? S.ent_p.get() : S.ent.get(),
               S.cand_j.get(),
               S.meta.get(),
               S.ljtype.get(),
               S.vdwpqa.get(),
               S.ntypes,
               reinterpret_cast<unsigned long long *>(S.fsort.get()),
               S.ssort,
               S.parts_ptr() + S.part.pair_off,
               S.vdwtab_d.get(),
               S.pair_blocks,
               pair_tab_lds_bytes<typename Vec2<CT>::T>(mode, S.ntypes),
               s,
               prune_out ? S.ent_p.get() : nullptr,
               prune_out ? S.items_p.get() : nullptr,
               (float)(iso ? (double)S.prune_r2 / ((S.box[0] / 4294967296.0) *
                                                   (S.box[0] / 4294967296.0))
                           : (double)S.prune_r2)};
  if (E && V)
    go_pair_iso<CT, true, true>(iso, mode, L, tab, sq, P, qfac, S.pc_d);
  else if (E)
    go_pair_iso<CT, true, false>(iso, mode, L, tab, sq, P, qfac, S.pc_d);
  else if (V)
    go_pair_iso<CT, false, true>(iso, mode, L, tab, sq, P, qfac, S.pc_d);
  else
    go_pair_iso<CT, false, false>(iso, mode, L, tab, sq, P, qfac, S.pc_d);
  APO_LAUNCH_CHECK();
}

void direct_launch(NbSystem &S, bool E, bool V, hipStream_t s, bool pruned, bool prune_out) {
  if (prune_out && (S.precision != 0 || E || V))
    throw InvalidArgument("fused list pruning runs in mixed-precision MD steps only");
  if (S.precision == 1)
    launch_pair_t<double>(S, S.vdwtab_d.get(), S.vdwsq_d.get(), S.pc_d, kCcelec, E, V, s, pruned,
                          false);
  else
    launch_pair_t<float>(S, S.vdwtab.get(), S.vdwsq.get(), S.pc_f, (float)kCcelec, E, V, s,
                         pruned, prune_out);
}

// Sorted -> global order (addSortedForce, src/CudaPMEDirectForce.cu:1375-1386).
// Each atom has exactly one sorted slot, so this is a plain add.
__device__ __forceinline__ void finalize_body(int bid, const int64_t *__restrict__ fsort, int ss,
                                              const int *__restrict__ loc2glo,
                                              const int *__restrict__ meta,
                                              int64_t *__restrict__ force, int stride) {
  int loc = bid * blockDim.x + threadIdx.x;
  if (loc >= meta[0] * kCell) return;
  int g = loc2glo[loc];
  if (g < 0) return;
  force[g] += fsort[loc];
  force[(size_t)stride + g] += fsort[ss + loc];
  force[2 * (size_t)stride + g] += fsort[2 * (size_t)ss + loc];
}
__global__ void k_direct_finalize(const int64_t *__restrict__ fsort, int ss,
                                  const int *__restrict__ loc2glo, const int *__restrict__ meta,
                                  int64_t *__restrict__ force, int stride) {
  finalize_body(blockIdx.x, fsort, ss, loc2glo, meta, force, stride);
}

void direct_finalize(NbSystem &S, int64_t *force, int stride, hipStream_t s) {
  k_direct_finalize<<<(S.ssort + 255) / 256, 256, 0, s>>>(S.fsort.get(), S.ssort, S.loc2glo.get(),
                                                          S.meta.get(), force, stride);
  APO_LAUNCH_CHECK();
}

// ---- fused launches over several systems ------------------------------------
void batch_add_pair(BatchLaunch<PairArgs> &b, NbSystem &S, int blocks) {
  if (S.precision != 0) throw InvalidArgument("fused batches run the mixed-precision path");
  PairArgs a{S.posq.get(), S.type_s.get(), S.items.get(), S.item_order.get(), S.ent.get(),
             S.cand_j.get(),
             S.meta.get(), S.vdwtab.get(), S.vdwsq.get(), S.ljtype.get(), S.vdwpqa.get(),
             S.ntypes, S.pc_f, (float)kCcelec,
             reinterpret_cast<unsigned long long *>(S.fsort.get()), S.ssort};
  // LDS of every LDS layout: the launch uses the one every system allows
  b.add(a, blocks, std::max({pair_tab_lds_bytes<float2>(TAB_SQUARE, S.ntypes),
                             pair_tab_lds_bytes<float2>(TAB_TRI, S.ntypes),
                             S.ntypes <= kMaxTypesPqa
                                 ? pair_tab_lds_bytes<float2>(TAB_PQA, S.ntypes)
                                 : (size_t)0}));
}
template <bool ISO>
static void batch_pair_t(BatchLaunch<PairArgs> &b, int mode, hipStream_t s) {
  if (mode == TAB_PQA)
    kb_pair<TAB_PQA, ISO><<<b.blocks(), 64, b.lds, s>>>(b.d_args.get(), b.d_start.get(), b.nsys());
  else if (mode == TAB_SQUARE)
    kb_pair<TAB_SQUARE, ISO><<<b.blocks(), 64, b.lds, s>>>(b.d_args.get(), b.d_start.get(),
                                                           b.nsys());
  else if (mode == TAB_TRI)
    kb_pair<TAB_TRI, ISO><<<b.blocks(), 64, b.lds, s>>>(b.d_args.get(), b.d_start.get(), b.nsys());
  else if (mode == TAB_COMBINE)
    kb_pair<TAB_COMBINE, ISO><<<b.blocks(), 64, 0, s>>>(b.d_args.get(), b.d_start.get(), b.nsys());
  else
    kb_pair<TAB_GLOBAL, ISO><<<b.blocks(), 64, 0, s>>>(b.d_args.get(), b.d_start.get(), b.nsys());
}
// mode: the table layout every system of the batch allows (pair_tab_mode);
// iso: every system's box is cubic
void batch_pair(BatchLaunch<PairArgs> &b, int mode, bool iso, hipStream_t s) {
  if (!b.blocks()) return;
  if (iso)
    batch_pair_t<true>(b, mode, s);
  else
    batch_pair_t<false>(b, mode, s);
  APO_LAUNCH_CHECK();
}
int pair_tab_lds(const NbSystem &S) { return pair_tab_mode<float2>(S.ntypes, S.lj_combine); }
bool pair_tab_iso(const NbSystem &S) { return pair_iso(S); }

}  // namespace apo
