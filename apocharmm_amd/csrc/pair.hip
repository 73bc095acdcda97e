// Direct-space PME pair kernel: erfc Coulomb + CHARMM force-switched LJ over
// 32x32 cell tiles (physics of calcForceKernel, include/CudaDirectForce_util.h
// :16-333, with pair_elec_force EWALD src/CudaDirectForceKernels.cu:476-490
// and pair_vdw_force VFSW :263-286).
//
// gfx950 design (not the reference's warp-rotation scheme):
//  * one wave64 = one work item (an i-cell and <= 16 of its tiles); lanes
//    0-31 hold the j-atoms of tile A, lanes 32-63 those of tile B;
//  * the 32 i-atoms are visited in a fully unrolled loop, each one
//    wave-uniform (scalar loads), so the i-data costs no LDS or VALU traffic;
//  * i-forces live in 96 registers per lane and are reduced once per item by
//    a transposed butterfly; j-forces are flushed per tile;
//  * coordinates are 32-bit fixed point per box edge: the int difference
//    wraps to the minimum image exactly, then converts to float;
//  * forces accumulate in FP32 inside a tile and are rounded to int64 2^-40
//    fixed point per flush, so results are bitwise reproducible.
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

// Sums v[0..31] over the 64 lanes; lanes 2t and 2t+1 return the total of v[t].
template <typename CT> __device__ __forceinline__ CT transpose_reduce(CT (&v)[32], int lane) {
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const bool up = lane & 32;
    CT send = up ? v[k] : v[k + 16];
    CT keep = up ? v[k + 16] : v[k];
    v[k] = keep + __shfl_xor(send, 32);
  }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const bool up = lane & 16;
    CT send = up ? v[k] : v[k + 8];
    CT keep = up ? v[k + 8] : v[k];
    v[k] = keep + __shfl_xor(send, 16);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const bool up = lane & 8;
    CT send = up ? v[k] : v[k + 4];
    CT keep = up ? v[k + 4] : v[k];
    v[k] = keep + __shfl_xor(send, 8);
  }
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const bool up = lane & 4;
    CT send = up ? v[k] : v[k + 2];
    CT keep = up ? v[k + 2] : v[k];
    v[k] = keep + __shfl_xor(send, 4);
  }
  {
    const bool up = lane & 2;
    CT send = up ? v[0] : v[1];
    CT keep = up ? v[1] : v[0];
    v[0] = keep + __shfl_xor(send, 2);
  }
  return v[0] + __shfl_xor(v[0], 1);
}

constexpr int kMaxTypesLds = 32;  // square vdW table in LDS up to 32 types

// TAB_LDS: vdW table square (ntypes x ntypes) in LDS, row offset carried with
// the i-atom; otherwise the triangular table is read from global memory.
template <typename CT, bool CALC_E, bool CALC_V, bool TAB_LDS>
__global__ __launch_bounds__(64) void k_pair(
    const int4 *__restrict__ posq, const int *__restrict__ types, const int4 *__restrict__ items,
    const int *__restrict__ tile_j, const uint32_t *__restrict__ tile_mask,
    const int *__restrict__ meta, const typename Vec2<CT>::T *__restrict__ vdwtab,
    const typename Vec2<CT>::T *__restrict__ vdwsq, int ntypes, const PairConsts<CT> P, CT qfac,
    unsigned long long *__restrict__ fsort, int ss, double *__restrict__ part) {
  using T2 = typename Vec2<CT>::T;
  __shared__ T2 sh_tab[TAB_LDS ? kMaxTypesLds * kMaxTypesLds : 1];
  __shared__ int4 sh_i[kCell];
  __shared__ int sh_row[kCell];
  __shared__ int sh_tj[kItemTiles];
  if (TAB_LDS)
    for (int k = threadIdx.x; k < ntypes * ntypes; k += 64) sh_tab[k] = vdwsq[k];

  const int lane = threadIdx.x;
  const int half = lane >> 5, l = lane & 31;
  const int nitem = meta[3];
  double acc[kPairPart];
#pragma unroll
  for (int k = 0; k < kPairPart; k++) acc[k] = 0.0;

#if APO_PAIR_XCD_MAP
  // consecutive items (neighbouring i-cells, shared j-cells) on one XCD:
  // blocks are dispatched round-robin over the 8 XCDs
  const int nb8 = gridDim.x >> 3;
  const int first = (gridDim.x & 7) == 0 ? (blockIdx.x & 7) * nb8 + (blockIdx.x >> 3) : blockIdx.x;
#else
  const int first = blockIdx.x;
#endif
  for (int item = first; item < nitem; item += gridDim.x) {
    const int4 it = items[item];
    const int ibase = it.x * kCell;
    __syncthreads();  // previous item's readers are done with sh_i
    if (lane < kCell) {
      sh_i[lane] = posq[ibase + lane];
      sh_row[lane] = TAB_LDS ? types[ibase + lane] * ntypes : types[ibase + lane];
    }
    if (lane < kItemTiles && it.y + lane < it.z) sh_tj[lane] = tile_j[it.y + lane];
    __syncthreads();
    CT fix[32], fiy[32], fiz[32];
#pragma unroll
    for (int t = 0; t < 32; t++) fix[t] = fiy[t] = fiz[t] = (CT)0;

    // j-data of the next tile pair is loaded while the current one computes
    auto jcell_of = [&](int tp) {
      const int tile = tp + half;
      return tile < it.z ? sh_tj[tile - it.y] : it.x;
    };
    int jc_n = jcell_of(it.y);
    int4 pj_n = posq[jc_n * kCell + l];
    int tj_n = types[jc_n * kCell + l];
    uint32_t m_n = it.y + half < it.z ? tile_mask[(size_t)(it.y + half) * kCell + l] : 0xffffffffu;

    for (int tp = it.y; tp < it.z; tp += 2) {
      // keep the 32 i-atom LDS reads inside the tile loop (not hoisted into
      // 160 registers)
      asm volatile("" ::: "memory");
      const int tile = tp + half;
      const bool tv = tile < it.z;
      const int jc = jc_n;
      const int jl = jc * kCell + l;
      const int4 pj = pj_n;
      const int tj = tj_n;
      const uint32_t m = m_n;
      if (tp + 2 < it.z) {
        jc_n = jcell_of(tp + 2);
        pj_n = posq[jc_n * kCell + l];
        tj_n = types[jc_n * kCell + l];
        m_n = tile + 2 < it.z ? tile_mask[(size_t)(tile + 2) * kCell + l] : 0xffffffffu;
      }
      // i-atoms with no pair inside the list cutoff in either tile are
      // skipped with a scalar branch
      uint32_t act = ~m;
      for (int off = 32; off > 0; off >>= 1) act |= __shfl_xor(act, off);
      act = __builtin_amdgcn_readfirstlane(act);
      const CT qj = charge_of<CT>(pj.w) * qfac;
      CT fjx = 0, fjy = 0, fjz = 0;
      // energies accumulate per pair in FP64 (large cancelling sums); the
      // virial per tile in CT
      double el = 0, ev = 0;
      CT v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
#pragma unroll
      for (int t = 0; t < 32; t++) {
        if (!(act & (1u << t))) continue;
        const int4 pi = sh_i[t];
        const CT dx = (CT)(int)(pi.x - pj.x) * P.sx;
        const CT dy = (CT)(int)(pi.y - pj.y) * P.sy;
        const CT dz = (CT)(int)(pi.z - pj.z) * P.sz;
        const CT r2 = dx * dx + dy * dy + dz * dz;
        if (!(m & (1u << t)) && r2 < P.roff2) {
          const CT rinv = rsqrt_ct(r2);
          const CT r = r2 * rinv;
          const CT qq = charge_of<CT>(pi.w) * qj;
          CT u_el;
          const CT f_el = ewald_pair(r2, r, rinv, qq, P, u_el);
          T2 c;
          if (TAB_LDS) {
            c = sh_tab[sh_row[t] + tj];
          } else {
            const int ti = sh_row[t];
            const int aa = ti > tj ? ti : tj, bb = ti > tj ? tj : ti;
            c = vdwtab[aa * (aa + 1) / 2 + bb];
          }
          const CT rinv2 = rinv * rinv;
          const CT rinv3 = rinv * rinv2;
          const CT rinv6 = rinv3 * rinv3;
          const bool outer = r2 > P.ron2;
          const CT A6 = outer ? P.k6 : (CT)1, B6 = outer ? P.roffinv3 : (CT)0;
          const CT A12 = outer ? P.k12 : (CT)1, B12 = outer ? P.roffinv6 : (CT)0;
          const CT t6 = rinv3 - B6, t12 = rinv6 - B12;
          const CT f_vdw = c.x * A6 * t6 * rinv3 - c.y * A12 * t12 * rinv6;
          const CT fij = (f_vdw + f_el) * rinv2;
          const CT gx = fij * dx, gy = fij * dy, gz = fij * dz;
          fjx -= gx;
          fjy -= gy;
          fjz -= gz;
          fix[t] += gx;
          fiy[t] += gy;
          fiz[t] += gz;
          if (CALC_E) {
            const CT C6 = outer ? (CT)0 : P.dv6, C12 = outer ? (CT)0 : P.dv12;
            el += (double)u_el;
            ev += (double)(c.y * ((CT)1 / (CT)12) * (A12 * t12 * t12 + C12) -
                           c.x * ((CT)1 / (CT)6) * (A6 * t6 * t6 + C6));
          }
          if (CALC_V) {
            v0 -= gx * dx;
            v1 -= gx * dy;
            v2 -= gx * dz;
            v3 -= gy * dy;
            v4 -= gy * dz;
            v5 -= gz * dz;
          }
        }
      }
      if (tv && m != 0xffffffffu) {
        atomicAdd(&fsort[jl], fx_round(fjx));
        atomicAdd(&fsort[ss + jl], fx_round(fjy));
        atomicAdd(&fsort[2 * (size_t)ss + jl], fx_round(fjz));
      }
      if (CALC_E) {
        acc[0] += el;
        acc[1] += ev;
      }
      if (CALC_V) {
        acc[2] += (double)v0;
        acc[3] += (double)v1;
        acc[4] += (double)v2;
        acc[5] += (double)v3;
        acc[6] += (double)v4;
        acc[7] += (double)v5;
      }
    }
    const CT sx = transpose_reduce<CT>(fix, lane);
    const CT sy = transpose_reduce<CT>(fiy, lane);
    const CT sz = transpose_reduce<CT>(fiz, lane);
    if (!(lane & 1)) {
      const int il = ibase + (lane >> 1);
      if (sx != (CT)0 || sy != (CT)0 || sz != (CT)0) {
        atomicAdd(&fsort[il], fx_round(sx));
        atomicAdd(&fsort[ss + il], fx_round(sy));
        atomicAdd(&fsort[2 * (size_t)ss + il], fx_round(sz));
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
      for (int k = 0; k < kPairPart; k++) part[blockIdx.x * kPairPart + k] = acc[k];
  }
}

template <typename CT>
static void launch_pair_t(NbSystem &S, const typename Vec2<CT>::T *tab,
                          const typename Vec2<CT>::T *sq, const PairConsts<CT> &P, CT qfac, bool E,
                          bool V, hipStream_t s) {
  auto *f = reinterpret_cast<unsigned long long *>(S.fsort.get());
  double *part = S.parts.get() + S.part.pair_off;
  const int G = S.pair_blocks;
  const bool lds = S.ntypes <= kMaxTypesLds;
#define APO_PAIR_ARGS                                                                   \
  S.posq.get(), S.type_s.get(), S.items.get(), S.tile_j.get(), S.tile_mask.get(),     \
      S.meta.get(), tab, sq, S.ntypes, P, qfac, f, S.ssort, part
#define APO_PAIR_GO(EE, VV)                                                             \
  if (lds)                                                                              \
    k_pair<CT, EE, VV, true><<<G, 64, 0, s>>>(APO_PAIR_ARGS);                          \
  else                                                                                  \
    k_pair<CT, EE, VV, false><<<G, 64, 0, s>>>(APO_PAIR_ARGS);
  if (E && V) {
    APO_PAIR_GO(true, true)
  } else if (E) {
    APO_PAIR_GO(true, false)
  } else if (V) {
    APO_PAIR_GO(false, true)
  } else {
    APO_PAIR_GO(false, false)
  }
#undef APO_PAIR_GO
#undef APO_PAIR_ARGS
  APO_LAUNCH_CHECK();
}

void direct_launch(NbSystem &S, bool E, bool V, hipStream_t s) {
  if (S.precision == 1)
    launch_pair_t<double>(S, S.vdwtab_d.get(), S.vdwsq_d.get(), S.pc_d, kCcelec, E, V, s);
  else
    launch_pair_t<float>(S, S.vdwtab.get(), S.vdwsq.get(), S.pc_f, (float)kCcelec, E, V, s);
}

// Sorted -> global order (addSortedForce, src/CudaPMEDirectForce.cu:1375-1386).
// Each atom has exactly one sorted slot, so this is a plain add.
__global__ void k_direct_finalize(const int64_t *__restrict__ fsort, int ss,
                                  const int *__restrict__ loc2glo, const int *__restrict__ meta,
                                  int64_t *__restrict__ force, int stride) {
  int loc = blockIdx.x * blockDim.x + threadIdx.x;
  if (loc >= meta[0] * kCell) return;
  int g = loc2glo[loc];
  if (g < 0) return;
  force[g] += fsort[loc];
  force[(size_t)stride + g] += fsort[ss + loc];
  force[2 * (size_t)stride + g] += fsort[2 * (size_t)ss + loc];
}

void direct_finalize(NbSystem &S, int64_t *force, int stride, hipStream_t s) {
  k_direct_finalize<<<(S.ssort + 255) / 256, 256, 0, s>>>(S.fsort.get(), S.ssort, S.loc2glo.get(),
                                                          S.meta.get(), force, stride);
  APO_LAUNCH_CHECK();
}

}  // namespace apo
