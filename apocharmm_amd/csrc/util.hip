// Small device utilities: fills, single-block scans, fixed-point conversion,
// self energy and the final fixed-order energy/virial reduction.
#include <algorithm>
#include <set>

#include "apocharmm_amd.h"
#include "system.hpp"

namespace apo {

// Fills (kernels instead of hipMemsetAsync so the step stays graph-safe).
__global__ void k_fill_i64(int64_t *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
static int fill_blocks(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 4096); }
__global__ void k_fill_f64(double *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

__global__ void k_fill_i32(int *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
void launch_fill_i32(int *p, size_t n, hipStream_t s) {
  if (!n) return;
  int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
  k_fill_i32<<<nb, 256, 0, s>>>(p, n);
  APO_LAUNCH_CHECK();
}

void launch_fill_i64(int64_t *p, size_t n, hipStream_t s) {
  if (!n) return;
  int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
  k_fill_i64<<<nb, 256, 0, s>>>(p, n);
  APO_LAUNCH_CHECK();
}
void launch_fill_f64(double *p, size_t n, hipStream_t s) {
  if (!n) return;
  int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
  k_fill_f64<<<nb, 256, 0, s>>>(p, n);
  APO_LAUNCH_CHECK();
}

// Exclusive scan of in[0..n) into out[0..n], out[n] = total, one block per
// array; the total also goes to *total_dst when given.  n = *n_dev (clamped
// to n_max) when n_dev is given, else n_max.  Each thread scans 8
// consecutive elements in registers, the waves' sums are scanned across the
// block, and chunks of 8192 elements follow each other with a carry.
struct ScanArgs {
  const int *in;
  int *out;
  const int *n_dev;
  int n_max;
  int *total_dst;
};
constexpr int kScanThreads = 1024, kScanPer = 8;
__global__ __launch_bounds__(kScanThreads) void k_scan(ScanArgs a0, ScanArgs a1) {
  const ScanArgs &a = blockIdx.x == 0 ? a0 : a1;
  __shared__ int wsum[kScanThreads / 64];
  const int n = a.n_dev ? min(*a.n_dev, a.n_max) : a.n_max;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int carry = 0;
  for (int base = 0; base < n; base += kScanThreads * kScanPer) {
    const int i0 = base + tid * kScanPer;
    int v[kScanPer];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
      v[k] = i0 + k < n ? a.in[i0 + k] : 0;
      sum += v[k];
    }
    int x = sum;  // inclusive scan of the thread sums within the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (wave == 0) {
      int w = lane < kScanThreads / 64 ? wsum[lane] : 0;
#pragma unroll
      for (int off = 1; off < kScanThreads / 64; off <<= 1) {
        const int y = __shfl_up(w, off, 64);
        if (lane >= off) w += y;
      }
      if (lane < kScanThreads / 64) wsum[lane] = w;
    }
    __syncthreads();
    int e = carry + x - sum + (wave > 0 ? wsum[wave - 1] : 0);
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
      if (i0 + k < n) a.out[i0 + k] = e;
      e += v[k];
    }
    carry += wsum[kScanThreads / 64 - 1];
    __syncthreads();  // wsum is rewritten by the next chunk
  }
  if (tid == 0) {
    a.out[n] = carry;
    if (a.total_dst) *a.total_dst = carry;
  }
}

void launch_scan(const int *in, int *out, const int *n_dev, int n_max, hipStream_t s,
                 int *total_dst) {
  const ScanArgs a{in, out, n_dev, n_max, total_dst};
  k_scan<<<1, kScanThreads, 0, s>>>(a, a);
  APO_LAUNCH_CHECK();
}
void launch_scan2(const int *in0, int *out0, const int *n0_dev, int n0_max, int *total0,
                  const int *in1, int *out1, const int *n1_dev, int n1_max, int *total1,
                  hipStream_t s) {
  k_scan<<<2, kScanThreads, 0, s>>>(ScanArgs{in0, out0, n0_dev, n0_max, total0},
                                    ScanArgs{in1, out1, n1_dev, n1_max, total1});
  APO_LAUNCH_CHECK();
}

// Force<long long> -> double (src/reduce.cu:132-143).
__global__ void k_force_to_double(const int64_t *f, int natom, int stride, double *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natom) return;
  for (int c = 0; c < 3; c++)
    out[(size_t)c * stride + i] = (double)f[(size_t)c * stride + i] * kInvForceScale;
}

void launch_force_to_double(const int64_t *f, int natom, int stride, double *out,
                            hipStream_t s) {
  k_force_to_double<<<(natom + 255) / 256, 256, 0, s>>>(f, natom, stride, out);
  APO_LAUNCH_CHECK();
}

// Plugin gradient into the total (add_force, src/reduce.cu:177-189).
__global__ void k_add_force(const int64_t *src, int natom, int stride, int64_t *dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natom) return;
  for (int c = 0; c < 3; c++) dst[(size_t)c * stride + i] += src[(size_t)c * stride + i];
}
void launch_add_force(const int64_t *src, int natom, int stride, int64_t *dst, hipStream_t s) {
  k_add_force<<<(natom + 255) / 256, 256, 0, s>>>(src, natom, stride, dst);
  APO_LAUNCH_CHECK();
}
// Plugin energy term 0 into the potential (src/ForceManager.cu:774-778).
__global__ void k_add_energy(const double *src, double *energy) {
  if (threadIdx.x == 0) energy[APO_E_EPOT] += src[0];
}
void launch_add_energy(const double *src, double *energy, hipStream_t s) {
  k_add_energy<<<1, 64, 0, s>>>(src, energy);
  APO_LAUNCH_CHECK();
}

// Block-level fixed-order double reduction (tree over LDS).
template <int NT>
__device__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = NT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// Sum of q^2 per block (Ewald self energy, include/cuda_pme_selfe.h:14-53).
__global__ __launch_bounds__(256) void k_self(const float4 *xyzq, int n, double *part) {
  __shared__ double sh[256];
  int i = blockIdx.x * 256 + threadIdx.x;
  double q = 0.0;
  if (i < n) q = (double)xyzq[i].w;
  double s = block_sum<256>(q * q, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

void self_launch(NbSystem &S, const float4 *x, hipStream_t s) {
  k_self<<<S.part.self_n, 256, 0, s>>>(x, S.natom, S.parts_ptr() + S.part.self_off);
  APO_LAUNCH_CHECK();
}

// Final reduction: every component is summed over its block partials in a
// fixed order (deterministic bits run to run).
struct ReduceArgs {
  PartLayout L;
  int terms;
  int E, V;
  int acc;  // add into energy / virial instead of overwriting
  double self_fac;
};

__global__ __launch_bounds__(256) void k_reduce_energy(const double *parts, ReduceArgs a,
                                                        double *energy, double *virial) {
  __shared__ double sh[256];
  const PartLayout &L = a.L;
  auto region_sum = [&](int off, int nblk, int width, int comp) -> double {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += parts[off + b * width + comp];
    return block_sum<256>(s, sh);
  };
  bool dir = a.terms & APO_TERM_DIRECT, p14 = a.terms & APO_TERM_PAIRS14,
       bnd = a.terms & APO_TERM_BONDED, rec = a.terms & APO_TERM_RECIP,
       slf = a.terms & APO_TERM_SELF;
  if (a.E) {
    double e[APO_NUM_ENERGY] = {0};
    for (int k = 0; k < 5; k++)
      e[k] = bnd ? region_sum(L.bond_off, L.bond_n, kBondPart, k) : 0.0;
    e[APO_E_EWKS] = rec ? region_sum(L.recip_off, L.recip_n, kRecipPart, 0) : 0.0;
    e[APO_E_EWSE] = slf ? a.self_fac * region_sum(L.self_off, L.self_n, 1, 0) : 0.0;
    e[APO_E_EWEX] = p14 ? region_sum(L.p14_off, L.p14_n, kP14Part, 2) : 0.0;
    double el = 0, vd = 0;
    if (dir) {
      el += region_sum(L.pair_off, L.pair_n, kPairPart, 0);
      vd += region_sum(L.pair_off, L.pair_n, kPairPart, 1);
    }
    if (p14) {
      el += region_sum(L.p14_off, L.p14_n, kP14Part, 0);
      vd += region_sum(L.p14_off, L.p14_n, kP14Part, 1);
    }
    e[APO_E_ELEC] = el;
    e[APO_E_VDW] = vd;
    double tot = 0;
    for (int k = 0; k < APO_E_EPOT; k++) tot += e[k];
    e[APO_E_EPOT] = tot;
    if (threadIdx.x == 0)
      for (int k = 0; k < APO_NUM_ENERGY; k++) energy[k] = a.acc ? energy[k] + e[k] : e[k];
  }
  if (a.V) {
    double v6[6] = {0};
    for (int c = 0; c < 6; c++) {
      double s = 0;
      if (dir) s += region_sum(L.pair_off, L.pair_n, kPairPart, 2 + c);
      if (p14) s += region_sum(L.p14_off, L.p14_n, kP14Part, 3 + c);
      if (bnd) s += region_sum(L.bond_off, L.bond_n, kBondPart, 5 + c);
      if (rec) s += region_sum(L.recip_off, L.recip_n, kRecipPart, 1 + c);
      v6[c] = s;
    }
    if (threadIdx.x == 0) {
      const double w[9] = {v6[0], v6[1], v6[2], v6[1], v6[3], v6[4], v6[2], v6[4], v6[5]};
      for (int k = 0; k < 9; k++) virial[k] = a.acc ? virial[k] + w[k] : w[k];
    }
  }
}

void reduce_energy_launch(NbSystem &S, int terms, bool E, bool V, double *energy,
                          double *virial, hipStream_t s, bool accumulate) {
  ReduceArgs a;
  a.L = S.part;
  a.terms = terms;
  a.E = E;
  a.V = V;
  a.acc = accumulate ? 1 : 0;
  a.self_fac = -S.kappa * kCcelec / std::sqrt(kPi);
  k_reduce_energy<<<1, 256, 0, s>>>(S.parts_ptr(), a, energy, virial);
  APO_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Topology: 1-2/1-3/1-4 sets from the bond graph and the exclusion CSR /
// in14 / ex14 lists (CharmmPSF::buildTopologicalExclusions and
// getInclusionExclusionLists, src/CharmmPSF.cu:235-385).
// ---------------------------------------------------------------------------
void build_exclusions(int natom, int nbond, const int32_t *bonds,
                      std::vector<int32_t> &excl_ptr, std::vector<int32_t> &excl_idx,
                      std::vector<int32_t> &in14, std::vector<int32_t> &ex14) {
  std::vector<std::set<int>> c12(natom), c13(natom), c14(natom);
  for (int b = 0; b < nbond; b++) {
    int i = bonds[2 * b], j = bonds[2 * b + 1];
    c12[i].insert(j);
    c12[j].insert(i);
  }
  for (int b = 0; b < nbond; b++) {
    int i = bonds[2 * b], j = bonds[2 * b + 1];
    for (int a : c12[j])
      if (a != i) c13[i].insert(a);
    for (int a : c12[i])
      if (a != j) c13[j].insert(a);
  }
  for (int i = 0; i < natom; i++)
    for (int j : c13[i])
      for (int k : c12[j])
        if (!c12[i].count(k)) c14[i].insert(k);
  excl_ptr.assign(1, 0);
  excl_idx.clear();
  in14.clear();
  ex14.clear();
  for (int i = 0; i < natom; i++) {
    std::set<int> all;
    all.insert(c12[i].begin(), c12[i].end());
    all.insert(c13[i].begin(), c13[i].end());
    all.insert(c14[i].begin(), c14[i].end());
    all.erase(i);
    for (int j : all) excl_idx.push_back(j);
    excl_ptr.push_back((int32_t)excl_idx.size());
    for (int j : c14[i])
      if (j > i && !c12[i].count(j) && !c13[i].count(j)) {
        in14.push_back(i);
        in14.push_back(j);
      }
    std::set<int> ex(c12[i].begin(), c12[i].end());
    ex.insert(c13[i].begin(), c13[i].end());
    for (int j : ex)
      if (j > i) {
        ex14.push_back(i);
        ex14.push_back(j);
      }
  }
}

}  // namespace apo
