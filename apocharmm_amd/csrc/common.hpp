// Shared definitions of the MI355X (gfx950) CHARMM force path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace apo {

// Gradient fixed point: value * 2^40 in int64 (include/gpu_utils.h:21-23).
constexpr double kForceScale = 1099511627776.0;  // 2^40
constexpr double kInvForceScale = 1.0 / 1099511627776.0;
// PME charge grid fixed point (int64, exact, order independent).
constexpr double kGridScale = 1099511627776.0;  // 2^40
constexpr double kCcelec = 332.0716;            // src/CudaDirectForceKernels.cu:59
constexpr double kPi = 3.14159265358979323846;
// Cell (tile edge) size; pair-kernel i-group size (slots of one cell) and
// 64-lane j steps per work item.
constexpr int kCell = 32;
constexpr int kGroup = 16;
constexpr int kGroupsPerCell = kCell / kGroup;
#ifndef APO_ITEM_STEPS
#define APO_ITEM_STEPS 16
#endif
constexpr int kItemSteps = APO_ITEM_STEPS;
// a j entry is (candidate tile index within its i-cell << 5 | slot in the
// j-cell) in the low 16 bits and the 16-bit skip mask in the high bits
constexpr int kMaxCellTiles = 2048;
// Types up to which the pair kernel reads the switched vdW coefficients
// (P, Q, A12) of both regions from an LDS table: 2 x 11^2 float4 = 3.9 KB,
// small enough for 32 one-wave blocks per CU.
constexpr int kMaxTypesPqa = 11;
// Pair-kernel grid: blocks (one wave each) per CU.  The kernel loops over
// work items, so this only sets how many waves each SIMD can interleave.
#ifndef APO_PAIR_BLOCKS_PER_CU
#define APO_PAIR_BLOCKS_PER_CU 32
#endif
constexpr int kPairBlocksPerCu = APO_PAIR_BLOCKS_PER_CU;
// PME spread + x FFT: slab of whole x rows, by x bz rows (y, z) per block.
#ifndef APO_SLAB_BY
#define APO_SLAB_BY 4
#endif
#ifndef APO_SLAB_BZ
#define APO_SLAB_BZ 4
#endif
constexpr int kSlabBy = APO_SLAB_BY;
constexpr int kSlabBz = APO_SLAB_BZ;
// FFT passes: real-row pairs per block (x inverse) and columns per block (y, z),
// one wave per block by default: while the pair kernel's one-wave blocks hold
// most wave slots, a one-wave block finds a free slot where a four-wave block
// waits for four on one CU (measured: the C2 MD step 0.338 -> 0.330 ms).
#ifndef APO_FFT_ROWS_PER_BLOCK
#define APO_FFT_ROWS_PER_BLOCK 4
#endif
#ifndef APO_FFT_COLS_PER_BLOCK
#define APO_FFT_COLS_PER_BLOCK 4
#endif
#ifndef APO_FFT_THREADS
#define APO_FFT_THREADS 64
#endif
constexpr int kFftRowsPerBlock = APO_FFT_ROWS_PER_BLOCK;
constexpr int kFftThreads = APO_FFT_THREADS;  // threads per block of the y, z and x-inverse passes
constexpr int kFftColsPerBlock = APO_FFT_COLS_PER_BLOCK;
// Wave issue priority (s_setprio, 0-3) the PME kernels take when the force
// system's pme_prio is set.  The PME chain runs beside the pair kernel, whose
// waves stay at 0: on a SIMD shared with pair waves, a ready PME wave issues
// first.  That shortens the step where the PME chain is the critical path
// (measured: 98,304-atom water box 0.310 -> 0.300 ms) and lengthens it where
// the pair kernel is (23,558-atom DHFR 0.158 -> 0.162 ms).  It orders issue
// only, so results do not change.
#ifndef APO_PME_PRIO
#define APO_PME_PRIO 3
#endif
__device__ __forceinline__ void pme_wave_priority() { __builtin_amdgcn_s_setprio(APO_PME_PRIO); }
// Largest xy column population the LDS column sort handles.
constexpr int kMaxColumn = 2048;
// Pair-kernel per-block partials: elec, vdw, virial xx xy xz yy yz zz.
constexpr int kPairPart = 8;
// 1-4/exclusion partials: elec, vdw, ewex, virial(6).
constexpr int kP14Part = 9;
// Bonded partials: bond, angle, ureyb, dihe, imdihe, virial(6).
constexpr int kBondPart = 11;
// Reciprocal partials: ewks, virial(6).
constexpr int kRecipPart = 7;

// ---- fused launches over several systems -------------------------------------
// A batched kernel's grid is the concatenation of the per-system grids:
// bstart[s] is the first block of system s (nsys + 1 entries).  The block
// finds its system (a wave-uniform binary search), loads that system's
// argument struct and runs the single-system body with its local block index.
__device__ __forceinline__ int batch_system(const int *__restrict__ bstart, int nsys) {
  const int b = blockIdx.x;
  int lo = 0, hi = nsys;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bstart[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct InvalidArgument : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct RuntimeError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct GpuError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define APO_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::apo::GpuError(std::string(#call) + ": " +                        \
                            hipGetErrorString(e_) + " (" __FILE__ ":" +        \
                            std::to_string(__LINE__) + ")");                   \
  } while (0)

#define APO_LAUNCH_CHECK() APO_HIP(hipGetLastError())

// Owning device buffer (no copies; moves allowed).
template <typename T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // (Re)allocates to hold at least `count` elements; contents are lost.
  void reserve(size_t count) {
    if (count <= n && p) return;
    release();
    size_t c = count ? count : 1;
    APO_HIP(hipMalloc(&p, c * sizeof(T)));
    n = c;
  }
  // Copies count elements from host memory, ordered on stream s; returns once
  // the copy has finished, so the host source may be reused or freed at once
  // (uploads happen at set-up and at table rebuilds, never per step).
  void upload(const T *h, size_t count, hipStream_t s) {
    reserve(count);
    if (count) {
      APO_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
      APO_HIP(hipStreamSynchronize(s));
    }
  }
  void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
  T *get() const { return p; }
};

// Host side of one fused launch: the per-system argument structs and block
// ranges, uploaded once per (re)capture of the batch step.
template <typename A> struct BatchLaunch {
  std::vector<A> args;
  std::vector<int> start{0};
  DevBuf<A> d_args;
  DevBuf<int> d_start;
  size_t lds = 0;
  void clear() {
    args.clear();
    start.assign(1, 0);
    lds = 0;
  }
  void add(const A &a, int nblocks, size_t lds_bytes = 0) {
    args.push_back(a);
    start.push_back(start.back() + nblocks);
    if (lds_bytes > lds) lds = lds_bytes;
  }
  int nsys() const { return (int)args.size(); }
  int blocks() const { return start.back(); }
  void upload(hipStream_t s) {
    d_args.upload(args, s);
    d_start.upload(start, s);
  }
};

// Pair-kernel constants (VFSW + Ewald), CT = float (mixed) or double (FP64).
// Switching constants follow src/CudaPMEDirectForce.cu:321-331.
template <typename CT> struct PairConsts {
  CT sx, sy, sz;          // box / 2^32: fixed-point coordinate unit
  // cubic boxes: the pair kernel forms r^2 in fixed-point units and converts
  // once (s2 = sx^2); roff2u = roff^2 in those units
  CT s2, roff2u;
  CT roff2, ron2, rlist2;
  CT k6, k12, dv6, dv12, roffinv3, roffinv6;
  CT kappa, kappa2, k2log2e, two_kappa_sqrtpi;
  CT erfc_p;              // t = 1/(1 + erfc_p * kappa r)
  CT erfc_c[9];           // erfc(x) ~ exp(-x^2) * sum_k c_k t^k
};
struct PairArgs {
  const int4 *posq;
  const int *types;
  const int4 *items;
  const int *iorder;  // processing order of the items
  const uint32_t *ent;
  const int *cand_j, *meta;
  const float2 *vdwtab, *vdwsq;
  const float2 *ljtype;  // per-type (sqrt(eps), Rmin/2) when the table follows the rule
  const float4 *vdwpqa;  // switched vdW coefficients (P, Q, A12) per type pair, few types
  int ntypes;
  PairConsts<float> P;
  float qfac;
  unsigned long long *fsort;
  int ss;
};

// Offsets of per-block partial regions in the partials buffer.
struct PartLayout {
  int pair_off = 0, pair_n = 0;
  int p14_off = 0, p14_n = 0;
  int bond_off = 0, bond_n = 0;
  int recip_off = 0, recip_n = 0;
  int self_off = 0, self_n = 0;
  int total = 0;
};

// ---- kernel launchers (defined in the .hip translation units) -------------
struct NbSystem;

void launch_fill_i64(int64_t *p, size_t n, hipStream_t s);

// Argument structs of the fused (multi-system) launches of the step kernels.
struct SortArgs {
  const float4 *xyzq;
  const int *vdwtype, *loc2glo, *meta;
  double3 inv_len;
  float3 unit;
  int4 *posq;
  int *type_s;
  int64_t *fsort;
  int ss;
  float4 *bb;
  int64_t *force;  // gradient to zero (nforce entries), or null
  size_t nforce;
};
void launch_fill_f64(double *p, size_t n, hipStream_t s);
void launch_fill_i32(int *p, size_t n, hipStream_t s);
// exclusive scans, out[n] = total (also to *total_dst when given); scan2 runs
// two independent scans in one launch
void launch_scan(const int *in, int *out, const int *n_dev, int n_max, hipStream_t s,
                 int *total_dst = nullptr);
void launch_scan2(const int *in0, int *out0, const int *n0_dev, int n0_max, int *total0,
                  const int *in1, int *out1, const int *n1_dev, int n1_max, int *total1,
                  hipStream_t s);
void launch_force_to_double(const int64_t *f, int natom, int stride,
                            double *out, hipStream_t s);
void launch_add_force(const int64_t *src, int natom, int stride, int64_t *dst, hipStream_t s);
void launch_add_energy(const double *src, double *energy, hipStream_t s);

void nblist_build(NbSystem &sys, const float4 *d_xyzq, hipStream_t s, bool sizing);
void list_export_check(NbSystem &sys, hipStream_t s, bool allow_growth);
void list_reserve(NbSystem &sys, int ncand, int nent, int nitem);
// Sorts the coordinates into cell order (and the cell boxes); when force is
// given, also zeroes its nforce entries.
void nblist_sort_coords(NbSystem &sys, const float4 *d_xyzq, hipStream_t s,
                        int64_t *force = nullptr, size_t nforce = 0);
void nblist_stats(NbSystem &sys, int64_t *out6, hipStream_t s);
// Prunes the list at radius sqrt(prune_r2) on the current sorted coordinates
// into ent_p / items_p.
void nblist_prune(NbSystem &sys, hipStream_t s);
// calc_force flags for internal callers (no public flag uses these bits): the
// pair kernel reads the pruned list (which a build in the same call, or
// kCalcPrune, refreshes after the coordinate sort).
constexpr int kCalcUsePruned = 1 << 27;
constexpr int kCalcPrune = 1 << 28;
// ... and leave the sorted pair gradient in fsort: the caller adds it (the
// integrator's step kernel does, as it reads each atom's gradient).
constexpr int kCalcNoFinalize = 1 << 26;

// pruned: the pair kernel reads the pruned list copy; prune_out: it reads the
// build's list and writes the pruned copy for the next steps (fused pruning,
// mixed-precision MD steps)
void direct_launch(NbSystem &sys, bool energy, bool virial, hipStream_t s, bool pruned = false,
                   bool prune_out = false);
void direct_finalize(NbSystem &sys, int64_t *d_force, int stride,
                     hipStream_t s);
void pairs14_launch(NbSystem &sys, const float4 *d_xyzq, int64_t *d_force,
                    int stride, bool energy, bool virial, hipStream_t s);
void bonded_launch(NbSystem &sys, const float4 *d_xyzq, int64_t *d_force,
                   int stride, bool energy, bool virial, hipStream_t s);
void recip_launch(NbSystem &sys, const float4 *d_xyzq, int64_t *d_force,
                  int stride, bool energy, bool virial, hipStream_t s);
void self_launch(NbSystem &sys, const float4 *d_xyzq, hipStream_t s);
void reduce_energy_launch(NbSystem &sys, int term_mask, bool energy,
                          bool virial, double *d_energy, double *d_virial,
                          hipStream_t s, bool accumulate = false);

// Host topology helper shared by the C ABI.
void build_exclusions(int natom, int nbond, const int32_t *bonds,
                      std::vector<int32_t> &excl_ptr,
                      std::vector<int32_t> &excl_idx,
                      std::vector<int32_t> &in14, std::vector<int32_t> &ex14);

}  // namespace apo
