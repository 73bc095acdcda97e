// Smooth PME reciprocal space on gfx950.
//
// Physics and grid conventions of CudaPMERecip (spread
// include/cuda_pme_spread.h:24-142, scalar_sum_ortho_kernel
// src/CudaPMERecip.cu:673-929, gather include/cuda_pme_gather.h:18-320,
// order-4 B-splines src/CudaPMERecip.cu:290-465):
//   fractional coordinate fr = K (w - (floor(w + 1/2) - 1/2)), w = x/L + 2,
//   grid point ix + t carries theta[t], x fastest in memory.
// gfx950 design:
//   * fr is formed in FP64 per atom (one conversion), B-splines in CT;
//   * the charge grid is int64 fixed point (2^-40): order-independent sums;
//   * the 3-D R2C/C2R FFT is five LDS passes of a hand-written mixed-radix
//     (4, 2, 3, 5, 7, 11, 13) Stockham FFT: x (two real rows packed per complex
//     transform, run by the spread kernel on the rows it owns), y, [z forward
//     -> k-space convolution + energy/virial -> z inverse] fused in one kernel,
//     y inverse, x inverse;
//   * gather uses 16 lanes per atom and writes int64 2^-40 gradients.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "system.hpp"

namespace apo {

template <typename CT> struct Cx;
template <> struct Cx<float> { using T = float2; };
template <> struct Cx<double> { using T = double2; };

// Division by a small positive divisor d as a high multiply: with
// m = ceil(2^32 / d), floor(x m / 2^32) = floor(x / d) for 0 <= x < 2^32 / d
// (the rounding excess x (m d - 2^32) / (d 2^32) stays below 1/d, and
// x / d sits at least 1/d below the next integer).  m = 0 stands for d = 1.
// Every index divided here stays below 2^20 and every divisor below 2^12.
static inline unsigned magic_of(int d) { return d > 1 ? 0xffffffffu / (unsigned)d + 1u : 0u; }
__device__ __forceinline__ int fdiv(int x, unsigned m) {
  return m ? (int)__umulhi((unsigned)x, m) : x;
}
__device__ __forceinline__ unsigned magic_dev(int d) {
  return d > 1 ? 0xffffffffu / (unsigned)d + 1u : 0u;
}

// Whether a transform length needs a radix above 5 (the FFT kernels' variant
// with the radix-7, 11 and 13 stages).
static bool plan_big(const FftDim &d) {
  for (int i = 0; i < d.nrad; i++)
    if (d.rad[i] > 5) return true;
  return false;
}

// Radix plan of one transform length, with the divisors each Stockham stage
// needs (butterflies per row, and the span Ns of the stage) as magic numbers.
struct FftPlan {
  int n, nrad;
  int rad[16];
  unsigned mbf[16], mns[16];
};

static FftPlan plan_of(const FftDim &d) {
  FftPlan p;
  p.n = d.n;
  p.nrad = d.nrad;
  int ns = 1;
  for (int i = 0; i < 16; i++) {
    p.rad[i] = d.rad[i];
    p.mbf[i] = p.mns[i] = 0;
    if (i < d.nrad) {
      p.mbf[i] = magic_of(d.n / d.rad[i]);
      p.mns[i] = magic_of(ns);
      ns *= d.rad[i];
    }
  }
  return p;
}

template <typename C2> __device__ __forceinline__ C2 cmul(C2 a, C2 b) {
  C2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C2> __device__ __forceinline__ C2 cadd(C2 a, C2 b) {
  C2 r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename C2> __device__ __forceinline__ C2 csub(C2 a, C2 b) {
  C2 r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
template <typename C2> __device__ __forceinline__ C2 twid(const C2 *tw, int e, bool inv) {
  C2 w = tw[e];
  if (inv) w.y = -w.y;
  return w;
}

// Small DFT of length R on registers (forward: e^{-2 pi i r q / R}).
template <typename C2, int R>
__device__ __forceinline__ void small_dft(C2 (&v)[R], const C2 *tw, int n, bool inv) {
  if constexpr (R == 2) {
    C2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    C2 a = v[0], b = v[1], c = v[2], d = v[3];
    C2 s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = csub(b, d);
    // -i*s3 (forward) or +i*s3 (inverse)
    C2 js3;
    if (inv) { js3.x = -s3.y; js3.y = s3.x; } else { js3.x = s3.y; js3.y = -s3.x; }
    v[0] = cadd(s0, s2);
    v[2] = csub(s0, s2);
    v[1] = cadd(s1, js3);
    v[3] = csub(s1, js3);
  } else {
    C2 o[R];
    const int st = n / R;
#pragma unroll
    for (int q = 0; q < R; q++) {
      C2 s = v[0];
#pragma unroll
      for (int r = 1; r < R; r++) s = cadd(s, cmul(v[r], twid(tw, ((r * q) % R) * st, inv)));
      o[q] = s;
    }
#pragma unroll
    for (int q = 0; q < R; q++) v[q] = o[q];
  }
}

// One Stockham autosort stage of radix R over nrow rows of length n.
// mbf, mns: magic numbers of n / R and Ns (fdiv).
template <typename C2, int R>
__device__ __forceinline__ void fft_stage(const C2 *in, C2 *out, int n, int nrow, int Ns,
                                          unsigned mbf, unsigned mns, const C2 *tw, bool inv) {
  const int nbf = n / R;
  const int step = n / (Ns * R);
  for (int w = threadIdx.x; w < nbf * nrow; w += blockDim.x) {
    const int row = fdiv(w, mbf), j = w - row * nbf;
    const C2 *ri = in + row * n;
    C2 *ro = out + row * n;
    const int jq = fdiv(j, mns), k = j - jq * Ns;
    C2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = ri[j + r * nbf];
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = cmul(v[r], twid(tw, r * k * step, inv));
    small_dft<C2, R>(v, tw, n, inv);
    const int idxD = jq * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) ro[idxD + r * Ns] = v[r];
  }
}

// nrow unnormalised FFTs of length p.n held in LDS buffer a (b = scratch).
// Returns the buffer holding the result.
// Always inlined: as a called function it makes every caller allocate the
// callee's worst-case register budget (262 VGPRs, one wave per SIMD).
// BIG: the plan may hold radices 7, 11 or 13.  Their generic DFT sets the
// register budget of every kernel that contains it (the y pass needs 94
// VGPRs with them, 44 without), so the launchers pick the variant without
// them whenever every radix of the plans is at most 5 (plan_big).
template <typename C2, bool BIG>
__device__ __forceinline__ C2 *lds_fft(C2 *a, C2 *b, const FftPlan &p, int nrow, const C2 *tw,
                                       bool inv) {
  int Ns = 1;
  for (int s = 0; s < p.nrad; s++) {
    const int R = p.rad[s];
    const unsigned mb = p.mbf[s], mn = p.mns[s];
    switch (R) {
      case 2: fft_stage<C2, 2>(a, b, p.n, nrow, Ns, mb, mn, tw, inv); break;
      case 3: fft_stage<C2, 3>(a, b, p.n, nrow, Ns, mb, mn, tw, inv); break;
      case 4: fft_stage<C2, 4>(a, b, p.n, nrow, Ns, mb, mn, tw, inv); break;
      case 5: fft_stage<C2, 5>(a, b, p.n, nrow, Ns, mb, mn, tw, inv); break;
      default:
        if constexpr (BIG) {
          if (R == 7) fft_stage<C2, 7>(a, b, p.n, nrow, Ns, mb, mn, tw, inv);
          else if (R == 11) fft_stage<C2, 11>(a, b, p.n, nrow, Ns, mb, mn, tw, inv);
          else fft_stage<C2, 13>(a, b, p.n, nrow, Ns, mb, mn, tw, inv);
        }
        break;
    }
    __syncthreads();
    C2 *t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// Block-wide copy of `total` elements, global -> LDS (or any element map):
// each thread issues up to kStage loads before it writes any of them, so
// their latencies overlap instead of each load waiting for the previous
// element's store.
constexpr int kStage = 8;
template <typename T, typename Load, typename Store>
__device__ __forceinline__ void staged_copy(int total, int tid, int nthreads, Load &&ld,
                                            Store &&st) {
  for (int e0 = tid; e0 < total; e0 += kStage * nthreads) {
    T v[kStage];
#pragma unroll
    for (int u = 0; u < kStage; u++) {
      const int e = e0 + u * nthreads;
      if (e < total) v[u] = ld(e);
    }
#pragma unroll
    for (int u = 0; u < kStage; u++) {
      const int e = e0 + u * nthreads;
      if (e < total) st(e, v[u]);
    }
  }
}

// ---------------------------------------------------------------------------
// B-splines and fractional coordinates
// ---------------------------------------------------------------------------
__device__ __forceinline__ void frac_coord(float x, double inv_len, int K, int &ix, double &w) {
  const double ww = (double)x * inv_len + 2.0;
  double fr = (double)K * (ww - (floor(ww + 0.5) - 0.5));
  // fr = K is the periodic image of 0; non-finite input maps to 0 so every
  // grid index stays in range
  if (!(fr >= 0.0 && fr < (double)K)) fr = 0.0;
  ix = (int)fr;
  w = fr - (double)ix;
}

// Order-4 Cardinal B-spline weights (and derivatives) at offset w.
template <typename CT>
__device__ __forceinline__ void bspline4(CT w, CT (&th)[4], CT (&dth)[4]) {
  th[3] = 0;
  th[1] = w;
  th[0] = (CT)1 - w;
  th[2] = (CT)0.5 * w * th[1];
  th[1] = (CT)0.5 * ((w + (CT)1) * th[0] + ((CT)2 - w) * th[1]);
  th[0] = (CT)0.5 * ((CT)1 - w) * th[0];
  dth[0] = -th[0];
  dth[1] = th[0] - th[1];
  dth[2] = th[1] - th[2];
  dth[3] = th[2] - th[3];
  const CT third = (CT)1 / (CT)3;
  th[3] = third * w * th[2];
  th[2] = third * ((w + (CT)1) * th[1] + ((CT)3 - w) * th[2]);
  th[1] = third * ((w + (CT)2) * th[0] + ((CT)2 - w) * th[1]);
  th[0] = third * ((CT)1 - w) * th[0];
}

struct GridGeom {
  int K0, K1, K2;
  double inv0, inv1, inv2;
};

// Spread fused with the forward x FFT (no global atomics, no real grid in
// HBM).  One block owns a slab of whole x rows, by x bz rows in (y, z), and
// accumulates in LDS int64 2^-40 fixed point every stencil contribution that
// lands in it, from the atoms of the cells whose (per-step) bounding box
// reaches the slab.  The integer sums are exact, so the charges are bitwise
// those of any order of accumulation, and no zero fill or HBM atomic is
// needed.  The block then converts its rows to CT, packs rows (y, y + 1) of one
// z plane into one complex transform, runs the x FFT in LDS and stores the
// half-complex rows the y pass reads.
//
// Work split: each lane prepares one candidate atom (grid base index and
// fractional offsets, FP64); the wave then walks the atoms that reach the
// slab one at a time, lane l adding the stencil point (l & 3, (l >> 2) & 3,
// l >> 4) of that atom, so the 64 LDS atomics of one instruction hit 64
// distinct points.
struct BrickGeom {
  int K0, K1, K2;
  int by, bz;    // rows per slab in y (even) and z
  int nby, nbz;  // slabs per axis
  float L0, L1, L2;
};

// LDS of one slab block, all of it dynamic so that the int64 rows start at
// offset 0 (64-bit LDS atomics need 8-byte alignment, and the dynamic region
// would otherwise begin right after any static arrays): int64 rows, twiddles,
// two complex FFT buffers, then the candidate cells and their count.
constexpr int kSlabCand = 1024;  // cells tested per round (a multiple of kSpreadThreads)
#ifndef APO_SPREAD_THREADS
#define APO_SPREAD_THREADS 256
#endif
constexpr int kSpreadThreads = APO_SPREAD_THREADS;  // threads per slab block
// Row stride of the int64 slab, >= K0 and = 4 (mod 32) elements: the 64 lanes
// of one stencil (4 x points, 4 y rows, 4 z rows) then fall on distinct LDS
// banks in each half wave; a stride that is a multiple of 16 puts all 16 rows
// of a lane group on the same banks.
__host__ __device__ __forceinline__ int slab_stride(int K0) { return K0 + ((36 - K0 % 32) % 32); }
// Atoms whose stencils a wave walks are staged as records in LDS, kSpreadRec
// per wave at a time: the twelve FP64 B-spline weights (the charge folded into
// the x weights) and the packed stencil base, followed by the list of the
// records' stencil rows (y, z) that lie inside the slab (16-bit entries:
// record, y row, z row).
constexpr int kSpreadRec = 32;
// Register budget of the spread kernels: five waves per SIMD, so a slab
// block still finds room beside the pair kernel's waves
#ifndef APO_SPREAD_WAVES_ATTR
#define APO_SPREAD_WAVES_ATTR __attribute__((amdgpu_waves_per_eu(5)))
#endif
constexpr size_t kSpreadRecBytes =
    (size_t)kSpreadRec * (12 * sizeof(double) + sizeof(int) + 16 * sizeof(unsigned short));
static_assert(kSpreadRec <= 32 && kSpreadRecBytes % 8 == 0, "record index and alignment");
template <typename C2> static size_t slab_lds(int K0, int by, int bz) {
  const size_t rows = (size_t)by * bz;
  return rows * slab_stride(K0) * sizeof(unsigned long long) + (size_t)K0 * sizeof(C2) +
         rows * K0 * sizeof(C2) + (kSlabCand + 4) * sizeof(int) +
         (kSpreadThreads / 64) * kSpreadRecBytes;
}

// Cyclic-interval overlap on a circle of length K (iK = 1/K).  The reduced
// offsets may be off by one rounding of K, which the callers' slack of two
// grid points covers.
__device__ __forceinline__ bool cyc_overlap(float a, float la, float b, float lb, float K,
                                            float iK) {
  float d1 = b - a;
  d1 -= K * floorf(d1 * iK);
  float d2 = a - b;
  d2 -= K * floorf(d2 * iK);
  return d1 < la || d2 < lb;
}

// round-to-nearest-even of v as an integer, for |v| < 2^51: adding 1.5 * 2^52
// leaves the rounded integer in the low mantissa bits
__device__ __forceinline__ long long rint_i64(double v) {
  const double magic = 6755399441055744.0;
  return __double_as_longlong(v + magic) - __double_as_longlong(magic);
}

// fr = K frac(u + 1/2) from the 32-bit fixed-point box coordinate u, in
// integers: with v = u + 2^31 (mod 2^32) and K v < 2^43, ix = floor(K v / 2^32)
// and w = (K v mod 2^32) / 2^32.  Every step of the FP64 form
// K (u/2^32 + 1/2 - floor(u/2^32 + 1/2)) is exact, so the bits are the same.
__device__ __forceinline__ void frac_fixed(int ufix, int K, int &ix, double &w) {
  const unsigned v = (unsigned)ufix + 0x80000000u;
  const unsigned long long p = (unsigned long long)v * (unsigned)K;
  ix = (int)(p >> 32);
  w = (double)(unsigned)(p & 0xffffffffull) * (1.0 / 4294967296.0);
}

// 6 x the order-4 B-spline weight of stencil point t at offset w:
// (1-w)^3, 4 - 6w^2 + 3w^3, 1 + 3w + 3w^2 - 3w^3, w^3 for t = 0..3 (the
// closed form of the recursion in bspline4).
struct Theta6 {
  double c0, c1, c2, c3;
};
__device__ __forceinline__ Theta6 theta6_coef(int t) {
  switch (t) {
    case 0: return Theta6{1.0, -3.0, 3.0, -1.0};
    case 1: return Theta6{4.0, 0.0, -6.0, 3.0};
    case 2: return Theta6{1.0, 3.0, 3.0, -3.0};
    default: return Theta6{0.0, 0.0, 0.0, 1.0};
  }
}
__device__ __forceinline__ double theta6(const Theta6 &c, double w) {
  return fma(fma(fma(c.c3, w, c.c2), w, c.c1), w, c.c0);
}

__device__ __forceinline__ double readlane_f64(double v, int k) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <typename CT, bool BIG>
__device__ __forceinline__ void spread_fftx_body(int bid, const int4 *__restrict__ posq,
                                                 const float4 *__restrict__ bb,
                                                 const int *__restrict__ meta, const BrickGeom &g,
                                                 double qunscale, const FftPlan &p,
                                                 const typename Cx<CT>::T *__restrict__ twg,
                                                 typename Cx<CT>::T *__restrict__ out) {
  using C2 = typename Cx<CT>::T;
  extern __shared__ char smem[];  // no static LDS in this kernel: smem is at offset 0
  const int n = g.K0, by = g.by, bz = g.bz, nrow = by * bz;
  unsigned long long *slab = reinterpret_cast<unsigned long long *>(smem);
  const int ss = slab_stride(n);
  C2 *tw = reinterpret_cast<C2 *>(slab + (size_t)nrow * ss);
  C2 *a = tw + n;
  C2 *b = a + (nrow / 2) * n;
  int *cand = reinterpret_cast<int *>(b + (nrow / 2) * n);
  int &ncand = cand[kSlabCand];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // this wave's atom records: x weights times charge, y and z weights (four
  // stencil points each), packed base (ix | (dy + 3) << 16 | (dz + 3) << 24)
  double *rtx = reinterpret_cast<double *>(cand + kSlabCand + 4) +
                (size_t)wave * (kSpreadRecBytes / sizeof(double));
  double *rty = rtx + 4 * kSpreadRec, *rtz = rty + 4 * kSpreadRec;
  int *rpos = reinterpret_cast<int *>(rtz + 4 * kSpreadRec);
  unsigned short *rlist = reinterpret_cast<unsigned short *>(rpos + kSpreadRec);
  const int byi = bid % g.nby, bzi = bid / g.nby;
  const int y0 = byi * by, z0 = bzi * bz;
  for (int k = tid; k < nrow * ss; k += kSpreadThreads) slab[k] = 0ull;
  staged_copy<C2>(n, tid, kSpreadThreads, [&](int e) { return twg[e]; },
                  [&](int e, C2 v) { tw[e] = v; });
  const int ncell = meta[0];
  // atoms whose stencil base lies in [p0 - 3, p0 + B) reach the slab in y and z
  const float by0 = (float)(y0 - 3), bz0 = (float)(z0 - 3);
  const float K1 = (float)g.K1, K2 = (float)g.K2;
  const float iK1 = 1.f / K1, iK2 = 1.f / K2;
  const float gy = K1 / g.L1, gz = K2 / g.L2;  // grid points per A
  // cells are tested kSlabCand at a time, four per thread with their loads in
  // flight together; the atoms of the cells that pass are walked by the four
  // waves, each prefetching its next atom while it adds the current hits
  for (int c0 = 0; c0 < ncell; c0 += kSlabCand) {
    if (tid == 0) ncand = 0;
    __syncthreads();
    float4 cc[kSlabCand / kSpreadThreads], hh[kSlabCand / kSpreadThreads];
#pragma unroll
    for (int u = 0; u < kSlabCand / kSpreadThreads; u++) {
      const int c = c0 + u * kSpreadThreads + tid;
      if (c < ncell) {
        cc[u] = bb[2 * c];
        hh[u] = bb[2 * c + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < kSlabCand / kSpreadThreads; u++) {
      const int c = c0 + u * kSpreadThreads + tid;
      if (c < ncell) {
        // cell extent in grid units (x/L + 1/2 origin), +2 for the floor and
        // single-precision slack
        float ay = fmaf(cc[u].y - hh[u].y, gy, 0.5f * K1), lay = fmaf(2.f * hh[u].y, gy, 2.f);
        float az = fmaf(cc[u].z - hh[u].z, gz, 0.5f * K2), laz = fmaf(2.f * hh[u].z, gz, 2.f);
        if (cyc_overlap(ay - 1.f, lay, by0, (float)(by + 3), K1, iK1) &&
            cyc_overlap(az - 1.f, laz, bz0, (float)(bz + 3), K2, iK2)) {
          int slot = atomicAdd(&ncand, 1);
          cand[slot] = c;
        }
      }
    }
    __syncthreads();
    const int na = ncand * kCell;
    int4 nxt = make_int4(0, 0, 0, 0);
    if (wave * 64 + lane < na) {
      const int w = wave * 64 + lane;
      nxt = posq[cand[w >> 5] * kCell + (w & 31)];
    }
    for (int base = wave * 64; base < na; base += kSpreadThreads) {
      // lane prepares atom base + lane
      const int w = base + lane;
      const int4 q4 = nxt;
      if (w + kSpreadThreads < na)
        nxt = posq[cand[(w + kSpreadThreads) >> 5] * kCell + ((w + kSpreadThreads) & 31)];
      int ix = 0, dy = by, dz = bz;
      double wx = 0.0, wy = 0.0, wz = 0.0, qs = 0.0;
      if (w < na) {
        const double q = (double)__int_as_float(q4.w) * qunscale;
        if (q != 0.0) {
          int iy, iz;
          frac_fixed(q4.x, g.K0, ix, wx);
          frac_fixed(q4.y, g.K1, iy, wy);
          frac_fixed(q4.z, g.K2, iz, wz);
          // offsets of the stencil base relative to the slab origin, reduced
          // into [-3, K-4] (dy + 3 lies in [4 - K, K + 2])
          dy = iy - y0;
          dz = iz - z0;
          dy += dy + 3 < 0 ? g.K1 : (dy + 3 >= g.K1 ? -g.K1 : 0);
          dz += dz + 3 < 0 ? g.K2 : (dz + 3 >= g.K2 ? -g.K2 : 0);
          qs = q * (kGridScale / 216.0);
        }
      }
      const bool ishit = dy < by && dz < bz;
      const unsigned long long hit = __ballot(ishit);
      const int nhit = __popcll(hit);
      const int rank = __popcll(hit & ((1ull << lane) - 1ull));
      // the hits' weights are formed once per atom (by the atom's own lane)
      // and staged in LDS, with the list of the atom's stencil rows (y, z)
      // inside the slab; a lane then takes one listed row and adds its four x
      // points, each product formed as (q wx) wy wz, the order in which a
      // lane forming it itself would
      for (int h0 = 0; h0 < nhit; h0 += kSpreadRec) {
        const bool mine = ishit && rank >= h0 && rank < h0 + kSpreadRec;
        const int r = rank - h0;
        unsigned ymask = 0, zmask = 0;
        if (mine) {
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const Theta6 c = theta6_coef(t);
            rtx[4 * r + t] = qs * theta6(c, wx);
            rty[4 * r + t] = theta6(c, wy);
            rtz[4 * r + t] = theta6(c, wz);
            if (dy + t >= 0 && dy + t < by) ymask |= 1u << t;
            if (dz + t >= 0 && dz + t < bz) zmask |= 1u << t;
          }
          rpos[r] = ix | ((dy + 3) << 16) | ((dz + 3) << 24);
        }
        const int cnt = __popc(ymask) * __popc(zmask);
        int incl = cnt;  // inclusive prefix sum over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int v = __shfl_up(incl, o, 64);
          if (lane >= o) incl += v;
        }
        const int total = __shfl(incl, 63, 64);
        if (cnt) {
          int at = incl - cnt;
#pragma unroll
          for (int t = 0; t < 4; t++)
#pragma unroll
            for (int u = 0; u < 4; u++)
              if (((zmask >> t) & (ymask >> u) & 1u) != 0u)
                rlist[at++] = (unsigned short)(r | (u << 5) | (t << 7));
        }
        // the records are this wave's own: in-order LDS within a wave, no
        // block barrier (waves walk different numbers of chunks)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int e = lane; e < total; e += 64) {
          const int ent = rlist[e];
          const int rr = ent & 31, sy = (ent >> 5) & 3, sz = ent >> 7;
          const int pos = rpos[rr];
          const int px0 = pos & 0xffff;
          const int py = ((pos >> 16) & 0xff) - 3 + sy;
          const int pz = (int)((unsigned)pos >> 24) - 3 + sz;
          unsigned long long *row = slab + (size_t)(pz * by + py) * ss;
          const double wyv = rty[4 * rr + sy], wzv = rtz[4 * rr + sz];
#pragma unroll
          for (int a = 0; a < 4; a++) {
            int px = px0 + a;
            if (px >= n) px -= n;
            const long long iv = rint_i64(rtx[4 * rr + a] * wyv * wzv);
            if (iv) atomicAdd(&row[px], (unsigned long long)iv);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
  }
  // rows (y, y + 1) of one z plane -> one complex row; rows past the grid
  // (edge slabs) hold only wrapped duplicates and enter as zero
  const int hb = by >> 1, ncx = nrow >> 1, H = n / 2 + 1;
  const unsigned mn = magic_dev(n), mh = magic_dev(H), mhb = magic_dev(hb);
  for (int e = tid; e < ncx * n; e += kSpreadThreads) {
    const int r = fdiv(e, mn), x = e - r * n;
    const int pz = fdiv(r, mhb), py = 2 * (r - pz * hb);
    const int y = y0 + py, z = z0 + pz;
    const unsigned long long *row = slab + (size_t)(pz * by + py) * ss;
    C2 v;
    v.x = (y < g.K1 && z < g.K2) ? (CT)((double)(long long)row[x] * (1.0 / kGridScale)) : (CT)0;
    v.y = (y + 1 < g.K1 && z < g.K2) ? (CT)((double)(long long)row[ss + x] * (1.0 / kGridScale))
                                     : (CT)0;
    a[e] = v;
  }
  __syncthreads();
  C2 *res = lds_fft<C2, BIG>(a, b, p, ncx, tw, false);
  for (int e = tid; e < ncx * H; e += kSpreadThreads) {
    const int r = fdiv(e, mh), k = e - r * H;
    const int pz = fdiv(r, mhb), py = 2 * (r - pz * hb);
    const int y = y0 + py, z = z0 + pz;
    const C2 Z = res[r * n + k], Zm = res[r * n + (k ? n - k : 0)];
    C2 A, B;
    A.x = (CT)0.5 * (Z.x + Zm.x);
    A.y = (CT)0.5 * (Z.y - Zm.y);
    B.x = (CT)0.5 * (Z.y + Zm.y);
    B.y = (CT)0.5 * (Zm.x - Z.x);
    const size_t o = ((size_t)z * g.K1 + y) * H + k;
    if (z < g.K2 && y < g.K1) out[o] = A;
    if (z < g.K2 && y + 1 < g.K1) out[o + H] = B;
  }
}
template <typename CT, bool PRIO, bool BIG>
__global__ __launch_bounds__(kSpreadThreads) APO_SPREAD_WAVES_ATTR
void k_spread_fftx(const int4 *__restrict__ posq,
                                                     const float4 *__restrict__ bb,
                                                     const int *__restrict__ meta, BrickGeom g,
                                                     double qunscale, FftPlan p,
                                                     const typename Cx<CT>::T *__restrict__ twg,
                                                     typename Cx<CT>::T *__restrict__ out) {
  if constexpr (PRIO) pme_wave_priority();
  spread_fftx_body<CT, BIG>(blockIdx.x, posq, bb, meta, g, qunscale, p, twg, out);
}

// Gather in cell order: 16 lanes per atom; the gradient goes straight into
// the global-order output with int64 atomics (three per atom), so the gather
// can run while the pair kernel and the bonded terms write the same atoms.
template <typename CT>
__device__ __forceinline__ void gather_body(int bid, const int4 *__restrict__ posq,
                                            const int *__restrict__ meta,
                                            const int *__restrict__ loc2glo, const GridGeom &g,
                                            double qunscale, const CT *__restrict__ phi,
                                            unsigned long long *__restrict__ force, int stride) {
  const int gt = bid * 256 + threadIdx.x;
  const int loc = gt >> 4, row = gt & 15;
  const bool valid = loc < meta[0] * kCell;
  const int4 p = valid ? posq[loc] : make_int4(0, 0, 0, 0);
  const double q = (double)__int_as_float(p.w) * qunscale;
  int ix, iy, iz;
  double wx, wy, wz;
  frac_fixed(p.x, g.K0, ix, wx);
  frac_fixed(p.y, g.K1, iy, wy);
  frac_fixed(p.z, g.K2, iz, wz);
  CT tx[4], ty[4], tz[4], dx[4], dy[4], dz[4];
  bspline4<CT>((CT)wx, tx, dx);
  bspline4<CT>((CT)wy, ty, dy);
  bspline4<CT>((CT)wz, tz, dz);
  const int a = row & 3, b = row >> 2;
  int y = iy + a, z = iz + b;
  if (y >= g.K1) y -= g.K1;
  if (z >= g.K2) z -= g.K2;
  CT f1 = 0, f2 = 0, f3 = 0;
  if (valid && q != 0.0) {
    const size_t base = (size_t)g.K0 * ((size_t)y + (size_t)g.K1 * z);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int x = ix + c;
      if (x >= g.K0) x -= g.K0;
      const CT v = phi[base + x];
      f1 += dx[c] * ty[a] * tz[b] * v;
      f2 += tx[c] * dy[a] * tz[b] * v;
      f3 += tx[c] * ty[a] * dz[b] * v;
    }
  }
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) {
    f1 += __shfl_xor(f1, off);
    f2 += __shfl_xor(f2, off);
    f3 += __shfl_xor(f3, off);
  }
  if (valid && q != 0.0 && row == 0) {
    const int gi = loc2glo[loc];
    if (gi >= 0) {
      const double qc = q * kCcelec;
      auto fx = [](double v) { return (unsigned long long)(long long)__builtin_rint(v); };
      atomicAdd(&force[gi], fx(qc * g.inv0 * g.K0 * (double)f1 * kForceScale));
      atomicAdd(&force[stride + gi], fx(qc * g.inv1 * g.K1 * (double)f2 * kForceScale));
      atomicAdd(&force[2 * (size_t)stride + gi], fx(qc * g.inv2 * g.K2 * (double)f3 * kForceScale));
    }
  }
}
template <typename CT, bool PRIO>
__global__ __launch_bounds__(256) void k_gather_sorted(const int4 *__restrict__ posq,
                                                       const int *__restrict__ meta,
                                                       const int *__restrict__ loc2glo, GridGeom g,
                                                       double qunscale, const CT *__restrict__ phi,
                                                       unsigned long long *__restrict__ force,
                                                       int stride) {
  if constexpr (PRIO) pme_wave_priority();
  gather_body<CT>(blockIdx.x, posq, meta, loc2glo, g, qunscale, phi, force, stride);
}


// x inverse: complex half rows -> real rows (C2R semantics: the imaginary
// parts of the k = 0 and Nyquist bins are ignored).
template <typename CT, bool BIG>
__device__ __forceinline__ void fftx_inv_body(int bid, const typename Cx<CT>::T *__restrict__ in,
                                              int nrows, const FftPlan &p, int rb,
                                              const typename Cx<CT>::T *__restrict__ twg,
                                              CT *__restrict__ phi) {
  using C2 = typename Cx<CT>::T;
  extern __shared__ char smem[];
  const int n = p.n, H = n / 2 + 1;
  C2 *tw = reinterpret_cast<C2 *>(smem);
  C2 *a = tw + n;
  C2 *b = a + rb * n;
  staged_copy<C2>(n, threadIdx.x, blockDim.x, [&](int e) { return twg[e]; },
                  [&](int e, C2 v) { tw[e] = v; });
  const int r0 = bid * rb;
  const unsigned mn = magic_dev(n);
  struct Pair2 {
    C2 A, B;
  };
  staged_copy<Pair2>(
      rb * n, threadIdx.x, blockDim.x,
      [&](int e) {
        const int r = fdiv(e, mn), k = e - r * n;
        const int ra = 2 * (r0 + r), rbw = ra + 1;
        const int kk = k > n / 2 ? n - k : k;
        Pair2 q;
        q.A = C2{0, 0};
        q.B = C2{0, 0};
        if (ra < nrows) q.A = in[(size_t)ra * H + kk];
        if (rbw < nrows) q.B = in[(size_t)rbw * H + kk];
        return q;
      },
      [&](int e, Pair2 q) {
        const int k = e - fdiv(e, mn) * n;
        const bool mirror = k > n / 2;
        const int kk = mirror ? n - k : k;
        C2 A = q.A, B = q.B;
        if (kk == 0 || 2 * kk == n) {
          A.y = 0;
          B.y = 0;
        }
        if (mirror) {
          A.y = -A.y;
          B.y = -B.y;
        }
        C2 Z;
        Z.x = A.x - B.y;
        Z.y = A.y + B.x;
        a[e] = Z;
      });
  __syncthreads();
  C2 *res = lds_fft<C2, BIG>(a, b, p, rb, tw, true);
  for (int e = threadIdx.x; e < rb * n; e += blockDim.x) {
    const int r = fdiv(e, mn), x = e - r * n;
    const int ra = 2 * (r0 + r), rbw = ra + 1;
    if (ra < nrows) phi[(size_t)ra * n + x] = res[e].x;
    if (rbw < nrows) phi[(size_t)rbw * n + x] = res[e].y;
  }
}
template <typename CT, bool PRIO, bool BIG>
__global__ __launch_bounds__(256) void k_fftx_inv(const typename Cx<CT>::T *__restrict__ in,
                                                  int nrows, FftPlan p, int rb,
                                                  const typename Cx<CT>::T *__restrict__ twg,
                                                  CT *__restrict__ phi) {
  if constexpr (PRIO) pme_wave_priority();
  fftx_inv_body<CT, BIG>(blockIdx.x, in, nrows, p, rb, twg, phi);
}


// y pass (either direction): block = (z, chunk of W kx columns).
template <typename CT, bool BIG>
__device__ __forceinline__ void ffty_body(int bid, typename Cx<CT>::T *__restrict__ data, int H,
                                          int K1, int W, const FftPlan &p,
                                          const typename Cx<CT>::T *__restrict__ twg, int inv) {
  using C2 = typename Cx<CT>::T;
  extern __shared__ char smem[];
  const int n = p.n;
  C2 *tw = reinterpret_cast<C2 *>(smem);
  C2 *a = tw + n;
  C2 *b = a + W * n;
  staged_copy<C2>(n, threadIdx.x, blockDim.x, [&](int e) { return twg[e]; },
                  [&](int e, C2 v) { tw[e] = v; });
  const int nchunk = (H + W - 1) / W;
  const int z = bid / nchunk, kx0 = (bid % nchunk) * W;
  const int wc = min(W, H - kx0);
  const unsigned mw = magic_dev(W);
  staged_copy<C2>(
      W * n, threadIdx.x, blockDim.x,
      [&](int e) {
        const int y = fdiv(e, mw), c = e - y * W;
        return c < wc ? data[((size_t)z * K1 + y) * H + kx0 + c] : C2{0, 0};
      },
      [&](int e, C2 v) {
        const int y = fdiv(e, mw), c = e - y * W;
        if (c < wc) a[c * n + y] = v;
      });
  __syncthreads();
  C2 *res = lds_fft<C2, BIG>(a, b, p, wc, tw, inv != 0);
  for (int e = threadIdx.x; e < W * n; e += blockDim.x) {
    const int y = fdiv(e, mw), c = e - y * W;
    if (c < wc) data[((size_t)z * K1 + y) * H + kx0 + c] = res[c * n + y];
  }
}
template <typename CT, bool PRIO, bool BIG>
__global__ __launch_bounds__(256) void k_ffty(typename Cx<CT>::T *__restrict__ data, int H, int K1,
                                              int W, FftPlan p,
                                              const typename Cx<CT>::T *__restrict__ twg,
                                              int inv) {
  if constexpr (PRIO) pme_wave_priority();
  ffty_body<CT, BIG>(blockIdx.x, data, H, K1, W, p, twg, inv);
}


struct ConvArgs {
  int K0, K1, K2, H, W;
  double r0, r1, r2;   // 1/L per axis
  double fac, piv_inv;  // pi^2/kappa^2, 1/(pi V)
  double half_ccelec;
};

template <typename CT> __device__ __forceinline__ CT exp_k(CT x);
template <> __device__ __forceinline__ float exp_k<float>(float x) { return expf(x); }
template <> __device__ __forceinline__ double exp_k<double>(double x) { return exp(x); }

// z pass with the k-space convolution in the middle: block = (y, W columns).
template <typename CT, bool CALC_EV, bool BIG>
__device__ __forceinline__ void fftz_conv_body(int bid, typename Cx<CT>::T *__restrict__ data,
                                               const ConvArgs &c, const FftPlan &p,
                                               const typename Cx<CT>::T *__restrict__ twg,
                                               const CT *__restrict__ prefac,
                                               double *__restrict__ part) {
  using C2 = typename Cx<CT>::T;
  extern __shared__ char smem[];
  const int n = p.n, H = c.H, W = c.W;
  C2 *tw = reinterpret_cast<C2 *>(smem);
  C2 *a = tw + n;
  C2 *b = a + W * n;
  CT *pre = reinterpret_cast<CT *>(b + W * n);
  __shared__ double red[4 * 8];
  staged_copy<C2>(n, threadIdx.x, blockDim.x, [&](int e) { return twg[e]; },
                  [&](int e, C2 v) { tw[e] = v; });
  staged_copy<CT>(c.K0 + c.K1 + c.K2, threadIdx.x, blockDim.x, [&](int e) { return prefac[e]; },
                  [&](int e, CT v) { pre[e] = v; });
  const int nchunk = (H + W - 1) / W;
  const int y = bid / nchunk, kx0 = (bid % nchunk) * W;
  const int wc = min(W, H - kx0);
  const size_t zs = (size_t)c.K1 * H;
  const unsigned mw = magic_dev(W), mn = magic_dev(n);
  staged_copy<C2>(
      W * n, threadIdx.x, blockDim.x,
      [&](int e) {
        const int z = fdiv(e, mw), col = e - z * W;
        return col < wc ? data[(size_t)z * zs + (size_t)y * H + kx0 + col] : C2{0, 0};
      },
      [&](int e, C2 v) {
        const int z = fdiv(e, mw), col = e - z * W;
        if (col < wc) a[col * n + z] = v;
      });
  __syncthreads();
  C2 *res = lds_fft<C2, BIG>(a, b, p, wc, tw, false);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int nf0 = c.K0 / 2 + (c.K0 % 2), nf1 = c.K1 / 2 + (c.K1 % 2), nf2 = c.K2 / 2 + (c.K2 % 2);
  for (int e = threadIdx.x; e < wc * n; e += blockDim.x) {
    const int col = fdiv(e, mn), z = e - col * n;
    const int kx = kx0 + col;
    C2 q = res[e];
    if (kx == 0 && y == 0 && z == 0) {
      q.x = 0;
      q.y = 0;
    } else {
      const int m0 = kx >= nf0 ? kx - c.K0 : kx;
      const int m1 = y >= nf1 ? y - c.K1 : y;
      const int m2 = z >= nf2 ? z - c.K2 : z;
      const CT mh0 = (CT)c.r0 * (CT)m0, mh1 = (CT)c.r1 * (CT)m1, mh2 = (CT)c.r2 * (CT)m2;
      const CT msq = mh0 * mh0 + mh1 * mh1 + mh2 * mh2;
      const CT msq_inv = (CT)1 / msq;
      const CT eterm = exp_k<CT>(-(CT)c.fac * msq) * (CT)c.piv_inv * pre[kx] *
                       pre[c.K0 + y] * pre[c.K0 + c.K1 + z] * msq_inv;
      if (CALC_EV) {
        // the x half grid stands for both Hermitian halves: every kx >= 1
        // column counts twice (the reference's mirror bookkeeping)
        const CT mult = kx >= 1 ? (CT)2 : (CT)1;
        const CT t1 = mult * eterm * (q.x * q.x + q.y * q.y);
        const CT vterm = (CT)2 * ((CT)c.fac + msq_inv);
        const CT t2 = t1 * vterm;
        acc[0] += (double)t1;
        acc[1] += (double)(t1 * (vterm * mh0 * mh0 - (CT)1));
        acc[2] += (double)(t2 * mh0 * mh1);
        acc[3] += (double)(t2 * mh0 * mh2);
        acc[4] += (double)(t1 * (vterm * mh1 * mh1 - (CT)1));
        acc[5] += (double)(t2 * mh1 * mh2);
        acc[6] += (double)(t1 * (vterm * mh2 * mh2 - (CT)1));
      }
      q.x *= eterm;
      q.y *= eterm;
    }
    res[e] = q;
  }
  __syncthreads();
  C2 *other = res == a ? b : a;
  C2 *out = lds_fft<C2, BIG>(res, other, p, wc, tw, true);
  for (int e = threadIdx.x; e < W * n; e += blockDim.x) {
    const int z = fdiv(e, mw), col = e - z * W;
    if (col < wc) data[(size_t)z * zs + (size_t)y * H + kx0 + col] = out[col * n + z];
  }
  if (CALC_EV) {
    for (int k = 0; k < 7; k++) {
      double v = acc[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * 8 + k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
      double s = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w * 8 + threadIdx.x];
      // energy +1/2 ccelec sum, virial -1/2 ccelec sum
      part[bid * kRecipPart + threadIdx.x] =
          threadIdx.x == 0 ? s * c.half_ccelec : -s * c.half_ccelec;
    }
  }
}
template <typename CT, bool CALC_EV, bool PRIO, bool BIG>
__global__ __launch_bounds__(256) void k_fftz_conv(typename Cx<CT>::T *__restrict__ data, ConvArgs c,
                                                   FftPlan p,
                                                   const typename Cx<CT>::T *__restrict__ twg,
                                                   const CT *__restrict__ prefac,
                                                   double *__restrict__ part) {
  if constexpr (PRIO) pme_wave_priority();
  fftz_conv_body<CT, CALC_EV, BIG>(blockIdx.x, data, c, p, twg, prefac, part);
}


// Launch shapes of the reciprocal passes for one system (shared by the
// single-system launcher and the fused multi-system launch).
template <typename CT> struct PmeShapes {
  GridGeom g;
  BrickGeom bg;
  int nslabs, nrows, npair, rb, wy, ncy, wz, ncz;
  int ft;  // threads per block of the y, z and x-inverse passes
  size_t shs, shx, shy, shz;
  ConvArgs c;
};

// rows: x-inverse row pairs per block; cols: y/z-pass columns per block (both
// capped by 64 KiB of LDS); threads: block size of those passes (64..256).
// The spread slab is kSlabBy x kSlabBz rows, halved (z first, then y down to
// 2) while the grid gives fewer slab blocks than CUs (small grids: a 32^3 grid
// has 64 4x4 slabs) or the LDS does not fit; every halving doubles the atoms
// each block walks, so one block per CU is kept, not more.  Any slab
// shape sums every grid point exactly and pairs the same rows for the x FFT,
// so the grid is bitwise the same.
template <typename CT>
// batched: the system is one of a fused batch, whose launch is filled by the
// other systems, so its slabs are not narrowed for parallelism
static PmeShapes<CT> pme_shapes(const NbSystem &S, int rows = kFftRowsPerBlock,
                                int cols = kFftColsPerBlock, int threads = kFftThreads,
                                bool batched = false) {
  using C2 = typename Cx<CT>::T;
  PmeShapes<CT> P;
  if (threads < 64 || threads > 256 || threads % 64)
    throw InvalidArgument("FFT threads per block must be 64, 128, 192 or 256");
  P.ft = threads;
  const int K0 = S.K[0], K1 = S.K[1], K2 = S.K[2], H = K0 / 2 + 1;
  P.g = GridGeom{K0, K1, K2, 1.0 / S.box[0], 1.0 / S.box[1], 1.0 / S.box[2]};
  const size_t lds_cap = 64 * 1024;
  int by = std::max(2, kSlabBy & ~1), bz = std::max(1, kSlabBz);
  auto slabs = [&](int y, int z) { return ((K1 + y - 1) / y) * ((K2 + z - 1) / z); };
  while ((slab_lds<C2>(K0, by, bz) > lds_cap - 2048 || (!batched && slabs(by, bz) < S.ncu)) &&
         (bz > 1 || by > 2)) {
    if (bz >= by && bz > 1) bz /= 2;
    else by = std::max(2, (by / 2) & ~1);
  }
  if (slab_lds<C2>(K0, by, bz) > lds_cap - 2048)
    throw InvalidArgument("PME grid x dimension too large for the spread slab");
  P.bg = BrickGeom{K0, K1, K2, by, bz, (K1 + by - 1) / by, (K2 + bz - 1) / bz, (float)S.box[0],
                   (float)S.box[1], (float)S.box[2]};
  P.nslabs = P.bg.nby * P.bg.nbz;
  P.shs = slab_lds<C2>(K0, by, bz);
  P.nrows = K1 * K2;
  P.rb = (int)std::max<size_t>(
      1, std::min<size_t>((size_t)std::max(rows, 1), (lds_cap / sizeof(C2) - K0) / (2 * K0)));
  P.shx = (size_t)(K0 + 2 * P.rb * K0) * sizeof(C2);
  P.npair = (P.nrows + 1) / 2;
  const int W = std::max(cols, 1);
  P.wy = (int)std::max<size_t>(1, std::min<size_t>(W, (lds_cap / sizeof(C2) - K1) / (2 * K1)));
  P.shy = (size_t)(K1 + 2 * P.wy * K1) * sizeof(C2);
  P.ncy = (H + P.wy - 1) / P.wy;
  P.wz = (int)std::max<size_t>(
      1, std::min<size_t>(W, (lds_cap / sizeof(C2) - K2 - (K0 + K1 + K2)) / (2 * K2)));
  P.shz = (size_t)(K2 + 2 * P.wz * K2) * sizeof(C2) + (size_t)(K0 + K1 + K2) * sizeof(CT);
  P.ncz = (H + P.wz - 1) / P.wz;
  P.c = ConvArgs{K0, K1, K2, H, P.wz, 1.0 / S.box[0], 1.0 / S.box[1], 1.0 / S.box[2],
                 kPi * kPi / (S.kappa * S.kappa),
                 1.0 / (S.box[0] * S.box[1] * S.box[2]) / kPi, 0.5 * kCcelec};
  return P;
}

template <typename CT, bool PRIO, bool BIG>
static void recip_t(NbSystem &S, const float4 *x, int64_t *force, int stride, bool EV,
                    typename Cx<CT>::T *cgrid, CT *phi, const typename Cx<CT>::T *twx,
                    const typename Cx<CT>::T *twy, const typename Cx<CT>::T *twz, const CT *pre,
                    hipStream_t s) {
  (void)x;
  const PmeShapes<CT> P = pme_shapes<CT>(S);
  const int K1 = S.K[1], K2 = S.K[2], H = S.K[0] / 2 + 1;
  const double qunscale = 1.0;  // sorted coordinates carry the raw charge
  hipEvent_t t0 = nullptr;
  S.tic(1, s, t0);
  k_spread_fftx<CT, PRIO, BIG><<<P.nslabs, kSpreadThreads, P.shs, s>>>(
      S.posq.get(), S.cell_bb.get(), S.meta.get(), P.bg, qunscale, plan_of(S.fx), twx, cgrid);
  APO_LAUNCH_CHECK();
  S.toc(1, s, t0);
  S.tic(2, s, t0);
  k_ffty<CT, PRIO, BIG><<<K2 * P.ncy, P.ft, P.shy, s>>>(cgrid, H, K1, P.wy, plan_of(S.fy), twy, 0);
  APO_LAUNCH_CHECK();
  if (K1 * P.ncz > S.part.recip_n) throw RuntimeError("reciprocal partial buffer too small");
  double *part = S.parts_ptr() + S.part.recip_off;
  if (EV) {
    launch_fill_f64(part, (size_t)S.part.recip_n * kRecipPart, s);
    k_fftz_conv<CT, true, PRIO, BIG><<<K1 * P.ncz, P.ft, P.shz, s>>>(cgrid, P.c, plan_of(S.fz), twz,
                                                                pre, part);
  } else {
    k_fftz_conv<CT, false, PRIO, BIG><<<K1 * P.ncz, P.ft, P.shz, s>>>(cgrid, P.c, plan_of(S.fz), twz,
                                                                 pre, part);
  }
  APO_LAUNCH_CHECK();
  k_ffty<CT, PRIO, BIG><<<K2 * P.ncy, P.ft, P.shy, s>>>(cgrid, H, K1, P.wy, plan_of(S.fy), twy, 1);
  APO_LAUNCH_CHECK();
  k_fftx_inv<CT, PRIO, BIG><<<(P.npair + P.rb - 1) / P.rb, P.ft, P.shx, s>>>(
      cgrid, P.nrows, plan_of(S.fx), P.rb, twx, phi);
  APO_LAUNCH_CHECK();
  S.toc(2, s, t0);
  S.tic(3, s, t0);
  k_gather_sorted<CT, PRIO><<<(S.ssort * 16 + 255) / 256, 256, 0, s>>>(
      S.posq.get(), S.meta.get(), S.loc2glo.get(), P.g, qunscale, phi,
      reinterpret_cast<unsigned long long *>(force), stride);
  APO_LAUNCH_CHECK();
  S.toc(3, s, t0);
}

void recip_launch(NbSystem &S, const float4 *x, int64_t *force, int stride, bool E, bool V,
                  hipStream_t s) {
  // the issue priority is a template parameter: as a kernel argument, every
  // wave waited for its load before issuing the kernel's own first loads
  auto run = [&](auto prio, auto big) {
    constexpr bool P = decltype(prio)::value, B = decltype(big)::value;
    if (S.precision == 1)
      recip_t<double, P, B>(S, x, force, stride, E || V, S.cgrid_d.get(), S.phi_d.get(),
                            S.fx.tw64.get(), S.fy.tw64.get(), S.fz.tw64.get(), S.prefac_d.get(),
                            s);
    else
      recip_t<float, P, B>(S, x, force, stride, E || V, S.cgrid.get(), S.phi.get(),
                           S.fx.tw.get(), S.fy.tw.get(), S.fz.tw.get(), S.prefac.get(), s);
  };
  const bool big = plan_big(S.fx) || plan_big(S.fy) || plan_big(S.fz);
  if (S.pme_prio) {
    if (big) run(std::true_type{}, std::true_type{});
    else run(std::true_type{}, std::false_type{});
  } else {
    if (big) run(std::false_type{}, std::true_type{});
    else run(std::false_type{}, std::false_type{});
  }
}

// ---- fused launches over several systems (mixed precision, no energy) ------
struct SpreadArgs {
  const int4 *posq;
  const float4 *bb;
  const int *meta;
  BrickGeom g;
  FftPlan p;
  const float2 *tw;
  float2 *cout;  // complex half rows out
};
struct FftxArgs {
  const float2 *cin;  // complex half rows in
  int nrows;
  FftPlan p;
  int rb;
  const float2 *tw;
  float *phi;         // potential out
};
struct FftyArgs {
  float2 *data;
  int H, K1, W;
  FftPlan p;
  const float2 *tw;
};
struct FftzArgs {
  float2 *data;
  ConvArgs c;
  FftPlan p;
  const float2 *tw;
  const float *prefac;
};
struct GatherArgs {
  const int4 *posq;
  const int *meta, *loc2glo;
  GridGeom g;
  const float *phi;
  unsigned long long *force;
  int stride;
};

template <bool BIG>
__global__ __launch_bounds__(kSpreadThreads) APO_SPREAD_WAVES_ATTR
void kb_spread_fftx(const SpreadArgs *__restrict__ tab,
                                                      const int *__restrict__ bstart, int nsys) {
  const int s = batch_system(bstart, nsys);
  const SpreadArgs &a = tab[s];
  spread_fftx_body<float, BIG>(blockIdx.x - bstart[s], a.posq, a.bb, a.meta, a.g, 1.0, a.p, a.tw,
                          a.cout);
}
template <bool BIG>
__global__ __launch_bounds__(256) void kb_fftx_inv(const FftxArgs *__restrict__ tab,
                                                   const int *__restrict__ bstart, int nsys) {
  const int s = batch_system(bstart, nsys);
  const FftxArgs &a = tab[s];
  fftx_inv_body<float, BIG>(blockIdx.x - bstart[s], a.cin, a.nrows, a.p, a.rb, a.tw, a.phi);
}
template <bool BIG>
__global__ __launch_bounds__(256) void kb_ffty(const FftyArgs *__restrict__ tab,
                                               const int *__restrict__ bstart, int nsys, int inv) {
  const int s = batch_system(bstart, nsys);
  const FftyArgs &a = tab[s];
  ffty_body<float, BIG>(blockIdx.x - bstart[s], a.data, a.H, a.K1, a.W, a.p, a.tw, inv);
}
template <bool BIG>
__global__ __launch_bounds__(256) void kb_fftz_conv(const FftzArgs *__restrict__ tab,
                                                    const int *__restrict__ bstart, int nsys) {
  const int s = batch_system(bstart, nsys);
  const FftzArgs &a = tab[s];
  fftz_conv_body<float, false, BIG>(blockIdx.x - bstart[s], a.data, a.c, a.p, a.tw, a.prefac,
                                    nullptr);
}
__global__ __launch_bounds__(256) void kb_gather(const GatherArgs *__restrict__ tab,
                                                 const int *__restrict__ bstart, int nsys) {
  const int s = batch_system(bstart, nsys);
  const GatherArgs &a = tab[s];
  gather_body<float>(blockIdx.x - bstart[s], a.posq, a.meta, a.loc2glo, a.g, 1.0, a.phi, a.force,
                     a.stride);
}

struct PmeBatch {
  BatchLaunch<SpreadArgs> spread;
  BatchLaunch<FftxArgs> fx_inv;
  BatchLaunch<FftyArgs> fy;
  BatchLaunch<FftzArgs> fz;
  BatchLaunch<GatherArgs> gather;
  int ft = kFftThreads;  // threads per block of the y, z and x-inverse passes
  bool big = false;      // some system's plans need radices above 5 (plan_big)
};

PmeBatch *pme_batch_create() { return new PmeBatch(); }
void pme_batch_destroy(PmeBatch *b) { delete b; }
void pme_batch_clear(PmeBatch &b) {
  b.spread.clear();
  b.fx_inv.clear();
  b.fy.clear();
  b.fz.clear();
  b.gather.clear();
  b.big = false;
}

void pme_batch_add(PmeBatch &b, NbSystem &S, int64_t *force, int stride, int fft_rows,
                   int fft_cols, int fft_threads) {
  if (S.precision != 0) throw InvalidArgument("fused batches run the mixed-precision path");
  const PmeShapes<float> P = pme_shapes<float>(S, fft_rows, fft_cols, fft_threads, true);
  b.ft = P.ft;
  b.big = b.big || plan_big(S.fx) || plan_big(S.fy) || plan_big(S.fz);
  if (S.K[1] * P.ncz > S.part.recip_n) throw RuntimeError("reciprocal partial buffer too small");
  const int H = S.K[0] / 2 + 1;
  b.spread.add(SpreadArgs{S.posq.get(), S.cell_bb.get(), S.meta.get(), P.bg, plan_of(S.fx),
                          S.fx.tw.get(), S.cgrid.get()},
               P.nslabs, P.shs);
  b.fx_inv.add(FftxArgs{S.cgrid.get(), P.nrows, plan_of(S.fx), P.rb, S.fx.tw.get(), S.phi.get()},
               (P.npair + P.rb - 1) / P.rb, P.shx);
  b.fy.add(FftyArgs{S.cgrid.get(), H, S.K[1], P.wy, plan_of(S.fy), S.fy.tw.get()},
           S.K[2] * P.ncy, P.shy);
  b.fz.add(FftzArgs{S.cgrid.get(), P.c, plan_of(S.fz), S.fz.tw.get(), S.prefac.get()},
           S.K[1] * P.ncz, P.shz);
  b.gather.add(GatherArgs{S.posq.get(), S.meta.get(), S.loc2glo.get(), P.g, S.phi.get(),
                          reinterpret_cast<unsigned long long *>(force), stride},
               (S.ssort * 16 + 255) / 256);
}

void pme_batch_upload(PmeBatch &b, hipStream_t s) {
  b.spread.upload(s);
  b.fx_inv.upload(s);
  b.fy.upload(s);
  b.fz.upload(s);
  b.gather.upload(s);
}

// spread + x FFT, the other four FFT passes with the convolution, gather: one
// launch each for all systems (cell boxes must be current).  Without the gather the
// launches write only the batch's grids (the autotuner times them so).
template <bool BIG> static void pme_batch_launch_t(PmeBatch &b, hipStream_t s, bool gather) {
  kb_spread_fftx<BIG><<<b.spread.blocks(), kSpreadThreads, b.spread.lds, s>>>(
      b.spread.d_args.get(), b.spread.d_start.get(), b.spread.nsys());
  kb_ffty<BIG><<<b.fy.blocks(), b.ft, b.fy.lds, s>>>(b.fy.d_args.get(), b.fy.d_start.get(),
                                                    b.fy.nsys(), 0);
  kb_fftz_conv<BIG><<<b.fz.blocks(), b.ft, b.fz.lds, s>>>(b.fz.d_args.get(), b.fz.d_start.get(),
                                                         b.fz.nsys());
  kb_ffty<BIG><<<b.fy.blocks(), b.ft, b.fy.lds, s>>>(b.fy.d_args.get(), b.fy.d_start.get(),
                                                    b.fy.nsys(), 1);
  kb_fftx_inv<BIG><<<b.fx_inv.blocks(), b.ft, b.fx_inv.lds, s>>>(
      b.fx_inv.d_args.get(), b.fx_inv.d_start.get(), b.fx_inv.nsys());
  if (gather)
    kb_gather<<<b.gather.blocks(), 256, 0, s>>>(b.gather.d_args.get(), b.gather.d_start.get(),
                                                b.gather.nsys());
  APO_LAUNCH_CHECK();
}
void pme_batch_launch(PmeBatch &b, hipStream_t s, bool gather) {
  if (!b.spread.nsys()) return;
  if (b.big) pme_batch_launch_t<true>(b, s, gather);
  else pme_batch_launch_t<false>(b, s, gather);
}

}  // namespace apo
