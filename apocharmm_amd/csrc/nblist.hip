// Neighbor list: xy-column binning, deterministic z sort per column in LDS,
// 32-atom cells with bounding boxes, cell-pair tiles inside the list cutoff,
// per-tile 32x32 exclusion/prune masks, then per 16-slot i-group the packed
// j slots that have a live pair, and the pair-kernel work items.
//
// Reference behaviour (re-designed, not translated):
//   CudaNeighborListSort   src/CudaNeighborListSort.cu:31-716 (columns of edge
//                          (V*32/N)^(1/3), z-sorted 32-atom cells, bbox)
//   CudaNeighborListBuild  src/CudaNeighborListBuild.cu:336-969 (cell pairs
//                          within the list cutoff, 32x32 exclusion masks,
//                          diagonal tile keeps i<j only)
// Differences by design: coordinates are 32-bit fixed point per box edge, so
// every pair uses its minimum image through integer wrap-around and a tile is
// one cell pair (no 27 image shifts); the column sort orders (z, atom index)
// so the list -- and every result -- is bitwise reproducible.
#include <algorithm>
#include <cmath>

#include "system.hpp"
#include "thermo.hpp"

namespace apo {

__device__ __forceinline__ uint32_t to_fixed(double x, double inv_len) {
  double f = x * inv_len;
  f -= floor(f);
  // [0,1) -> [0, 2^32); 2^32 wraps to 0 which is the same periodic point
  return (uint32_t)(int64_t)llrint(f * 4294967296.0);
}

// 1. column of each atom (wrapped xy) and its slot in the column.  Atoms
// that are neighbours in index order usually share a column, so the lanes of
// a wave that hit the same column take their slots with one atomic (the
// first such lane adds the group's size, the others take consecutive slots).
// The coordinates the list is built from are kept (xref) for the integrator's
// displacement check.
__global__ void k_assign(const float4 *__restrict__ xyzq, int n, double3 inv_len, double3 len,
                         int ncx, int ncy, int *__restrict__ col_count, int *__restrict__ atom_col,
                         int *__restrict__ atom_slot, float *__restrict__ zkey,
                         int *__restrict__ meta, float4 *__restrict__ xref) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = xyzq[i];
  xref[i] = p;
  double fx = (double)p.x * inv_len.x, fy = (double)p.y * inv_len.y, fz = (double)p.z * inv_len.z;
  fx -= floor(fx);
  fy -= floor(fy);
  fz -= floor(fz);
  // non-finite input: flag it (the host raises) and keep every index in range
  if (!(fx >= 0.0 && fx <= 1.0 && fy >= 0.0 && fy <= 1.0 && fz >= 0.0 && fz <= 1.0)) {
    atomicOr(&meta[6], 1);
    fx = fy = fz = 0.0;
  }
  int cx = max(0, min((int)(fx * ncx), ncx - 1)), cy = max(0, min((int)(fy * ncy), ncy - 1));
  int col = cx + ncx * cy;
  atom_col[i] = col;
  zkey[i] = (float)(fz * len.z);
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(true);
  int slot = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lcol = __shfl(col, leader, 64);
    const unsigned long long peers = __ballot(col == lcol) & todo;
    int base = 0;
    if (lane == leader) base = atomicAdd(&col_count[lcol], __popcll(peers));
    base = __shfl(base, leader, 64);
    if (col == lcol && ((todo >> lane) & 1ull))
      slot = base + __popcll(peers & ((1ull << lane) - 1ull));
    todo &= ~peers;
  }
  atom_slot[i] = slot;
}

// 2. cells per column; largest column.
__global__ void k_col_prep(const int *__restrict__ col_count, int ncol, int *__restrict__ col_ncell,
                           int *__restrict__ meta) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  int cnt = col_count[c];
  col_ncell[c] = (cnt + kCell - 1) / kCell;
  atomicMax(&meta[4], cnt);
}

// 3. scatter atoms into column order (slot order is arbitrary here; step 4
// makes it deterministic).
__global__ void k_scatter(int n, const int *__restrict__ atom_col, const int *__restrict__ atom_slot,
                          const int *__restrict__ col_start, const float *__restrict__ zkey,
                          unsigned long long *__restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int pos = col_start[atom_col[i]] + atom_slot[i];
  // (z, atom) packed so that an unsigned compare orders by z then index;
  // z >= 0, so its float bits are monotone as unsigned.
  keys[pos] = ((unsigned long long)__float_as_uint(zkey[i]) << 32) | (unsigned)i;
}

// 4. per-column bitonic sort in LDS, then cell layout.
__global__ __launch_bounds__(256) void k_colsort(const unsigned long long *__restrict__ keys,
                                                 const int *__restrict__ col_count,
                                                 const int *__restrict__ col_start,
                                                 const int *__restrict__ col_cellstart,
                                                 int *__restrict__ loc2glo, int *__restrict__ glo2loc,
                                                 int *__restrict__ meta) {
  __shared__ unsigned long long sh[kMaxColumn];
  const int c = blockIdx.x;
  const int cnt = col_count[c], st = col_start[c];
  if (cnt > kMaxColumn) {
    // leave the column's slots empty (the host raises after the build)
    if (threadIdx.x == 0) atomicOr(&meta[5], 1);
    const int cs0 = col_cellstart[c], ncl0 = (cnt + kCell - 1) / kCell;
    for (int k = threadIdx.x; k < ncl0 * kCell; k += blockDim.x) loc2glo[cs0 * kCell + k] = -1;
    return;
  }
  int P = 1;
  while (P < cnt) P <<= 1;
  for (int k = threadIdx.x; k < P; k += blockDim.x) sh[k] = k < cnt ? keys[st + k] : ~0ull;
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = threadIdx.x; k < P; k += blockDim.x) {
        int partner = k ^ stride;
        if (partner > k) {
          bool up = (k & size) == 0;
          unsigned long long a = sh[k], b = sh[partner];
          if ((a > b) == up) {
            sh[k] = b;
            sh[partner] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  const int cs = col_cellstart[c];
  const int ncl = (cnt + kCell - 1) / kCell;
  for (int k = threadIdx.x; k < ncl * kCell; k += blockDim.x) {
    int loc = cs * kCell + k;
    if (k < cnt) {
      int a = (int)(unsigned)(sh[k] & 0xffffffffull);
      loc2glo[loc] = a;
      glo2loc[a] = loc;
    } else {
      loc2glo[loc] = -1;
    }
  }
}

// Per-step gather of xyzq into cell order as 32-bit fixed-point coordinates
// (setSortedCoords, src/CudaPMEDirectForce.cu:1389-1400), fused with the
// clear of the sorted force buffer and the cell bounding boxes (centre and
// half extent in A of the fixed-point coordinates; the list build and the
// spread read them).  A block of 256 slots covers 8 whole cells, so each
// cell's box is a reduction over one half wave.
__device__ __forceinline__ void sort_coords_body(int bid, const float4 *__restrict__ xyzq,
                                                 const int *__restrict__ vdwtype,
                                                 const int *__restrict__ loc2glo,
                                                 const int *__restrict__ meta, double3 inv_len,
                                                 float qscale, float3 unit,
                                                 int4 *__restrict__ posq,
                                                 int *__restrict__ type_s,
                                                 int64_t *__restrict__ fsort, int ss,
                                                 float4 *__restrict__ bb,
                                                 int64_t *__restrict__ force, size_t nforce,
                                                 int nb) {
  const int loc = bid * 256 + threadIdx.x;
  // the step's gradient starts from zero (every term adds into it after the sort)
  if (force)
    for (size_t k = loc; k < nforce; k += (size_t)nb * 256) force[k] = 0;
  const int ncell = meta[0];
  const bool in = loc < ncell * kCell;
  const int g = in ? loc2glo[loc] : -1;
  int4 o = make_int4(0, 0, 0, 0);
  int ty = 0;
  if (g >= 0) {
    float4 p = xyzq[g];
    o.x = (int)to_fixed((double)p.x, inv_len.x);
    o.y = (int)to_fixed((double)p.y, inv_len.y);
    o.z = (int)to_fixed((double)p.z, inv_len.z);
    o.w = __float_as_int(p.w * qscale);
    ty = vdwtype[g];
  }
  if (in) {
    posq[loc] = o;
    type_s[loc] = ty;
    fsort[loc] = 0;
    fsort[ss + loc] = 0;
    fsort[2 * (size_t)ss + loc] = 0;
  }
  const bool valid = g >= 0;
  const float x = (float)(uint32_t)o.x * unit.x, y = (float)(uint32_t)o.y * unit.y,
              z = (float)(uint32_t)o.z * unit.z;
  float mnx = valid ? x : 3e38f, mny = valid ? y : 3e38f, mnz = valid ? z : 3e38f;
  float mxx = valid ? x : -3e38f, mxy = valid ? y : -3e38f, mxz = valid ? z : -3e38f;
  for (int off = 16; off > 0; off >>= 1) {
    mnx = fminf(mnx, __shfl_xor(mnx, off));
    mny = fminf(mny, __shfl_xor(mny, off));
    mnz = fminf(mnz, __shfl_xor(mnz, off));
    mxx = fmaxf(mxx, __shfl_xor(mxx, off));
    mxy = fmaxf(mxy, __shfl_xor(mxy, off));
    mxz = fmaxf(mxz, __shfl_xor(mxz, off));
  }
  const int cell = loc >> 5;
  if (cell < ncell && (loc & 31) == 0) {
    bb[2 * cell] = make_float4(0.5f * (mnx + mxx), 0.5f * (mny + mxy), 0.5f * (mnz + mxz), 0.f);
    bb[2 * cell + 1] = make_float4(0.5f * (mxx - mnx), 0.5f * (mxy - mny), 0.5f * (mxz - mnz), 0.f);
  }
}
// with_fin: the last block runs an integrator's pending kinetic-energy
// reduction instead (thermo.hpp)
__global__ __launch_bounds__(256) void k_sort_coords(const float4 *__restrict__ xyzq,
                                                     const int *__restrict__ vdwtype,
                                                     const int *__restrict__ loc2glo,
                                                     const int *__restrict__ meta, double3 inv_len,
                                                     float qscale, float3 unit,
                                                     int4 *__restrict__ posq,
                                                     int *__restrict__ type_s,
                                                     int64_t *__restrict__ fsort, int ss,
                                                     float4 *__restrict__ bb,
                                                     int64_t *__restrict__ force, size_t nforce,
                                                     LtFinArgs fin, int with_fin) {
  const int nsort = gridDim.x - with_fin;
  if ((int)blockIdx.x >= nsort) {
    lt_finalize_body(fin.part, fin.npart, fin.ndof, fin.thermo, fin.counters, fin.dpart, fin.meta,
                     fin.hdisp, fin.pend);
    return;
  }
  sort_coords_body(blockIdx.x, xyzq, vdwtype, loc2glo, meta, inv_len, qscale, unit, posq, type_s,
                   fsort, ss, bb, force, nforce, nsort);
}
__global__ __launch_bounds__(256) void kb_sort_coords(const SortArgs *__restrict__ tab,
                                                      const int *__restrict__ bstart, int nsys) {
  const int s = batch_system(bstart, nsys);
  const SortArgs a = tab[s];
  sort_coords_body(blockIdx.x - bstart[s], a.xyzq, a.vdwtype, a.loc2glo, a.meta, a.inv_len, 1.0f,
                   a.unit, a.posq, a.type_s, a.fsort, a.ss, a.bb, a.force, a.nforce,
                   bstart[s + 1] - bstart[s]);
}

__device__ __forceinline__ float bb_gap2(float4 ci, float4 hi, float4 cj, float4 hj, float3 L) {
  float d2 = 0.f;
  float d, g;
  d = ci.x - cj.x; d -= L.x * rintf(d / L.x); g = fmaxf(fabsf(d) - hi.x - hj.x, 0.f); d2 += g * g;
  d = ci.y - cj.y; d -= L.y * rintf(d / L.y); g = fmaxf(fabsf(d) - hi.y - hj.y, 0.f); d2 += g * g;
  d = ci.z - cj.z; d -= L.z * rintf(d / L.z); g = fmaxf(fabsf(d) - hi.z - hj.z, 0.f); d2 += g * g;
  return d2;
}

// 6. candidate tiles: cell pairs (i <= j) whose boxes are within the list
// cutoff.  emit = 0 counts per i-cell, emit = 1 writes in j order.
__global__ __launch_bounds__(64) void k_tiles(const float4 *__restrict__ bb,
                                              const int *__restrict__ meta, float3 L, float rl2,
                                              int *__restrict__ icell_cnt,
                                              const int *__restrict__ icell_start,
                                              int *__restrict__ cand_i, int *__restrict__ cand_j,
                                              int emit, int cand_cap) {
  const int icell = blockIdx.x;
  const int ncell = meta[0];
  if (icell >= ncell) return;
  const int lane = threadIdx.x;
  const float4 ci = bb[2 * icell], hi = bb[2 * icell + 1];
  int count = 0;
  const int base = emit ? icell_start[icell] : 0;
  for (int j0 = icell; j0 < ncell; j0 += 64) {
    int j = j0 + lane;
    bool hit = false;
    if (j < ncell) hit = bb_gap2(ci, hi, bb[2 * j], bb[2 * j + 1], L) < rl2;
    unsigned long long m = __ballot(hit);
    if (emit && hit) {
      int pos = base + count + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < cand_cap) {  // overflow is flagged by k_list_guard_export
        cand_i[pos] = icell;
        cand_j[pos] = j;
      }
    }
    count += __popcll(m);
  }
  if (!emit && lane == 0) icell_cnt[icell] = count;
}

// 7. 32x32 masks: bit t of lane l set <=> pair (i-cell atom t, j-cell atom l)
// is skipped: padding, diagonal i >= j, topological exclusion (1-2/1-3/1-4),
// or farther than the list cutoff at build time.  Four waves per i-cell walk
// its candidate tiles two at a time (lanes 0-31 and 32-63); the i-cell's atoms
// are staged in LDS once per block and read as wave-uniform values.
// ISO (cubic box): the differences stay in fixed-point units and r^2 is
// compared with rl2 in those units (rl2u), with no per-axis multiplies.
constexpr int kMaskWaves = 4;
template <bool ISO>
__global__ __launch_bounds__(64 * kMaskWaves) void k_tile_mask(
    const int *__restrict__ meta, const int *__restrict__ icell_start,
    const int *__restrict__ cand_j, const int4 *__restrict__ posq,
    const int *__restrict__ loc2glo, const int *__restrict__ glo2loc,
    const int *__restrict__ excl_ptr, const int *__restrict__ excl_idx, float3 unit, float rl2,
    float rl2u, uint32_t *__restrict__ cand_mask, int cand_cap) {
  const int ic = blockIdx.x;
  if (ic >= meta[0]) return;
  const int lane = threadIdx.x & 63, l = lane & 31, half = lane >> 5;
  const int wave = threadIdx.x >> 6;
  // the i-cell's coordinates, read by every wave as LDS broadcasts (one
  // 16-byte read per i atom instead of three cross-lane reads)
  __shared__ int4 sh_pi[kCell];
  if (threadIdx.x < kCell) sh_pi[threadIdx.x] = posq[ic * kCell + threadIdx.x];
  __syncthreads();
  const int gil = loc2glo[ic * kCell + l];
  uint32_t ipad = 0;  // bit t: i slot t is padding
  {
    const unsigned long long b = __ballot(gil < 0);
    ipad = (uint32_t)b;
  }
  const int s = icell_start[ic], e = min(icell_start[ic + 1], cand_cap);
  for (int k0 = s + 2 * wave; k0 < e; k0 += 2 * kMaskWaves) {
    const int k = k0 + half;
    const bool tv = k < e;
    const int jc = tv ? cand_j[k] : ic;
    const int jl = jc * kCell + l;
    const int gj = loc2glo[jl];
    uint32_t m = 0xffffffffu;
    const int4 pj = posq[jl];
    if (tv && gj >= 0) m = ipad | (jc == ic ? ~((1u << l) - 1u) : 0u);
#pragma unroll 8
    for (int t = 0; t < kCell; t++) {
      const int4 pi = sh_pi[t];
      const int px = pi.x, py = pi.y, pz = pi.z;
      const float dx = ISO ? (float)(int)(px - pj.x) : (float)(int)(px - pj.x) * unit.x;
      const float dy = ISO ? (float)(int)(py - pj.y) : (float)(int)(py - pj.y) * unit.y;
      const float dz = ISO ? (float)(int)(pz - pj.z) : (float)(int)(pz - pj.z) * unit.z;
      if (dx * dx + dy * dy + dz * dz > (ISO ? rl2u : rl2)) m |= 1u << t;
    }
    if (tv && gj >= 0)
      for (int x = excl_ptr[gj]; x < excl_ptr[gj + 1]; x++) {
        const int lp = glo2loc[excl_idx[x]];
        if ((lp >> 5) == ic) m |= 1u << (lp & 31);
      }
    if (tv) cand_mask[(size_t)k * kCell + l] = m;
  }
}

// 8. per i-group (kGroup consecutive slots of a cell) the number of j slots
// with at least one live pair in the 32x32 masks, rounded up to whole 64-lane
// steps, and the number of work items of <= kItemSteps steps.  Cells beyond
// the live count write zeros so one scan covers every group.
__global__ __launch_bounds__(64 * kGroupsPerCell) void k_grp_count(const int *__restrict__ meta,
                                                  const int *__restrict__ icell_start,
                                                  const uint32_t *__restrict__ cand_mask,
                                                  int *__restrict__ grp_cnt,
                                                  int *__restrict__ grp_nitem,
                                                  int *__restrict__ meta_w, int cand_cap,
                                                  int isteps) {
  constexpr uint32_t kFull = (1u << kGroup) - 1u;
  const int icell = blockIdx.x;
  const int lane = threadIdx.x & 63, l = lane & 31, half = lane >> 5;
  const int h = threadIdx.x >> 6;  // one wave per group
  int cnt = 0;
  if (icell < meta[0]) {
    const int s = icell_start[icell], e = icell_start[icell + 1];
    // the entry format indexes at most kMaxCellTiles tiles per cell
    if (e - s > kMaxCellTiles && threadIdx.x == 0) atomicOr(&meta_w[7], 1);
    for (int k0 = s; k0 < e; k0 += 2) {
      const int k = k0 + half;
      const uint32_t m = k < e && k < cand_cap ? cand_mask[(size_t)k * kCell + l] : 0xffffffffu;
      cnt += __popcll(__ballot(((m >> (kGroup * h)) & kFull) != kFull));
    }
  }
  if (lane == 0) {
    const int c = (cnt + 63) / 64 * 64;
    grp_cnt[icell * kGroupsPerCell + h] = c;
    grp_nitem[icell * kGroupsPerCell + h] = (c + 64 * isteps - 1) / (64 * isteps);
  }
}

// 9. entries (deterministic order: candidate tiles in j order, slots
// ascending), null padding (full skip mask on the cell's own first slot) and
// work items.
__global__ __launch_bounds__(64 * kGroupsPerCell) void k_grp_emit(
    const int *__restrict__ meta, const int *__restrict__ icell_start,
    const uint32_t *__restrict__ cand_mask, const int *__restrict__ grp_start,
    const int *__restrict__ grp_itemstart, uint32_t *__restrict__ ent, int4 *__restrict__ items,
    int isteps) {
  constexpr uint32_t kFull = (1u << kGroup) - 1u;
  const int icell = blockIdx.x;
  if (icell >= meta[0]) return;
  const int lane = threadIdx.x & 63, l = lane & 31, half = lane >> 5;
  const int h = threadIdx.x >> 6;  // one wave per group
  if (meta[8]) return;  // the list overflowed a capacity: k_list_guard_export emptied it
  const int s = icell_start[icell], e = icell_start[icell + 1];
  if (e - s > kMaxCellTiles) return;  // flagged in k_grp_count; the host raises
  const int g = icell * kGroupsPerCell + h;
  const int st = grp_start[g], en = grp_start[g + 1];
  int run = 0;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int k0 = s; k0 < e; k0 += 2) {
    const int k = k0 + half;
    const uint32_t m = k < e ? cand_mask[(size_t)k * kCell + l] : 0xffffffffu;
    const uint32_t mg = (m >> (kGroup * h)) & kFull;
    const bool live = mg != kFull;
    const unsigned long long b = __ballot(live);
    if (live) ent[st + run + __popcll(b & below)] = (uint32_t)((k - s) << 5 | l) | (mg << 16);
    run += __popcll(b);
  }
  for (int p = st + run + lane; p < en; p += 64) ent[p] = kFull << 16;
  const int i0 = grp_itemstart[g], i1 = grp_itemstart[g + 1];
  for (int it = i0 + lane; it < i1; it += 64) {
    const int a = st + (it - i0) * 64 * isteps;
    items[it] = make_int4(g, a, min(a + 64 * isteps, en), s);
  }
}

// 10. Processing order of the work items for the pair kernel: by step count,
// longest first, items of one length in index order (a stable counting sort
// in one block).  The kernel's blocks take items in this order, so the items
// still running when the launch drains are short ones (list scheduling,
// longest processing time first).  order[p] = item index.
constexpr int kOrderThreads = 1024;
__global__ __launch_bounds__(kOrderThreads) void k_item_order(const int *__restrict__ meta,
                                                              const int4 *__restrict__ items,
                                                              int *__restrict__ order,
                                                              int isteps) {
  constexpr int kW = kOrderThreads / 64;
  __shared__ int cnt[kItemSteps + 1], run[kItemSteps + 1];
  __shared__ int wcnt[kW][kItemSteps + 1];
  const int n = meta[3];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  auto len_of = [&](int i) {
    const int4 it = items[i];
    return min(max((it.z - it.y) >> 6, 1), isteps);
  };
  if (tid <= kItemSteps) cnt[tid] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += kOrderThreads) atomicAdd(&cnt[len_of(i)], 1);
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int L = isteps; L >= 1; L--) {
      run[L] = acc;
      acc += cnt[L];
    }
  }
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += kOrderThreads) {
    const int i = c0 + tid;
    const int L = i < n ? len_of(i) : 0;
    int rank = 0;
    for (int b = 1; b <= isteps; b++) {
      const unsigned long long m = __ballot(L == b);
      if (L == b) rank = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) wcnt[w][b] = __popcll(m);
    }
    __syncthreads();
    if (i < n) {
      int p = run[L] + rank;
      for (int v = 0; v < w; v++) p += wcnt[v][L];
      order[p] = i;
    }
    __syncthreads();
    if (tid >= 1 && tid <= isteps) {
      int t = 0;
      for (int v = 0; v < kW; v++) t += wcnt[v][tid];
      run[tid] += t;
    }
    __syncthreads();
  }
}

// List pruning (no counterpart in the reference; the idea of GROMACS's
// dynamic pruning).  Between builds, every K steps, each group's entries of
// the build's list are tested against the current coordinates with the
// radius r_p = r_off + (cutnb - r_off) K / nbfreq, which assumes the same
// largest relative motion per step as the build's own buffer; pairs beyond
// r_p gain their skip bit, entries left with no live pair are dropped, the
// rest are packed to the front of the group's region of ent_p (same region
// as in ent, null padding after them) and the group's items are shortened.
// Every index stays inside the group's region, so an item that becomes
// empty still points at valid entries.  One wave per group.
__global__ __launch_bounds__(64 * kGroupsPerCell) void k_prune(
    const int *__restrict__ meta, const int *__restrict__ icell_start,
    const int *__restrict__ grp_start, const int *__restrict__ grp_itemstart,
    const int *__restrict__ cand_j, const int4 *__restrict__ posq,
    const uint32_t *__restrict__ ent, const int4 *__restrict__ items, float3 unit, float rp2,
    uint32_t *__restrict__ ent_p, int4 *__restrict__ items_p) {
  constexpr uint32_t kFull = (1u << kGroup) - 1u;
  const int icell = blockIdx.x;
  if (icell >= meta[0] || meta[8]) return;
  const int lane = threadIdx.x & 63, h = threadIdx.x >> 6;
  const int g = icell * kGroupsPerCell + h;
  const int st = grp_start[g], en = grp_start[g + 1];
  const int *tiles = cand_j + icell_start[icell];
  const int4 pil = posq[g * kGroup + (lane & (kGroup - 1))];
  unsigned pix[kGroup], piy[kGroup], piz[kGroup];
#pragma unroll
  for (int k = 0; k < kGroup; k++) {
    pix[k] = (unsigned)__builtin_amdgcn_readlane(pil.x, k);
    piy[k] = (unsigned)__builtin_amdgcn_readlane(pil.y, k);
    piz[k] = (unsigned)__builtin_amdgcn_readlane(pil.z, k);
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  int run = 0;
  // the next chunk's entry and j coordinates load while this one is tested
  auto jpos = [&](uint32_t e) { return posq[tiles[(e & 0xffffu) >> 5] * kCell + (int)(e & 31u)]; };
  uint32_t e_n = st < en ? ent[st + lane] : 0u;
  int4 pj_n = st < en ? jpos(e_n) : make_int4(0, 0, 0, 0);
  for (int p0 = st; p0 < en; p0 += 64) {  // en - st is a multiple of 64
    const uint32_t e = e_n;
    const int4 pj = pj_n;
    if (p0 + 64 < en) {
      e_n = ent[p0 + 64 + lane];
      pj_n = jpos(e_n);
    }
    uint32_t m = e >> 16;
#pragma unroll
    for (int k = 0; k < kGroup; k++) {
      const float dx = (float)(int)(pix[k] - (unsigned)pj.x) * unit.x;
      const float dy = (float)(int)(piy[k] - (unsigned)pj.y) * unit.y;
      const float dz = (float)(int)(piz[k] - (unsigned)pj.z) * unit.z;
      if (!(dx * dx + dy * dy + dz * dz < rp2)) m |= 1u << k;
    }
    const bool live = m != kFull;
    const unsigned long long b = __ballot(live);
    if (live) ent_p[st + run + __popcll(b & below)] = (e & 0xffffu) | (m << 16);
    run += __popcll(b);
  }
  for (int p = st + run + lane; p < en; p += 64) ent_p[p] = kFull << 16;
  const int end = st + (run + 63) / 64 * 64;
  for (int it = grp_itemstart[g] + lane; it < grp_itemstart[g + 1]; it += 64) {
    int4 q = items[it];
    q.z = max(q.y, min(q.z, end));
    items_p[it] = q;
  }
}

void nblist_prune(NbSystem &S, hipStream_t s) {
  const float3 unit = make_float3((float)(S.box[0] / 4294967296.0), (float)(S.box[1] / 4294967296.0),
                                  (float)(S.box[2] / 4294967296.0));
  k_prune<<<S.ncell_cap, 64 * kGroupsPerCell, 0, s>>>(
      S.meta.get(), S.icell_start.get(), S.grp_start.get(), S.grp_itemstart.get(),
      S.cand_j.get(), S.posq.get(), S.ent.get(), S.items.get(), unit, S.prune_r2, S.ent_p.get(),
      S.items_p.get());
  APO_LAUNCH_CHECK();
}

// 11. After the counts are known: any capacity overflow or input error
// leaves the list empty (no pair kernel work, no out-of-range write) and sets
// meta[8]; then the build's counts and flags go to pinned host memory with
// the running build number in slot 15 (meta[16] survives the per-build
// clear), where the host learns of errors and growth.
__global__ void k_list_guard_export(int *__restrict__ meta, int cand_cap, int ent_cap,
                                    int item_cap, int *__restrict__ host) {
  if (threadIdx.x != 0) return;
  const bool over = meta[1] > cand_cap || meta[2] > ent_cap || meta[3] > item_cap;
  if (over || meta[5] || meta[6] || meta[7]) {
    meta[8] = over ? 1 : 2;
    meta[3] = 0;
  }
  const int seq = meta[16] + 1;
  meta[16] = seq;
  for (int k = 0; k < 15; k++) host[k] = meta[k];
  __threadfence_system();
  host[15] = seq;
  __threadfence_system();
}

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// Host side of the exported counts: errors of a finished build raise here;
// with allow_growth (only right before a rebuild, whose lists replace the
// current ones) capacities that the last finished build used above 85 % grow:
// the device is drained first and captured step graphs see the new
// generation.  Without it a pending growth stays pending for the next check.
void list_export_check(NbSystem &S, hipStream_t s, bool allow_growth) {
  const volatile int *h = S.h_meta;
  if (!h || h[15] == 0) return;
  int m[16];
  for (int k = 0; k < 16; k++) m[k] = h[k];
  if (m[15] == S.meta_checked) return;
  const bool grow = m[1] > 0.85 * S.cand_cap || m[2] > 0.85 * S.ent_cap || m[3] > 0.85 * S.item_cap;
  const bool error = m[5] || m[6] || m[7] || m[8] == 1;
  if (!error && grow && !allow_growth) return;
  S.meta_checked = m[15];
  if (m[6]) throw RuntimeError("non-finite coordinates in xyzq at a neighbor-list rebuild");
  if (m[5]) throw RuntimeError("neighbor list: an xy column exceeds the LDS sort capacity");
  if (m[7]) throw RuntimeError("neighbor list: a cell has more than 2048 candidate tiles");
  if (m[8] == 1)
    throw RuntimeError("neighbor list: a rebuild exceeded the list capacity; the forces of the "
                       "steps after it are invalid (capacity was sized from earlier builds)");
  if (grow) {
    APO_HIP(hipStreamSynchronize(s));
    APO_HIP(hipDeviceSynchronize());
    list_reserve(S, m[1], m[2], m[3]);
  }
}

void list_reserve(NbSystem &S, int ncand, int nent, int nitem) {
  if ((size_t)ncand > S.cand_cap * 0.85 || S.cand_cap == 0) {
    S.cand_cap = (size_t)(ncand * 1.5) + 64;
    S.cand_i.reserve(S.cand_cap);
    S.cand_j.reserve(S.cand_cap);
    S.cand_mask.reserve(S.cand_cap * kCell);
    S.generation++;
  }
  if ((size_t)nent > S.ent_cap * 0.85 || S.ent_cap == 0) {
    S.ent_cap = (size_t)(nent * 1.5) + 64 * kItemSteps;
    S.ent.reserve(S.ent_cap);
    S.ent_p.reserve(S.ent_cap);
    S.generation++;
  }
  if ((size_t)nitem > S.item_cap * 0.85 || S.item_cap == 0) {
    S.item_cap = (size_t)(nitem * 1.5) + 64;
    S.items.reserve(S.item_cap);
    S.items_p.reserve(S.item_cap);
    S.item_order.reserve(S.item_cap);
    S.generation++;
  }
}

// Builds the list from xyzq.  sizing = true synchronises twice to size the
// buffers exactly and raises input errors at once (the first build and
// apo_nb_reset_list); otherwise the build is host-sync free, guarded by the
// capacities, and its errors surface at the next check of the exported
// counts (so a rebuild can be part of a captured step graph).
void nblist_build(NbSystem &S, const float4 *x, hipStream_t s, bool sizing) {
  const int n = S.natom;
  const double3 inv = make_double3(1.0 / S.box[0], 1.0 / S.box[1], 1.0 / S.box[2]);
  const double3 len = make_double3(S.box[0], S.box[1], S.box[2]);
  const float3 unit = make_float3((float)(S.box[0] / 4294967296.0), (float)(S.box[1] / 4294967296.0),
                                  (float)(S.box[2] / 4294967296.0));
  const float3 Lf = make_float3((float)S.box[0], (float)S.box[1], (float)S.box[2]);
  const float rl2 = (float)(S.cutnb * S.cutnb);
  int *meta = S.meta.get();
  sizing = sizing || S.cand_cap == 0 || S.ent_cap == 0 || S.item_cap == 0;
  if (sizing && !S.may_allocate)
    throw InvalidArgument("the neighbor list has no capacity yet: build it once with an "
                          "allocating call (apo_nb_reset_list or apo_nb_calc_force)");
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  APO_HIP(hipStreamIsCapturing(s, &cap));
  if (sizing && cap != hipStreamCaptureStatusNone)
    throw RuntimeError("the first neighbor-list build cannot be captured in a graph");
  if (!sizing && cap == hipStreamCaptureStatusNone) list_export_check(S, s, S.may_allocate);

  launch_fill_i32(meta, 16, s);
  launch_fill_i32(S.col_count.get(), S.ncol + 1, s);
  k_assign<<<nblk(n, 256), 256, 0, s>>>(x, n, inv, len, S.ncx, S.ncy, S.col_count.get(),
                                         S.atom_col.get(), S.atom_slot.get(), S.zkey.get(), meta,
                                         S.xref.get());
  APO_LAUNCH_CHECK();
  k_col_prep<<<nblk(S.ncol, 256), 256, 0, s>>>(S.col_count.get(), S.ncol, S.col_ncell.get(), meta);
  APO_LAUNCH_CHECK();
  // column starts of the atoms and of the cells; the cell total is meta[0]
  launch_scan2(S.col_count.get(), S.col_start.get(), nullptr, S.ncol, nullptr, S.col_ncell.get(),
               S.col_cellstart.get(), nullptr, S.ncol, meta + 0, s);
  unsigned long long *kb = S.colkeys.get();
  k_scatter<<<nblk(n, 256), 256, 0, s>>>(n, S.atom_col.get(), S.atom_slot.get(), S.col_start.get(),
                                          S.zkey.get(), kb);
  APO_LAUNCH_CHECK();
  k_colsort<<<S.ncol, 256, 0, s>>>(kb, S.col_count.get(), S.col_start.get(), S.col_cellstart.get(),
                                   S.loc2glo.get(), S.glo2loc.get(), meta);
  APO_LAUNCH_CHECK();
  nblist_sort_coords(S, x, s);  // also the cell boxes
  k_tiles<<<S.ncell_cap, 64, 0, s>>>(S.cell_bb.get(), meta, Lf, rl2, S.icell_cnt.get(), nullptr,
                                     nullptr, nullptr, 0, 0);
  APO_LAUNCH_CHECK();
  launch_scan(S.icell_cnt.get(), S.icell_start.get(), meta + 0, S.ncell_cap, s, meta + 1);
  int hmeta[8];
  if (sizing) {
    APO_HIP(hipMemcpyAsync(hmeta, meta, sizeof(hmeta), hipMemcpyDeviceToHost, s));
    APO_HIP(hipStreamSynchronize(s));
    if (hmeta[6]) {
      S.list_valid = false;
      throw RuntimeError("non-finite coordinates in xyzq");
    }
    if (hmeta[5]) {
      S.list_valid = false;
      throw RuntimeError("neighbor list: an xy column exceeds the LDS sort capacity");
    }
    S.host_ncell = hmeta[0];
    S.host_ncand = hmeta[1];
    list_reserve(S, hmeta[1], 0, 0);
    if (S.item_steps == 0) {
      // the longest items that still give every other pair-kernel wave one
      // (measured: 98,304 atoms keep 16 steps, 23,558 atoms run faster on 4);
      // a candidate tile holds ~32 live entries, half a 64-entry step
      const double steps = 0.5 * hmeta[1];
      int k = kItemSteps;
      while (k > 2 && steps / k < 0.5 * S.pair_blocks) k /= 2;
      S.item_steps = k;
    }
  }
  if (S.item_steps == 0) S.item_steps = kItemSteps;
  const int ccap = (int)S.cand_cap;
  k_tiles<<<S.ncell_cap, 64, 0, s>>>(S.cell_bb.get(), meta, Lf, rl2, nullptr, S.icell_start.get(),
                                     S.cand_i.get(), S.cand_j.get(), 1, ccap);
  APO_LAUNCH_CHECK();
  {
    // cubic boxes compare in fixed-point units (the pair kernel's form)
    const bool iso = S.box[0] == S.box[1] && S.box[1] == S.box[2];
    const double u = S.box[0] / 4294967296.0;
    const float rl2u = (float)(S.cutnb * S.cutnb / (u * u));
    if (iso)
      k_tile_mask<true><<<S.ncell_cap, 64 * kMaskWaves, 0, s>>>(
          meta, S.icell_start.get(), S.cand_j.get(), S.posq.get(), S.loc2glo.get(),
          S.glo2loc.get(), S.excl_ptr.get(), S.excl_idx.get(), unit, rl2, rl2u, S.cand_mask.get(),
          ccap);
    else
      k_tile_mask<false><<<S.ncell_cap, 64 * kMaskWaves, 0, s>>>(
          meta, S.icell_start.get(), S.cand_j.get(), S.posq.get(), S.loc2glo.get(),
          S.glo2loc.get(), S.excl_ptr.get(), S.excl_idx.get(), unit, rl2, rl2u, S.cand_mask.get(),
          ccap);
  }
  APO_LAUNCH_CHECK();
  const int ngrp = S.ncell_cap * kGroupsPerCell;
  k_grp_count<<<S.ncell_cap, 64 * kGroupsPerCell, 0, s>>>(
      meta, S.icell_start.get(), S.cand_mask.get(), S.grp_cnt.get(), S.grp_nitem.get(), meta, ccap,
      S.item_steps);
  APO_LAUNCH_CHECK();
  launch_scan2(S.grp_cnt.get(), S.grp_start.get(), nullptr, ngrp, meta + 2, S.grp_nitem.get(),
               S.grp_itemstart.get(), nullptr, ngrp, meta + 3, s);
  if (sizing) {
    APO_HIP(hipMemcpyAsync(hmeta, meta, sizeof(hmeta), hipMemcpyDeviceToHost, s));
    APO_HIP(hipStreamSynchronize(s));
    if (hmeta[7]) {
      S.list_valid = false;
      throw RuntimeError("neighbor list: a cell has more than 2048 candidate tiles");
    }
    S.host_nent = hmeta[2];
    S.host_nitem = hmeta[3];
    list_reserve(S, 0, hmeta[2], hmeta[3]);
  }
  k_list_guard_export<<<1, 64, 0, s>>>(meta, ccap, (int)S.ent_cap, (int)S.item_cap, S.h_meta);
  APO_LAUNCH_CHECK();
  k_grp_emit<<<S.ncell_cap, 64 * kGroupsPerCell, 0, s>>>(meta, S.icell_start.get(), S.cand_mask.get(),
                                        S.grp_start.get(), S.grp_itemstart.get(), S.ent.get(),
                                        S.items.get(), S.item_steps);
  APO_LAUNCH_CHECK();
  k_item_order<<<1, kOrderThreads, 0, s>>>(meta, S.items.get(), S.item_order.get(), S.item_steps);
  APO_LAUNCH_CHECK();
  S.builds_queued++;
  S.list_valid = true;
}

void nblist_sort_coords(NbSystem &S, const float4 *x, hipStream_t s, int64_t *force,
                        size_t nforce) {
  const double3 inv = make_double3(1.0 / S.box[0], 1.0 / S.box[1], 1.0 / S.box[2]);
  const float qscale = 1.0f;  // raw charge; ccelec is applied per tile in the pair kernel
  const float3 unit = make_float3((float)(S.box[0] / 4294967296.0), (float)(S.box[1] / 4294967296.0),
                                  (float)(S.box[2] / 4294967296.0));
  const int with_fin = S.sort_fin ? 1 : 0;
  const LtFinArgs fin = S.sort_fin ? *S.sort_fin : LtFinArgs{};
  k_sort_coords<<<nblk(S.ssort, 256) + with_fin, 256, 0, s>>>(
      x, S.vdwtype.get(), S.loc2glo.get(), S.meta.get(), inv, qscale, unit, S.posq.get(),
      S.type_s.get(), S.fsort.get(), S.ssort, S.cell_bb.get(), force, nforce, fin, with_fin);
  APO_LAUNCH_CHECK();
}

// Pairs inside the list cutoff covered by the list (clear skip-mask bits).
__global__ void k_count_pairs(const uint32_t *__restrict__ ent, int nent,
                              unsigned long long *__restrict__ out) {
  size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  unsigned c = 0;
  if (k < (size_t)nent) c = __popc(~(ent[k] >> 16) & ((1u << kGroup) - 1u));
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, (unsigned long long)c);
}

void nblist_stats(NbSystem &S, int64_t *out6, hipStream_t s) {
  int hmeta[16];
  APO_HIP(hipMemcpyAsync(hmeta, S.meta.get(), sizeof(hmeta), hipMemcpyDeviceToHost, s));
  APO_HIP(hipStreamSynchronize(s));
  list_export_check(S, s, false);
  const int nent = hmeta[8] ? 0 : std::min(hmeta[2], (int)S.ent_cap);
  DevBuf<unsigned long long> cnt;
  cnt.reserve(1);
  APO_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(unsigned long long), s));
  if (S.list_valid && nent > 0)
    k_count_pairs<<<nblk((long)nent, 256), 256, 0, s>>>(S.ent.get(), nent, cnt.get());
  APO_LAUNCH_CHECK();
  unsigned long long pairs = 0;
  APO_HIP(hipMemcpyAsync(&pairs, cnt.get(), sizeof(pairs), hipMemcpyDeviceToHost, s));
  APO_HIP(hipStreamSynchronize(s));
  out6[0] = hmeta[0];
  out6[1] = hmeta[2];
  out6[2] = hmeta[3];
  out6[3] = (int64_t)pairs;
  out6[4] = S.ncol;
  out6[5] = hmeta[4];
}

// ---- fused launches over several systems ------------------------------------
void batch_add_sort(BatchLaunch<SortArgs> &b, NbSystem &S, const float4 *x, int64_t *force,
                    size_t nforce) {
  SortArgs a{x, S.vdwtype.get(), S.loc2glo.get(), S.meta.get(),
             make_double3(1.0 / S.box[0], 1.0 / S.box[1], 1.0 / S.box[2]),
             make_float3((float)(S.box[0] / 4294967296.0), (float)(S.box[1] / 4294967296.0),
                         (float)(S.box[2] / 4294967296.0)),
             S.posq.get(), S.type_s.get(), S.fsort.get(), S.ssort, S.cell_bb.get(), force,
             nforce};
  b.add(a, nblk(S.ssort, 256));
}
void batch_sort(BatchLaunch<SortArgs> &b, hipStream_t s) {
  if (!b.blocks()) return;
  kb_sort_coords<<<b.blocks(), 256, 0, s>>>(b.d_args.get(), b.d_start.get(), b.nsys());
  APO_LAUNCH_CHECK();
}

}  // namespace apo
