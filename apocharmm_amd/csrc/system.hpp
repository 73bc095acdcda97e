// One system's device state: the MI355X counterpart of what ForceManager
// owns on the force path (CudaPMEDirectForce + CudaNeighborList,
// CudaPMEReciprocalForce, CudaBondedForce, CudaEnergyVirial).
#pragma once

#include "apocharmm_amd.h"
#include "common.hpp"

namespace apo {

struct LtFinArgs;
struct BondT { int i, j; float b0, kb; };
struct AngleT { int i, j, k; float t0, kt; };
struct DiheT { int i, j, k, l; float n, kc, sd, cd; float delta; int pad[3]; };

// FFT plan of one dimension: radices and twiddles exp(-2 pi i e / n).
struct FftDim {
  int n = 0;
  int nrad = 0;
  int rad[16] = {0};
  DevBuf<float2> tw;
  DevBuf<double2> tw64;
};

struct NbSystem {
  // ---- parameters ----------------------------------------------------------
  int natom = 0;
  double box[3] = {0, 0, 0};
  double kappa = 0, cutnb = 0, ctofnb = 0, ctonnb = 0, e14fac = 1;
  int K[3] = {0, 0, 0};
  int order = 4;
  int precision = 0;
  int device = 0;
  int ncu = 256;

  // ---- topology ------------------------------------------------------------
  int ntypes = 0;
  DevBuf<int> vdwtype;
  DevBuf<float2> vdwtab;      // (6*c6, 12*c12)
  DevBuf<float2> vdwtab14;
  DevBuf<double2> vdwtab_d;   // FP64 instantiation
  DevBuf<double2> vdwtab14_d;
  DevBuf<float2> vdwsq;       // square (ntypes x ntypes) copy for <= 32 types
  DevBuf<double2> vdwsq_d;
  // per-type (sqrt(eps), Rmin/2) when every table entry follows the CHARMM
  // combination rule (no NBFIX among the types present): the pair kernel
  // then combines per pair instead of reading a table too large for LDS
  bool lj_combine = false;
  DevBuf<float2> ljtype;
  // Switched vdW force in one form for both regions, per type pair (few
  // types): r dU/dr / r^2 contribution = (P r^-6 - Q r^-3 - A12 r^-12) r^-2,
  // with (P, Q, A12) = (6c6, 0, 12c12) inside r_on and (k6 6c6 + k12 12c12
  // r_off^-6, k6 6c6 r_off^-3, k12 12c12) beyond it; [ti][tj][inner | outer]
  DevBuf<float4> vdwpqa;
  DevBuf<int> excl_ptr, excl_idx;
  int nin14 = 0, nex14 = 0;
  DevBuf<int2> in14, ex14;
  int nbond = 0, nub = 0, nangle = 0, ndihe = 0, nimdihe = 0;
  DevBuf<BondT> bonds, ureyb;
  DevBuf<AngleT> angles;
  DevBuf<DiheT> dihe, imdihe;

  // ---- neighbor list (cell sort + tiles) ------------------------------------
  int ncx = 1, ncy = 1, ncol = 1;
  int ncell_cap = 0;  // ceil(natom/32) + ncol + 1
  bool list_valid = false;
  // bumped whenever a buffer or a launch constant of the per-step path
  // changes (list buffers regrown, new box); captured step graphs compare it
  unsigned long long generation = 0;
  // Launch settings that change no result bits (launch constants of captured
  // graphs; the integrator's first list cycles pick them by timing):
  // pme_prio 1 = the PME kernels raise their wave issue priority;
  // pair_first 1 = the pair kernel is launched (captured) before PME and
  // bonded, so a graph dispatches it first from the caller's queue.
  int pme_prio = 0;
  int pair_first = 0;
  // integrators bound to this system (it must not be destroyed before them)
  int n_integrators = 0;
  std::vector<int> h_bonds;  // 2 * nbond, host copy for molecule grouping
  DevBuf<int> meta;   // [0] ncell [1] ncand [2] nent [3] nitem [4] maxcol
  DevBuf<int> col_count, col_start, col_ncell, col_cellstart;
  DevBuf<int> atom_col, atom_slot;
  DevBuf<unsigned long long> colkeys;
  DevBuf<float> zkey;
  DevBuf<int> loc2glo, glo2loc;
  DevBuf<float4> cell_bb;  // 2 per cell: center, half extent
  DevBuf<int> icell_cnt, icell_start;
  DevBuf<int> cand_i, cand_j;
  DevBuf<uint32_t> cand_mask;
  // pair list: per i-group (kGroup consecutive slots of a cell) the j slots
  // with at least one live pair, packed as (candidate tile index << 5 | slot)
  // | skip-mask << 16 and padded to whole 64-lane steps; items = (group,
  // first entry, end entry, first candidate tile of the group's cell)
  DevBuf<int> grp_cnt, grp_start, grp_nitem, grp_itemstart;
  DevBuf<uint32_t> ent;
  DevBuf<int4> items;
  DevBuf<int> item_order;  // items by step count, longest first (pair-kernel order)
  // pruned copy of the list (same layout and regions), read by the pair
  // kernel in calc_force calls with kCalcUsePruned; pruned_build is the build
  // (builds_queued) it was last derived from, kept by its user
  DevBuf<uint32_t> ent_p;
  DevBuf<int4> items_p;
  float prune_r2 = 0.f;
  unsigned long long pruned_build = ~0ull;
  size_t cand_cap = 0, ent_cap = 0, item_cap = 0;
  int host_ncell = 0, host_ncand = 0, host_nent = 0, host_nitem = 0;
  // 64-entry steps per pair-kernel work item (power of two, 2..kItemSteps),
  // fixed by the first sizing build so the items roughly fill the pair
  // kernel's waves (small systems get short items)
  int item_steps = 0;
  // set by an integrator around its force calls: the coordinate sort runs
  // that integrator's pending kinetic-energy reduction in one extra block
  const LtFinArgs *sort_fin = nullptr;
  // counts and flags of each build, exported to pinned host memory (slot 15
  // = build number) and checked lazily on the host
  int *h_meta = nullptr;
  int meta_checked = 0;
  // list builds enqueued so far; the device numbers its builds the same way
  // (meta[16], exported with the counts)
  unsigned long long builds_queued = 0;
  DevBuf<float4> xref;  // coordinates of the last list build (global order)

  // ---- per-step direct-space buffers -----------------------------------------
  DevBuf<int4> posq;       // fixed-point coords (x,y,z as uint32) + scaled q
  DevBuf<int> type_s;
  DevBuf<int64_t> fsort;   // sorted-order gradient, 3 * ssort
  int ssort = 0;           // stride of fsort (= ncell_cap*32)
  int pair_blocks = 0;

  // ---- PME ------------------------------------------------------------------
  FftDim fx, fy, fz;
  DevBuf<float2> cgrid;      // (K0/2+1)*K1*K2 complex half grid
  DevBuf<float> phi;         // K0*K1*K2 potential grid
  DevBuf<double2> cgrid_d;   // FP64 instantiation
  DevBuf<double> phi_d;
  DevBuf<float> prefac;      // K0 + K1 + K2 (1/|b|^2 per axis)
  DevBuf<double> prefac_d;
  int zconv_blocks = 0;

  // ---- energies / virial --------------------------------------------------------
  PartLayout part;
  DevBuf<double> parts;
  // when set, the partials live here instead (the no-allocation entry point
  // lends its static device array for the length of one call)
  double *parts_ext = nullptr;
  double *parts_ptr() const { return parts_ext ? parts_ext : parts.get(); }
  // false: list builds must fit the current capacities (no allocation)
  bool may_allocate = true;

  // calc_force runs PME and the bonded/1-4 terms on two side streams that
  // fork from and join back into the caller's stream
  hipStream_t side[2] = {nullptr, nullptr};
  bool single_stream = false;  // everything on the caller's stream (batched systems)
  hipEvent_t ev_fork = nullptr, ev_sorted = nullptr, ev_pme = nullptr, ev_bond = nullptr;

  // HIP-event kernel timers (apo_nb_profile)
  struct TimerRec { int slot; hipEvent_t a, b; };
  bool prof_on = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  int prof_open = 0;  // tics not yet closed by their toc
  std::vector<TimerRec> prof_recs;
  double prof_ms[APO_PROFILE_SLOTS] = {0};
  long long prof_cnt[APO_PROFILE_SLOTS] = {0};
  hipEvent_t prof_event();
  void tic(int slot, hipStream_t s, hipEvent_t &a);
  void toc(int slot, hipStream_t s, hipEvent_t a);
  void profile_flush();
  ~NbSystem();

  PairConsts<float> pc_f;
  PairConsts<double> pc_d;

  void make_pair_consts();
  void setup_pme();
  void setup_columns();
};

}  // namespace apo
