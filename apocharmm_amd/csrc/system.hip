// System set-up and the calc_force orchestration (ForceManager::calcForce
// parts 1-3, src/ForceManager.cu:557-819, re-designed: one stream, no host
// synchronisation, device-side fixed-order reductions).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "system.hpp"

namespace apo {

// ---------------------------------------------------------------------------
// erfc approximation for the FP32 pair kernel.
// erfc(x) = exp(-x^2) erfcx(x); erfcx is interpolated at Chebyshev nodes in
// t = 1/(1 + p x) on [0, xfit] (xfit >= kappa*r_off), then expanded to a
// monomial series in t evaluated by Horner in the kernel.  Project-fitted at
// set-up for the system's own interval (no third-party constants).
// ---------------------------------------------------------------------------
static double erfcx_host(double x) { return std::erfc(x) * std::exp(x * x); }

static void fit_erfc(double xfit, double p, double *coef /*9*/) {
  const int deg = 8, nn = deg + 1;
  double tmin = 1.0 / (1.0 + p * xfit);
  // Chebyshev coefficients of f(u), u in [-1,1] <-> t = tmin + (u+1)(1-tmin)/2
  double c[nn];
  for (int k = 0; k < nn; k++) {
    double s = 0;
    for (int j = 0; j < nn; j++) {
      double u = std::cos(kPi * (j + 0.5) / nn);
      double t = tmin + (u + 1.0) * (1.0 - tmin) * 0.5;
      double x = (1.0 / t - 1.0) / p;
      s += erfcx_host(x) * std::cos(kPi * k * (j + 0.5) / nn);
    }
    c[k] = (k == 0 ? 1.0 : 2.0) * s / nn;
  }
  // Chebyshev -> monomial in u (T_{k+1} = 2u T_k - T_{k-1})
  double Tprev[nn] = {0}, Tcur[nn] = {0}, pu[nn] = {0};
  Tprev[0] = 1.0;            // T0
  Tcur[1] = 1.0;             // T1
  pu[0] += c[0];
  for (int i = 0; i < nn; i++) pu[i] += c[1] * Tcur[i];
  for (int k = 2; k < nn; k++) {
    double Tn[nn] = {0};
    for (int i = 0; i < nn; i++) {
      if (i > 0) Tn[i] += 2.0 * Tcur[i - 1];
      Tn[i] -= Tprev[i];
    }
    for (int i = 0; i < nn; i++) {
      pu[i] += c[k] * Tn[i];
      Tprev[i] = Tcur[i];
      Tcur[i] = Tn[i];
    }
  }
  // compose with u = a t + b
  double a = 2.0 / (1.0 - tmin), b = -2.0 * tmin / (1.0 - tmin) - 1.0;
  double pt[nn] = {0};
  double powab[nn] = {0};  // (a t + b)^k coefficients
  powab[0] = 1.0;
  for (int k = 0; k < nn; k++) {
    for (int i = 0; i < nn; i++) pt[i] += pu[k] * powab[i];
    double nxt[nn] = {0};
    for (int i = 0; i < nn; i++) {
      nxt[i] += b * powab[i];
      if (i + 1 < nn) nxt[i + 1] += a * powab[i];
    }
    std::memcpy(powab, nxt, sizeof(nxt));
  }
  for (int i = 0; i < nn; i++) coef[i] = pt[i];
}

void NbSystem::make_pair_consts() {
  double roff = ctofnb, ron = ctonnb;
  double roff3 = roff * roff * roff, ron3 = ron * ron * ron;
  double roff6 = roff3 * roff3, ron6 = ron3 * ron3;
  double k6, k12, dv6, dv12;
  if (ron < roff) {
    k6 = roff3 / (roff3 - ron3);
    k12 = roff6 / (roff6 - ron6);
    dv6 = -1.0 / (ron3 * roff3);
    dv12 = -1.0 / (ron6 * roff6);
  } else {
    k6 = 1.0;
    k12 = 1.0;
    dv6 = -1.0 / roff6;
    dv12 = -1.0 / (roff6 * roff6);
  }
  double s32 = 4294967296.0;
  double ep = 0.47, coef[9];
  fit_erfc(std::max(kappa * roff * 1.01, 4.0), ep, coef);
  auto fill = [&](auto &pc) {
    using T = std::remove_reference_t<decltype(pc.sx)>;
    pc.sx = (T)(box[0] / s32);
    pc.sy = (T)(box[1] / s32);
    pc.sz = (T)(box[2] / s32);
    const double u = box[0] / s32;
    pc.s2 = (T)(u * u);
    pc.roff2u = (T)(roff * roff / (u * u));
    pc.roff2 = (T)(roff * roff);
    pc.ron2 = (T)(ron * ron);
    pc.rlist2 = (T)(cutnb * cutnb);
    pc.k6 = (T)k6;
    pc.k12 = (T)k12;
    pc.dv6 = (T)dv6;
    pc.dv12 = (T)dv12;
    pc.roffinv3 = (T)(1.0 / roff3);
    pc.roffinv6 = (T)(1.0 / roff6);
    pc.kappa = (T)kappa;
    pc.kappa2 = (T)(kappa * kappa);
    pc.k2log2e = (T)(kappa * kappa * 1.4426950408889634);
    pc.two_kappa_sqrtpi = (T)(2.0 * kappa / std::sqrt(kPi));
    pc.erfc_p = (T)(ep * kappa);
    for (int i = 0; i < 9; i++) pc.erfc_c[i] = (T)coef[i];
  };
  fill(pc_f);
  fill(pc_d);
}

// ---------------------------------------------------------------------------
// PME set-up: Euler-spline moduli, FFT plans, grids.
// ---------------------------------------------------------------------------

// Cardinal B-spline values at the integer knots for offset w = 0 (order-n
// recursion), used by the moduli below.
static void bspline_knots(int order, std::vector<double> &m) {
  // M_2(x) on knots, then M_n(x) = (x M_{n-1}(x) + (n-x) M_{n-1}(x-1))/(n-1)
  // evaluated at x = 1..order-1 (M_n vanishes at 0 and n).
  std::vector<double> prev = {0.0, 1.0, 0.0};  // M_2 at x = 0,1,2
  for (int n = 3; n <= order; n++) {
    std::vector<double> cur(n + 1, 0.0);
    for (int x = 0; x <= n; x++) {
      double a = (x <= n - 1) ? prev[x] : 0.0;
      double b = (x >= 1 && x - 1 <= n - 1) ? prev[x - 1] : 0.0;
      cur[x] = (x * a + (n - x) * b) / (n - 1);
    }
    prev = cur;
  }
  m = prev;  // M_order at x = 0..order
}

// prefac[k] = 1/|sum_j M(j+1) e^{2 pi i j k / K}|^2, with the reference's
// guard for vanishing moduli (dftmod, src/CudaPMERecip.cu:2656-2678): the
// knot values M(1..order-1) sit at j = 0..order-2 (bsp_arr[1..] of
// calc_prefac, src/CudaPMERecip.cu:2724-2728, shifted by the knot offset).
static std::vector<double> euler_moduli(int order, int K) {
  std::vector<double> m;
  bspline_knots(order, m);
  std::vector<double> arr(std::max(K, order + 2), 0.0);
  for (int x = 1; x <= order - 1; x++) arr[x] = m[x];  // arr[1..order-1]
  std::vector<double> mod(K);
  for (int k = 0; k < K; k++) {
    double s1 = 0, s2 = 0;
    for (int j = 0; j < K - 1; j++) {
      double arg = 2.0 * kPi * (double)k * (double)j / (double)K;
      s1 += arr[j] * std::cos(arg);
      s2 += arr[j] * std::sin(arg);
    }
    mod[k] = s1 * s1 + s2 * s2;
  }
  for (int k = 0; k < K; k++)
    if (mod[k] < 1e-10) mod[k] = 0.5 * (mod[(k + K - 1) % K] + mod[(k + 1) % K]);
  for (int k = 0; k < K; k++) mod[k] = 1.0 / mod[k];
  return mod;
}

static void make_fft_dim(FftDim &d, int n, hipStream_t s) {
  d.n = n;
  d.nrad = 0;
  int m = n;
  const int pref[] = {4, 2, 3, 5, 7, 11, 13};
  while (m > 1) {
    bool found = false;
    for (int r : pref) {
      if (m % r == 0) {
        if (d.nrad >= 16) throw InvalidArgument("FFT size has too many factors");
        d.rad[d.nrad++] = r;
        m /= r;
        found = true;
        break;
      }
    }
    if (!found)
      throw InvalidArgument("FFT grid size " + std::to_string(n) +
                            " must factor into 2, 3, 5, 7, 11, 13");
  }
  std::vector<float2> tw(n);
  std::vector<double2> tw64(n);
  for (int e = 0; e < n; e++) {
    double a = -2.0 * kPi * (double)e / (double)n;
    tw64[e] = make_double2(std::cos(a), std::sin(a));
    tw[e] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  d.tw.upload(tw, s);
  d.tw64.upload(tw64, s);
}

void NbSystem::setup_pme() {
  if (order != 4) throw InvalidArgument("only PME spline order 4 is supported");
  for (int d = 0; d < 3; d++)
    if (K[d] < 4 || K[d] > 1024) throw InvalidArgument("FFT grid size out of range [4, 1024]");
  make_fft_dim(fx, K[0], 0);
  make_fft_dim(fy, K[1], 0);
  make_fft_dim(fz, K[2], 0);
  size_t ng = (size_t)K[0] * K[1] * K[2];
  size_t nc = (size_t)(K[0] / 2 + 1) * K[1] * K[2];
  if (precision == 1) {
    cgrid_d.reserve(nc);
    phi_d.reserve(ng);
  } else {
    cgrid.reserve(nc);
    phi.reserve(ng);
  }
  std::vector<double> pre;
  for (int d = 0; d < 3; d++) {
    auto m = euler_moduli(order, K[d]);
    pre.insert(pre.end(), m.begin(), m.end());
  }
  std::vector<float> pref(pre.begin(), pre.end());
  prefac.upload(pref, 0);
  prefac_d.upload(pre, 0);
}

void NbSystem::setup_columns() {
  // xy column edge for ~32-atom cells: (V*32/N)^(1/3) (the reference's
  // choice, src/CudaNeighborListSort.cu:899-903), narrowed so that the
  // expected column population stays <= kMaxColumn/2 for the LDS sort.
  double V = box[0] * box[1] * box[2];
  double rho = natom / V;
  double edge = std::cbrt(V * kCell / std::max(natom, 1));
  double cap_edge = std::sqrt((kMaxColumn / 2) / std::max(rho * box[2], 1e-30));
  edge = std::min(edge, cap_edge);
  ncx = std::max(1, (int)std::floor(box[0] / edge));
  ncy = std::max(1, (int)std::floor(box[1] / edge));
  ncol = ncx * ncy;
  ncell_cap = (natom + kCell - 1) / kCell + ncol + 1;
}

}  // namespace apo

// ---------------------------------------------------------------------------
// creation / calc_force
// ---------------------------------------------------------------------------
#include "apocharmm_amd.h"

namespace apo {

static void check_params(const apo_nb_params &p, int natom) {
  for (int d = 0; d < 3; d++)
    if (!(p.box[d] > 0)) throw InvalidArgument("box dimensions must be positive");
  if (!(p.ctofnb > 0) || !(p.cutnb >= p.ctofnb))
    throw InvalidArgument("need 0 < ctofnb <= cutnb");
  if (!(p.ctonnb >= 0)) throw InvalidArgument("ctonnb must be >= 0");
  if (!(p.kappa > 0)) throw InvalidArgument("kappa must be positive");
  // minimum image must be unique inside r_off (src/ForceManager.cu:443-447
  // applies the same rule to the cutoff)
  for (int d = 0; d < 3; d++)
    if (p.ctofnb >= 0.5 * p.box[d])
      throw InvalidArgument("ctofnb must be smaller than half the box");
  if (natom <= 0) throw InvalidArgument("natom must be positive");
  if (p.precision != 0 && p.precision != 1) throw InvalidArgument("unknown precision");
}

NbSystem *create_system(const apo_nb_params &p, const apo_nb_topology &t) {
  check_params(p, t.natom);
  if (t.ntypes <= 0 || !t.vdwtype || !t.vdwtab)
    throw InvalidArgument("topology needs vdwtype and vdwtab");
  auto *S = new NbSystem();
  try {
    S->natom = t.natom;
    for (int d = 0; d < 3; d++) S->box[d] = p.box[d];
    S->kappa = p.kappa;
    S->cutnb = p.cutnb;
    S->ctofnb = p.ctofnb;
    S->ctonnb = p.ctonnb;
    S->e14fac = p.e14fac;
    for (int d = 0; d < 3; d++) S->K[d] = p.fft_grid[d];
    S->order = p.pme_order ? p.pme_order : 4;
    S->precision = p.precision;
    APO_HIP(hipGetDevice(&S->device));
    hipDeviceProp_t prop;
    APO_HIP(hipGetDeviceProperties(&prop, S->device));
    S->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    for (hipStream_t &st : S->side) APO_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (hipEvent_t *e : {&S->ev_fork, &S->ev_sorted, &S->ev_pme, &S->ev_bond})
      APO_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));

    const int n = t.natom;
    for (int i = 0; i < n; i++)
      if (t.vdwtype[i] < 0 || t.vdwtype[i] >= t.ntypes)
        throw InvalidArgument("vdwtype out of range");
    S->ntypes = t.ntypes;
    S->vdwtype.upload(t.vdwtype, n, 0);
    int npair = t.ntypes * (t.ntypes + 1) / 2;
    std::vector<float2> tab(npair), tab14(npair);
    std::vector<double2> tabd(npair), tab14d(npair);
    const float *t14 = t.vdwtab14 ? t.vdwtab14 : t.vdwtab;
    for (int k = 0; k < npair; k++) {
      // stored as (6 c6, 12 c12) like setup_vdwparam
      // (src/CudaPMEDirectForce.cu:534-540)
      tab[k] = make_float2(6.0f * t.vdwtab[2 * k], 12.0f * t.vdwtab[2 * k + 1]);
      tab14[k] = make_float2(6.0f * t14[2 * k], 12.0f * t14[2 * k + 1]);
      tabd[k] = make_double2(6.0 * (double)t.vdwtab[2 * k], 12.0 * (double)t.vdwtab[2 * k + 1]);
      tab14d[k] = make_double2(6.0 * (double)t14[2 * k], 12.0 * (double)t14[2 * k + 1]);
    }
    if (t.ntypes <= 32) {
      std::vector<float2> sq((size_t)t.ntypes * t.ntypes);
      std::vector<double2> sqd((size_t)t.ntypes * t.ntypes);
      for (int a = 0; a < t.ntypes; a++)
        for (int b = 0; b < t.ntypes; b++) {
          int hi = a > b ? a : b, lo = a > b ? b : a;
          sq[(size_t)a * t.ntypes + b] = tab[hi * (hi + 1) / 2 + lo];
          sqd[(size_t)a * t.ntypes + b] = tabd[hi * (hi + 1) / 2 + lo];
        }
      S->vdwsq.upload(sq, 0);
      S->vdwsq_d.upload(sqd, 0);
    }
    // does the (c6, c12) table follow eps_ij = sqrt(eps_i eps_j), Rmin_ij =
    // Rmin_i/2 + Rmin_j/2 (c6 = 2 eps Rmin^6, c12 = eps Rmin^12)?  Per-type
    // values from the diagonal: Rmin_ii^6 = 2 c12/c6, eps_i = c6^2 / (4 c12).
    {
      const int nt = t.ntypes;
      std::vector<double> se(nt), rh(nt);
      bool ok = true;
      for (int a = 0; a < nt && ok; a++) {
        const double c6 = t.vdwtab[2 * (a * (a + 1) / 2 + a)];
        const double c12 = t.vdwtab[2 * (a * (a + 1) / 2 + a) + 1];
        if (c6 == 0.0 && c12 == 0.0) {
          se[a] = 0.0;
          rh[a] = 0.0;
        } else if (c6 > 0.0 && c12 > 0.0) {
          se[a] = std::sqrt(c6 * c6 / (4.0 * c12));
          rh[a] = 0.5 * std::pow(2.0 * c12 / c6, 1.0 / 6.0);
        } else {
          ok = false;
        }
      }
      for (int a = 0; a < nt && ok; a++)
        for (int b = 0; b <= a && ok; b++) {
          const double e = se[a] * se[b], R = rh[a] + rh[b], R6 = std::pow(R, 6);
          const double c6 = 2.0 * e * R6, c12 = e * R6 * R6;
          const double t6 = t.vdwtab[2 * (a * (a + 1) / 2 + b)],
                       t12 = t.vdwtab[2 * (a * (a + 1) / 2 + b) + 1];
          ok = std::fabs(c6 - t6) <= 1e-6 * std::fabs(t6) + 1e-30 &&
               std::fabs(c12 - t12) <= 1e-6 * std::fabs(t12) + 1e-30;
        }
      S->lj_combine = ok;
      std::vector<float2> lt(nt);
      for (int a = 0; a < nt; a++) lt[a] = make_float2((float)se[a], (float)rh[a]);
      S->ljtype.upload(lt, 0);
    }
    S->vdwtab.upload(tab, 0);
    S->vdwtab14.upload(tab14, 0);
    S->vdwtab_d.upload(tabd, 0);
    S->vdwtab14_d.upload(tab14d, 0);

    // exclusions from the bond connectivity
    for (int b = 0; b < 2 * t.nbond; b++)
      if (t.bonds[b] < 0 || t.bonds[b] >= n) throw InvalidArgument("bond index out of range");
    std::vector<int32_t> eptr, eidx, in14, ex14;
    build_exclusions(n, t.nbond, t.bonds, eptr, eidx, in14, ex14);
    S->h_bonds.assign(t.bonds, t.bonds + 2 * (size_t)t.nbond);
    S->excl_ptr.upload(eptr.data(), eptr.size(), 0);
    S->excl_idx.upload(eidx.data(), eidx.size(), 0);
    S->nin14 = (int)in14.size() / 2;
    S->nex14 = (int)ex14.size() / 2;
    S->in14.upload(reinterpret_cast<const int2 *>(in14.data()), S->nin14, 0);
    S->ex14.upload(reinterpret_cast<const int2 *>(ex14.data()), S->nex14, 0);

    // bonded lists
    auto chk = [&](const int32_t *idx, int cnt, int w, const char *what) {
      if (cnt < 0) throw InvalidArgument(std::string("negative count: ") + what);
      if (cnt > 0 && !idx) throw InvalidArgument(std::string("missing list: ") + what);
      for (int k = 0; k < cnt * w; k++)
        if (idx[k] < 0 || idx[k] >= n)
          throw InvalidArgument(std::string("atom index out of range in ") + what);
    };
    chk(t.bonds, t.nbond, 2, "bonds");
    chk(t.ureyb, t.nureyb, 2, "ureyb");
    chk(t.angles, t.nangle, 3, "angles");
    chk(t.dihe, t.ndihe, 4, "dihedrals");
    chk(t.imdihe, t.nimdihe, 4, "impropers");
    std::vector<BondT> hb(t.nbond), hu(t.nureyb);
    for (int k = 0; k < t.nbond; k++)
      hb[k] = {t.bonds[2 * k], t.bonds[2 * k + 1], t.bond_par[2 * k], t.bond_par[2 * k + 1]};
    for (int k = 0; k < t.nureyb; k++)
      hu[k] = {t.ureyb[2 * k], t.ureyb[2 * k + 1], t.ub_par[2 * k], t.ub_par[2 * k + 1]};
    std::vector<AngleT> ha(t.nangle);
    for (int k = 0; k < t.nangle; k++)
      ha[k] = {t.angles[3 * k], t.angles[3 * k + 1], t.angles[3 * k + 2], t.angle_par[2 * k],
               t.angle_par[2 * k + 1]};
    auto mkd = [](const int32_t *idx, const float *par, int cnt) {
      std::vector<DiheT> v(cnt);
      for (int k = 0; k < cnt; k++) {
        DiheT d{};
        d.i = idx[4 * k];
        d.j = idx[4 * k + 1];
        d.k = idx[4 * k + 2];
        d.l = idx[4 * k + 3];
        d.n = par[3 * k];
        d.kc = par[3 * k + 1];
        d.delta = par[3 * k + 2];
        d.sd = (float)std::sin((double)par[3 * k + 2]);
        d.cd = (float)std::cos((double)par[3 * k + 2]);
        v[k] = d;
      }
      return v;
    };
    S->nbond = t.nbond;
    S->nub = t.nureyb;
    S->nangle = t.nangle;
    S->ndihe = t.ndihe;
    S->nimdihe = t.nimdihe;
    S->bonds.upload(hb, 0);
    S->ureyb.upload(hu, 0);
    S->angles.upload(ha, 0);
    S->dihe.upload(mkd(t.dihe, t.dihe_par, t.ndihe), 0);
    S->imdihe.upload(mkd(t.imdihe, t.imdihe_par, t.nimdihe), 0);

    S->make_pair_consts();
    if (t.ntypes <= kMaxTypesPqa) {
      // (P, Q, A12) of the inner and outer switch regions per type pair,
      // formed in FP64 from the (6 c6, 12 c12) table
      const int nt = t.ntypes;
      const double r3 = 1.0 / (S->ctofnb * S->ctofnb * S->ctofnb), r6 = r3 * r3;
      const double k6 = S->pc_d.k6, k12 = S->pc_d.k12;
      std::vector<float4> pqa(2 * (size_t)nt * nt);
      for (int a = 0; a < nt; a++)
        for (int b = 0; b < nt; b++) {
          const int hi = a > b ? a : b, lo = a > b ? b : a;
          const double c6 = tabd[hi * (hi + 1) / 2 + lo].x, c12 = tabd[hi * (hi + 1) / 2 + lo].y;
          const size_t k = 2 * ((size_t)a * nt + b);
          pqa[k] = make_float4((float)c6, 0.f, (float)c12, 0.f);
          pqa[k + 1] = make_float4((float)(k6 * c6 + k12 * c12 * r6), (float)(k6 * c6 * r3),
                                   (float)(k12 * c12), 0.f);
        }
      S->vdwpqa.upload(pqa, 0);
    }
    S->setup_columns();
    S->setup_pme();

    // per-step direct buffers
    S->ssort = S->ncell_cap * kCell;
    S->posq.reserve(S->ssort);
    S->type_s.reserve(S->ssort);
    S->fsort.reserve(3 * (size_t)S->ssort);
    S->loc2glo.reserve(S->ssort);
    S->glo2loc.reserve(n);
    S->xref.reserve(n);
    launch_fill_i32(S->glo2loc.get(), n, 0);
    S->cell_bb.reserve(2 * (size_t)S->ncell_cap);
    S->meta.reserve(32);
    launch_fill_i32(S->meta.get(), 32, 0);
    APO_HIP(hipHostMalloc(reinterpret_cast<void **>(&S->h_meta), 16 * sizeof(int)));
    for (int k = 0; k < 16; k++) S->h_meta[k] = 0;
    S->col_count.reserve(S->ncol + 1);
    S->col_start.reserve(S->ncol + 1);
    S->col_ncell.reserve(S->ncol + 1);
    S->col_cellstart.reserve(S->ncol + 1);
    S->atom_col.reserve(n);
    S->atom_slot.reserve(n);
    S->colkeys.reserve(n);
    S->zkey.reserve(n);
    for (auto *b : {&S->icell_cnt, &S->icell_start}) b->reserve(S->ncell_cap + 1);
    for (auto *b : {&S->grp_cnt, &S->grp_start, &S->grp_nitem, &S->grp_itemstart})
      b->reserve((size_t)S->ncell_cap * kGroupsPerCell + 1);

    // partials layout (fixed block counts -> deterministic reductions)
    S->pair_blocks = kPairBlocksPerCu * S->ncu;
    PartLayout &L = S->part;
    int off = 0;
    L.pair_off = off; L.pair_n = S->pair_blocks; off += L.pair_n * kPairPart;
    L.p14_off = off; L.p14_n = std::max(1, (S->nin14 + S->nex14 + 255) / 256); off += L.p14_n * kP14Part;
    int nbt = S->nbond + S->nub + S->nangle + S->ndihe + S->nimdihe;
    L.bond_off = off; L.bond_n = std::max(1, (nbt + 255) / 256); off += L.bond_n * kBondPart;
    // z-pass blocks are K1 * ceil(H / W) with W <= 16 set by LDS size; size
    // for the worst case W = 1 (unused slots are zero-filled)
    S->zconv_blocks = S->K[1] * (S->K[0] / 2 + 1);
    L.recip_off = off; L.recip_n = S->zconv_blocks; off += L.recip_n * kRecipPart;
    L.self_off = off; L.self_n = std::max(1, (n + 255) / 256); off += L.self_n;
    L.total = off;
    S->parts.reserve(off);
    APO_HIP(hipDeviceSynchronize());
  } catch (...) {
    delete S;
    throw;
  }
  return S;
}

void set_box(NbSystem &S, const double *b) {
  for (int d = 0; d < 3; d++)
    if (!(b[d] > 0) || S.ctofnb >= 0.5 * b[d])
      throw InvalidArgument("box too small for ctofnb or not positive");
  for (int d = 0; d < 3; d++) S.box[d] = b[d];
  S.make_pair_consts();
  S.list_valid = false;
  S.generation++;
}

void calc_force(NbSystem &S, const float *xyzq, int flags, int64_t *force, int stride,
                double *energy, double *virial, hipStream_t s) {
  if (!xyzq || !force) throw InvalidArgument("xyzq and force must be device pointers");
  int min_stride = ((S.natom + 31) * 8 + 255) / 256 * 256 / 8;
  if (stride < S.natom) throw InvalidArgument("stride smaller than natom");
  (void)min_stride;
  bool E = flags & APO_CALC_ENERGY, V = flags & APO_CALC_VIRIAL;
  if (E && !energy) throw InvalidArgument("APO_CALC_ENERGY needs an energy buffer");
  if (V && !virial) throw InvalidArgument("APO_CALC_VIRIAL needs a virial buffer");
  int terms = flags & APO_TERM_ALL;
  if (!terms) terms = APO_TERM_ALL;
  const float4 *x4 = reinterpret_cast<const float4 *>(xyzq);
  hipEvent_t t_all = nullptr, t0 = nullptr, t1 = nullptr;
  S.tic(5, s, t_all);
  const bool dir = terms & APO_TERM_DIRECT, rec = terms & APO_TERM_RECIP;
  // the coordinate sort zeroes `force` (unless the call accumulates); every
  // term adds into it after the sort
  const bool acc = flags & APO_ACCUMULATE;
  if (!(dir || rec) && !acc) launch_fill_i64(force, 3 * (size_t)stride, s);
  const bool bnd = terms & (APO_TERM_PAIRS14 | APO_TERM_BONDED);
  // Three branches that meet before the sorted pair gradient is folded into
  // `force`:  bonded + 1-4 (side[1], int64 atomics into force) | pair kernel
  // (caller's stream, atomics into the sorted buffer) | PME spread, FFT,
  // gather (side[0], atomics into force).
  const bool fork = !S.single_stream;
  hipStream_t sb = fork ? S.side[1] : s, sp = fork ? S.side[0] : s;
  bool rebuilt = false;
  if (dir || rec) {
    // both the pair kernel and PME work on the cell-sorted coordinates
    if ((flags & APO_RESET_LIST) || !S.list_valid) {
      rebuilt = true;
      S.tic(4, s, t0);
      nblist_build(S, x4, s, false);
      S.toc(4, s, t0);
    }
    S.tic(7, s, t0);
    nblist_sort_coords(S, x4, s, acc ? nullptr : force, acc ? 0 : 3 * (size_t)stride);
    S.toc(7, s, t0);
  }
  // Both side branches fork right after the sort.  The capture order of the
  // three branches is a launch setting (pair_first): a graph keeps the first
  // captured child of the sort on the caller's queue and dispatches it first.
  if (fork && rec) {
    APO_HIP(hipEventRecord(S.ev_sorted, s));
    APO_HIP(hipStreamWaitEvent(sp, S.ev_sorted, 0));
  }
  if (fork && bnd) {
    APO_HIP(hipEventRecord(S.ev_fork, s));
    APO_HIP(hipStreamWaitEvent(sb, S.ev_fork, 0));
  }
  auto pair_branch = [&] {
    // a pruned list is refreshed after every build and when asked: by the
    // pair kernel itself in mixed-precision steps without energies (it tests
    // every listed pair anyway), else by the separate prune kernel first
    const bool pruned = flags & kCalcUsePruned;
    const bool refresh = pruned && (rebuilt || (flags & kCalcPrune));
    const bool fused = refresh && S.precision == 0 && !E && !V;
    if (refresh && !fused) nblist_prune(S, s);
    S.tic(0, s, t0);
    direct_launch(S, E, V, s, pruned, fused);
    S.toc(0, s, t0);
  };
  if (dir && S.pair_first) pair_branch();
  if (rec) {
    recip_launch(S, x4, force, stride, E, V, sp);
    if (fork) APO_HIP(hipEventRecord(S.ev_pme, sp));
  }
  // bonded after PME (measured: launched first, its blocks delay the sort
  // and the PME chain)
  if (bnd) {
    S.tic(6, sb, t1);
    if (terms & APO_TERM_PAIRS14) pairs14_launch(S, x4, force, stride, E, V, sb);
    if (terms & APO_TERM_BONDED) bonded_launch(S, x4, force, stride, E, V, sb);
    S.toc(6, sb, t1);
    if (fork) APO_HIP(hipEventRecord(S.ev_bond, sb));
  }
  if (dir && !S.pair_first) pair_branch();
  if (fork && rec) APO_HIP(hipStreamWaitEvent(s, S.ev_pme, 0));
  if (fork && bnd) APO_HIP(hipStreamWaitEvent(s, S.ev_bond, 0));
  if ((dir || rec) && !(flags & kCalcNoFinalize)) {
    S.tic(7, s, t0);
    direct_finalize(S, force, stride, s);
    S.toc(7, s, t0);
  }
  if ((terms & APO_TERM_SELF) && E) self_launch(S, x4, s);
  if (E || V) reduce_energy_launch(S, terms, E, V, energy, virial, s, acc);
  S.toc(5, s, t_all);
}

}  // namespace apo

namespace apo {

hipEvent_t NbSystem::prof_event() {
  if (ev_used == ev_pool.size()) {
    hipEvent_t e;
    APO_HIP(hipEventCreate(&e));
    ev_pool.push_back(e);
  }
  return ev_pool[ev_used++];
}

void NbSystem::tic(int, hipStream_t s, hipEvent_t &a) {
  if (!prof_on) return;
  a = prof_event();
  APO_HIP(hipEventRecord(a, s));
  prof_open++;
}

// Events are recycled only when no timer is open: an enclosing tic's event
// must not be re-recorded before its toc.
void NbSystem::toc(int slot, hipStream_t s, hipEvent_t a) {
  if (!prof_on) return;
  hipEvent_t b = prof_event();
  APO_HIP(hipEventRecord(b, s));
  prof_recs.push_back({slot, a, b});
  if (prof_open > 0) prof_open--;
  if (prof_open == 0 && ev_used > 4096) profile_flush();
}

void NbSystem::profile_flush() {
  for (auto &r : prof_recs) {
    APO_HIP(hipEventSynchronize(r.b));
    float ms = 0.f;
    APO_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    prof_ms[r.slot] += ms;
    prof_cnt[r.slot] += 1;
  }
  prof_recs.clear();
  ev_used = 0;
  prof_open = 0;
}

NbSystem::~NbSystem() {
  // list builds write the pinned h_meta from whatever stream the caller gave
  // them, and PME / bonded work runs on the side streams: drain the device
  // before any of them is released
  (void)hipDeviceSynchronize();
  if (h_meta) (void)hipHostFree(h_meta);
  for (auto e : ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : {ev_fork, ev_sorted, ev_pme, ev_bond})
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t st : side)
    if (st) (void)hipStreamDestroy(st);
}

}  // namespace apo
