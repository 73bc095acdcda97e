// A C++ caller of the force path through the ForceManager mirror
// (include/apocharmm_amd/ForceManager.hpp) with a plugin force subscribed,
// as a reference caller attaches HarmonicRestraintForce
// (test/unittests/unittest-restraintForce.cpp:70-71).  Checks, bit for bit:
//   total gradient with the plugin = core gradient + plugin gradient,
//   every core energy term unchanged and epot = core epot + plugin energy,
//   the plugin's gradient = 2 k (x - x0) rounded to 2^-40,
//   APO_ACCUMULATE on a buffer holding the plugin's gradient = the total.
// Exit status 0 when every check holds.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <vector>

#include "apocharmm_amd/ForceManager.hpp"

namespace {

#define HIPCHECK(x)                                                                 \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
      std::exit(2);                                                                 \
    }                                                                               \
  } while (0)

constexpr double kScale = 1099511627776.0;  // 2^40

// Harmonic positional restraint: E = sum_i k_i |x_i - x0_i|^2 (the energy form
// of HarmonicRestraintForce, include/HarmonicRestraintForce.h).
__global__ void k_restraint(const float4 *xyzq, const double4 *ref, int n, int stride,
                            long long *f, double *energy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < n) {
    const double4 r = ref[i];
    const double dx = (double)xyzq[i].x - r.x, dy = (double)xyzq[i].y - r.y,
                 dz = (double)xyzq[i].z - r.z;
    f[i] = llrint(2.0 * r.w * dx * kScale);
    f[stride + i] = llrint(2.0 * r.w * dy * kScale);
    f[2 * stride + i] = llrint(2.0 * r.w * dz * kScale);
    e = r.w * (dx * dx + dy * dy + dz * dz);
  }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
  if ((threadIdx.x & 63) == 0 && e != 0.0) atomicAdd(energy, e);
}
__global__ void k_zero(long long *f, size_t n, double *e) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    f[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) e[0] = 0.0;
}

class HarmonicRestraint {
 public:
  HarmonicRestraint(hipStream_t s, std::vector<double4> ref) : s_(s), ref_h_(std::move(ref)) {}
  ~HarmonicRestraint() {
    for (void *p : {(void *)f_, (void *)e_, (void *)ref_}) (void)hipFree(p);
  }
  void initialize(int n, const std::vector<double> &) {
    n_ = n;
    stride_ = ((n + 31) * 8 + 255) / 256 * 256 / 8;  // Force<long long> stride
    HIPCHECK(hipMalloc(reinterpret_cast<void **>(&f_), 3 * (size_t)stride_ * sizeof(long long)));
    HIPCHECK(hipMalloc(reinterpret_cast<void **>(&e_), sizeof(double)));
    HIPCHECK(hipMalloc(reinterpret_cast<void **>(&ref_), n * sizeof(double4)));
    HIPCHECK(hipMemcpy(ref_, ref_h_.data(), n * sizeof(double4), hipMemcpyHostToDevice));
  }
  void clear() { k_zero<<<64, 256, 0, s_>>>(f_, 3 * (size_t)stride_, e_); }
  void calcForce(const float4 *xyzq, bool, bool) {
    k_restraint<<<(n_ + 255) / 256, 256, 0, s_>>>(xyzq, ref_, n_, stride_, f_, e_);
  }
  void setBoxDimensions(const std::vector<double> &) {}
  int64_t *getForce() { return reinterpret_cast<int64_t *>(f_); }
  double *getEnergy() { return e_; }

 private:
  hipStream_t s_;
  std::vector<double4> ref_h_;
  int n_ = 0, stride_ = 0;
  long long *f_ = nullptr;
  double *e_ = nullptr;
  double4 *ref_ = nullptr;
};

// 216 rigid TIP3P waters on a 6^3 lattice (3.1 A) with a deterministic jitter.
struct Water {
  apocharmm_amd::Topology topo;
  std::vector<float4> xyzq;
  double L = 18.6;
};
Water make_water() {
  Water w;
  const int m = 6, nw = m * m * m, n = 3 * nw;
  auto &t = w.topo;
  t.natom = n;
  t.ntypes = 2;  // 0 = HT, 1 = OT
  const double eps[2] = {-0.046, -0.1521}, rh[2] = {0.2245, 1.7682};
  for (int a = 0; a < 2; a++)
    for (int b = 0; b <= a; b++) {
      const double e = std::sqrt(eps[a] * eps[b]), r = rh[a] + rh[b];
      t.vdwtab.push_back((float)(2.0 * e * std::pow(r, 6)));
      t.vdwtab.push_back((float)(e * std::pow(r, 12)));
    }
  const double th = 104.52 * M_PI / 180.0, b = 0.9572;
  uint32_t rng = 12345u;
  auto jit = [&]() {
    rng = rng * 1664525u + 1013904223u;
    return ((rng >> 8) / 16777216.0 - 0.5) * 0.3;
  };
  for (int k = 0; k < nw; k++) {
    const double ox = (k % m) * 3.1 + jit(), oy = ((k / m) % m) * 3.1 + jit(),
                 oz = (k / (m * m)) * 3.1 + jit();
    const double c = std::cos(0.7 * k), s = std::sin(0.7 * k);  // in-plane turn
    const double h1[3] = {b * c, b * s, 0.0};
    const double h2[3] = {b * (c * std::cos(th) - s * std::sin(th)),
                          b * (s * std::cos(th) + c * std::sin(th)), 0.0};
    w.xyzq.push_back(make_float4((float)ox, (float)oy, (float)oz, -0.834f));
    w.xyzq.push_back(make_float4((float)(ox + h1[0]), (float)(oy + h1[1]), (float)oz, 0.417f));
    w.xyzq.push_back(make_float4((float)(ox + h2[0]), (float)(oy + h2[1]), (float)oz, 0.417f));
    t.vdwtype.insert(t.vdwtype.end(), {1, 0, 0});
    const int o = 3 * k;
    t.bonds.insert(t.bonds.end(), {o, o + 1, o, o + 2});
    t.bond_par.insert(t.bond_par.end(), {0.9572f, 450.0f, 0.9572f, 450.0f});
    t.angles.insert(t.angles.end(), {o + 1, o, o + 2});
    t.angle_par.insert(t.angle_par.end(), {(float)th, 55.0f});
  }
  return w;
}

void setup(apocharmm_amd::ForceManager &fm, double L) {
  fm.setBoxDimensions({L, L, L});
  fm.setFFTGrid(20, 20, 20);
  fm.setKappa(0.34);
  fm.setCutoff(9.0);
  fm.setCtofnb(8.0);
  fm.setCtonnb(7.0);
}

template <typename T> std::vector<T> fetch(const T *d, size_t n) {
  std::vector<T> h(n);
  HIPCHECK(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
  return h;
}

}  // namespace

int main() {
  Water w = make_water();
  const int n = w.topo.natom;
  hipStream_t s;
  HIPCHECK(hipStreamCreate(&s));
  float4 *xyzq = nullptr;
  HIPCHECK(hipMalloc(&xyzq, n * sizeof(float4)));
  HIPCHECK(hipMemcpy(xyzq, w.xyzq.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  // restraint on the oxygens, to points 0.2-0.4 A away
  std::vector<double4> ref(n);
  for (int i = 0; i < n; i++) {
    const double k = i % 3 == 0 ? 5.0 + (i % 7) : 0.0;
    ref[i] = make_double4(w.xyzq[i].x + 0.2, w.xyzq[i].y - 0.3, w.xyzq[i].z + 0.4, k);
  }

  int bad = 0;
  auto expect = [&](bool ok, const char *what) {
    if (!ok) {
      std::printf("FAIL: %s\n", what);
      bad++;
    }
  };

  apocharmm_amd::ForceManager core(w.topo);
  setup(core, w.L);
  core.initialize(s);
  core.calcForce(xyzq, true, true, true);
  const int stride = core.getForceStride();
  HIPCHECK(hipStreamSynchronize(s));
  const auto g0 = fetch(core.getForcesFixed(), 3 * (size_t)stride);
  const auto e0 = core.getEnergyComponents();
  const auto v0 = core.getVirial();

  apocharmm_amd::ForceManager fm(w.topo);
  setup(fm, w.L);
  auto restraint = std::make_shared<HarmonicRestraint>(s, ref);
  fm.subscribe(restraint, "harm");
  fm.initialize(s);
  fm.calcForce(xyzq, true, true, true);
  HIPCHECK(hipStreamSynchronize(s));
  const auto g1 = fetch(fm.getForcesFixed(), 3 * (size_t)stride);
  const auto e1 = fm.getEnergyComponents();
  const auto v1 = fm.getVirial();
  const auto gp = fetch(restraint->getForce(), 3 * (size_t)stride);
  const double ep = fetch(restraint->getEnergy(), 1)[0];

  // the plugin's own output against its closed form
  double ep_host = 0.0;
  bool gp_ok = true;
  for (int i = 0; i < n; i++) {
    const double d[3] = {(double)w.xyzq[i].x - ref[i].x, (double)w.xyzq[i].y - ref[i].y,
                         (double)w.xyzq[i].z - ref[i].z};
    for (int c = 0; c < 3; c++)
      gp_ok = gp_ok && gp[(size_t)c * stride + i] == llrint(2.0 * ref[i].w * d[c] * kScale);
    ep_host += ref[i].w * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  }
  expect(gp_ok, "plugin gradient = 2 k (x - x0) at 2^-40");
  expect(std::fabs(ep - ep_host) <= 1e-12 * ep_host && ep > 0, "plugin energy closed form");

  bool sum_ok = true;
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < n; i++) {
      const size_t k = (size_t)c * stride + i;
      sum_ok = sum_ok && g1[k] == g0[k] + gp[k];
    }
  expect(sum_ok, "total gradient = core + plugin (int64, every atom)");
  for (const auto &kv : e0)
    if (kv.first != "epot") expect(e1.at(kv.first) == kv.second, "core energy term unchanged");
  expect(e1.at("epot") == e0.at("epot") + ep, "epot = core epot + plugin energy");
  for (int k = 0; k < 9; k++) expect(v1[k] == v0[k], "virial without the plugin's term");

  // APO_ACCUMULATE: the core path added onto buffers that already hold the
  // plugin's gradient (and energy in the epot slot)
  int64_t *acc = nullptr;
  double *eacc = nullptr, *vacc = nullptr;
  HIPCHECK(hipMalloc(&acc, 3 * (size_t)stride * sizeof(int64_t)));
  HIPCHECK(hipMalloc(&eacc, APO_NUM_ENERGY * sizeof(double)));
  HIPCHECK(hipMalloc(&vacc, 9 * sizeof(double)));
  HIPCHECK(hipMemcpy(acc, gp.data(), gp.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  std::vector<double> eh(APO_NUM_ENERGY, 0.0), vh(9, 0.0);
  eh[APO_E_EPOT] = ep;
  HIPCHECK(hipMemcpy(eacc, eh.data(), eh.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(vacc, vh.data(), vh.size() * sizeof(double), hipMemcpyHostToDevice));
  apocharmm_amd::check(apo_nb_calc_force(core.handle(), reinterpret_cast<const float *>(xyzq),
                                         APO_CALC_ENERGY | APO_CALC_VIRIAL | APO_ACCUMULATE,
                                         acc, stride, eacc, vacc, s));
  HIPCHECK(hipStreamSynchronize(s));
  const auto ga = fetch(acc, 3 * (size_t)stride);
  const auto ea = fetch(eacc, APO_NUM_ENERGY);
  const auto va = fetch(vacc, 9);
  bool acc_ok = true;
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < n; i++) acc_ok = acc_ok && ga[(size_t)c * stride + i] == g1[(size_t)c * stride + i];
  expect(acc_ok, "APO_ACCUMULATE gradient = total with the plugin");
  expect(ea[APO_E_EPOT] == e1.at("epot"), "APO_ACCUMULATE epot");
  expect(ea[APO_E_ELEC] == e0.at("elec") && ea[APO_E_VDW] == e0.at("vdw"), "APO_ACCUMULATE terms");
  for (int k = 0; k < 9; k++) expect(va[k] == v0[k], "APO_ACCUMULATE virial");

  // a second evaluation with the subscribed plugin gives the same bits
  fm.calcForce(xyzq, false, true, false);
  HIPCHECK(hipStreamSynchronize(s));
  expect(fetch(fm.getForcesFixed(), 3 * (size_t)stride) == g1, "repeat evaluation bitwise");

  std::printf("fm_plugin: natom=%d plugins=%zu epot core %.10f + plugin %.10f = %.10f; %s\n", n,
              fm.getNumSubscribedForces(), e0.at("epot"), ep, e1.at("epot"),
              bad ? "FAILED" : "all checks passed");
  (void)hipFree(acc);
  (void)hipFree(eacc);
  (void)hipFree(vacc);
  (void)hipFree(xyzq);
  return bad ? 1 : 0;
}
