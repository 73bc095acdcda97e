// Replays the host-side call sequences of the serving loops against the HIP
// stand-in (mockhip.cpp), under AddressSanitizer: many water-box jobs
// created, advanced in chunks on their own streams, read back and destroyed
// (the stream server), then jobs one after another (the plain run), then
// jobs in fused batches re-formed at every membership change (the batch
// server), then systems in the many-system benchmark's order.  With APO_STANDIN_GROW=1 every list build reports more entries
// than the last, so capacities grow and captured graphs are recaptured.
// Exit status 0 = every call succeeded and nothing leaked; ASAN aborts on a
// memory error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "apocharmm_amd.h"

extern "C" size_t apo_standin_live_device_allocs(void);
extern "C" size_t apo_standin_live_handles(void);
extern "C" void apo_standin_counts(size_t *out4);

namespace {

#define CHECK(call)                                                                    \
  do {                                                                                 \
    apo_status st_ = (call);                                                           \
    if (st_ != APO_OK) {                                                               \
      std::fprintf(stderr, "%s:%d: %s -> %d: %s\n", __FILE__, __LINE__, #call, (int)st_, \
                   apo_last_error());                                                  \
      std::exit(2);                                                                    \
    }                                                                                  \
  } while (0)

// A TIP3P box of nw waters on a cubic lattice (geometry and parameters of
// the library's synthetic workloads), with its own PME grid.
struct Box {
  int nw = 0, n = 0;
  double L = 0;
  int K = 0;
  std::vector<double> xyz, mass, vel;
  std::vector<float> q, vdwtab, bond_par, angle_par;
  std::vector<int32_t> type, bonds, angles, waters;
};

int smooth(double L) {
  int k = std::max(4, (int)std::ceil(L - 1e-9));
  for (;; k++) {
    int m = k;
    for (int p : {2, 3, 5, 7})
      while (m % p == 0) m /= p;
    if (m == 1) return k;
  }
}

Box water_box(int nw, unsigned seed) {
  Box b;
  b.nw = nw;
  b.n = 3 * nw;
  b.L = std::cbrt(nw / 0.0334);
  b.K = smooth(b.L);
  const int m = (int)std::ceil(std::cbrt((double)nw));
  const double a = b.L / m, th = 104.52 * M_PI / 180.0, r = 0.9572;
  std::mt19937 rng(seed);
  std::normal_distribution<double> nrm(0.0, 1.0);
  for (int w = 0; w < nw; w++) {
    const double o[3] = {(w % m + 0.5) * a, (w / m % m + 0.5) * a, (w / (m * m) + 0.5) * a};
    const double h1[3] = {o[0] + r, o[1], o[2]};
    const double h2[3] = {o[0] + r * std::cos(th), o[1] + r * std::sin(th), o[2]};
    for (const double *p : {o, h1, h2})
      for (int d = 0; d < 3; d++) b.xyz.push_back(p[d]);
    const int i = 3 * w;
    b.q.insert(b.q.end(), {-0.834f, 0.417f, 0.417f});
    b.mass.insert(b.mass.end(), {15.9994, 1.008, 1.008});
    b.type.insert(b.type.end(), {1, 0, 0});
    b.bonds.insert(b.bonds.end(), {i, i + 1, i, i + 2, i + 1, i + 2});
    b.bond_par.insert(b.bond_par.end(), {0.9572f, 450.f, 0.9572f, 450.f, 1.5139f, 0.f});
    b.angles.insert(b.angles.end(), {i + 1, i, i + 2});
    b.angle_par.insert(b.angle_par.end(), {(float)th, 55.f});
    b.waters.insert(b.waters.end(), {i, i + 1, i + 2});
  }
  for (int k = 0; k < 3 * b.n; k++) b.vel.push_back(0.01 * nrm(rng));
  // (c6, c12) for the type pairs HH, OH, OO from eps and Rmin/2
  const double eps[2] = {0.046, 0.1521}, rh[2] = {0.2245, 1.7682};
  for (int x = 0; x < 2; x++)
    for (int y = 0; y <= x; y++) {
      const double e = std::sqrt(eps[x] * eps[y]), R = rh[x] + rh[y], R6 = std::pow(R, 6);
      b.vdwtab.push_back((float)(2.0 * e * R6));
      b.vdwtab.push_back((float)(e * R6 * R6));
    }
  return b;
}

struct Job {
  Box box;
  apo_nb_system *nb = nullptr;
  apo_lt_integrator *lt = nullptr;
  hipStream_t stream = nullptr;
  int nsteps = 0, done = 0;
};

void start(Job &j, bool own_stream) {
  const Box &b = j.box;
  apo_nb_params p{};
  p.box[0] = p.box[1] = p.box[2] = b.L;
  p.kappa = 0.34;
  p.cutnb = 14.0;
  p.ctofnb = 12.0;
  p.ctonnb = 10.0;
  p.e14fac = 1.0;
  p.fft_grid[0] = p.fft_grid[1] = p.fft_grid[2] = b.K;
  p.pme_order = 4;
  apo_nb_topology t{};
  t.natom = b.n;
  t.ntypes = 2;
  t.vdwtype = b.type.data();
  t.vdwtab = b.vdwtab.data();
  t.vdwtab14 = b.vdwtab.data();
  t.nbond = (int)b.bonds.size() / 2;
  t.bonds = b.bonds.data();
  t.bond_par = b.bond_par.data();
  t.nangle = (int)b.angles.size() / 3;
  t.angles = b.angles.data();
  t.angle_par = b.angle_par.data();
  CHECK(apo_nb_create(&p, &t, &j.nb));
  apo_lt_params lp{};
  lp.timestep = 0.002;
  lp.temperature = 300.0;
  lp.friction = 1.0;
  lp.seed = 5100 + b.nw;
  lp.nbfreq = 20;
  lp.nwater = b.nw;
  lp.use_graph = 1;
  lp.waters = b.waters.data();
  lp.mass = b.mass.data();
  lp.charge = b.q.data();
  lp.launch_setting = 0;  // the launch trial runs (timed cycles)
  CHECK(apo_lt_create(j.nb, &lp, b.xyz.data(), b.vel.data(), &j.lt));
  if (own_stream && hipStreamCreateWithFlags(&j.stream, hipStreamNonBlocking) != hipSuccess)
    std::exit(3);
  CHECK(apo_lt_initialize(j.lt, j.stream));
}

void finish(Job &j) {
  std::vector<double> x(3 * (size_t)j.box.n), v(x.size()), g(x.size());
  double thermo[6];
  int64_t step = 0;
  CHECK(apo_lt_get_state(j.lt, x.data(), v.data(), nullptr, g.data(), thermo, &step));
  if (step != j.done) {
    std::fprintf(stderr, "step count %lld != %d\n", (long long)step, j.done);
    std::exit(4);
  }
  CHECK(apo_lt_destroy(j.lt));
  CHECK(apo_nb_destroy(j.nb));
  if (j.stream) (void)hipStreamDestroy(j.stream);
  j = Job{};
}

std::vector<Job> make_jobs(int count, unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_int_distribution<int> waters(1000, 3000), chunks(2, 8);
  std::vector<Job> out(count);
  for (int k = 0; k < count; k++) {
    out[k].box = water_box(waters(rng), seed + k);
    out[k].nsteps = 100 * chunks(rng);
  }
  return out;
}

}  // namespace

int main(int argc, char **argv) {
  const int njobs = argc > 1 ? std::atoi(argv[1]) : 40;
  const int slots = 8, chunk = 100;
  size_t base_allocs = apo_standin_live_device_allocs();
  size_t base_handles = apo_standin_live_handles();

  // 1. stream server: up to `slots` jobs at once, each on its own stream
  {
    std::vector<Job> pending = make_jobs(njobs, 11);
    std::vector<Job> active;
    size_t next = 0;
    int served = 0;
    while (next < pending.size() || !active.empty()) {
      while (next < pending.size() && (int)active.size() < slots) {
        active.push_back(std::move(pending[next++]));
        start(active.back(), true);
      }
      for (Job &j : active) CHECK(apo_lt_propagate(j.lt, chunk, j.stream));
      for (Job &j : active) j.done += chunk;
      for (size_t k = 0; k < active.size();) {
        if (active[k].done >= active[k].nsteps) {
          finish(active[k]);
          active.erase(active.begin() + (long)k);
          served++;
        } else {
          k++;
        }
      }
    }
    std::printf("[standin] stream server: %d jobs\n", served);
  }
  // 2. plain run: jobs one after another on the null stream
  {
    std::vector<Job> jobs = make_jobs(6, 23);
    for (Job &j : jobs) {
      start(j, false);
      while (j.done < j.nsteps) {
        CHECK(apo_lt_propagate(j.lt, chunk, nullptr));
        j.done += chunk;
      }
      finish(j);
    }
    std::printf("[standin] plain run: %zu jobs\n", jobs.size());
  }
  // 3. batch server: the fused batch re-formed whenever membership changes
  {
    std::vector<Job> pending = make_jobs(njobs / 2, 37);
    std::vector<Job> active;
    size_t next = 0;
    apo_lt_batch *batch = nullptr;
    int served = 0, reforms = 0;
    while (next < pending.size() || !active.empty()) {
      bool changed = false;
      while (next < pending.size() && (int)active.size() < slots) {
        active.push_back(std::move(pending[next++]));
        start(active.back(), false);
        changed = true;
      }
      if (changed || !batch) {
        if (batch) CHECK(apo_lt_batch_destroy(batch));
        std::vector<apo_lt_integrator *> lts;
        for (Job &j : active) lts.push_back(j.lt);
        CHECK(apo_lt_batch_create(lts.data(), (int)lts.size(), 4, &batch));
        reforms++;
      }
      CHECK(apo_lt_batch_propagate(batch, chunk, nullptr));
      bool any = false;
      for (Job &j : active) {
        j.done += chunk;
        any = any || j.done >= j.nsteps;
      }
      if (any) {
        CHECK(apo_lt_batch_destroy(batch));
        batch = nullptr;
        for (size_t k = 0; k < active.size();) {
          if (active[k].done >= active[k].nsteps) {
            finish(active[k]);
            active.erase(active.begin() + (long)k);
            served++;
          } else {
            k++;
          }
        }
      }
    }
    if (batch) CHECK(apo_lt_batch_destroy(batch));
    std::printf("[standin] batch server: %d jobs, %d batches formed\n", served, reforms);
  }
  // 4. the many-system benchmark's order: every system created first, each
  // warmed and then advanced alone (launch trial included), then all of them
  // in one fused batch with its launch-shape autotuning, then destroyed
  {
    std::vector<Job> jobs = make_jobs(16, 53);
    for (Job &j : jobs) start(j, false);
    for (Job &j : jobs) CHECK(apo_lt_propagate(j.lt, 30, nullptr));
    for (Job &j : jobs) CHECK(apo_lt_propagate(j.lt, 400, nullptr));
    std::vector<apo_lt_integrator *> lts;
    for (Job &j : jobs) {
      j.done = 430;
      lts.push_back(j.lt);
    }
    apo_lt_batch *batch = nullptr;
    CHECK(apo_lt_batch_create(lts.data(), (int)lts.size(), 4, &batch));
    double report[16] = {0};
    CHECK(apo_lt_batch_autotune(batch, nullptr, report));
    CHECK(apo_lt_batch_propagate(batch, 100, nullptr));
    CHECK(apo_lt_batch_destroy(batch));
    for (Job &j : jobs) {
      j.done += 100;
      finish(j);
    }
    std::printf("[standin] many-system bench order: %zu systems\n", jobs.size());
  }
  const size_t allocs = apo_standin_live_device_allocs(), handles = apo_standin_live_handles();
  size_t c[4];
  apo_standin_counts(c);
  std::printf("[standin] live device allocations %zu (start %zu), live handles %zu (start %zu)\n",
              allocs, base_allocs, handles, base_handles);
  std::printf("[standin] device allocations %zu, graphs instantiated %zu, graph launches %zu, "
              "list builds exported %zu\n", c[0], c[1], c[2], c[3]);
  if (allocs != base_allocs || handles != base_handles) return 5;
  return 0;
}
