// The reference's existing C ABI (include/apocharmm_amd_capi.h) over this
// library's force path (NbSystem) and integrator (LtSystem): PSF / CRD /
// parameter handles, ForceManager, CharmmContext, the integrator with its
// subscribers and restart files.  Behaviour follows the reference classes
// behind capi/src/apocharmm_c/*.cpp: CharmmPSF / CharmmCrd /
// CharmmParameters (src/CharmmPSF.cu, src/CharmmCrd.cu,
// src/CharmmParameters.cu), ForceManager::initialize
// (src/ForceManager.cu:436-550), CharmmContext (src/CharmmContext.cu:27-48,
// 96-125, 478-527, 731-741, 919-945), CudaIntegrator::propagate /
// reportIfNeeded (src/CudaIntegrator.cu:129-212, 347+), RestartSubscriber and
// DcdSubscriber (src/RestartSubscriber.cu:73-445, src/DcdSubscriber.cu).
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "apocharmm_amd.h"
#include "apocharmm_amd_capi.h"
#include "capi_guard.hpp"
#include "charmm_files.hpp"
#include "system.hpp"

namespace apo {
NbSystem *create_system(const apo_nb_params &p, const apo_nb_topology &t);
void set_box(NbSystem &S, const double *b);
struct LtSystem;
LtSystem *create_lt(NbSystem *nb, const apo_lt_params &p, const double *xyz, const double *vel);
void destroy_lt(LtSystem *L);
void lt_propagate(LtSystem &L, int nsteps, hipStream_t caller, bool fresh);
void lt_get_state(LtSystem &L, double *xyz, double *vel, double *dprev, double *grad,
                  double *thermo, int64_t *step);
void lt_set_state(LtSystem &L, const double *xyz, const double *vel, const double *dprev,
                  unsigned long long step);
}  // namespace apo

namespace {

constexpr double kBoltzAkma = 1.987191e-03;  // include/Constants.h:17
constexpr double kUnset = -9999.9999;

struct FmObj {
  std::shared_ptr<const apo::files::Psf> psf;
  std::shared_ptr<const apo::files::Params> prm;
  double box[3] = {kUnset, kUnset, kUnset};
  double kappa = 0.34, cutnb = 14.0, ctonnb = 12.0, ctofnb = 10.0;
  int order = 4;
  int grid[3] = {-1, -1, -1};
  apo::files::Resolved topo;
  apo::NbSystem *sys = nullptr;
  apo_nb_system view{nullptr, true};  // lent out by apo_force_manager_get_nb_system
  ~FmObj() { delete sys; }

  bool initialized() const { return sys != nullptr; }
  // ForceManager::initialize (src/ForceManager.cu:436-550)
  void initialize() {
    if (sys) return;
    if (box[0] == kUnset || box[1] == kUnset || box[2] == kUnset)
      throw apo::InvalidArgument("Error: Box dimension was not set");
    if (cutnb <= 0.0 || cutnb > box[0] / 2.0)
      throw apo::InvalidArgument("Error: Cutoff value (" + std::to_string(cutnb) +
                                 ") not valid (boxx: " + std::to_string(box[0]) + ")");
    topo = apo::files::resolve(*psf, *prm);
    apo_nb_params p{};
    for (int d = 0; d < 3; d++) {
      p.box[d] = box[d];
      // computeFFTGridSize (src/ForceManager.cu:923-943) unless set
      p.fft_grid[d] = grid[d] > 0 ? grid[d] : std::max(2, 2 * ((int)box[d] / 2));
    }
    p.kappa = kappa;
    p.cutnb = cutnb;
    p.ctofnb = ctofnb;
    p.ctonnb = ctonnb;
    p.e14fac = 1.0;
    p.pme_order = order;
    p.precision = APO_PRECISION_MIXED;
    apo_nb_topology t{};
    t.natom = psf->natom();
    t.ntypes = topo.ntypes;
    t.vdwtype = topo.vdwtype.data();
    t.vdwtab = topo.vdwtab.data();
    t.vdwtab14 = topo.vdwtab14.data();
    t.nbond = (int32_t)topo.bonds.size() / 2;
    t.bonds = topo.bonds.data();
    t.bond_par = topo.bond_par.data();
    t.nureyb = (int32_t)topo.ureyb.size() / 2;
    t.ureyb = topo.ureyb.data();
    t.ub_par = topo.ub_par.data();
    t.nangle = (int32_t)topo.angles.size() / 3;
    t.angles = topo.angles.data();
    t.angle_par = topo.angle_par.data();
    t.ndihe = (int32_t)topo.dihe.size() / 4;
    t.dihe = topo.dihe.data();
    t.dihe_par = topo.dihe_par.data();
    t.nimdihe = (int32_t)topo.imdihe.size() / 4;
    t.imdihe = topo.imdihe.data();
    t.imdihe_par = topo.imdihe_par.data();
    sys = apo::create_system(p, t);
    view.sys = sys;
  }
};

struct IntObj;

struct CtxObj {
  std::shared_ptr<FmObj> fm;
  int natom = -1;
  std::vector<double> xyz, vel;  // atom-major; the integrator's state once one is bound
  uint64_t seed = 0;
  bool holonomic = true;
  float temperature = 300.0f;
  IntObj *bound = nullptr;  // integrator that holds the current state
  // CharmmContext::useHolonomicConstraints (src/CharmmContext.cu:919-945): P1
  // removes 3; constraints remove one each (3 per SETTLE water, 1 per SHAKE bond)
  int ndof() const {
    int n = 3 * natom - 3;
    if (holonomic) {
      n -= 3 * (int)(fm->topo.waters.size() / 3);
      for (size_t g = 0; g < fm->topo.shake_groups.size(); g += 4)
        for (int k = 1; k < 4; k++) n -= fm->topo.shake_groups[g + k] >= 0 ? 1 : 0;
    }
    return n;
  }
  void pull_state();
};

struct SubObj {
  enum Kind { kRestart, kDcd } kind;
  std::string path;
  int freq = 1000;
  int frames = 0;
  bool header = false;
};

struct IntObj {
  double dt_ps = 0.001;
  double friction = 0.0, temperature = 300.0;
  uint64_t seed = 0;
  std::shared_ptr<CtxObj> ctx;
  apo::LtSystem *lt = nullptr;
  std::vector<std::shared_ptr<SubObj>> subs;
  unsigned long long total_steps = 0;  // NPRIV
  int last_nsteps = 0;                 // NSTEP
  ~IntObj() { release(); }
  void release() {
    if (lt) {
      if (ctx && ctx->bound == this) {
        ctx->pull_state();
        ctx->bound = nullptr;
      }
      apo::destroy_lt(lt);
      lt = nullptr;
    }
  }
  // the integrator over the context's current state (created on first use
  // and again after a parameter change)
  apo::LtSystem &ensure() {
    if (!ctx) throw apo::NotInitialized("No CharmmContext object was set for this CudaIntegrator");
    if (lt) return *lt;
    CtxObj &c = *ctx;
    if (c.natom < 0) throw apo::NotInitialized("the context has no coordinates (setCoordinates)");
    if (c.bound && c.bound != this) c.bound->release();
    const auto &psf = *c.fm->psf;
    std::vector<double> mass(psf.masses);
    std::vector<float> charge(psf.charges.begin(), psf.charges.end());
    apo_lt_params p{};
    p.timestep = dt_ps;
    p.temperature = temperature;
    p.friction = friction;
    p.seed = seed;
    p.nbfreq = 20;
    p.comfreq = 1000;
    const auto &w = c.fm->topo.waters;
    p.nwater = c.holonomic ? (int32_t)(w.size() / 3) : 0;
    p.use_graph = 1;
    p.waters = w.data();
    p.mass = mass.data();
    p.charge = charge.data();
    p.launch_setting = 1;  // fixed launch order: no host-blocking timing trial
    if (c.holonomic && !c.fm->topo.shake_groups.empty()) {
      p.nshake = (int32_t)(c.fm->topo.shake_groups.size() / 4);
      p.shake_groups = c.fm->topo.shake_groups.data();
      p.shake_r2 = c.fm->topo.shake_r2.data();
    }
    lt = apo::create_lt(c.fm->sys, p, c.xyz.data(), c.vel.data());
    c.bound = this;
    return *lt;
  }
  void report(int istep);
};

void CtxObj::pull_state() {
  if (!bound || !bound->lt) return;
  int64_t step = 0;
  apo::lt_get_state(*bound->lt, xyz.data(), vel.data(), nullptr, nullptr, nullptr, &step);
}

// ---- subscribers -----------------------------------------------------------
void write_restart(IntObj &I, SubObj &s) {
  CtxObj &c = *I.ctx;
  apo::files::Restart r;
  for (int d = 0; d < 3; d++) r.box[d] = c.fm->box[d];
  r.natom = c.natom;
  r.npriv = I.total_steps;
  r.nstep = I.last_nsteps;
  r.nsavc = s.freq;
  r.ndegf = c.ndof();
  r.seed = I.seed;
  r.x.resize(3 * (size_t)c.natom);
  r.v.resize(3 * (size_t)c.natom);
  r.dprev.resize(3 * (size_t)c.natom);
  double thermo[6] = {0};
  int64_t step = 0;
  apo::lt_get_state(*I.lt, r.x.data(), r.v.data(), r.dprev.data(), nullptr, thermo, &step);
  r.rng_state = (unsigned long long)step;
  r.avetem = thermo[4];
  apo::files::write_restart(s.path, r);
}

// CHARMM DCD frames of the float coordinates (DcdSubscriber::writeHeader /
// writeXtalData / writeCoordData, src/DcdSubscriber.cu)
void write_dcd(IntObj &I, SubObj &s) {
  CtxObj &c = *I.ctx;
  const int n = c.natom;
  std::fstream f;
  if (!s.header) {
    f.open(s.path, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f.is_open()) throw apo::RuntimeError("could not open DCD file \"" + s.path + "\"");
    auto rec = [&](const void *p, int bytes) {
      f.write(reinterpret_cast<const char *>(&bytes), 4);
      f.write(reinterpret_cast<const char *>(p), bytes);
      f.write(reinterpret_cast<const char *>(&bytes), 4);
    };
    int32_t h[21] = {0};
    h[0] = 1146244931;  // 'C','O','R','D'
    h[3] = s.freq;
    h[8] = c.ndof();
    const float delta = (float)(I.dt_ps / 0.0488882129);
    std::memcpy(&h[10], &delta, 4);
    h[11] = 1;  // unit cell present
    h[20] = 49;
    rec(h, 84);
    char title[164];
    std::memset(title, ' ', sizeof(title));
    const int32_t ntitle = 2;
    std::memcpy(title, &ntitle, 4);
    const char *t1 = "* GENERATED BY APOCHARMM_AMD";
    std::memcpy(title + 4, t1, std::strlen(t1));
    title[84] = '*';
    rec(title, 164);
    const int32_t na = n;
    rec(&na, 4);
    s.header = true;
  } else {
    f.open(s.path, std::ios::in | std::ios::out | std::ios::binary);
    if (!f.is_open()) throw apo::RuntimeError("could not open DCD file \"" + s.path + "\"");
    f.seekp(0, std::ios::end);
  }
  auto rec = [&](const void *p, int bytes) {
    f.write(reinterpret_cast<const char *>(&bytes), 4);
    f.write(reinterpret_cast<const char *>(p), bytes);
    f.write(reinterpret_cast<const char *>(&bytes), 4);
  };
  double xtl[6] = {c.fm->box[0], 0.0, c.fm->box[1], 0.0, 0.0, c.fm->box[2]};
  rec(xtl, 48);
  std::vector<double> x(3 * (size_t)n);
  apo::lt_get_state(*I.lt, x.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
  std::vector<float> comp(n);
  for (int d = 0; d < 3; d++) {
    for (int i = 0; i < n; i++) comp[i] = (float)x[3 * (size_t)i + d];
    rec(comp.data(), 4 * n);
  }
  s.frames++;
  const int32_t nfile = s.frames, nstep = s.frames * s.freq;
  f.seekp(8);
  f.write(reinterpret_cast<const char *>(&nfile), 4);
  f.seekp(20);
  f.write(reinterpret_cast<const char *>(&nstep), 4);
  if (!f.good()) throw apo::RuntimeError("writing DCD file \"" + s.path + "\" failed");
}

void IntObj::report(int istep) {
  for (auto &s : subs)
    if (s->freq > 0 && istep % s->freq == 0) {
      if (s->kind == SubObj::kRestart) write_restart(*this, *s);
      else write_dcd(*this, *s);
    }
}

}  // namespace

// ---- handles ------------------------------------------------------------------
struct apo_charmm_psf {
  std::shared_ptr<const apo::files::Psf> object;
};
struct apo_charmm_crd {
  std::shared_ptr<const apo::files::Crd> object;
};
struct apo_charmm_parameters {
  std::shared_ptr<const apo::files::Params> object;
};
struct apo_force_manager {
  std::shared_ptr<FmObj> object;
};
struct apo_charmm_context {
  std::shared_ptr<CtxObj> object;
};
struct apo_subscriber {
  std::shared_ptr<SubObj> object;
};
struct apo_restart_subscriber {
  apo_subscriber base;
};
struct apo_dcd_subscriber {
  apo_subscriber base;
};
struct apo_cuda_integrator {
  std::shared_ptr<IntObj> object;
};
struct apo_cuda_langevin_thermostat_integrator {
  apo_cuda_integrator base;
};

namespace {
using apo::guard;

template <typename H> auto &obj(H *h, const char *what) {
  if (!h || !h->object) throw apo::InvalidArgument(std::string(what) + " object is NULL");
  return *h->object;
}
void need(const void *p, const char *what) {
  if (!p) throw apo::InvalidArgument(std::string(what) + " pointer is NULL");
}
void need_path(const char *p, const char *what) {
  if (!p || !*p) throw apo::InvalidArgument(std::string(what) + " is NULL or empty");
}
void need_len(size_t len, size_t req, const char *what) {
  if (len < req) throw apo::InvalidArgument(std::string(what) + " is too small");
}
void put8(char *out, const std::vector<std::string> &v) {
  for (size_t i = 0; i < v.size(); i++)
    for (size_t j = 0; j < 8; j++) out[i * 8 + j] = j < v[i].size() ? v[i][j] : ' ';
}
template <typename H>
apo_status make_sub(H **out, const char *path, int freq, SubObj::Kind kind) {
  return guard([&] {
    need(out, "out");
    *out = nullptr;
    need_path(path, "subscriber file path");
    if (freq <= 0) throw apo::InvalidArgument("report frequency must be positive");
    auto h = std::make_unique<H>();
    h->base.object = std::make_shared<SubObj>();
    h->base.object->kind = kind;
    h->base.object->path = path;
    h->base.object->freq = freq;
    *out = h.release();
  });
}
}  // namespace

extern "C" {

// ---- CharmmPsf -----------------------------------------------------------------
APO_API apo_status apo_charmm_psf_create(apo_charmm_psf **out, const char *path) {
  return guard([&] {
    need(out, "out");
    *out = nullptr;
    need_path(path, "PSF path");
    auto h = std::make_unique<apo_charmm_psf>();
    h->object = std::make_shared<apo::files::Psf>(apo::files::read_psf(path));
    *out = h.release();
  });
}
APO_API void apo_charmm_psf_destroy(apo_charmm_psf *psf) { delete psf; }

#define APO_PSF_COUNT(fn, expr)                                               \
  APO_API apo_status fn(size_t *num, const apo_charmm_psf *psf) {            \
    return guard([&] {                                                        \
      need(num, "count");                                                     \
      *num = 0;                                                               \
      const auto &p = obj(psf, "CharmmPsf");                                  \
      *num = (size_t)(expr);                                                  \
    });                                                                       \
  }
APO_PSF_COUNT(apo_charmm_psf_get_num_atoms, p.natom())
APO_PSF_COUNT(apo_charmm_psf_get_num_bonds, p.bonds.size() / 2)
APO_PSF_COUNT(apo_charmm_psf_get_num_angles, p.angles.size() / 3)
APO_PSF_COUNT(apo_charmm_psf_get_num_dihedrals, p.dihedrals.size() / 4)
APO_PSF_COUNT(apo_charmm_psf_get_num_impropers, p.impropers.size() / 4)
APO_PSF_COUNT(apo_charmm_psf_get_num_cross_terms, p.crossterms.size() / 8)
#undef APO_PSF_COUNT

#define APO_PSF_STRINGS(fn, member, what)                                     \
  APO_API apo_status fn(char *out, const size_t len, const apo_charmm_psf *psf) { \
    return guard([&] {                                                        \
      const auto &p = obj(psf, "CharmmPsf");                                  \
      need(out, what);                                                        \
      need_len(len, 8 * p.member.size(), what);                               \
      put8(out, p.member);                                                    \
    });                                                                       \
  }
APO_PSF_STRINGS(apo_charmm_psf_get_segment_identifiers, segids, "Segment identifier buffer")
APO_PSF_STRINGS(apo_charmm_psf_get_residue_names, resnames, "Residue name buffer")
APO_PSF_STRINGS(apo_charmm_psf_get_atom_names, names, "Atom name buffer")
APO_PSF_STRINGS(apo_charmm_psf_get_atom_types, types, "Atom type buffer")
#undef APO_PSF_STRINGS

APO_API apo_status apo_charmm_psf_get_residue_identifiers(int *resis, const size_t len,
                                                          const apo_charmm_psf *psf) {
  return guard([&] {
    const auto &p = obj(psf, "CharmmPsf");
    need(resis, "Residue identifier buffer");
    need_len(len, p.resids.size(), "Residue identifier buffer");
    std::copy(p.resids.begin(), p.resids.end(), resis);
  });
}
APO_API apo_status apo_charmm_psf_get_charges(double *charges, const size_t len,
                                              const apo_charmm_psf *psf) {
  return guard([&] {
    const auto &p = obj(psf, "CharmmPsf");
    need(charges, "Charge buffer");
    need_len(len, p.charges.size(), "Charge buffer");
    std::copy(p.charges.begin(), p.charges.end(), charges);
  });
}
APO_API apo_status apo_charmm_psf_get_masses(double *masses, const size_t len,
                                             const apo_charmm_psf *psf) {
  return guard([&] {
    const auto &p = obj(psf, "CharmmPsf");
    need(masses, "Mass buffer");
    need_len(len, p.masses.size(), "Mass buffer");
    std::copy(p.masses.begin(), p.masses.end(), masses);
  });
}
APO_API apo_status apo_charmm_psf_get_net_charge(double *net_charge, const apo_charmm_psf *psf) {
  return guard([&] {
    need(net_charge, "net_charge");
    double s = 0.0;
    for (double q : obj(psf, "CharmmPsf").charges) s += q;
    *net_charge = s;
  });
}
APO_API apo_status apo_charmm_psf_get_total_mass(double *total_mass, const apo_charmm_psf *psf) {
  return guard([&] {
    need(total_mass, "total_mass");
    double s = 0.0;
    for (double m : obj(psf, "CharmmPsf").masses) s += m;
    *total_mass = s;
  });
}
APO_API apo_status apo_charmm_psf_get_file_name(char *file_name, const size_t len,
                                                const apo_charmm_psf *psf) {
  return guard([&] {
    const auto &p = obj(psf, "CharmmPsf");
    need(file_name, "File name buffer");
    need_len(len, p.file_name.size() + 1, "File name buffer");
    std::memcpy(file_name, p.file_name.c_str(), p.file_name.size() + 1);
  });
}

// ---- CharmmCrd -------------------------------------------------------------------
APO_API apo_status apo_charmm_crd_create(apo_charmm_crd **out, const char *path) {
  return guard([&] {
    need(out, "out");
    *out = nullptr;
    need_path(path, "CRD path");
    auto h = std::make_unique<apo_charmm_crd>();
    h->object = std::make_shared<apo::files::Crd>(apo::files::read_crd(path));
    *out = h.release();
  });
}
APO_API void apo_charmm_crd_destroy(apo_charmm_crd *crd) { delete crd; }
APO_API apo_status apo_charmm_crd_get_num_atoms(size_t *num_atoms, const apo_charmm_crd *crd) {
  return guard([&] {
    need(num_atoms, "num_atoms");
    *num_atoms = (size_t)obj(crd, "CharmmCrd").natom();
  });
}
APO_API apo_status apo_charmm_crd_get_coordinates(double *xyz, const size_t xyz_len,
                                                  const apo_charmm_crd *crd) {
  return guard([&] {
    const auto &c = obj(crd, "CharmmCrd");
    need(xyz, "Coordinate buffer");
    need_len(xyz_len, c.xyz.size(), "Coordinate buffer");
    std::copy(c.xyz.begin(), c.xyz.end(), xyz);
  });
}

// ---- CharmmParameters --------------------------------------------------------------
APO_API apo_status apo_charmm_parameters_create(apo_charmm_parameters **out, const char *path) {
  return apo_charmm_parameters_create_from_files(out, &path, 1);
}
APO_API apo_status apo_charmm_parameters_create_from_files(apo_charmm_parameters **out,
                                                           const char *const *paths,
                                                           const size_t num_paths) {
  return guard([&] {
    need(out, "out");
    *out = nullptr;
    need(paths, "paths");
    if (num_paths == 0) throw apo::InvalidArgument("no parameter files given");
    auto prm = std::make_shared<apo::files::Params>();
    for (size_t i = 0; i < num_paths; i++) {
      need_path(paths[i], "parameter file path");
      prm->read(paths[i]);
    }
    auto h = std::make_unique<apo_charmm_parameters>();
    h->object = prm;
    *out = h.release();
  });
}
APO_API void apo_charmm_parameters_destroy(apo_charmm_parameters *parameters) {
  delete parameters;
}

// ---- ForceManager -------------------------------------------------------------------
APO_API apo_status apo_force_manager_create(apo_force_manager **out, const apo_charmm_psf *psf,
                                            const apo_charmm_parameters *parameters) {
  return guard([&] {
    need(out, "out");
    *out = nullptr;
    auto fm = std::make_shared<FmObj>();
    obj(psf, "CharmmPsf");
    obj(parameters, "CharmmParameters");
    fm->psf = psf->object;
    fm->prm = parameters->object;
    auto h = std::make_unique<apo_force_manager>();
    h->object = fm;
    *out = h.release();
  });
}
APO_API void apo_force_manager_destroy(apo_force_manager *force_manager) {
  delete force_manager;
}
APO_API apo_status apo_force_manager_set_box_dimensions(apo_force_manager *force_manager,
                                                        const double x, const double y,
                                                        const double z) {
  return guard([&] {
    FmObj &fm = obj(force_manager, "ForceManager");
    if (!(x > 0 && y > 0 && z > 0))
      throw apo::InvalidArgument("box dimensions must be positive");
    const double b[3] = {x, y, z};
    if (fm.sys) apo::set_box(*fm.sys, b);
    for (int d = 0; d < 3; d++) fm.box[d] = b[d];
  });
}


#define APO_FM_SET(fn, field, check)                                             \
  APO_API apo_status fn(apo_force_manager *force_manager, const double v) {      \
    return guard([&] {                                                           \
      FmObj &fm = obj(force_manager, "ForceManager");                            \
      if (fm.sys) throw apo::InvalidArgument("the ForceManager is already initialized"); \
      if (!(check)) throw apo::InvalidArgument(#field " out of range");          \
      fm.field = v;                                                              \
    });                                                                          \
  }
APO_FM_SET(apo_force_manager_set_kappa, kappa, v > 0)
APO_FM_SET(apo_force_manager_set_cutoff, cutnb, v > 0)
APO_FM_SET(apo_force_manager_set_ctonnb, ctonnb, v >= 0)
APO_FM_SET(apo_force_manager_set_ctofnb, ctofnb, v > 0)
#undef APO_FM_SET

APO_API apo_status apo_force_manager_set_fft_grid(apo_force_manager *force_manager, const int nx,
                                                  const int ny, const int nz) {
  return guard([&] {
    FmObj &fm = obj(force_manager, "ForceManager");
    if (fm.sys) throw apo::InvalidArgument("the ForceManager is already initialized");
    fm.grid[0] = nx;
    fm.grid[1] = ny;
    fm.grid[2] = nz;
  });
}

APO_API apo_status apo_force_manager_get_nb_system(apo_force_manager *force_manager,
                                                   apo_nb_system **out) {
  return guard([&] {
    need(out, "out");
    *out = nullptr;
    FmObj &fm = obj(force_manager, "ForceManager");
    if (!fm.sys)
      throw apo::NotInitialized("the ForceManager is not initialized (create a context on it)");
    *out = &fm.view;
  });
}

// ---- CharmmContext ----------------------------------------------------------------------
APO_API apo_status apo_charmm_context_create(apo_charmm_context **out,
                                             const apo_force_manager *force_manager) {
  return guard([&] {
    need(out, "out");
    *out = nullptr;
    obj(force_manager, "ForceManager");
    auto ctx = std::make_shared<CtxObj>();
    ctx->fm = force_manager->object;
    ctx->fm->initialize();  // the constructor initialises its manager
    ctx->seed = std::random_device{}();
    auto h = std::make_unique<apo_charmm_context>();
    h->object = ctx;
    *out = h.release();
  });
}
APO_API void apo_charmm_context_destroy(apo_charmm_context *context) { delete context; }

APO_API apo_status apo_charmm_context_set_coordinates(apo_charmm_context *context,
                                                      const apo_charmm_crd *crd) {
  return guard([&] {
    CtxObj &c = obj(context, "CharmmContext");
    const auto &x = obj(crd, "CharmmCrd");
    if (x.natom() != c.fm->psf->natom())
      throw apo::InvalidArgument("coordinate count " + std::to_string(x.natom()) +
                                 " does not match the PSF's " +
                                 std::to_string(c.fm->psf->natom()));
    if (c.bound) c.bound->release();  // a new state: integrators start again from it
    c.natom = x.natom();
    c.xyz = x.xyz;
    c.vel.assign(3 * (size_t)c.natom, 0.0);
  });
}

APO_API apo_status apo_charmm_context_set_random_seed_for_velocities(apo_charmm_context *context,
                                                                     const uint64_t seed) {
  return guard([&] { obj(context, "CharmmContext").seed = seed; });
}

APO_API apo_status apo_charmm_context_use_holonomic_constraints(apo_charmm_context *context,
                                                                const bool useHolonomicConstraints) {
  return guard([&] {
    CtxObj &c = obj(context, "CharmmContext");
    if (c.bound) c.bound->release();
    c.holonomic = useHolonomicConstraints;
  });
}

// CharmmContext::assignVelocitiesAtTemperature (src/CharmmContext.cu:478-527):
// per atom a fresh normal distribution of width sqrt(kB T / m) draws x, y, z
// from one mt19937 seeded with the context's seed; the centre-of-mass motion
// is left as drawn (removeCenterOfMassMotion has no effect there, :450-476).
APO_API apo_status apo_charmm_context_assign_velocities_at_temperature(
    apo_charmm_context *context, const double temperature) {
  return guard([&] {
    CtxObj &c = obj(context, "CharmmContext");
    if (c.natom < 0)
      throw apo::InvalidArgument("numAtoms = -1 in assignVelocitiesAtTemperature: set the "
                                 "coordinates first");
    if (c.bound) c.bound->release();
    c.temperature = (float)temperature;
    const double boltz = kBoltzAkma * c.temperature;
    std::mt19937 gen{(std::mt19937::result_type)c.seed};
    const auto &m = c.fm->psf->masses;
    for (int i = 0; i < c.natom; i++) {
      const double sd = std::sqrt(boltz * (1.0 / m[i]));
      std::normal_distribution<> d(0, sd);
      c.vel[3 * (size_t)i] = d(gen);
      c.vel[3 * (size_t)i + 1] = d(gen);
      c.vel[3 * (size_t)i + 2] = d(gen);
    }
  });
}

// CharmmContext::computeTemperature (src/CharmmContext.cu:731-741)
APO_API apo_status apo_charmm_context_compute_temperature(double *temperature,
                                                          apo_charmm_context *context) {
  return guard([&] {
    need(temperature, "temperature");
    CtxObj &c = obj(context, "CharmmContext");
    if (c.natom < 0) throw apo::InvalidArgument("No atoms in the system");
    c.pull_state();
    const auto &m = c.fm->psf->masses;
    double ke = 0.0;
    for (int i = 0; i < c.natom; i++) {
      const double *v = &c.vel[3 * (size_t)i];
      ke += m[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    ke *= 0.5;
    *temperature = ke / (0.5 * c.ndof() * kBoltzAkma);
  });
}

// ---- subscribers --------------------------------------------------------------------------
APO_API apo_status apo_subscriber_set_report_frequency(apo_subscriber *subscriber,
                                                       const int report_frequency) {
  return guard([&] {
    if (report_frequency <= 0) throw apo::InvalidArgument("report frequency must be positive");
    obj(subscriber, "Subscriber").freq = report_frequency;
  });
}
APO_API apo_status apo_subscriber_get_report_frequency(int *report_frequency,
                                                       const apo_subscriber *subscriber) {
  return guard([&] {
    need(report_frequency, "report_frequency");
    *report_frequency = obj(subscriber, "Subscriber").freq;
  });
}

APO_API apo_status apo_restart_subscriber_create(apo_restart_subscriber **out, const char *path) {
  return make_sub(out, path, 1000, SubObj::kRestart);
}
APO_API apo_status apo_restart_subscriber_create_with_report_frequency(
    apo_restart_subscriber **out, const char *path, const int report_frequency) {
  return make_sub(out, path, report_frequency, SubObj::kRestart);
}
APO_API void apo_restart_subscriber_destroy(apo_restart_subscriber *subscriber) {
  delete subscriber;
}
APO_API apo_status apo_restart_subscriber_as_subscriber(apo_subscriber **out,
                                                        apo_restart_subscriber *subscriber) {
  return guard([&] {
    need(out, "out");
    need(subscriber, "RestartSubscriber");
    *out = &subscriber->base;
  });
}
APO_API apo_status apo_dcd_subscriber_create(apo_dcd_subscriber **out, const char *path) {
  return make_sub(out, path, 1000, SubObj::kDcd);
}
APO_API apo_status apo_dcd_subscriber_create_with_report_frequency(apo_dcd_subscriber **out,
                                                                   const char *path,
                                                                   const int report_frequency) {
  return make_sub(out, path, report_frequency, SubObj::kDcd);
}
APO_API void apo_dcd_subscriber_destroy(apo_dcd_subscriber *subscriber) { delete subscriber; }
APO_API apo_status apo_dcd_subscriber_as_subscriber(apo_subscriber **out,
                                                    apo_dcd_subscriber *subscriber) {
  return guard([&] {
    need(out, "out");
    need(subscriber, "DcdSubscriber");
    *out = &subscriber->base;
  });
}

// ---- CudaIntegrator ----------------------------------------------------------------------
APO_API apo_status apo_cuda_integrator_set_time_step(apo_cuda_integrator *integrator,
                                                     const double time_step) {
  return guard([&] {
    IntObj &I = obj(integrator, "CudaIntegrator");
    if (!(time_step > 0)) throw apo::InvalidArgument("time step must be positive (ps)");
    I.release();
    I.dt_ps = time_step;
  });
}
APO_API apo_status apo_cuda_integrator_set_charmm_context(apo_cuda_integrator *integrator,
                                                          apo_charmm_context *context) {
  return guard([&] {
    IntObj &I = obj(integrator, "CudaIntegrator");
    obj(context, "CharmmContext");
    I.release();
    I.ctx = context->object;
  });
}
APO_API apo_status apo_cuda_integrator_subscribe(apo_cuda_integrator *integrator,
                                                 apo_subscriber *subscriber) {
  return guard([&] {
    IntObj &I = obj(integrator, "CudaIntegrator");
    obj(subscriber, "Subscriber");
    I.subs.push_back(subscriber->object);
  });
}
// CudaIntegrator::propagate (src/CudaIntegrator.cu:129-212): one list reset
// and centering at the start, steps 1..n, each subscriber updated after the
// steps that are multiples of its frequency; the steps between reports run as
// one fused propagate piece, and later pieces continue the list cadence.
APO_API apo_status apo_cuda_integrator_propagate(apo_cuda_integrator *integrator,
                                                 const int num_steps) {
  return guard([&] {
    IntObj &I = obj(integrator, "CudaIntegrator");
    if (num_steps < 0) throw apo::InvalidArgument("num_steps must be >= 0");
    apo::LtSystem &L = I.ensure();
    I.last_nsteps = num_steps;
    int done = 0;
    while (done < num_steps) {
      int next = num_steps;
      for (auto &s : I.subs)
        if (s->freq > 0) next = std::min(next, (done / s->freq + 1) * s->freq);
      apo::lt_propagate(L, next - done, nullptr, done == 0);
      I.total_steps += (unsigned long long)(next - done);
      done = next;
      I.report(done);
    }
    APO_HIP(hipStreamSynchronize(nullptr));
  });
}
APO_API apo_status apo_cuda_integrator_initialize_from_restart_file(
    apo_cuda_integrator *integrator, const char *path) {
  return guard([&] {
    IntObj &I = obj(integrator, "CudaIntegrator");
    need_path(path, "restart file path");
    if (!I.ctx)
      throw apo::RuntimeError("CharmmContext must be set before initializing from a restart file");
    const apo::files::Restart r = apo::files::read_restart(path);
    CtxObj &c = *I.ctx;
    if (r.natom != c.natom)
      throw apo::InvalidArgument("NATOM mismatch in restart file \"" + std::string(path) + "\"");
    if (r.ndegf != c.ndof())
      throw apo::InvalidArgument("NDEGF mismatch in restart file \"" + std::string(path) + "\"");
    I.release();
    const double b[3] = {r.box[0], r.box[1], r.box[2]};
    apo::set_box(*c.fm->sys, b);
    for (int d = 0; d < 3; d++) c.fm->box[d] = b[d];
    I.seed = r.seed;
    I.total_steps = r.npriv;
    I.last_nsteps = r.nstep;
    c.xyz = r.x;
    c.vel = r.v;
    apo::LtSystem &L = I.ensure();
    apo::lt_set_state(L, r.x.data(), r.v.data(), r.dprev.data(), r.rng_state);
  });
}
APO_API apo_status apo_cuda_integrator_get_state(apo_cuda_integrator *integrator, double *xyz,
                                                 double *vel, double *dprev,
                                                 unsigned long long *total_steps) {
  return guard([&] {
    IntObj &I = obj(integrator, "CudaIntegrator");
    apo::LtSystem &L = I.ensure();
    apo::lt_get_state(L, xyz, vel, dprev, nullptr, nullptr, nullptr);
    if (total_steps) *total_steps = I.total_steps;
  });
}

// ---- CudaLangevinThermostatIntegrator -------------------------------------------------
APO_API apo_status apo_cuda_langevin_thermostat_integrator_create(
    apo_cuda_langevin_thermostat_integrator **out, const double time_step) {
  return guard([&] {
    need(out, "out");
    *out = nullptr;
    if (!(time_step > 0)) throw apo::InvalidArgument("time step must be positive (ps)");
    auto h = std::make_unique<apo_cuda_langevin_thermostat_integrator>();
    h->base.object = std::make_shared<IntObj>();
    h->base.object->dt_ps = time_step;
    *out = h.release();
  });
}
APO_API void apo_cuda_langevin_thermostat_integrator_destroy(
    apo_cuda_langevin_thermostat_integrator *integrator) {
  delete integrator;
}
#define APO_LT_SET(fn, type, field, check)                                        \
  APO_API apo_status fn(apo_cuda_langevin_thermostat_integrator *integrator, const type v) { \
    return guard([&] {                                                            \
      need(integrator, "CudaLangevinThermostatIntegrator");                       \
      IntObj &I = obj(&integrator->base, "CudaLangevinThermostatIntegrator");      \
      if (!(check)) throw apo::InvalidArgument(#field " out of range");           \
      I.release();                                                                \
      I.field = v;                                                                \
    });                                                                           \
  }
APO_LT_SET(apo_cuda_langevin_thermostat_integrator_set_friction, double, friction, v >= 0)
APO_LT_SET(apo_cuda_langevin_thermostat_integrator_set_reference_temperature, double, temperature,
           v >= 0)
APO_LT_SET(apo_cuda_langevin_thermostat_integrator_set_rng_seed, uint64_t, seed, true)
#undef APO_LT_SET
APO_API apo_status apo_cuda_langevin_thermostat_integrator_as_cuda_integrator(
    apo_cuda_integrator **out, apo_cuda_langevin_thermostat_integrator *integrator) {
  return guard([&] {
    need(out, "out");
    need(integrator, "CudaLangevinThermostatIntegrator");
    *out = &integrator->base;
  });
}

}  // extern "C"
