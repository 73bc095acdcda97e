// C++ host mirror of the reference's ForceManager surface
// (include/ForceManager.h:200-476) over the C ABI in apocharmm_amd.h.
//
// Method names, argument meaning and error behaviour follow the reference:
// setters before initialize(), calcForce(xyzq, reset, calcEnergy, calcVirial)
// with device xyzq, forces as the int64 2^-40 gradient (Force<long long>),
// std::invalid_argument / std::runtime_error on failure, plugin forces
// through subscribe().  Header-only; link against libapocharmm_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../apocharmm_amd.h"

namespace apocharmm_amd {

inline void check(apo_status s) {
  if (s == APO_OK) return;
  std::string msg = apo_last_error();
  if (s == APO_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// Resolved topology/parameters (what the reference pulls from CharmmPSF +
// CharmmParameters in ForceManager::initialize, src/ForceManager.cu:463-527).
struct Topology {
  int natom = 0;
  int ntypes = 0;
  std::vector<int32_t> vdwtype;
  std::vector<float> vdwtab, vdwtab14;  // (c6, c12) per lower-triangular type pair
  std::vector<int32_t> bonds, ureyb, angles, dihe, imdihe;
  std::vector<float> bond_par, ub_par, angle_par, dihe_par, imdihe_par;
};

// Type-erased plugin force (ForceView, include/ForceManager.h:37-105).  A
// plugin type F provides
//   void initialize(int numAtoms, const std::vector<double> &box);
//   void clear();
//   void calcForce(const float4 *xyzq, bool calcEnergy, bool calcVirial);
//   void setBoxDimensions(const std::vector<double> &box);
//   int64_t *getForce();         // device int64 2^-40 gradient, laid out with
//                                // the manager's getForceStride()
//   double *getEnergy();         // device double[1]: the plugin's energy
// and enqueues its work on the manager's stream (getStream(), e.g. handed to
// the plugin when it is constructed), so the manager's sums are ordered after
// it without a host synchronisation.
class ForceView {
 public:
  template <typename F>
  explicit ForceView(std::shared_ptr<F> f)
      : obj_(f),
        initialize_([](void *p, int n, const std::vector<double> &b) {
          static_cast<F *>(p)->initialize(n, b);
        }),
        clear_([](void *p) { static_cast<F *>(p)->clear(); }),
        calc_([](void *p, const float4 *x, bool e, bool v) {
          static_cast<F *>(p)->calcForce(x, e, v);
        }),
        box_([](void *p, const std::vector<double> &b) {
          static_cast<F *>(p)->setBoxDimensions(b);
        }),
        force_([](void *p) -> int64_t * { return static_cast<F *>(p)->getForce(); }),
        energy_([](void *p) -> double * { return static_cast<F *>(p)->getEnergy(); }) {}
  void initialize(int n, const std::vector<double> &box) { initialize_(obj_.get(), n, box); }
  void clear() { clear_(obj_.get()); }
  void calcForce(const float4 *x, bool e, bool v) { calc_(obj_.get(), x, e, v); }
  void setBoxDimensions(const std::vector<double> &box) { box_(obj_.get(), box); }
  int64_t *getForce() { return force_(obj_.get()); }
  double *getEnergy() { return energy_(obj_.get()); }

 private:
  std::shared_ptr<void> obj_;
  void (*initialize_)(void *, int, const std::vector<double> &);
  void (*clear_)(void *);
  void (*calc_)(void *, const float4 *, bool, bool);
  void (*box_)(void *, const std::vector<double> &);
  int64_t *(*force_)(void *);
  double *(*energy_)(void *);
};

class ForceManager {
 public:
  explicit ForceManager(Topology topo) : topo_(std::move(topo)) {
    params_.kappa = 0.34;
    params_.cutnb = 14.0;
    params_.ctofnb = 12.0;
    params_.ctonnb = 10.0;
    params_.e14fac = 1.0;
    params_.pme_order = 4;
    params_.precision = APO_PRECISION_MIXED;
  }
  ~ForceManager() { release(); }
  ForceManager(const ForceManager &) = delete;
  ForceManager &operator=(const ForceManager &) = delete;

  // ---- setters (include/ForceManager.h:200-231) ----
  void setBoxDimensions(const std::vector<double> &box) {
    for (int d = 0; d < 3; d++) params_.box[d] = box.at(d);
    if (sys_) check(apo_nb_set_box(sys_, params_.box));
    for (auto &v : views_) v.setBoxDimensions(getBoxDimensions());
  }
  std::vector<double> getBoxDimensions() const {
    return {params_.box[0], params_.box[1], params_.box[2]};
  }
  void setFFTGrid(int nx, int ny, int nz) {
    params_.fft_grid[0] = nx;
    params_.fft_grid[1] = ny;
    params_.fft_grid[2] = nz;
  }
  void setKappa(double kappa) { params_.kappa = kappa; }
  void setCutoff(double cutnb) { params_.cutnb = cutnb; }
  void setCtonnb(double ctonnb) { params_.ctonnb = ctonnb; }
  void setCtofnb(double ctofnb) { params_.ctofnb = ctofnb; }
  void setPmeSplineOrder(int order) { params_.pme_order = order; }
  void setPrecision(int precision) { params_.precision = precision; }

  // ---- set-up (ForceManager::initialize, src/ForceManager.cu:436-550) ----
  void initialize(hipStream_t stream = nullptr) {
    release();
    stream_ = stream;
    apo_nb_topology t{};
    t.natom = topo_.natom;
    t.ntypes = topo_.ntypes;
    t.vdwtype = topo_.vdwtype.data();
    t.vdwtab = topo_.vdwtab.data();
    t.vdwtab14 = topo_.vdwtab14.empty() ? nullptr : topo_.vdwtab14.data();
    t.nbond = (int32_t)topo_.bonds.size() / 2;
    t.bonds = topo_.bonds.data();
    t.bond_par = topo_.bond_par.data();
    t.nureyb = (int32_t)topo_.ureyb.size() / 2;
    t.ureyb = topo_.ureyb.data();
    t.ub_par = topo_.ub_par.data();
    t.nangle = (int32_t)topo_.angles.size() / 3;
    t.angles = topo_.angles.data();
    t.angle_par = topo_.angle_par.data();
    t.ndihe = (int32_t)topo_.dihe.size() / 4;
    t.dihe = topo_.dihe.data();
    t.dihe_par = topo_.dihe_par.data();
    t.nimdihe = (int32_t)topo_.imdihe.size() / 4;
    t.imdihe = topo_.imdihe.data();
    t.imdihe_par = topo_.imdihe_par.data();
    check(apo_nb_create(&params_, &t, &sys_));
    check(apo_nb_force_stride(sys_, &stride_));
    alloc(reinterpret_cast<void **>(&force_), sizeof(int64_t) * 3 * (size_t)stride_);
    alloc(reinterpret_cast<void **>(&force_d_), sizeof(double) * 3 * (size_t)stride_);
    alloc(reinterpret_cast<void **>(&energy_), sizeof(double) * APO_NUM_ENERGY);
    alloc(reinterpret_cast<void **>(&virial_), sizeof(double) * 9);
    for (auto &v : views_) v.initialize(topo_.natom, getBoxDimensions());
  }

  // ForceManager::subscribe (include/ForceManager.h:427-444): the plugin's
  // gradient and energy are added to the totals after every calcForce
  // (src/ForceManager.cu:678-682, 774-778).  Plugins subscribed before
  // initialize() are initialised with it.
  template <typename F>
  void subscribe(std::shared_ptr<F> force, const std::string &tag = "") {
    if (!force) throw std::invalid_argument("subscribe: null force");
    views_.emplace_back(std::move(force));
    tags_.push_back(tag);
    if (sys_) views_.back().initialize(topo_.natom, getBoxDimensions());
  }
  size_t getNumSubscribedForces() const { return views_.size(); }
  hipStream_t getStream() const { return stream_; }

  void resetNeighborList(const float4 *xyzq) {
    check(apo_nb_reset_list(sys_, reinterpret_cast<const float *>(xyzq), stream_));
  }

  // ForceManager::calcForce (include/ForceManager.h:424)
  void calcForce(const float4 *xyzq, bool reset = false, bool calcEnergy = false,
                 bool calcVirial = false) {
    int32_t flags = (reset ? APO_RESET_LIST : 0) | (calcEnergy ? APO_CALC_ENERGY : 0) |
                    (calcVirial ? APO_CALC_VIRIAL : 0);
    for (auto &v : views_) {
      v.clear();
      v.calcForce(xyzq, calcEnergy, calcVirial);
    }
    check(apo_nb_calc_force(sys_, reinterpret_cast<const float *>(xyzq), flags, force_, stride_,
                            energy_, virial_, stream_));
    for (auto &v : views_) {
      check(apo_nb_add_force(v.getForce(), topo_.natom, stride_, force_, stream_));
      if (calcEnergy) check(apo_nb_add_energy(v.getEnergy(), energy_, stream_));
    }
  }

  int getForceStride() const { return stride_; }
  int getNumAtoms() const { return topo_.natom; }
  // The C-ABI handle of the force path (for apo_nb_* calls of its own).
  apo_nb_system *handle() const { return sys_; }
  double *getEnergyBuffer() const { return energy_; }
  double *getVirialBuffer() const { return virial_; }
  // Device int64 gradient (Force<long long>).
  int64_t *getForcesFixed() const { return force_; }
  // Device double gradient, SoA with getForceStride() (Force<double>).
  double *getForces() {
    check(apo_nb_force_to_double(force_, topo_.natom, stride_, force_d_, stream_));
    return force_d_;
  }
  // Device double[1]-style pointer to the total potential energy.
  double *getPotentialEnergy() const { return energy_ + APO_E_EPOT; }
  std::map<std::string, double> getEnergyComponents() const {
    static const char *names[APO_NUM_ENERGY] = {"bond", "angle", "ureyb", "dihe",
                                                 "imdihe", "ewks", "ewse", "ewex",
                                                 "elec", "vdw", "epot"};
    std::array<double, APO_NUM_ENERGY> h{};
    if (hipStreamSynchronize(stream_) != hipSuccess ||
        hipMemcpy(h.data(), energy_, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("energy copy failed");
    std::map<std::string, double> out;
    for (int k = 0; k < APO_NUM_ENERGY; k++) out[names[k]] = h[k];
    return out;
  }
  std::vector<double> getVirial() const {
    std::vector<double> v(9);
    if (hipStreamSynchronize(stream_) != hipSuccess ||
        hipMemcpy(v.data(), virial_, 9 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("virial copy failed");
    return v;
  }

 private:
  static void alloc(void **p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed");
  }
  void release() {
    for (void *p : {(void *)force_, (void *)force_d_, (void *)energy_, (void *)virial_})
      if (p) (void)hipFree(p);
    force_ = nullptr;
    force_d_ = nullptr;
    energy_ = nullptr;
    virial_ = nullptr;
    if (sys_) (void)apo_nb_destroy(sys_);
    sys_ = nullptr;
  }

  Topology topo_;
  std::vector<ForceView> views_;
  std::vector<std::string> tags_;
  apo_nb_params params_{};
  apo_nb_system *sys_ = nullptr;
  hipStream_t stream_ = nullptr;
  int32_t stride_ = 0;
  int64_t *force_ = nullptr;
  double *force_d_ = nullptr;
  double *energy_ = nullptr;
  double *virial_ = nullptr;
};

}  // namespace apocharmm_amd
