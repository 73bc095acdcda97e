// C ABI entry points (include/apocharmm_amd.h).  Error handling mirrors the
// reference's C ABI: every body runs inside guard(), which maps exceptions to
// apo_status and keeps the message in a thread-local string
// (capi/private_include/apocharmm_c/detail/ErrorInternal.h:27-40,
// capi/src/apocharmm_c/Error.cpp:18-22).
#include <cstring>
#include <string>

#include "apocharmm_amd.h"
#include "capi_guard.hpp"
#include "system.hpp"

namespace apo {
NbSystem *create_system(const apo_nb_params &p, const apo_nb_topology &t);
void set_box(NbSystem &S, const double *b);
void calc_force(NbSystem &S, const float *xyzq, int flags, int64_t *force, int stride,
                double *energy, double *virial, hipStream_t s);
void calc_force_noalloc(NbSystem &S, const float *xyzq, int flags, int64_t *force, int stride,
                        double *energy, double *virial, hipStream_t s);
}  // namespace apo

namespace apo {
struct LtSystem;
LtSystem *create_lt(NbSystem *nb, const apo_lt_params &p, const double *xyz, const double *vel);
void destroy_lt(LtSystem *L);
void lt_initialize(LtSystem &L, hipStream_t caller);
void lt_propagate(LtSystem &L, int nsteps, hipStream_t caller);
void lt_propagate_batch(LtSystem **Ls, int n, int nsteps, hipStream_t caller);
struct LtBatch;
LtBatch *create_batch(LtSystem **Ls, int n, int lanes);
void destroy_batch(LtBatch *B);
void batch_propagate(LtBatch &B, int nsteps, hipStream_t caller);
void lt_get_state(LtSystem &L, double *xyz, double *vel, double *dprev, double *grad,
                  double *thermo, int64_t *step);
const float *lt_xyzq(LtSystem &L);
void lt_list_info(LtSystem &L, double *out4);
void lt_launch_info(LtSystem &L, double *out6);
void batch_autotune(LtBatch &B, hipStream_t caller, double *report9);
}  // namespace apo

struct apo_lt_integrator {
  apo::LtSystem *lt;
};

struct apo_lt_batch {
  apo::LtBatch *b;
};

namespace apo {
thread_local std::string g_last_error;
void capi_set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace apo

namespace {
using apo::guard;

apo::NbSystem &get(apo_nb_system *h) {
  if (!h || !h->sys) throw apo::InvalidArgument("null system handle");
  return *h->sys;
}
const apo::NbSystem &get(const apo_nb_system *h) {
  if (!h || !h->sys) throw apo::InvalidArgument("null system handle");
  return *h->sys;
}
apo::LtSystem &get(apo_lt_integrator *h) {
  if (!h || !h->lt) throw apo::InvalidArgument("null integrator handle");
  return *h->lt;
}
}  // namespace

extern "C" {

APO_API const char *apo_last_error(void) { return apo::g_last_error.c_str(); }

APO_API const char *apo_version(void) { return "apocharmm_amd 0.1 (gfx950)"; }

APO_API apo_status apo_nb_create(const apo_nb_params *params, const apo_nb_topology *topology,
                                 apo_nb_system **out) {
  return guard([&] {
    if (!params || !topology || !out) throw apo::InvalidArgument("null argument");
    *out = nullptr;
    apo::NbSystem *S = apo::create_system(*params, *topology);
    *out = new apo_nb_system{S};
  });
}

APO_API apo_status apo_nb_destroy(apo_nb_system *sys) {
  return guard([&] {
    if (!sys) return;
    if (sys->borrowed)
      throw apo::InvalidArgument("this force system belongs to a ForceManager: destroy the "
                                 "manager instead");
    if (sys->sys && sys->sys->n_integrators > 0)
      throw apo::InvalidArgument("an integrator is still bound to this force system: destroy "
                                 "the integrator first");
    delete sys->sys;
    delete sys;
  });
}

APO_API apo_status apo_nb_set_box(apo_nb_system *sys, const double *box3) {
  return guard([&] {
    if (!box3) throw apo::InvalidArgument("null box");
    apo::set_box(get(sys), box3);
  });
}

APO_API apo_status apo_nb_reset_list(apo_nb_system *sys, const float *d_xyzq, void *stream) {
  return guard([&] {
    if (!d_xyzq) throw apo::InvalidArgument("null xyzq");
    apo::nblist_build(get(sys), reinterpret_cast<const float4 *>(d_xyzq),
                      reinterpret_cast<hipStream_t>(stream), true);
  });
}

APO_API apo_status apo_nb_calc_force(apo_nb_system *sys, const float *d_xyzq, int32_t flags,
                                     int64_t *d_force, int32_t stride, double *d_energy,
                                     double *d_virial, void *stream) {
  return guard([&] {
    constexpr int32_t known =
        APO_CALC_ENERGY | APO_CALC_VIRIAL | APO_RESET_LIST | APO_ACCUMULATE | APO_TERM_ALL;
    if (flags & ~known) throw apo::InvalidArgument("unknown calc_force flag bits");
    apo::calc_force(get(sys), d_xyzq, flags, d_force, stride, d_energy, d_virial,
                    reinterpret_cast<hipStream_t>(stream));
  });
}

APO_API apo_status apo_nb_calc_force_noalloc(apo_nb_system *sys, const float *d_xyzq,
                                             int32_t flags, int64_t *d_force, int32_t stride,
                                             double *d_energy, double *d_virial, void *stream) {
  return guard([&] {
    constexpr int32_t known =
        APO_CALC_ENERGY | APO_CALC_VIRIAL | APO_RESET_LIST | APO_ACCUMULATE | APO_TERM_ALL;
    if (flags & ~known) throw apo::InvalidArgument("unknown calc_force flag bits");
    apo::calc_force_noalloc(get(sys), d_xyzq, flags, d_force, stride, d_energy, d_virial,
                            reinterpret_cast<hipStream_t>(stream));
  });
}

APO_API apo_status apo_nb_add_force(const int64_t *d_src, int32_t natom, int32_t stride,
                                    int64_t *d_dst, void *stream) {
  return guard([&] {
    if (!d_src || !d_dst || natom < 0 || stride < natom)
      throw apo::InvalidArgument("bad force addition arguments");
    if (natom)
      apo::launch_add_force(d_src, natom, stride, d_dst, reinterpret_cast<hipStream_t>(stream));
  });
}

APO_API apo_status apo_nb_add_energy(const double *d_plugin_energy, double *d_energy,
                                     void *stream) {
  return guard([&] {
    if (!d_plugin_energy || !d_energy) throw apo::InvalidArgument("null energy pointer");
    apo::launch_add_energy(d_plugin_energy, d_energy, reinterpret_cast<hipStream_t>(stream));
  });
}

APO_API apo_status apo_nb_force_stride(const apo_nb_system *sys, int32_t *stride) {
  return guard([&] {
    if (!stride) throw apo::InvalidArgument("null stride");
    const int n = get(sys).natom;
    *stride = ((n + 31) * 8 + 255) / 256 * 256 / 8;
  });
}

APO_API apo_status apo_nb_num_atoms(const apo_nb_system *sys, int32_t *natom) {
  return guard([&] {
    if (!natom) throw apo::InvalidArgument("null natom");
    *natom = get(sys).natom;
  });
}

APO_API apo_status apo_nb_force_to_double(const int64_t *d_force, int32_t natom, int32_t stride,
                                          double *d_out, void *stream) {
  return guard([&] {
    if (!d_force || !d_out || natom < 0 || stride < natom)
      throw apo::InvalidArgument("bad force conversion arguments");
    if (natom)
      apo::launch_force_to_double(d_force, natom, stride, d_out,
                                  reinterpret_cast<hipStream_t>(stream));
  });
}

APO_API apo_status apo_nb_list_stats(apo_nb_system *sys, int64_t *out6, void *stream) {
  return guard([&] {
    if (!out6) throw apo::InvalidArgument("null output");
    apo::nblist_stats(get(sys), out6, reinterpret_cast<hipStream_t>(stream));
  });
}

APO_API apo_status apo_nb_profile(apo_nb_system *sys, int32_t enable) {
  return guard([&] {
    apo::NbSystem &S = get(sys);
    S.profile_flush();
    S.prof_on = enable != 0;
    for (int k = 0; k < APO_PROFILE_SLOTS; k++) {
      S.prof_ms[k] = 0;
      S.prof_cnt[k] = 0;
    }
  });
}

APO_API apo_status apo_nb_profile_read(apo_nb_system *sys, double *ms, int64_t *count) {
  return guard([&] {
    if (!ms || !count) throw apo::InvalidArgument("null output");
    apo::NbSystem &S = get(sys);
    S.profile_flush();
    for (int k = 0; k < APO_PROFILE_SLOTS; k++) {
      ms[k] = S.prof_ms[k];
      count[k] = S.prof_cnt[k];
    }
  });
}

APO_API apo_status apo_topology_exclusions(int32_t natom, int32_t nbond, const int32_t *bonds,
                                           int32_t *counts3, int32_t *excl_ptr,
                                           int32_t *excl_idx, int32_t *in14, int32_t *ex14) {
  return guard([&] {
    if (natom < 0 || nbond < 0 || !counts3 || (nbond > 0 && !bonds))
      throw apo::InvalidArgument("bad topology arguments");
    for (int k = 0; k < 2 * nbond; k++)
      if (bonds[k] < 0 || bonds[k] >= natom) throw apo::InvalidArgument("bond index out of range");
    std::vector<int32_t> p, x, i14, e14;
    apo::build_exclusions(natom, nbond, bonds, p, x, i14, e14);
    counts3[0] = (int32_t)x.size();
    counts3[1] = (int32_t)i14.size() / 2;
    counts3[2] = (int32_t)e14.size() / 2;
    if (excl_ptr) std::memcpy(excl_ptr, p.data(), p.size() * sizeof(int32_t));
    if (excl_idx) std::memcpy(excl_idx, x.data(), x.size() * sizeof(int32_t));
    if (in14) std::memcpy(in14, i14.data(), i14.size() * sizeof(int32_t));
    if (ex14) std::memcpy(ex14, e14.data(), e14.size() * sizeof(int32_t));
  });
}

APO_API apo_status apo_lt_create(apo_nb_system *sys, const apo_lt_params *params,
                                 const double *xyz, const double *vel, apo_lt_integrator **out) {
  return guard([&] {
    if (!params || !out) throw apo::InvalidArgument("null argument");
    *out = nullptr;
    apo::LtSystem *L = apo::create_lt(&get(sys), *params, xyz, vel);
    *out = new apo_lt_integrator{L};
  });
}

APO_API apo_status apo_lt_destroy(apo_lt_integrator *lt) {
  return guard([&] {
    if (!lt) return;
    apo::destroy_lt(lt->lt);
    delete lt;
  });
}

APO_API apo_status apo_lt_initialize(apo_lt_integrator *lt, void *stream) {
  return guard([&] { apo::lt_initialize(get(lt), static_cast<hipStream_t>(stream)); });
}

APO_API apo_status apo_lt_propagate(apo_lt_integrator *lt, int32_t nsteps, void *stream) {
  return guard([&] { apo::lt_propagate(get(lt), nsteps, static_cast<hipStream_t>(stream)); });
}

APO_API apo_status apo_lt_propagate_batch(apo_lt_integrator *const *lts, int32_t n,
                                          int32_t nsteps, void *stream) {
  return guard([&] {
    if (n > 0 && !lts) throw apo::InvalidArgument("null integrator array");
    std::vector<apo::LtSystem *> v((size_t)(n > 0 ? n : 0));
    for (int i = 0; i < n; i++) v[i] = &get(lts[i]);
    apo::lt_propagate_batch(v.data(), n, nsteps, static_cast<hipStream_t>(stream));
  });
}

APO_API apo_status apo_lt_batch_create(apo_lt_integrator *const *lts, int32_t n, int32_t lanes,
                                       apo_lt_batch **out) {
  return guard([&] {
    if (!out || (n > 0 && !lts)) throw apo::InvalidArgument("null argument");
    *out = nullptr;
    std::vector<apo::LtSystem *> v((size_t)(n > 0 ? n : 0));
    for (int i = 0; i < n; i++) v[i] = &get(lts[i]);
    apo::LtBatch *B = apo::create_batch(v.data(), n, lanes);
    *out = new apo_lt_batch{B};
  });
}

APO_API apo_status apo_lt_batch_propagate(apo_lt_batch *b, int32_t nsteps, void *stream) {
  return guard([&] {
    if (!b || !b->b) throw apo::InvalidArgument("null batch handle");
    apo::batch_propagate(*b->b, nsteps, static_cast<hipStream_t>(stream));
  });
}

APO_API apo_status apo_lt_batch_autotune(apo_lt_batch *b, void *stream, double *report9) {
  return guard([&] {
    if (!b || !b->b) throw apo::InvalidArgument("null batch handle");
    apo::batch_autotune(*b->b, static_cast<hipStream_t>(stream), report9);
  });
}

APO_API apo_status apo_lt_batch_destroy(apo_lt_batch *b) {
  return guard([&] {
    if (!b) return;
    apo::destroy_batch(b->b);
    delete b;
  });
}

APO_API apo_status apo_lt_get_state(apo_lt_integrator *lt, double *xyz, double *vel,
                                    double *dprev, double *grad, double *thermo6,
                                    int64_t *step) {
  return guard([&] { apo::lt_get_state(get(lt), xyz, vel, dprev, grad, thermo6, step); });
}

APO_API apo_status apo_lt_list_info(apo_lt_integrator *lt, double *out4) {
  return guard([&] {
    if (!out4) throw apo::InvalidArgument("null output");
    apo::lt_list_info(get(lt), out4);
  });
}

APO_API apo_status apo_lt_launch_info(apo_lt_integrator *lt, double *out6) {
  return guard([&] {
    if (!out6) throw apo::InvalidArgument("null output");
    apo::lt_launch_info(get(lt), out6);
  });
}

APO_API apo_status apo_lt_xyzq(apo_lt_integrator *lt, const float **d_xyzq) {
  return guard([&] {
    if (!d_xyzq) throw apo::InvalidArgument("null output");
    *d_xyzq = apo::lt_xyzq(get(lt));
  });
}

}  // extern "C"
