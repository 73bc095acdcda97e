// Error mapping of every C-ABI entry: the body runs inside guard(), which
// turns exceptions into apo_status codes and keeps the message in a
// thread-local string read by apo_last_error() (the reference's C ABI does
// the same: capi/private_include/apocharmm_c/detail/ErrorInternal.h:27-40,
// capi/src/apocharmm_c/Error.cpp:18-22).
#pragma once

#include <stdexcept>
#include <string>

#include "apocharmm_amd.h"
#include "common.hpp"

namespace apo {
struct NbSystem;
}  // namespace apo

// The force-system handle of apocharmm_amd.h: a borrowed or owned NbSystem.
struct apo_nb_system {
  apo::NbSystem *sys;
  bool borrowed = false;  // owned by a force manager (apo_force_manager_get_nb_system)
};

namespace apo {

void capi_set_error(const std::string &msg);  // empty: no error

struct NotImplemented : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NotInitialized : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <typename F> apo_status guard(F &&f) {
  try {
    f();
    capi_set_error(std::string());
    return APO_OK;
  } catch (const InvalidArgument &e) {
    capi_set_error(e.what());
    return APO_INVALID_ARGUMENT;
  } catch (const GpuError &e) {
    capi_set_error(e.what());
    return APO_GPU_ERROR;
  } catch (const NotInitialized &e) {
    capi_set_error(e.what());
    return APO_NOT_INITIALIZED;
  } catch (const NotImplemented &e) {
    capi_set_error(e.what());
    return APO_NOT_IMPLEMENTED;
  } catch (const std::invalid_argument &e) {
    capi_set_error(e.what());
    return APO_INVALID_ARGUMENT;
  } catch (const std::exception &e) {
    capi_set_error(e.what());
    return APO_RUNTIME_ERROR;
  } catch (...) {
    capi_set_error("unknown error");
    return APO_RUNTIME_ERROR;
  }
}

}  // namespace apo
