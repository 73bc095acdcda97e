// calc_force without device allocation (apo_nb_calc_force_noalloc).
//
// The energy / virial partial sums of one force call (per-block FP64 partials
// of the pair, 1-4, bonded, reciprocal and self terms, reduced by
// k_reduce_energy) go to `scratch`, a file-scope device array of
// kScratchDoubles doubles instead of the system's own buffer, and the call
// never grows the neighbor-list buffers.  Every slot the reduction reads is
// written by the same call, so nothing carries over between calls; calls that
// overlap on one GPU share the array and race.
#include <mutex>

#include "system.hpp"

namespace apo {

void calc_force(NbSystem &S, const float *xyzq, int flags, int64_t *force, int stride,
                double *energy, double *virial, hipStream_t s);

constexpr size_t kScratchDoubles = size_t(1) << 17;  // 1 MiB per GPU

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) on its first use.
static double *scratch_ptr() {
  int dev = 0;
  APO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw RuntimeError("device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    APO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw RuntimeError("hipGetSymbolAddress returned no address for the scratch array");
    g_ptr[dev] = static_cast<double *>(p);
  }
  return g_ptr[dev];
}

size_t noalloc_scratch_doubles() { return kScratchDoubles; }

void calc_force_noalloc(NbSystem &S, const float *xyzq, int flags, int64_t *force, int stride,
                        double *energy, double *virial, hipStream_t s) {
  if ((size_t)S.part.total > kScratchDoubles)
    throw InvalidArgument("this system's energy/virial partials need " +
                          std::to_string(S.part.total) + " doubles; the no-allocation entry "
                          "holds at most " + std::to_string(kScratchDoubles));
  if (S.prof_on) throw InvalidArgument("kernel timers allocate events: turn profiling off");
  if (S.cand_cap == 0 || S.ent_cap == 0 || S.item_cap == 0)
    throw InvalidArgument("the neighbor list has no capacity yet: build it once with an "
                          "allocating call (apo_nb_reset_list or apo_nb_calc_force)");
  double *p = scratch_ptr();
  struct Restore {
    NbSystem &S;
    ~Restore() {
      S.parts_ext = nullptr;
      S.may_allocate = true;
    }
  } restore{S};
  S.parts_ext = p;
  S.may_allocate = false;
  calc_force(S, xyzq, flags, force, stride, energy, virial, s);
}

}  // namespace apo
