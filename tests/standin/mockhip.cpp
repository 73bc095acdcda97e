// A host-only stand-in for the HIP runtime calls the library makes, for CPU
// tests of its host logic under AddressSanitizer (no GPU, no kernels).
//
// Model: every operation runs at once on the calling thread.  Device memory
// is host memory (calloc), so host code that touches freed device or pinned
// memory is caught by ASAN.  Kernels are no-ops, except the list build's
// count export (k_list_guard_export), which is emulated so that the host sees
// builds complete; with APO_STANDIN_GROW=1 the exported counts grow at every
// export, which drives the capacity-growth path (list_export_check ->
// list_reserve -> new generation -> graph recapture).  Stream capture records
// memcpy/memset operations and emulated kernels; a launched graph replays
// them.  Handles (streams, events, graphs) are checked on every use: a call
// with a destroyed handle aborts with a message.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

namespace {

struct Op {
  enum Kind { kMemcpy, kMemset, kMetaExport } kind;
  void *dst = nullptr;
  const void *src = nullptr;
  size_t bytes = 0;
  int value = 0;
};
struct Graph {
  std::vector<Op> ops;
};
struct Stream {
  Graph *capture = nullptr;
};
struct Event {
  Graph *capture = nullptr;
};

std::mutex mu;
std::set<Stream *> streams;
std::set<Event *> events;
std::set<Graph *> graphs;        // captured graphs
std::set<Graph *> execs;         // instantiated graphs
std::map<const void *, std::string> kernel_names;
std::map<const void *, size_t> var_sizes;
std::map<void *, size_t> device_allocs, host_allocs;
std::map<const int *, int> grow_count;  // per list-count buffer: exports so far
size_t n_malloc = 0, n_instantiate = 0, n_launch = 0, n_export = 0;
dim3 cfg_grid, cfg_block;
size_t cfg_shmem = 0;
hipStream_t cfg_stream = nullptr;

[[noreturn]] void fail(const char *what, const void *h) {
  std::fprintf(stderr, "[standin] %s: %p\n", what, h);
  std::fflush(stderr);
  std::abort();
}
Stream *as_stream(hipStream_t s) {
  if (!s) return nullptr;  // the null stream
  auto *p = reinterpret_cast<Stream *>(s);
  if (!streams.count(p)) fail("use of a stream that does not exist (destroyed?)", s);
  return p;
}
Event *as_event(hipEvent_t e) {
  auto *p = reinterpret_cast<Event *>(e);
  if (!events.count(p)) fail("use of an event that does not exist (destroyed?)", e);
  return p;
}
bool grow_on() {
  const char *g = std::getenv("APO_STANDIN_GROW");
  return g && g[0] == '1';
}

// k_list_guard_export(int *meta, int cand_cap, int ent_cap, int item_cap,
// int *host): the build's counts and its number to
// the pinned host copy (host[15] last); with growth on, the counts ncand,
// nent, nitem (host[1..3]) report more than the last build used
void meta_export(int *meta, int *host) {
  n_export++;
  const int seq = meta[16] + 1;
  meta[16] = seq;
  for (int k = 0; k < 15; k++) host[k] = meta[k];
  if (grow_on()) {
    int &g = grow_count[meta];
    if (g < 6) {
      const int scale = 64 << g;
      host[1] = 40 * scale;
      host[2] = 900 * scale;
      host[3] = 12 * scale;
    }
    g++;
  }
  host[15] = seq;
}

void run(const Op &op) {
  switch (op.kind) {
    case Op::kMemcpy: std::memmove(op.dst, op.src, op.bytes); break;
    case Op::kMemset: std::memset(op.dst, op.value, op.bytes); break;
    case Op::kMetaExport:
      meta_export(static_cast<int *>(op.dst), const_cast<int *>(static_cast<const int *>(op.src)));
      break;
  }
}
void submit(hipStream_t s, const Op &op) {
  Stream *st = as_stream(s);
  if (st && st->capture) st->capture->ops.push_back(op);
  else run(op);
}

}  // namespace

extern "C" {

// ---- registration (emitted by the compiler for every translation unit) ----
void **__hipRegisterFatBinary(const void *) {
  static void *handle = nullptr;
  return &handle;
}
void __hipRegisterFunction(void **, const void *hostFunction, char *, const char *deviceName,
                           unsigned int, void *, void *, void *, void *, int *) {
  std::lock_guard<std::mutex> lk(mu);
  kernel_names[hostFunction] = deviceName ? deviceName : "";
}
void __hipRegisterVar(void **, void *var, char *, char *, int, size_t size, int, int) {
  std::lock_guard<std::mutex> lk(mu);
  var_sizes[var] = size;
}
void __hipUnregisterFatBinary(void **) {}

hipError_t __hipPushCallConfiguration(dim3 grid, dim3 block, size_t shmem, hipStream_t s) {
  cfg_grid = grid;
  cfg_block = block;
  cfg_shmem = shmem;
  cfg_stream = s;
  return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3 *grid, dim3 *block, size_t *shmem, hipStream_t *s) {
  *grid = cfg_grid;
  *block = cfg_block;
  *shmem = cfg_shmem;
  *s = cfg_stream;
  return hipSuccess;
}
hipError_t hipLaunchKernel(const void *f, dim3, dim3, void **args, size_t, hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu);
  auto it = kernel_names.find(f);
  if (it == kernel_names.end()) fail("launch of an unregistered kernel", f);
  if (it->second.find("k_list_guard_export") != std::string::npos) {
    Op op{Op::kMetaExport};
    op.dst = *static_cast<int **>(args[0]);
    op.src = *static_cast<int **>(args[4]);
    submit(s, op);
  } else {
    (void)as_stream(s);
  }
  return hipSuccess;
}

// ---- errors, device ---------------------------------------------------------
hipError_t hipGetLastError(void) { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "stand-in error"; }
hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600 *prop, int) {
  std::memset(prop, 0, sizeof(*prop));
  std::strcpy(prop->name, "host stand-in");
  prop->multiProcessorCount = 8;
  prop->totalGlobalMem = size_t(64) << 30;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }

// ---- memory -----------------------------------------------------------------
hipError_t hipMalloc(void **p, size_t bytes) {
  std::lock_guard<std::mutex> lk(mu);
  *p = std::calloc(1, bytes ? bytes : 1);
  device_allocs[*p] = bytes;
  n_malloc++;
  return hipSuccess;
}
hipError_t hipFree(void *p) {
  std::lock_guard<std::mutex> lk(mu);
  if (!p) return hipSuccess;
  if (!device_allocs.erase(p)) fail("hipFree of a pointer hipMalloc did not return", p);
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned int) {
  std::lock_guard<std::mutex> lk(mu);
  *p = std::calloc(1, bytes ? bytes : 1);
  host_allocs[*p] = bytes;
  return hipSuccess;
}
hipError_t hipHostFree(void *p) {
  std::lock_guard<std::mutex> lk(mu);
  if (!p) return hipSuccess;
  if (!host_allocs.erase(p)) fail("hipHostFree of a pointer hipHostMalloc did not return", p);
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind) {
  std::memmove(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind,
                          hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu);
  Op op{Op::kMemcpy};
  op.dst = dst;
  op.src = src;
  op.bytes = bytes;
  submit(s, op);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void *dst, int value, size_t bytes, hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu);
  Op op{Op::kMemset};
  op.dst = dst;
  op.value = value;
  op.bytes = bytes;
  submit(s, op);
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
  std::lock_guard<std::mutex> lk(mu);
  if (!var_sizes.count(symbol)) fail("hipGetSymbolAddress of an unregistered variable", symbol);
  *p = const_cast<void *>(symbol);  // the host shadow stands for the device copy
  return hipSuccess;
}

// ---- streams and events -------------------------------------------------------
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int) {
  std::lock_guard<std::mutex> lk(mu);
  auto *p = new Stream();
  streams.insert(p);
  *s = reinterpret_cast<hipStream_t>(p);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu);
  Stream *p = as_stream(s);
  if (p && p->capture) fail("destroying a stream that is capturing", s);
  streams.erase(p);
  delete p;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu);
  Stream *p = as_stream(s);
  if (p && p->capture) fail("synchronizing a capturing stream", s);
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) {
  std::lock_guard<std::mutex> lk(mu);
  auto *p = new Event();
  events.insert(p);
  *e = reinterpret_cast<hipEvent_t>(p);
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned int) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  std::lock_guard<std::mutex> lk(mu);
  Event *p = as_event(e);
  events.erase(p);
  delete p;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu);
  Event *ev = as_event(e);
  Stream *st = as_stream(s);
  ev->capture = st ? st->capture : nullptr;
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  std::lock_guard<std::mutex> lk(mu);
  if (as_event(e)->capture) fail("synchronizing an event recorded during capture", e);
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  std::lock_guard<std::mutex> lk(mu);
  (void)as_event(a);
  (void)as_event(b);
  *ms = 0.01f;
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int) {
  std::lock_guard<std::mutex> lk(mu);
  Stream *st = as_stream(s);
  Event *ev = as_event(e);
  if (ev->capture && st) st->capture = ev->capture;  // the stream joins the capture
  return hipSuccess;
}

// ---- graphs -------------------------------------------------------------------
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  std::lock_guard<std::mutex> lk(mu);
  Stream *st = as_stream(s);
  if (!st) fail("capture on the null stream", s);
  if (st->capture) fail("nested capture", s);
  auto *g = new Graph();
  graphs.insert(g);
  st->capture = g;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *out) {
  std::lock_guard<std::mutex> lk(mu);
  Stream *st = as_stream(s);
  if (!st || !st->capture) fail("end of a capture that was not begun", s);
  Graph *g = st->capture;
  for (Stream *o : streams)
    if (o->capture == g) o->capture = nullptr;
  *out = reinterpret_cast<hipGraph_t>(g);
  return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *status) {
  std::lock_guard<std::mutex> lk(mu);
  Stream *st = as_stream(s);
  *status = st && st->capture ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *exec, hipGraph_t graph, hipGraphNode_t *, char *,
                               size_t) {
  std::lock_guard<std::mutex> lk(mu);
  auto *g = reinterpret_cast<Graph *>(graph);
  if (!graphs.count(g)) fail("instantiating a graph that does not exist", graph);
  auto *x = new Graph(*g);
  execs.insert(x);
  n_instantiate++;
  *exec = reinterpret_cast<hipGraphExec_t>(x);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t graph) {
  std::lock_guard<std::mutex> lk(mu);
  auto *g = reinterpret_cast<Graph *>(graph);
  if (!graphs.erase(g)) fail("destroying a graph that does not exist", graph);
  delete g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
  std::lock_guard<std::mutex> lk(mu);
  auto *x = reinterpret_cast<Graph *>(exec);
  if (!execs.erase(x)) fail("destroying a graph exec that does not exist", exec);
  delete x;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu);
  auto *x = reinterpret_cast<Graph *>(exec);
  if (!execs.count(x)) fail("launching a graph exec that does not exist", exec);
  Stream *st = as_stream(s);
  if (st && st->capture) fail("graph launch inside a capture", s);
  for (const Op &op : x->ops) run(op);
  n_launch++;
  return hipSuccess;
}

// ---- what the driver reads back ------------------------------------------------
size_t apo_standin_live_device_allocs(void) {
  std::lock_guard<std::mutex> lk(mu);
  return device_allocs.size();
}
void apo_standin_counts(size_t *out4) {
  std::lock_guard<std::mutex> lk(mu);
  out4[0] = n_malloc;
  out4[1] = n_instantiate;
  out4[2] = n_launch;
  out4[3] = n_export;
}
size_t apo_standin_live_handles(void) {
  std::lock_guard<std::mutex> lk(mu);
  return streams.size() + events.size() + graphs.size() + execs.size() + host_allocs.size();
}

}  // extern "C"
