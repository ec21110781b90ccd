// extern "C" entry points of libravesim_hip.so (declared in include/ravesim.h).
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "rsim_fft_cfg.hpp"
#include "rsim_internal.hpp"
#include "rsim_tables.hpp"

namespace rsim {

int point_source(void* u, int nx, int ny, int dtype, double dx, double dy, double z, double xs,
                 double ys, double wl, hipStream_t s);
int transmission(void* u, int nx, int ny, int dtype, double dx, double dy, double wl,
                 const rsim_volume* vol, int slice, double px, double py, double pz, double xpos,
                 double ypos, double coord_scale, hipStream_t s);
int tilted_plane(void* u, int nx, int ny, int dtype, double dx, double dy, double wl, double dz,
                 double xs, double ys, hipStream_t s);
int detector_area_v1(const void* u, int nx, int ny, int dtype, double dx, double dy, int ox,
                     int oy, double dsx, double dsy, double z, double weight, bool accumulate,
                     double* out_dev, hipStream_t s);
int paint_spheres(rsim_volume* vol, const double* sph, const int* mats, int n, hipStream_t s);
int propagate(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
              double dz, double freq, bool skip_fft, hipStream_t s);
int detector_area_v1_adjoint(const void* v, void* g, int nx, int ny, int dtype, double dx,
                             double dy, int ox, int oy, double dsx, double dsy, double z,
                             const double* gimg_dev, hipStream_t s);
int transmission_db_2d(void* u, int nx, int ny, int dtype, double wl, double pz,
                       const float* delta, const float* beta, hipStream_t s);
int multislice_db(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                  double freq, const float* delta, const float* beta, double pz, int n,
                  void* ckpt, int every, int engine, int flags, hipStream_t s);
int multislice_db_adjoint(void* g, void* U, int nx, int ny, int dtype, double dx, double dy,
                          double wl, double freq, const float* delta, const float* beta,
                          double pz, int n, const void* ckpt, int every, void* work,
                          float* gdelta, float* gbeta, int engine, int flags, hipStream_t s);
int batch_create(rsim_batch** out, const rsim_batch_instance* inst, int count, int first,
                 int n, int flags, hipStream_t s);
int batch_run(rsim_batch* b, hipStream_t s);
int batch_info(const rsim_batch* b, int* groups, int* kernels);
int batch_destroy(rsim_batch* b);
int rsim_release_tables();
int release_tables();
int release_multislice_tables();
int release_tune();
int release_detector_plans();
int release_staging();
int release_volume_tables(unsigned long long serial);
int fused_schedule_lookup(int nx, int ny, int* rv, int* cv, int* tile);
int fill_zero(void* p, size_t bytes, hipStream_t s);
int stream_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
int detector_area_v1_noalloc(const void* u, int nx, int ny, int dtype, double dx, double dy,
                             int ox, int oy, double dsx, double dsy, double z, double weight,
                             double* out_dev, hipStream_t s);
size_t detector_noalloc_capacity();

namespace {
thread_local std::string t_err;
std::mutex g_mu;
std::map<int, hipStream_t> g_streams;
std::map<int, hipStream_t> g_side_streams;
std::atomic<int> g_cancel{0};
rsim_progress_fn g_progress = nullptr;
void* g_progress_user = nullptr;
}  // namespace

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_err = buf;
  return code;
}

void clear_error() { t_err.clear(); }
const char* last_error_cstr() { return t_err.c_str(); }
void set_cancel(int v) { g_cancel.store(v); }
void set_progress(rsim_progress_fn fn, void* user) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_progress = fn;
  g_progress_user = user;
}
void release_streams() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& kv : g_streams) hipStreamDestroy(kv.second);
  g_streams.clear();
  for (auto& kv : g_side_streams) hipStreamDestroy(kv.second);
  g_side_streams.clear();
}

hipStream_t side_stream() {
  std::lock_guard<std::mutex> lk(g_mu);
  const int dev = current_device();
  auto it = g_side_streams.find(dev);
  if (it != g_side_streams.end()) return it->second;
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return nullptr;
  g_side_streams[dev] = st;
  return st;
}

int current_device() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return 0;
  return d;
}

hipStream_t capture_stream();

hipStream_t resolve_stream(void* s) {
  if (s) return (hipStream_t)s;
  // inside a capture, calls without a stream go to the capturing stream
  if (hipStream_t cs = capture_stream()) return cs;
  std::lock_guard<std::mutex> lk(g_mu);
  const int dev = current_device();
  auto it = g_streams.find(dev);
  if (it != g_streams.end()) return it->second;
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return nullptr;
  g_streams[dev] = st;
  return st;
}

size_t max_lds_per_block() {
  static std::map<int, size_t> cache;
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  const int dev = current_device();
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  int v = 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    v = 65536;
  cache[dev] = (size_t)v;
  return (size_t)v;
}


// ---- kernel timing ----
namespace {
struct KEv {
  int kind;
  hipEvent_t a, b;
};
std::mutex g_kt_mu;
bool g_kt_on = false;
int g_kt_every = 1;
long long g_kt_seen[3] = {0, 0, 0};
std::vector<KEv> g_kt_done;
std::vector<hipEvent_t> g_kt_pool;
KEv g_kt_open[3];
bool g_kt_has_open[3] = {false, false, false};

hipEvent_t kt_event() {
  if (!g_kt_pool.empty()) {
    hipEvent_t e = g_kt_pool.back();
    g_kt_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}
}  // namespace

bool ktime_on() { return g_kt_on; }

namespace {
thread_local bool t_kt_suspended = false;
}
void ktime_suspend(bool on) { t_kt_suspended = on; }

void ktime_begin(int kind, hipStream_t s) {
  if (!g_kt_on || t_kt_suspended) return;
  std::lock_guard<std::mutex> lk(g_kt_mu);
  // sample one launch in g_kt_every of each kind: an event pair costs a few
  // microseconds of dispatch gap, which would skew the run it measures
  if (g_kt_seen[kind]++ % g_kt_every != 0) return;
  KEv ev{kind, kt_event(), kt_event()};
  hipEventRecord(ev.a, s);
  g_kt_open[kind] = ev;
  g_kt_has_open[kind] = true;
}

void ktime_end(int kind, hipStream_t s) {
  if (!g_kt_on || t_kt_suspended) return;
  std::lock_guard<std::mutex> lk(g_kt_mu);
  if (!g_kt_has_open[kind]) return;
  hipEventRecord(g_kt_open[kind].b, s);
  g_kt_done.push_back(g_kt_open[kind]);
  g_kt_has_open[kind] = false;
}

int ktime_enable(int every) {
  if (every < 0) return fail(RSIM_E_INVALID_ARG, "rsim_kernel_timing: every must be >= 0");
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kt_on = every != 0;
  g_kt_every = every > 0 ? every : 1;
  for (int k = 0; k < 3; ++k) g_kt_seen[k] = 0;
  return RSIM_OK;
}

int ktime_read(double* ms3, int* n3) {
  std::lock_guard<std::mutex> lk(g_kt_mu);
  double ms[3] = {0, 0, 0};
  int n[3] = {0, 0, 0};
  for (auto& e : g_kt_done) {
    RSIM_HIP(hipEventSynchronize(e.b));
    float t = 0.f;
    RSIM_HIP(hipEventElapsedTime(&t, e.a, e.b));
    ms[e.kind] += t;
    n[e.kind] += 1;
    g_kt_pool.push_back(e.a);
    g_kt_pool.push_back(e.b);
  }
  g_kt_done.clear();
  for (int k = 0; k < 3; ++k) {
    if (ms3) ms3[k] = ms[k];
    if (n3) n3[k] = n[k];
  }
  return RSIM_OK;
}

// ---- graph capture ----
namespace {
struct CaptureState {
  hipStream_t s = nullptr;
  int device = 0;
  std::vector<const void*> used;   // cached tables the captured work reads
  std::vector<void*> owned;        // scratch the graph owns
};
thread_local CaptureState* t_cap = nullptr;
}  // namespace

bool capture_active() { return t_cap != nullptr; }
hipStream_t capture_stream() { return t_cap ? t_cap->s : nullptr; }

int capture_alloc(size_t bytes, void** p) {
  if (!t_cap) return fail(RSIM_E_INVALID_ARG, "capture_alloc outside a capture");
  *p = nullptr;
  RSIM_HIP(hipMalloc(p, bytes ? bytes : 1));
  t_cap->owned.push_back(*p);
  return RSIM_OK;
}

int capture_miss(const char* what) {
  if (!t_cap) return RSIM_OK;
  return fail(RSIM_E_UNSUPPORTED,
              "graph capture: %s are not built yet; run the same calls once before "
              "rsim_graph_begin so that everything the graph reads exists", what);
}

}  // namespace rsim

// A captured sequence of library calls (rsim_graph_begin .. rsim_graph_end).
struct rsim_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t nodes = 0;
  unsigned long long generation = 0;
  std::vector<const void*> pinned;
  std::vector<void*> owned;
};

// A run control: the cancel flag, progress callback and slice count of the
// runs started on the thread it is bound to.
struct rsim_control {
  std::atomic<int> cancel{0};
  rsim_progress_fn fn = nullptr;
  void* user = nullptr;
  int every = 1;
  std::atomic<long long> completed{0};
  std::atomic<long long> total{0};
};

namespace rsim {
namespace {
thread_local rsim_control* t_ctl = nullptr;
}

bool cancel_requested() {
  if (t_ctl) return t_ctl->cancel.load() != 0;
  return g_cancel.load() != 0;
}

bool progress_due(long long completed, long long total) {
  if (t_ctl) {
    if (!t_ctl->fn) return false;
    return completed == total || completed % t_ctl->every == 0;
  }
  return g_progress != nullptr;
}

void report_progress(const char* stage, long long completed, long long total) {
  if (t_ctl) {
    t_ctl->completed.store(completed);
    t_ctl->total.store(total);
    if (t_ctl->fn && (completed == total || completed % t_ctl->every == 0))
      t_ctl->fn(stage, completed, total, t_ctl->user);
    return;
  }
  if (g_progress) g_progress(stage, completed, total, g_progress_user);
}

void note_done(long long completed, long long total) {
  if (t_ctl) {
    t_ctl->completed.store(completed);
    t_ctl->total.store(total);
  }
}

}  // namespace rsim

using namespace rsim;

namespace {

int check_field(const void* p, int nx, int ny, int dtype, const char* name) {
  if (!p) return fail(RSIM_E_INVALID_ARG, "%s is NULL", name);
  if (nx <= 0 || ny <= 0) return fail(RSIM_E_INVALID_ARG, "nx and ny must be positive, got nx=%d, ny=%d", nx, ny);
  if (!valid_dtype(dtype)) return fail(RSIM_E_INVALID_ARG, "dtype must be RSIM_C64 or RSIM_C128, got %d", dtype);
  return RSIM_OK;
}

int check_fft_sizes(int nx, int ny) {
  if (!fft_size_supported(nx) || !fft_size_supported(ny))
    return fail(RSIM_E_INVALID_ARG, "FFT sizes must factor into 2, 3, 5: nx=%d, ny=%d", nx, ny);
  return RSIM_OK;
}

}  // namespace


extern "C" {

const char* rsim_last_error(void) { return last_error_cstr(); }

const char* rsim_version(void) { return "ravesim-mi355x 0.1 (gfx950)"; }

int rsim_device_count(int* count) {
  if (!count) return fail(RSIM_E_INVALID_ARG, "count is NULL");
  RSIM_HIP(hipGetDeviceCount(count));
  return RSIM_OK;
}

int rsim_set_device(int device) {
  RSIM_HIP(hipSetDevice(device));
  return RSIM_OK;
}

int rsim_malloc(void** dptr, size_t bytes) {
  if (!dptr) return fail(RSIM_E_INVALID_ARG, "dptr is NULL");
  *dptr = nullptr;
  if (bytes == 0) bytes = 1;
  hipError_t e = hipMalloc(dptr, bytes);
  if (e == hipErrorOutOfMemory) return fail(RSIM_E_NOMEM, "hipMalloc(%zu) out of device memory", bytes);
  if (e != hipSuccess) return fail(RSIM_E_HIP, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return RSIM_OK;
}

int rsim_free(void* dptr) {
  if (dptr) RSIM_HIP(hipFree(dptr));
  return RSIM_OK;
}

int rsim_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return RSIM_OK;
  if (!dst || !src) return fail(RSIM_E_INVALID_ARG, "NULL pointer in memcpy");
  hipStream_t s = resolve_stream(stream);
  RSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  RSIM_HIP(hipStreamSynchronize(s));
  return RSIM_OK;
}

int rsim_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return RSIM_OK;
  if (!dst || !src) return fail(RSIM_E_INVALID_ARG, "NULL pointer in memcpy");
  hipStream_t s = resolve_stream(stream);
  RSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  RSIM_HIP(hipStreamSynchronize(s));
  return RSIM_OK;
}

int rsim_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return RSIM_OK;
  if (!dst || !src) return fail(RSIM_E_INVALID_ARG, "NULL pointer in memcpy");
  RSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, resolve_stream(stream)));
  return RSIM_OK;
}

int rsim_memset(void* dptr, int value, size_t bytes, void* stream) {
  if (bytes == 0) return RSIM_OK;
  if (!dptr) return fail(RSIM_E_INVALID_ARG, "dptr is NULL");
  if (capture_active())
    return fail(RSIM_E_UNSUPPORTED, "rsim_memset inside a graph capture; use rsim_fill_zero");
  RSIM_HIP(hipMemsetAsync(dptr, value, bytes, resolve_stream(stream)));
  return RSIM_OK;
}

int rsim_synchronize(void* stream) {
  RSIM_HIP(hipStreamSynchronize(resolve_stream(stream)));
  return RSIM_OK;
}

int rsim_stream_create(void** stream) {
  if (!stream) return fail(RSIM_E_INVALID_ARG, "stream is NULL");
  hipStream_t s;
  RSIM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return RSIM_OK;
}

int rsim_stream_destroy(void* stream) {
  if (stream) RSIM_HIP(hipStreamDestroy((hipStream_t)stream));
  return RSIM_OK;
}

int rsim_release_caches(void) {
  RSIM_HIP(hipDeviceSynchronize());
  rsim_release_tables();
  release_tables();
  release_multislice_tables();
  release_tune();
  release_detector_plans();
  release_staging();
  {
    std::lock_guard<std::recursive_mutex> lk(table_mu());
    table_forget_all();
  }
  bump_cache_generation();
  return RSIM_OK;
}

int rsim_kernel_timing(int every) { return ktime_enable(every); }

int rsim_kernel_timing_read(double* ms3, int* launches3) { return ktime_read(ms3, launches3); }

int rsim_set_progress(rsim_progress_fn fn, void* user) {
  set_progress(fn, user);
  return RSIM_OK;
}

int rsim_request_cancel(void) {
  set_cancel(1);
  return RSIM_OK;
}

int rsim_clear_cancel(void) {
  set_cancel(0);
  return RSIM_OK;
}

int rsim_graph_begin(void* stream) {
  if (t_cap) return fail(RSIM_E_INVALID_ARG, "a graph capture is already open on this thread");
  hipStream_t s = resolve_stream(stream);
  if (!s) return fail(RSIM_E_HIP, "no stream");
  // the work queued on the stream before the capture is not part of it
  RSIM_HIP(hipStreamSynchronize(s));
  RSIM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  t_cap = new CaptureState();
  t_cap->s = s;
  t_cap->device = current_device();
  table_use_log(&t_cap->used);
  ktime_suspend(true);
  return RSIM_OK;
}

int rsim_graph_end(void* stream, rsim_graph** out) {
  if (!out) return fail(RSIM_E_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (!t_cap) return fail(RSIM_E_INVALID_ARG, "no graph capture is open on this thread");
  hipStream_t s = resolve_stream(stream);
  if (s != t_cap->s)
    return fail(RSIM_E_INVALID_ARG, "rsim_graph_end on another stream than rsim_graph_begin");
  CaptureState* cap = t_cap;
  t_cap = nullptr;
  table_use_log(nullptr);
  ktime_suspend(false);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &graph);
  auto drop_owned = [&] {
    for (void* p : cap->owned) (void)hipFree(p);
  };
  if (e != hipSuccess || !graph) {
    (void)hipGetLastError();
    if (graph) hipGraphDestroy(graph);
    drop_owned();
    delete cap;
    return fail(RSIM_E_HIP, "graph capture failed: %s (a call inside it synchronised, "
                "allocated or failed)", hipGetErrorString(e));
  }
  rsim_graph* g = new rsim_graph();
  g->device = cap->device;
  g->graph = graph;
  g->owned = std::move(cap->owned);
  g->generation = cache_generation();
  for (const void* p : cap->used)
    if (table_pin(p, +1)) g->pinned.push_back(p);
  delete cap;
  (void)hipGraphGetNodes(graph, nullptr, &g->nodes);
  const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    rsim_graph_destroy(g);
    return fail(RSIM_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
  }
  *out = g;
  return RSIM_OK;
}

int rsim_graph_launch(rsim_graph* g, void* stream) {
  if (!g) return fail(RSIM_E_INVALID_ARG, "graph is NULL");
  if (t_cap) return fail(RSIM_E_INVALID_ARG, "rsim_graph_launch inside a capture");
  if (current_device() != g->device)
    return fail(RSIM_E_INVALID_ARG, "graph belongs to device %d", g->device);
  if (g->generation != cache_generation())
    return fail(RSIM_E_INVALID_ARG,
                "graph was captured before rsim_release_caches freed its tables; capture it again");
  RSIM_HIP(hipGraphLaunch(g->exec, resolve_stream(stream)));
  return RSIM_OK;
}

int rsim_graph_info(const rsim_graph* g, size_t* nodes) {
  if (!g) return fail(RSIM_E_INVALID_ARG, "graph is NULL");
  if (nodes) *nodes = g->nodes;
  return RSIM_OK;
}

int rsim_graph_destroy(rsim_graph* g) {
  if (!g) return RSIM_OK;
  if (g->generation == cache_generation())
    for (const void* p : g->pinned) table_pin(p, -1);
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  for (void* p : g->owned) (void)hipFree(p);
  delete g;
  return RSIM_OK;
}

int rsim_stream_copy(void* dst, const void* src, size_t bytes, void* stream) {
  if ((!dst || !src) && bytes) return fail(RSIM_E_INVALID_ARG, "NULL pointer in stream copy");
  if (bytes == 0) return RSIM_OK;
  return stream_copy(dst, src, bytes, resolve_stream(stream));
}

int rsim_fill_zero(void* dptr, size_t bytes, void* stream) {
  if (!dptr && bytes) return fail(RSIM_E_INVALID_ARG, "dptr is NULL");
  return fill_zero(dptr, bytes, resolve_stream(stream));
}

int rsim_cache_info(size_t* bytes, int* entries) {
  table_cache_stats(bytes, entries);
  return RSIM_OK;
}

int rsim_control_create(rsim_control** ctl) {
  if (!ctl) return fail(RSIM_E_INVALID_ARG, "ctl is NULL");
  *ctl = new (std::nothrow) rsim_control();
  if (!*ctl) return fail(RSIM_E_NOMEM, "out of host memory");
  return RSIM_OK;
}

int rsim_control_destroy(rsim_control* ctl) {
  if (!ctl) return RSIM_OK;
  if (t_ctl == ctl) t_ctl = nullptr;
  delete ctl;
  return RSIM_OK;
}

int rsim_control_set_progress(rsim_control* ctl, rsim_progress_fn fn, void* user, int every) {
  if (!ctl) return fail(RSIM_E_INVALID_ARG, "ctl is NULL");
  if (every < 1) return fail(RSIM_E_INVALID_ARG, "progress interval must be >= 1, got %d", every);
  ctl->fn = fn;
  ctl->user = user;
  ctl->every = every;
  return RSIM_OK;
}

int rsim_control_request_cancel(rsim_control* ctl) {
  if (!ctl) return fail(RSIM_E_INVALID_ARG, "ctl is NULL");
  ctl->cancel.store(1);
  return RSIM_OK;
}

int rsim_control_clear_cancel(rsim_control* ctl) {
  if (!ctl) return fail(RSIM_E_INVALID_ARG, "ctl is NULL");
  ctl->cancel.store(0);
  return RSIM_OK;
}

int rsim_control_bind(rsim_control* ctl) {
  t_ctl = ctl;
  return RSIM_OK;
}

int rsim_control_completed(const rsim_control* ctl, int64_t* completed, int64_t* total) {
  if (!ctl) return fail(RSIM_E_INVALID_ARG, "ctl is NULL");
  if (completed) *completed = ctl->completed.load();
  if (total) *total = ctl->total.load();
  return RSIM_OK;
}

int rsim_fft2(void* dst, const void* src, int nx, int ny, int dtype, int direction, void* stream) {
  RSIM_CHECK(check_field(src, nx, ny, dtype, "src"));
  RSIM_CHECK(check_field(dst, nx, ny, dtype, "dst"));
  RSIM_CHECK(check_fft_sizes(nx, ny));
  if (direction != RSIM_FORWARD && direction != RSIM_INVERSE)
    return fail(RSIM_E_INVALID_ARG, "direction must be RSIM_FORWARD or RSIM_INVERSE");
  return fft2(dtype, direction, src, dst, nx, ny, resolve_stream(stream));
}

int rsim_fft1d_rows(void* dst, const void* src, int n, int batch, int dtype, int direction,
                    void* stream) {
  RSIM_CHECK(check_field(src, n, batch, dtype, "src"));
  RSIM_CHECK(check_field(dst, n, batch, dtype, "dst"));
  RSIM_CHECK(check_fft_sizes(n, 1));
  if (direction != RSIM_FORWARD && direction != RSIM_INVERSE)
    return fail(RSIM_E_INVALID_ARG, "direction must be RSIM_FORWARD or RSIM_INVERSE");
  const double scale = direction > 0 ? 1.0 / n : 1.0;
  return fft_axis(dtype, direction, src, dst, n, 1, n, batch, 1, 0, scale, resolve_stream(stream));
}

int rsim_apply_frequency_cutoff_2d(void* U, int nx, int ny, int dtype, double freq, double dx,
                                   double dy, void* stream) {
  RSIM_CHECK(check_field(U, nx, ny, dtype, "U"));
  if (!(dx > 0) || !(dy > 0)) return fail(RSIM_E_INVALID_ARG, "dx and dy must be positive");
  return chirp_cutoff(U, nx, ny, dtype, dx, dy, 1.0, 0.0, freq, false, resolve_stream(stream));
}

int rsim_propagate_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                      double dz, double cutoff_freq, int skip_fft, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  RSIM_CHECK(check_field(U, nx, ny, dtype, "U"));
  RSIM_CHECK(check_fft_sizes(nx, ny));
  if (!(dx > 0) || !(dy > 0) || !(wl > 0))
    return fail(RSIM_E_INVALID_ARG, "dx, dy and wl must be positive");
  return propagate(u, U, nx, ny, dtype, dx, dy, wl, dz, cutoff_freq, skip_fft != 0,
                   resolve_stream(stream));
}

int rsim_propagate_analytically_2d(void* u, int nx, int ny, int dtype, double dx, double dy,
                                   double z, double x_source, double y_source, double wl,
                                   void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  if (!(wl > 0)) return fail(RSIM_E_INVALID_ARG, "wl must be positive");
  return point_source(u, nx, ny, dtype, dx, dy, z, x_source, y_source, wl, resolve_stream(stream));
}

// Shape and material-table checks of a new volume (nothing allocated).
static int volume_args(rsim_volume** vol, int nz, int ny_g, int nx_g, const double* db, int n_db) {
  if (!vol) return fail(RSIM_E_INVALID_ARG, "vol is NULL");
  *vol = nullptr;
  if (nz <= 0 || ny_g < 2 || nx_g < 2)
    return fail(RSIM_E_INVALID_ARG, "volume must have nz >= 1, ny_g >= 2, nx_g >= 2 (got %d, %d, %d)", nz, ny_g, nx_g);
  if (!db || n_db < 1 || n_db > 256)
    return fail(RSIM_E_INVALID_ARG, "deltabeta table must hold 1..256 entries (got %d)", n_db);
  if (db[0] != 0.0 || db[1] != 0.0)
    return fail(RSIM_E_INVALID_ARG, "deltabeta entry 0 (vacuum) must be 0");
  return RSIM_OK;
}

static int volume_alloc(rsim_volume** vol, int nz, int ny_g, int nx_g, const double* db, int n_db) {
  RSIM_CHECK(volume_args(vol, nz, ny_g, nx_g, db, n_db));
  static std::atomic<unsigned long long> serial{0};
  rsim_volume* v = new rsim_volume();
  v->serial = ++serial;
  v->device = current_device();
  v->nz = nz;
  v->ny_g = ny_g;
  v->nx_g = nx_g;
  v->n_db = n_db;
  v->db_host = new double[2 * (size_t)n_db];
  std::memcpy(v->db_host, db, sizeof(double) * 2 * n_db);
  const size_t nvox = (size_t)nz * ny_g * nx_g;
  hipError_t e = hipMalloc((void**)&v->ids, nvox);
  if (e != hipSuccess) {
    delete[] v->db_host;
    delete v;
    return fail(e == hipErrorOutOfMemory ? RSIM_E_NOMEM : RSIM_E_HIP, "volume allocation of %zu bytes failed: %s", nvox, hipGetErrorString(e));
  }
  e = hipMalloc((void**)&v->db_dev, sizeof(double) * 2 * n_db);
  if (e == hipSuccess) e = hipMemcpy(v->db_dev, db, sizeof(double) * 2 * n_db, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(v->ids);
    delete[] v->db_host;
    delete v;
    return fail(RSIM_E_HIP, "volume table upload failed: %s", hipGetErrorString(e));
  }
  *vol = v;
  return RSIM_OK;
}

static inline uint32_t host_id(const void* ids, int id_bytes, size_t i) {
  if (id_bytes == 1) return ((const uint8_t*)ids)[i];
  if (id_bytes == 2) return ((const uint16_t*)ids)[i];
  return ((const uint32_t*)ids)[i];
}

int rsim_volume_create(rsim_volume** vol, const void* ids, int id_bytes, int nz, int ny_g, int nx_g,
                       const double* deltabeta, int n_db) {
  if (!ids) return fail(RSIM_E_INVALID_ARG, "ids is NULL");
  if (id_bytes != 1 && id_bytes != 2 && id_bytes != 4)
    return fail(RSIM_E_INVALID_ARG, "id_bytes must be 1, 2 or 4 (got %d)", id_bytes);
  RSIM_CHECK(volume_args(vol, nz, ny_g, nx_g, deltabeta, n_db));
  // every id is checked on the host before any device memory is taken
  const size_t plane = (size_t)ny_g * nx_g;
  const size_t total = (size_t)nz * plane;
  for (size_t i = 0; i < total; ++i) {
    const uint32_t id = host_id(ids, id_bytes, i);
    if (id >= (uint32_t)n_db)
      return fail(RSIM_E_INVALID_ARG, "material id %u at voxel %zu is outside the %d-entry deltabeta table", id, i, n_db);
  }
  RSIM_CHECK(volume_alloc(vol, nz, ny_g, nx_g, deltabeta, n_db));
  rsim_volume* v = *vol;
  std::vector<uint8_t> buf(plane);
  for (int z = 0; z < nz; ++z) {
    const size_t off = (size_t)z * plane;
    for (size_t i = 0; i < plane; ++i) buf[i] = (uint8_t)host_id(ids, id_bytes, off + i);
    hipError_t e = hipMemcpy(v->ids + off, buf.data(), plane, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      rsim_volume_destroy(v);
      *vol = nullptr;
      return fail(RSIM_E_HIP, "volume upload failed: %s", hipGetErrorString(e));
    }
  }
  return RSIM_OK;
}

int rsim_volume_create_empty(rsim_volume** vol, int nz, int ny_g, int nx_g, const double* deltabeta,
                             int n_db) {
  RSIM_CHECK(volume_alloc(vol, nz, ny_g, nx_g, deltabeta, n_db));
  RSIM_HIP(hipMemset((*vol)->ids, 0, (size_t)nz * ny_g * nx_g));
  return RSIM_OK;
}

int rsim_volume_destroy(rsim_volume* vol) {
  if (!vol) return RSIM_OK;
  release_volume_tables(vol->serial);
  hipFree(vol->ids);
  hipFree(vol->db_dev);
  delete[] vol->db_host;
  delete vol;
  return RSIM_OK;
}

int rsim_volume_info(const rsim_volume* vol, void** dev_ids, int* nz, int* ny_g, int* nx_g) {
  if (!vol) return fail(RSIM_E_INVALID_ARG, "vol is NULL");
  if (dev_ids) *dev_ids = vol->ids;
  if (nz) *nz = vol->nz;
  if (ny_g) *ny_g = vol->ny_g;
  if (nx_g) *nx_g = vol->nx_g;
  return RSIM_OK;
}

int rsim_volume_paint_spheres(rsim_volume* vol, const double* spheres, const int* mats, int n,
                              void* stream) {
  if (!vol) return fail(RSIM_E_INVALID_ARG, "vol is NULL");
  if (n < 0 || (n > 0 && (!spheres || !mats))) return fail(RSIM_E_INVALID_ARG, "bad sphere list");
  return paint_spheres(vol, spheres, mats, n, resolve_stream(stream));
}

static int check_sample(const rsim_volume* vol, int slice0, int count, double px, double py,
                        double pz) {
  if (!vol) return fail(RSIM_E_INVALID_ARG, "vol is NULL");
  if (!(px > 0) || !(py > 0) || !(pz > 0))
    return fail(RSIM_E_INVALID_ARG, "pixel sizes must be positive");
  if (slice0 < 0 || count < 0 || slice0 + count > vol->nz)
    return fail(RSIM_E_INVALID_ARG, "invalid Sample slice range %d..%d of %d", slice0, slice0 + count, vol->nz);
  if (vol->device != current_device())
    return fail(RSIM_E_INVALID_ARG, "volume lives on device %d, current device is %d", vol->device, current_device());
  return RSIM_OK;
}

int rsim_transmission_2d(void* u, int nx, int ny, int dtype, double dx, double dy, double wl,
                         const rsim_volume* vol, int slice, double pixel_size_x,
                         double pixel_size_y, double pixel_size_z, double x_pos, double y_pos,
                         void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  RSIM_CHECK(check_sample(vol, slice, 1, pixel_size_x, pixel_size_y, pixel_size_z));
  if (!(wl > 0)) return fail(RSIM_E_INVALID_ARG, "wl must be positive");
  return transmission(u, nx, ny, dtype, dx, dy, wl, vol, slice, pixel_size_x, pixel_size_y,
                      pixel_size_z, x_pos, y_pos, 1.0, resolve_stream(stream));
}

int rsim_transmission_2d_scaled(void* u, int nx, int ny, int dtype, double dx, double dy,
                                double wl, const rsim_volume* vol, int slice, double pixel_size_x,
                                double pixel_size_y, double pixel_size_z, double x_pos,
                                double y_pos, double coord_scale, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  RSIM_CHECK(check_sample(vol, slice, 1, pixel_size_x, pixel_size_y, pixel_size_z));
  if (!(wl > 0)) return fail(RSIM_E_INVALID_ARG, "wl must be positive");
  if (!(coord_scale > 0) || !std::isfinite(coord_scale))
    return fail(RSIM_E_INVALID_ARG, "coord_scale must be positive and finite");
  return transmission(u, nx, ny, dtype, dx, dy, wl, vol, slice, pixel_size_x, pixel_size_y,
                      pixel_size_z, x_pos, y_pos, coord_scale, resolve_stream(stream));
}

int rsim_tilted_plane_2d(void* u, int nx, int ny, int dtype, double dx, double dy, double wl,
                         double dz, double x_source, double y_source, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  if (!(wl > 0) || !(dz > 0) || !std::isfinite(dz))
    return fail(RSIM_E_INVALID_ARG, "Fresnel source-to-sample distance must be positive");
  return tilted_plane(u, nx, ny, dtype, dx, dy, wl, dz, x_source, y_source, resolve_stream(stream));
}

int rsim_sample_apply_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                         double wl, double cutoff_freq, const rsim_volume* vol,
                         double pixel_size_x, double pixel_size_y, double pixel_size_z,
                         double x_pos, double y_pos, int first_slice, int n_slices, int engine,
                         int flags, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  RSIM_CHECK(check_field(U, nx, ny, dtype, "U"));
  RSIM_CHECK(check_fft_sizes(nx, ny));
  RSIM_CHECK(check_sample(vol, first_slice, n_slices, pixel_size_x, pixel_size_y, pixel_size_z));
  if (engine != RSIM_PLAIN && engine != RSIM_FUSED)
    return fail(RSIM_E_INVALID_ARG, "engine must be RSIM_PLAIN or RSIM_FUSED");
  if (!(wl > 0) || !(dx > 0) || !(dy > 0)) return fail(RSIM_E_INVALID_ARG, "dx, dy and wl must be positive");
  return sample_apply(u, U, nx, ny, dtype, dx, dy, wl, cutoff_freq, vol, pixel_size_x,
                      pixel_size_y, pixel_size_z, x_pos, y_pos, first_slice, n_slices, nullptr,
                      nullptr, engine, flags, resolve_stream(stream));
}

namespace {
int check_chain_steps(const rsim_chain_step* steps, int n_steps, int inst) {
  if (n_steps < 0 || (n_steps > 0 && !steps))
    return fail(RSIM_E_INVALID_ARG, "a chain needs n_steps >= 0 and its steps");
  for (int j = 0; j < n_steps; ++j) {
    const rsim_chain_step& st = steps[j];
    if (st.vol) {
      const int rc = check_sample(st.vol, st.first_slice, st.n_slices, st.pixel_size_x,
                                  st.pixel_size_y, st.pixel_size_z);
      if (rc != RSIM_OK) {
        const std::string why = rsim_last_error();
        if (inst >= 0) return fail(rc, "instance %d, chain step %d: %s", inst, j, why.c_str());
        return fail(rc, "chain step %d: %s", j, why.c_str());
      }
    } else if (!std::isfinite(st.dz)) {
      if (inst >= 0)
        return fail(RSIM_E_INVALID_ARG, "instance %d, chain step %d: free-space distance must be finite",
                    inst, j);
      return fail(RSIM_E_INVALID_ARG, "chain step %d: free-space distance must be finite", j);
    }
  }
  return RSIM_OK;
}
}  // namespace

int rsim_chain_apply_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                        double wl, double cutoff_freq, int band_limit,
                        const rsim_chain_step* steps, int n_steps, int engine, int flags,
                        void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  RSIM_CHECK(check_field(U, nx, ny, dtype, "U"));
  RSIM_CHECK(check_fft_sizes(nx, ny));
  if (engine != RSIM_PLAIN && engine != RSIM_FUSED)
    return fail(RSIM_E_INVALID_ARG, "engine must be RSIM_PLAIN or RSIM_FUSED");
  if (!(wl > 0) || !(dx > 0) || !(dy > 0)) return fail(RSIM_E_INVALID_ARG, "dx, dy and wl must be positive");
  RSIM_CHECK(check_chain_steps(steps, n_steps, -1));
  const rsim_chain_instance one{u, U, dx, dy, wl, cutoff_freq, steps};
  return chain_apply(&one, 1, nx, ny, dtype, band_limit != 0, n_steps, engine, flags,
                     resolve_stream(stream));
}

int rsim_chain_apply_2d_batch(const rsim_chain_instance* instances, int count, int nx, int ny,
                              int dtype, int band_limit, int n_steps, int engine, int flags,
                              void* stream) {
  if (count < 0 || (count > 0 && !instances))
    return fail(RSIM_E_INVALID_ARG, "a batched chain needs count >= 0 and its instances");
  RSIM_CHECK(check_fft_sizes(nx, ny));
  if (engine != RSIM_PLAIN && engine != RSIM_FUSED)
    return fail(RSIM_E_INVALID_ARG, "engine must be RSIM_PLAIN or RSIM_FUSED");
  for (int i = 0; i < count; ++i) {
    const rsim_chain_instance& q = instances[i];
    RSIM_CHECK(check_field(q.u, nx, ny, dtype, "u"));
    RSIM_CHECK(check_field(q.U, nx, ny, dtype, "U"));
    if (!(q.wl > 0) || !(q.dx > 0) || !(q.dy > 0))
      return fail(RSIM_E_INVALID_ARG, "instance %d: dx, dy and wl must be positive", i);
    RSIM_CHECK(check_chain_steps(q.steps, n_steps, i));
    for (int j = 0; j < n_steps; ++j) {
      const rsim_chain_step& a = q.steps[j];
      const rsim_chain_step& b = instances[0].steps[j];
      if ((a.vol == nullptr) != (b.vol == nullptr) || (a.vol && a.n_slices != b.n_slices))
        return fail(RSIM_E_INVALID_ARG,
                    "instance %d, chain step %d: every instance needs the same step kinds and "
                    "slice counts as instance 0",
                    i, j);
    }
    for (int k = 0; k < i; ++k)
      if (instances[k].u == q.u || instances[k].U == q.U || instances[k].u == q.U ||
          instances[k].U == q.u)
        return fail(RSIM_E_INVALID_ARG, "instances %d and %d share a field", k, i);
  }
  return chain_apply(instances, count, nx, ny, dtype, band_limit != 0, n_steps, engine, flags,
                     resolve_stream(stream));
}

}  // extern "C"

namespace {
int history_spec(const rsim_history_frames* h, int n_slices, HistSpec* out) {
  if (!h) return RSIM_OK;
  if (!h->frames && n_slices > 0) return fail(RSIM_E_INVALID_ARG, "history frames buffer is NULL");
  if (h->out_x <= 0 || h->out_y <= 0)
    return fail(RSIM_E_INVALID_ARG, "2D detector output must be non-empty, got (%d, %d)", h->out_y,
                h->out_x);
  if (!(h->pixel_size_x > 0) || !(h->pixel_size_y > 0))
    return fail(RSIM_E_INVALID_ARG, "2D detector grid and pixel spacings must be positive");
  if (!std::isfinite(h->z_start)) return fail(RSIM_E_INVALID_ARG, "history z_start must be finite");
  out->ox = h->out_x;
  out->oy = h->out_y;
  out->dsx = h->pixel_size_x;
  out->dsy = h->pixel_size_y;
  out->z_start = h->z_start;
  out->frames = h->frames;
  return RSIM_OK;
}

int sample_apply_checked(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                         double wl, double cutoff_freq, const rsim_volume* vol,
                         double pixel_size_x, double pixel_size_y, double pixel_size_z,
                         double x_pos, double y_pos, int first_slice, int n_slices,
                         const double* coord_scale, const double* slice_dz, int engine, int flags,
                         const rsim_history_frames* history, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  RSIM_CHECK(check_field(U, nx, ny, dtype, "U"));
  RSIM_CHECK(check_fft_sizes(nx, ny));
  RSIM_CHECK(check_sample(vol, first_slice, n_slices, pixel_size_x, pixel_size_y, pixel_size_z));
  if (engine != RSIM_PLAIN && engine != RSIM_FUSED)
    return fail(RSIM_E_INVALID_ARG, "engine must be RSIM_PLAIN or RSIM_FUSED");
  if (!(wl > 0) || !(dx > 0) || !(dy > 0)) return fail(RSIM_E_INVALID_ARG, "dx, dy and wl must be positive");
  for (int k = 0; k < n_slices; ++k) {
    if (coord_scale && (!(coord_scale[k] > 0) || !std::isfinite(coord_scale[k])))
      return fail(RSIM_E_INVALID_ARG, "coord_scale[%d] must be positive and finite", k);
    if (slice_dz && !std::isfinite(slice_dz[k]))
      return fail(RSIM_E_INVALID_ARG, "slice_dz[%d] must be finite", k);
  }
  HistSpec hs;
  RSIM_CHECK(history_spec(history, n_slices, &hs));
  return sample_apply(u, U, nx, ny, dtype, dx, dy, wl, cutoff_freq, vol, pixel_size_x,
                      pixel_size_y, pixel_size_z, x_pos, y_pos, first_slice, n_slices,
                      coord_scale, slice_dz, engine, flags, resolve_stream(stream),
                      history ? &hs : nullptr);
}
}  // namespace

extern "C" {

int rsim_sample_apply_2d_scaled(void* u, void* U, int nx, int ny, int dtype, double dx,
                                double dy, double wl, double cutoff_freq, const rsim_volume* vol,
                                double pixel_size_x, double pixel_size_y, double pixel_size_z,
                                double x_pos, double y_pos, int first_slice, int n_slices,
                                const double* coord_scale, const double* slice_dz, int engine,
                                int flags, void* stream) {
  return sample_apply_checked(u, U, nx, ny, dtype, dx, dy, wl, cutoff_freq, vol, pixel_size_x,
                              pixel_size_y, pixel_size_z, x_pos, y_pos, first_slice, n_slices,
                              coord_scale, slice_dz, engine, flags, nullptr, stream);
}

int rsim_sample_apply_2d_history(void* u, void* U, int nx, int ny, int dtype, double dx,
                                 double dy, double wl, double cutoff_freq,
                                 const rsim_volume* vol, double pixel_size_x,
                                 double pixel_size_y, double pixel_size_z, double x_pos,
                                 double y_pos, int first_slice, int n_slices,
                                 const double* coord_scale, const double* slice_dz, int engine,
                                 int flags, const rsim_history_frames* history, void* stream) {
  if (!history) return fail(RSIM_E_INVALID_ARG, "history is NULL");
  return sample_apply_checked(u, U, nx, ny, dtype, dx, dy, wl, cutoff_freq, vol, pixel_size_x,
                              pixel_size_y, pixel_size_z, x_pos, y_pos, first_slice, n_slices,
                              coord_scale, slice_dz, engine, flags, history, stream);
}

int rsim_fused_schedule(int nx, int ny, int* row_variant, int* col_variant) {
  if (!row_variant || !col_variant) return fail(RSIM_E_INVALID_ARG, "NULL output pointer");
  int tile = 0;
  return fused_schedule_lookup(nx, ny, row_variant, col_variant, &tile);
}

int rsim_fused_tile(int nx, int ny, int* tile_width) {
  if (!tile_width) return fail(RSIM_E_INVALID_ARG, "NULL output pointer");
  int rv = 0, cv = 0;
  return fused_schedule_lookup(nx, ny, &rv, &cv, tile_width);
}

int rsim_detector_area_v1(const void* u, int nx, int ny, int dtype, double dx, double dy,
                          int out_x, int out_y, double ds_x, double ds_y, double z,
                          double* out_host, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  if (!out_host) return fail(RSIM_E_INVALID_ARG, "out_host is NULL");
  if (!(dx > 0) || !(dy > 0) || !(ds_x > 0) || !(ds_y > 0))
    return fail(RSIM_E_INVALID_ARG, "2D detector grid and pixel spacings must be positive");
  hipStream_t s = resolve_stream(stream);
  const size_t bytes = sizeof(double) * (size_t)out_x * (size_t)(out_y > 0 ? out_y : 0);
  double* d = nullptr;
  if (out_x > 0 && out_y > 0) RSIM_HIP(hipMallocAsync((void**)&d, bytes, s));
  int rc = detector_area_v1(u, nx, ny, dtype, dx, dy, out_x, out_y, ds_x, ds_y, z, 1.0, false, d, s);
  if (rc == RSIM_OK) {
    RSIM_HIP(hipMemcpyAsync(out_host, d, bytes, hipMemcpyDeviceToHost, s));
    RSIM_HIP(hipStreamSynchronize(s));
  }
  if (d) hipFreeAsync(d, s);
  return rc;
}

int rsim_detector_area_v1_accumulate(const void* u, int nx, int ny, int dtype, double dx, double dy,
                                     int out_x, int out_y, double ds_x, double ds_y, double z,
                                     double weight, double* out_dev, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  if (!out_dev) return fail(RSIM_E_INVALID_ARG, "out_dev is NULL");
  if (!(dx > 0) || !(dy > 0) || !(ds_x > 0) || !(ds_y > 0))
    return fail(RSIM_E_INVALID_ARG, "2D detector grid and pixel spacings must be positive");
  return detector_area_v1(u, nx, ny, dtype, dx, dy, out_x, out_y, ds_x, ds_y, z, weight, true,
                          out_dev, resolve_stream(stream));
}

int rsim_detector_area_v1_accumulate_noalloc(const void* u, int nx, int ny, int dtype, double dx,
                                             double dy, int out_x, int out_y, double ds_x,
                                             double ds_y, double z, double weight,
                                             double* out_dev, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  if (!out_dev) return fail(RSIM_E_INVALID_ARG, "out_dev is NULL");
  if (!(dx > 0) || !(dy > 0) || !(ds_x > 0) || !(ds_y > 0))
    return fail(RSIM_E_INVALID_ARG, "2D detector grid and pixel spacings must be positive");
  return detector_area_v1_noalloc(u, nx, ny, dtype, dx, dy, out_x, out_y, ds_x, ds_y, z, weight,
                                  out_dev, resolve_stream(stream));
}

int rsim_detector_noalloc_capacity(size_t* bytes) {
  if (!bytes) return fail(RSIM_E_INVALID_ARG, "bytes is NULL");
  *bytes = detector_noalloc_capacity();
  return RSIM_OK;
}

int rsim_detector_area_v1_adjoint(const void* v, void* g, int nx, int ny, int dtype, double dx,
                                  double dy, int out_x, int out_y, double ds_x, double ds_y,
                                  double z, const double* grad_img, void* stream) {
  RSIM_CHECK(check_field(v, nx, ny, dtype, "v"));
  RSIM_CHECK(check_field(g, nx, ny, dtype, "g"));
  if (!grad_img) return fail(RSIM_E_INVALID_ARG, "grad_img is NULL");
  if (!(dx > 0) || !(dy > 0) || !(ds_x > 0) || !(ds_y > 0))
    return fail(RSIM_E_INVALID_ARG, "2D detector grid and pixel spacings must be positive");
  return detector_area_v1_adjoint(v, g, nx, ny, dtype, dx, dy, out_x, out_y, ds_x, ds_y, z,
                                  grad_img, resolve_stream(stream));
}

static int check_engine(int engine) {
  if (engine != RSIM_PLAIN && engine != RSIM_FUSED)
    return fail(RSIM_E_INVALID_ARG, "engine must be RSIM_PLAIN or RSIM_FUSED");
  return RSIM_OK;
}

// butterfly arithmetic of the fused differentiable chain: at most one mode
static int check_db_flags(int flags) {
  const int math = RSIM_FLAG_MATH_FP32 | RSIM_FLAG_MATH_FP64 | RSIM_FLAG_MATH_COMP32;
  if (flags & ~math) return fail(RSIM_E_INVALID_ARG, "flags may only select the butterfly arithmetic");
  const int m = flags & math;
  if (m & (m - 1)) return fail(RSIM_E_INVALID_ARG, "at most one RSIM_FLAG_MATH_* flag");
  return RSIM_OK;
}

static int check_db_args(int nx, int ny, double dx, double dy, double wl, double pz,
                         const float* delta, const float* beta, int n, int every) {
  RSIM_CHECK(check_fft_sizes(nx, ny));
  if (!(dx > 0) || !(dy > 0) || !(wl > 0)) return fail(RSIM_E_INVALID_ARG, "dx, dy and wl must be positive");
  if (!(pz > 0) || !std::isfinite(pz)) return fail(RSIM_E_INVALID_ARG, "pz must be positive and finite");
  if (!delta || !beta) return fail(RSIM_E_INVALID_ARG, "delta and beta volumes must not be NULL");
  if (n < 1) return fail(RSIM_E_INVALID_ARG, "n_slices must be >= 1");
  if (every < 1) return fail(RSIM_E_INVALID_ARG, "ckpt_every must be >= 1");
  return RSIM_OK;
}

int rsim_transmission_db_2d(void* u, int nx, int ny, int dtype, double wl, double pz,
                            const float* delta, const float* beta, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  if (!delta || !beta) return fail(RSIM_E_INVALID_ARG, "delta and beta must not be NULL");
  if (!(wl > 0) || !std::isfinite(pz)) return fail(RSIM_E_INVALID_ARG, "wl must be positive, pz finite");
  return transmission_db_2d(u, nx, ny, dtype, wl, pz, delta, beta, resolve_stream(stream));
}

int rsim_multislice_db_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                          double wl, double cutoff_freq, const float* delta, const float* beta,
                          double pz, int n_slices, void* ckpt, int ckpt_every, int engine,
                          int flags, void* stream) {
  RSIM_CHECK(check_field(u, nx, ny, dtype, "u"));
  RSIM_CHECK(check_field(U, nx, ny, dtype, "U"));
  RSIM_CHECK(check_db_args(nx, ny, dx, dy, wl, pz, delta, beta, n_slices, ckpt ? ckpt_every : 1));
  RSIM_CHECK(check_engine(engine));
  RSIM_CHECK(check_db_flags(flags));
  return multislice_db(u, U, nx, ny, dtype, dx, dy, wl, cutoff_freq, delta, beta, pz, n_slices,
                       ckpt, ckpt ? ckpt_every : 1, engine, flags, resolve_stream(stream));
}

int rsim_multislice_adjoint_workspace(int nx, int ny, int dtype, int n_slices, int ckpt_every,
                                      size_t* ckpt_bytes, size_t* work_bytes) {
  if (!valid_dtype(dtype)) return fail(RSIM_E_INVALID_ARG, "unknown dtype %d", dtype);
  if (nx < 1 || ny < 1 || n_slices < 1 || ckpt_every < 1)
    return fail(RSIM_E_INVALID_ARG, "sizes, n_slices and ckpt_every must be >= 1");
  const size_t fb = (size_t)nx * ny * elem_bytes(dtype);
  const int seg = ckpt_every < n_slices ? ckpt_every : n_slices;
  const int nseg = (n_slices + ckpt_every - 1) / ckpt_every;
  // complex64: the fused engine rebuilds the next segment while the adjoint
  // runs through the current one (two slot sets)
  const size_t sets = (dtype == RSIM_C64 && nseg >= 2) ? 2 : 1;
  if (ckpt_bytes) *ckpt_bytes = (size_t)nseg * fb;
  if (work_bytes) *work_bytes = sets * (size_t)(seg - 1) * fb;
  return RSIM_OK;
}

int rsim_multislice_db_adjoint_2d(void* g, void* U, int nx, int ny, int dtype, double dx,
                                  double dy, double wl, double cutoff_freq, const float* delta,
                                  const float* beta, double pz, int n_slices, const void* ckpt,
                                  int ckpt_every, void* work, float* grad_delta,
                                  float* grad_beta, int engine, int flags, void* stream) {
  RSIM_CHECK(check_field(g, nx, ny, dtype, "g"));
  RSIM_CHECK(check_field(U, nx, ny, dtype, "U"));
  RSIM_CHECK(check_db_args(nx, ny, dx, dy, wl, pz, delta, beta, n_slices, ckpt_every));
  if (!ckpt) return fail(RSIM_E_INVALID_ARG, "ckpt is NULL");
  if (!grad_delta || !grad_beta) return fail(RSIM_E_INVALID_ARG, "gradient outputs must not be NULL");
  if (!work && ckpt_every > 1 && n_slices > 1)
    return fail(RSIM_E_INVALID_ARG, "work is NULL but ckpt_every > 1");
  RSIM_CHECK(check_engine(engine));
  RSIM_CHECK(check_db_flags(flags));
  return multislice_db_adjoint(g, U, nx, ny, dtype, dx, dy, wl, cutoff_freq, delta, beta, pz,
                               n_slices, ckpt, ckpt_every, work, grad_delta, grad_beta, engine,
                               flags, resolve_stream(stream));
}

int rsim_batch_create(rsim_batch** batch, const rsim_batch_instance* instances, int count,
                      int first_slice, int n_slices, int flags, void* stream) {
  if (!batch) return fail(RSIM_E_INVALID_ARG, "batch is NULL");
  *batch = nullptr;
  if (!instances || count < 1) return fail(RSIM_E_INVALID_ARG, "a batch needs at least one instance");
  if (first_slice < 0 || n_slices < 1)
    return fail(RSIM_E_INVALID_ARG, "first_slice must be >= 0 and n_slices >= 1");
  for (int i = 0; i < count; ++i) {
    const rsim_batch_instance& q = instances[i];
    RSIM_CHECK(check_field(q.u, q.nx, q.ny, RSIM_C64, "instance u"));
    RSIM_CHECK(check_field(q.U, q.nx, q.ny, RSIM_C64, "instance U"));
    if (!q.vol) return fail(RSIM_E_INVALID_ARG, "instance %d: vol is NULL", i);
    if (first_slice + n_slices > q.vol->nz)
      return fail(RSIM_E_INVALID_ARG, "instance %d: slices [%d, %d) exceed the volume depth %d", i,
                  first_slice, first_slice + n_slices, q.vol->nz);
    if (!(q.dx > 0) || !(q.dy > 0) || !(q.wl > 0) || !(q.pixel_size_x > 0) ||
        !(q.pixel_size_y > 0) || !(q.pixel_size_z > 0))
      return fail(RSIM_E_INVALID_ARG, "instance %d: spacings and wl must be positive", i);
  }
  return batch_create(batch, instances, count, first_slice, n_slices, flags, resolve_stream(stream));
}

int rsim_batch_run(rsim_batch* batch, void* stream) {
  if (!batch) return fail(RSIM_E_INVALID_ARG, "batch is NULL");
  return batch_run(batch, resolve_stream(stream));
}

int rsim_batch_info(const rsim_batch* batch, int* groups, int* kernels) {
  if (!batch) return fail(RSIM_E_INVALID_ARG, "batch is NULL");
  return batch_info(batch, groups, kernels);
}

int rsim_batch_destroy(rsim_batch* batch) {
  if (!batch) return RSIM_OK;
  return batch_destroy(batch);
}

}  // extern "C"
