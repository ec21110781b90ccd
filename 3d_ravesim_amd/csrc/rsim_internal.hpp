// Host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <functional>
#include <string>

#include "../../include/ravesim.h"

namespace rsim {

int fail(int code, const char* fmt, ...);
void clear_error();

#define RSIM_HIP(call)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::rsim::fail(RSIM_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                          __FILE__, __LINE__);                                             \
  } while (0)

#define RSIM_CHECK(expr)          \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != RSIM_OK) return rc_; \
  } while (0)

#define RSIM_LAUNCHED()                                                                        \
  do {                                                                                         \
    hipError_t e_ = hipGetLastError();                                                         \
    if (e_ != hipSuccess)                                                                      \
      return ::rsim::fail(RSIM_E_HIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                          __FILE__, __LINE__);                                                 \
  } while (0)

hipStream_t resolve_stream(void* s);
// a second library stream per device (default priority) for work that runs
// beside the caller's stream, ordered with it by events
hipStream_t side_stream();
int current_device();
size_t max_lds_per_block();

inline size_t elem_bytes(int dtype) { return dtype == RSIM_C128 ? 16 : 8; }
// internal table kind: compensated FP32 (hi + lo) twiddles
constexpr int kTwComp = 2;
inline bool valid_dtype(int dtype) { return dtype == RSIM_C64 || dtype == RSIM_C128; }

// ---- FFT (rsim_fft_host.hip) ----
bool fft_size_supported(int n);
// One batch of 1D transforms of length N along an axis.  Sequence q in
// [0, nseq) of batch b starts at b*bs + q*ss; its elements are es apart.
int fft_axis(int dtype, int dir, const void* in, void* out, int N, long long es, long long ss,
             int nseq, int batch, long long bs, double scale, hipStream_t s);
int fft2(int dtype, int dir, const void* in, void* out, int nx, int ny, hipStream_t s);
// Per-stage twiddle tables of the register schedule for N (nullptr entries
// for stages without twiddles); false when N has no register schedule.
bool reg_stage_tables(int dtype, int N, const void* out[4]);
// W_N^m, m in [0, N), forward sign, in dtype's table format (nullptr on error).
const void* fft_full_table(int dtype, int N);

// ---- launchers implemented in the per-precision register FFT units ----
bool launch_reg_fft_c64(int dir, int N, bool colmode, const void* in, void* out, long long es,
                        long long ss, long long bs, int nseq, int batch, const void* const* tw,
                        double scale, hipStream_t s);
bool launch_reg_fft_c128(int dir, int N, bool colmode, const void* in, void* out, long long es,
                         long long ss, long long bs, int nseq, int batch, const void* const* tw,
                         double scale, hipStream_t s);

// ---- pointwise (rsim_pointwise.hip) ----
int chirp_cutoff(void* U, int nx, int ny, int dtype, double dx, double dy, double wl, double dz,
                 double freq, bool chirp, hipStream_t s);

// ---- multislice over a voxel source (rsim_multislice.hip) ----
// Where slice transmissions come from: material ids with a delta-beta table
// (Sample), or a per-voxel delta-beta field (PlasmaSample).  ``plain`` applies
// slice k's FP64 transmission to a field (plain engine); ``field`` is the
// per-voxel (delta, beta) FP32 volume the fused engine interpolates.
struct VoxSrc {
  int nz = 0, ny_g = 0, nx_g = 0;
  const struct ::rsim_volume* vol = nullptr;
  const float2* field = nullptr;
  std::function<int(void* u, int dtype, int slice, double scale, hipStream_t s)> plain;
};
// History frames of a slice loop (optical_element.py:548-551): frame k is the
// area_v1 image of the field entering slice first + k, taken at
// z = z_start + (first + k) * pz, into frames + k * oy * ox (device).
struct HistSpec {
  int ox = 0, oy = 0;
  double dsx = 0.0, dsy = 0.0, z_start = 0.0;
  double* frames = nullptr;
};
// Slices [first, first + count) of a sample applied to u (optical_element.py:
// 502-596): transmission, then propagation over dzs[k] (pz when null) with the
// voxel map scaled by scales[k] (1 when null).  U = FFT2(u) afterwards unless
// RSIM_FLAG_NO_SPECTRUM.
int slices_apply(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                 double freq, const VoxSrc& src, double px, double py, double pz, double xpos,
                 double ypos, int first, int count, const double* scales, const double* dzs,
                 int engine, int flags, hipStream_t s, const HistSpec* hist = nullptr);
int detector_area_v1(const void* u, int nx, int ny, int dtype, double dx, double dy, int ox,
                     int oy, double dsx, double dsy, double z, double weight, bool accumulate,
                     double* out_dev, hipStream_t s);
// area_v1 split at the field rows.  Stage 1 (inside a row pass): field row j
// in [j0, j0 + nj) and detector column p give the x-integral
//   xs[(j - j0) * ox + p] = dx * (sum_{xi0 <= i < xi1} |u|^2
//                                 + xf1 |u[xi1]|^2 - xf0 |u[xi0]|^2)
// (end terms only for cells inside the field).  Stage 2 (detector_ybin)
// sums them over each detector row's overlapping field rows.
struct XBinArgs {
  const int* xi0;
  const int* xi1;
  const double* xf0;
  const double* xf1;
  double* xs;   // nj x ox, set by the caller
  int ox, j0, nj;
  double dx;
};
// The cached plan's stage-1 arguments; *ok false when the detector reads no
// field row or its column ranges are not monotone (then use detector_area_v1).
int detector_xbin_plan(int nx, int ny, double dx, double dy, int ox, int oy, double dsx,
                       double dsy, XBinArgs* xb, bool* ok);
int detector_ybin(const double* xs, int nx, int ny, double dx, double dy, int ox, int oy,
                  double dsx, double dsy, double z, double weight, bool accumulate,
                  double* out_dev, hipStream_t s);
bool fused_supported(int nx, int ny, int dtype);
// per-voxel delta/beta multislice and its adjoint on the fused engine
bool fused_db_supported(int nx, int ny, int dtype);
int fused_multislice_db(void* u, void* U, int nx, int ny, double dx, double dy, double wl,
                        double freq, const float* delta, const float* beta, double pz, int n,
                        void* ckpt, int every, int flags, hipStream_t s);
int fused_multislice_db_adjoint(void* g, void* U, int nx, int ny, double dx, double dy,
                                double wl, double freq, const float* delta, const float* beta,
                                double pz, int n, const void* ckpt, int every, void* work,
                                float* gdelta, float* gbeta, int flags, hipStream_t s);
int propagate(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
              double dz, double freq, bool skip_fft, hipStream_t s);
int sample_apply(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                 double freq, const rsim_volume* vol, double px, double py, double pz,
                 double xpos, double ypos, int first, int count, const double* scales,
                 const double* dzs, int engine, int flags, hipStream_t s,
                 const HistSpec* hist = nullptr);
// rsim_chain_apply_2d[_batch] after their argument checks
int chain_apply(const rsim_chain_instance* in, int count, int nx, int ny, int dtype,
                bool band_limit, int n_steps, int engine, int flags, hipStream_t s);

// ---- kernel timing (kinds: 0 row pass, 1 column pass, 2 pointwise) ----
enum { KT_ROW = 0, KT_COL = 1, KT_PW = 2 };
bool ktime_on();
void ktime_begin(int kind, hipStream_t s);
void ktime_end(int kind, hipStream_t s);
// suspend kernel timing on this thread (while a stream is being captured)
void ktime_suspend(bool on);

// ---- graph capture (rsim_graph_begin .. rsim_graph_end on this thread) ----
bool capture_active();
// device scratch that a captured graph reads on every replay: allocated now,
// owned by the graph being captured and freed with it
int capture_alloc(size_t bytes, void** p);
// RSIM_OK outside a capture; inside, RSIM_E_UNSUPPORTED naming what would
// have to be built (tables, tuning and scratch are built by running the same
// calls once before the capture)
int capture_miss(const char* what);

// ---- cancel / progress (the run control bound to the calling thread, or
// the process-wide flag and callback when none is bound) ----
bool cancel_requested();
// true when slice `completed` of `total` is to be reported: the slice loop
// then waits for the GPU first, so a report means the work is done
bool progress_due(long long completed, long long total);
void report_progress(const char* stage, long long completed, long long total);
// record how many slices a loop finished (on return or on cancel)
void note_done(long long completed, long long total);

}  // namespace rsim

struct rsim_volume {
  unsigned long long serial;   // unique per created volume (cache key; addresses are reused)
  int device;
  int nz, ny_g, nx_g;
  uint8_t* ids;      // device, nz*ny_g*nx_g
  int n_db;
  double* db_host;   // 2*n_db (re, im), entry 0 = vacuum
  double* db_dev;    // device copy of db_host
};
