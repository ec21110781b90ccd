/*
 * ravesim.h -- C ABI of the MI355X-native RAVE-SIM 2D Fresnel multislice engine
 * (libravesim_hip.so, built from 3d_ravesim_amd/csrc for gfx950).
 *
 * Conventions
 *  - Every entry point returns RSIM_OK (0) or a negative RSIM_E_* code; the
 *    message of the last failure on the calling thread is rsim_last_error().
 *    Nothing in the library calls exit() (the reference's fast-wave does,
 *    fast-wave/fwcuda/types.hpp:226-239).
 *  - Fields are device buffers of nx*ny complex values, row-major (ny rows of
 *    nx), interleaved re/im: complex64 (dtype RSIM_C64) or complex128
 *    (RSIM_C128) -- the numpy layout big-wave uses (big-wave/vector.py:319-385)
 *    and fast-wave's DevComplex (fast-wave/fwcuda/types.hpp:14-39).
 *  - `stream` is a hipStream_t; NULL selects the library's own stream for the
 *    current device.  Calls are asynchronous on that stream unless stated.
 *  - Physical quantities are SI (metres, 1/metres).
 */
#ifndef RAVESIM_H
#define RAVESIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RSIM_OK = 0,
  RSIM_E_INVALID_ARG = -1, /* -> ValueError */
  RSIM_E_HIP = -2,         /* -> RuntimeError */
  RSIM_E_CANCELLED = -3,   /* -> InterruptedError */
  RSIM_E_NOMEM = -4,       /* -> MemoryError */
  RSIM_E_UNSUPPORTED = -5  /* -> NotImplementedError */
};

enum { RSIM_C64 = 0, RSIM_C128 = 1 };
enum { RSIM_FORWARD = -1, RSIM_INVERSE = 1 };
/* multislice engines: PLAIN = unfused kernels in the field precision;
 * FUSED = complex64 fused row/column kernels with FP64 phase tables. */
enum { RSIM_PLAIN = 0, RSIM_FUSED = 1 };
/* rsim_sample_apply_2d flags */
enum {
  RSIM_FLAG_NO_SPECTRUM = 1, /* do not materialise U = FFT2(u) */
  /* FUSED butterfly arithmetic.  RSIM_FLAG_MATH_FP32: FP32 with norm-matched
   * inverse rotation constants.  RSIM_FLAG_MATH_COMP32: FP32 data, hi+lo
   * compensated constants and twiddles (16 % slower; its error grows as the
   * square root of the slice count instead of linearly).
   * RSIM_FLAG_MATH_FP64: FP64 butterflies.  Without any of them: FP32 up
   * to the depth at which its measured detector error (per slice, per
   * transform length) reaches half of 1e-5 -- 320 slices at 4096^2, 113 at
   * 512^2 -- and COMP32 for deeper volumes. */
  RSIM_FLAG_MATH_FP32 = 2,
  RSIM_FLAG_MATH_FP64 = 4,
  RSIM_FLAG_MATH_COMP32 = 8
};

typedef struct rsim_volume rsim_volume;

/* ---- library / memory plumbing (no reference counterpart: device buffers) ---- */
const char* rsim_last_error(void);
const char* rsim_version(void);
int rsim_device_count(int* count);
int rsim_set_device(int device);
int rsim_malloc(void** dptr, size_t bytes);
int rsim_free(void* dptr);
int rsim_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream); /* synchronous */
int rsim_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream); /* synchronous */
int rsim_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int rsim_memset(void* dptr, int value, size_t bytes, void* stream);
int rsim_synchronize(void* stream);
int rsim_stream_create(void** stream);
int rsim_stream_destroy(void* stream);
/* Release cached FFT plans, tables and workspaces of the current device.
 * Batches created before the call return RSIM_E_INVALID_ARG from
 * rsim_batch_run afterwards (their graphs point into the freed tables). */
int rsim_release_caches(void);
/* Bytes and entries held by the bounded device table caches (axis maps,
 * chirp tables and cone-beam runs' per-slice table sets of every device; at
 * most 256 MiB per device, least recently used entries that no batch, graph
 * or running call pins are freed first). */
int rsim_cache_info(size_t* bytes, int* entries);

/* Cooperative cancel + progress (big-wave/propagation.py:669-688 detector
 * tokens, big-fourier/src/lib.rs:132-187 bfpy's progress callback and cancel
 * token, whose behaviour this follows): a slice loop polls for cancel at every
 * slice boundary and returns RSIM_E_CANCELLED with the field valid after the
 * slices it finished (U then does not hold FFT2(u)); progress is reported as
 * (stage, completed, total) after a slice has finished on the GPU.
 *
 * Per run: an rsim_control bound to the calling thread with
 * rsim_control_bind serves every library call that thread makes until it is
 * unbound (bind NULL) or destroyed; runs on other threads are not affected.
 * rsim_control_request_cancel may be called from any thread.  A progress
 * callback is called on the bound thread after every `every`-th slice and
 * after the last; each report waits for the GPU (hipStreamSynchronize), so a
 * large `every` keeps the slice loop asynchronous.  rsim_control_completed
 * returns the slices the last slice loop under the control finished, and its
 * length.  Threads with no bound control use the process-wide flag and
 * callback below (rsim_set_progress reports every slice). */
typedef void (*rsim_progress_fn)(const char* stage, int64_t completed, int64_t total, void* user);
typedef struct rsim_control rsim_control;
int rsim_control_create(rsim_control** ctl);
int rsim_control_destroy(rsim_control* ctl);
int rsim_control_set_progress(rsim_control* ctl, rsim_progress_fn fn, void* user, int every);
int rsim_control_request_cancel(rsim_control* ctl);
int rsim_control_clear_cancel(rsim_control* ctl);
int rsim_control_bind(rsim_control* ctl);
int rsim_control_completed(const rsim_control* ctl, int64_t* completed, int64_t* total);
int rsim_set_progress(rsim_progress_fn fn, void* user);
int rsim_request_cancel(void);
int rsim_clear_cancel(void);

/* ---- Captured pipelines (no reference counterpart: fast-wave launches every
 * kernel from the host, with a host synchronisation per slice,
 * fast-wave/fwcuda/kernels.cu:365) ----
 * rsim_graph_begin starts recording the calls this thread makes on `stream`
 * (NULL: the library's stream; calls made with a NULL stream during the
 * capture are recorded on the capturing stream); rsim_graph_end turns them
 * into one graph, replayed by rsim_graph_launch with a single launch.  The
 * recorded calls must be stream-ordered device work with device pointers
 * (no host outputs, no rsim_memset -- use rsim_fill_zero) and must have been
 * made once before with the same arguments, so that every table, detector
 * plan and launch-shape choice they use exists; a call that would build one
 * fails with RSIM_E_UNSUPPORTED and rsim_graph_end then fails too.  Buffers
 * must stay allocated while the graph lives; scratch the graph needs is owned
 * by it.  rsim_release_caches invalidates graphs (rsim_graph_launch then
 * returns RSIM_E_INVALID_ARG).  Progress callbacks and cancel do not apply to
 * replays. */
typedef struct rsim_graph rsim_graph;
int rsim_graph_begin(void* stream);
int rsim_graph_end(void* stream, rsim_graph** graph);
int rsim_graph_launch(rsim_graph* graph, void* stream);
int rsim_graph_info(const rsim_graph* graph, size_t* nodes);
int rsim_graph_destroy(rsim_graph* graph);
/* Zero `bytes` (a multiple of 8, 16-byte aligned) of device memory with a
 * kernel, on `stream` (usable inside a capture). */
int rsim_fill_zero(void* dptr, size_t bytes, void* stream);
/* Device-to-device copy by a plain HIP kernel (one contiguous 128 KiB tile
 * per workgroup, 8-byte accesses; 8-byte aligned buffers and size): the
 * achievable-bandwidth ceiling the benchmark reports beside the spec peak and
 * hipMemcpy. */
int rsim_stream_copy(void* dst, const void* src, size_t bytes, void* stream);

/* Kernel timing (the reference's CudaTimerSection, fast-wave/src/simulation.cpp:14-33):
 * every > 0 brackets one launch in `every` of each kind (row pass -- of the
 * slice loops, the two-transform MID passes only --, column pass, pointwise
 * launch of the FFT and multislice engines) with hipEvents on
 * its own stream (every = 1: all launches); 0 disables.  Each event pair adds
 * a few microseconds of dispatch gap, hence the sampling.
 * rsim_kernel_timing_read waits for them and returns, per kind
 * (0 = row pass, 1 = column pass, 2 = pointwise), the summed kernel time in
 * ms and the number of timed launches since the last read, then resets. */
int rsim_kernel_timing(int every);
int rsim_kernel_timing_read(double* ms3, int* launches3);

/* ---- FFT ----
 * Replaces NumpyVector.fft2 / ifft2 (big-wave/vector.py:341-347) and fast-wave
 * FFT<S>::forward / inverse + scale_kernel (fast-wave/src/fft.cpp:35-57,
 * fast-wave/fwcuda/kernels.cu:861-872).  direction RSIM_FORWARD is
 * unnormalised, RSIM_INVERSE multiplies by 1/(nx*ny).  src == dst allowed.
 * Sizes: any nx, ny whose prime factors are 2, 3, 5. */
int rsim_fft2(void* dst, const void* src, int nx, int ny, int dtype, int direction, void* stream);
/* Batched 1D transforms along the contiguous axis: `batch` rows of length n. */
int rsim_fft1d_rows(void* dst, const void* src, int n, int batch, int dtype, int direction,
                    void* stream);

/* ---- Fresnel propagation ---- */
/* Replaces apply_frequency_cutoff_2d (big-wave/propagation.py:298-326):
 * zero where kx^2 + ky^2 > 2 f^2, in place on a spectrum. */
int rsim_apply_frequency_cutoff_2d(void* U, int nx, int ny, int dtype, double freq, double dx,
                                   double dy, void* stream);
/* Replaces propagate_2d (big-wave/propagation.py:396-442; fast-wave
 * simulation.cpp:472-480 + kernels.cu:166-212): U = FFT2(u) unless skip_fft,
 * U *= exp(2 pi i dz/wl) exp(-i pi wl dz (kx^2+ky^2)), circular cutoff,
 * u = IFFT2(U).  On return U holds FFT2(u). */
int rsim_propagate_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                      double dz, double cutoff_freq, int skip_fft, void* stream);
/* Replaces propagate_analytically_2d (big-wave/propagation.py:458-484;
 * fast-wave kernels.cu:107-153): u = exp(2 pi i r/wl)/r. */
int rsim_propagate_analytically_2d(void* u, int nx, int ny, int dtype, double dx, double dy,
                                   double z, double x_source, double y_source, double wl,
                                   void* stream);

/* ---- Sample volumes ----
 * Replaces the device upload of fast-wave apply_sample_2d
 * (fast-wave/src/simulation.cpp:651-665) and Sample.store_deltabetas
 * (big-wave/optical_element.py:433-440).  ids: host array (nz, ny_g, nx_g),
 * C order, id_bytes = 1, 2 or 4 (uint8/uint16/uint32 or int32); every id must
 * be < n_db.  deltabeta: n_db complex values as (re, im) pairs, entry 0 is
 * vacuum and must be 0.  The device copy is uint8. */
int rsim_volume_create(rsim_volume** vol, const void* ids, int id_bytes, int nz, int ny_g,
                       int nx_g, const double* deltabeta, int n_db);
/* Same, allocating an all-vacuum volume to be filled on the device. */
int rsim_volume_create_empty(rsim_volume** vol, int nz, int ny_g, int nx_g,
                             const double* deltabeta, int n_db);
int rsim_volume_destroy(rsim_volume* vol);
/* Device pointer of the uint8 ids and the volume shape. */
int rsim_volume_info(const rsim_volume* vol, void** dev_ids, int* nz, int* ny_g, int* nx_g);
/* Paint spheres into the volume in order (later spheres overwrite earlier
 * ones).  spheres: n x (cz, cy, cx, radius) in voxel units; mats: n ids. */
int rsim_volume_paint_spheres(rsim_volume* vol, const double* spheres, const int* mats, int n,
                              void* stream);

/* Replaces the transmission step of Sample._apply_2d
 * (big-wave/optical_element.py:533-584; fast-wave kernels.cu:215-407):
 * u *= exp(-2 pi i pz/wl conj(db)) with db bilinear in delta-beta space from
 * the four voxel corners of slice `slice`, vacuum outside the voxel box. */
int rsim_transmission_2d(void* u, int nx, int ny, int dtype, double dx, double dy, double wl,
                         const rsim_volume* vol, int slice, double pixel_size_x,
                         double pixel_size_y, double pixel_size_z, double x_pos, double y_pos,
                         void* stream);

/* Replaces Sample._apply_2d (big-wave/optical_element.py:502-596) and fast-wave
 * apply_sample_2d (fast-wave/src/simulation.cpp:598-733): for each slice in
 * [first_slice, first_slice + n_slices): transmission, then propagate_2d over
 * pixel_size_z.  engine RSIM_PLAIN or RSIM_FUSED (FUSED needs RSIM_C64; other
 * sizes fall back to PLAIN kernels).  Unless flags has RSIM_FLAG_NO_SPECTRUM,
 * U holds FFT2(u) on return (OpticalElement.apply contract,
 * big-wave/optical_element.py:97).  Polls the cancel flag between slices. */
int rsim_sample_apply_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                         double wl, double cutoff_freq, const rsim_volume* vol,
                         double pixel_size_x, double pixel_size_y, double pixel_size_z,
                         double x_pos, double y_pos, int first_slice, int n_slices, int engine,
                         int flags, void* stream);

/* One step of a propagation chain: n_slices slices of vol (each a
 * transmission, then propagation over pixel_size_z, as rsim_sample_apply_2d),
 * or, with vol NULL, free space over dz (as rsim_propagate_2d). */
typedef struct rsim_chain_step {
  const rsim_volume* vol;
  double pixel_size_x, pixel_size_y, pixel_size_z;
  double x_pos, y_pos;
  int first_slice, n_slices;
  double dz;
} rsim_chain_step;

/* Replaces the element loop of big-wave's run_simulation for one source, at
 * one cutoff frequency for every step (big-wave/wavesim.py:199-300: the gap
 * to each element, its apply, the last leg, each its own FFT2 ... IFFT2):
 * with band_limit set, first U = FFT2(u), the circular cutoff, u = IFFT2(U)
 * (the point source's step, big-wave/source.py:97-107); then the steps in
 * order.  With engine RSIM_FUSED and RSIM_C64 at a fused size (columns of up
 * to 8192) the chain is one sequence of fused row and column passes: the
 * propagations with no transmission between them (band limit, free space,
 * the slice before it) fold into one column pass over their summed distance,
 * P(a) P(b) = P(a + b) with one cutoff.  Otherwise every step runs as its own
 * call.  Unless flags has RSIM_FLAG_NO_SPECTRUM, U holds FFT2(u) on return.
 * The cancel flag is polled once, before the chain starts. */
int rsim_chain_apply_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                        double wl, double cutoff_freq, int band_limit,
                        const rsim_chain_step* steps, int n_steps, int engine, int flags,
                        void* stream);

/* One instance of a batched chain: its fields, pixel pitch, wavelength,
 * cutoff and its own n_steps steps. */
typedef struct rsim_chain_instance {
  void* u;
  void* U;
  double dx, dy, wl, cutoff_freq;
  const rsim_chain_step* steps;
} rsim_chain_instance;

/* rsim_chain_apply_2d for count independent instances of one nx x ny grid
 * (the multi-source runner's instances, big-wave/multisim.py:258-519, each the
 * element loop above): every instance has n_steps steps of the same kinds
 * with the same slice counts, and its own volumes, placements, distances,
 * wavelength and cutoff.  On the fused engine each pass of the chain is one
 * launch for up to 16 instances (per transmission mode), and the results
 * equal count calls of rsim_chain_apply_2d bit for bit; otherwise the
 * instances run one after another. */
int rsim_chain_apply_2d_batch(const rsim_chain_instance* instances, int count, int nx, int ny,
                              int dtype, int band_limit, int n_steps, int engine, int flags,
                              void* stream);

/* Fresnel-scaled / cone-beam BPM forms (big-wave/propagation.py:114-196,
 * big-wave/optical_element.py:533-592, big-wave/source.py:62-90).
 * rsim_sample_apply_2d_scaled: slice k samples the voxel grid at field
 * coordinates times coord_scale[k] (the local cone-beam magnification at the
 * slice midpoint) and propagates over slice_dz[k] (effective_slice_dz); the
 * transmission keeps the physical pixel_size_z.  NULL arrays mean 1 and
 * pixel_size_z.  rsim_transmission_2d_scaled: one slice at coord_scale.
 * rsim_tilted_plane_2d: the Fresnel-scaled point source, u = exp(-i k (x xs
 * + y ys) / dz) / dz with k = 2 pi / wl (FFT + cutoff + IFFT follow in the
 * caller, as in the reference). */
int rsim_sample_apply_2d_scaled(void* u, void* U, int nx, int ny, int dtype, double dx,
                                double dy, double wl, double cutoff_freq, const rsim_volume* vol,
                                double pixel_size_x, double pixel_size_y, double pixel_size_z,
                                double x_pos, double y_pos, int first_slice, int n_slices,
                                const double* coord_scale, const double* slice_dz, int engine,
                                int flags, void* stream);
int rsim_transmission_2d_scaled(void* u, int nx, int ny, int dtype, double dx, double dy,
                                double wl, const rsim_volume* vol, int slice, double pixel_size_x,
                                double pixel_size_y, double pixel_size_z, double x_pos,
                                double y_pos, double coord_scale, void* stream);
int rsim_tilted_plane_2d(void* u, int nx, int ny, int dtype, double dx, double dy, double wl,
                         double dz, double x_source, double y_source, void* stream);

/* ---- History frames of a slice loop (big-wave/optical_element.py:548-551,
 * big-wave/plasma_sample.py:172-176; fast-wave keeps its history on the device
 * too, fast-wave/src/simulation.cpp:43-102) ----
 * Frame k of a call over slices [first, first + n) is the area_v1 image
 * (out_y x out_x, float64, physical detector pixels) of the field entering
 * slice first + k, at z = z_start + (first + k) * pixel_size_z, written to
 * frames + k * out_y * out_x (a device buffer of n * out_y * out_x doubles).
 * The frames are stream-ordered with the loop; nothing synchronises. */
typedef struct rsim_history_frames {
  int out_x, out_y;
  double pixel_size_x, pixel_size_y;
  double z_start;
  double* frames;
} rsim_history_frames;
int rsim_sample_apply_2d_history(void* u, void* U, int nx, int ny, int dtype, double dx,
                                 double dy, double wl, double cutoff_freq,
                                 const rsim_volume* vol, double pixel_size_x,
                                 double pixel_size_y, double pixel_size_z, double x_pos,
                                 double y_pos, int first_slice, int n_slices,
                                 const double* coord_scale, const double* slice_dz, int engine,
                                 int flags, const rsim_history_frames* history, void* stream);

/* ---- PlasmaSample (big-wave/plasma_sample.py:126-233, big-wave/plasma.py:380-434) ----
 * A plasma volume: four (nz, ny_g, nx_g) grids -- electron density n_e and ion
 * density n_i [cm^-3], electron temperature T_e [eV] and mean ionisation Z* --
 * of float32 (elem_size 4) or float64 (8), copied to the device.  ``optics``
 * is the per-atom Chantler quadruple (f1bar, f2eff_bar, zbar, z2bar) of the
 * element or compound at ``energy`` [eV] (plasma.py:238-262; the
 * plasma_optics_table of a prepared subconfig).
 * rsim_plasma_deltabeta: (delta, beta) per voxel of slices [first, first + n)
 * as complex128 into the device buffer ``out`` (plasma_delta_beta_grid_multi,
 * plasma.py:380-434, FP64 in its operation order).
 * rsim_plasma_transmission_2d: one slice's transmission, bilinear in (delta,
 * beta), vacuum outside the grid (plasma_sample.py:160-218).
 * rsim_plasma_apply_2d: PlasmaSample._apply_2d over slices [first, first + n)
 * -- transmission then propagate_2d per slice, optional cone-beam scale and
 * distance arrays as in rsim_sample_apply_2d_scaled; engine RSIM_PLAIN or
 * RSIM_FUSED (complex64; the FP32 (delta, beta) volume of each (plasma,
 * energy, optics) is built once and cached until rsim_plasma_destroy). */
typedef struct rsim_plasma rsim_plasma;
int rsim_plasma_create(rsim_plasma** out, const void* ne, const void* ni, const void* te,
                       const void* zstar, int elem_size, int nz, int ny_g, int nx_g);
int rsim_plasma_destroy(rsim_plasma* plasma);
int rsim_plasma_deltabeta(const rsim_plasma* plasma, int first_slice, int n_slices, double energy,
                          const double* optics, void* out, void* stream);
int rsim_plasma_transmission_2d(void* u, int nx, int ny, int dtype, double dx, double dy,
                                double wl, const rsim_plasma* plasma, int slice, double energy,
                                const double* optics, double pixel_size_x, double pixel_size_y,
                                double pixel_size_z, double x_pos, double y_pos,
                                double coord_scale, void* stream);
int rsim_plasma_apply_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                         double wl, double cutoff_freq, const rsim_plasma* plasma, double energy,
                         const double* optics, double pixel_size_x, double pixel_size_y,
                         double pixel_size_z, double x_pos, double y_pos, int first_slice,
                         int n_slices, const double* coord_scale, const double* slice_dz,
                         int engine, int flags, void* stream);
/* rsim_plasma_apply_2d with history frames, as rsim_sample_apply_2d_history */
int rsim_plasma_apply_2d_history(void* u, void* U, int nx, int ny, int dtype, double dx,
                                 double dy, double wl, double cutoff_freq,
                                 const rsim_plasma* plasma, double energy, const double* optics,
                                 double pixel_size_x, double pixel_size_y, double pixel_size_z,
                                 double x_pos, double y_pos, int first_slice, int n_slices,
                                 const double* coord_scale, const double* slice_dz, int engine,
                                 int flags, const rsim_history_frames* history, void* stream);

/* ---- Batched fused multislice (no counterpart in the reference: the
 * reference runs each (source, energy) instance as its own process,
 * fast-wave/Readme.md:27) ----
 * One instance = the arguments of one RSIM_FUSED complex64
 * rsim_sample_apply_2d call.  rsim_batch_create groups the instances by
 * (nx, ny) into launches of up to 16 fields (grid.y = instance), builds every
 * table and autotunes each shape, then captures the whole slice loop of all
 * instances (rows, columns and the final row pass of each slice) as one
 * hipGraph.  rsim_batch_run replays it with a single graph launch: every
 * instance's u goes through its n_slices slices and, unless
 * RSIM_FLAG_NO_SPECTRUM, U = FFT2(u) on return.  Fields are read from and
 * written to the instances' own u / U buffers, which must stay allocated for
 * the batch's life.  Sizes: the fused engine's (ny <= 8192).  flags: the
 * arithmetic flags of rsim_sample_apply_2d (the same depth-dependent default).
 * The cached tables the graph reads are pinned until rsim_batch_destroy;
 * rsim_release_caches invalidates existing batches (rsim_batch_run then
 * returns RSIM_E_INVALID_ARG).  Volumes must outlive the batches using them. */
typedef struct rsim_batch rsim_batch;
typedef struct {
  void* u;
  void* U;
  int nx, ny;
  double dx, dy, wl, cutoff_freq;
  const rsim_volume* vol;
  double pixel_size_x, pixel_size_y, pixel_size_z, x_pos, y_pos;
} rsim_batch_instance;
int rsim_batch_create(rsim_batch** batch, const rsim_batch_instance* instances, int count,
                      int first_slice, int n_slices, int flags, void* stream);
int rsim_batch_run(rsim_batch* batch, void* stream);
/* Number of shape groups and of kernel nodes in the captured graph. */
int rsim_batch_info(const rsim_batch* batch, int* groups, int* kernels);
int rsim_batch_destroy(rsim_batch* batch);

/* Launch-shape autotuner of the fused engine (no counterpart in the
 * reference): the first RSIM_FUSED call for an (nx, ny) times each row-pass
 * and column-pass variant on a scratch field, then the leading row x column
 * pairs as a slice runs them, both in place and with the column pass storing
 * column tiles that the next row pass reads (two buffers, u and U), and
 * keeps the fastest (the tiled form unless an in-place pair is more than 1 %
 * faster: it moves fewer HBM bytes); the environment variable
 * RAVESIM_FUSED_VARIANT="row,col[,tiled]" overrides it.  Returns the variants
 * of the voxel slice loop for (nx, ny) (material-id lookup, FP32 butterflies,
 * when it has run; else the first tuned form), or -1 before the first fused
 * call. */
int rsim_fused_schedule(int nx, int ny, int* row_variant, int* col_variant);
/* Width of the column tiles between the passes of the voxel slice loop for
 * (nx, ny): 0 = in place, row-major; -1 before the first fused call (no
 * counterpart in the reference). */
int rsim_fused_tile(int nx, int ny, int* tile_width);

/* ---- Detector ----
 * Replaces square_and_downsample_2d with the area_v1 integrator
 * (big-wave/propagation.py:995-1130; fast-wave kernels.cu:525-623):
 * out[q,p] = sum |u|^2 overlap_x overlap_y z/sqrt(x_p^2+y_q^2+z^2), FP64.
 * out_host: out_y*out_x float64, written before return (synchronous). */
int rsim_detector_area_v1(const void* u, int nx, int ny, int dtype, double dx, double dy,
                          int out_x, int out_y, double ds_x, double ds_y, double z,
                          double* out_host, void* stream);
/* Same with a device output buffer (float64) and accumulation weight:
 * out_dev += weight * detector (for spectral sums). */
int rsim_detector_area_v1_accumulate(const void* u, int nx, int ny, int dtype, double dx, double dy,
                                     int out_x, int out_y, double ds_x, double ds_y, double z,
                                     double weight, double* out_dev, void* stream);
/* rsim_detector_area_v1_accumulate without device allocations -- no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (serving loops, harnesses that forbid allocation).  The detector
 * plan and the x-integrals (out_y-covering field rows x out_x doubles) live
 * in one static device array of rsim_detector_noalloc_capacity() bytes
 * (64 MiB) per GPU, held until the process ends; larger detectors return
 * RSIM_E_INVALID_ARG.  Each call uploads its plan and waits for the stream
 * before launching (the result stays stream-ordered); every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU race.  Same
 * results as rsim_detector_area_v1_accumulate, bit for bit. */
int rsim_detector_area_v1_accumulate_noalloc(const void* u, int nx, int ny, int dtype, double dx,
                                             double dy, int out_x, int out_y, double ds_x,
                                             double ds_y, double z, double weight,
                                             double* out_dev, void* stream);
int rsim_detector_noalloc_capacity(size_t* bytes);

/* ---- Differentiable multislice (per-voxel delta/beta volumes) ----
 * The reference has no adjoint; these entry points are the gradient path of
 * the same physics.  Volumes are device float32 arrays (n_slices, ny, nx),
 * C-order, one voxel per field pixel: the direct per-voxel form of
 * fast-wave's complex delta-beta grid (fast-wave/fwcuda/kernels.cu:411-491).
 * Slice s multiplies u by material_factor(delta_s + i beta_s, pz, wl)
 * = exp(-2 pi pz/wl (beta_s + i delta_s)) (big-wave/optical_element.py:39-42)
 * and then runs propagate_2d over pz (big-wave/propagation.py:396-442). */

/* One transmission slice (the voxel multiply of Sample._apply_2d,
 * big-wave/optical_element.py:580-582, with per-voxel delta/beta). */
int rsim_transmission_db_2d(void* u, int nx, int ny, int dtype, double wl, double pz,
                            const float* delta, const float* beta, void* stream);
/* Forward multislice over n_slices.  U is scratch (holds FFT2(u) on return).
 * ckpt (optional): receives u before slice k for every k % ckpt_every == 0,
 * ceil(n_slices / ckpt_every) fields.  engine RSIM_PLAIN, or RSIM_FUSED: the
 * fused complex64 chain (row lengths 256..16384, column lengths 256..8192 of
 * the fused set; other shapes and complex128 run plain), whose butterfly
 * arithmetic flags selects (one RSIM_FLAG_MATH_*; none = FP32).  The adjoint
 * must use the engine and flags that wrote the checkpoints. */
int rsim_multislice_db_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                          double wl, double cutoff_freq, const float* delta, const float* beta,
                          double pz, int n_slices, void* ckpt, int ckpt_every, int engine,
                          int flags, void* stream);
/* Bytes of the checkpoint and work buffers for (n_slices, ckpt_every);
 * ckpt_every = ceil(sqrt(n_slices)) minimises their sum.  The work buffer
 * holds the rebuilt fields of one segment (ckpt_every - 1 fields), two for
 * complex64 with more than one segment: the fused engine rebuilds the next
 * segment while the adjoint runs through the current one. */
int rsim_multislice_adjoint_workspace(int nx, int ny, int dtype, int n_slices, int ckpt_every,
                                      size_t* ckpt_bytes, size_t* work_bytes);
/* Adjoint multislice.  On entry g = dL/du* after the last slice, on return
 * dL/du* before the first.  grad_delta / grad_beta (n_slices, ny, nx) receive
 * dL/d delta and dL/d beta.  ckpt as written by rsim_multislice_db_2d with the
 * same ckpt_every, engine and flags; work as sized by
 * rsim_multislice_adjoint_workspace; U is scratch.  The fused engine runs the
 * segment rebuilds on a second library stream ordered with `stream` by
 * events; all of its work is ordered before later work on `stream`. */
int rsim_multislice_db_adjoint_2d(void* g, void* U, int nx, int ny, int dtype, double dx,
                                  double dy, double wl, double cutoff_freq, const float* delta,
                                  const float* beta, double pz, int n_slices, const void* ckpt,
                                  int ckpt_every, void* work, float* grad_delta,
                                  float* grad_beta, int engine, int flags, void* stream);
/* Adjoint of rsim_detector_area_v1 (big-wave/propagation.py:995-1130): given
 * grad_img = dL/dI (out_y, out_x) float64 on the device, writes
 * g = dL/dv* = v * A^T grad_img, where I = A |v|^2. */
int rsim_detector_area_v1_adjoint(const void* v, void* g, int nx, int ny, int dtype, double dx,
                                  double dy, int out_x, int out_y, double ds_x, double ds_y,
                                  double z, const double* grad_img, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RAVESIM_H */
