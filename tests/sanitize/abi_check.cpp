// Host-sanitizer check of the C ABI.
//
// Built with -fsanitize=address,undefined against libravesim_hip_san.so (the
// library's host code instrumented the same way, device code as usual), so
// that heap misuse, out-of-bounds host accesses, use after free and undefined
// behaviour in the library's host layer abort the run with a report.
//
//   abi_check cpu   argument validation, error strings, controls, caches:
//                   every call returns before touching a device
//   abi_check gpu   the same, then real work on device 0: volumes, the fused
//                   and plain slice loops, the cone-beam tables built by host
//                   threads, chains, batches, captured graphs, per-run cancel
//                   and progress, the detectors (allocating and static
//                   scratch), the adjoint with its checkpoint workspace,
//                   cache release invalidating batches and graphs; then FFTs,
//                   the point source, per-voxel and plasma transmissions,
//                   history frames in both engines, the detector adjoint and
//                   batched chains
//
// Prints "abi_check <mode>: OK" and exits 0 when every expectation holds.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ravesim.h"

namespace {

int g_fail = 0;

void expect(bool ok, const char* what, int line) {
  if (!ok) {
    std::fprintf(stderr, "abi_check: FAILED at line %d: %s (last error: %s)\n", line, what,
                 rsim_last_error());
    ++g_fail;
  }
}
#define EXPECT(c) expect((c), #c, __LINE__)
#define EXPECT_RC(call, rc) expect((call) == (rc), #call " == " #rc, __LINE__)
#define STEP(name) std::fprintf(stderr, "abi_check: %s\n", name)

const double kDb[8] = {0, 0, 3.5e-7, 2e-10, 1.1e-6, 3e-9, 2.6e-7, 1.5e-10};

void check_arguments() {
  STEP("argument checks");
  EXPECT(std::strstr(rsim_version(), "gfx950") != nullptr);
  EXPECT_RC(rsim_fft2(nullptr, nullptr, 4, 4, RSIM_C64, RSIM_FORWARD, nullptr), RSIM_E_INVALID_ARG);
  EXPECT(std::strlen(rsim_last_error()) > 0);
  void* p = reinterpret_cast<void*>(16);
  EXPECT_RC(rsim_fft2(p, p, 7, 4, RSIM_C64, RSIM_FORWARD, nullptr), RSIM_E_INVALID_ARG);
  EXPECT_RC(rsim_fft2(p, p, 4, 4, 9, RSIM_FORWARD, nullptr), RSIM_E_INVALID_ARG);
  EXPECT_RC(rsim_propagate_2d(p, p, 4, 4, RSIM_C128, NAN, 1e-6, 1e-10, 0.1, 1e5, 0, nullptr),
            RSIM_E_INVALID_ARG);
  // vacuum entry must be zero; ids must index the table
  rsim_volume* vol = nullptr;
  uint8_t ids[8] = {0, 1, 2, 3, 0, 1, 2, 3};
  const double bad_db[4] = {1e-7, 0, 3.5e-7, 2e-10};
  EXPECT_RC(rsim_volume_create(&vol, ids, 1, 2, 2, 2, bad_db, 2), RSIM_E_INVALID_ARG);
  EXPECT_RC(rsim_volume_create(&vol, ids, 1, 2, 2, 2, kDb, 2), RSIM_E_INVALID_ARG);
  EXPECT_RC(rsim_volume_create(&vol, ids, 3, 2, 2, 2, kDb, 4), RSIM_E_INVALID_ARG);
  EXPECT(vol == nullptr);
  EXPECT_RC(rsim_kernel_timing(-1), RSIM_E_INVALID_ARG);
  rsim_batch* b = nullptr;
  EXPECT_RC(rsim_batch_create(&b, nullptr, 0, 0, 1, 0, nullptr), RSIM_E_INVALID_ARG);
  EXPECT_RC(rsim_batch_destroy(nullptr), RSIM_OK);
  rsim_plasma* pl = nullptr;
  EXPECT_RC(rsim_plasma_create(&pl, p, p, p, p, 2, 1, 2, 4), RSIM_E_INVALID_ARG);
  EXPECT_RC(rsim_plasma_destroy(nullptr), RSIM_OK);
  EXPECT_RC(rsim_tilted_plane_2d(p, 4, 4, RSIM_C64, 1e-6, 1e-6, 1e-10, 0.0, 0.0, 0.0, nullptr),
            RSIM_E_INVALID_ARG);
  EXPECT_RC(rsim_chain_apply_2d(p, p, 4, 4, RSIM_C64, 1e-6, 1e-6, 1e-10, 1e5, 1, nullptr, 1,
                                RSIM_FUSED, 0, nullptr),
            RSIM_E_INVALID_ARG);
  size_t cap = 0;
  EXPECT_RC(rsim_detector_noalloc_capacity(&cap), RSIM_OK);
  EXPECT(cap == (size_t)64 << 20);
  size_t bytes = 1;
  int entries = -1;
  EXPECT_RC(rsim_cache_info(&bytes, &entries), RSIM_OK);
  size_t ck = 0, wk = 0;
  EXPECT_RC(rsim_multislice_adjoint_workspace(256, 256, RSIM_C64, 16, 4, &ck, &wk), RSIM_OK);
  EXPECT(ck == 4 * (size_t)256 * 256 * 8);
}

struct ProgressLog {
  int64_t last = -1, total = -1;
  int calls = 0;
  rsim_control* cancel_ctl = nullptr;
  int64_t cancel_at = -1;
};

void on_progress(const char*, int64_t done, int64_t total, void* user) {
  auto* log = static_cast<ProgressLog*>(user);
  log->last = done;
  log->total = total;
  ++log->calls;
  if (log->cancel_ctl && done == log->cancel_at) rsim_control_request_cancel(log->cancel_ctl);
}

void check_controls() {
  STEP("controls");
  rsim_control* ctl = nullptr;
  EXPECT_RC(rsim_control_create(&ctl), RSIM_OK);
  ProgressLog log;
  EXPECT_RC(rsim_control_set_progress(ctl, on_progress, &log, 4), RSIM_OK);
  EXPECT_RC(rsim_control_request_cancel(ctl), RSIM_OK);
  EXPECT_RC(rsim_control_clear_cancel(ctl), RSIM_OK);
  EXPECT_RC(rsim_control_bind(ctl), RSIM_OK);
  EXPECT_RC(rsim_control_bind(nullptr), RSIM_OK);
  int64_t c = -1, t = -1;
  EXPECT_RC(rsim_control_completed(ctl, &c, &t), RSIM_OK);
  EXPECT_RC(rsim_control_destroy(ctl), RSIM_OK);
  EXPECT_RC(rsim_set_progress(nullptr, nullptr), RSIM_OK);
  EXPECT_RC(rsim_request_cancel(), RSIM_OK);
  EXPECT_RC(rsim_clear_cancel(), RSIM_OK);
}

double rel_l2(const std::vector<std::complex<float>>& a, const std::vector<std::complex<float>>& b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += std::norm(std::complex<double>(a[i]) - std::complex<double>(b[i]));
    den += std::norm(std::complex<double>(b[i]));
  }
  return std::sqrt(num / den);
}

struct Field {
  void* d = nullptr;
  size_t bytes = 0;
  explicit Field(size_t n) : bytes(n * 8) { EXPECT_RC(rsim_malloc(&d, bytes), RSIM_OK); }
  ~Field() { rsim_free(d); }
  void set(const std::vector<std::complex<float>>& h) {
    EXPECT_RC(rsim_memcpy_h2d(d, h.data(), bytes, nullptr), RSIM_OK);
  }
  std::vector<std::complex<float>> get() const {
    std::vector<std::complex<float>> h(bytes / 8);
    EXPECT_RC(rsim_memcpy_d2h(h.data(), d, bytes, nullptr), RSIM_OK);
    return h;
  }
};

void check_device() {
  const int n = 256, nz = 6;
  const double dx = 0.5e-6, wl = 0.5e-10, pz = 0.5e-6, cut = 0.4 / dx;
  const size_t pts = (size_t)n * n;
  std::vector<uint8_t> grid((size_t)nz * pts, 0);
  for (int k = 0; k < nz; ++k)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int dy = y - 100 - 5 * k, dxp = x - 140;
        if (dy * dy + dxp * dxp < 40 * 40) grid[(size_t)k * pts + (size_t)y * n + x] = 1 + k % 3;
      }
  std::vector<std::complex<float>> u0(pts);
  for (size_t i = 0; i < pts; ++i) u0[i] = std::polar(1.0f, 0.05f * (float)((i * 2654435761u) % 97) / 97.f);

  STEP("volume and slice loops");
  rsim_volume* vol = nullptr;
  EXPECT_RC(rsim_volume_create(&vol, grid.data(), 1, nz, n, n, kDb, 4), RSIM_OK);
  Field u(pts), U(pts), u2(pts), U2(pts);
  u.set(u0);
  u2.set(u0);
  EXPECT_RC(rsim_sample_apply_2d(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, vol, dx, dx, pz, 0.0,
                                 0.0, 0, nz, RSIM_FUSED, 0, nullptr),
            RSIM_OK);
  EXPECT_RC(rsim_sample_apply_2d(u2.d, U2.d, n, n, RSIM_C64, dx, dx, wl, cut, vol, dx, dx, pz, 0.0,
                                 0.0, 0, nz, RSIM_PLAIN, 0, nullptr),
            RSIM_OK);
  const auto fused = u.get(), plain = u2.get();
  EXPECT(rel_l2(fused, plain) < 1e-5);

  STEP("cone-beam tables");
  std::vector<double> scale(nz), dz(nz);
  for (int k = 0; k < nz; ++k) {
    scale[k] = 1.0 + 1e-3 * k;
    dz[k] = pz / (scale[k] * scale[k]);
  }
  for (int rep = 0; rep < 2; ++rep) {   // the second call reuses the cached table set
    u.set(u0);
    EXPECT_RC(rsim_sample_apply_2d_scaled(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, vol, dx, dx,
                                          pz, 0.0, 0.0, 0, nz, scale.data(), dz.data(),
                                          RSIM_FUSED, 0, nullptr),
              RSIM_OK);
  }

  STEP("chain");
  rsim_chain_step steps[3] = {};
  steps[0].vol = vol;
  steps[0].pixel_size_x = steps[0].pixel_size_y = dx;
  steps[0].pixel_size_z = pz;
  steps[0].n_slices = 3;
  steps[1].vol = nullptr;
  steps[1].dz = 0.01;
  steps[2] = steps[0];
  steps[2].first_slice = 3;
  u.set(u0);
  EXPECT_RC(rsim_chain_apply_2d(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, 1, steps, 3, RSIM_FUSED,
                                0, nullptr),
            RSIM_OK);

  STEP("per-run cancel and progress");
  rsim_control* ctl = nullptr;
  EXPECT_RC(rsim_control_create(&ctl), RSIM_OK);
  ProgressLog log;
  log.cancel_ctl = ctl;
  log.cancel_at = 2;
  EXPECT_RC(rsim_control_set_progress(ctl, on_progress, &log, 1), RSIM_OK);
  EXPECT_RC(rsim_control_bind(ctl), RSIM_OK);
  u.set(u0);
  EXPECT_RC(rsim_sample_apply_2d(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, vol, dx, dx, pz, 0.0,
                                 0.0, 0, nz, RSIM_FUSED, 0, nullptr),
            RSIM_E_CANCELLED);
  int64_t done = -1, total = -1;
  EXPECT_RC(rsim_control_completed(ctl, &done, &total), RSIM_OK);
  EXPECT(done == 2 && total == nz);
  EXPECT_RC(rsim_control_bind(nullptr), RSIM_OK);
  EXPECT_RC(rsim_control_destroy(ctl), RSIM_OK);

  STEP("detectors");
  const int ox = n / 4, oy = n / 4;
  std::vector<double> img((size_t)ox * oy), acc((size_t)ox * oy), acc2((size_t)ox * oy);
  EXPECT_RC(rsim_detector_area_v1(u2.d, n, n, RSIM_C64, dx, dx, ox, oy, 4 * dx, 4 * dx, 0.5,
                                  img.data(), nullptr),
            RSIM_OK);
  void* dacc = nullptr;
  void* dacc2 = nullptr;
  EXPECT_RC(rsim_malloc(&dacc, acc.size() * 8), RSIM_OK);
  EXPECT_RC(rsim_malloc(&dacc2, acc.size() * 8), RSIM_OK);
  EXPECT_RC(rsim_fill_zero(dacc, acc.size() * 8, nullptr), RSIM_OK);
  EXPECT_RC(rsim_fill_zero(dacc2, acc.size() * 8, nullptr), RSIM_OK);
  EXPECT_RC(rsim_detector_area_v1_accumulate(u2.d, n, n, RSIM_C64, dx, dx, ox, oy, 4 * dx, 4 * dx,
                                             0.5, 0.25, static_cast<double*>(dacc), nullptr),
            RSIM_OK);
  EXPECT_RC(rsim_detector_area_v1_accumulate_noalloc(u2.d, n, n, RSIM_C64, dx, dx, ox, oy, 4 * dx,
                                                     4 * dx, 0.5, 0.25,
                                                     static_cast<double*>(dacc2), nullptr),
            RSIM_OK);
  EXPECT_RC(rsim_memcpy_d2h(acc.data(), dacc, acc.size() * 8, nullptr), RSIM_OK);
  EXPECT_RC(rsim_memcpy_d2h(acc2.data(), dacc2, acc2.size() * 8, nullptr), RSIM_OK);
  EXPECT(std::memcmp(acc.data(), acc2.data(), acc.size() * 8) == 0);
  double worst = 0;
  for (size_t i = 0; i < img.size(); ++i) worst = std::fmax(worst, std::fabs(acc[i] - 0.25 * img[i]));
  EXPECT(worst <= 1e-12 * 0.25 * (1 + std::fabs(img[0])));
  rsim_free(dacc);
  rsim_free(dacc2);

  STEP("captured graph");
  void* s = nullptr;
  EXPECT_RC(rsim_stream_create(&s), RSIM_OK);
  u.set(u0);
  EXPECT_RC(rsim_sample_apply_2d(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, vol, dx, dx, pz, 0.0,
                                 0.0, 0, nz, RSIM_FUSED, 0, s),
            RSIM_OK);
  EXPECT_RC(rsim_synchronize(s), RSIM_OK);
  EXPECT_RC(rsim_graph_begin(s), RSIM_OK);
  EXPECT_RC(rsim_sample_apply_2d(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, vol, dx, dx, pz, 0.0,
                                 0.0, 0, nz, RSIM_FUSED, 0, s),
            RSIM_OK);
  rsim_graph* graph = nullptr;
  EXPECT_RC(rsim_graph_end(s, &graph), RSIM_OK);
  u.set(u0);
  EXPECT_RC(rsim_graph_launch(graph, s), RSIM_OK);
  EXPECT_RC(rsim_synchronize(s), RSIM_OK);
  EXPECT(std::memcmp(u.get().data(), fused.data(), pts * 8) == 0);

  STEP("batch");
  Field ub(2 * pts), Ub(2 * pts);
  rsim_batch_instance inst[2] = {};
  for (int i = 0; i < 2; ++i) {
    inst[i].u = static_cast<char*>(ub.d) + i * pts * 8;
    inst[i].U = static_cast<char*>(Ub.d) + i * pts * 8;
    inst[i].nx = inst[i].ny = n;
    inst[i].dx = inst[i].dy = dx;
    inst[i].wl = wl * (1 + 0.1 * i);
    inst[i].cutoff_freq = cut;
    inst[i].vol = vol;
    inst[i].pixel_size_x = inst[i].pixel_size_y = dx;
    inst[i].pixel_size_z = pz;
  }
  rsim_batch* batch = nullptr;
  EXPECT_RC(rsim_batch_create(&batch, inst, 2, 0, nz, 0, s), RSIM_OK);
  EXPECT_RC(rsim_batch_run(batch, s), RSIM_OK);
  EXPECT_RC(rsim_synchronize(s), RSIM_OK);
  int groups = 0, kernels = 0;
  EXPECT_RC(rsim_batch_info(batch, &groups, &kernels), RSIM_OK);
  EXPECT(groups == 1 && kernels > 0);

  STEP("adjoint workspace");
  const int ns = 8, every = 3;
  size_t ckb = 0, wkb = 0;
  EXPECT_RC(rsim_multislice_adjoint_workspace(n, n, RSIM_C64, ns, every, &ckb, &wkb), RSIM_OK);
  void *dl = nullptr, *bt = nullptr, *ckp = nullptr, *wk = nullptr, *gd = nullptr, *gb = nullptr;
  const size_t vb = (size_t)ns * pts * 4;
  EXPECT_RC(rsim_malloc(&dl, vb), RSIM_OK);
  EXPECT_RC(rsim_malloc(&bt, vb), RSIM_OK);
  EXPECT_RC(rsim_malloc(&gd, vb), RSIM_OK);
  EXPECT_RC(rsim_malloc(&gb, vb), RSIM_OK);
  EXPECT_RC(rsim_malloc(&ckp, ckb), RSIM_OK);
  EXPECT_RC(rsim_malloc(&wk, wkb), RSIM_OK);
  std::vector<float> hd((size_t)ns * pts), hb((size_t)ns * pts);
  for (size_t i = 0; i < hd.size(); ++i) {
    hd[i] = 1e-6f * (float)((i * 7919u) % 101) / 101.f;
    hb[i] = 1e-9f * (float)((i * 104729u) % 89) / 89.f;
  }
  EXPECT_RC(rsim_memcpy_h2d(dl, hd.data(), vb, nullptr), RSIM_OK);
  EXPECT_RC(rsim_memcpy_h2d(bt, hb.data(), vb, nullptr), RSIM_OK);
  for (int engine : {RSIM_PLAIN, RSIM_FUSED}) {
    u.set(u0);
    EXPECT_RC(rsim_multislice_db_2d(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut,
                                    static_cast<const float*>(dl), static_cast<const float*>(bt),
                                    pz, ns, ckp, every, engine, 0, nullptr),
              RSIM_OK);
    EXPECT_RC(rsim_multislice_db_adjoint_2d(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut,
                                            static_cast<const float*>(dl),
                                            static_cast<const float*>(bt), pz, ns, ckp, every, wk,
                                            static_cast<float*>(gd), static_cast<float*>(gb),
                                            engine, 0, nullptr),
              RSIM_OK);
  }
  EXPECT_RC(rsim_synchronize(nullptr), RSIM_OK);
  for (void* q : {dl, bt, gd, gb, ckp, wk}) rsim_free(q);

  STEP("cache release invalidates batches and graphs");
  size_t cbytes = 0;
  int centries = 0;
  EXPECT_RC(rsim_cache_info(&cbytes, &centries), RSIM_OK);
  EXPECT(centries > 0);
  EXPECT_RC(rsim_release_caches(), RSIM_OK);
  EXPECT_RC(rsim_batch_run(batch, s), RSIM_E_INVALID_ARG);
  EXPECT_RC(rsim_graph_launch(graph, s), RSIM_E_INVALID_ARG);
  EXPECT_RC(rsim_batch_destroy(batch), RSIM_OK);
  EXPECT_RC(rsim_graph_destroy(graph), RSIM_OK);
  // the library rebuilds what it needs after a release
  u.set(u0);
  EXPECT_RC(rsim_sample_apply_2d(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, vol, dx, dx, pz, 0.0,
                                 0.0, 0, nz, RSIM_FUSED, 0, s),
            RSIM_OK);
  EXPECT_RC(rsim_synchronize(s), RSIM_OK);
  EXPECT(std::memcmp(u.get().data(), fused.data(), pts * 8) == 0);
  EXPECT_RC(rsim_stream_destroy(s), RSIM_OK);
  EXPECT_RC(rsim_volume_destroy(vol), RSIM_OK);
  EXPECT_RC(rsim_release_caches(), RSIM_OK);
}

// The remaining entry points with host-side work of their own: FFTs (1D rows,
// a four-step length), the point source and the cutoff, per-voxel and plasma
// transmissions (the plasma volume's tables cached per energy), history frames
// in both engines, the detector adjoint, and batched chains.
void check_device_more() {
  const int n = 256, nz = 4;
  const double dx = 0.5e-6, wl = 0.5e-10, pz = 0.5e-6, cut = 0.4 / dx;
  const size_t pts = (size_t)n * n;
  Field u(pts), U(pts);

  STEP("transforms, point source, cutoff");
  EXPECT_RC(rsim_propagate_analytically_2d(u.d, n, n, RSIM_C64, dx, dx, 0.5, 1e-6, -2e-6, wl,
                                           nullptr),
            RSIM_OK);
  EXPECT_RC(rsim_fft2(U.d, u.d, n, n, RSIM_C64, RSIM_FORWARD, nullptr), RSIM_OK);
  EXPECT_RC(rsim_apply_frequency_cutoff_2d(U.d, n, n, RSIM_C64, cut, dx, dx, nullptr), RSIM_OK);
  EXPECT_RC(rsim_fft2(u.d, U.d, n, n, RSIM_C64, RSIM_INVERSE, nullptr), RSIM_OK);
  EXPECT_RC(rsim_fft1d_rows(U.d, u.d, n, n, RSIM_C64, RSIM_FORWARD, nullptr), RSIM_OK);
  {
    const int nl = 16384, rows = 4;   // a four-step length: per-stream scratch
    Field a((size_t)nl * rows), b((size_t)nl * rows);
    EXPECT_RC(rsim_fill_zero(a.d, a.bytes, nullptr), RSIM_OK);
    EXPECT_RC(rsim_fft1d_rows(b.d, a.d, nl, rows, RSIM_C64, RSIM_FORWARD, nullptr), RSIM_OK);
    Field c(4096 * 8), d(4096 * 8);
    EXPECT_RC(rsim_stream_copy(d.d, c.d, c.bytes, nullptr), RSIM_OK);
  }
  EXPECT_RC(rsim_tilted_plane_2d(u.d, n, n, RSIM_C64, dx, dx, wl, 0.02, 1e-6, 0.0, nullptr), RSIM_OK);

  STEP("per-voxel transmission and detector adjoint");
  void *dl = nullptr, *bt = nullptr, *gimg = nullptr, *g = nullptr;
  EXPECT_RC(rsim_malloc(&dl, pts * 4), RSIM_OK);
  EXPECT_RC(rsim_malloc(&bt, pts * 4), RSIM_OK);
  EXPECT_RC(rsim_fill_zero(dl, pts * 4, nullptr), RSIM_OK);
  EXPECT_RC(rsim_fill_zero(bt, pts * 4, nullptr), RSIM_OK);
  EXPECT_RC(rsim_transmission_db_2d(u.d, n, n, RSIM_C64, wl, pz, static_cast<const float*>(dl),
                                    static_cast<const float*>(bt), nullptr),
            RSIM_OK);
  const int ox = n / 4, oy = n / 4;
  EXPECT_RC(rsim_malloc(&gimg, (size_t)ox * oy * 8), RSIM_OK);
  EXPECT_RC(rsim_malloc(&g, pts * 8), RSIM_OK);
  EXPECT_RC(rsim_fill_zero(gimg, (size_t)ox * oy * 8, nullptr), RSIM_OK);
  EXPECT_RC(rsim_detector_area_v1_adjoint(u.d, g, n, n, RSIM_C64, dx, dx, ox, oy, 4 * dx, 4 * dx,
                                          0.5, static_cast<const double*>(gimg), nullptr),
            RSIM_OK);
  for (void* q : {dl, bt, gimg, g}) rsim_free(q);

  STEP("plasma volume");
  const size_t nvox = (size_t)nz * pts;
  std::vector<double> ne(nvox, 0.0), ni(nvox, 0.0), te(nvox, 0.0), zs(nvox, 0.0);
  for (int k = 0; k < nz; ++k)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int dy = y - 128, dxp = x - 120 - 4 * k;
        if (dy * dy + dxp * dxp < 60 * 60) {
          const size_t i = (size_t)k * pts + (size_t)y * n + x;
          ne[i] = 1e22 * (1 + 0.01 * (x % 7));
          zs[i] = (x % 3) ? 8.0 : 0.0;
          ni[i] = ne[i] / 8.0;
          te[i] = (y % 2) ? 300.0 : 0.0;
        }
      }
  const double optics[4] = {13.31, 0.187, 13.0, 169.0};
  rsim_plasma* pl = nullptr;
  EXPECT_RC(rsim_plasma_create(&pl, ne.data(), ni.data(), te.data(), zs.data(), 8, nz, n, n),
            RSIM_OK);
  void* db = nullptr;
  EXPECT_RC(rsim_malloc(&db, nvox * 16), RSIM_OK);
  EXPECT_RC(rsim_plasma_deltabeta(pl, 0, nz, 8000.0, optics, db, nullptr), RSIM_OK);
  rsim_free(db);
  EXPECT_RC(rsim_plasma_transmission_2d(u.d, n, n, RSIM_C64, dx, dx, wl, pl, 1, 8000.0, optics, dx,
                                        dx, pz, 0.0, 0.0, 1.0, nullptr),
            RSIM_OK);
  for (int engine : {RSIM_PLAIN, RSIM_FUSED})
    for (double energy : {8000.0, 8500.0}) {   // a second energy builds a second cached volume
      EXPECT_RC(rsim_plasma_apply_2d(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, pl, energy,
                                     optics, dx, dx, pz, 0.0, 0.0, 0, nz, nullptr, nullptr, engine,
                                     0, nullptr),
                RSIM_OK);
    }

  STEP("history frames");
  rsim_volume* vol = nullptr;
  std::vector<uint8_t> grid(nvox, 0);
  for (size_t i = 0; i < nvox; ++i) grid[i] = (uint8_t)((i / 97) % 4 == 0 ? 1 + (i % 3) : 0);
  EXPECT_RC(rsim_volume_create(&vol, grid.data(), 1, nz, n, n, kDb, 4), RSIM_OK);
  rsim_history_frames hf = {};
  hf.out_x = ox;
  hf.out_y = oy;
  hf.pixel_size_x = hf.pixel_size_y = 4 * dx;
  hf.z_start = 0.1;
  EXPECT_RC(rsim_malloc(reinterpret_cast<void**>(&hf.frames), (size_t)nz * ox * oy * 8), RSIM_OK);
  for (int engine : {RSIM_PLAIN, RSIM_FUSED}) {
    EXPECT_RC(rsim_sample_apply_2d_history(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, vol, dx, dx,
                                           pz, 0.0, 0.0, 0, nz, nullptr, nullptr, engine, 0, &hf,
                                           nullptr),
              RSIM_OK);
    EXPECT_RC(rsim_plasma_apply_2d_history(u.d, U.d, n, n, RSIM_C64, dx, dx, wl, cut, pl, 8000.0,
                                           optics, dx, dx, pz, 0.0, 0.0, 0, nz, nullptr, nullptr,
                                           engine, 0, &hf, nullptr),
              RSIM_OK);
  }
  EXPECT_RC(rsim_synchronize(nullptr), RSIM_OK);
  rsim_free(hf.frames);

  STEP("batched chains");
  Field ub(3 * pts), Ub(3 * pts);
  std::vector<rsim_chain_step> steps(3 * 2);
  std::vector<rsim_chain_instance> inst(3);
  for (int i = 0; i < 3; ++i) {
    rsim_chain_step* s = &steps[2 * i];
    s[0] = rsim_chain_step{};
    s[0].vol = vol;
    s[0].pixel_size_x = s[0].pixel_size_y = dx;
    s[0].pixel_size_z = pz;
    s[0].x_pos = 1e-6 * i;
    s[0].n_slices = nz;
    s[1] = rsim_chain_step{};
    s[1].dz = 0.05 + 0.01 * i;
    inst[i].u = static_cast<char*>(ub.d) + i * pts * 8;
    inst[i].U = static_cast<char*>(Ub.d) + i * pts * 8;
    inst[i].dx = inst[i].dy = dx;
    inst[i].wl = wl * (1 + 0.05 * i);
    inst[i].cutoff_freq = cut;
    inst[i].steps = s;
  }
  EXPECT_RC(rsim_fill_zero(ub.d, ub.bytes, nullptr), RSIM_OK);
  EXPECT_RC(rsim_chain_apply_2d_batch(inst.data(), 3, n, n, RSIM_C64, 1, 2, RSIM_FUSED, 0, nullptr),
            RSIM_OK);
  EXPECT_RC(rsim_synchronize(nullptr), RSIM_OK);
  EXPECT_RC(rsim_plasma_destroy(pl), RSIM_OK);
  EXPECT_RC(rsim_volume_destroy(vol), RSIM_OK);
  EXPECT_RC(rsim_release_caches(), RSIM_OK);
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  if (mode != "cpu" && mode != "gpu") {
    std::fprintf(stderr, "usage: abi_check cpu|gpu\n");
    return 2;
  }
  check_arguments();
  check_controls();
  if (mode == "gpu") {
    int count = 0;
    if (rsim_device_count(&count) != RSIM_OK || count < 1) {
      std::fprintf(stderr, "abi_check: no HIP device\n");
      return 3;
    }
    check_device();
    check_device_more();
  }
  if (g_fail) {
    std::fprintf(stderr, "abi_check %s: %d expectation(s) failed\n", mode.c_str(), g_fail);
    return 1;
  }
  std::printf("abi_check %s: OK\n", mode.c_str());
  return 0;
}
