// FFT planning and dispatch: twiddle tables (FP64-computed, rounded once to
// the transform precision), schedule selection and 2D row/column passes.
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "rsim_fft_cfg.hpp"
#include "rsim_internal.hpp"

namespace rsim {

// Generic LDS kernel instantiations.
template __global__ void fft_lds_kernel<float, -1>(const cpx<float>*, cpx<float>*, AxisDesc);
template __global__ void fft_lds_kernel<float, 1>(const cpx<float>*, cpx<float>*, AxisDesc);
template __global__ void fft_lds_kernel<double, -1>(const cpx<double>*, cpx<double>*, AxisDesc);
template __global__ void fft_lds_kernel<double, 1>(const cpx<double>*, cpx<double>*, AxisDesc);

namespace {

struct Tables {
  void* full = nullptr;          // W_N^m, m in [0, N)
  void* stage[4] = {nullptr, nullptr, nullptr, nullptr};
  bool reg = false;
};

std::mutex g_mu;
std::map<std::tuple<int, int, int>, Tables> g_tables;  // (device, dtype, N)
// scratch field of the out-of-LDS schedule, one per (device, stream) so that
// transforms on different streams never share it (grown on demand, freed by
// rsim_release_caches)
std::map<std::pair<int, hipStream_t>, std::pair<void*, size_t>> g_scratch;

// exp(-2*pi*i*m/M) with the angle reduced to the first octant so the FP64
// value is correct to a few ulp for every m.
void twiddle(long long m, long long M, double* re, double* im) {
  m %= M;
  if (m < 0) m += M;
  // use symmetry: angle = 2*pi*m/M; reduce via (8m)/M octant index
  const double pi = 3.14159265358979323846264338327950288;
  long long m8 = 8 * m;
  long long oct = m8 / M;            // 0..7
  long long rem = m8 - oct * M;      // 8m mod M
  double a = (pi / 4.0) * ((double)rem / (double)M);   // in [0, pi/4)
  double c, s;
  // angle = oct*pi/4 + a; for odd octants use the complement for accuracy
  if (oct & 1) {
    double b = (pi / 4.0) * ((double)(M - rem) / (double)M);  // pi/4 - a
    double cb = std::cos(b), sb = std::sin(b);
    // angle = (oct+1)*pi/4 - b
    long long q = (oct + 1);
    // cos/sin of q*pi/4 - b for q even (2,4,6,8)
    switch (q % 8) {
      case 2: c = sb; s = cb; break;           // pi/2 - b
      case 4: c = -cb; s = sb; break;          // pi - b
      case 6: c = -sb; s = -cb; break;         // 3pi/2 - b
      default: c = cb; s = -sb; break;         // 2pi - b
    }
  } else {
    double ca = std::cos(a), sa = std::sin(a);
    switch (oct) {
      case 0: c = ca; s = sa; break;
      case 2: c = -sa; s = ca; break;          // pi/2 + a
      case 4: c = -ca; s = -sa; break;         // pi + a
      default: c = sa; s = -ca; break;         // 3pi/2 + a
    }
  }
  *re = c;
  *im = -s;
}

// W_M^m for m in [0, M) when R == 0; otherwise the [k][j] layout of a
// register-FFT stage: entry k*R + j = W_M^(j*k), M = LS*R.
int upload_table(int dtype, long long M, void** out, int R = 0) {
  std::vector<double> d(2 * M);
  if (R == 0) {
    for (long long m = 0; m < M; ++m) twiddle(m, M, &d[2 * m], &d[2 * m + 1]);
  } else {
    const long long LS = M / R;
    for (long long k = 0; k < LS; ++k)
      for (long long j = 0; j < R; ++j)
        twiddle(j * k, M, &d[2 * (k * R + j)], &d[2 * (k * R + j) + 1]);
  }
  size_t bytes;
  if (dtype == kTwComp) {
    std::vector<float> f(4 * M);
    for (long long i = 0; i < M; ++i) {
      const float hr = (float)d[2 * i], hi = (float)d[2 * i + 1];
      f[4 * i] = hr;
      f[4 * i + 1] = hi;
      f[4 * i + 2] = (float)(d[2 * i] - (double)hr);
      f[4 * i + 3] = (float)(d[2 * i + 1] - (double)hi);
    }
    bytes = 16 * M;
    RSIM_HIP(hipMalloc(out, bytes));
    RSIM_HIP(hipMemcpy(*out, f.data(), bytes, hipMemcpyHostToDevice));
  } else if (dtype == RSIM_C128) {
    bytes = 16 * M;
    RSIM_HIP(hipMalloc(out, bytes));
    RSIM_HIP(hipMemcpy(*out, d.data(), bytes, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(2 * M);
    for (long long i = 0; i < 2 * M; ++i) f[i] = (float)d[i];
    bytes = 8 * M;
    RSIM_HIP(hipMalloc(out, bytes));
    RSIM_HIP(hipMemcpy(*out, f.data(), bytes, hipMemcpyHostToDevice));
  }
  return RSIM_OK;
}

int get_tables(int dtype, int N, Tables** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(current_device(), dtype, N);
  auto it = g_tables.find(key);
  if (it != g_tables.end()) {
    *out = &it->second;
    return RSIM_OK;
  }
  RSIM_CHECK(capture_miss("FFT twiddle tables"));
  Tables t;
  RSIM_CHECK(upload_table(dtype, N, &t.full));
  SchedInfo si;
  if (sched_info(N, &si)) {
    t.reg = true;
    int ls = 1;
    for (int s = 0; s < si.nst; ++s) {
      if (s > 0) RSIM_CHECK(upload_table(dtype, (long long)ls * si.radix[s], &t.stage[s], si.radix[s]));
      ls *= si.radix[s];
    }
  }
  auto res = g_tables.emplace(key, t);
  *out = &res.first->second;
  return RSIM_OK;
}

// Greedy mixed-radix factorization for the generic kernel.
bool factorize(int n, std::vector<int>* r) {
  r->clear();
  int m = n;
  while (m % 16 == 0 && m > 16) { r->push_back(16); m /= 16; }
  const int cand[] = {16, 8, 4, 2, 5, 3};
  while (m > 1) {
    bool ok = false;
    for (int c : cand) {
      if (m % c == 0) {
        r->push_back(c);
        m /= c;
        ok = true;
        break;
      }
    }
    if (!ok) return false;
  }
  if (r->empty()) r->push_back(1);
  return true;
}

}  // namespace

bool fft_size_supported(int n) {
  if (n < 1) return false;
  std::vector<int> r;
  return factorize(n, &r);
}

const void* fft_full_table(int dtype, int N) {
  Tables* t = nullptr;
  if (get_tables(dtype, N, &t) != RSIM_OK) return nullptr;
  return t->full;
}

bool reg_stage_tables(int dtype, int N, const void* out[4]) {
  Tables* t = nullptr;
  if (get_tables(dtype, N, &t) != RSIM_OK || !t->reg) return false;
  for (int i = 0; i < 4; ++i) out[i] = t->stage[i];
  return true;
}

int rsim_release_tables() {
  std::lock_guard<std::mutex> lk(g_mu);
  int dev = current_device();
  for (auto it = g_tables.begin(); it != g_tables.end();) {
    if (std::get<0>(it->first) == dev) {
      hipFree(it->second.full);
      for (int i = 0; i < 4; ++i)
        if (it->second.stage[i]) hipFree(it->second.stage[i]);
      it = g_tables.erase(it);
    } else {
      ++it;
    }
  }
  for (auto sc = g_scratch.begin(); sc != g_scratch.end();) {
    if (sc->first.first == dev) {
      hipFree(sc->second.first);
      sc = g_scratch.erase(sc);
    } else {
      ++sc;
    }
  }
  return RSIM_OK;
}

namespace {

// Stage list of the generic kernel for N.
int axis_desc(int N, const Tables* t, AxisDesc* d) {
  std::vector<int> r;
  if (!factorize(N, &r)) return fail(RSIM_E_INVALID_ARG, "FFT length %d has a prime factor > 5", N);
  if ((int)r.size() > kMaxStages) return fail(RSIM_E_UNSUPPORTED, "FFT length %d needs too many stages", N);
  d->N = N;
  d->nst = (r.size() == 1 && r[0] == 1) ? 0 : (int)r.size();
  int ls = 1;
  for (int i = 0; i < d->nst; ++i) {
    d->radix[i] = r[i];
    d->ls[i] = ls;
    ls *= r[i];
  }
  d->tw = t->full;
  return RSIM_OK;
}

int launch_lds(int dtype, int dir, const void* in, void* out, const AxisDesc& d, int batch,
               hipStream_t s) {
  const size_t shmem = (size_t)2 * d.C * d.N * elem_bytes(dtype);
  if (shmem > max_lds_per_block())
    return fail(RSIM_E_UNSUPPORTED, "FFT length %d exceeds the single-pass LDS capacity", d.N);
  if (batch > 65535) return fail(RSIM_E_UNSUPPORTED, "FFT batch %d exceeds the grid limit", batch);
  dim3 grid((d.nseq + d.C - 1) / d.C, batch);
  dim3 block(256);
  if (dtype == RSIM_C64) {
    if (dir < 0) {
      hipFuncSetAttribute((const void*)fft_lds_kernel<float, -1>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
      hipLaunchKernelGGL((fft_lds_kernel<float, -1>), grid, block, shmem, s,
                         (const cpx<float>*)in, (cpx<float>*)out, d);
    } else {
      hipFuncSetAttribute((const void*)fft_lds_kernel<float, 1>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
      hipLaunchKernelGGL((fft_lds_kernel<float, 1>), grid, block, shmem, s,
                         (const cpx<float>*)in, (cpx<float>*)out, d);
    }
  } else {
    if (dir < 0) {
      hipFuncSetAttribute((const void*)fft_lds_kernel<double, -1>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
      hipLaunchKernelGGL((fft_lds_kernel<double, -1>), grid, block, shmem, s,
                         (const cpx<double>*)in, (cpx<double>*)out, d);
    } else {
      hipFuncSetAttribute((const void*)fft_lds_kernel<double, 1>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
      hipLaunchKernelGGL((fft_lds_kernel<double, 1>), grid, block, shmem, s,
                         (const cpx<double>*)in, (cpx<double>*)out, d);
    }
  }
  RSIM_LAUNCHED();
  return RSIM_OK;
}

int axis_scratch(size_t bytes, void** out, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto& e = g_scratch[std::make_pair(current_device(), s)];
  if (e.second < bytes) {
    RSIM_CHECK(capture_miss("the FFT scratch field of this stream"));
    if (e.first) {
      RSIM_HIP(hipStreamSynchronize(s));
      RSIM_HIP(hipFree(e.first));
      e.first = nullptr;
      e.second = 0;
    }
    RSIM_HIP(hipMalloc(&e.first, bytes));
    e.second = bytes;
  }
  *out = e.first;
  return RSIM_OK;
}

// Largest divisor of n that is <= sqrt(n) and leaves both factors
// factorizable into radices 2..16, 3, 5 (0 if none).
int split_length(int n) {
  std::vector<int> r;
  for (int a = (int)std::sqrt((double)n); a >= 2; --a)
    if (n % a == 0 && factorize(a, &r) && factorize(n / a, &r)) return a;
  return 0;
}

}  // namespace

// Out-of-LDS transform of length N = N1*N2 (four-step): x[N2*n1 + n2]
//   step 1: Y[k1, n2] = W_N^(n2 k1) * sum_n1 x[N2 n1 + n2] W_N1^(n1 k1)
//   step 2: X[k1 + N1 k2] = sum_n2 Y[k1, n2] W_N2^(n2 k2)
// through a scratch copy of the axis data (so in == out is allowed).  Rows
// (es == 1) keep the sequence index on the contiguous n2; columns (ss == 1)
// keep it on the contiguous column index and put n2 / k1 on the grid's
// batch dimension.  One batch of sequences only.
int big_axis(int dtype, int dir, const void* in, void* out, int N, long long es, long long ss,
             int nseq, int batch, long long bs, double scale, hipStream_t s) {
  if (batch != 1) return fail(RSIM_E_UNSUPPORTED, "out-of-LDS FFT of length %d needs batch 1", N);
  const bool rows = es == 1;
  if (!rows && ss != 1) return fail(RSIM_E_UNSUPPORTED, "out-of-LDS FFT layout (es=%lld, ss=%lld)", es, ss);
  const int N2 = split_length(N);
  if (N2 == 0) return fail(RSIM_E_UNSUPPORTED, "FFT length %d has no out-of-LDS split", N);
  const int N1 = N / N2;
  Tables *t1 = nullptr, *t2 = nullptr, *tN = nullptr;
  RSIM_CHECK(get_tables(dtype, N1, &t1));
  RSIM_CHECK(get_tables(dtype, N2, &t2));
  RSIM_CHECK(get_tables(dtype, N, &tN));
  const size_t eb = elem_bytes(dtype);
  void* scr = nullptr;
  RSIM_CHECK(axis_scratch((size_t)nseq * N * eb, &scr, s));
  const int kind = rows ? KT_ROW : KT_COL;
  AxisDesc a{}, b{};
  RSIM_CHECK(axis_desc(N1, t1, &a));
  RSIM_CHECK(axis_desc(N2, t2, &b));
  a.scale = 1.0;
  b.scale = scale;
  a.ptw_P = N;
  a.ptw = tN->full;
  a.C = b.C = 16;
  while (a.C > 1 && (size_t)2 * a.C * N1 * eb > max_lds_per_block()) a.C /= 2;
  while (b.C > 1 && (size_t)2 * b.C * N2 * eb > max_lds_per_block()) b.C /= 2;
  int ga, gb;
  if (rows) {
    // step 1: sequences n2 (stride 1), elements n1 (stride N2), batch q (row)
    a.es_in = N2; a.ss_in = 1; a.bs_in = ss; a.nseq = N2;
    a.es_out = N2; a.ss_out = 1; a.bs_out = N;       // scratch[q][N2 k1 + n2]
    a.ptw_mode = 1;
    // step 2: sequences k1 (stride N2), elements n2 (stride 1), batch q
    b.es_in = 1; b.ss_in = N2; b.bs_in = N; b.nseq = N1;
    b.es_out = N1; b.ss_out = 1; b.bs_out = ss;      // out[q][k1 + N1 k2]
    ga = gb = nseq;
  } else {
    // step 1: sequences q (stride 1), elements n1 (stride es*N2), batch n2
    a.es_in = es * N2; a.ss_in = 1; a.bs_in = es; a.nseq = nseq;
    a.es_out = (long long)N2 * nseq; a.ss_out = 1; a.bs_out = nseq;  // scratch[N2 k1 + n2][q]
    a.ptw_mode = 2;
    // step 2: sequences q, elements n2 (stride nseq), batch k1
    b.es_in = nseq; b.ss_in = 1; b.bs_in = (long long)N2 * nseq; b.nseq = nseq;
    b.es_out = es * N1; b.ss_out = 1; b.bs_out = es;  // out[(k1 + N1 k2) es + q]
    ga = N2;
    gb = N1;
  }
  ktime_begin(kind, s);
  RSIM_CHECK(launch_lds(dtype, dir, in, scr, a, ga, s));
  RSIM_CHECK(launch_lds(dtype, dir, scr, out, b, gb, s));
  ktime_end(kind, s);
  return RSIM_OK;
}

int fft_axis(int dtype, int dir, const void* in, void* out, int N, long long es, long long ss,
             int nseq, int batch, long long bs, double scale, hipStream_t s) {
  if (N == 1) {
    // length-1 transform is the identity (times scale)
    if (in != out || scale != 1.0) {
      // handled by the generic kernel below (cheap and rare)
    } else {
      return RSIM_OK;
    }
  }
  Tables* t = nullptr;
  RSIM_CHECK(get_tables(dtype, N, &t));
  const bool col = es != 1;
  const int kind = col ? KT_COL : KT_ROW;
  if (t->reg) {
    // complex64 register transforms take the compensated (hi + lo) tables
    Tables* tr = t;
    if (dtype == RSIM_C64) RSIM_CHECK(get_tables(kTwComp, N, &tr));
    ktime_begin(kind, s);
    const void* tw[4] = {tr->stage[0], tr->stage[1], tr->stage[2], tr->stage[3]};
    bool ok = dtype == RSIM_C64
                  ? launch_reg_fft_c64(dir, N, col, in, out, es, ss, bs, nseq, batch, tw, scale, s)
                  : launch_reg_fft_c128(dir, N, col, in, out, es, ss, bs, nseq, batch, tw, scale, s);
    if (ok) {
      ktime_end(kind, s);
      RSIM_LAUNCHED();
      return RSIM_OK;
    }
  }
  if ((size_t)2 * N * elem_bytes(dtype) > max_lds_per_block())
    return big_axis(dtype, dir, in, out, N, es, ss, nseq, batch, bs, scale, s);
  AxisDesc d{};
  RSIM_CHECK(axis_desc(N, t, &d));
  d.es_in = d.es_out = es;
  d.ss_in = d.ss_out = ss;
  d.bs_in = d.bs_out = bs;
  d.nseq = nseq;
  d.scale = scale;
  const size_t eb = elem_bytes(dtype);
  const size_t lds_cap = max_lds_per_block();
  // tile: enough sequences for >= 64 B contiguous runs on strided axes,
  // bounded by LDS (two ping-pong buffers)
  int C = col ? 8 : 1;
  while (C > 1 && (size_t)2 * C * N * eb > lds_cap) C /= 2;
  if (!col) {
    // several short rows per workgroup
    while (C * N < 1024 && C < 64 && (size_t)4 * C * N * eb <= lds_cap) C *= 2;
  }
  d.C = C;
  ktime_begin(kind, s);
  RSIM_CHECK(launch_lds(dtype, dir, in, out, d, batch, s));
  ktime_end(kind, s);
  return RSIM_OK;
}

int fft2(int dtype, int dir, const void* in, void* out, int nx, int ny, hipStream_t s) {
  const double scale = dir > 0 ? 1.0 / ((double)nx * (double)ny) : 1.0;
  // rows (contiguous) then columns (stride nx); normalisation on the last pass
  RSIM_CHECK(fft_axis(dtype, dir, in, out, nx, 1, nx, ny, 1, 0, 1.0, s));
  RSIM_CHECK(fft_axis(dtype, dir, out, out, ny, nx, 1, nx, 1, 0, scale, s));
  return RSIM_OK;
}

}  // namespace rsim
