// Device-side building blocks: complex arithmetic and small DFT butterflies.
//
// Sign convention: DIR = -1 is the forward (unnormalised) transform
// X[k] = sum_n x[n] exp(-2*pi*i*n*k/N); DIR = +1 the inverse kernel
// exp(+2*pi*i*n*k/N).  This is scipy.fft's convention, which big-wave's
// NumpyVector uses (big-wave/vector.py:341-347).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace rsim {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
  R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cpx<R> mk(R a, R b) {
  cpx<R> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename R>
__device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R>
__device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename R>
__device__ __forceinline__ cpx<R> cmulc(cpx<R> a, cpx<R> b) {
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) { return mk(a.x * s, a.y * s); }
// a * (c + i s)
template <typename R>
__device__ __forceinline__ cpx<R> cmulcs(cpx<R> a, R c, R s) {
  return mk(a.x * c - a.y * s, a.x * s + a.y * c);
}
// a * (DIR * i)
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> rot(cpx<R> a) {
  return DIR < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}
// twiddle application: forward uses the table value w, inverse conj(w)
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> twmul(cpx<R> a, cpx<R> w) {
  return DIR < 0 ? cmul(a, w) : cmulc(a, w);
}

// ---------------------------------------------------------------------------
// Butterflies: in-place DFT of size R on v[0..R-1] (natural order in and out).
// ---------------------------------------------------------------------------

template <int DIR, typename R>
__device__ __forceinline__ void dft2(cpx<R>* v) {
  cpx<R> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft3(cpx<R>* v) {
  // W3 = -1/2 + DIR i sqrt(3)/2; in FP32 the sqrt(3)/2 product is carried
  // hi + lo (an FP32-rounded constant alone would make |W3| != 1 by 3e-8 in
  // every transform)
  constexpr double kS3 = 0.86602540378443864676372317075294;
  const R c = R(-0.5);
  cpx<R> a = v[1] + v[2];
  cpx<R> b = v[1] - v[2];
  cpx<R> m = mk(v[0].x + c * a.x, v[0].y + c * a.y);
  cpx<R> sb;
  if constexpr (std::is_same<R, float>::value) {
    const float sh = (float)kS3, sl = (float)(kS3 - (double)(float)kS3);
    sb = mk(__builtin_fmaf(b.x, sh, b.x * sl), __builtin_fmaf(b.y, sh, b.y * sl));
  } else {
    sb = mk(b.x * R(kS3), b.y * R(kS3));
  }
  const cpx<R> t = DIR < 0 ? mk(sb.y, -sb.x) : mk(-sb.y, sb.x);  // DIR * i * s * b
  v[0] = v[0] + a;
  v[1] = m + t;
  v[2] = m - t;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft4(cpx<R>* v) {
  cpx<R> t0 = v[0] + v[2], t1 = v[0] - v[2];
  cpx<R> t2 = v[1] + v[3], t3 = rot<DIR>(v[1] - v[3]);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + t3;
  v[3] = t1 - t3;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft5(cpx<R>* v) {
  const R c1 = R(0.30901699437494742410229341718282);   // cos(2pi/5)
  const R c2 = R(-0.80901699437494742410229341718282);  // cos(4pi/5)
  const R s1 = R(DIR) * R(0.95105651629515357211643933337938);
  const R s2 = R(DIR) * R(0.58778525229247312916870595463907);
  cpx<R> a1 = v[1] + v[4], b1 = v[1] - v[4];
  cpx<R> a2 = v[2] + v[3], b2 = v[2] - v[3];
  cpx<R> x0 = v[0];
  cpx<R> m1 = mk(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  cpx<R> m2 = mk(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  cpx<R> n1 = mk(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
  cpx<R> n2 = mk(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
  // i * n
  cpx<R> in1 = mk(-n1.y, n1.x), in2 = mk(-n2.y, n2.x);
  v[0] = x0 + a1 + a2;
  v[1] = m1 + in1;
  v[4] = m1 - in1;
  v[2] = m2 + in2;
  v[3] = m2 - in2;
}

// hi/lo splits of the rotation constants (exact decimal values rounded in
// two steps: hi = fl32(v), lo = fl32(v - hi))
constexpr double kC1 = 0.92387953251128675612818318939679;  // cos(pi/8)
constexpr double kS1 = 0.38268343236508977172845998403040;  // sin(pi/8)
constexpr double kH = 0.70710678118654752440084436210485;   // sqrt(1/2)
constexpr double kC16 = 0.98078528040323044912618223613424;  // cos(pi/16)
constexpr double kS16 = 0.19509032201612826784828486847702;  // sin(pi/16)
constexpr double kC316 = 0.83146961230254523707878837761791; // cos(3pi/16)
constexpr double kS316 = 0.55557023301960222474283081394853; // sin(3pi/16)
#define RSIM_HI(v) ((float)(v))
#define RSIM_LO(v) ((float)((v) - (double)(float)(v)))

// Rotation constants of the radix-8/16 butterflies.  FP32 roundings of
// cos(pi/8) and sin(pi/8) leave |W16|^2 - 1 = -5.7e-8, the same in every
// transform; over a multislice run (a forward and an inverse transform per
// axis and slice) that biases the field norm linearly.  The inverse
// direction therefore uses FP32 values within an ulp, chosen so that
// |W_inverse|^2 * |W_forward|^2 - 1 = -4.0e-9: the pair's norm errors cancel.
template <int DIR, typename R>
struct RotK {
  static constexpr R c1 = R(kC1);
  static constexpr R s1 = R(kS1);
};
template <>
struct RotK<1, float> {
  static constexpr float c1 = 0.9238795638084412f;   // fl32(cos(pi/8)) + 1 ulp
  static constexpr float s1 = 0.3826834261417389f;   // fl32(sin(pi/8))
};

// v * W8^(-DIR) = v * (1 + DIR i) / sqrt(2): the sum and difference first,
// then one scaling by sqrt(1/2), carried hi + lo in FP32 (an FP32-rounded
// sqrt(1/2) alone would change |v| by -1.7e-8 at every rotation).
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> rot45(cpx<R> v) {
  const R a = v.x - R(DIR) * v.y;
  const R b = v.y + R(DIR) * v.x;
  if constexpr (std::is_same<R, float>::value) {
    const float hh = RSIM_HI(kH), hl = RSIM_LO(kH);
    return mk(__builtin_fmaf(a, hh, a * hl), __builtin_fmaf(b, hh, b * hl));
  } else {
    return mk(a * R(kH), b * R(kH));
  }
}

template <int DIR, typename R>
__device__ __forceinline__ void dft8(cpx<R>* v) {
  cpx<R> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<R> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // W8^1 = (h, DIR*h), W8^2 = DIR*i, W8^3 = W8^2 * W8^1
  cpx<R> o1 = rot45<DIR>(o[1]);
  cpx<R> o2 = rot<DIR>(o[2]);
  cpx<R> o3 = rot<DIR>(rot45<DIR>(o[3]));
  v[0] = e[0] + o[0];
  v[4] = e[0] - o[0];
  v[1] = e[1] + o1;
  v[5] = e[1] - o1;
  v[2] = e[2] + o2;
  v[6] = e[2] - o2;
  v[3] = e[3] + o3;
  v[7] = e[3] - o3;
}

template <int DIR, typename R>
__device__ __forceinline__ void dft16(cpx<R>* v) {
  // 16 = 4 x 4: Y[n2][k1] = DFT4 over n1 of x[4 n1 + n2], twiddle W16^(n2 k1),
  // X[k1 + 4 k2] = DFT4 over n2.
  const R c1 = RotK<DIR, R>::c1;
  const R s1 = RotK<DIR, R>::s1;
  cpx<R> y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    y[n2][0] = v[n2];
    y[n2][1] = v[4 + n2];
    y[n2][2] = v[8 + n2];
    y[n2][3] = v[12 + n2];
    dft4<DIR>(y[n2]);
  }
  const R d = R(DIR);
  // n2 = 1: W^1, W^2, W^3
  y[1][1] = cmulcs(y[1][1], c1, d * s1);
  y[1][2] = rot45<DIR>(y[1][2]);
  y[1][3] = cmulcs(y[1][3], s1, d * c1);
  // n2 = 2: W^2, W^4, W^6
  y[2][1] = rot45<DIR>(y[2][1]);
  y[2][2] = rot<DIR>(y[2][2]);
  y[2][3] = rot<DIR>(rot45<DIR>(y[2][3]));
  // n2 = 3: W^3, W^6, W^9 = -W^1
  y[3][1] = cmulcs(y[3][1], s1, d * c1);
  y[3][2] = rot<DIR>(rot45<DIR>(y[3][2]));
  y[3][3] = cmulcs(y[3][3], -c1, -d * s1);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cpx<R> z[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4<DIR>(z);
    v[k1] = z[0];
    v[k1 + 4] = z[1];
    v[k1 + 8] = z[2];
    v[k1 + 12] = z[3];
  }
}

// Rotation constants of the radix-32 butterfly: cos and sin of pi/16 and
// 3 pi/16.  As for radix 16, the inverse direction takes FP32 values within
// an ulp that match the forward ones' norm error: |W_inv|^2 |W_fwd|^2 - 1 is
// 1.6e-9 (pi/16) and 2.9e-9 (3 pi/16), where plain roundings leave -1.2e-7
// and -3.0e-8 per transform pair.
template <int DIR, typename R>
struct RotK32 {
  static constexpr R a1 = R(kC16), b1 = R(kS16), a3 = R(kC316), b3 = R(kS316);
};
template <>
struct RotK32<1, float> {
  static constexpr float a1 = 0x1.f6297ep-1f;   // fl32(cos(pi/16)) + 1 ulp
  static constexpr float b1 = 0x1.8f8b84p-3f;   // fl32(sin(pi/16))
  static constexpr float a3 = 0x1.a9b664p-1f;   // fl32(cos(3pi/16)) + 1 ulp
  static constexpr float b3 = 0x1.1c73b2p-1f;   // fl32(sin(3pi/16)) - 1 ulp
};

// Radix 32 = 2 x 16 (decimation in time): E = DFT16 of the even inputs,
// O = DFT16 of the odd ones, X[k] = E[k] + W32^k O[k], X[k + 16] =
// E[k] - W32^k O[k], W32 = exp(DIR 2 pi i / 32).
template <int DIR, typename R>
__device__ __forceinline__ void dft32(cpx<R>* v) {
  cpx<R> e[16], o[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  dft16<DIR>(e);
  dft16<DIR>(o);
  const R c1 = RotK<DIR, R>::c1, s1 = RotK<DIR, R>::s1;
  const R a1 = RotK32<DIR, R>::a1, b1 = RotK32<DIR, R>::b1;
  const R a3 = RotK32<DIR, R>::a3, b3 = RotK32<DIR, R>::b3;
  const R d = R(DIR);
  // W32^k = (cos(k pi/16), DIR sin(k pi/16))
  o[1] = cmulcs(o[1], a1, d * b1);
  o[2] = cmulcs(o[2], c1, d * s1);
  o[3] = cmulcs(o[3], a3, d * b3);
  o[4] = rot45<DIR>(o[4]);
  o[5] = cmulcs(o[5], b3, d * a3);
  o[6] = cmulcs(o[6], s1, d * c1);
  o[7] = cmulcs(o[7], b1, d * a1);
  o[8] = rot<DIR>(o[8]);
  o[9] = cmulcs(o[9], -b1, d * a1);
  o[10] = cmulcs(o[10], -s1, d * c1);
  o[11] = cmulcs(o[11], -b3, d * a3);
  o[12] = rot<DIR>(rot45<DIR>(o[12]));
  o[13] = cmulcs(o[13], -a3, d * b3);
  o[14] = cmulcs(o[14], -c1, d * s1);
  o[15] = cmulcs(o[15], -a1, d * b1);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    v[k] = e[k] + o[k];
    v[k + 16] = e[k] - o[k];
  }
}

// ---------------------------------------------------------------------------
// Compensated single precision: constants and twiddles carried as FP32
// hi + lo pairs (value = hi + lo to ~2^-48), products formed with FMA so that
// every product carries only its own rounding.  FP32-rounded rotation
// constants make |W| != 1 by ~1e-8, a bias that repeats identically in every
// transform and compounds linearly over a multislice run; compensation
// removes it while keeping FP32 data and FP32 register pressure.
// ---------------------------------------------------------------------------
struct twc {
  float x, y, lx, ly;   // hi.re, hi.im, lo.re, lo.im
};

// a * (c + i s) with c = ch + cl, s = sh + sl
__device__ __forceinline__ cpx<float> cmul_comp(cpx<float> a, float ch, float cl, float sh,
                                                float sl) {
  const float re_lo = a.x * cl - a.y * sl;
  const float im_lo = a.x * sl + a.y * cl;
  return mk(__builtin_fmaf(a.x, ch, __builtin_fmaf(-a.y, sh, re_lo)),
            __builtin_fmaf(a.x, sh, __builtin_fmaf(a.y, ch, im_lo)));
}
template <int DIR>
__device__ __forceinline__ cpx<float> twmul_comp(cpx<float> a, const twc& w) {
  // forward: a * w; inverse: a * conj(w)
  if (DIR < 0) return cmul_comp(a, w.x, w.lx, w.y, w.ly);
  return cmul_comp(a, w.x, w.lx, -w.y, -w.ly);
}


template <int DIR>
__device__ __forceinline__ void dft8_comp(cpx<float>* v) {
  const float hh = RSIM_HI(kH), hl = RSIM_LO(kH);
  const float d = (float)DIR;
  cpx<float> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<float> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  cpx<float> o1 = cmul_comp(o[1], hh, hl, d * hh, d * hl);
  cpx<float> o2 = rot<DIR>(o[2]);
  cpx<float> o3 = rot<DIR>(cmul_comp(o[3], hh, hl, d * hh, d * hl));
  v[0] = e[0] + o[0];
  v[4] = e[0] - o[0];
  v[1] = e[1] + o1;
  v[5] = e[1] - o1;
  v[2] = e[2] + o2;
  v[6] = e[2] - o2;
  v[3] = e[3] + o3;
  v[7] = e[3] - o3;
}

template <int DIR>
__device__ __forceinline__ void dft16_comp(cpx<float>* v) {
  const float ch = RSIM_HI(kC1), cl = RSIM_LO(kC1);
  const float sh = RSIM_HI(kS1), sl = RSIM_LO(kS1);
  const float hh = RSIM_HI(kH), hl = RSIM_LO(kH);
  const float d = (float)DIR;
  cpx<float> y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    y[n2][0] = v[n2];
    y[n2][1] = v[4 + n2];
    y[n2][2] = v[8 + n2];
    y[n2][3] = v[12 + n2];
    dft4<DIR>(y[n2]);
  }
  y[1][1] = cmul_comp(y[1][1], ch, cl, d * sh, d * sl);
  y[1][2] = cmul_comp(y[1][2], hh, hl, d * hh, d * hl);
  y[1][3] = cmul_comp(y[1][3], sh, sl, d * ch, d * cl);
  y[2][1] = cmul_comp(y[2][1], hh, hl, d * hh, d * hl);
  y[2][2] = rot<DIR>(y[2][2]);
  y[2][3] = rot<DIR>(cmul_comp(y[2][3], hh, hl, d * hh, d * hl));
  y[3][1] = cmul_comp(y[3][1], sh, sl, d * ch, d * cl);
  y[3][2] = rot<DIR>(cmul_comp(y[3][2], hh, hl, d * hh, d * hl));
  y[3][3] = cmul_comp(y[3][3], -ch, -cl, -d * sh, -d * sl);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cpx<float> z[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4<DIR>(z);
    v[k1] = z[0];
    v[k1 + 4] = z[1];
    v[k1 + 8] = z[2];
    v[k1 + 12] = z[3];
  }
}

// Radix 32 in compensated FP32 (dft32 with hi + lo constants).
template <int DIR>
__device__ __forceinline__ void dft32_comp(cpx<float>* v) {
  cpx<float> e[16], o[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  dft16_comp<DIR>(e);
  dft16_comp<DIR>(o);
  const float d = (float)DIR;
  const float ch = RSIM_HI(kC1), cl = RSIM_LO(kC1), sh = RSIM_HI(kS1), sl = RSIM_LO(kS1);
  const float hh = RSIM_HI(kH), hl = RSIM_LO(kH);
  const float ah = RSIM_HI(kC16), al = RSIM_LO(kC16), bh = RSIM_HI(kS16), bl = RSIM_LO(kS16);
  const float gh = RSIM_HI(kC316), gl = RSIM_LO(kC316), qh = RSIM_HI(kS316), ql = RSIM_LO(kS316);
  o[1] = cmul_comp(o[1], ah, al, d * bh, d * bl);
  o[2] = cmul_comp(o[2], ch, cl, d * sh, d * sl);
  o[3] = cmul_comp(o[3], gh, gl, d * qh, d * ql);
  o[4] = cmul_comp(o[4], hh, hl, d * hh, d * hl);
  o[5] = cmul_comp(o[5], qh, ql, d * gh, d * gl);
  o[6] = cmul_comp(o[6], sh, sl, d * ch, d * cl);
  o[7] = cmul_comp(o[7], bh, bl, d * ah, d * al);
  o[8] = rot<DIR>(o[8]);
  o[9] = cmul_comp(o[9], -bh, -bl, d * ah, d * al);
  o[10] = cmul_comp(o[10], -sh, -sl, d * ch, d * cl);
  o[11] = cmul_comp(o[11], -qh, -ql, d * gh, d * gl);
  o[12] = rot<DIR>(cmul_comp(o[12], hh, hl, d * hh, d * hl));
  o[13] = cmul_comp(o[13], -gh, -gl, d * qh, d * ql);
  o[14] = cmul_comp(o[14], -ch, -cl, d * sh, d * sl);
  o[15] = cmul_comp(o[15], -ah, -al, d * bh, d * bl);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    v[k] = e[k] + o[k];
    v[k + 16] = e[k] - o[k];
  }
}

// Prime-factor (Good-Thomas) radix-6 and radix-12 butterflies: for
// N = N1 * N2 with coprime factors, x[(N2 n1 + N1 n2) mod N] -> DFT_N2 over n2
// -> DFT_N1 over n1 -> X[k], k = k1 (mod N1), k = k2 (mod N2), with no
// twiddles between the two sub-transforms.
template <int DIR, typename R>
__device__ __forceinline__ void dft6(cpx<R>* v) {
  // N1 = 2, N2 = 3: inputs (3 n1 + 2 n2) mod 6, outputs (3 k1 + 4 k2) mod 6
  cpx<R> a[3] = {v[0], v[2], v[4]};
  cpx<R> b[3] = {v[3], v[5], v[1]};
  dft3<DIR>(a);
  dft3<DIR>(b);
  v[0] = a[0] + b[0];
  v[3] = a[0] - b[0];
  v[4] = a[1] + b[1];
  v[1] = a[1] - b[1];
  v[2] = a[2] + b[2];
  v[5] = a[2] - b[2];
}

template <int DIR, typename R>
__device__ __forceinline__ void dft12(cpx<R>* v) {
  // N1 = 4, N2 = 3: inputs (3 n1 + 4 n2) mod 12, outputs (9 k1 + 4 k2) mod 12
  cpx<R> y[4][3];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) y[n1][n2] = v[(3 * n1 + 4 * n2) % 12];
    dft3<DIR>(y[n1]);
  }
#pragma unroll
  for (int k2 = 0; k2 < 3; ++k2) {
    cpx<R> z[4] = {y[0][k2], y[1][k2], y[2][k2], y[3][k2]};
    dft4<DIR>(z);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) v[(9 * k1 + 4 * k2) % 12] = z[k1];
  }
}

template <int RADIX, int DIR, typename R>
__device__ __forceinline__ void dft(cpx<R>* v) {
  if constexpr (RADIX == 2) dft2<DIR>(v);
  else if constexpr (RADIX == 3) dft3<DIR>(v);
  else if constexpr (RADIX == 4) dft4<DIR>(v);
  else if constexpr (RADIX == 5) dft5<DIR>(v);
  else if constexpr (RADIX == 6) dft6<DIR>(v);
  else if constexpr (RADIX == 8) dft8<DIR>(v);
  else if constexpr (RADIX == 12) dft12<DIR>(v);
  else if constexpr (RADIX == 16) dft16<DIR>(v);
  else if constexpr (RADIX == 32) dft32<DIR>(v);
}

template <int RADIX, int DIR>
__device__ __forceinline__ void dft_comp(cpx<float>* v) {
  if constexpr (RADIX == 8) dft8_comp<DIR>(v);
  else if constexpr (RADIX == 16) dft16_comp<DIR>(v);
  else if constexpr (RADIX == 32) dft32_comp<DIR>(v);
  else dft<RADIX, DIR>(v);
}


// ---------------------------------------------------------------------------
// Buffer-resource global I/O: the base address lives in SGPRs (a V#), each
// lane supplies a 32-bit byte offset and each element a uniform SGPR offset,
// so a thread's E loads/stores cost one address VGPR instead of E 64-bit
// pointers.  num_records bounds every access (out-of-range loads return 0,
// stores are dropped).  Word 3 0x00020000 is the gfx9-family raw-buffer
// format.
// ---------------------------------------------------------------------------
template <typename R>
struct Gio {
  __amdgpu_buffer_rsrc_t rs;
  __device__ __forceinline__ Gio(const void* base, uint32_t nbytes)
      : rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)nbytes,
                                             0x00020000)) {}
  __device__ __forceinline__ cpx<R> ld(uint32_t voff, uint32_t soff) const {
    if constexpr (sizeof(R) == 4) {
      return __builtin_bit_cast(cpx<R>, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
    } else {
      return __builtin_bit_cast(cpx<R>, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0));
    }
  }
  __device__ __forceinline__ void st(cpx<R> v, uint32_t voff, uint32_t soff) const {
    if constexpr (sizeof(R) == 4) {
      typedef unsigned int u2 __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rs, voff, soff, 0);
    } else {
      typedef unsigned int u4 __attribute__((ext_vector_type(4)));
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), rs, voff, soff, 0);
    }
  }
};

// Compiler-only memory fence: keeps the twiddle loads of one transform from
// being merged with (and held live across) those of the next one.
__device__ __forceinline__ void compiler_fence() { __asm__ __volatile__("" ::: "memory"); }

// XCD-aware tile order: consecutive workgroup ids are dispatched round-robin
// over the 8 XCDs, so map them to tiles such that each XCD walks a contiguous
// run of tiles (neighbouring column tiles share 128 B lines in its L2).
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  if (ntiles % 8 != 0) return b;
  return (b % 8) * (ntiles / 8) + b / 8;
}

// exp(-kb) * (cos kd, -sin kd) in FP32: the factor exp(-i*K*(delta - i*beta))
// of a bilinearly interpolated voxel (K*delta = kd, K*beta = kb).  The phase
// is reduced in FP64 by the nearest multiple of pi/2 (two-term Cody-Waite
// split of pi/2; exact to far beyond any slice phase), sin and cos of the
// FP32 remainder (|r| <= pi/4) are their Taylor series to r^9 / r^10
// (truncation below 2e-9), and exp(-kb) is the hardware base-2 exponential
// of -kb*log2(e).  Branch-free; max error against FP64 about 1.2e-7, the
// same as sincosf and expf.
__device__ __forceinline__ cpx<float> voxel_factor(float kd, float kb) {
  const float att = __builtin_amdgcn_exp2f(-kb * 1.44269504088896341f);
  const double dk = (double)kd;
  const double q = rint(dk * 0.63661977236758134308);
  double rd = fma(-q, 1.5707963267948966192, dk);
  rd = fma(-q, 6.1232339957367658e-17, rd);
  const float r = (float)rd;
  const int qi = (int)(q - 4.0 * floor(q * 0.25));   // q mod 4
  const float r2 = r * r;
  const float sp = r2 * fmaf(r2, fmaf(r2, fmaf(r2, 2.75573192239858907e-06f,
                                               -1.98412698412698413e-04f),
                                      8.33333333333333333e-03f),
                             -1.66666666666666667e-01f);
  const float sr = fmaf(r, sp, r);
  const float cr = fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, -2.75573192239858907e-07f,
                                                            2.48015873015873016e-05f),
                                                   -1.38888888888888889e-03f),
                                          4.16666666666666667e-02f),
                                 -0.5f), 1.0f);
  const float a = (qi & 1) ? cr : sr;   // sin(kd) up to sign
  const float b = (qi & 1) ? sr : cr;   // cos(kd) up to sign
  const float s = (qi & 2) ? -a : a;
  const float c = ((qi + 1) & 2) ? -b : b;
  return mk(att * c, -(att * s));
}

// sincos of a phase already reduced to [-pi, pi] (single precision output).
__device__ __forceinline__ void sincos_reduced(double ph, float* s, float* c) {
  sincosf((float)ph, s, c);
}

// Reduce an FP64 phase to [-pi, pi] before a single-precision sincos.
__device__ __forceinline__ double reduce_2pi(double ph) {
  const double two_pi = 6.283185307179586476925286766559;
  return ph - two_pi * rint(ph / two_pi);
}

}  // namespace rsim
