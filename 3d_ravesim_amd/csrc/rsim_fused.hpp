// Fused complex64 multislice kernels and their launch-shape variants.
//
// Row pass (one or more rows y per workgroup, positions x):
//   FIRST  V = FFT_x(T_s * u)
//   MID    V = FFT_x(T_s * IFFT_x(W))      (IFFT unnormalised; 1/N is in M)
//   LAST   u = IFFT_x(W)
//   ADJ    V = FFT_x(conj(T_s) * h), h = IFFT_x(W), with the slice's
//          gradients from h and the stored field u_s (adjoint chain)
//   ADJ_LAST  g = conj(T_s) * IFFT_x(W), with the gradients
// Column pass (C columns x per workgroup, positions y):
//   W = IFFT_y(M * FFT_y(V)),  M = exp(i(theta_g+theta_x+theta_y)) * cutoff / (nx*ny)
// optionally storing the spectrum M*FFT2 (= FFT2 of the new field) on the
// last slice.  T_s is the voxel transmission of slice s (optical_element.py:
// 552-584); in the aligned case (every voxel coordinate an exact integer,
// e.g. voxel pitch == field pitch) it is a per-material lookup.
#pragma once

#include "rsim_fft_cfg.hpp"
#include "rsim_tables.hpp"

namespace rsim {

// Transmission source of the fused row pass: bilinear over material ids,
// per-material lookup (integral voxel map), bilinear over a per-voxel
// delta-beta field (plasma samples), separate delta and beta planes on the
// field grid itself (differentiable path), or none (a plain row FFT).
enum { TM_BILINEAR = 0, TM_LUT = 1, TM_FIELD = 2, TM_PLANAR = 3, TM_NONE = 4 };

struct TransArgsF {
  const uint8_t* ids;     // slice base (ny_g x nx_g)
  const float2* dbv;      // per-voxel (delta, beta) slice base, TM_FIELD
  const float* dpl;       // delta and beta slice bases, TM_PLANAR (row stride nxg = nx)
  const float* bpl;
  int nxg;
  AxisMapDev xm, ym;
  const float4* lut;      // per material (hi.re, hi.im, lo.re, lo.im) of the FP64 factor
  const float2* dbf;      // per material (delta, beta), bilinear mode
  float K;                // 2*pi*pz/wl
  // LUT mode (integral voxel map): voxel x = ax*x + cx for x in [xlo, xhi)
  int ax, cx, xlo, xhi, ay, cy, ylo, yhi;
};

// Material tables staged in LDS (256 entries); filled by the whole workgroup:
// the LUT factors (TM_LUT) or the (delta, beta) pairs (TM_BILINEAR).
template <typename T>
__device__ __forceinline__ void table_to_lds(T* lds_tab, const T* tab) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) lds_tab[i] = tab[i];
}

// v times a material factor held as FP32 hi + lo of the FP64 value.
__device__ __forceinline__ cpx<float> lut_mul(cpx<float> v, float4 t) {
  const float rr = v.x * t.x - v.y * t.y + (v.x * t.z - v.y * t.w);
  const float ii = v.x * t.y + v.y * t.x + (v.x * t.w + v.y * t.z);
  return mk(rr, ii);
}

// Transmission of one element.  TM_LUT: the voxel map is integral, hence
// affine (x0 = ax*x + cx); the factor is a per-material lookup (LDS) of the
// FP64 value split hi+lo.  Otherwise bilinear delta-beta interpolation from
// the four corners (material table or per-voxel field) and an FP32
// exponential.  row0 = ids row of this field row (TM_LUT).
template <int TM>
__device__ __forceinline__ cpx<float> transmit(cpx<float> v, const TransArgsF& a, int x, int y0,
                                               float fy, bool iny, const uint8_t* row0,
                                               const float4* lds_lut, const float2* lds_dbf) {
  if constexpr (TM == TM_NONE) {
    return v;
  } else if constexpr (TM == TM_PLANAR) {
    const long long i = (long long)y0 * a.nxg + x;
    const cpx<float> t = voxel_factor(a.K * a.dpl[i], a.K * a.bpl[i]);
    return mk(v.x * t.x - v.y * t.y, v.x * t.y + v.y * t.x);
  } else if constexpr (TM == TM_LUT) {
    if (!(iny && x >= a.xlo && x < a.xhi)) return v;
    return lut_mul(v, lds_lut[row0[a.ax * x + a.cx]]);
  } else {
    if (!(iny && a.xm.inside[x])) return v;
    const int x0 = a.xm.i0[x];
    const float fx = a.xm.fracf[x];
    float2 d00, d01, d10, d11;
    if constexpr (TM == TM_FIELD) {
      const float2* r0 = a.dbv + (long long)y0 * a.nxg + x0;
      const float2* r1 = r0 + a.nxg;
      d00 = r0[0];
      d01 = r0[1];
      d10 = r1[0];
      d11 = r1[1];
    } else {
      const uint8_t* r0 = a.ids + (long long)y0 * a.nxg + x0;
      const uint8_t* r1 = r0 + a.nxg;
      d00 = lds_dbf[r0[0]];
      d01 = lds_dbf[r0[1]];
      d10 = lds_dbf[r1[0]];
      d11 = lds_dbf[r1[1]];
    }
    const float gx = 1.f - fx, gy = 1.f - fy;
    const float de = (d00.x * gx + d01.x * fx) * gy + (d10.x * gx + d11.x * fx) * fy;
    const float be = (d00.y * gx + d01.y * fx) * gy + (d10.y * gx + d11.y * fx) * fy;
    const cpx<float> t = voxel_factor(a.K * de, a.K * be);
    return mk(v.x * t.x - v.y * t.y, v.x * t.y + v.y * t.x);
  }
}

// A copy of a lane index that the compiler cannot prove equal to the
// original (LD), or the index itself.  With LD the second transform of a
// kernel recomputes its LDS exchange and twiddle addresses instead of keeping
// the first transform's live through the pointwise step: fewer registers
// (no spills in the 2048, 3072, 6144, 8192 and 16384 shapes), a few more
// VALU instructions.  Both forms are launch-shape variants.  The empty asm
// statement only hides the value from the optimiser; it emits no
// instruction.  The mask restores what the optimiser knew of the index, that
// it lies in [0, 1024), so offsets still fold into instruction immediates.
template <bool LD>
__device__ __forceinline__ int opaque_index(int x) {
  if constexpr (LD) {
    asm volatile("" : "+v"(x));
    x &= 1023;
  }
  return x;
}

// Per-row transmission state.
struct RowT {
  int y0;
  float fy;
  bool iny;
  const uint8_t* row0;
};
template <int TM>
__device__ __forceinline__ RowT row_state(const TransArgsF& a, int y) {
  RowT r;
  if constexpr (TM == TM_LUT) {
    r.iny = y >= a.ylo && y < a.yhi;
    r.y0 = a.ay * y + a.cy;
    r.fy = 0.f;
    r.row0 = a.ids + (long long)(r.iny ? r.y0 : 0) * a.nxg;
  } else if constexpr (TM == TM_PLANAR || TM == TM_NONE) {
    r.y0 = y;
    r.fy = 0.f;
    r.iny = true;
    r.row0 = a.ids;
  } else {
    r.y0 = a.ym.i0[y];
    r.fy = a.ym.fracf[y];
    r.iny = a.ym.inside[y];
    r.row0 = a.ids;
  }
  return r;
}

// ROW_MID_OUT is ROW_MID with its optional outputs (AdjArgs fld_out, xb);
// the launcher picks it when a launch asks for one, so the plain MID kernel
// carries no code for them.
enum { ROW_FIRST = 0, ROW_MID = 1, ROW_LAST = 2, ROW_ADJ = 3, ROW_ADJ_LAST = 4, ROW_MID_OUT = 5 };

// Extra buffers of single-field launches.
//   ROW_MID: fld_out (optional) receives the field IFFT_x(W) entering the
//     slice (history frames; the differentiable path's checkpoints and
//     rebuilt fields);
//   ROW_ADJ / ROW_ADJ_LAST: fld_in = the field u_s entering slice s, and
//     gd / gb (ny x nx) receive dL/d delta_s and dL/d beta_s.
//   ROW_MID: xb.xs (optional) receives the x-integrals of |IFFT_x(W)|^2 over
//     the area_v1 detector columns (history frames), rows [j0, j0 + nj) of
//     the field to xs[(y - j0) * ox + p]: detector column p covers cells
//     [xi0, xi1] with fractions xf0 / xf1 taken off / added at the ends.
struct AdjArgs {
  const cpx<float>* fld_in;
  cpx<float>* fld_out;
  float* gd;
  float* gb;
  XBinArgs xb;
};

// Per-instance launch arguments.  A launch covers up to kMaxBatch fields of
// one shape (grid.y = instance); the descriptors travel by value in the
// kernel arguments, so a captured graph keeps each slice's pointers.
constexpr int kMaxBatch = 16;
struct RowInst {
  const cpx<float>* in;
  cpx<float>* out;
  TransArgsF ta;   // ta.ids: this launch's voxel slice
};
// in_tile: layout of the launch's inputs.  0: row-major; w > 0: column tiles
// of width w as a tiled column pass leaves them (element (y, x) at
// ((x / w) * ny + y) * w + x % w), w a power of two dividing the shape's TC.
struct RowBatch {
  RowInst inst[kMaxBatch];
  int in_tile;
};
struct ColInst {
  const cpx<float>* in;
  cpx<float>* out;
  cpx<float>* spec;
  const double2* Bx;
  const double2* Ay;
  const int* Rmax;
};
// out_tile: the launch stores its result in column tiles of its own width C
// (each workgroup writes one contiguous block of C x ny elements), for a row
// pass that reads them with in_tile = C; 0: row-major.
struct ColBatch {
  ColInst inst[kMaxBatch];
  int out_tile;
};

// Occupancy request of a row launch shape: LD = 2 asks the compiler for three
// workgroups of up to 512 threads per CU (waves per SIMD = 3 x the
// workgroup's share), LD = 3 for two, i.e. a register budget that lets them
// in; LD = 4 for four waves per SIMD whatever the workgroup size (128
// registers: four 256-thread or two 512-thread workgroups per CU).
template <class F, int C, int LD>
constexpr int row_waves_per_eu() {
  return (LD == 2 && F::TC * C <= 512)   ? 3 * ((F::TC * C + 255) / 256)
         : (LD == 3 && F::TC * C <= 512) ? 2 * ((F::TC * C + 255) / 256)
         : (LD == 4 && F::TC * C <= 512) ? 4
                                         : 1;
}

// Where a MID row keeps the voxel ids it loads with the field until the
// transmission between its transforms: 1 = packed four per register, or
// 2 = packed four per word in LDS for the three-workgroup register budget
// (LD = 2) where three workgroups still fit beside them: no registers held
// through the first transform (the 4096 row's spills go away, MID 93.2 ->
// 90.8 us).  Shapes with registers to spare run faster with 1 (2048 row
// 24.7 us against 26.7 with 2).  The four-waves budget (LD = 4) stashes them
// in LDS too where its workgroups (16 waves per CU) still fit.
template <class F, int C, int PS, bool SPLIT, int LD>
constexpr int row_id_mode() {
  const long long lds = (long long)LdsSize<F, C, PS>::slots * (SPLIT ? 4 : 8) + 256 * 16 +
                        (long long)((F::E + 3) / 4) * 4 * F::TC * C;
  const long long wgs4 = 16 / ((F::TC * C + 63) / 64);
  return (LD == 2 && 3 * lds <= 163840) || (LD == 4 && wgs4 >= 1 && wgs4 * lds <= 163840) ? 2 : 1;
}

template <class F, int C, int PS, bool SPLIT, int LD, int MODE, int TM, typename TW>
__global__ void __launch_bounds__(F::TC * C)
    __attribute__((amdgpu_waves_per_eu(row_waves_per_eu<F, C, LD>())))
    ms_row_kernel(RowBatch rb, int ny, uint32_t nbytes, StageTw<TW> tw, AdjArgs aj) {
  constexpr bool LUT = TM == TM_LUT;
  constexpr int N = F::N;
  constexpr bool PRE = (MODE == ROW_MID || MODE == ROW_MID_OUT) && LUT;
  constexpr int IDM = PRE ? row_id_mode<F, C, PS, SPLIT, LD>() : 0;
  constexpr int NW = (F::E + 3) / 4;
  constexpr int XB = LdsSize<F, C, PS>::slots * (SPLIT ? 4 : 8);
  // the exchange buffer, then (IDM = 2) the packed ids
  __shared__ unsigned char lds[XB + (IDM == 2 ? NW * 4 * F::TC * C : 0)];
  __shared__ float4 lds_lut[(MODE != ROW_LAST && LUT) ? 256 : 1];
  __shared__ float2 lds_dbf[(MODE != ROW_LAST && TM == TM_BILINEAR) ? 256 : 1];
  uint32_t* const lds_ids = reinterpret_cast<uint32_t*>(lds + XB);
  const RowInst& I = rb.inst[blockIdx.y];
  const cpx<float>* in = I.in;
  cpx<float>* out = I.out;
  const TransArgsF& ta = I.ta;
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  const int tile = rb.in_tile;   // uniform
  // tiled input: rows that share 128 B lines of a tile go to one XCD's L2
  const int y = (tile ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x) * C + c;
  const bool live = y < ny;
  const int yy = live ? y : 0;
  const Gio<float> gi(in, nbytes), go(out, nbytes);
  const uint32_t vo = ((uint32_t)yy * N + t) * 8u;
  // element x = t + m (m a multiple of TC, and so of the tile width) of a
  // tiled row sits m * ny elements after element t
  uint32_t vi = vo, es = 8u;
  if (tile) {
    const int ts = __builtin_ctz(tile);
    vi = (((((uint32_t)t >> ts) * (uint32_t)ny + (uint32_t)yy) << ts) + ((uint32_t)t & (tile - 1))) * 8u;
    es = (uint32_t)ny * 8u;
  }
  // The field's loads go out first: the material tables' copy into LDS
  // waits on its own loads, and would otherwise hold them back.
  cpx<float> v[F::E];
#pragma unroll
  for (int e = 0; e < F::E; ++e) v[e] = gi.ld(vi, first_off<F>(e, es));
  // MID with material lookups: the row's voxel ids are loaded with the field
  // rather than between the two transforms, where their latency would stall
  // the row (each element's load waited on its inside-the-volume test)
  uint32_t idw[IDM ? NW : 1];
  if constexpr (IDM != 0) {
    // byte loads through a buffer resource over the row's nxg ids: a column
    // outside the volume (its factor is not applied) may give an offset
    // outside the row, which reads 0 instead of faulting
    const RowT r0 = row_state<TM>(ta, yy);
    const __amdgpu_buffer_rsrc_t rid =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(r0.row0), (short)0, ta.nxg, 0x00020000);
    const uint32_t ib = (uint32_t)(ta.ax * t + ta.cx);
    uint32_t b[F::E];
#pragma unroll
    for (int e = 0; e < F::E; ++e)
      b[e] = __builtin_amdgcn_raw_buffer_load_b8(rid, ib + (uint32_t)(ta.ax * (last_pos<F>(t, e) - t)), 0, 0);
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      uint32_t w = 0;
#pragma unroll
      for (int q = 0; q < 4 && 4 * k + q < F::E; ++q) w |= b[4 * k + q] << (8 * q);
      idw[k] = w;
    }
    if constexpr (IDM == 2) {
#pragma unroll
      for (int k = 0; k < NW; ++k) lds_ids[k * F::TC * C + tid] = idw[k];
    }
  }
  if constexpr (MODE != ROW_LAST && LUT) table_to_lds(lds_lut, ta.lut);
  if constexpr (MODE != ROW_LAST && TM == TM_BILINEAR) table_to_lds(lds_dbf, ta.dbf);
  if constexpr (MODE == ROW_FIRST) {
    const RowT r = row_state<TM>(ta, yy);
    if constexpr (LUT || TM == TM_BILINEAR) __syncthreads();
#pragma unroll
    for (int e = 0; e < F::E; ++e) {
      v[e] = transmit<TM>(v[e], ta, first_pos<F>(t, e), r.y0, r.fy, r.iny, r.row0, lds_lut, lds_dbf);
      if ((e & 3) == 3) compiler_fence();
    }
    reg_fft<F, -1, C, PS, SPLIT>(v, tw, lds, t, c);
  } else if constexpr (MODE == ROW_MID || MODE == ROW_MID_OUT) {
    reg_fft<F, 1, C, PS, SPLIT>(v, tw, lds, t, c);  // its barriers publish the LDS tables
    compiler_fence();
    if constexpr (MODE == ROW_MID_OUT) {
      if (aj.fld_out && live) {   // uniform: one field per launch
        const Gio<float> gf(aj.fld_out, nbytes);
#pragma unroll
        for (int e = 0; e < F::E; ++e) gf.st(v[e], vo, last_off<F>(e, 8u));
      }
      if (aj.xb.xs) {   // uniform
        // |u|^2 of the block's rows through LDS (free between the two FFTs),
        // then each detector column's x-integral in the cells' order
        // four detector columns per thread and round; their plan entries are
        // loaded (clamped to the last column) before the LDS phase
        float* e2 = reinterpret_cast<float*>(lds);
        const int ox = aj.xb.ox;
        const int jj = y - aj.xb.j0;
        const bool rowin = live && jj >= 0 && jj < aj.xb.nj;
        int p0 = t;
        int i0[4], i1[4];
        double f0[4], f1[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int pc = min(p0 + k * F::TC, ox - 1);
          i0[k] = aj.xb.xi0[pc];
          i1[k] = aj.xb.xi1[pc];
          f0[k] = aj.xb.xf0[pc];
          f1[k] = aj.xb.xf1[pc];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < F::E; ++e)
          e2[c * N + last_pos<F>(t, e)] = v[e].x * v[e].x + v[e].y * v[e].y;
        __syncthreads();
        const float* r = e2 + c * N;
        const double dx = aj.xb.dx;
        while (rowin && p0 < ox) {
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int p = p0 + k * F::TC;
            if (p < ox) {
              double sum = 0.0;
              for (int i = i0[k]; i < i1[k]; ++i) sum += (double)r[i];
              double xi = sum * dx;
              if (i1[k] < N) xi += f1[k] * dx * (double)r[i1[k]];
              if (i0[k] < N) xi -= f0[k] * dx * (double)r[i0[k]];
              aj.xb.xs[(long long)jj * ox + p] = xi;
            }
          }
          p0 += 4 * F::TC;
          if (p0 < ox) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const int pc = min(p0 + k * F::TC, ox - 1);
              i0[k] = aj.xb.xi0[pc];
              i1[k] = aj.xb.xi1[pc];
              f0[k] = aj.xb.xf0[pc];
              f1[k] = aj.xb.xf1[pc];
            }
          }
        }
        __syncthreads();
      }
    }
    const RowT r = row_state<TM>(ta, yy);
    if constexpr (IDM == 2) {
#pragma unroll
      for (int k = 0; k < NW; ++k) idw[k] = lds_ids[k * F::TC * C + tid];
    }
#pragma unroll
    for (int e = 0; e < F::E; ++e) {
      if constexpr (IDM != 0) {
        const int x = last_pos<F>(t, e);
        if (r.iny && x >= ta.xlo && x < ta.xhi)
          v[e] = lut_mul(v[e], lds_lut[(idw[e / 4] >> (8 * (e % 4))) & 255u]);
      } else {
        v[e] = transmit<TM>(v[e], ta, last_pos<F>(t, e), r.y0, r.fy, r.iny, r.row0, lds_lut, lds_dbf);
      }
      if ((e & 3) == 3) compiler_fence();
    }
    reg_fft<F, -1, C, PS, SPLIT>(v, tw, lds, opaque_index<(LD != 0)>(t), opaque_index<(LD != 0)>(c));
  } else if constexpr (MODE == ROW_ADJ || MODE == ROW_ADJ_LAST) {
    // h = IFFT_x(W) is dL/du* after the transmission of slice s; with
    // w = T_s u_s: dL/d delta = 2K Im(conj(h) w), dL/d beta = -2K Re(conj(h) w),
    // and dL/du_s* = conj(T_s) h
    reg_fft<F, 1, C, PS, SPLIT>(v, tw, lds, t, c);
    compiler_fence();
    const Gio<float> gu(aj.fld_in, nbytes);
    const long long row = (long long)yy * ta.nxg;
    const float k2 = 2.f * ta.K;
#pragma unroll
    for (int e = 0; e < F::E; ++e) {
      const long long i = row + last_pos<F>(t, e);
      const cpx<float> tt = voxel_factor(ta.K * ta.dpl[i], ta.K * ta.bpl[i]);
      const cpx<float> us = gu.ld(vo, last_off<F>(e, 8u));
      const cpx<float> h = v[e];
      const float wr = us.x * tt.x - us.y * tt.y, wi = us.x * tt.y + us.y * tt.x;
      const float cr = h.x * wr + h.y * wi;
      const float ci = h.x * wi - h.y * wr;
      if (live) {
        aj.gd[i] = k2 * ci;
        aj.gb[i] = -(k2 * cr);
      }
      v[e] = mk(tt.x * h.x + tt.y * h.y, tt.x * h.y - tt.y * h.x);
      if ((e & 3) == 3) compiler_fence();
    }
    if constexpr (MODE == ROW_ADJ)
      reg_fft<F, -1, C, PS, SPLIT>(v, tw, lds, opaque_index<(LD != 0)>(t), opaque_index<(LD != 0)>(c));
  } else {
    reg_fft<F, 1, C, PS, SPLIT>(v, tw, lds, t, c);
  }
  if (live) {
#pragma unroll
    for (int e = 0; e < F::E; ++e) go.st(v[e], vo, last_off<F>(e, 8u));
  }
}

template <class F, int C, int PS, bool SPLIT, int LD, int SPEC, bool TILE, typename TW>
__global__ void __launch_bounds__(F::TC * C) ms_col_kernel(ColBatch cb, int nx, uint32_t nbytes,
                                                           StageTw<TW> tw, float spec_scale) {
  constexpr int N = F::N;
  __shared__ unsigned char lds[LdsSize<F, C, PS>::slots * (SPLIT ? 4 : 8)];
  const ColInst& I = cb.inst[blockIdx.y];
  const cpx<float>* in = I.in;
  cpx<float>* out = I.out;
  cpx<float>* spec = I.spec;
  const double2* __restrict__ Bx = I.Bx;
  const double2* __restrict__ Ay = I.Ay;
  const int* __restrict__ Rmax = I.Rmax;
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  const int g = xcd_tile(blockIdx.x, gridDim.x);
  const int x = g * C + c;
  const bool live = x < nx;
  const int xx = live ? x : 0;
  const Gio<float> gi(in, nbytes), go(out, nbytes), gs(spec, nbytes);
  const uint32_t row_b = (uint32_t)nx * 8u;
  const uint32_t vo = (uint32_t)t * row_b + (uint32_t)xx * 8u;
  cpx<float> v[F::E];
#pragma unroll
  for (int e = 0; e < F::E; ++e) v[e] = gi.ld(vo, first_off<F>(e, row_b));
  // the column's chirp factor and cutoff go out with the field
  const double2 bx = Bx[xx];
  const int rmax = Rmax[xx];
  reg_fft<F, -1, C, PS, SPLIT>(v, tw, lds, t, c);
  compiler_fence();
  constexpr int half = (N - 1) / 2 + 1;
  // groups of 4 with a compiler fence in between bound the number of FP64
  // table values in flight (register pressure)
#pragma unroll
  for (int e = 0; e < F::E; ++e) {
    const int y = last_pos<F>(t, e);
    const int sy = y < half ? y : y - N;
    const int ay = sy < 0 ? -sy : sy;
    const double2 a = Ay[ay];   // Ay[y] = Ay[ny - y]: a table half the size in the caches
    const double mr = bx.x * a.x - bx.y * a.y;
    const double mi = bx.x * a.y + bx.y * a.x;
    const double vr = (double)v[e].x, vi = (double)v[e].y;
    const cpx<float> w = mk((float)(vr * mr - vi * mi), (float)(vr * mi + vi * mr));
    v[e] = ay > rmax ? mk(0.f, 0.f) : w;
    if constexpr (SPEC) {
      if (live) gs.st(mk(v[e].x * spec_scale, v[e].y * spec_scale), vo, last_off<F>(e, row_b));
    }
    if ((e & 3) == 3) compiler_fence();
  }
  reg_fft<F, 1, C, PS, SPLIT>(v, tw, lds, opaque_index<(LD != 0)>(t), opaque_index<(LD != 0)>(c));
  if constexpr (TILE) {
    // tile g = this workgroup's C columns, rows in order: full 128 B lines.
    // The row stride of a tile comes from the launch (out_tile == C), so the
    // per-element offsets stay scalar as with the row-major stride.
    const uint32_t tb = (uint32_t)cb.out_tile * 8u;
    const uint32_t vt = ((uint32_t)g * N + (uint32_t)t) * tb + (uint32_t)c * 8u;
    if (live) {
#pragma unroll
      for (int e = 0; e < F::E; ++e) go.st(v[e], vt, last_off<F>(e, tb));
    }
  } else if (live) {
#pragma unroll
    for (int e = 0; e < F::E; ++e) go.st(v[e], vo, last_off<F>(e, row_b));
  }
}

// ---------------------------------------------------------------------------
// Launch-shape variants.  A variant fixes E (elements per thread) and C
// (sequences per workgroup) for one size; the radix schedule (and so the
// twiddle tables) is Sched<N>'s.
// ---------------------------------------------------------------------------
template <int N, int E>
struct FacE {
  using B = typename Sched<N>::F;
  using F = RegFac<N, E, B::R0, B::R1, B::R2, B::R3>;
};

// Exchange padding of a launch shape: C pad slots after every 2^shift slots
// (C sequences interleaved).  Lanes of the first exchange's writes sit
// R0 * C slots apart (R0 = the first radix), so 2^shift = 16 * C spreads them
// over the 32 banks of 4 bytes for radix-16 schedules.  The radix-32 16384
// schedule needs 32 * C (its split exchanges took 2 LDS cycles per access,
// row MID 2471 -> 2427 us).  The radix-8 first stage of the 2^a * 3 lengths
// is conflict-free with 8 * C for two sequences in 8-byte slots (3072 row
// 67.8 -> 63.7 us) and for four or more in 4-byte slots; the latter only
// where one workgroup per CU is all the LDS holds either way (6144 columns
// 305 -> 298 us): 3072's four-column buffer grows from 52 to 55 KiB, which
// leaves room for two workgroups instead of three (62 -> 75 us).  Chosen with
// a model of every exchange instruction's bank accesses; each choice keeps
// the exchange addresses linear in the element index (the per-element part
// folds into the instruction offset) wherever 16 * C did.
constexpr int exchange_shift(int n, int r0, int C, bool split) {
  const int lc = (C >= 2) + (C >= 4) + (C >= 8) + (C >= 16);
  if (r0 >= 32) return 5 + lc;
  if (r0 == 8 && C == 2 && !split) return 3 + lc;
  if (r0 == 8 && C >= 4 && split && (long long)n * C * 4 > 81920) return 3 + lc;
  return 4 + lc;
}

template <int N, int E, int C, bool SPLIT_ = ((size_t)N * C * 8 > 65536), int LD_ = 0>
struct Shape {
  static constexpr int LD = LD_;
  using F = typename FacE<N, E>::F;
  static constexpr int CC = C;
  static constexpr bool SPLIT = SPLIT_;
  static constexpr int PS = pad_code(exchange_shift(N, F::R0, C, SPLIT_), C);
  static constexpr int threads = F::TC * C;
};

template <class S, typename TW>
struct RowLaunch {
  template <int MODE, int TM>
  static bool launch(const RowBatch& b, int count, int ny, uint32_t nb, const StageTw<TW>& tw,
                     hipStream_t s, const AdjArgs& aj) {
    using F = typename S::F;
    constexpr int C = S::CC;
    if (b.in_tile && (F::TC % b.in_tile != 0 || (b.in_tile & (b.in_tile - 1)) != 0)) return false;
    dim3 grid((ny + C - 1) / C, count), block(S::threads);
    hipLaunchKernelGGL((ms_row_kernel<F, C, S::PS, S::SPLIT, S::LD, MODE, TM, TW>), grid, block, 0, s, b, ny, nb, tw, aj);
    return true;
  }
  template <int MODE>
  static bool go_tm(int tm, const RowBatch& b, int count, int ny, uint32_t nb,
                    const StageTw<TW>& tw, hipStream_t s, const AdjArgs& aj) {
    switch (tm) {
      case TM_LUT: return launch<MODE, TM_LUT>(b, count, ny, nb, tw, s, aj);
      case TM_FIELD: return launch<MODE, TM_FIELD>(b, count, ny, nb, tw, s, aj);
      case TM_BILINEAR: return launch<MODE, TM_BILINEAR>(b, count, ny, nb, tw, s, aj);
      case TM_PLANAR: return launch<MODE, TM_PLANAR>(b, count, ny, nb, tw, s, aj);
      default: return false;
    }
  }
  static bool go(int mode, int tm, const RowBatch& b, int count, int ny, uint32_t nb,
                 const StageTw<TW>& tw, hipStream_t s, const AdjArgs& aj) {
    switch (mode) {
      case ROW_FIRST:
        if (tm == TM_NONE) return launch<ROW_FIRST, TM_NONE>(b, count, ny, nb, tw, s, aj);
        return go_tm<ROW_FIRST>(tm, b, count, ny, nb, tw, s, aj);
      case ROW_MID:
        if (aj.fld_out || aj.xb.xs) return go_tm<ROW_MID_OUT>(tm, b, count, ny, nb, tw, s, aj);
        return go_tm<ROW_MID>(tm, b, count, ny, nb, tw, s, aj);
      // no transmission in the last pass: one instantiation serves every source
      case ROW_LAST: return launch<ROW_LAST, TM_LUT>(b, count, ny, nb, tw, s, aj);
      case ROW_ADJ:
        return tm == TM_PLANAR && launch<ROW_ADJ, TM_PLANAR>(b, count, ny, nb, tw, s, aj);
      case ROW_ADJ_LAST:
        return tm == TM_PLANAR && launch<ROW_ADJ_LAST, TM_PLANAR>(b, count, ny, nb, tw, s, aj);
      default: return false;
    }
  }
};

template <class S, typename TW>
struct ColLaunch {
  static bool go(bool spec_out, const ColBatch& b, int count, int nx, uint32_t nb,
                 const StageTw<TW>& tw, float spec_scale, hipStream_t s) {
    using F = typename S::F;
    constexpr int C = S::CC;
    dim3 grid((nx + C - 1) / C, count), block(S::threads);
    if (b.out_tile) {
      // whole tiles only; the spectrum output is row-major
      if (b.out_tile != C || nx % C != 0 || spec_out) return false;
      hipLaunchKernelGGL((ms_col_kernel<F, C, S::PS, S::SPLIT, S::LD, 0, true, TW>), grid, block, 0, s, b, nx, nb, tw, spec_scale);
    } else if (spec_out) {
      hipLaunchKernelGGL((ms_col_kernel<F, C, S::PS, S::SPLIT, S::LD, 1, false, TW>), grid, block, 0, s, b, nx, nb, tw, spec_scale);
    } else {
      hipLaunchKernelGGL((ms_col_kernel<F, C, S::PS, S::SPLIT, S::LD, 0, false, TW>), grid, block, 0, s, b, nx, nb, tw, spec_scale);
    }
    return true;
  }
};

// Variant lists: X(N, id, E, C, SPLIT).  Variant 0 is the default.  E1 is
// the schedule's own element count (16, 24 for 2^a * 3 lengths, 32 for
// 16384), E2 = 2*E1 (E1 for 16384).
template <int N> constexpr int vE1() { return Sched<N>::F::E; }
// (schedules of 32 elements per thread have no wider variant: 64 complex
// values per thread leave no registers for the butterflies)
template <int N> constexpr int vE2() { return Sched<N>::F::E >= 32 ? Sched<N>::F::E : 2 * Sched<N>::F::E; }
constexpr int vmax1(int v) { return v > 0 ? v : 1; }
constexpr int vmin(int a, int b) { return a < b ? a : b; }
constexpr bool vsplit(int n, int c, int cap) { return (long long)n * c * 8 > cap; }
constexpr int vpow2floor(int v) { return v >= 16 ? 16 : v >= 8 ? 8 : v >= 4 ? 4 : v >= 2 ? 2 : 1; }
// most sequences whose split (4-byte) exchange buffer fits LDS beside the LUT
constexpr int vLdsC(int n) { return vpow2floor(vmax1((163840 - 8192) / (n * 4))); }
template <int N> constexpr int vRowC(int E, int t) { return vmin(vmin(vmax1(t / (N / E)), 16), vLdsC(N)); }
template <int N> constexpr int vColC(int E, int t) { return vmin(vmin(vmax1(t / (N / E)), 16), vLdsC(N)); }
// X(N, id, E, C, SPLIT, LD).  The base lists keep the index of the second
// transform of a kernel as it is (LD 0); the LD lists are copies of some
// base shapes whose second transform recomputes its addresses
// (opaque_index; LD 1, or LD 2 / LD 3 / LD 4 with the three- / two-workgroup
// or four-waves register budget of row_waves_per_eu).  The two sets are compiled in separate translation units
// (rsim_fused_<N>.hip and rsim_fused_<N>_ld.hip): kernels of the other kind
// in the same module change the code generated for the base kernels.
// Variant ids: base 0..kRowBase-1, then the LD list.
#define RSIM_ROW_VARIANTS(N, X)                                                                 \
  X(N, 0, vE1<N>(), vRowC<N>(vE1<N>(), 256), vsplit(N, vRowC<N>(vE1<N>(), 256), 65536), 0)   \
  X(N, 1, vE1<N>(), vRowC<N>(vE1<N>(), 512), vsplit(N, vRowC<N>(vE1<N>(), 512), 65536), 0)   \
  X(N, 2, vE1<N>(), vRowC<N>(vE1<N>(), 256), true, 0)                                       \
  X(N, 3, vE2<N>(), vRowC<N>(vE2<N>(), 256), vsplit(N, vRowC<N>(vE2<N>(), 256), 65536), 0)
#define RSIM_ROW_VARIANTS_LD(N, X)                                                              \
  X(N, 0, vE1<N>(), vRowC<N>(vE1<N>(), 256), vsplit(N, vRowC<N>(vE1<N>(), 256), 65536), 1)    \
  X(N, 1, vE1<N>(), vRowC<N>(vE1<N>(), 512), vsplit(N, vRowC<N>(vE1<N>(), 512), 65536), 1)    \
  X(N, 2, vE2<N>(), vRowC<N>(vE2<N>(), 256), vsplit(N, vRowC<N>(vE2<N>(), 256), 65536), 1)    \
  X(N, 3, vE1<N>(), vRowC<N>(vE1<N>(), 512), true, 2)                                        \
  X(N, 4, vE2<N>(), vRowC<N>(vE2<N>(), 512), true, 3)                                        \
  X(N, 5, vE1<N>(), vRowC<N>(vE1<N>(), 256), true, 4)
#define RSIM_COL_VARIANTS(N, X)                                                                 \
  X(N, 0, vE1<N>(), vColC<N>(vE1<N>(), 1024), true, 0)                                      \
  X(N, 1, vE1<N>(), vColC<N>(vE1<N>(), 1024), vsplit(N, vColC<N>(vE1<N>(), 1024), 163840 - 8192), 0) \
  X(N, 2, vE2<N>(), vColC<N>(vE2<N>(), 512), true, 0)                                       \
  X(N, 3, vE1<N>(), vmax1(vColC<N>(vE1<N>(), 1024) / 2), true, 0)                           \
  X(N, 4, vE2<N>(), vColC<N>(vE2<N>(), 1024), true, 0)
#define RSIM_COL_VARIANTS_LD(N, X)                                                              \
  X(N, 0, vE1<N>(), vColC<N>(vE1<N>(), 1024), true, 1)                                       \
  X(N, 1, vE1<N>(), vColC<N>(vE1<N>(), 1024), vsplit(N, vColC<N>(vE1<N>(), 1024), 163840 - 8192), 1) \
  X(N, 2, vE2<N>(), vColC<N>(vE2<N>(), 512), true, 1)                                       \
  X(N, 3, vE2<N>(), vColC<N>(vE2<N>(), 1024), true, 1)
constexpr int kRowBase = 4;
constexpr int kColBase = 5;
constexpr int kRowVariants = kRowBase + 6;
constexpr int kColVariants = kColBase + 4;

// Per-size launchers (explicitly instantiated in rsim_fused_<N>.hip and
// rsim_fused_<N>_ld.hip); count instances (1..kMaxBatch) of shape (N, ny) or
// (nx, N) in one launch.
template <int N, typename TW>
bool row_go(int v, int mode, int tm, const RowBatch& b, int count, int ny,
            const StageTw<TW>& tw, hipStream_t s, const AdjArgs& aj = AdjArgs{});
template <int N, typename TW>
bool col_go(int v, bool spec_out, const ColBatch& b, int count, int nx, const StageTw<TW>& tw,
            float ss, hipStream_t s);
template <int N, typename TW>
bool row_go_ld(int v, int mode, int tm, const RowBatch& b, int count, int ny,
               const StageTw<TW>& tw, hipStream_t s, const AdjArgs& aj);
template <int N, typename TW>
bool col_go_ld(int v, bool spec_out, const ColBatch& b, int count, int nx, const StageTw<TW>& tw,
               float ss, hipStream_t s);

#define RSIM_ROW_CASE(N, V, E, C, SP, LD) \
  case V:                                \
    return RowLaunch<Shape<N, E, C, SP, LD>, TW>::go(mode, tm, b, count, ny, nb, tw, s, aj);
#define RSIM_COL_CASE(N, V, E, C, SP, LD) \
  case V:                                \
    return ColLaunch<Shape<N, E, C, SP, LD>, TW>::go(spec_out, b, count, nx, nb, tw, ss, s);

// Columns per workgroup (= the tile width of a tiled output) of column
// variant v of length N.
#define RSIM_COL_WIDTH_CASE(N, V, E, C, SP, LD) \
  case V:                                      \
    return C;
template <int N>
constexpr int col_width(int v) {
  if (v >= kColBase) {
    switch (v - kColBase) {
      RSIM_COL_VARIANTS_LD(N, RSIM_COL_WIDTH_CASE)
      default: return 0;
    }
  }
  switch (v) {
    RSIM_COL_VARIANTS(N, RSIM_COL_WIDTH_CASE)
    default: return 0;
  }
}

// Launcher over one variant list (NAME: row_go or row_go_ld); the base one
// hands the LD ids to row_go_ld (when HAS_LD).
#define RSIM_ROW_LAUNCHER(NAME, N, LIST, HAS_LD)                                             \
  template <int NN, typename TW>                                                            \
  bool NAME##_impl(int v, int mode, int tm, const RowBatch& b, int count, int ny,           \
                   const StageTw<TW>& tw, hipStream_t s, const AdjArgs& aj) {               \
    if (HAS_LD && v >= kRowBase)                                                            \
      return row_go_ld<NN, TW>(v - kRowBase, mode, tm, b, count, ny, tw, s, aj);            \
    const uint32_t nb = (uint32_t)((long long)NN * ny * 8);                                 \
    switch (v) {                                                                            \
      LIST(NN, RSIM_ROW_CASE)                                                               \
      default: return false;                                                                \
    }                                                                                       \
  }                                                                                         \
  template <>                                                                               \
  bool NAME<N, float>(int v, int mode, int tm, const RowBatch& b, int count, int ny,        \
                      const StageTw<float>& tw, hipStream_t s, const AdjArgs& aj) {         \
    return NAME##_impl<N, float>(v, mode, tm, b, count, ny, tw, s, aj);                     \
  }                                                                                         \
  template <>                                                                               \
  bool NAME<N, double>(int v, int mode, int tm, const RowBatch& b, int count, int ny,       \
                       const StageTw<double>& tw, hipStream_t s, const AdjArgs& aj) {       \
    return NAME##_impl<N, double>(v, mode, tm, b, count, ny, tw, s, aj);                    \
  }                                                                                         \
  template <>                                                                               \
  bool NAME<N, twc>(int v, int mode, int tm, const RowBatch& b, int count, int ny,          \
                    const StageTw<twc>& tw, hipStream_t s, const AdjArgs& aj) {             \
    return NAME##_impl<N, twc>(v, mode, tm, b, count, ny, tw, s, aj);                       \
  }

#define RSIM_COL_LAUNCHER(NAME, N, LIST, HAS_LD)                                             \
  template <int NN, typename TW>                                                            \
  bool NAME##_impl(int v, bool spec_out, const ColBatch& b, int count, int nx,              \
                   const StageTw<TW>& tw, float ss, hipStream_t s) {                        \
    if (HAS_LD && v >= kColBase)                                                            \
      return col_go_ld<NN, TW>(v - kColBase, spec_out, b, count, nx, tw, ss, s);            \
    const uint32_t nb = (uint32_t)((long long)NN * nx * 8);                                 \
    switch (v) {                                                                            \
      LIST(NN, RSIM_COL_CASE)                                                               \
      default: return false;                                                                \
    }                                                                                       \
  }                                                                                         \
  template <>                                                                               \
  bool NAME<N, float>(int v, bool spec_out, const ColBatch& b, int count, int nx,           \
                      const StageTw<float>& tw, float ss, hipStream_t s) {                  \
    return NAME##_impl<N, float>(v, spec_out, b, count, nx, tw, ss, s);                     \
  }                                                                                         \
  template <>                                                                               \
  bool NAME<N, double>(int v, bool spec_out, const ColBatch& b, int count, int nx,          \
                       const StageTw<double>& tw, float ss, hipStream_t s) {                \
    return NAME##_impl<N, double>(v, spec_out, b, count, nx, tw, ss, s);                    \
  }                                                                                         \
  template <>                                                                               \
  bool NAME<N, twc>(int v, bool spec_out, const ColBatch& b, int count, int nx,             \
                    const StageTw<twc>& tw, float ss, hipStream_t s) {                      \
    return NAME##_impl<N, twc>(v, spec_out, b, count, nx, tw, ss, s);                       \
  }

#define RSIM_FUSED_INSTANTIATE_ROWS(N) RSIM_ROW_LAUNCHER(row_go, N, RSIM_ROW_VARIANTS, true)
#define RSIM_FUSED_INSTANTIATE_COLS(N) RSIM_COL_LAUNCHER(col_go, N, RSIM_COL_VARIANTS, true)
#define RSIM_FUSED_INSTANTIATE(N) RSIM_FUSED_INSTANTIATE_ROWS(N) RSIM_FUSED_INSTANTIATE_COLS(N)
#define RSIM_FUSED_INSTANTIATE_ROWS_LD(N) RSIM_ROW_LAUNCHER(row_go_ld, N, RSIM_ROW_VARIANTS_LD, false)
#define RSIM_FUSED_INSTANTIATE_COLS_LD(N) RSIM_COL_LAUNCHER(col_go_ld, N, RSIM_COL_VARIANTS_LD, false)
#define RSIM_FUSED_INSTANTIATE_LD(N) RSIM_FUSED_INSTANTIATE_ROWS_LD(N) RSIM_FUSED_INSTANTIATE_COLS_LD(N)

// ---------------------------------------------------------------------------
// Three-pass column schedule for ny = N1 * N2 beyond one LDS pass
// (y = N2*n1 + n2), each pass on tiles of C neighbouring columns:
//   A: FFT over n1 (rows N2 apart) for fixed n2, times W_ny^(n2 k1), in place
//   B: for fixed k1 (N2 contiguous rows): FFT over n2 -> k2, times the chirp
//      and cutoff at ky = k1 + N1 k2, IFFT over k2 -> n2, times W_ny^-(n2 k1)
//   C: IFFT over k1 (rows N2 apart) for fixed n2, in place: natural order
// The sub-transforms (Sched<128>) start and end with the same radix, so B
// chains FFT -> chirp -> IFFT in registers.  Every row segment a wavefront
// touches is C * 8 contiguous bytes.
// ---------------------------------------------------------------------------
// The A and B twiddles of an element are W and conj(W) of the same index in
// every slice, so an FP32-rounded |W| != 1 would scale that element
// identically in each slice; FP32 arithmetic modes therefore take the hi + lo
// table for them (8 twiddles per thread and pass).
template <typename TW>
using BigTwElem = typename std::conditional<std::is_same<TW, double>::value, cpx<double>, twc>::type;

template <int DIR>
__device__ __forceinline__ cpx<float> big_tw(cpx<float> v, const twc* T, int i) {
  return twmul_comp<DIR>(v, T[i]);
}
template <int DIR>
__device__ __forceinline__ cpx<float> big_tw(cpx<float> v, const cpx<float>* T, int i) {
  return twmul<DIR>(v, T[i]);
}
template <int DIR>
__device__ __forceinline__ cpx<float> big_tw(cpx<float> v, const cpx<double>* T, int i) {
  const cpx<double> w = twmul<DIR>(mk((double)v.x, (double)v.y), T[i]);
  return mk((float)w.x, (float)w.y);
}

// Passes A (DIR = -1, with the W_ny twiddle) and C (DIR = +1).  grid.y = n2.
template <class F, int C, int PS, bool SPLIT, int DIR, typename TW>
__global__ void __launch_bounds__(F::TC * C) ms_colAC_kernel(cpx<float>* buf, int nx, int N2,
                                                             uint32_t nbytes, StageTw<TW> tw,
                                                             const BigTwElem<TW>* __restrict__ wN) {
  __shared__ unsigned char lds[LdsSize<F, C, PS>::slots * (SPLIT ? 4 : 8)];
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  const int x = xcd_tile(blockIdx.x, gridDim.x) * C + c;
  const bool live = x < nx;
  const int xx = live ? x : 0;
  const int n2 = blockIdx.y;
  const Gio<float> g(buf, nbytes);
  const uint32_t row_b = (uint32_t)nx * 8u;
  const uint32_t es_b = (uint32_t)N2 * row_b;
  const uint32_t vo = ((uint32_t)N2 * (uint32_t)t + (uint32_t)n2) * row_b + (uint32_t)xx * 8u;
  cpx<float> v[F::E];
#pragma unroll
  for (int e = 0; e < F::E; ++e) v[e] = g.ld(vo, first_off<F>(e, es_b));
  reg_fft<F, DIR, C, PS, SPLIT>(v, tw, lds, t, c);
  if constexpr (DIR < 0) {
#pragma unroll
    for (int e = 0; e < F::E; ++e) v[e] = big_tw<-1>(v[e], wN, n2 * last_pos<F>(t, e));
  }
  if (live) {
#pragma unroll
    for (int e = 0; e < F::E; ++e) g.st(v[e], vo, last_off<F>(e, es_b));
  }
}

// Pass B.  grid.y = k1; ny = N1 * F::N.
template <class F, int C, int PS, bool SPLIT, typename TW>
__global__ void __launch_bounds__(F::TC * C) ms_colB_kernel(cpx<float>* buf, int nx, int N1,
                                                            uint32_t nbytes, StageTw<TW> tw,
                                                            const BigTwElem<TW>* __restrict__ wN,
                                                            const double2* __restrict__ Bx,
                                                            const double2* __restrict__ Ay,
                                                            const int* __restrict__ Rmax) {
  constexpr int N2 = F::N;
  __shared__ unsigned char lds[LdsSize<F, C, PS>::slots * (SPLIT ? 4 : 8)];
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  const int x = xcd_tile(blockIdx.x, gridDim.x) * C + c;
  const bool live = x < nx;
  const int xx = live ? x : 0;
  const int k1 = blockIdx.y;
  const int ny = N1 * N2;
  const Gio<float> g(buf, nbytes);
  const uint32_t row_b = (uint32_t)nx * 8u;
  const uint32_t vo = ((uint32_t)N2 * (uint32_t)k1 + (uint32_t)t) * row_b + (uint32_t)xx * 8u;
  cpx<float> v[F::E];
#pragma unroll
  for (int e = 0; e < F::E; ++e) v[e] = g.ld(vo, first_off<F>(e, row_b));
  reg_fft<F, -1, C, PS, SPLIT>(v, tw, lds, t, c);
  compiler_fence();
  const double2 bx = Bx[xx];
  const int rmax = Rmax[xx];
  const int half = (ny - 1) / 2 + 1;
#pragma unroll
  for (int e = 0; e < F::E; ++e) {
    const int y = k1 + N1 * last_pos<F>(t, e);
    const int sy = y < half ? y : y - ny;
    const int ay = sy < 0 ? -sy : sy;
    const double2 a = Ay[y];
    const double mr = bx.x * a.x - bx.y * a.y;
    const double mi = bx.x * a.y + bx.y * a.x;
    const double vr = (double)v[e].x, vi = (double)v[e].y;
    const cpx<float> w = mk((float)(vr * mr - vi * mi), (float)(vr * mi + vi * mr));
    v[e] = ay > rmax ? mk(0.f, 0.f) : w;
    if ((e & 3) == 3) compiler_fence();
  }
  reg_fft<F, 1, C, PS, SPLIT>(v, tw, lds, t, c);
#pragma unroll
  for (int e = 0; e < F::E; ++e) v[e] = big_tw<1>(v[e], wN, k1 * last_pos<F>(t, e));
  if (live) {
#pragma unroll
    for (int e = 0; e < F::E; ++e) g.st(v[e], vo, last_off<F>(e, row_b));
  }
}

// Launcher of the three passes on buf (in place) in launch shape v (tiles of
// 16, 32 or 64 columns: 128-, 256- or 512-byte row segments); false if ny has
// no such schedule or v no such shape.
constexpr int kBigColVariants = 3;
template <typename TW>
bool big_col_go(int v, void* buf, int nx, int ny, const StageTw<TW>& tw128,
                const BigTwElem<TW>* wN, const double2* Bx, const double2* Ay, const int* Rmax,
                hipStream_t s);

}  // namespace rsim
