// MI355X (gfx950) element-wise kernels of the MFT path: operand packing,
// pupil phasors, detector accumulation, Field products and fills.
#include "mftx_device.h"

namespace mftx {

// --------------------------------------------------------------------------
// operand packing
// --------------------------------------------------------------------------

// dtype codes: 0 complex64, 1 float32, 2 complex128, 3 float64
__device__ __forceinline__ float2 load_elem(const void* p, int dtype, long long idx) {
  switch (dtype) {
    case 0: return ((const float2*)p)[idx];
    case 1: return make_float2(((const float*)p)[idx], 0.f);
    case 2: {
      const double2 v = ((const double2*)p)[idx];
      return make_float2((float)v.x, (float)v.y);
    }
    default: return make_float2((float)((const double*)p)[idx], 0.f);
  }
}

__device__ __forceinline__ void store_packed16(int prec, void* dst, long long lda, long long row, int kb,
                                               const float2 (&v)[16]) {
  if (prec == PREC_BF16X3) {
    uint32_t rh[8], rl[8], ih[8], il[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      split2(v[2 * i].x, v[2 * i + 1].x, rh[i], rl[i]);
      split2(v[2 * i].y, v[2 * i + 1].y, ih[i], il[i]);
    }
    uint4* d = (uint4*)((__bf16*)dst + row * lda + (long long)kb * 64);
    d[0] = make_uint4(rh[0], rh[1], rh[2], rh[3]);
    d[1] = make_uint4(rh[4], rh[5], rh[6], rh[7]);
    d[2] = make_uint4(rl[0], rl[1], rl[2], rl[3]);
    d[3] = make_uint4(rl[4], rl[5], rl[6], rl[7]);
    d[4] = make_uint4(ih[0], ih[1], ih[2], ih[3]);
    d[5] = make_uint4(ih[4], ih[5], ih[6], ih[7]);
    d[6] = make_uint4(il[0], il[1], il[2], il[3]);
    d[7] = make_uint4(il[4], il[5], il[6], il[7]);
  } else {
    float4* d = (float4*)((float2*)dst + row * lda + (long long)kb * 16);
#pragma unroll
    for (int i = 0; i < 8; ++i) d[i] = make_float4(v[2 * i].x, v[2 * i].y, v[2 * i + 1].x, v[2 * i + 1].y);
  }
}

// One thread per (row, 16-wide K block) of the packed destination.  The
// source is m x n with row stride lds; packed element (row, k) is source
// (row, k), or source (k, row) when trans (the packed operand is then n x m),
// times the folds cis(fold * (x0 + k) + fold2 * (x02 + row)).  Everything
// outside the source is zero.
__global__ void pack_kernel(int prec, const void* __restrict__ src, int dtype, long long m, long long n,
                            long long lds, int trans, double fold, double x0, double fold2, double x02,
                            void* __restrict__ dst, long long rows_pad, int kpad) {
  const int kblocks = kpad / 16;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows_pad * kblocks) return;
  const long long row = t / kblocks;
  const int kb = (int)(t % kblocks);
  const bool folded = fold != 0.0 || fold2 != 0.0;
  const double row_phase = fold2 * (x02 + (double)row);
  float2 v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const long long k = (long long)kb * 16 + i;
    const long long r = trans ? k : row, c = trans ? row : k;
    if (r < m && c < n) {
      const float2 e = load_elem(src, dtype, r * lds + c);
      v[i] = folded ? cmul(e, cis_turns(__builtin_fma(fold, x0 + (double)k, row_phase))) : e;
    } else {
      v[i] = make_float2(0.f, 0.f);
    }
  }
  store_packed16(prec, dst, packed_lda(prec, kpad), row, kb, v);
}

hipError_t launch_pack(int prec, const void* src, int src_dtype, long long m, long long n, long long lds,
                       int trans, double fold, double x0, double fold2, double x02, void* dst, long long rows_pad,
                       int kpad, hipStream_t s) {
  const long long total = rows_pad * (kpad / 16);
  if (total <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, prec, src,
                     src_dtype, m, n, lds, trans, fold, x0, fold2, x02, dst, rows_pad, kpad);
  return hipGetLastError();
}

// Index of the op that block b0 (its first thread) belongs to: ops start on
// block boundaries, so the search runs on block-uniform values (scalar loads).
template <typename Op>
__device__ __forceinline__ int block_op(const Op* __restrict__ ops, int nops, long long b0) {
  int lo = 0, hi = nops - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ops[mid].begin <= b0) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Grouped form of pack_kernel with one thread per (op, row, pair of K
// values): lanes run along K (along the packed row) for a row-major source and
// along the packed rows (along the source row) for a transposed one, so the
// source reads of a wave are contiguous; each thread writes 4 bytes of every
// BF16 plane (or one 16-byte complex64 pair).
__global__ void pack_group_kernel(int prec, const PackOp* __restrict__ ops, int nops, long long total) {
  const long long b0 = (long long)blockIdx.x * kGroupBlock;
  const PackOp& op = ops[block_op(ops, nops, b0)];
  const long long u = b0 + threadIdx.x - op.begin;
  const int npairs = op.kpad / 2;
  if (u >= op.rows_pad * npairs) return;
  const long long row = op.trans ? u % op.rows_pad : u / npairs;
  const int kp = (int)(op.trans ? u / op.rows_pad : u % npairs);
  const bool folded = op.fold != 0.0 || op.fold2 != 0.0;
  const double row_phase = op.fold2 * (op.x02 + (double)row);
  float2 v[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const long long k = 2LL * kp + i;
    const long long r = op.trans ? k : row, c = op.trans ? row : k;
    if (r < op.m && c < op.n) {
      const float2 e = load_elem(op.src, op.dtype, r * op.lds + c);
      v[i] = folded ? cmul(e, cis_turns(__builtin_fma(op.fold, op.x0 + (double)k, row_phase))) : e;
    } else {
      v[i] = make_float2(0.f, 0.f);
    }
  }
  const long long lda = packed_lda(prec, op.kpad);
  const int k0 = 2 * kp;
  if (prec == PREC_BF16X3) {
    uint32_t rh, rl, ih, il;
    split2(v[0].x, v[1].x, rh, rl);
    split2(v[0].y, v[1].y, ih, il);
    uint32_t* d = (uint32_t*)((__bf16*)op.dst + row * lda + (long long)(k0 / 16) * 64 + (k0 % 16));
    d[0] = rh;    // plane 0 (re hi) at +0, plane p at + p * 16 bf16 = 8 words
    d[8] = rl;
    d[16] = ih;
    d[24] = il;
  } else {
    *(float4*)((float2*)op.dst + row * lda + k0) = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
  }
}

hipError_t launch_pack_group(int prec, const PackOp* ops, int nops, long long total, hipStream_t s) {
  if (total <= 0 || nops <= 0) return hipSuccess;
  hipLaunchKernelGGL(pack_group_kernel, dim3((unsigned)((total + kGroupBlock - 1) / kGroupBlock)), dim3(kGroupBlock),
                     0, s, prec, ops, nops, total);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// pupil phasor  f = amp * exp(2*pi*i*opd/lambda)   (plane.py:507)
// --------------------------------------------------------------------------

__device__ __forceinline__ double load_real(const void* p, int dtype, long long idx) {
  return dtype == 3 ? ((const double*)p)[idx] : (double)((const float*)p)[idx];
}

__device__ __forceinline__ PhasorWave load_wave(cptr<const PhasorWave> p) {
  PhasorWave w;
  w.inv_wl = p->inv_wl;
  w.fold = p->fold;
  w.x0 = p->x0;
  w.fold2 = p->fold2;
  w.x02 = p->x02;
  return w;
}

// One thread per (op, row, half of a 16-wide K block, group of PH_WAVES
// wavelengths), every phasor operand of a plan in one launch (the op is found
// by a block-uniform search).  The 8 amp/opd values are read once and the
// packed operand of the m x n tile (n x m when trans) is written for each
// wavelength of the group, with that wavelength's stage-1 fold along K and
// stage-2 fold along the rows r in the same phase:
// amp * cis(opd * inv_wl + fold * (x0 + k) + fold2 * (x02 + r)).  Consecutive threads
// take consecutive source columns (half blocks when not trans, rows of the
// transposed operand when trans), so the source reads of a wave are
// contiguous; the two halves of a block store one 32-byte plane together.
__global__ void phasor_group_kernel(int prec, const PhasorOp* __restrict__ ops, int nops, long long total) {
  const long long b0 = (long long)blockIdx.x * kGroupBlock;
  const PhasorOp& op = ops[block_op(ops, nops, b0)];
  const long long u = b0 + threadIdx.x - op.begin;
  if (u >= phasor_threads(op.rows_pad, op.kpad, op.nwave)) return;
  const int khalves = op.kpad / 8;
  const long long per_group = op.rows_pad * khalves;
  // the wavelength group is wave-uniform: rows_pad is a multiple of 128 and
  // op.begin of kGroupBlock, so a group spans whole waves; its descriptors
  // are then scalar loads
  const int g = __builtin_amdgcn_readfirstlane((int)(u / per_group));
  const long long r = u % per_group;
  const long long row = op.trans ? r % op.rows_pad : r / khalves;
  const int kh = (int)(op.trans ? r / op.rows_pad : r % khalves);
  const int kb = kh >> 1, i0 = (kh & 1) * 8;
  // Every load is issued before any math and none sits behind a branch, so
  // they are in flight together: an element outside the m x n tile reads
  // element 0 and is zeroed afterwards, the dtypes are tested once, and the
  // thread's wavelength descriptors are read up front.
  long long ia[8], io[8];
  bool in[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const long long k = (long long)kb * 16 + i0 + i;
    const long long sr = op.trans ? k : row, sc = op.trans ? row : k;
    in[i] = sr < op.m && sc < op.n;
    ia[i] = in[i] ? sr * op.ld_amp + sc : 0;
    io[i] = in[i] ? sr * op.ld_opd + sc : 0;
  }
  float a[8];
  double o[8];
  if (op.amp_dtype == 3) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = (float)((gptr<const double>)op.amp)[ia[i]];
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = ((gptr<const float>)op.amp)[ia[i]];
  }
  if (op.opd_dtype == 3) {
#pragma unroll
    for (int i = 0; i < 8; ++i) o[i] = ((gptr<const double>)op.opd)[io[i]];
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) o[i] = (double)((gptr<const float>)op.opd)[io[i]];
  }
  const int w0 = g * PH_WAVES, nw = min(op.nwave - w0, PH_WAVES);
  PhasorWave pws[PH_WAVES];
#pragma unroll
  for (int j = 0; j < PH_WAVES; ++j) pws[j] = load_wave((cptr<const PhasorWave>)op.waves + w0 + min(j, nw - 1));
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[i] = in[i] ? a[i] : 0.f;
    o[i] = in[i] ? o[i] : 0.0;
  }
  const long long lda = packed_lda(prec, op.kpad);
  // The precision is tested once, outside the wavelength loop, so each
  // loop is straight-line code whose stores are never waited for.
  auto store_all = [&](auto bf16x3) {
#pragma unroll
    for (int j = 0; j < PH_WAVES; ++j) {
      if (j >= nw) break;
      const PhasorWave pw = pws[j];
      // both folds' phase at K = 16 kb of this row: the stage-1 fold along
      // K, the stage-2 fold along the rows
      const double f0 = __builtin_fma(pw.fold, pw.x0 + (double)kb * 16, pw.fold2 * (pw.x02 + (double)row));
      float2 v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const float2 c = cis_turns(__builtin_fma(o[i], pw.inv_wl, __builtin_fma(pw.fold, (double)(i0 + i), f0)));
        v[i] = make_float2(a[i] * c.x, a[i] * c.y);
      }
      char* dst = (char*)op.dst + (long long)(w0 + j) * op.wave_stride;
      if constexpr (decltype(bf16x3)::value) {
        uint32_t rh[4], rl[4], ih[4], il[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          split2(v[2 * i].x, v[2 * i + 1].x, rh[i], rl[i]);
          split2(v[2 * i].y, v[2 * i + 1].y, ih[i], il[i]);
        }
        gptr<v4u32> d = (gptr<v4u32>)((__bf16*)dst + row * lda + (long long)kb * 64 + i0);  // plane p at + 16 p bf16
        d[0] = v4u32{rh[0], rh[1], rh[2], rh[3]};
        d[2] = v4u32{rl[0], rl[1], rl[2], rl[3]};
        d[4] = v4u32{ih[0], ih[1], ih[2], ih[3]};
        d[6] = v4u32{il[0], il[1], il[2], il[3]};
      } else {
        gptr<v4f32> d = (gptr<v4f32>)((float2*)dst + row * lda + (long long)kb * 16 + i0);
#pragma unroll
        for (int i = 0; i < 4; ++i) d[i] = v4f32{v[2 * i].x, v[2 * i].y, v[2 * i + 1].x, v[2 * i + 1].y};
      }
    }
  };
  if (prec == PREC_BF16X3)
    store_all(std::true_type{});
  else
    store_all(std::false_type{});
}

hipError_t launch_phasor_group(int prec, const PhasorOp* ops, int nops, long long total, hipStream_t s) {
  if (total <= 0 || nops <= 0) return hipSuccess;
  hipLaunchKernelGGL(phasor_group_kernel, dim3((unsigned)((total + kGroupBlock - 1) / kGroupBlock)),
                     dim3(kGroupBlock), 0, s, prec, ops, nops, total);
  return hipGetLastError();
}

// Plain complex64 phasor Field (one wavelength), row-major m x n.
__global__ void phasor_field_kernel(const void* __restrict__ amp, int amp_dtype, long long ld_amp,
                                    const void* __restrict__ opd, int opd_dtype, long long ld_opd, int m,
                                    int n, double inv_wl, float2* __restrict__ dst) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)m * n) return;
  const long long row = t / n, k = t % n;
  const float a = (float)load_real(amp, amp_dtype, row * ld_amp + k);
  const float2 c = cis_turns(load_real(opd, opd_dtype, row * ld_opd + k) * inv_wl);
  dst[t] = make_float2(a * c.x, a * c.y);
}

hipError_t launch_phasor_field(const void* amp, int amp_dtype, long long ld_amp, const void* opd, int opd_dtype,
                               long long ld_opd, int m, int n, double inv_wl, void* dst, hipStream_t s) {
  const long long total = (long long)m * n;
  if (total <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(phasor_field_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, amp,
                     amp_dtype, ld_amp, opd, opd_dtype, ld_opd, m, n, inv_wl, (float2*)dst);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// detector accumulation
// --------------------------------------------------------------------------

// out[r0 + r][c0 + c] (+)= sum_g partial[g][r][c], clipped to the output
// grid.  The sum over groups runs in FP64 and is rounded once.
__global__ void reduce_partials_kernel(const float* __restrict__ partial, long long gstride, int ngroups, int rows,
                                       int cols, long long ldp, void* __restrict__ out, int out_dtype,
                                       long long ldo, int out_rows, int out_cols, int r0, int c0, int accumulate) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rows * cols) return;
  const int r = (int)(t / cols), c = (int)(t % cols);
  const int orr = r0 + r, oc = c0 + c;
  if (orr < 0 || oc < 0 || orr >= out_rows || oc >= out_cols) return;
  double acc = 0.0;
  for (int g = 0; g < ngroups; ++g) acc += (double)partial[(long long)g * gstride + (long long)r * ldp + c];
  const long long o = (long long)orr * ldo + oc;
  if (out_dtype == 3) {
    double* d = (double*)out;
    d[o] = accumulate ? d[o] + acc : acc;
  } else {
    float* d = (float*)out;
    d[o] = accumulate ? (float)((double)d[o] + acc) : (float)acc;
  }
}

hipError_t launch_reduce_partials(const float* partial, long long gstride, int ngroups, int rows, int cols,
                                  long long ldp, void* out, int out_dtype, long long ldo, int out_rows,
                                  int out_cols, int r0, int c0, int accumulate, hipStream_t s) {
  const long long total = (long long)rows * cols;
  if (total <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, partial,
                     gstride, ngroups, rows, cols, ldp, out, out_dtype, ldo, out_rows, out_cols, r0, c0,
                     accumulate);
  return hipGetLastError();
}

// out[or0 + r][oc0 + c] += weight * |E|^2  or  weight * E   (field.py:301-304)
// for E = field[fr0 + r][fc0 + c], r < h, c < w (already clipped by the host).
__global__ void insert_kernel(const float2* __restrict__ field, long long ldf, int fr0, int fc0, int h, int w,
                              void* __restrict__ out, int out_dtype, long long ldo, int or0, int oc0,
                              int intensity, double weight) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)h * w) return;
  const int r = (int)(t / w), c = (int)(t % w);
  const float2 e = field[(long long)(fr0 + r) * ldf + (fc0 + c)];
  const long long o = (long long)(or0 + r) * ldo + (oc0 + c);
  if (intensity) {
    const double i2 = ((double)e.x * e.x + (double)e.y * e.y) * weight;
    if (out_dtype == 3) ((double*)out)[o] += i2;
    else ((float*)out)[o] += (float)i2;
  } else {
    if (out_dtype == 2) {
      double2* d = (double2*)out;
      d[o].x += (double)e.x * weight;
      d[o].y += (double)e.y * weight;
    } else {
      float2* d = (float2*)out;
      d[o].x += (float)(e.x * weight);
      d[o].y += (float)(e.y * weight);
    }
  }
}

hipError_t launch_insert(const void* field, long long ldf, int fr0, int fc0, int h, int w, void* out,
                         int out_dtype, long long ldo, int or0, int oc0, int intensity, double weight,
                         hipStream_t s) {
  const long long total = (long long)h * w;
  if (total <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(insert_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s,
                     (const float2*)field, ldf, fr0, fc0, h, w, out, out_dtype, ldo, or0, oc0, intensity, weight);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// phase-retrieval backward pass (no counterpart in the reference)
// --------------------------------------------------------------------------

// out[r][c] = scale * w[r][c] * F[r][c]: the cotangent w * dL/dI * F of a
// weighted intensity image (d|F|^2 / d conj(F) = F).
__global__ void weight_field_kernel(const float2* __restrict__ F, long long ldF, const void* __restrict__ w,
                                    int w_dtype, long long ldw, double scale, float2* __restrict__ out,
                                    long long ldo, int h, int wd) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)h * wd) return;
  const long long r = t / wd, c = t % wd;
  const double g = (w_dtype == 3 ? ((const double*)w)[r * ldw + c] : (double)((const float*)w)[r * ldw + c]) * scale;
  const float2 v = F[r * ldF + c];
  out[r * ldo + c] = make_float2((float)(g * v.x), (float)(g * v.y));
}

hipError_t launch_weight_field(const void* F, long long ldF, const void* w, int w_dtype, long long ldw, double scale,
                               void* out, long long ldo, int h, int wd, hipStream_t s) {
  const long long total = (long long)h * wd;
  if (total <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(weight_field_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, (const float2*)F,
                     ldF, w, w_dtype, ldw, scale, (float2*)out, ldo, h, wd);
  return hipGetLastError();
}

// Given D_w = dL/d conj(f_w) for f_w = amp * exp(2 pi i opd / lambda_w),
// w = 0 .. nwave-1 (D_w at D + w * strideD):
//   grad_opd += sum_w (2 pi / lambda_w) * (-2) Im(conj(D_w) f_w)
//   grad_amp += sum_w 2 Re(conj(D_w) exp(2 pi i opd / lambda_w))      (if grad_amp)
// f is recomputed from amp and opd (opd already includes any diversity); the
// 1/lambda values are inv_wls[w] on the device, or inv_wl0 when inv_wls is null.
__global__ void phase_grad_kernel(const void* __restrict__ amp, int amp_dtype, long long ld_amp,
                                  const void* __restrict__ opd, int opd_dtype, long long ld_opd, int m, int n,
                                  const double* __restrict__ inv_wls, double inv_wl0, int nwave,
                                  const float2* __restrict__ D, long long ldD, long long strideD,
                                  double* __restrict__ g_opd, long long ldg, double* __restrict__ g_amp,
                                  long long ldga) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)m * n) return;
  const long long r = t / n, c = t % n;
  const double a = amp_dtype == 3 ? ((const double*)amp)[r * ld_amp + c] : (double)((const float*)amp)[r * ld_amp + c];
  const double o = opd_dtype == 3 ? ((const double*)opd)[r * ld_opd + c] : (double)((const float*)opd)[r * ld_opd + c];
  double go = 0.0, ga = 0.0;
  for (int w = 0; w < nwave; ++w) {
    const double inv_wl = inv_wls ? inv_wls[w] : inv_wl0;
    const float2 e = cis_turns(o * inv_wl);
    const float2 d = D[(long long)w * strideD + r * ldD + c];
    // conj(D) * e = (d.x - i d.y)(e.x + i e.y)
    const double re = (double)d.x * e.x + (double)d.y * e.y;
    const double im = (double)d.x * e.y - (double)d.y * e.x;
    go += -4.0 * 3.14159265358979323846 * inv_wl * a * im;
    ga += 2.0 * re;
  }
  g_opd[r * ldg + c] += go;
  if (g_amp) g_amp[r * ldga + c] += ga;
}

hipError_t launch_phase_grad(const void* amp, int amp_dtype, long long ld_amp, const void* opd, int opd_dtype,
                             long long ld_opd, int m, int n, const double* inv_wls, double inv_wl0, int nwave,
                             const void* D, long long ldD, long long strideD, double* g_opd, long long ldg,
                             double* g_amp, long long ldga, hipStream_t s) {
  const long long total = (long long)m * n;
  if (total <= 0 || nwave <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(phase_grad_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, amp, amp_dtype, ld_amp,
                     opd, opd_dtype, ld_opd, m, n, inv_wls, inv_wl0, nwave, (const float2*)D, ldD, strideD, g_opd,
                     ldg, g_amp, ldga);
  return hipGetLastError();
}

// out[r][c] = a[r][c] * b[r][c] (complex64), the Field product of field.py:142.
__global__ void multiply_kernel(const float2* __restrict__ a, long long lda, const float2* __restrict__ b,
                                long long ldb, float2* __restrict__ out, long long ldo, int h, int w) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)h * w) return;
  const long long r = t / w, c = t % w;
  out[r * ldo + c] = cmul(a[r * lda + c], b[r * ldb + c]);
}

hipError_t launch_multiply(const void* a, long long lda, const void* b, long long ldb, void* out, long long ldo,
                           int h, int w, hipStream_t s) {
  const long long total = (long long)h * w;
  if (total <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(multiply_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, (const float2*)a, lda,
                     (const float2*)b, ldb, (float2*)out, ldo, h, w);
  return hipGetLastError();
}

__global__ void zero_kernel(uint4* __restrict__ p, long long n16, unsigned char* __restrict__ tail, int ntail) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n16) p[t] = make_uint4(0, 0, 0, 0);
  if (t < ntail) tail[t] = 0;
}

// Zero rows x cols elements of esz bytes (4 or 8) at row stride ld elements;
// the bytes between rows and after the last row are left alone.
__global__ void zero_rows_kernel(void* __restrict__ p, int esz, long long rows, long long cols, long long ld) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * cols) return;
  const long long o = (t / cols) * ld + t % cols;
  if (esz == 8) ((unsigned long long*)p)[o] = 0ull;
  else ((unsigned int*)p)[o] = 0u;
}

hipError_t launch_zero_rows(void* p, int esz, long long rows, long long cols, long long ld, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  if (esz != 4 && esz != 8) return hipErrorInvalidValue;
  if (ld == cols && ((uintptr_t)p & 15) == 0) return launch_zero(p, (size_t)esz * rows * cols, s);
  const long long total = rows * cols;
  const int bs = 256;
  hipLaunchKernelGGL(zero_rows_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, p, esz, rows, cols,
                     ld);
  return hipGetLastError();
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  const long long n16 = (long long)(bytes / 16);
  const int ntail = (int)(bytes % 16);
  const long long total = n16 > ntail ? n16 : ntail;
  const int bs = 256;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, (uint4*)p, n16,
                     (unsigned char*)p + n16 * 16, ntail);
  return hipGetLastError();
}

}  // namespace mftx
