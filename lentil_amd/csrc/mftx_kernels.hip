// MI355X (gfx950) kernels for the matrix Fourier transform path.
//
// The MFT F = c * E1 . f . E2 (lentil fourier.py:97) is evaluated as two
// complex GEMMs of the form C = A . B where A is data in HBM and B is a
// twiddle matrix generated on the fly into LDS:
//     stage 1:  G   = f   . E2      (stored transposed, as G^T)
//     stage 2:  F^T = G^T . E1^T    (|F|^2 accumulated, or F stored)
// Twiddles B[k][j] = cis(alpha * x_k * y_j), cis(t) = exp(2*pi*i*t), with the
// product alpha*x*y and its range reduction in FP64 and the sine/cosine of
// the reduced angle in FP32 (fourier.py:108-109 builds the same matrices in
// complex128).
//
// Precision forms
//   PREC_FP32   : complex64 operands, v_mfma_f32_32x32x2_f32, 4 real MFMAs per
//                 complex product.
//   PREC_BF16X3 : every FP32 operand x is split x = hi + lo (two BF16, round
//                 to nearest even); a real product is hi*hi + hi*lo + lo*hi on
//                 v_mfma_f32_32x32x16_bf16 with FP32 accumulation, 12 MFMAs per
//                 complex product.
#include "mftx_internal.h"

namespace mftx {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// --------------------------------------------------------------------------
// scalar helpers
// --------------------------------------------------------------------------

// cis(t) = (cos 2*pi*t, sin 2*pi*t) for t in turns.  t is reduced in FP64 to
// a quarter-turn index n and a remainder f in [-1/2, 1/2] quarter turns; the
// remainder angle (|a| <= pi/4) goes through Taylor series to degree 9 (sin)
// and 10 (cos), whose truncation error is below 3e-8 on that interval.
__device__ __forceinline__ float2 cis_turns(double t) {
  const double q = t * 4.0;
  const double n = __builtin_rint(q);
  const float f = (float)(q - n);
  const int quad = (int)(long long)n;
  const float a = f * 1.57079632679489662f;
  const float a2 = a * a;
  float s = 2.75573192239858907e-6f;           // 1/9!
  s = __builtin_fmaf(s, a2, -1.98412698412698413e-4f);  // -1/7!
  s = __builtin_fmaf(s, a2, 8.33333333333333333e-3f);   // 1/5!
  s = __builtin_fmaf(s, a2, -1.66666666666666667e-1f);  // -1/3!
  s = a + a * a2 * s;
  float c = -2.75573192239858907e-7f;          // -1/10!
  c = __builtin_fmaf(c, a2, 2.48015873015873016e-5f);   // 1/8!
  c = __builtin_fmaf(c, a2, -1.38888888888888889e-3f);  // -1/6!
  c = __builtin_fmaf(c, a2, 4.16666666666666667e-2f);   // 1/4!
  c = __builtin_fmaf(c, a2, -0.5f);
  c = __builtin_fmaf(c, a2, 1.0f);
  // multiply by i^quad
  const int k = quad & 3;
  const float re = (k == 0) ? c : (k == 1) ? -s : (k == 2) ? -c : s;
  const float im = (k == 0) ? s : (k == 1) ? c : (k == 2) ? -s : -c;
  return make_float2(re, im);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(__builtin_fmaf(a.x, b.x, -a.y * b.y), __builtin_fmaf(a.x, b.y, a.y * b.x));
}

// Round-to-nearest-even pack of two floats into two bf16 (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  bf16x2_t v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ float bf16_lo_to_f32(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16_hi_to_f32(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }

// Split two floats into (hi pair, lo pair) with x = hi + lo + O(2^-17 |x|).
__device__ __forceinline__ void split2(float a, float b, uint32_t& hi, uint32_t& lo) {
  hi = pack_bf16(a, b);
  lo = pack_bf16(a - bf16_lo_to_f32(hi), b - bf16_hi_to_f32(hi));
}

// --------------------------------------------------------------------------
// GEMM
// --------------------------------------------------------------------------

template <int PREC> struct Lds;

// bf16x3: 4 planes (re_hi, re_lo, im_hi, im_lo) of [row][BK + 8] bf16; the 8
// padding elements make a 48-byte row stride, which spreads the 16-byte
// fragment reads of 8 consecutive lanes over distinct banks.
template <> struct Lds<PREC_BF16X3> {
  static constexpr int RS = BK + 8;
  __bf16 a[2][4][BM][RS];
  __bf16 b[2][4][BN][RS];
};
// fp32: complex64 [row][BK + 2]; the 144-byte row stride keeps the 8-byte
// fragment reads of 16 consecutive lanes on distinct banks.
template <> struct Lds<PREC_FP32> {
  static constexpr int RS = BK + 2;
  float2 a[2][BM][RS];
  float2 b[2][BN][RS];
};

template <int PREC, int EPI>
__global__ __launch_bounds__(NTHREADS, 1) void gemm_kernel(const Work* __restrict__ works,
                                                           const Problem* __restrict__ probs,
                                                           const Segment* __restrict__ segs) {
  __shared__ Lds<PREC> lds;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wr = wid >> 1, wc = wid & 1;
  const Work wk = works[blockIdx.x];

  // loader / generator coordinates: one A row or B column, one half of BK
  const int ld_row = tid >> 1;
  const int ld_half = tid & 1;

  f32x16 cr[2][2], ci[2][2];
  f32x16 iacc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cr[i][j] = f32x16{};
      ci[i][j] = f32x16{};
      iacc[i][j] = f32x16{};
    }

  // ---- generation / load cursor -------------------------------------------
  int gp = wk.pbeg;   // problem of the next chunk to stage
  int gs = 0;         // chunk index within gp
  int gseg = 0;       // segment of (gp, gs)
  float2 q[8];        // per-problem factors cis(alpha * i * y) for i = 0..7
  double g_alpha = 0.0, g_y = 0.0;

  auto begin_problem = [&](int p) {
    const Problem& P = probs[p];
    g_alpha = P.alpha;
    g_y = P.y0 + (double)(wk.tc * BN + ld_row);
    gseg = P.seg_begin;
#pragma unroll
    for (int i = 0; i < 8; ++i) q[i] = cis_turns(g_alpha * ((double)i * g_y));
  };

  uint4 areg[4];

  auto load_a = [&](int p, int s) {
    const Problem& P = probs[p];
    const long long row = (long long)wk.tr * BM + ld_row;
    if (PREC == PREC_BF16X3) {
      const __bf16* src = (const __bf16*)P.a + row * P.lda + (long long)s * 64 + ld_half * 32;
      const uint4* s4 = (const uint4*)src;
#pragma unroll
      for (int i = 0; i < 4; ++i) areg[i] = s4[i];
    } else {
      const float2* src = (const float2*)P.a + row * P.lda + (long long)s * BK + ld_half * 8;
      const uint4* s4 = (const uint4*)src;
#pragma unroll
      for (int i = 0; i < 4; ++i) areg[i] = s4[i];
    }
  };

  auto store_a = [&](int buf) {
    if (PREC == PREC_BF16X3) {
      // areg holds planes 2*half and 2*half+1, 16 bf16 each
      uint4* d0 = (uint4*)&((Lds<PREC_BF16X3>*)&lds)->a[buf][2 * ld_half][ld_row][0];
      uint4* d1 = (uint4*)&((Lds<PREC_BF16X3>*)&lds)->a[buf][2 * ld_half + 1][ld_row][0];
      d0[0] = areg[0];
      d0[1] = areg[1];
      d1[0] = areg[2];
      d1[1] = areg[3];
    } else {
      uint4* d = (uint4*)&((Lds<PREC_FP32>*)&lds)->a[buf][ld_row][ld_half * 8];
#pragma unroll
      for (int i = 0; i < 4; ++i) d[i] = areg[i];
    }
  };

  auto gen_b = [&](int p, int s, int buf) {
    const int k0 = s * BK;
    while (k0 >= segs[gseg].kend) ++gseg;
    const Segment& S = segs[gseg];
    const double X = S.x0 + (double)(k0 - S.kbeg + ld_half * 8);
    const float2 pf = cis_turns(g_alpha * (X * g_y));
    float2 t[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) t[i] = cmul(pf, q[i]);
    if (PREC == PREC_BF16X3) {
      uint32_t rh[4], rl[4], ih[4], il[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        split2(t[2 * i].x, t[2 * i + 1].x, rh[i], rl[i]);
        split2(t[2 * i].y, t[2 * i + 1].y, ih[i], il[i]);
      }
      auto* L = (Lds<PREC_BF16X3>*)&lds;
      *(uint4*)&L->b[buf][0][ld_row][ld_half * 8] = make_uint4(rh[0], rh[1], rh[2], rh[3]);
      *(uint4*)&L->b[buf][1][ld_row][ld_half * 8] = make_uint4(rl[0], rl[1], rl[2], rl[3]);
      *(uint4*)&L->b[buf][2][ld_row][ld_half * 8] = make_uint4(ih[0], ih[1], ih[2], ih[3]);
      *(uint4*)&L->b[buf][3][ld_row][ld_half * 8] = make_uint4(il[0], il[1], il[2], il[3]);
    } else {
      auto* L = (Lds<PREC_FP32>*)&lds;
      float4* d = (float4*)&L->b[buf][ld_row][ld_half * 8];
#pragma unroll
      for (int i = 0; i < 4; ++i) d[i] = make_float4(t[2 * i].x, t[2 * i].y, t[2 * i + 1].x, t[2 * i + 1].y);
    }
  };

  auto advance = [&]() {
    ++gs;
    if (gs * BK >= probs[gp].kpad) {
      gs = 0;
      ++gp;
      if (gp < wk.pend) begin_problem(gp);
    }
  };

  auto compute = [&](int buf) {
    if (PREC == PREC_BF16X3) {
      auto* L = (Lds<PREC_BF16X3>*)&lds;
      const int koff = (lane >> 5) * 8;
      bf16x8_t a[2][4], b[2][4], nb[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int row = wr * 64 + i * 32 + (lane & 31);
        const int col = wc * 64 + i * 32 + (lane & 31);
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          a[i][p] = *(const bf16x8_t*)&L->a[buf][p][row][koff];
          b[i][p] = *(const bf16x8_t*)&L->b[buf][p][col][koff];
        }
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          uint4 u = __builtin_bit_cast(uint4, b[i][2 + p]);
          u.x ^= 0x80008000u; u.y ^= 0x80008000u; u.z ^= 0x80008000u; u.w ^= 0x80008000u;
          nb[i][p] = __builtin_bit_cast(bf16x8_t, u);
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          f32x16 r = cr[i][j], m = ci[i][j];
          // Cr += Ar*Br - Ai*Bi ; Ci += Ar*Bi + Ai*Br  (each product as hh + hl + lh)
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][0], r, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][2], m, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][1], r, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][3], m, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][1], b[j][0], r, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][1], b[j][2], m, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][2], nb[j][0], r, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][2], b[j][0], m, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][2], nb[j][1], r, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][2], b[j][1], m, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][3], nb[j][0], r, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][3], b[j][0], m, 0, 0, 0);
          cr[i][j] = r;
          ci[i][j] = m;
        }
    } else {
      auto* L = (Lds<PREC_FP32>*)&lds;
#pragma unroll
      for (int ks = 0; ks < BK / 2; ++ks) {
        const int kk = 2 * ks + (lane >> 5);
        float2 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          a[i] = L->a[buf][wr * 64 + i * 32 + (lane & 31)][kk];
          b[i] = L->b[buf][wc * 64 + i * 32 + (lane & 31)][kk];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            f32x16 r = cr[i][j], m = ci[i][j];
            r = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].x, b[j].x, r, 0, 0, 0);
            m = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].x, b[j].y, m, 0, 0, 0);
            r = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].y, -b[j].y, r, 0, 0, 0);
            m = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].y, b[j].x, m, 0, 0, 0);
            cr[i][j] = r;
            ci[i][j] = m;
          }
      }
    }
  };

  // C element (block i, j; accumulator register e) -> (row, col) of C
  auto c_row = [&](int i, int e) {
    return wk.tr * BM + wr * 64 + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
  };
  auto c_col = [&](int j) { return wk.tc * BN + wc * 64 + j * 32 + (lane & 31); };

  auto epilogue = [&](int p) {
    const Problem& P = probs[p];
    if (EPI == EPI_INTENSITY) {
      const float w = P.scale;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 16; ++e)
            iacc[i][j][e] = __builtin_fmaf(w, __builtin_fmaf(cr[i][j][e], cr[i][j][e], ci[i][j][e] * ci[i][j][e]),
                                           iacc[i][j][e]);
    } else if (EPI == EPI_STORE_CPLX) {
      float2* dst = (float2*)P.dst;
      const float sc = P.scale;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int col = c_col(j);
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            const int row = c_row(i, e);
            if (row < P.rows && col < P.cols) {
              const float2 v = make_float2(cr[i][j][e] * sc, ci[i][j][e] * sc);
              if (P.trans) dst[(long long)col * P.ldd + row] = v;
              else dst[(long long)row * P.ldd + col] = v;
            }
          }
        }
    } else {  // EPI_STORE_PACKED_T: destination row = C col, destination k = C row
      const float sc = P.scale;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const long long col = c_col(j);
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const int row0 = c_row(i, 4 * g);  // 4 consecutive rows row0..row0+3
            if (row0 >= P.limit) continue;
            float re[4], im[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              re[e] = cr[i][j][4 * g + e] * sc;
              im[e] = ci[i][j][4 * g + e] * sc;
            }
            if (PREC == PREC_BF16X3) {
              uint32_t rh0, rl0, rh1, rl1, ih0, il0, ih1, il1;
              split2(re[0], re[1], rh0, rl0);
              split2(re[2], re[3], rh1, rl1);
              split2(im[0], im[1], ih0, il0);
              split2(im[2], im[3], ih1, il1);
              __bf16* d = (__bf16*)P.dst + col * P.ldd + (long long)(row0 >> 4) * 64 + (row0 & 15);
              *(uint2*)(d + 0) = make_uint2(rh0, rh1);
              *(uint2*)(d + 16) = make_uint2(rl0, rl1);
              *(uint2*)(d + 32) = make_uint2(ih0, ih1);
              *(uint2*)(d + 48) = make_uint2(il0, il1);
            } else {
              float4* d = (float4*)((float2*)P.dst + col * P.ldd + row0);
              d[0] = make_float4(re[0], im[0], re[1], im[1]);
              d[1] = make_float4(re[2], im[2], re[3], im[3]);
            }
          }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        cr[i][j] = f32x16{};
        ci[i][j] = f32x16{};
      }
  };

  // ---- prologue: stage chunk 0 of the first problem ------------------------
  if (gp >= wk.pend) return;
  begin_problem(gp);
  load_a(gp, 0);
  gen_b(gp, 0, 0);
  store_a(0);
  advance();
  __syncthreads();

  int it = 0;
  for (int pc = wk.pbeg; pc < wk.pend; ++pc) {
    const int nch = probs[pc].kpad / BK;
    for (int s = 0; s < nch; ++s, ++it) {
      const int cur = it & 1;
      const bool more = gp < wk.pend;
      if (more) {
        load_a(gp, gs);
        gen_b(gp, gs, cur ^ 1);
      }
      compute(cur);
      if (more) {
        store_a(cur ^ 1);
        advance();
      }
      __syncthreads();
    }
    epilogue(pc);
  }

  if (EPI == EPI_INTENSITY) {
    const int trans = probs[wk.pbeg].trans;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = c_col(j);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int row = c_row(i, e);
          const int pr = trans ? col : row;
          const int pcn = trans ? row : col;
          if (pr < wk.prows && pcn < wk.pcols) wk.partial[(long long)pr * wk.ldp + pcn] = iacc[i][j][e];
        }
      }
  }
}

hipError_t launch_gemm(int prec, int epi, const Work* work, int nwork, const Problem* probs,
                       const Segment* segs, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  dim3 grid(nwork), block(NTHREADS);
#define MFTX_LAUNCH(P, E) hipLaunchKernelGGL((gemm_kernel<P, E>), grid, block, 0, s, work, probs, segs)
  if (prec == PREC_BF16X3) {
    if (epi == EPI_STORE_CPLX) MFTX_LAUNCH(PREC_BF16X3, EPI_STORE_CPLX);
    else if (epi == EPI_STORE_PACKED_T) MFTX_LAUNCH(PREC_BF16X3, EPI_STORE_PACKED_T);
    else MFTX_LAUNCH(PREC_BF16X3, EPI_INTENSITY);
  } else {
    if (epi == EPI_STORE_CPLX) MFTX_LAUNCH(PREC_FP32, EPI_STORE_CPLX);
    else if (epi == EPI_STORE_PACKED_T) MFTX_LAUNCH(PREC_FP32, EPI_STORE_PACKED_T);
    else MFTX_LAUNCH(PREC_FP32, EPI_INTENSITY);
  }
#undef MFTX_LAUNCH
  return hipGetLastError();
}

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

// One thread per (row, 16-wide K block) of the packed destination.
__global__ void pack_kernel(int prec, const void* __restrict__ src, int dtype, long long m, long long n,
                            long long lds, void* __restrict__ dst, long long rows_pad, int kpad) {
  const int kblocks = kpad / 16;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows_pad * kblocks) return;
  const long long row = t / kblocks;
  const int kb = (int)(t % kblocks);
  float2 v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const long long k = (long long)kb * 16 + i;
    v[i] = (row < m && k < n) ? load_elem(src, dtype, row * lds + k) : make_float2(0.f, 0.f);
  }
  store_packed16(prec, dst, packed_lda(prec, kpad), row, kb, v);
}

hipError_t launch_pack(int prec, const void* src, int src_dtype, long long m, long long n, long long lds,
                       void* dst, long long rows_pad, int kpad, hipStream_t s) {
  const long long total = rows_pad * (kpad / 16);
  if (total <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, prec, src,
                     src_dtype, m, n, lds, dst, rows_pad, kpad);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// pupil phasor  f = amp * exp(2*pi*i*opd/lambda)   (plane.py:507)
// --------------------------------------------------------------------------

__device__ __forceinline__ double load_real(const void* p, int dtype, long long idx) {
  return dtype == 3 ? ((const double*)p)[idx] : (double)((const float*)p)[idx];
}

// One thread per (row, 16-wide K block, wavelength); writes the packed operand.
__global__ void phasor_kernel(int prec, const void* __restrict__ amp, int amp_dtype, long long ld_amp,
                              const void* __restrict__ opd, int opd_dtype, long long ld_opd, int m, int n,
                              const double* __restrict__ inv_wl, int nwave, void* __restrict__ dst,
                              long long wave_stride, long long rows_pad, int kpad) {
  const int kblocks = kpad / 16;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per_wave = rows_pad * kblocks;
  if (t >= per_wave * nwave) return;
  const int w = (int)(t / per_wave);
  const long long r = t % per_wave;
  const long long row = r / kblocks;
  const int kb = (int)(r % kblocks);
  const double iw = inv_wl[w];
  float2 v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const long long k = (long long)kb * 16 + i;
    if (row < m && k < n) {
      const float a = (float)load_real(amp, amp_dtype, row * ld_amp + k);
      const float2 c = cis_turns(load_real(opd, opd_dtype, row * ld_opd + k) * iw);
      v[i] = make_float2(a * c.x, a * c.y);
    } else {
      v[i] = make_float2(0.f, 0.f);
    }
  }
  char* base = (char*)dst + (long long)w * wave_stride;
  store_packed16(prec, base, packed_lda(prec, kpad), row, kb, v);
}

hipError_t launch_phasor(int prec, const void* amp, int amp_dtype, long long ld_amp, const void* opd,
                         int opd_dtype, long long ld_opd, int m, int n, const double* inv_wl, int nwave,
                         void* dst, long long wave_stride, long long rows_pad, int kpad, hipStream_t s) {
  const long long total = rows_pad * (kpad / 16) * nwave;
  if (total <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(phasor_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, s, prec, amp,
                     amp_dtype, ld_amp, opd, opd_dtype, ld_opd, m, n, inv_wl, nwave, dst, wave_stride,
                     rows_pad, kpad);
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
