// Complex GEMM C = A . B of the two MFT stages on gfx950 MFMA.
//
// A is data in HBM (packed, see mftx_internal.h); B is a twiddle matrix
// B[k][j] = cis(alpha * x_k * y_j) generated on the fly into LDS:
//     stage 1:  G   = f   . E2      (stored transposed, packed, as G^T)
//     stage 2:  F^T = G^T . E1^T    (|F|^2 accumulated, or F stored)
// which is the triple product of lentil fourier.py:97 with the twiddles of
// fourier.py:108-109.
//
// Workgroup: 256 threads = 4 waves in 2 x 2, each wave a 64 x 64 block of C
// as 2 x 2 MFMA 32x32 tiles; C tile 128 x 128, K in chunks of 16 staged in
// double-buffered LDS.  One chunk ahead, each thread stages one A row half
// (global -> registers -> LDS) and generates one B column half: 8 twiddles
// t_i = P * Q_i with P = cis(alpha * x * y) (FP64 argument, FP32 sin/cos)
// per chunk and the per-problem table Q_i = cis(alpha * i * y), i = 0..7.
// The twiddle VALU work is interleaved with the MFMAs of the current chunk
// (sched_group_barrier), and the loop body is one basic block.
//
// Precision forms
//   PREC_FP32   : complex64 operands, v_mfma_f32_32x32x2_f32, 4 real MFMAs per
//                 complex product.
//   PREC_BF16X3 : every FP32 operand x is split x = hi + lo (two BF16, round
//                 to nearest even); a real product is hi*hi + hi*lo + lo*hi on
//                 v_mfma_f32_32x32x16_bf16 with FP32 accumulation, 12 MFMAs per
//                 complex product.
#include "mftx_device.h"

namespace mftx {

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

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const u32x4 gu32x4;  // global-memory view

// Staged chunk in registers: A prefetch and the generated twiddles
// (bf16x3: planes re_hi, re_lo, im_hi, im_lo of 8 K values; fp32: 8 complex).
// Native vector types keep these in VGPRs (no private-memory copies).
struct Staged {
  u32x4 a0, a1, a2, a3;
  u32x4 b0, b1, b2, b3;
};

// Per-problem staging parameters (uniform) and the per-thread Q table.
struct StageProb {
  const char* a;
  long long lda;
  double alpha4;   // 4 * alpha: twiddle arguments in quarter turns
  double y;        // y of this thread's B column
  int nch;         // chunks in the problem
  int cb;          // chunk_begin
};

__device__ __forceinline__ void set_stage(const Problem& P, int col, StageProb& sp, float2 (&q)[8]) {
  sp.a = (const char*)P.a;
  sp.lda = P.lda;
  sp.alpha4 = 4.0 * P.alpha;
  sp.y = P.y0 + (double)col;
  sp.nch = P.kpad / BK;
  sp.cb = P.chunk_begin;
#pragma unroll
  for (int i = 0; i < 8; ++i) q[i] = cis_quarter(sp.alpha4 * ((double)i * sp.y));
}

template <int PREC>
__device__ __forceinline__ void load_a(const StageProb& sp, long long row, int chunk, int half, Staged& st) {
  gu32x4* src;
  if (PREC == PREC_BF16X3)
    src = (gu32x4*)((const __bf16*)sp.a + row * sp.lda + (long long)chunk * 64 + half * 32);
  else
    src = (gu32x4*)((const float2*)sp.a + row * sp.lda + (long long)chunk * BK + half * 8);
  st.a0 = src[0];
  st.a1 = src[1];
  st.a2 = src[2];
  st.a3 = src[3];
}

// Twiddles of K indices x, x+1, ..., x+7 for this thread's column.
template <int PREC>
__device__ __forceinline__ void gen_b(double x, const StageProb& sp, const float2 (&q)[8], Staged& st) {
  const float2 pf = cis_quarter(sp.alpha4 * (x * sp.y));
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
    st.b0 = u32x4{rh[0], rh[1], rh[2], rh[3]};
    st.b1 = u32x4{rl[0], rl[1], rl[2], rl[3]};
    st.b2 = u32x4{ih[0], ih[1], ih[2], ih[3]};
    st.b3 = u32x4{il[0], il[1], il[2], il[3]};
  } else {
    auto bits = [](float v) { return __float_as_uint(v); };
    st.b0 = u32x4{bits(t[0].x), bits(t[0].y), bits(t[1].x), bits(t[1].y)};
    st.b1 = u32x4{bits(t[2].x), bits(t[2].y), bits(t[3].x), bits(t[3].y)};
    st.b2 = u32x4{bits(t[4].x), bits(t[4].y), bits(t[5].x), bits(t[5].y)};
    st.b3 = u32x4{bits(t[6].x), bits(t[6].y), bits(t[7].x), bits(t[7].y)};
  }
}

template <int PREC>
__device__ __forceinline__ void write_stage(Lds<PREC>& L, int buf, int row, int half, const Staged& st) {
  if (PREC == PREC_BF16X3) {
    auto& LB = reinterpret_cast<Lds<PREC_BF16X3>&>(L);
    *(u32x4*)&LB.b[buf][0][row][half * 8] = st.b0;
    *(u32x4*)&LB.b[buf][1][row][half * 8] = st.b1;
    *(u32x4*)&LB.b[buf][2][row][half * 8] = st.b2;
    *(u32x4*)&LB.b[buf][3][row][half * 8] = st.b3;
    // A: this thread's 64 bytes are planes 2*half and 2*half+1 of its row
    *(u32x4*)&LB.a[buf][2 * half][row][0] = st.a0;
    *(u32x4*)&LB.a[buf][2 * half][row][8] = st.a1;
    *(u32x4*)&LB.a[buf][2 * half + 1][row][0] = st.a2;
    *(u32x4*)&LB.a[buf][2 * half + 1][row][8] = st.a3;
  } else {
    auto& LF = reinterpret_cast<Lds<PREC_FP32>&>(L);
    u32x4* b = (u32x4*)&LF.b[buf][row][half * 8];
    b[0] = st.b0;
    b[1] = st.b1;
    b[2] = st.b2;
    b[3] = st.b3;
    u32x4* a = (u32x4*)&LF.a[buf][row][half * 8];
    a[0] = st.a0;
    a[1] = st.a1;
    a[2] = st.a2;
    a[3] = st.a3;
  }
}

#define MFMA_BF16(a, b, c) __builtin_amdgcn_mfma_f32_32x32x16_bf16((a), (b), (c), 0, 0, 0)
#define SGB(mask, n) __builtin_amdgcn_sched_group_barrier((mask), (n), 0)
constexpr int SG_VALU = 0x002, SG_MFMA = 0x008, SG_VMEM_RD = 0x020, SG_DSR = 0x100;

__device__ __forceinline__ bf16x8_t neg_bf16x8(bf16x8_t v) {
  return __builtin_bit_cast(bf16x8_t, __builtin_bit_cast(u32x4, v) ^ 0x80008000u);
}

// One complex 32x32 block, K = 16: Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br,
// each real product as hi*hi + hi*lo + lo*hi.  a/b: planes (rh, rl, ih, il);
// nbi: (-Bi_hi, -Bi_lo).
__device__ __forceinline__ void cblock_bf16(const bf16x8_t (&a)[4], const bf16x8_t (&b)[4], const bf16x8_t (&nbi)[2],
                                            f32x16& r, f32x16& m) {
  r = MFMA_BF16(a[0], b[0], r);
  m = MFMA_BF16(a[0], b[2], m);
  r = MFMA_BF16(a[0], b[1], r);
  m = MFMA_BF16(a[0], b[3], m);
  r = MFMA_BF16(a[1], b[0], r);
  m = MFMA_BF16(a[1], b[2], m);
  r = MFMA_BF16(a[2], nbi[0], r);
  m = MFMA_BF16(a[2], b[0], m);
  r = MFMA_BF16(a[2], nbi[1], r);
  m = MFMA_BF16(a[2], b[1], m);
  r = MFMA_BF16(a[3], nbi[0], r);
  m = MFMA_BF16(a[3], b[0], m);
}

template <int PREC, int EPI>
__global__ __launch_bounds__(NTHREADS, 1) void gemm_kernel(const Work* __restrict__ works,
                                                           const Problem* __restrict__ probs,
                                                           const double* __restrict__ chunk_x) {
  __shared__ Lds<PREC> lds;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wr = wid >> 1, wc = wid & 1;
  const Work wk = works[blockIdx.x];
  if (wk.pbeg >= wk.pend) return;

  // staging coordinates: one A row / B column, one half of the 16-wide chunk
  const int st_row = tid >> 1;
  const int st_half = tid & 1;
  const long long a_row = (long long)wk.tr * BM + st_row;
  const int b_col = wk.tc * BN + st_row;

  f32x16 cr[2][2], ci[2][2], iacc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cr[i][j] = f32x16{};
      ci[i][j] = f32x16{};
      iacc[i][j] = f32x16{};
    }

  // ---- prologue: stage chunk 0 of the first problem into buffer 0 ----------
  int sp = wk.pbeg, ss = 0;  // staging cursor: next (problem, chunk) to stage
  StageProb sgp;
  float2 q[8];
  Staged st;
  set_stage(probs[sp], b_col, sgp, q);
  double sx = chunk_x[sgp.cb] + 8.0 * st_half;
  load_a<PREC>(sgp, a_row, 0, st_half, st);
  gen_b<PREC>(sx, sgp, q, st);
  write_stage<PREC>(lds, 0, st_row, st_half, st);
  if (++ss == sgp.nch) {
    ss = 0;
    if (++sp < wk.pend) set_stage(probs[sp], b_col, sgp, q);
  }
  sx = chunk_x[sgp.cb + ss] + 8.0 * st_half;
  __syncthreads();

  int cur = 0;
  for (int pc = wk.pbeg; pc < wk.pend; ++pc) {
    const int nch = probs[pc].kpad / BK;
    for (int s = 0; s < nch; ++s) {
      // stage the next chunk (after the last one: restage a valid chunk into
      // the idle buffer, never read)
      const bool more = sp < wk.pend;
      const int ts = more ? ss : 0;
      load_a<PREC>(sgp, a_row, ts, st_half, st);
      gen_b<PREC>(sx, sgp, q, st);

      if (PREC == PREC_BF16X3) {
        auto& L = reinterpret_cast<Lds<PREC_BF16X3>&>(lds);
        const int koff = (lane >> 5) * 8;
        const int r0 = wr * 64 + (lane & 31), c0 = wc * 64 + (lane & 31);
        bf16x8_t a0[4], a1[4], b0[4], b1[4], nb0[2], nb1[2];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          a0[p] = *(const bf16x8_t*)&L.a[cur][p][r0][koff];
          b0[p] = *(const bf16x8_t*)&L.b[cur][p][c0][koff];
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) b1[p] = *(const bf16x8_t*)&L.b[cur][p][c0 + 32][koff];
#pragma unroll
        for (int p = 0; p < 4; ++p) a1[p] = *(const bf16x8_t*)&L.a[cur][p][r0 + 32][koff];
        nb0[0] = neg_bf16x8(b0[2]);
        nb0[1] = neg_bf16x8(b0[3]);
        nb1[0] = neg_bf16x8(b1[2]);
        nb1[1] = neg_bf16x8(b1[3]);
        cblock_bf16(a0, b0, nb0, cr[0][0], ci[0][0]);
        cblock_bf16(a0, b1, nb1, cr[0][1], ci[0][1]);
        cblock_bf16(a1, b0, nb0, cr[1][0], ci[1][0]);
        cblock_bf16(a1, b1, nb1, cr[1][1], ci[1][1]);
        // schedule: next chunk's global loads first (consumed only by the LDS
        // writes after the MFMAs), fragment reads, then the twiddle VALU work
        // threaded between the 48 MFMAs
        SGB(SG_VMEM_RD, 4);
        SGB(SG_DSR, 8);
#pragma unroll
        for (int i = 0; i < 12; ++i) {
          SGB(SG_MFMA, 1);
          SGB(SG_VALU, 3);
        }
        SGB(SG_DSR, 8);
#pragma unroll
        for (int i = 0; i < 36; ++i) {
          SGB(SG_MFMA, 1);
          SGB(SG_VALU, 2);
        }
      } else {
        auto& L = reinterpret_cast<Lds<PREC_FP32>&>(lds);
#pragma unroll
        for (int ks = 0; ks < BK / 2; ++ks) {
          const int kk = 2 * ks + (lane >> 5);
          float2 a[2], b[2];
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            a[i] = L.a[cur][wr * 64 + i * 32 + (lane & 31)][kk];
            b[i] = L.b[cur][wc * 64 + i * 32 + (lane & 31)][kk];
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

      write_stage<PREC>(lds, cur ^ 1, st_row, st_half, st);
      __syncthreads();
      cur ^= 1;
      if (more) {
        if (++ss == sgp.nch) {
          ss = 0;
          if (++sp < wk.pend) set_stage(probs[sp], b_col, sgp, q);
        }
        sx = chunk_x[sgp.cb + (sp < wk.pend ? ss : 0)] + 8.0 * st_half;
      }
    }

    // ---- problem epilogue ----------------------------------------------------
    const Problem& P = probs[pc];
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
    } else {
      const float sc = P.scale;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int col = wk.tc * BN + wc * 64 + j * 32 + (lane & 31);
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            // rows row0 .. row0+3 of C in this lane (32x32 accumulator layout)
            const int row0 = wk.tr * BM + wr * 64 + i * 32 + 8 * g + 4 * (lane >> 5);
            float re[4], im[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              re[e] = cr[i][j][4 * g + e] * sc;
              im[e] = ci[i][j][4 * g + e] * sc;
            }
            if (EPI == EPI_STORE_CPLX) {
              float2* dst = (float2*)P.dst;
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                const int row = row0 + e;
                if (row < P.rows && col < P.cols) {
                  const float2 v = make_float2(re[e], im[e]);
                  if (P.trans) dst[(long long)col * P.ldd + row] = v;
                  else dst[(long long)row * P.ldd + col] = v;
                }
              }
            } else if (row0 < P.limit) {  // EPI_STORE_PACKED_T: destination row = C col, K = C row
              if (PREC == PREC_BF16X3) {
                uint32_t rh0, rl0, rh1, rl1, ih0, il0, ih1, il1;
                split2(re[0], re[1], rh0, rl0);
                split2(re[2], re[3], rh1, rl1);
                split2(im[0], im[1], ih0, il0);
                split2(im[2], im[3], ih1, il1);
                __bf16* d = (__bf16*)P.dst + (long long)col * P.ldd + (long long)(row0 >> 4) * 64 + (row0 & 15);
                *(uint2*)(d + 0) = make_uint2(rh0, rh1);
                *(uint2*)(d + 16) = make_uint2(rl0, rl1);
                *(uint2*)(d + 32) = make_uint2(ih0, ih1);
                *(uint2*)(d + 48) = make_uint2(il0, il1);
              } else {
                float4* d = (float4*)((float2*)P.dst + (long long)col * P.ldd + row0);
                d[0] = make_float4(re[0], im[0], re[1], im[1]);
                d[1] = make_float4(re[2], im[2], re[3], im[3]);
              }
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
  }

  if (EPI == EPI_INTENSITY) {
    const int trans = probs[wk.pbeg].trans;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = wk.tc * BN + wc * 64 + j * 32 + (lane & 31);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int row = wk.tr * BM + wr * 64 + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
          const int pr = trans ? col : row;
          const int pcn = trans ? row : col;
          if (pr < wk.prows && pcn < wk.pcols) wk.partial[(long long)pr * wk.ldp + pcn] = iacc[i][j][e];
        }
      }
  }
}

hipError_t launch_gemm(int prec, int epi, const Work* work, int nwork, const Problem* probs, const double* chunk_x,
                       hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  dim3 grid(nwork), block(NTHREADS);
#define MFTX_LAUNCH(P, E) hipLaunchKernelGGL((gemm_kernel<P, E>), grid, block, 0, s, work, probs, chunk_x)
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

}  // namespace mftx
