// Complex GEMM C = A . B of the two MFT stages on gfx950 MFMA.
//
// A is data in HBM (packed, see mftx_internal.h); B is a twiddle matrix
// B[k][j] = cis(alpha * x_k * y_j) generated on the fly into LDS:
//     stage 1:  G   = f   . E2      (stored transposed, packed, as G^T)
//     stage 2:  F^T = G^T . E1^T    (|F|^2 accumulated, or F stored)
// which is the triple product of lentil fourier.py:97 with the twiddles of
// fourier.py:108-109.
//
// Workgroup: 2 x WN waves, each wave a 64 x 32 block of C as two MFMA 32x32
// tiles; C tile BM x BN = 128 x 32 WN.  K runs in chunks of BKC = 16 KS staged
// in double-buffered LDS, one barrier per chunk.  Two configurations:
//   WN = 2, KS = 1: 64-column tiles, 48 KiB LDS, two workgroups per CU;
//   WN = 4, KS = 2: 128-column tiles, 32-wide chunks, 128 KiB LDS, one
//                   workgroup (8 waves) per CU.
// One chunk ahead, each thread stages A items (32 bytes of one row: a BF16
// plane or four complex64 values; global -> registers -> LDS) and generates
// 4 KS consecutive twiddles of one B column, t_i = P * Q_i with
// P = cis(alpha * x * y) (FP64 argument and range reduction, FP32 sin/cos)
// per chunk and the per-problem table Q_i = cis(alpha * i * y).  The twiddle
// VALU work is interleaved with the MFMAs of the current chunk
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

// Geometry of one tile configuration.
template <int WN, int KS> struct Cfg {
  static constexpr int BN = 32 * WN;             // C cols per workgroup
  static constexpr int BKC = 16 * KS;            // K per chunk
  static constexpr int NT = 64 * WAVES_M * WN;   // threads per workgroup
  static constexpr int A_ITEMS = 512 * KS / NT;  // A staging items per thread
  static constexpr int TW = 4 * KS;              // twiddles per thread per chunk
  static constexpr int WG_PER_CU = WN <= 2 ? 2 : 1;
  static_assert((512 * KS) % NT == 0, "A items must divide evenly");
  static_assert(BN * BKC == NT * TW, "one twiddle item per thread");
};

template <int PREC, int WN, int KS> struct Lds;

// bf16x3: 4 planes (re_hi, re_lo, im_hi, im_lo) of [row][BKC] bf16, made of
// 16-byte units of 8 K values.  Units are swizzled per row,
// unit' = unit ^ swz_key(row), so the 16-byte accesses of 16 consecutive
// lanes (16 consecutive rows, one unit) cover all 64 banks once.
template <int WN, int KS> struct Lds<PREC_BF16X3, WN, KS> {
  __bf16 a[2][4][BM][16 * KS];
  __bf16 b[2][4][Cfg<WN, KS>::BN][16 * KS];
};
template <int KS> __device__ __forceinline__ int swz_key(int row) {
  return KS == 1 ? (row >> 3) & 1 : (row >> 2) & 3;
}
// element offset of logical unit u of a row
template <int KS> __device__ __forceinline__ int swz(int row, int u) { return 8 * (u ^ swz_key<KS>(row)); }

// fp32: complex64 [row][BKC + 2]; the padded row stride keeps the 16-byte
// stores of 16 consecutive rows and the 8-byte fragment reads on distinct banks.
template <int WN, int KS> struct Lds<PREC_FP32, WN, KS> {
  static constexpr int RS = 16 * KS + 2;
  float2 a[2][BM][RS];
  float2 b[2][Cfg<WN, KS>::BN][RS];
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const u32x4 gu32x4;  // global-memory view

// A staging: 512 KS items of 32 bytes per chunk (128 rows x 4 parts x KS
// 16-wide K steps; a part is a BF16 plane or four complex64 values).  Item i
// is row (i & 15) + 16 ((i >> 6) & 7), part (i >> 4) & 3, K step i >> 9, so
// 16 consecutive lanes store 16 consecutive rows of one part (conflict-free
// LDS stores).  Thread t stages items t + j * NT.
// B generation: BN x 4 items (column, group of TW K values), one per thread:
// column (t & 15) + 16 (t >> 6), group (t >> 4) & 3.
__device__ __forceinline__ int item_row(int i) { return (i & 15) + 16 * ((i >> 6) & 7); }
__device__ __forceinline__ int item_part(int i) { return (i >> 4) & 3; }
__device__ __forceinline__ int item_kstep(int i) { return i >> 9; }
__device__ __forceinline__ int bitem_col(int t) { return (t & 15) + 16 * (t >> 6); }

// Staged chunk in registers: A prefetch and the generated twiddles.
// bf16x3: b[p] holds plane p of the TW K values (TW / 2 words); fp32: the TW
// complex values, 2 per 16 bytes.  Native vector types keep these in VGPRs.
template <int WN, int KS> struct Staged {
  u32x4 a[Cfg<WN, KS>::A_ITEMS][2];
  u32x4 b[4];
};

// Per-problem staging parameters (uniform) and the per-thread Q table.
struct StageProb {
  const char* a;
  long long lda;
  double alpha4;   // 4 * alpha: twiddle arguments in quarter turns
  double y;        // y of this thread's B column
  int nch;         // chunks in the problem
  int cb;          // chunk_begin (index of the problem's first 16-wide K block in chunk_x)
};

template <int KS, int NQ>
__device__ __forceinline__ void set_stage(const Problem& P, int col, StageProb& sp, float2 (&q)[NQ]) {
  sp.a = (const char*)P.a;
  sp.lda = P.lda;
  sp.alpha4 = 4.0 * P.alpha;
  sp.y = P.y0 + (double)col;
  sp.nch = P.kpad / (16 * KS);
  sp.cb = P.chunk_begin;
#pragma unroll
  for (int i = 0; i < NQ; ++i) q[i] = cis_quarter(sp.alpha4 * ((double)i * sp.y));
}

// x of the first twiddle of group g of chunk c (chunk_x holds one origin per
// 16-wide K block).
template <int KS>
__device__ __forceinline__ double twiddle_x(const double* __restrict__ chunk_x, const StageProb& sp, int c, int g) {
  const int k = 4 * KS * g;
  return chunk_x[sp.cb + KS * c + (k >> 4)] + (double)(k & 15);
}

template <int PREC, int WN, int KS>
__device__ __forceinline__ void load_a(const StageProb& sp, long long row0, int tid, int chunk,
                                       Staged<WN, KS>& st) {
#pragma unroll
  for (int j = 0; j < Cfg<WN, KS>::A_ITEMS; ++j) {
    const int it = tid + j * Cfg<WN, KS>::NT;
    const long long row = row0 + item_row(it);
    const long long blk = (long long)chunk * KS + item_kstep(it);  // 16-wide K block
    gu32x4* src;
    if (PREC == PREC_BF16X3)
      src = (gu32x4*)((const __bf16*)sp.a + row * sp.lda + blk * 64 + item_part(it) * 16);
    else
      src = (gu32x4*)((const float2*)sp.a + row * sp.lda + blk * 16 + item_part(it) * 4);
    st.a[j][0] = src[0];
    st.a[j][1] = src[1];
  }
}

// TW consecutive twiddles from K origin x for this thread's column.
template <int PREC, int WN, int KS>
__device__ __forceinline__ void gen_b(double x, const StageProb& sp, const float2 (&q)[4 * KS],
                                      Staged<WN, KS>& st) {
  constexpr int TW = 4 * KS;
  const float2 pf = cis_quarter(sp.alpha4 * (x * sp.y));
  float2 t[TW];
#pragma unroll
  for (int i = 0; i < TW; ++i) t[i] = cmul(pf, q[i]);
  if (PREC == PREC_BF16X3) {
    uint32_t rh[4] = {}, rl[4] = {}, ih[4] = {}, il[4] = {};
#pragma unroll
    for (int i = 0; i < TW / 2; ++i) {
      split2(t[2 * i].x, t[2 * i + 1].x, rh[i], rl[i]);
      split2(t[2 * i].y, t[2 * i + 1].y, ih[i], il[i]);
    }
    st.b[0] = u32x4{rh[0], rh[1], rh[2], rh[3]};
    st.b[1] = u32x4{rl[0], rl[1], rl[2], rl[3]};
    st.b[2] = u32x4{ih[0], ih[1], ih[2], ih[3]};
    st.b[3] = u32x4{il[0], il[1], il[2], il[3]};
  } else {
    auto bits = [](float v) { return __float_as_uint(v); };
#pragma unroll
    for (int i = 0; i < TW / 2; ++i)
      st.b[i] = u32x4{bits(t[2 * i].x), bits(t[2 * i].y), bits(t[2 * i + 1].x), bits(t[2 * i + 1].y)};
  }
}

template <int PREC, int WN, int KS>
__device__ __forceinline__ void write_b(Lds<PREC, WN, KS>& L, int buf, int col, int g, const Staged<WN, KS>& st) {
  if (PREC == PREC_BF16X3) {
    auto& LB = reinterpret_cast<Lds<PREC_BF16X3, WN, KS>&>(L);
    if (KS == 1) {  // 4 K values: half of unit g >> 1
      const int k = swz<KS>(col, g >> 1) + 4 * (g & 1);
#pragma unroll
      for (int p = 0; p < 4; ++p) *(u32x2*)&LB.b[buf][p][col][k] = u32x2{st.b[p][0], st.b[p][1]};
    } else {        // 8 K values: unit g
      const int k = swz<KS>(col, g);
#pragma unroll
      for (int p = 0; p < 4; ++p) *(u32x4*)&LB.b[buf][p][col][k] = st.b[p];
    }
  } else {
    auto& LF = reinterpret_cast<Lds<PREC_FP32, WN, KS>&>(L);
    u32x4* b = (u32x4*)&LF.b[buf][col][g * 4 * KS];
#pragma unroll
    for (int i = 0; i < 2 * KS; ++i) b[i] = st.b[i];
  }
}

template <int PREC, int WN, int KS>
__device__ __forceinline__ void write_a(Lds<PREC, WN, KS>& L, int buf, int tid, const Staged<WN, KS>& st) {
#pragma unroll
  for (int j = 0; j < Cfg<WN, KS>::A_ITEMS; ++j) {
    const int it = tid + j * Cfg<WN, KS>::NT;
    const int row = item_row(it), part = item_part(it), ks = item_kstep(it);
    if (PREC == PREC_BF16X3) {
      auto& LB = reinterpret_cast<Lds<PREC_BF16X3, WN, KS>&>(L);
      *(u32x4*)&LB.a[buf][part][row][swz<KS>(row, 2 * ks)] = st.a[j][0];
      *(u32x4*)&LB.a[buf][part][row][swz<KS>(row, 2 * ks + 1)] = st.a[j][1];
    } else {
      auto& LF = reinterpret_cast<Lds<PREC_FP32, WN, KS>&>(L);
      u32x4* a = (u32x4*)&LF.a[buf][row][16 * ks + part * 4];
      a[0] = st.a[j][0];
      a[1] = st.a[j][1];
    }
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

// End of one problem: intensity accumulates w * |C|^2 in registers; the
// store forms write C * scale (complex, or packed transposed as the next
// stage's A operand).  Accumulator element e of a lane's 32x32 tile i is C
// row 8 (e / 4) + 4 (lane / 32) + e % 4, column lane % 32.
template <int PREC, int EPI>
__device__ __forceinline__ void problem_epilogue(const Problem& P, const Work& wk, int BN, int wr, int wc, int lane,
                                                 const f32x16 (&cr)[2], const f32x16 (&ci)[2], f32x16 (&iacc)[2]) {
  if (EPI == EPI_INTENSITY) {
    const float w = P.scale;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 16; ++e)
        iacc[i][e] = __builtin_fmaf(w, __builtin_fmaf(cr[i][e], cr[i][e], ci[i][e] * ci[i][e]), iacc[i][e]);
  } else {
    const float sc = P.scale;
    const int col = wk.tc * BN + wc * 32 + (lane & 31);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        // rows row0 .. row0+3 of C in this lane (32x32 accumulator layout)
        const int row0 = wk.tr * BM + wr * 64 + i * 32 + 8 * g + 4 * (lane >> 5);
        float re[4], im[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          re[e] = cr[i][4 * g + e] * sc;
          im[e] = ci[i][4 * g + e] * sc;
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
}

// The workgroup's summed intensity tile into its partial image (transposed
// like the problems' trans flag), clipped to the partial extent.
__device__ __forceinline__ void store_partial(const Work& wk, int trans, int BN, int wr, int wc, int lane,
                                              const f32x16 (&iacc)[2]) {
  const int col = wk.tc * BN + wc * 32 + (lane & 31);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int row = wk.tr * BM + wr * 64 + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
      const int pr = trans ? col : row;
      const int pcn = trans ? row : col;
      if (pr < wk.prows && pcn < wk.pcols) wk.partial[(long long)pr * wk.ldp + pcn] = iacc[i][e];
    }
  }
}

template <int PREC, int EPI, int WN, int KS>
__global__ __launch_bounds__((Cfg<WN, KS>::NT), (Cfg<WN, KS>::WG_PER_CU)) void gemm_kernel(
    const Work* __restrict__ works, const Problem* __restrict__ probs, const double* __restrict__ chunk_x) {
  using C = Cfg<WN, KS>;
  constexpr int BN = C::BN;
  __shared__ Lds<PREC, WN, KS> lds;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wr = wid / WN, wc = wid % WN;
  const Work wk = works[blockIdx.x];
  if (wk.pbeg >= wk.pend) return;

  // staging coordinates: A rows from a_row0, one B column / K group
  const long long a_row0 = (long long)wk.tr * BM;
  const int st_col = bitem_col(tid);
  const int st_g = item_part(tid);
  const int b_col = wk.tc * BN + st_col;

  f32x16 cr[2], ci[2], iacc[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    cr[i] = f32x16{};
    ci[i] = f32x16{};
    iacc[i] = f32x16{};
  }

  // ---- prologue: stage chunk 0 of the first problem into buffer 0 ----------
  int sp = wk.pbeg, ss = 0;  // staging cursor: next (problem, chunk) to stage
  StageProb sgp;
  float2 q[C::TW];
  Staged<WN, KS> st;
  set_stage<KS>(probs[sp], b_col, sgp, q);
  double sx = twiddle_x<KS>(chunk_x, sgp, 0, st_g);
  load_a<PREC, WN, KS>(sgp, a_row0, tid, 0, st);
  gen_b<PREC, WN, KS>(sx, sgp, q, st);
  write_b<PREC, WN, KS>(lds, 0, st_col, st_g, st);
  write_a<PREC, WN, KS>(lds, 0, tid, st);
  if (++ss == sgp.nch) {
    ss = 0;
    if (++sp < wk.pend) set_stage<KS>(probs[sp], b_col, sgp, q);
  }
  sx = twiddle_x<KS>(chunk_x, sgp, ss, st_g);
  __syncthreads();

  int cur = 0;
  for (int pc = wk.pbeg; pc < wk.pend; ++pc) {
    const int nch = probs[pc].kpad / C::BKC;
    for (int s = 0; s < nch; ++s) {
      // stage the next chunk (after the last one: restage a valid chunk into
      // the idle buffer, never read)
      const bool more = sp < wk.pend;
      const int ts = more ? ss : 0;
      load_a<PREC, WN, KS>(sgp, a_row0, tid, ts, st);
      gen_b<PREC, WN, KS>(sx, sgp, q, st);

      if (PREC == PREC_BF16X3) {
        auto& L = reinterpret_cast<Lds<PREC_BF16X3, WN, KS>&>(lds);
        const int r0 = wr * 64 + (lane & 31), c0 = wc * 32 + (lane & 31);
#pragma unroll
        for (int j = 0; j < KS; ++j) {
          // fragment rows are 32-aligned + (lane & 31): one swizzle key per lane
          const int koff = swz<KS>(lane & 31, 2 * j + (lane >> 5));
          bf16x8_t a0[4], a1[4], b[4], nb[2];
#pragma unroll
          for (int p = 0; p < 4; ++p) {
            a0[p] = *(const bf16x8_t*)&L.a[cur][p][r0][koff];
            b[p] = *(const bf16x8_t*)&L.b[cur][p][c0][koff];
          }
#pragma unroll
          for (int p = 0; p < 4; ++p) a1[p] = *(const bf16x8_t*)&L.a[cur][p][r0 + 32][koff];
          nb[0] = neg_bf16x8(b[2]);
          nb[1] = neg_bf16x8(b[3]);
          cblock_bf16(a0, b, nb, cr[0], ci[0]);
          cblock_bf16(a1, b, nb, cr[1], ci[1]);
        }
        // schedule: next chunk's global loads first (consumed only by the LDS
        // stores after the MFMAs), fragment reads, then the twiddle VALU work
        // threaded between the MFMAs
        SGB(SG_VMEM_RD, 2 * C::A_ITEMS);
        SGB(SG_DSR, 8);
#pragma unroll
        for (int i = 0; i < 12; ++i) {
          SGB(SG_MFMA, 1);
          SGB(SG_VALU, 4);
        }
        SGB(SG_DSR, 4);
#pragma unroll
        for (int i = 0; i < 12; ++i) {
          SGB(SG_MFMA, 1);
          SGB(SG_VALU, 2);
        }
        if (KS == 2) {
          SGB(SG_DSR, 12);
#pragma unroll
          for (int i = 0; i < 24; ++i) {
            SGB(SG_MFMA, 1);
            SGB(SG_VALU, 2);
          }
        }
      } else {
        auto& L = reinterpret_cast<Lds<PREC_FP32, WN, KS>&>(lds);
#pragma unroll
        for (int ks = 0; ks < C::BKC / 2; ++ks) {
          const int kk = 2 * ks + (lane >> 5);
          float2 a[2];
#pragma unroll
          for (int i = 0; i < 2; ++i) a[i] = L.a[cur][wr * 64 + i * 32 + (lane & 31)][kk];
          const float2 b = L.b[cur][wc * 32 + (lane & 31)][kk];
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            f32x16 r = cr[i], m = ci[i];
            r = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].x, b.x, r, 0, 0, 0);
            m = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].x, b.y, m, 0, 0, 0);
            r = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].y, -b.y, r, 0, 0, 0);
            m = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].y, b.x, m, 0, 0, 0);
            cr[i] = r;
            ci[i] = m;
          }
        }
      }

      write_b<PREC, WN, KS>(lds, cur ^ 1, st_col, st_g, st);
      write_a<PREC, WN, KS>(lds, cur ^ 1, tid, st);
      __syncthreads();
      cur ^= 1;
      if (more) {
        if (++ss == sgp.nch) {
          ss = 0;
          if (++sp < wk.pend) set_stage<KS>(probs[sp], b_col, sgp, q);
        }
        sx = twiddle_x<KS>(chunk_x, sgp, sp < wk.pend ? ss : 0, st_g);
      }
    }
    problem_epilogue<PREC, EPI>(probs[pc], wk, BN, wr, wc, lane, cr, ci, iacc);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      cr[i] = f32x16{};
      ci[i] = f32x16{};
    }
  }
  if (EPI == EPI_INTENSITY) store_partial(wk, probs[wk.pbeg].trans, BN, wr, wc, lane, iacc);
}

template <int WN, int KS>
static hipError_t launch_cfg(int prec, int epi, const Work* work, int nwork, const Problem* probs,
                             const double* chunk_x, hipStream_t s) {
  dim3 grid(nwork), block(Cfg<WN, KS>::NT);
#define MFTX_LAUNCH(P, E) \
  hipLaunchKernelGGL((gemm_kernel<P, E, WN, KS>), grid, block, 0, s, work, probs, chunk_x)
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

hipError_t launch_gemm(int prec, int epi, int bn, const Work* work, int nwork, const Problem* probs,
                       const double* chunk_x, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  if (bn == 64) return launch_cfg<2, 1>(prec, epi, work, nwork, probs, chunk_x, s);
  if (bn == 128) return launch_cfg<4, 2>(prec, epi, work, nwork, probs, chunk_x, s);
  return hipErrorInvalidValue;
}

}  // namespace mftx
