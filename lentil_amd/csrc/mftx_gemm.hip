// Complex x twiddle GEMM of the two MFT stages on gfx950 MFMA.
//
// A is data in HBM (packed, see mftx_internal.h); the twiddles
// cis(alpha * x_k * Y_s) are generated on the fly into LDS:
//     stage 1:  G   = f   . E2      (stored transposed, packed, as G^T)
//     stage 2:  F^T = G^T . E1^T    (|F|^2 accumulated, or F stored)
// which is the triple product of lentil fourier.py:97 with the twiddles of
// fourier.py:108-109.
//
// Centred output grid.  With the shift folded into A by its producer, each
// stage is C[i][j] = sum_k A'[i][k] cis(alpha x_k Y_j) with Y symmetric about
// 0, so one GEMM over the slots Y_s >= 0 with the real twiddle matrices
// cos and sin gives P = A' cos, Q = A' sin and both output columns
// C(+Y_s) = P + iQ and C(-Y_s) = P - iQ: half the MFMAs and half the
// twiddles of the direct complex product.
//
// Workgroup: 4 x WN waves, each wave a 32-row x 32-slot block (two output
// columns per slot) with accumulators P_re, P_im, Q_re, Q_im; C tile 128 rows
// x 32 WN slots.  K runs in chunks of 16 KS staged in double-buffered LDS, one
// barrier per chunk.  Two configurations:
//   WN = 1, KS = 1: 32 slots (64 output columns), 40 KiB LDS, two workgroups per CU;
//   WN = 2, KS = 2: 64 slots (128 output columns), 32-wide K chunks, 96 KiB LDS,
//                   one workgroup (8 waves) per CU.
// Two chunks ahead, in one of two register sets that alternate by chunk, each
// thread stages A items (32 bytes of one row: a BF16 plane or four complex64
// values; global -> registers -> LDS) and generates 2 KS consecutive twiddles
// of one slot, t_i = P * Q_i with P = cis(alpha * x * Y) (FP64 argument and
// range reduction, FP32 sin/cos) per chunk and the per-problem table
// Q_i = cis(alpha * i * Y).  The compiler schedules the loop (the twiddle VALU
// work lands after the chunk's barrier, next to the LDS stores that use it;
// scheduling hints and hand-threading it between the MFMAs measured no better).
//
// Precision forms
//   PREC_FP32   : complex64 A, v_mfma_f32_32x32x2_f32: 4 real MFMAs per
//                 (A element, twiddle) pair.
//   PREC_BF16X3 : every FP32 operand x is split x = hi + lo (two BF16, round
//                 to nearest even); a real product is hi*hi + hi*lo + lo*hi on
//                 v_mfma_f32_32x32x16_bf16 with FP32 accumulation, 12 MFMAs per
//                 32x32 slot block and 16-wide K step.
#include "mftx_device.h"

namespace mftx {

// Geometry of one tile configuration.
template <int WN, int KS> struct Cfg {
  static constexpr int BS = 32 * WN;               // slots per workgroup
  static constexpr int BKC = 16 * KS;              // K per chunk
  static constexpr int NT = 64 * WAVES_M * WN;     // threads per workgroup
  static constexpr int A_ITEMS = 512 * KS / NT;    // A staging items per thread
  static constexpr int TW = 2 * KS;                // twiddles per thread per chunk
  static constexpr int WG_PER_CU = WN == 1 ? 2 : 1;
  static_assert((512 * KS) % NT == 0, "A items must divide evenly");
  static_assert(BS * BKC == NT * TW, "one twiddle item per thread");
};

template <int PREC, int WN, int KS> struct Lds;

// bf16x3: 4 planes of [row][16 KS] bf16 (A: re_hi, re_lo, im_hi, im_lo; B:
// cos_hi, cos_lo, sin_hi, sin_lo), made of 16-byte units of 8 K values.  Units
// are swizzled per row, unit' = unit ^ swz_key(row).  For KS = 2 (64-byte
// rows) the key ((row >> 1) ^ (row >> 2)) & 3 keeps the 16-byte accesses of
// 8 consecutive rows on distinct banks of a 32-bank (128-byte) LDS line and
// those of 16 consecutive rows on distinct banks of a 64-bank line, so the
// fragment reads and the A staging stores are conflict-free whichever way the
// hardware groups the lanes of a 16-byte access (checked on the CPU for 8 and
// 16 lanes per cycle); (row >> 2) & 3 was conflict-free only for 16.
template <int WN, int KS> struct Lds<PREC_BF16X3, WN, KS> {
  __bf16 a[2][4][BM][16 * KS];
  __bf16 b[2][4][Cfg<WN, KS>::BS][16 * KS];
};
template <int KS> __device__ __forceinline__ int swz_key(int row) {
  return KS == 1 ? (row >> 3) & 1 : ((row >> 1) ^ (row >> 2)) & 3;
}
// element offset of logical unit u of a row
template <int KS> __device__ __forceinline__ int swz(int row, int u) { return 8 * (u ^ swz_key<KS>(row)); }

// fp32: complex64 [row][BKC + 2] (B: (cos, sin) pairs); the padded row stride
// keeps the 16-byte stores of 16 consecutive rows and the 8-byte fragment
// reads on distinct banks.
template <int WN, int KS> struct Lds<PREC_FP32, WN, KS> {
  static constexpr int RS = 16 * KS + 2;
  float2 a[2][BM][RS];
  float2 b[2][Cfg<WN, KS>::BS][RS];
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const u32x4 gu32x4;  // global-memory view

// A staging: 512 KS items of 32 bytes per chunk (128 rows x 4 parts x KS
// 16-wide K steps; a part is a BF16 plane or four complex64 values).  Item i
// is row (i & 15) + 16 ((i >> 6) & 7), part (i >> 4) & 3, K step i >> 9, so
// 16 consecutive lanes store 16 consecutive rows of one part (conflict-free
// LDS stores).  Thread t stages items t + j * NT.
// B generation: BS x 8 items (slot, group of TW K values), one per thread.
// KS = 1: slot (t & 15) + 16 (t >> 7), group (t >> 4) & 7.  KS = 2 (8-byte
// stores of 4 K values): lane l of wave w takes slot (l & 7) + 8 ((l >> 4) & 1)
// + 16 (w >> 1) and group ((l >> 3) & 1) + 2 ((l >> 5) & 1) + 4 (w & 1), so 16
// consecutive lanes store 8 slots x the two halves of one 16-byte unit: 128
// bytes on distinct banks (and 32 lanes 256 bytes), with the KS = 2 key.
__device__ __forceinline__ int item_row(int i) { return (i & 15) + 16 * ((i >> 6) & 7); }
__device__ __forceinline__ int item_part(int i) { return (i >> 4) & 3; }
__device__ __forceinline__ int item_kstep(int i) { return i >> 9; }
template <int KS> __device__ __forceinline__ int bitem_slot(int t) {
  if (KS == 1) return (t & 15) + 16 * (t >> 7);
  const int l = t & 63;
  return (l & 7) + 8 * ((l >> 4) & 1) + 16 * (t >> 7);
}
template <int KS> __device__ __forceinline__ int bitem_group(int t) {
  if (KS == 1) return (t >> 4) & 7;
  const int l = t & 63;
  return ((l >> 3) & 1) + 2 * ((l >> 5) & 1) + 4 * ((t >> 6) & 1);
}

// Staged chunk in registers: A prefetch and the generated twiddles.
// bf16x3: b[p] holds plane p of the TW K values (TW / 2 words); fp32: the TW
// (cos, sin) pairs, two per 16 bytes.  Native vector types keep these in VGPRs.
template <int WN, int KS> struct Staged {
  u32x4 a[Cfg<WN, KS>::A_ITEMS][2];
  u32x4 b[4];
};

// Per-problem staging parameters (uniform) and the per-thread Q table.
struct StageProb {
  const char* a;
  long long lda;
  double alpha4;   // 4 * alpha: twiddle arguments in quarter turns
  double y;        // Y of this thread's slot
  int nch;         // chunks in the problem
  int cb;          // chunk_begin (index of the problem's first 16-wide K block in chunk_x)
};

template <int KS, int NQ>
__device__ __forceinline__ void set_stage(const Problem& P, int slot, StageProb& sp, float2 (&q)[NQ]) {
  sp.a = (const char*)P.a;
  sp.lda = P.lda;
  sp.alpha4 = 4.0 * P.alpha;
  sp.y = P.y0 + (double)slot;
  sp.nch = P.kpad / (16 * KS);
  sp.cb = P.chunk_begin;
#pragma unroll
  for (int i = 0; i < NQ; ++i) q[i] = cis_quarter(sp.alpha4 * ((double)i * sp.y));
}

// x of the first twiddle of group g of chunk c (chunk_x holds one origin per
// 16-wide K block).  The block origins are read at a wave-uniform address
// (scalar loads through the scalar cache) and selected per lane: a per-lane
// index would make this a vector load whose latency sits on every chunk's
// critical path.
__device__ __forceinline__ double uniform_double(const double* p) {
  const long long b = *(const long long*)p;
  const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int KS>
__device__ __forceinline__ double twiddle_x(const double* __restrict__ chunk_x, const StageProb& sp, int c, int g) {
  const int k = 2 * KS * g;
  const double* px = chunk_x + __builtin_amdgcn_readfirstlane(sp.cb + KS * c);  // uniform by construction
  const double x0 = uniform_double(px);
  const double xb = KS == 2 ? ((k >> 4) ? uniform_double(px + 1) : x0) : x0;
  return xb + (double)(k & 15);
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

// TW consecutive twiddles from K origin x for this thread's slot.
template <int PREC, int WN, int KS>
__device__ __forceinline__ void gen_b(double x, const StageProb& sp, const float2 (&q)[2 * KS],
                                      Staged<WN, KS>& st) {
  constexpr int TW = 2 * KS;
  const float2 pf = cis_quarter(sp.alpha4 * (x * sp.y));
  float2 t[TW];
#pragma unroll
  for (int i = 0; i < TW; ++i) t[i] = cmul(pf, q[i]);
  if (PREC == PREC_BF16X3) {
    uint32_t ch[2] = {}, cl[2] = {}, sh[2] = {}, sl[2] = {};
#pragma unroll
    for (int i = 0; i < TW / 2; ++i) {
      split2(t[2 * i].x, t[2 * i + 1].x, ch[i], cl[i]);
      split2(t[2 * i].y, t[2 * i + 1].y, sh[i], sl[i]);
    }
    st.b[0] = u32x4{ch[0], ch[1], 0u, 0u};
    st.b[1] = u32x4{cl[0], cl[1], 0u, 0u};
    st.b[2] = u32x4{sh[0], sh[1], 0u, 0u};
    st.b[3] = u32x4{sl[0], sl[1], 0u, 0u};
  } else {
    auto bits = [](float v) { return __float_as_uint(v); };
#pragma unroll
    for (int i = 0; i < TW / 2; ++i)
      st.b[i] = u32x4{bits(t[2 * i].x), bits(t[2 * i].y), bits(t[2 * i + 1].x), bits(t[2 * i + 1].y)};
  }
}

template <int PREC, int WN, int KS>
__device__ __forceinline__ void write_b(Lds<PREC, WN, KS>& L, int buf, int slot, int g, const Staged<WN, KS>& st) {
  constexpr int TW = 2 * KS;
  if (PREC == PREC_BF16X3) {
    auto& LB = reinterpret_cast<Lds<PREC_BF16X3, WN, KS>&>(L);
    const int k = TW * g;  // first K of the group: unit k / 8, offset k % 8
    const int e = swz<KS>(slot, k >> 3) + (k & 7);
    if (KS == 1) {
#pragma unroll
      for (int p = 0; p < 4; ++p) *(uint32_t*)&LB.b[buf][p][slot][e] = st.b[p][0];
    } else {
#pragma unroll
      for (int p = 0; p < 4; ++p) *(u32x2*)&LB.b[buf][p][slot][e] = u32x2{st.b[p][0], st.b[p][1]};
    }
  } else {
    auto& LF = reinterpret_cast<Lds<PREC_FP32, WN, KS>&>(L);
    u32x4* b = (u32x4*)&LF.b[buf][slot][g * TW];
#pragma unroll
    for (int i = 0; i < TW / 2; ++i) b[i] = st.b[i];
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

// One 32x32 slot block, K = 16: P += A cos, Q += A sin for A = (re, im), each
// real product as hi*hi + hi*lo + lo*hi.  a: planes (re_hi, re_lo, im_hi,
// im_lo); b: planes (cos_hi, cos_lo, sin_hi, sin_lo).  acc: P_re, P_im, Q_re, Q_im.
__device__ __forceinline__ void sblock_bf16(const bf16x8_t (&a)[4], const bf16x8_t (&b)[4], f32x16 (&acc)[4]) {
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int ia = t == 2 ? 1 : 0;   // hi, hi, lo  of A
    const int ib = t == 1 ? 1 : 0;   // hi, lo, hi  of the twiddle
    acc[0] = MFMA_BF16(a[ia], b[ib], acc[0]);
    acc[1] = MFMA_BF16(a[2 + ia], b[ib], acc[1]);
    acc[2] = MFMA_BF16(a[ia], b[2 + ib], acc[2]);
    acc[3] = MFMA_BF16(a[2 + ia], b[2 + ib], acc[3]);
  }
}

// Weight-map values at map offsets idx[0..n-1] (1 without a map).  No load
// sits behind a branch, so they are in flight together.
template <int NW>
__device__ __forceinline__ void load_weights(const Problem& P, const long long (&idx)[NW], float (&wv)[NW]) {
  if (!P.wmap) {
#pragma unroll
    for (int t = 0; t < NW; ++t) wv[t] = 1.0f;
  } else if (P.wdtype == 3) {
#pragma unroll
    for (int t = 0; t < NW; ++t) wv[t] = (float)((gptr<const double>)P.wmap)[idx[t]];
  } else {
#pragma unroll
    for (int t = 0; t < NW; ++t) wv[t] = ((gptr<const float>)P.wmap)[idx[t]];
  }
}

// End of one problem.  Accumulator element e of a lane is C row
// 8 (e / 4) + 4 (lane / 32) + e % 4 of its 32-row block and slot lane % 32;
// the slot's two output columns are jp = cp + s (P + iQ) and jm = cm - s
// (P - iQ).  Intensity accumulates w |C|^2 for both in registers; the store
// forms write C * scale (complex, or packed transposed as the next stage's A';
// the next stage's fold is already in C, from the phase of this stage's
// operand rows).
template <int PREC, int EPI>
__device__ __forceinline__ void problem_epilogue(const Problem& P, const Work& wk, int BS, int wr, int wc, int lane,
                                                 const f32x16 (&acc)[4], f32x16 (&ip), f32x16 (&im)) {
  const int s = wk.tc * BS + wc * 32 + (lane & 31);
  const int N = P.ncols;
  const int jp = P.cp + s, jm = P.cm - s;
  if (EPI == EPI_INTENSITY) {
    const float w = P.scale;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const float pr = acc[0][e], pi = acc[1][e], qr = acc[2][e], qi = acc[3][e];
      const float ar = pr - qi, ai = pi + qr;   // P + iQ
      const float br = pr + qi, bi = pi - qr;   // P - iQ
      ip[e] = __builtin_fmaf(w, __builtin_fmaf(ar, ar, ai * ai), ip[e]);
      im[e] = __builtin_fmaf(w, __builtin_fmaf(br, br, bi * bi), im[e]);
    }
    return;
  }
  const float sc = P.scale;
  const bool has_p = jp < N, has_m = jm >= 0 && jm < N && jm != jp;
  // the first C row of this wave's block, opaque to the compiler: the
  // indices below then stay in the epilogue instead of being hoisted out of
  // the K loop and spilled
  int rblk = wk.tr * BM + wr * 32;
  asm volatile("" : "+v"(rblk));
  if (EPI == EPI_STORE_CPLX && P.trans) {
    // Transposed complex store: C row r of output column j goes to
    // dst[j * ldd + r].  As for the packed store, one lane swap per word
    // gives lane l rows 0-7 and lane l+32 rows 8-15 of each 16-row block, so
    // a lane stores 64 contiguous bytes as 16-byte pieces.
    float2* dst = (float2*)P.dst;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float v[2][2][8];  // [side][group of the pair][re, im of rows e = 0..3]
#pragma unroll
      for (int gg = 0; gg < 2; ++gg) {
        const int g = 2 * h + gg;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float pr = acc[0][4 * g + e], pi = acc[1][4 * g + e], qr = acc[2][4 * g + e], qi = acc[3][4 * g + e];
          v[0][gg][2 * e] = (pr - qi) * sc;
          v[0][gg][2 * e + 1] = (pi + qr) * sc;
          v[1][gg][2 * e] = (pr + qi) * sc;
          v[1][gg][2 * e + 1] = (pi - qr) * sc;
        }
      }
      const int r0 = rblk + 16 * h + 8 * (lane >> 5);  // first of this lane's 8 rows
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        float o[16];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[side][0][q]),
                                                          __float_as_uint(v[side][1][q]), false, false);
          o[q] = __uint_as_float(r[0]);
          o[8 + q] = __uint_as_float(r[1]);
        }
        if (!(side ? has_m : has_p)) continue;
        const int j = side ? jm : jp;
        // rows past P.rows (not stored) read the column's row 0
        long long wi[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) wi[t] = (long long)j * P.ldw + (r0 + t < P.rows ? r0 + t : 0);
        float wv[8];
        load_weights(P, wi, wv);
        float2* d = dst + (long long)j * P.ldd + r0;
        if (r0 + 8 <= P.rows && (((uintptr_t)d) & 15) == 0) {
#pragma unroll
          for (int t = 0; t < 4; ++t)
            ((gptr<v4f32>)d)[t] = v4f32{o[4 * t] * wv[2 * t], o[4 * t + 1] * wv[2 * t], o[4 * t + 2] * wv[2 * t + 1],
                                        o[4 * t + 3] * wv[2 * t + 1]};
        } else {
#pragma unroll
          for (int t = 0; t < 8; ++t)
            if (r0 + t < P.rows) ((gptr<v2f32>)d)[t] = v2f32{o[2 * t] * wv[t], o[2 * t + 1] * wv[t]};
        }
      }
    }
    return;
  }
  if (EPI == EPI_STORE_PACKED_T && PREC == PREC_BF16X3) {
    // Lane l holds K 0-3 (group 2h) and 8-11 (group 2h+1) of a 16-K block of
    // its output row, lane l+32 K 4-7 and 12-15.  One lane swap per word
    // (v_permlane32_swap: the high half of the first operand trades places
    // with the low half of the second) gives lane l K 0-7 and lane l+32 K
    // 8-15, so each lane stores 16 bytes of every plane.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int kb = wk.tr * BM + wr * 32 + 16 * h;  // first K of the block (wave-uniform)
      if (kb >= P.limit) continue;
      uint32_t w[2][2][4][2];  // [group of the pair][side][plane][word]
#pragma unroll
      for (int gg = 0; gg < 2; ++gg) {
        const int g = 2 * h + gg;
        float v[2][2][4];  // [side][re, im][e]
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float pr = acc[0][4 * g + e], pi = acc[1][4 * g + e], qr = acc[2][4 * g + e], qi = acc[3][4 * g + e];
          v[0][0][e] = (pr - qi) * sc;
          v[0][1][e] = (pi + qr) * sc;
          v[1][0][e] = (pr + qi) * sc;
          v[1][1][e] = (pi - qr) * sc;
        }
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          split2(v[side][0][0], v[side][0][1], w[gg][side][0][0], w[gg][side][1][0]);
          split2(v[side][0][2], v[side][0][3], w[gg][side][0][1], w[gg][side][1][1]);
          split2(v[side][1][0], v[side][1][1], w[gg][side][2][0], w[gg][side][3][0]);
          split2(v[side][1][2], v[side][1][3], w[gg][side][2][1], w[gg][side][3][1]);
        }
      }
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        uint32_t o[4][4];
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) {
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const auto r = __builtin_amdgcn_permlane32_swap(w[0][side][pl][q], w[1][side][pl][q], false, false);
            o[pl][q] = r[0];
            o[pl][2 + q] = r[1];
          }
        }
        if (side ? has_m : has_p) {
          __bf16* d = (__bf16*)P.dst + (long long)(side ? jm : jp) * P.ldd + (long long)(kb >> 4) * 64 + 8 * (lane >> 5);
#pragma unroll
          for (int pl = 0; pl < 4; ++pl)
            *(gptr<v4u32>)(d + 16 * pl) = v4u32{o[pl][0], o[pl][1], o[pl][2], o[pl][3]};
        }
      }
    }
    return;
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    // rows row0 .. row0+3 of C in this lane
    const int row0 = rblk + 8 * g + 4 * (lane >> 5);
    float pre[4], pim[4], mre[4], mim[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float pr = acc[0][4 * g + e], pi = acc[1][4 * g + e], qr = acc[2][4 * g + e], qi = acc[3][4 * g + e];
      pre[e] = (pr - qi) * sc;
      pim[e] = (pi + qr) * sc;
      mre[e] = (pr + qi) * sc;
      mim[e] = (pi - qr) * sc;
    }
    if (EPI == EPI_STORE_CPLX) {
      float2* dst = (float2*)P.dst;
      // optional real weight map, indexed like the destination; an element
      // that is not stored reads the map's first value
      long long wi[8];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = row0 + e;
        const bool in = row < P.rows;
        wi[e] = in && has_p ? (P.trans ? (long long)jp * P.ldw + row : (long long)row * P.ldw + jp) : 0;
        wi[4 + e] = in && has_m ? (P.trans ? (long long)jm * P.ldw + row : (long long)row * P.ldw + jm) : 0;
      }
      float wv[8];
      load_weights(P, wi, wv);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = row0 + e;
        if (row < P.rows) {
          if (has_p) {
            const long long o = P.trans ? (long long)jp * P.ldd + row : (long long)row * P.ldd + jp;
            ((gptr<v2f32>)dst)[o] = v2f32{pre[e] * wv[e], pim[e] * wv[e]};
          }
          if (has_m) {
            const long long o = P.trans ? (long long)jm * P.ldd + row : (long long)row * P.ldd + jm;
            ((gptr<v2f32>)dst)[o] = v2f32{mre[e] * wv[4 + e], mim[e] * wv[4 + e]};
          }
        }
      }
    } else if (row0 < P.limit) {  // EPI_STORE_PACKED_T: destination row = output column, K = C row
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int j = side ? jm : jp;
        if (side ? !has_m : !has_p) continue;
        const float* re = side ? mre : pre;
        const float* imv = side ? mim : pim;
        if (PREC == PREC_BF16X3) {
          uint32_t rh0, rl0, rh1, rl1, ih0, il0, ih1, il1;
          split2(re[0], re[1], rh0, rl0);
          split2(re[2], re[3], rh1, rl1);
          split2(imv[0], imv[1], ih0, il0);
          split2(imv[2], imv[3], ih1, il1);
          __bf16* d = (__bf16*)P.dst + (long long)j * P.ldd + (long long)(row0 >> 4) * 64 + (row0 & 15);
          *(gptr<v2u32>)(d + 0) = v2u32{rh0, rh1};
          *(gptr<v2u32>)(d + 16) = v2u32{rl0, rl1};
          *(gptr<v2u32>)(d + 32) = v2u32{ih0, ih1};
          *(gptr<v2u32>)(d + 48) = v2u32{il0, il1};
        } else {
          gptr<v4f32> d = (gptr<v4f32>)((float2*)P.dst + (long long)j * P.ldd + row0);
          d[0] = v4f32{re[0], imv[0], re[1], imv[1]};
          d[1] = v4f32{re[2], imv[2], re[3], imv[3]};
        }
      }
    }
  }
}

// The workgroup's summed intensity into its partial image (transposed like
// the problems' trans flag), clipped to the partial extent.
__device__ __forceinline__ void store_partial(const Work& wk, const Problem& P, int BS, int wr, int wc, int lane,
                                              const f32x16& ip, const f32x16& im) {
  const int s = wk.tc * BS + wc * 32 + (lane & 31);
  const int jp = P.cp + s, jm = P.cm - s;
  const bool has_p = jp < P.ncols, has_m = jm >= 0 && jm < P.ncols && jm != jp;
  const int trans = P.trans;
  if (trans) {
    // C rows run along the partial's rows, and a lane's elements 4g .. 4g+3
    // are 4 consecutive C rows: one aligned 16-byte store each (the partial
    // and its row stride are multiples of 128 floats).  Rows past pcols
    // land in the partial's padding columns, inside its allocation, which
    // the reduction never reads.
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int row = wk.tr * BM + wr * 32 + 8 * g + 4 * (lane >> 5);
      if (has_p && jp < wk.prows)
        *(gptr<v4f32>)&((gptr<float>)wk.partial)[(long long)jp * wk.ldp + row] =
            v4f32{ip[4 * g], ip[4 * g + 1], ip[4 * g + 2], ip[4 * g + 3]};
      if (has_m && jm < wk.prows)
        *(gptr<v4f32>)&((gptr<float>)wk.partial)[(long long)jm * wk.ldp + row] =
            v4f32{im[4 * g], im[4 * g + 1], im[4 * g + 2], im[4 * g + 3]};
    }
    return;
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int row = wk.tr * BM + wr * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
    if (has_p) {
      const int pr = trans ? jp : row, pc = trans ? row : jp;
      if (pr < wk.prows && pc < wk.pcols) ((gptr<float>)wk.partial)[(long long)pr * wk.ldp + pc] = ip[e];
    }
    if (has_m) {
      const int pr = trans ? jm : row, pc = trans ? row : jm;
      if (pr < wk.prows && pc < wk.pcols) ((gptr<float>)wk.partial)[(long long)pr * wk.ldp + pc] = im[e];
    }
  }
}

// The work item's loop.  MF = false for a wave whose 32 rows all lie past
// the problems' last row (the ragged last row tile of a small tile): it
// stages its share of the operands and keeps the barriers but issues no
// fragment reads or MFMAs, and its zero accumulators give the same zero rows
// the padding would, so the wave sharing its SIMD runs alone.
template <int PREC, int EPI, int WN, int KS, bool MF>
__device__ __forceinline__ void gemm_body(const Work& wk, const Problem* __restrict__ probs,
                                          const double* __restrict__ chunk_x, Lds<PREC, WN, KS>& lds) {
  using C = Cfg<WN, KS>;
  constexpr int BS = C::BS;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wr = wid / WN, wc = wid % WN;

  // staging coordinates: A rows from a_row0, one slot / K group of twiddles
  const long long a_row0 = (long long)wk.tr * BM;
  const int st_slot = bitem_slot<KS>(tid);
  const int st_g = bitem_group<KS>(tid);
  const int b_slot = wk.tc * BS + st_slot;

  f32x16 acc[4], ip, im;
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = f32x16{};
  ip = f32x16{};
  im = f32x16{};

  // ---- prologue: chunk 0 into LDS buffer 0, chunks 1 and 2 staged in the two
  // register sets (chunk k in set k & 1), so each global load has two chunks
  // of work to arrive in ----
  int sp = wk.pbeg, ss = 0;  // staging cursor: next (problem, chunk) to stage
  StageProb sgp;
  float2 q[C::TW];
  Staged<WN, KS> st0, st1;
  set_stage<KS>(probs[sp], b_slot, sgp, q);
  double sx = twiddle_x<KS>(chunk_x, sgp, 0, st_g);
  auto advance = [&]() {
    if (++ss == sgp.nch) {
      ss = 0;
      if (++sp < wk.pend) set_stage<KS>(probs[sp], b_slot, sgp, q);
    }
    sx = twiddle_x<KS>(chunk_x, sgp, sp < wk.pend ? ss : 0, st_g);
  };
  // stage the cursor's chunk (past the last chunk: a valid chunk restaged, never read)
  auto stage = [&](Staged<WN, KS>& X) {
    const bool more = sp < wk.pend;
    load_a<PREC, WN, KS>(sgp, a_row0, tid, more ? ss : 0, X);
    gen_b<PREC, WN, KS>(sx, sgp, q, X);
    if (more) advance();
  };
  stage(st0);
  write_b<PREC, WN, KS>(lds, 0, st_slot, st_g, st0);
  write_a<PREC, WN, KS>(lds, 0, tid, st0);
  stage(st1);
  stage(st0);
  __syncthreads();

  int cur = 0;
  int pc = wk.pbeg, cs = 0, cn = probs[pc].kpad / C::BKC;  // compute cursor: problem, chunk, chunks
  // Chunk s: store chunk s+1 from its register set X into the idle buffer (its
  // LDS stores then overlap this chunk's MFMAs), stage chunk s+3 into X,
  // compute chunk s, and close the problem after its last chunk.
  auto chunk = [&](Staged<WN, KS>& X) {
    write_b<PREC, WN, KS>(lds, cur ^ 1, st_slot, st_g, X);
    write_a<PREC, WN, KS>(lds, cur ^ 1, tid, X);
    stage(X);

    if (!MF) {
      // no fragment reads or MFMAs for this wave
    } else if (PREC == PREC_BF16X3) {
      auto& L = reinterpret_cast<Lds<PREC_BF16X3, WN, KS>&>(lds);
      const int r0 = wr * 32 + (lane & 31), c0 = wc * 32 + (lane & 31);
#pragma unroll
      for (int j = 0; j < KS; ++j) {
        // fragment rows are 16-aligned + (lane & 31): one swizzle key per lane
        const int koff = swz<KS>(lane & 31, 2 * j + (lane >> 5));
        bf16x8_t a[4], b[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          a[p] = *(const bf16x8_t*)&L.a[cur][p][r0][koff];
          b[p] = *(const bf16x8_t*)&L.b[cur][p][c0][koff];
        }
        sblock_bf16(a, b, acc);
      }
    } else {
      auto& L = reinterpret_cast<Lds<PREC_FP32, WN, KS>&>(lds);
#pragma unroll
      for (int ks = 0; ks < C::BKC / 2; ++ks) {
        const int kk = 2 * ks + (lane >> 5);
        const float2 a = L.a[cur][wr * 32 + (lane & 31)][kk];
        const float2 b = L.b[cur][wc * 32 + (lane & 31)][kk];
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b.x, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b.x, acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b.y, acc[2], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b.y, acc[3], 0, 0, 0);
      }
    }

    __syncthreads();
    cur ^= 1;
    if (++cs == cn) {
      problem_epilogue<PREC, EPI>(probs[pc], wk, BS, wr, wc, lane, acc, ip, im);
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = f32x16{};
      if (++pc < wk.pend) {
        cs = 0;
        cn = probs[pc].kpad / C::BKC;
      }
    }
  };
  // two chunks per trip, so each register set has a fixed name
  while (true) {
    chunk(st1);
    if (pc >= wk.pend) break;
    chunk(st0);
    if (pc >= wk.pend) break;
  }
  if (EPI == EPI_INTENSITY)
    store_partial(wk, probs[wk.pbeg], BS, wr, wc, lane, ip, im);
}

template <int PREC, int EPI, int WN, int KS>
__global__ __launch_bounds__((Cfg<WN, KS>::NT), (Cfg<WN, KS>::WG_PER_CU)) void gemm_kernel(
    const Work* __restrict__ works, const Problem* __restrict__ probs, const double* __restrict__ chunk_x) {
  __shared__ Lds<PREC, WN, KS> lds;
  const Work wk = works[blockIdx.x];
  if (wk.pbeg >= wk.pend) return;
  int rows = 0;
  for (int p = wk.pbeg; p < wk.pend; ++p) rows = max(rows, probs[p].rows);
  const int wr = (threadIdx.x >> 6) / WN;
  if (wk.tr * BM + wr * 32 < rows)  // wave-uniform
    gemm_body<PREC, EPI, WN, KS, true>(wk, probs, chunk_x, lds);
  else
    gemm_body<PREC, EPI, WN, KS, false>(wk, probs, chunk_x, lds);
}

// ---------------------------------------------------------------------------
// Warp-specialised form of the 128-row x 64-slot, 32-wide-K configuration
// (BF16x3).  Twelve waves per workgroup, three per SIMD:
//   consumers (waves 0-7): a 4 x 2 grid of 32-row x 32-slot blocks; each
//     issues the direct-to-LDS loads (global_load_lds_dwordx4, 16 bytes per
//     lane at M0 + 16 lane) of 16 rows of the A tile two chunks ahead, reads
//     its fragments from LDS and issues the 24 MFMAs of a chunk, and runs the
//     problem epilogues;
//   producers (waves 8-11): producer wave p generates the twiddles of K
//     group p (8 consecutive K values for each of the 64 slots) one chunk
//     ahead into LDS.
// The twiddle VALU work thus runs on a wave of its own on each SIMD while
// that SIMD's two consumer waves keep the MFMA pipe busy.  The A operand
// never passes through VGPRs, and its loads, which stall the issuing wave
// while the vector memory path is backed up, sit with the MFMA waves (the
// SIMD's other consumer keeps the pipe busy) rather than in series with the
// twiddle VALU.  LDS: A in three buffers (chunk s in s % 3), B in two
// (s % 2), 128 KiB.  An A buffer is row-major, a row's 256 bytes of the
// chunk (two 16-wide K blocks of 4 planes) in 16-byte units swizzled as
// unit ^ (row & 15): one load moves 4 whole rows (8 cache lines, every byte
// used), and the fragment reads of 16 (or 8) consecutive rows fall on
// distinct banks.  One barrier per chunk (in stage 2 before the chunk's last
// product group, so the next chunk's first fragment reads run under that
// group).  Consumers wait for their previous
// chunk's loads (vmcnt(4): the 4 loads just issued may stay in flight; loads
// complete in issue order, so epilogue stores in between do not weaken this)
// and their fragment reads, producers for their twiddle stores.  The barrier
// and waits are written out: an LDS fence would make the compiler drain every
// load, direct-to-LDS ones included, at each barrier.
namespace ws {
constexpr int NCONS = 8, NPROD = 4;
constexpr int NT = 64 * (NCONS + NPROD);
constexpr int BS = 64, BKC = 32;
constexpr int NA = 3, NB = 2;
// 16-byte alignment stated: the fragment reads are 16-byte LDS reads
// (ds_read_b128), which the compiler emits only where it can prove it.
struct alignas(16) Lds {
  __bf16 a[NA][BM][128];   // [row][16-byte unit ^ (row & 15)][8 K values]
  __bf16 b[NB][4][BS][32];
};
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) const void gvoid;
}  // namespace ws

// A consumer's A cursor: the chunk its next direct-to-LDS loads fetch.
// Consumer wave w loads rows 16 w .. 16 w + 15 of the tile, 4 rows per load:
// lane l takes row 16 w + 4 j + (l >> 4) of load j and LDS unit l & 15, which
// holds the row's logical unit (l & 15) ^ (row & 15) (row & 15 = 4 j + (l >> 4)).
struct WsACursor {
  int p, c, nch;              // problem, chunk in it, chunks in it
  const char* base;           // wave-uniform: the wave's first row (16 w) at chunk 0
  int voff[4];                // this lane's byte offset from base (at chunk 0) for load j
};

// The base is wave-uniform (row0 and w are), so it lives in SGPRs and each
// load adds only the chunk offset to it there; a lane's offset within the
// wave's 16 rows fits 32 bits (15 rows of a row stride below 128 MiB).
__device__ __forceinline__ void ws_a_set(WsACursor& ac, const Problem& P, long long row0, int lane, int w) {
  const long long ldb = P.lda * 2;  // bytes per A row
  const int rr = lane >> 4, pu = lane & 15;
  ac.nch = P.kpad / ws::BKC;
  ac.base = (const char*)P.a + (row0 + 16 * w) * ldb;
#pragma unroll
  for (int j = 0; j < 4; ++j) ac.voff[j] = (int)((4 * j + rr) * ldb) + 16 * (pu ^ (4 * j + rr));
}

// Direct-to-LDS load j (rows 16 w + 4 j .. + 3) of the cursor's chunk into A
// buffer ab: lane l's 16 bytes land at M0 + 16 l.  Written out rather than
// through the builtin: the compiler counts a builtin LDS DMA on the LDS/scalar
// counter too, and then waits for every fragment read at once (lgkmcnt(0))
// where a count would do; the load has no register result, and the waits for
// it (vmcnt) are explicit.  SALU writes of M0 need one wait state before an
// LDS DMA reads it.  The compiler reserves M0 and warns that it will not
// restore it across the statement; these loads are the only M0 readers and
// writers in the GEMM kernels (nothing else here uses M0), so nothing needs it
// restored.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void ws_a_load1(ws::Lds& L, int ab, int w, const WsACursor& ac, int j) {
  const char* src = ac.base + (long long)ac.c * 256;  // wave-uniform
  const unsigned dst = (unsigned)(size_t)(ws::lds_void*)&L.a[ab][16 * w + 4 * j][0];
  asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1"
               :: "v"(ac.voff[j]), "s"(src), "s"(dst) : "memory", "m0");
}
#pragma clang diagnostic pop

// All 4 loads of this wave's 16 rows of the cursor's chunk into A buffer ab.
__device__ __forceinline__ void ws_a_load(ws::Lds& L, int ab, int w, const WsACursor& ac) {
#pragma unroll
  for (int j = 0; j < 4; ++j) ws_a_load1(L, ab, w, ac, j);
}

// The producer's twiddle state for one problem: 8 consecutive K values of one
// slot per chunk, t_i = cis(alpha x Y) cis(alpha i Y), i = 0..7.
struct WsB {
  int p, c, nch, cb;
  double alpha4, y;
  float2 q[8];
};

// Per problem: the slot's phase rate and q[i] = cis(4 alpha i y), i = 0..7,
// from one FP64-phase evaluation and a product tree of depth 3 (about 3 ulp):
// eight full evaluations at every problem start made the twiddle waves late
// for that chunk's barrier.
__device__ __forceinline__ void ws_b_set(WsB& b, const Problem& P, int slot) {
  b.alpha4 = 4.0 * P.alpha;
  b.y = P.y0 + (double)slot;
  b.nch = P.kpad / ws::BKC;
  b.cb = P.chunk_begin;
  const float2 q1 = cis_quarter(b.alpha4 * b.y);
  const float2 q2 = cmul(q1, q1), q4 = cmul(q2, q2);
  b.q[0] = make_float2(1.0f, 0.0f);
  b.q[1] = q1;
  b.q[2] = q2;
  b.q[3] = cmul(q2, q1);
  b.q[4] = q4;
  b.q[5] = cmul(q4, q1);
  b.q[6] = cmul(q4, q2);
  b.q[7] = cmul(q4, b.q[3]);
}

// Twiddles of K group g (K 8g .. 8g+7 of the chunk) of this thread's slot into B buffer bb.
__device__ __forceinline__ void ws_b_write(ws::Lds& L, int bb, int slot, int g, const WsB& b,
                                           const double* __restrict__ chunk_x) {
  // chunk_x holds one origin per 16-wide K block; this group lies in block g >> 1
  const double* px = chunk_x + __builtin_amdgcn_readfirstlane(b.cb + 2 * b.c + (g >> 1));
  const double x = uniform_double(px) + (double)(8 * (g & 1));
  const float2 pf = cis_quarter(b.alpha4 * (x * b.y));
  uint32_t ch[4], cl[4], sh[4], sl[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float2 t0 = cmul(pf, b.q[2 * i]), t1 = cmul(pf, b.q[2 * i + 1]);
    split2(t0.x, t1.x, ch[i], cl[i]);
    split2(t0.y, t1.y, sh[i], sl[i]);
  }
  const int e = swz<2>(slot, g);
  *(u32x4*)&L.b[bb][0][slot][e] = u32x4{ch[0], ch[1], ch[2], ch[3]};
  *(u32x4*)&L.b[bb][1][slot][e] = u32x4{cl[0], cl[1], cl[2], cl[3]};
  *(u32x4*)&L.b[bb][2][slot][e] = u32x4{sh[0], sh[1], sh[2], sh[3]};
  *(u32x4*)&L.b[bb][3][slot][e] = u32x4{sl[0], sl[1], sl[2], sl[3]};
}

template <int EPI>
__device__ __forceinline__ void gemm_ws_body(const Work& wk, const Problem* __restrict__ probs,
                                             const double* __restrict__ chunk_x, ws::Lds& L) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform, in an SGPR
  int nchunks = 0;  // every problem's K loop back to back (uniform)
  for (int p = wk.pbeg; p < wk.pend; ++p) nchunks += probs[p].kpad / ws::BKC;

  if (wid >= ws::NCONS) {
    // ---------------- producer ----------------
    const int pw = wid - ws::NCONS;           // K group of the twiddles
    const int slot = lane;                    // this thread's twiddle slot in the tile
    WsB b;
    b.p = wk.pbeg;
    b.c = 0;
    ws_b_set(b, probs[b.p], wk.tc * ws::BS + slot);
    auto b_next = [&]() {
      if (++b.c == b.nch) {
        b.c = 0;
        if (++b.p < wk.pend) ws_b_set(b, probs[b.p], wk.tc * ws::BS + slot);
      }
    };
    ws_b_write(L, 0, slot, pw, b, chunk_x);
    b_next();
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    for (int s = 0; s < nchunks; ++s) {
      if (s + 1 < nchunks) {
        ws_b_write(L, (s + 1) & 1, slot, pw, b, chunk_x);
        b_next();
      }
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    }
    return;
  }

  // ---------------- consumer ----------------
  const int wr = wid >> 1, wc = wid & 1;
  int rows = 0;
  for (int p = wk.pbeg; p < wk.pend; ++p) rows = max(rows, probs[p].rows);
  const bool mf = wk.tr * BM + wr * 32 < rows;  // wave-uniform: past the last row a wave issues no MFMAs
  f32x16 acc[4], ip, im;
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = f32x16{};
  ip = f32x16{};
  im = f32x16{};
  const int r0 = wr * 32 + (lane & 31), c0 = wc * 32 + (lane & 31);
  const long long row0 = (long long)wk.tr * BM;
  WsACursor ac;
  ac.p = wk.pbeg;
  ac.c = 0;
  ws_a_set(ac, probs[ac.p], row0, lane, wid);
  // past the work's last chunk the cursor stays on it: the loads stay valid and nobody reads them
  auto a_next = [&]() {
    if (ac.c + 1 < ac.nch) {
      ++ac.c;
    } else if (ac.p + 1 < wk.pend) {
      ++ac.p;
      ac.c = 0;
      ws_a_set(ac, probs[ac.p], row0, lane, wid);
    }
  };
  ws_a_load(L, 0, wid, ac);
  a_next();
  ws_a_load(L, 1, wid, ac);
  a_next();
  asm volatile("s_waitcnt vmcnt(4)\n\ts_barrier" ::: "memory");
  int pc = wk.pbeg, cs = 0, cn = probs[pc].kpad / ws::BKC;
  int ab = 0, an = 2;  // A buffers of chunks s and s + 2
  // Operands of a 16-wide K block: A planes hi (re, im) and lo, twiddle
  // planes hi (cos, sin) and lo.
  const int ua = (lane >> 5) ^ (r0 & 15);  // A unit of (block 0, plane 0); block j, plane p: ua ^ (8 j + 2 p)
  auto rd_a = [&](int buf, int j, int p) { return *(const bf16x8_t*)&L.a[buf][r0][8 * (ua ^ (8 * j + 2 * p))]; };
  auto rd_b = [&](int buf, int j, int p) {
    return *(const bf16x8_t*)&L.b[buf][p][c0][swz<2>(lane & 31, 2 * j + (lane >> 5))];
  };
  auto mf4 = [&](const bf16x8_t (&x)[2], const bf16x8_t (&y)[2]) {
    acc[0] = MFMA_BF16(x[0], y[0], acc[0]);   // re cos
    acc[1] = MFMA_BF16(x[1], y[0], acc[1]);   // im cos
    acc[2] = MFMA_BF16(x[0], y[1], acc[2]);   // re sin
    acc[3] = MFMA_BF16(x[1], y[1], acc[3]);   // im sin
  };
  auto load = [&](int j) {
    __builtin_amdgcn_sched_barrier(0);
    ws_a_load1(L, an, wid, ac, j);
    __builtin_amdgcn_sched_barrier(0);
  };
  // Chunk s + 1's first fragments are read under chunk s's last product
  // group, across the chunk's barrier (the complex store, whose epilogue
  // needs those 16 registers at problem ends, reads them at the chunk's top).
  constexpr bool kAhead = EPI != EPI_STORE_CPLX;
  // Block 0's A hi and twiddle lo fragments of chunk s.
  bf16x8_t ah[2] = {}, bl[2] = {};
  // Nothing is outstanding on the LDS/scalar counter after this wait; the
  // compiler sees it (the barrier's own waits are in asm it cannot read), so
  // it waits for each fragment read below by count, not for all at once.
  auto drained = [&]() { __builtin_amdgcn_s_waitcnt(0xc07f); };  // lgkmcnt(0); vmcnt, expcnt unrestricted
  auto barrier = [&]() {
    // Wait for this wave's chunk s + 1 loads (the 4 just issued for chunk
    // s + 2 may stay in flight: loads complete in issue order, so epilogue
    // stores in between do not weaken this) and its fragment reads.
    asm volatile("s_waitcnt vmcnt(4) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    drained();
  };
  auto first_reads = [&](int abuf, int bbuf) {
    ah[0] = rd_a(abuf, 0, 0);
    ah[1] = rd_a(abuf, 0, 2);
    bl[0] = rd_b(bbuf, 0, 1);
    bl[1] = rd_b(bbuf, 0, 3);
  };
  drained();
  if (kAhead && mf) first_reads(0, 0);
  for (int s = 0; s < nchunks; ++s) {
    const int b0 = s & 1;
    if (mf) {
      if (!kAhead) first_reads(ab, b0);
      // The block's three products run in the order hi*lo, hi*hi, lo*hi, so
      // each operand pair is dead as early as possible and the next
      // fragments are read into its registers right after the MFMAs that
      // last read it: block 1's reads are in flight while block 0
      // multiplies, and the next chunk's first reads while the last group
      // multiplies, without extra registers.  The 4 direct-to-LDS loads of
      // chunk s + 2 sit between product groups, so the address unit works
      // through the workgroup's 32 loads while the MFMAs issue.  Reading
      // ahead, the chunk's barrier comes before its last group: by then
      // every fragment read of chunk s and every load of chunk s + 2 is
      // issued, and after it chunk s + 1 is complete in LDS.  Scheduling
      // barriers pin this order.
      bf16x8_t al[2] = {rd_a(ab, 0, 1), rd_a(ab, 0, 3)};
      bf16x8_t bh[2] = {rd_b(b0, 0, 0), rd_b(b0, 0, 2)};
      mf4(ah, bl);
      load(0);
      bl[0] = rd_b(b0, 1, 1);
      bl[1] = rd_b(b0, 1, 3);
      mf4(ah, bh);
      load(1);
      ah[0] = rd_a(ab, 1, 0);
      ah[1] = rd_a(ab, 1, 2);
      mf4(al, bh);
      __builtin_amdgcn_sched_barrier(0);
      bh[0] = rd_b(b0, 1, 0);
      bh[1] = rd_b(b0, 1, 2);
      al[0] = rd_a(ab, 1, 1);
      al[1] = rd_a(ab, 1, 3);
      __builtin_amdgcn_sched_barrier(0);
      mf4(ah, bl);
      load(2);
      mf4(ah, bh);
      load(3);
      if (kAhead) {
        barrier();
        __builtin_amdgcn_sched_barrier(0);
        // chunk s + 1's first fragments (past the last chunk: a buffer
        // nobody writes any more, read and unused)
        first_reads(ab == 2 ? 0 : ab + 1, b0 ^ 1);
        __builtin_amdgcn_sched_barrier(0);
      }
      mf4(al, bh);
      if (!kAhead) barrier();
    } else {
      ws_a_load(L, an, wid, ac);
      barrier();
    }
    a_next();
    an = an == 2 ? 0 : an + 1;
    ab = ab == 2 ? 0 : ab + 1;
    if (++cs == cn) {
      problem_epilogue<PREC_BF16X3, EPI>(probs[pc], wk, ws::BS, wr, wc, lane, acc, ip, im);
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = f32x16{};
      if (++pc < wk.pend) {
        cs = 0;
        cn = probs[pc].kpad / ws::BKC;
      }
    }
  }
  // no load may still be writing LDS when the workgroup ends
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  if (EPI == EPI_INTENSITY) store_partial(wk, probs[wk.pbeg], ws::BS, wr, wc, lane, ip, im);
}

template <int EPI>
__global__ __launch_bounds__(ws::NT, 1) void gemm_ws_kernel(const Work* __restrict__ works,
                                                            const Problem* __restrict__ probs,
                                                            const double* __restrict__ chunk_x) {
  __shared__ ws::Lds lds;
  const Work wk = works[blockIdx.x];
  if (wk.pbeg >= wk.pend) return;
  gemm_ws_body<EPI>(wk, probs, chunk_x, lds);
}

static hipError_t launch_ws(int epi, const Work* work, int nwork, const Problem* probs, const double* chunk_x,
                            hipStream_t s) {
  dim3 grid(nwork), block(ws::NT);
  if (epi == EPI_STORE_CPLX)
    hipLaunchKernelGGL((gemm_ws_kernel<EPI_STORE_CPLX>), grid, block, 0, s, work, probs, chunk_x);
  else if (epi == EPI_STORE_PACKED_T)
    hipLaunchKernelGGL((gemm_ws_kernel<EPI_STORE_PACKED_T>), grid, block, 0, s, work, probs, chunk_x);
  else
    hipLaunchKernelGGL((gemm_ws_kernel<EPI_INTENSITY>), grid, block, 0, s, work, probs, chunk_x);
  return hipGetLastError();
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
  if (bn == 64) return launch_cfg<1, 1>(prec, epi, work, nwork, probs, chunk_x, s);
  if (bn == 128) {
    if (prec == PREC_BF16X3) return launch_ws(epi, work, nwork, probs, chunk_x, s);
    return launch_cfg<2, 2>(prec, epi, work, nwork, probs, chunk_x, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace mftx
