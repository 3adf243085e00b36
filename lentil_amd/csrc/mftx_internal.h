// Internal declarations shared by the MFT kernels and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct mftx_ctx;

namespace mftx {

// Tile geometry of the complex GEMM: WAVES_M x WN waves per workgroup, each
// wave a 32-row x 32-slot block (64 output columns, see Problem); BM rows and
// 32 WN slots per workgroup.
constexpr int WAVES_M = 4;
constexpr int BM = 32 * WAVES_M;   // C rows per workgroup (A rows)
constexpr int BK = 16;             // K block: granule of the packed layout and of chunk_x
constexpr int BN_DEFAULT = 128;    // output columns per workgroup unless a plan picks another
// output columns per workgroup: 64 (32 slots, 16-wide K steps) or 128 (64 slots, 32-wide K steps)
inline bool valid_bn(int bn) { return bn == 64 || bn == 128; }
// K per GEMM pipeline stage for a tile width; every GEMM problem's kpad is a multiple of it.
inline int k_granule(int bn) { return bn == 128 ? 32 : 16; }
// slots per workgroup for a tile width
inline int slots_per_tile(int bn) { return bn / 2; }
static_assert(BM == 128, "staging assumes 128 A rows per workgroup");

enum Precision { PREC_FP32 = 0, PREC_BF16X3 = 1 };

// Epilogues of the GEMM.
enum Epilogue {
  EPI_STORE_CPLX = 0,   // dst complex64, (r, c) -> dst[c*ld + r] if trans else dst[r*ld + c]; * scale; bounds
  EPI_STORE_PACKED_T = 1,  // dst packed A-operand layout (precision dependent) of C^T, rows r < limit
  EPI_INTENSITY = 2,    // accumulate w*|C|^2 over the workgroup's problems into a partial image
};

// One complex GEMM problem: one MFT stage,
//   C[i][j] = sum_k A[i][k] cis(alpha x_k (j - N/2 - shift)),   0 <= j < N = ncols,
// with cis(t) = exp(2 pi i t) and N/2 the integer half, computed about a
// pairing centre c (integer or half-integer; the host picks it at the PSF core
// j = N/2 + shift or at the grid centre next to it).  The producer of A has
// multiplied A[i][k] by the fold cis(-alpha (shift + N/2 - c) x_k), so that
// with Y = j - c
//   C[i][j] = sum_k A'[i][k] cis(alpha x_k Y),
// whose cosine part is even and sine part odd in Y.  The GEMM runs over slots
// s = 0 .. cols-1 with Y_s = y0 + s (y0 = 0 for integer c, 1/2 otherwise):
// P = sum_k A' cos, Q = sum_k A' sin, and
//   C[i][cp + s] = P + i Q,    C[i][cm - s] = P - i Q
// (cp = ceil(c), cm = floor(c); outputs outside [0, N) are dropped, and the
// second is dropped where it equals the first).  c = 0 is the direct product
// (one output per slot).  Twiddle B[k][s] = cis(alpha * x_k * Y_s) with
// x_k = chunk_x[chunk_begin + k/BK] + (k % BK): x advances by one per K index
// inside each BK-wide block, and each block carries its own starting x (so
// several K segments with their own coordinate origins can be concatenated).
struct Problem {
  const void* a;        // packed A' operand (row-major, K-blocked)
  long long lda;        // A row stride in scalar elements (float2 for fp32, bf16 for bf16x3)
  double alpha;         // signed twiddle scale in turns per unit of x*y
  double y0;            // Y of slot 0
  int rows;             // valid C rows
  int cols;             // slots
  int ncols;            // N: output columns
  int cp, cm;           // output columns of slot 0 (P + iQ, P - iQ)
  int kpad;             // K (multiple of the launch's k_granule); A is zero beyond the valid K
  int chunk_begin;      // index of this problem's first entry in chunk_x
  float scale;          // store: C *= scale; intensity: weight applied to |C|^2
  int limit;            // EPI_STORE_PACKED_T: C rows written (K extent of the destination)
  int trans;            // EPI_STORE_CPLX / EPI_INTENSITY: write C transposed
  void* dst;            // store destination (EPI_STORE_*)
  long long ldd;        // destination row stride (elements of its type)
  const void* wmap;     // EPI_STORE_CPLX: optional real map multiplied into each stored value,
  long long ldw;        //   indexed like the destination (row stride ldw)
  int wdtype;           //   1 float32, 3 float64
  int pad2_;
};

// Per-wavelength parameters of a phasor launch: the phase opd * inv_wl, the
// stage-1 fold cis(fold * (x0 + k)) along the operand's K index k, and the
// stage-2 fold cis(fold2 * (x02 + r)) along its row index r (stage 1's C row r
// is the operand's row r times the row's phase, so the stage-1 output comes
// out folded for stage 2).
struct PhasorWave {
  double inv_wl;
  double fold;
  double x0;
  double fold2;
  double x02;
};

// A workgroup's unit of work: problems [pbeg, pend) on tile (tr, tc) of C.
// For EPI_INTENSITY the problems' |C|^2 are summed into partial[r][c]
// (row-major, stride ldp, transposed like the problems' trans flag).
struct Work {
  int pbeg, pend;
  int tr, tc;
  float* partial;
  long long ldp;
  int prows, pcols;     // partial image extent (bounds for the store)
};

// Packed A-operand layouts.
//  fp32   : complex64 row-major, element (r, k) at r*lda + k.
//  bf16x3 : per row, K in blocks of 16; each block holds 4 planes of 16 bf16
//           (re_hi, re_lo, im_hi, im_lo): element (r, k, p) at
//           r*lda + (k/16)*64 + p*16 + (k%16), lda = 4*kpad.
__host__ __device__ inline long long packed_lda(int prec, int kpad) { return prec == PREC_FP32 ? kpad : 4LL * kpad; }
__host__ __device__ inline size_t packed_bytes(int prec, long long rows, int kpad) {
  return (size_t)rows * (size_t)kpad * (prec == PREC_FP32 ? 8 : 8);
}

// One phasor operand of a plan: the packed (transposed when trans) operand
// of an m x n tile for nwave wavelengths, written wave_stride bytes apart.
// `begin` is the op's first thread in the grouped launch.
constexpr int PH_WAVES = 8;  // wavelengths per phasor thread
struct PhasorOp {
  const void* amp;
  const void* opd;
  int amp_dtype, opd_dtype;
  long long ld_amp, ld_opd;
  int m, n, trans, nwave;
  const PhasorWave* waves;
  void* dst;
  long long wave_stride, rows_pad;
  int kpad, pad_;
  long long begin;
};
__host__ __device__ inline long long phasor_threads(long long rows_pad, int kpad, int nwave) {
  return rows_pad * (kpad / 8) * ((nwave + PH_WAVES - 1) / PH_WAVES);
}

// Grouped launches (phasor and pack ops): blocks of kGroupBlock threads, and
// every op's first thread is a multiple of kGroupBlock, so a block belongs to
// one op and finds it with a block-uniform search.
constexpr int kGroupBlock = 256;
inline long long group_begin(long long total) { return (total + kGroupBlock - 1) / kGroupBlock * kGroupBlock; }

// One packed operand from complex data (mftx_dft2_many): src (or src^T) times
// the fold cis(fold * (x0 + k)) along K; rows_pad * kpad / 2 threads from `begin`.
struct PackOp {
  const void* src;
  int dtype, trans;
  long long m, n, lds;
  double fold, x0;      // stage-1 fold along K
  double fold2, x02;    // stage-2 fold along the packed rows (see PhasorWave)
  void* dst;
  long long rows_pad;
  int kpad, pad_;
  long long begin;
};

// launchers (mftx_gemm.hip, mftx_kernels.hip)
// bn: output columns per workgroup (64 or 128, see valid_bn).
hipError_t launch_gemm(int prec, int epi, int bn, const Work* work, int nwork, const Problem* probs,
                       const double* chunk_x, hipStream_t s);
// packed A' = src (or src^T) times the folds cis(fold * (x0 + k)) along K and
// cis(fold2 * (x02 + row)) along the packed rows
hipError_t launch_pack(int prec, const void* src, int src_dtype, long long m, long long n,
                       long long lds, int trans, double fold, double x0, double fold2, double x02, void* dst,
                       long long rows_pad, int kpad, hipStream_t s);
// every op of a plan in one launch (ops_dev on the device, sorted by begin)
hipError_t launch_phasor_group(int prec, const PhasorOp* ops_dev, int nops, long long total_threads, hipStream_t s);
// every complex pack op in one launch (ops_dev on the device, sorted by begin)
hipError_t launch_pack_group(int prec, const PackOp* ops_dev, int nops, long long total_threads, hipStream_t s);
hipError_t launch_phasor_field(const void* amp, int amp_dtype, long long ld_amp, const void* opd,
                               int opd_dtype, long long ld_opd, int m, int n, double inv_wavelength,
                               void* dst, hipStream_t s);
hipError_t launch_reduce_partials(const float* partial, long long group_stride, int ngroups,
                                  int rows, int cols, long long ldp, void* out, int out_dtype,
                                  long long ldo, int out_rows, int out_cols, int r0, int c0,
                                  int accumulate, hipStream_t s);
hipError_t launch_insert(const void* field, long long ldf, int fr0, int fc0, int h, int w,
                         void* out, int out_dtype, long long ldo, int or0, int oc0,
                         int intensity, double weight, hipStream_t s);
hipError_t launch_multiply(const void* a, long long lda, const void* b, long long ldb, void* out,
                           long long ldo, int h, int w, hipStream_t s);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
// rows x cols elements of esz (4 or 8) bytes at row stride ld elements
hipError_t launch_zero_rows(void* p, int esz, long long rows, long long cols, long long ld, hipStream_t s);
hipError_t launch_weight_field(const void* F, long long ldF, const void* w, int w_dtype, long long ldw,
                               double scale, void* out, long long ldo, int h, int wd, hipStream_t s);
// nwave wavelengths (inv_wls on the device, or inv_wl0 alone when null), D_w at D + w * strideD
hipError_t launch_phase_grad(const void* amp, int amp_dtype, long long ld_amp, const void* opd, int opd_dtype,
                             long long ld_opd, int m, int n, const double* inv_wls, double inv_wl0, int nwave,
                             const void* D, long long ldD, long long strideD, double* g_opd, long long ldg,
                             double* g_amp, long long ldga, hipStream_t s);

// detector-grid post-processing (mftx_detector.hip)
hipError_t launch_rebin(const void* src, int dtype, long long nimg, long long rows, long long cols, long long ld,
                        int factor, void* dst, int odt, hipStream_t s);
// *out = sum of a real array (part: one double per 256 elements)
hipError_t launch_sum(const void* x, int dtype, long long rows, long long cols, long long ld, double* part,
                      double* out, hipStream_t s);
// out = |F| (F complex64), *sum = sum of out (part as for launch_sum)
hipError_t launch_magnitude(const void* F, long long ldF, long long rows, long long cols, void* out, int odt,
                            long long ldo, double* part, double* sum, hipStream_t s);
// x *= *num / *den
hipError_t launch_scale_ratio(void* x, int dtype, long long rows, long long cols, long long ld, const double* num,
                              const double* den, hipStream_t s);
// map_coordinates(order 0/1/3, mode 'nearest') resampling with the mask and
// unitary scaling of rescale; P, T: (rows + 24) x (cols + 24) doubles; raw,
// mval: orows x ocols doubles; part: one double per 256 elements of the
// larger of the input and output; sums: 2 doubles
// The B-spline prefilter of a spline order as its two-sided impulse response
// h[|k|], |k| <= n (n = 0 for orders 0 and 1).
constexpr int kPrefilterTaps = 48;
struct PrefilterTaps {
  int n;
  double h[kPrefilterTaps + 1];
};
PrefilterTaps prefilter_taps(int order);
// rescale (util.py:262-349): map_coordinates(order 0..5, mode MFTX_MODE_*)
// of a real or complex (dtype 0 / 2) image, mask, unitary ratio; P and T hold
// the padded image (rows + 24) x (cols + 24) doubles each, raw / raw_im / mval
// the output size, part ceil(max(padded, output) / 256) doubles, sums 4.
hipError_t launch_rescale(const void* img, int dtype, long long rows, long long cols, long long ld, const void* msrc,
                          int mask_mode, int mdt, long long ldm, double eps, double scale, long long orows,
                          long long ocols, int order, int mode, int unitary, void* out, int odt, long long ldo,
                          double* P, double* T, double* raw, double* raw_im, double* mval, double* part, double* sums,
                          hipStream_t s);

// modal OPD operations (mftx_modal.hip)
// out[seg][9] = normal equations of the piston/tip/tilt fit; part: nseg * ceil(rows*cols/256) * 9 doubles
hipError_t launch_ptt_normal(const void* opd, int odt, long long ld_opd, const void* masks, int mdt, long long ld_mask,
                             long long mask_stride, int nseg, long long rows, long long cols, double dx_r, double dx_c,
                             double* part, double* out, hipStream_t s);
hipError_t launch_remove_tilt(const void* opd, int odt, long long ld_opd, const void* masks, int mdt,
                              long long ld_mask, long long mask_stride, int nseg, int multi, long long rows,
                              long long cols, double dx_r, double dx_c, const double* tilt, double* out,
                              long long ldo, hipStream_t s);
hipError_t launch_zernike(const double* rho, const double* theta, const void* mask, int mdt, long long npix,
                          int nmodes, const int* desc, const double* coef, const int* pw, const double* norm,
                          double* out, hipStream_t s);

// host side of the C ABI (mftx_api.cpp)
// mftx_dft2 on the caller's workspace ws (cap bytes), or on the context's own
// growable workspace when ws is null
int dft2_on_workspace(mftx_ctx* ctx, const void* f, int f_dtype, int64_t m, int64_t n, int64_t ldf, void* F,
                      int64_t M, int64_t N, int64_t ldF, double alpha_r, double alpha_c, double shift_r,
                      double shift_c, double offset_r, double offset_c, double scale, int conj, int precision,
                      void* ws, size_t cap);
// record msg as the context's last error and return code
int set_error(mftx_ctx* ctx, int code, const char* msg);
// the device a context was created on
int context_device(const mftx_ctx* ctx);
// device address of the no-allocation entry point's scratch array on `device`
// (current), looked up once per device (mftx_noalloc.hip)
hipError_t scratch_address(int device, void** out);

}  // namespace mftx
