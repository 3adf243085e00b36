/*
 * mftx — MI355X matrix Fourier transform library (C ABI).
 *
 * The drop-in boundary of the lentil Fraunhofer path.  The reference is pure
 * Python + NumPy and has no FFI, so each entry point here replaces the NumPy
 * call sites of one reference function; the Python package lentil_amd binds
 * them with ctypes and keeps lentil's own signatures above them.
 *
 * Conventions
 *   - every function returns 0 on success or a negative MFTX_E* code; the
 *     message of the last failure on a context is mftx_last_error(ctx);
 *   - all array arguments are device pointers, row-major, with explicit row
 *     strides in elements; the caller owns every buffer it passes;
 *   - work is queued on the context's stream (mftx_set_stream) and runs
 *     asynchronously; the library allocates only context- or plan-owned
 *     workspace, never a caller buffer;
 *   - dtype codes: 0 complex64, 1 float32, 2 complex128, 3 float64;
 *   - precision codes: 0 FP32 (complex64 operands, FP32 MFMA),
 *     1 BF16X3 (split hi/lo BF16 operands, three MFMA passes, FP32 accumulate).
 */
#ifndef MFTX_H
#define MFTX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFTX_OK 0
#define MFTX_EINVAL -1   /* bad argument: shape, dtype, pointer (maps to ValueError/TypeError) */
#define MFTX_ENOMEM -2   /* workspace allocation failed */
#define MFTX_EDEVICE -3  /* HIP runtime error */
#define MFTX_EHANDLE -4  /* null or destroyed handle */

#define MFTX_DTYPE_C64 0
#define MFTX_DTYPE_F32 1
#define MFTX_DTYPE_C128 2
#define MFTX_DTYPE_F64 3

#define MFTX_PREC_FP32 0
#define MFTX_PREC_BF16X3 1

typedef struct mftx_ctx mftx_ctx;
typedef struct mftx_plan mftx_plan;

/* Library version (major*10000 + minor*100 + patch). */
int mftx_version(void);

/* Context: owns a device, a stream binding and a growable workspace.
 * No counterpart in the reference (NumPy needs no context).  Plans keep a
 * pointer to the context they were created on: destroy them first. */
int mftx_create(int device, mftx_ctx** out);
int mftx_destroy(mftx_ctx* ctx);
const char* mftx_last_error(const mftx_ctx* ctx);
/* Queue subsequent work on `stream` (a hipStream_t; NULL = default stream). */
int mftx_set_stream(mftx_ctx* ctx, void* stream);
/* Block until the context's stream has drained. */
int mftx_synchronize(mftx_ctx* ctx);

/*
 * One matrix Fourier transform:
 *   F[u][v] = scale * sum_{r,s} f[r][s] * exp(sgn*2*pi*i*(ar*(R_r+or)*(U_u-sr) + ac*(S_s+oc)*(V_v-sc)))
 * with R_r = r - floor(m/2), S_s = s - floor(n/2), U_u = u - floor(M/2),
 * V_v = v - floor(N/2), sgn = -1 (conj == 0) or +1 (conj == 1).
 *
 * Replaces lentil fourier.dft2 (fourier.py:5-103): twiddles of
 * _dft2_matrices (fourier.py:106-110), the two np.dot calls (fourier.py:97)
 * and the unitary scale (fourier.py:100-101, passed in `scale`).  idft2
 * (fourier.py:124-198) is conj == 1 with scale = unitary/F.size; the adjoint
 * of dft2 is conj == 1 with the roles of (shift, offset) swapped and negated.
 *
 * f: m x n, dtype f_dtype (any of the four codes), row stride ldf.
 * F: M x N complex64, row stride ldF.  F must not overlap f.
 */
int mftx_dft2(mftx_ctx* ctx, const void* f, int f_dtype, int64_t m, int64_t n, int64_t ldf,
              void* F, int64_t M, int64_t N, int64_t ldF,
              double alpha_r, double alpha_c, double shift_r, double shift_c,
              double offset_r, double offset_c, double scale, int conj, int precision);

/*
 * mftx_dft2 without any device allocation.  No counterpart in the
 * reference: this is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as mftx_dft2, but the
 * packed operands live in a fixed array of MFTX_NOALLOC_SCRATCH_BYTES of
 * device memory that the library holds on every GPU it runs on until the
 * process ends (its address is looked up once per device, at mftx_create).
 * A transform whose operands need more than that (about a 1024^2 -> 2048^2
 * transform) fails with MFTX_EINVAL and a message giving both sizes.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (other host threads, or streams with no order between them) race on it;
 * calls queued on one stream are ordered and safe.
 */
#define MFTX_NOALLOC_SCRATCH_BYTES (32u << 20)
int mftx_dft2_noalloc(mftx_ctx* ctx, const void* f, int f_dtype, int64_t m, int64_t n, int64_t ldf,
                      void* F, int64_t M, int64_t N, int64_t ldF,
                      double alpha_r, double alpha_c, double shift_r, double shift_c,
                      double offset_r, double offset_c, double scale, int conj, int precision);

/*
 * Pupil phasor  out = amp * exp(2*pi*i*opd/wavelength)  (complex64, m x n, dense).
 * Replaces the phasor of Plane.__mul__ (plane.py:503-510).  amp and opd are
 * float32 or float64 (dtype codes 1 or 3) with row strides ld_amp, ld_opd.
 */
int mftx_phasor(mftx_ctx* ctx, const void* amp, int amp_dtype, int64_t ld_amp,
                const void* opd, int opd_dtype, int64_t ld_opd, int64_t m, int64_t n,
                double wavelength, void* out);

/*
 * out[or0 + r][oc0 + c] += weight * |E|^2   (intensity != 0, out float32/float64)
 * out[or0 + r][oc0 + c] += weight * E       (intensity == 0, out complex64/complex128)
 * for E = field[fr0 + r][fc0 + c], 0 <= r < h, 0 <= c < w (both slices within
 * their row strides: fc0 + w <= ldf, oc0 + w <= ldo; h, w <= 2^20).
 * Replaces the accumulation of field.insert (field.py:301-304); the caller
 * passes the clipped slices of field.py:259-299.  Coherent merging of
 * overlapping Fields (field._merge, field.py:339-342) is the complex form
 * into a zeroed union buffer.
 */
int mftx_insert(mftx_ctx* ctx, const void* field, int64_t ldf, int64_t fr0, int64_t fc0,
                int64_t h, int64_t w, void* out, int out_dtype, int64_t ldo,
                int64_t or0, int64_t oc0, int intensity, double weight);

/*
 * out = a * b elementwise over h x w complex64 arrays (row strides lda, ldb, ldo).
 * Replaces the Field product of Field.__mul__ (field.py:142) on the
 * intersection slices computed by the host (field.py:130-147).
 */
int mftx_multiply(mftx_ctx* ctx, const void* a, int64_t lda, const void* b, int64_t ldb,
                  void* out, int64_t ldo, int64_t h, int64_t w);

/* ---------------------------------------------------------------------------
 * Phase-retrieval backward pass.  The reference has no gradients; these
 * two element-wise steps plus mftx_dft2 with conj == 1 (the adjoint MFT)
 * give dL/d opd and dL/d amp of weighted broadband images, recomputing the
 * forward fields instead of storing them.
 * ------------------------------------------------------------------------- */

/* out = scale * w * F elementwise (F, out complex64; w float32/float64, h x wd):
 * the cotangent weight * dL/dI * F of I = weight * |F|^2. */
int mftx_weight_field(mftx_ctx* ctx, const void* F, int64_t ldF, const void* w, int w_dtype, int64_t ldw,
                      double scale, void* out, int64_t ldo, int64_t h, int64_t wd);

/* Accumulate the OPD (and optionally amplitude) gradient of
 * f = amp * exp(2*pi*i*opd/wavelength) given D = dL/d conj(f) (complex64):
 *   grad_opd += -(4*pi/wavelength) * amp * Im(conj(D) exp(i theta))
 *   grad_amp += 2 * Re(conj(D) exp(i theta))          (grad_amp may be NULL)
 * grad arrays are float64, m x n. */
int mftx_phase_grad(mftx_ctx* ctx, const void* amp, int amp_dtype, int64_t ld_amp, const void* opd, int opd_dtype,
                    int64_t ld_opd, int64_t m, int64_t n, double wavelength, const void* D, int64_t ldD,
                    double* grad_opd, int64_t ld_gopd, double* grad_amp, int64_t ld_gamp);

/* mftx_phase_grad summed over nwave wavelengths in one launch: D holds one
 * m x n complex64 cotangent per wavelength, D_w at D + w * strideD elements,
 * and wavelengths[w] (host memory) is its wavelength.  No reference
 * counterpart (the gradient of a broadband image for phase retrieval). */
int mftx_phase_grad_many(mftx_ctx* ctx, const void* amp, int amp_dtype, int64_t ld_amp, const void* opd,
                         int opd_dtype, int64_t ld_opd, int64_t m, int64_t n, const double* wavelengths,
                         int32_t nwave, const void* D, int64_t ldD, int64_t strideD, double* grad_opd,
                         int64_t ld_gopd, double* grad_amp, int64_t ld_gamp);

/* ---------------------------------------------------------------------------
 * Many independent MFTs in one grouped call (no reference counterpart: the
 * reference runs one fourier.dft2 per call).  Each entry is one mftx_dft2
 * (same arguments and meaning), with two additions:
 *   - when f is NULL the input is the pupil phasor
 *     amp * exp(2*pi*i*opd*inv_wl) (plane.py:503-510) generated in place;
 *   - when w is not NULL the output is multiplied elementwise by the real
 *     M x N map w (float32/float64, row stride ldw): the weighted cotangent
 *     of an intensity for the adjoint pass of phase retrieval.
 * All entries run in one pack launch and one launch per GEMM stage.
 * ------------------------------------------------------------------------- */
typedef struct {
  const void* f;       /* input m x n (f_dtype, row stride ldf), or NULL for a phasor */
  const void* amp;     /* phasor amplitude and OPD (dtype 1 or 3), used when f is NULL */
  const void* opd;
  int32_t f_dtype, amp_dtype, opd_dtype, conj;
  int64_t m, n, ldf, ld_amp, ld_opd;
  double inv_wl;       /* 1 / wavelength of the phasor */
  void* F;             /* output M x N complex64, row stride ldF */
  int64_t M, N, ldF;
  double alpha_r, alpha_c, shift_r, shift_c, offset_r, offset_c, scale;
  const void* w;       /* optional real weight map */
  int32_t w_dtype, pad_;
  int64_t ldw;
} mftx_mft;
int mftx_dft2_many(mftx_ctx* ctx, const mftx_mft* mfts, int32_t count, int32_t precision);

/* Persistent form of mftx_dft2_many for a list that runs many times (no
 * reference counterpart; iterative callers such as phase retrieval): the
 * descriptors and the operand workspace are built once in plan-owned device
 * memory, and each execute issues only the launches, on the context's
 * current stream.  The entries' arrays must stay allocated; their contents
 * may change between executes.  Plans on one context run in stream order. */
typedef struct mftx_mft_plan mftx_mft_plan;
int mftx_mft_plan_create(mftx_ctx* ctx, const mftx_mft* mfts, int32_t count, int32_t precision, mftx_mft_plan** out);
int mftx_mft_plan_execute(mftx_mft_plan* plan);
int mftx_mft_plan_destroy(mftx_mft_plan* plan);

/* ---------------------------------------------------------------------------
 * Detector-grid post-processing of an accumulated image (device pointers,
 * real float32 (1) / float64 (3) data, row strides in elements).
 * ------------------------------------------------------------------------- */

/* Integer block sums of nimg stacked rows x cols images (image i at
 * img + i * rows * ld) into nimg contiguous (rows/factor) x (cols/factor)
 * outputs: util.rebin (util.py:221-259).  factor must divide rows and cols
 * (MFTX_EINVAL otherwise, as the reference's reshape fails). */
int mftx_rebin(mftx_ctx* ctx, const void* img, int dtype, int64_t nimg, int64_t rows, int64_t cols, int64_t ld,
               int64_t factor, void* out, int out_dtype);

/* out = |ifft2(fft2(img) * kernel)|, times sum(img) / sum(out) when
 * normalize: the MTF convolution of detector.pixel (detector.py:216-225),
 * convolvable.jitter (convolvable.py:72-82) and convolvable.smear
 * (convolvable.py:152-169); kernel is rows x cols in numpy's fftfreq order.
 * Both transforms are MFTs with alpha = 1/N; the kernel multiplies the
 * spectrum in the forward transform's store. */
int mftx_fft_filter(mftx_ctx* ctx, const void* img, int dtype, int64_t rows, int64_t cols, int64_t ld,
                    const void* kernel, int kernel_dtype, int64_t ldk, int normalize, void* out, int out_dtype,
                    int64_t ldo, int precision);

/* Boundary modes of util.rescale's map_coordinates (util.py:340-342), named
 * as scipy.ndimage names them; cval is 0. */
#define MFTX_MODE_NEAREST 0
#define MFTX_MODE_REFLECT 1
#define MFTX_MODE_GRID_MIRROR 2
#define MFTX_MODE_MIRROR 3
#define MFTX_MODE_GRID_WRAP 4
#define MFTX_MODE_WRAP 5
#define MFTX_MODE_CONSTANT 6
#define MFTX_MODE_GRID_CONSTANT 7

/* util.rescale (util.py:262-349): resample img (float32/float64, or
 * complex64/complex128 whose real and imaginary parts are resampled
 * separately, util.py:339-342) to out_rows x out_cols at the coordinates
 * (i - out/2) / scale + in/2 with spline order 0..5 and boundary mode
 * MFTX_MODE_* (scipy map_coordinates semantics: for orders >= 2 the B-spline
 * prefilter, after 12 samples of padding in modes nearest and grid-constant);
 * the mask (img != 0 when mask is NULL; a mask is real) is resampled with
 * order 1, mode nearest, and zeroed below its dtype's eps; unitary scales by
 * sum(img) / sum(resampled) (complex sums for a complex image); the output
 * is the product with the mask, real for a real image and complex64 /
 * complex128 for a complex one. */
int mftx_rescale(mftx_ctx* ctx, const void* img, int dtype, int64_t rows, int64_t cols, int64_t ld, const void* mask,
                 int mask_dtype, int64_t ldm, double scale, int64_t out_rows, int64_t out_cols, int order, int mode,
                 int unitary, void* out, int out_dtype, int64_t ldo);

/* ---------------------------------------------------------------------------
 * Modal OPD operations (device pointers; masks: nseg arrays mask_stride
 * elements apart, float32 (1) / float64 (3)).
 * ------------------------------------------------------------------------- */

/* Normal equations of Plane.fit_tilt's least-squares piston/tip/tilt fit
 * (plane.py:521-611) per segment, over the columns
 * mask * (1, r * dx_r, -c * dx_c) of Plane.ptt_vector with the centred mesh
 * r, c (helper.py:7-17).  normal (host memory, nseg x 9) receives
 * G00 G01 G02 G11 G12 G22 b0 b1 b2 per segment; the call synchronises. */
int mftx_ptt_fit(mftx_ctx* ctx, const void* opd, int opd_dtype, int64_t ld_opd, const void* masks, int mask_dtype,
                 int64_t ld_mask, int64_t mask_stride, int32_t nseg, int64_t rows, int64_t cols, double dx_r,
                 double dx_c, double* normal);

/* Remove fitted tilt (plane.py:590-609): out = opd - m (r dx_r t1 - c dx_c t2)
 * for one segment (multi = 0), or sum_s (opd - m_s (...)) m_s for several;
 * tilt (host memory) holds (t1, t2) per segment.  out is float64. */
int mftx_remove_tilt(mftx_ctx* ctx, const void* opd, int opd_dtype, int64_t ld_opd, const void* masks, int mask_dtype,
                     int64_t ld_mask, int64_t mask_stride, int32_t nseg, int32_t multi, int64_t rows, int64_t cols,
                     double dx_r, double dx_c, const double* tilt, double* out, int64_t ldo);

/* Zernike basis (zernike.py:11-109,183-234): out[mode][pixel] for nmodes
 * (n, m) pairs in noll_nm (host memory), at float64 polar coordinates
 * rho/theta over npix pixels, times the mask's nonzero pixels; Noll's
 * normalisation when normalize. */
int mftx_zernike_basis(mftx_ctx* ctx, const double* rho, const double* theta, const void* mask, int mask_dtype,
                       int64_t npix, const int32_t* noll_nm, int32_t nmodes, int normalize, double* out);

/* Fill `bytes` bytes at device pointer p with zeros (a kernel, not a memset node). */
int mftx_zero(mftx_ctx* ctx, void* p, int64_t bytes);

/* ---------------------------------------------------------------------------
 * Batched broadband / segmented PSF plan.
 *
 * Defined as the reference user loop (docs/user/diffraction.rst:148-160):
 *   for each wavelength w:  W = Wavefront(w) * pupil [* tilts];
 *                           W = propagate_dft(W, ...);  image = W.insert(image, weight_w)
 * i.e. Plane.__mul__ (plane.py:477-516) -> propagate_dft (propagate.py:147-242)
 * -> Wavefront.insert (wavefront.py:145-165) with coherent merging of
 * overlapping Fields (field.py:413-461).  The host computes the per-Field
 * geometry of propagate_dft and passes one mftx_problem per (tile, wavelength)
 * that reaches the output; the plan runs every problem in a few grouped
 * launches and accumulates weight*|E|^2 onto the detector image.  A tile may
 * also appear at one wavelength in several problems (different placements
 * or shifts); each distinct output grid then gets an operand of its own.
 * ------------------------------------------------------------------------- */

typedef struct {
  const void* amp;   /* m x n tile amplitude (already multiplied by its segment mask) */
  const void* opd;   /* m x n tile OPD in metres */
  int32_t amp_dtype, opd_dtype;  /* 1 float32 or 3 float64 */
  int64_t ld_amp, ld_opd;
  int64_t m, n;
} mftx_tile;

typedef struct {
  int32_t tile;        /* index into tiles */
  int32_t wave;        /* index into wavelengths */
  int32_t group;       /* coherent group: problems with equal group are summed as fields before |.|^2 */
  int32_t M, N;        /* dft2 output shape (propagate.py:215) */
  double alpha_r, alpha_c;     /* propagate.py:232-234 */
  double shift_r, shift_c;     /* prop_shift + subpixel shift (propagate.py:235-236) */
  double offset_r, offset_c;   /* the Field offset (propagate.py:236) */
  int64_t out_r0, out_c0;      /* detector index of F[0][0] (may be outside the image; |.| <= 2^30) */
  double weight;               /* intensity weight of the wavelength */
} mftx_problem;

typedef struct {
  double flops;          /* algorithmic FLOPs of one execute (8 per complex MAC) */
  double gemm_flops[2];  /* per GEMM stage */
  double mfma_flops[2];  /* executed MFMA FLOPs per GEMM stage (split passes and tile padding included) */
  double bytes_hbm_min;  /* minimum HBM bytes of one execute (phasor + intermediates + image) */
  double bytes_phasor;   /* algorithmic bytes of the phasor launches (amp/opd read once, operands written) */
  double bytes_reduce;   /* algorithmic bytes of the detector reductions (partials read, image written) */
  int32_t n_launch;      /* kernel launches per execute */
  int32_t coherent_fast; /* 1 if every coherent group runs as one K-concatenated GEMM */
  int32_t n_stage1, n_stage2, n_work1, n_work2;
} mftx_plan_info;

/* Build a plan; validates every problem (MFTX_EINVAL on a bad index or shape;
 * image_rows, image_cols and every M, N at most 2^20)
 * and allocates plan-owned workspace.  Tile pointers are kept and re-read on
 * every execute, so amp/opd contents may change between executes.
 * Schedule knobs (no reference counterpart; lentil_amd.PSFPlan can pick
 * them by timing candidates): lambda_group = wavelengths whose |E|^2 one
 * GEMM workgroup sums in registers (0 = 8); tile_n = C columns per GEMM
 * workgroup, 64 (16-wide K steps) or 128 (32-wide K steps; every tile's K is
 * padded to a multiple of 32) (0 = 128). */
int mftx_plan_psf(mftx_ctx* ctx, const mftx_tile* tiles, int32_t ntiles,
                  const double* wavelengths, int32_t nwave,
                  const mftx_problem* problems, int32_t nprob,
                  int64_t image_rows, int64_t image_cols, int32_t precision,
                  int32_t lambda_group, int32_t tile_n, mftx_plan** out);
/* image: device float32 (dtype 1) or float64 (dtype 3), image_rows x image_cols,
 * row stride ldi.  accumulate != 0 adds to image (Wavefront.insert), else overwrites. */
int mftx_plan_execute(mftx_plan* plan, void* image, int32_t image_dtype, int64_t ldi,
                      int32_t accumulate);
/* Complex fields from the plan's last execute (no reference counterpart;
 * for gradients, which need the fields the image came from).  set_fields
 * registers, once, per problem (the plan's problem order): the M x N complex64
 * output F[u][v] (row stride ldd), an optional real weight map multiplying it
 * (same layout, row stride ldw; wmap may be NULL, or hold NULL entries) and a
 * scale.  A coherent group's field goes to its first problem's pointers.
 * execute_fields then runs one complex-output stage-2 GEMM over the stage-1
 * operands the last mftx_plan_execute left in the plan's workspace: call it
 * after that execute and before the next. */
int mftx_plan_set_fields(mftx_plan* plan, void* const* dst, int64_t ldd, const void* const* wmap, int32_t wmap_dtype,
                         int64_t ldw, const double* scale);
int mftx_plan_execute_fields(mftx_plan* plan);
/* Capture the plan's launches in a hipGraph for this image pointer and replay it
 * on later executes with the same pointer (use_graph != 0), or run eagerly. */
int mftx_plan_set_graph(mftx_plan* plan, int32_t use_graph);
int mftx_plan_get_info(const mftx_plan* plan, mftx_plan_info* info);
/* Time the plan's launch groups (phasors, stage-1 GEMM, stage-2 GEMM,
 * detector reduction) in the mode executes run in, `repeats` times each (at
 * most 4096).  When the plan replays a graph: one graph per prefix of the
 * groups (the first k, k = 1..4) is replayed `repeats` times back to back
 * between two stream events, after two full replays, the prefixes in the
 * order 1 2 3 4 4 3 2 1 so a slow clock drift cancels; a group's time is the
 * difference of the mean replay times of the prefixes that end with and
 * before it, so each launch runs right after its producer with no event in
 * between (the phasor time includes one graph launch's overhead).  Otherwise
 * the groups are queued eagerly with events between them.  ms[0..3] receive
 * the mean milliseconds per group, ms[4] is 1 for the graph mode and 0 for
 * the eager one (also when the runtime refuses a capture).  The image is
 * overwritten.  Measurement hook; no reference counterpart. */
int mftx_plan_profile(mftx_plan* plan, void* image, int32_t image_dtype, int64_t ldi,
                      int32_t repeats, double* ms);
int mftx_plan_destroy(mftx_plan* plan);

#ifdef __cplusplus
}
#endif
#endif /* MFTX_H */
