// MI355X (gfx950) kernels of the detector-grid post-processing: integer
// rebinning (util.py:221-259), the magnitude / weight-preserving passes of
// the FFT filters (detector.py:167-225, convolvable.py:4-169) and the cubic
// B-spline resampling of rescale (util.py:262-349).  All HBM-bound; the
// arithmetic is FP64 so results follow the reference's float64 numpy/scipy
// to rounding.
#include "../../include/mftx.h"
#include "mftx_device.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mftx {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double load_real(const void* p, int dtype, long long i) {
  return dtype == 3 ? ((const double*)p)[i] : (double)((const float*)p)[i];
}

__device__ __forceinline__ void store_real(void* p, int dtype, long long i, double v) {
  if (dtype == 3) ((double*)p)[i] = v;
  else ((float*)p)[i] = (float)v;
}

// Sum of v over the block into part[blockIdx.x] (fixed order: deterministic).
__device__ void block_partial(double v, double* part) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __shared__ double w[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) w[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) s += w[i];
    part[blockIdx.x] = s;
  }
}

unsigned blocks_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

// --------------------------------------------------------------------------
// rebin: out[i][r][c] = sum of the factor x factor block (util.py:251-257)
// --------------------------------------------------------------------------
__global__ void rebin_kernel(const void* __restrict__ src, int dtype, long long nimg, long long rows, long long cols,
                             long long ld, int factor, void* __restrict__ dst, int odt) {
  const long long orows = rows / factor, ocols = cols / factor;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nimg * orows * ocols) return;
  const long long img = t / (orows * ocols), rem = t % (orows * ocols);
  const long long r = rem / ocols, c = rem % ocols;
  const long long base = img * rows * ld;
  double s = 0.0;
  for (int i = 0; i < factor; ++i)
    for (int j = 0; j < factor; ++j) s += load_real(src, dtype, base + (r * factor + i) * ld + c * factor + j);
  store_real(dst, odt, t, s);
}

hipError_t launch_rebin(const void* src, int dtype, long long nimg, long long rows, long long cols, long long ld,
                        int factor, void* dst, int odt, hipStream_t s) {
  const long long n = nimg * (rows / factor) * (cols / factor);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(rebin_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, src, dtype, nimg, rows, cols, ld, factor,
                     dst, odt);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// reductions and the weight-preserving rescale of the filters
// --------------------------------------------------------------------------

// part[b] = sum of a real rows x cols array over block b's elements
__global__ void sum_kernel(const void* __restrict__ x, int dtype, long long rows, long long cols, long long ld,
                           double* __restrict__ part) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (t < rows * cols) v = load_real(x, dtype, (t / cols) * ld + t % cols);
  block_partial(v, part);
}

// out = |F| (F complex64), part[b] = block sums of out
__global__ void magnitude_kernel(const float2* __restrict__ F, long long ldF, long long rows, long long cols,
                                 void* __restrict__ out, int odt, long long ldo, double* __restrict__ part) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (t < rows * cols) {
    const long long r = t / cols, c = t % cols;
    const float2 f = F[r * ldF + c];
    v = sqrt((double)f.x * f.x + (double)f.y * f.y);
    store_real(out, odt, r * ldo + c, v);
    if (odt != 3) v = (double)(float)v;  // sum what was stored
  }
  block_partial(v, part);
}

// *out = sum of part[0 .. n) (one block, fixed order)
__global__ void finalize_sum_kernel(const double* __restrict__ part, long long n, double* __restrict__ out) {
  double v = 0.0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __shared__ double w[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) w[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) s += w[i];
    *out = s;
  }
}

// x *= num / den  (device scalars)
__global__ void scale_ratio_kernel(void* __restrict__ x, int dtype, long long rows, long long cols, long long ld,
                                   const double* __restrict__ num, const double* __restrict__ den) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * cols) return;
  const long long i = (t / cols) * ld + t % cols;
  store_real(x, dtype, i, load_real(x, dtype, i) * (*num / *den));
}

hipError_t launch_sum(const void* x, int dtype, long long rows, long long cols, long long ld, double* part,
                      double* out, hipStream_t s) {
  const unsigned nb = blocks_for(rows * cols);
  hipLaunchKernelGGL(sum_kernel, dim3(nb), dim3(kBlock), 0, s, x, dtype, rows, cols, ld, part);
  hipLaunchKernelGGL(finalize_sum_kernel, dim3(1), dim3(kBlock), 0, s, (const double*)part, (long long)nb, out);
  return hipGetLastError();
}

hipError_t launch_magnitude(const void* F, long long ldF, long long rows, long long cols, void* out, int odt,
                            long long ldo, double* part, double* sum, hipStream_t s) {
  const unsigned nb = blocks_for(rows * cols);
  hipLaunchKernelGGL(magnitude_kernel, dim3(nb), dim3(kBlock), 0, s, (const float2*)F, ldF, rows, cols, out, odt, ldo,
                     part);
  hipLaunchKernelGGL(finalize_sum_kernel, dim3(1), dim3(kBlock), 0, s, (const double*)part, (long long)nb, sum);
  return hipGetLastError();
}

hipError_t launch_scale_ratio(void* x, int dtype, long long rows, long long cols, long long ld, const double* num,
                              const double* den, hipStream_t s) {
  if (rows * cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(scale_ratio_kernel, dim3(blocks_for(rows * cols)), dim3(kBlock), 0, s, x, dtype, rows, cols, ld,
                     num, den);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// rescale: map_coordinates(order 0..5, any boundary mode) on the device
// --------------------------------------------------------------------------
// The algorithm of scipy.ndimage.map_coordinates as util.rescale calls it:
// orders >= 2 prefilter the input with the inverse B-spline filter (after
// 12 samples of padding in modes 'nearest' and 'grid-constant'), with the
// mode's boundary rule; the spline of order p is then sampled with p + 1
// taps per axis, taps past the line extended by the mode's rule; 'constant'
// gives 0 outside [0, n - 1], and unpadded modes move an outside coordinate
// inside by their symmetry first.  cval is 0 (rescale never passes one).

namespace {

// index extension rules
enum { EX_CLAMP = 0, EX_REFLECT = 1, EX_MIRROR = 2, EX_PERIODIC = 3, EX_ZERO = 4 };

__host__ __device__ inline int filter_rule(int mode) {
  return mode == MFTX_MODE_GRID_WRAP ? EX_PERIODIC
         : (mode == MFTX_MODE_NEAREST || mode == MFTX_MODE_REFLECT || mode == MFTX_MODE_GRID_MIRROR) ? EX_REFLECT
                                                                                                    : EX_MIRROR;
}
__device__ inline int tap_rule(int mode) {
  switch (mode) {
    case MFTX_MODE_NEAREST: return EX_CLAMP;
    case MFTX_MODE_REFLECT:
    case MFTX_MODE_GRID_MIRROR: return EX_REFLECT;
    case MFTX_MODE_GRID_WRAP: return EX_PERIODIC;
    case MFTX_MODE_GRID_CONSTANT: return EX_ZERO;
    default: return EX_MIRROR;  // mirror, wrap (after its coordinate map), constant
  }
}

// Index of sample i of an n-sample line under a rule; -1 reads 0 (EX_ZERO).
__device__ __forceinline__ long long extend_index(long long i, long long n, int rule) {
  if (i >= 0 && i < n) return i;
  switch (rule) {
    case EX_CLAMP: return i < 0 ? 0 : n - 1;
    case EX_REFLECT: {  // half-sample symmetric, period 2n
      const long long p = 2 * n;
      long long k = i % p;
      if (k < 0) k += p;
      return k >= n ? p - 1 - k : k;
    }
    case EX_MIRROR: {  // whole-sample symmetric, period 2n - 2
      if (n == 1) return 0;
      const long long p = 2 * n - 2;
      long long k = i % p;
      if (k < 0) k += p;
      return k >= n ? p - k : k;
    }
    case EX_PERIODIC: {
      long long k = i % n;
      return k < 0 ? k + n : k;
    }
    default: return -1;
  }
}

// A coordinate outside [0, n - 1] moved inside by the mode's symmetry (only
// for unpadded lines; the other modes keep it).  For orders >= 1 this is
// the same as extending the taps; for order 0 it decides where a
// half-integer coordinate rounds to.
__device__ __forceinline__ double map_coordinate(double x, long long n, int mode) {
  const double last = (double)(n - 1);
  if (x >= 0.0 && x <= last) return x;
  const bool sym = mode == MFTX_MODE_MIRROR || mode == MFTX_MODE_REFLECT || mode == MFTX_MODE_GRID_MIRROR ||
                   mode == MFTX_MODE_WRAP || mode == MFTX_MODE_GRID_WRAP || mode == MFTX_MODE_NEAREST;
  if (!sym) return x;
  if (n <= 1) return 0.0;
  switch (mode) {
    case MFTX_MODE_MIRROR: {
      const double p = 2.0 * n - 2.0;
      if (x < 0) {
        x += p * trunc(-x / p);
        return x <= 1.0 - n ? x + p : -x;
      }
      x -= p * trunc(x / p);
      return x >= n ? p - x : x;
    }
    case MFTX_MODE_REFLECT:
    case MFTX_MODE_GRID_MIRROR: {
      const double p = 2.0 * n;
      if (x < 0) {
        if (x < -p) x += p * trunc(-x / p);
        return x < -n ? x + p : (x > -1e-15 ? 1e-15 : -x) - 1.0;
      }
      x -= p * trunc(x / p);
      return x >= n ? p - x - 1.0 : x;
    }
    case MFTX_MODE_WRAP: {
      const double p = (double)(n - 1);
      return x < 0 ? x + p * (trunc(-x / p) + 1.0) : x - p * trunc(x / p);
    }
    case MFTX_MODE_GRID_WRAP:
      return x < 0 ? x + n * (trunc(-x / n) + 1.0) : x - n * trunc(x / n);
    default:  // nearest
      return x < 0 ? 0.0 : last;
  }
}

// The centred B-spline of degree `order` at t.
__device__ __forceinline__ double bspline(int order, double t) {
  t = fabs(t);
  switch (order) {
    case 0: return t <= 0.5 ? 1.0 : 0.0;
    case 1: return t < 1.0 ? 1.0 - t : 0.0;
    case 2: return t < 0.5 ? 0.75 - t * t : (t < 1.5 ? 0.5 * (1.5 - t) * (1.5 - t) : 0.0);
    case 3: return t < 1.0 ? 2.0 / 3.0 - t * t + 0.5 * t * t * t : (t < 2.0 ? (2.0 - t) * (2.0 - t) * (2.0 - t) / 6.0 : 0.0);
    case 4: {
      const double t2 = t * t;
      if (t < 0.5) return 115.0 / 192.0 - 5.0 * t2 / 8.0 + t2 * t2 / 4.0;
      if (t < 1.5) return 55.0 / 96.0 + 5.0 * t / 24.0 - 5.0 * t2 / 4.0 + 5.0 * t2 * t / 6.0 - t2 * t2 / 6.0;
      if (t < 2.5) { const double u = 2.5 - t; return u * u * u * u / 24.0; }
      return 0.0;
    }
    default: {
      const double t2 = t * t;
      if (t < 1.0) return 11.0 / 20.0 - t2 / 2.0 + t2 * t2 / 4.0 - t2 * t2 * t / 12.0;
      if (t < 2.0)
        return 17.0 / 40.0 + 5.0 * t / 8.0 - 7.0 * t2 / 4.0 + 5.0 * t2 * t / 4.0 - 3.0 * t2 * t2 / 8.0 +
               t2 * t2 * t / 24.0;
      if (t < 3.0) { const double u = 3.0 - t; return u * u * u * u * u / 120.0; }
      return 0.0;
    }
  }
}

__device__ __forceinline__ long long first_tap(double c, int order) {
  return order & 1 ? (long long)floor(c) - (order - 1) / 2 : (long long)floor(c + 0.5) - order / 2;
}

// component `comp` (0 real, 1 imaginary) of element i of a real or complex array
__device__ __forceinline__ double load_component(const void* p, int dtype, long long i, int comp) {
  switch (dtype) {
    case 0: return comp ? ((const float2*)p)[i].y : ((const float2*)p)[i].x;
    case 2: return comp ? ((const double2*)p)[i].y : ((const double2*)p)[i].x;
    default: return comp ? 0.0 : load_real(p, dtype, i);
  }
}
__device__ __forceinline__ bool nonzero(const void* p, int dtype, long long i) {
  return load_component(p, dtype, i, 0) != 0.0 || ((dtype == 0 || dtype == 2) && load_component(p, dtype, i, 1) != 0.0);
}

}  // namespace

// P (P0 x P1, double) = component comp of img (rows x cols) with npad samples
// of padding on every side: edge values (numpy.pad mode='edge') or zeros.
__global__ void rescale_load_kernel(const void* __restrict__ img, int dtype, int comp, long long rows, long long cols,
                                    long long ld, int npad, int pad_zero, double* __restrict__ P) {
  const long long P0 = rows + 2 * npad, P1 = cols + 2 * npad;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P0 * P1) return;
  const long long i = t / P1 - npad, j = t % P1 - npad;
  if (pad_zero && (i < 0 || i >= rows || j < 0 || j >= cols)) {
    P[t] = 0.0;
    return;
  }
  const long long r = min(max(i, 0LL), rows - 1), c = min(max(j, 0LL), cols - 1);
  P[t] = load_component(img, dtype, r * ld + c, comp);
}

// One axis of the B-spline prefilter as its two-sided impulse response,
// dst[i] = sum_{|k| <= taps} h[|k|] src[extend(i + k)], along axis 0
// (columns) or 1 (rows) of a P0 x P1 array.  h is the product over the
// order's poles z of (1 - z) / (1 + z) z^|k|, truncated where it falls
// below double rounding; the boundary rule is the mode's.
__global__ void spline_fir_kernel(const double* __restrict__ src, double* __restrict__ dst, long long P0, long long P1,
                                  int axis, int rule, PrefilterTaps taps) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P0 * P1) return;
  const long long i = t / P1, j = t % P1;
  const long long n = axis == 0 ? P0 : P1, pos = axis == 0 ? i : j;
  const long long stride = axis == 0 ? P1 : 1;
  const double* line = src + (axis == 0 ? j : i * P1);
  double s = taps.h[0] * line[pos * stride];
  for (int k = 1; k <= taps.n; ++k)
    s += taps.h[k] * (line[extend_index(pos - k, n, rule) * stride] + line[extend_index(pos + k, n, rule) * stride]);
  dst[t] = s;
}

// Resample at the coordinates of util.py:333-336,
//   y = (r - orows / 2) / scale + rows / 2,  x = (c - ocols / 2) / scale + cols / 2,
// from the coefficients C (P0 x P1, npad samples of padding) with spline
// order 0..5 and the mode's tap and coordinate rules.  With mval, the mask
// (mask_mode 1: img != 0; 2: a real mask array) is sampled with order 1,
// mode 'nearest', on the unpadded grid and zeroed below eps
// (util.py:338-339).  raw = interpolated value, part = block sums of raw.
__global__ void resample_kernel(const double* __restrict__ C, long long P0, long long P1, int npad, int order,
                                int mode, long long rows, long long cols, double scale, long long orows,
                                long long ocols, const void* __restrict__ msrc, int mask_mode, int mdt, long long ldm,
                                double eps, double* __restrict__ raw, double* __restrict__ mval,
                                double* __restrict__ part) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (t < orows * ocols) {
    const long long r = t / ocols, c = t % ocols;
    const double y = ((double)r - (double)orows / 2.0) / scale + (double)rows / 2.0;
    const double x = ((double)c - (double)ocols / 2.0) / scale + (double)cols / 2.0;
    const bool outside = mode == MFTX_MODE_CONSTANT && (y < 0.0 || y > (double)(rows - 1) || x < 0.0 ||
                                                       x > (double)(cols - 1));
    if (!outside) {
      const double cy = (npad ? y : map_coordinate(y, rows, mode)) + npad;
      const double cx = (npad ? x : map_coordinate(x, cols, mode)) + npad;
      const int rule = tap_rule(mode);
      const long long sy = first_tap(cy, order), sx = first_tap(cx, order);
      for (int a = 0; a <= order; ++a) {
        const long long rr = extend_index(sy + a, P0, rule);
        if (rr < 0) continue;
        double row = 0.0;
        for (int b = 0; b <= order; ++b) {
          const long long cc = extend_index(sx + b, P1, rule);
          if (cc >= 0) row += bspline(order, cx - (double)(sx + b)) * C[rr * P1 + cc];
        }
        v += bspline(order, cy - (double)(sy + a)) * row;
      }
    }
    raw[t] = v;
    if (mval) {
      const double my = fmin(fmax(y, 0.0), (double)(rows - 1)), mx = fmin(fmax(x, 0.0), (double)(cols - 1));
      const long long fy = (long long)floor(my), fx = (long long)floor(mx);
      const double ty = my - fy, tx = mx - fx;
      const long long r1 = min(fy + 1, rows - 1), c1 = min(fx + 1, cols - 1);
      auto m = [&](long long rr, long long cc) -> double {
        return mask_mode == 1 ? (nonzero(msrc, mdt, rr * ldm + cc) ? 1.0 : 0.0) : load_real(msrc, mdt, rr * ldm + cc);
      };
      double mv =
          (1 - ty) * ((1 - tx) * m(fy, fx) + tx * m(fy, c1)) + ty * ((1 - tx) * m(r1, fx) + tx * m(r1, c1));
      if (mv < eps) mv = 0.0;
      mval[t] = mv;
    }
  }
  block_partial(v, part);
}

// out = raw * (unitary ? sum(img) / sum(raw) : 1) * mval   (util.py:340-347);
// with raw_im, complex: the ratio of the complex sums, sums = (img_re,
// img_im, raw_re, raw_im), out complex64 (odt 0) or complex128 (odt 2).
__global__ void rescale_finish_kernel(const double* __restrict__ raw, const double* __restrict__ raw_im,
                                      const double* __restrict__ mval, long long orows, long long ocols, int unitary,
                                      const double* __restrict__ sums, void* __restrict__ out, int odt, long long ldo) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= orows * ocols) return;
  const long long o = (t / ocols) * ldo + t % ocols;
  if (!raw_im) {
    double v = raw[t];
    if (unitary) v *= sums[0] / sums[2];
    store_real(out, odt, o, v * mval[t]);
    return;
  }
  double re = raw[t], im = raw_im[t];
  if (unitary) {
    // (a + ib) / (c + id) with a, b = sum(img), c, d = sum(raw)
    const double a = sums[0], b = sums[1], c = sums[2], d = sums[3];
    const double den = c * c + d * d;
    const double qr = (a * c + b * d) / den, qi = (b * c - a * d) / den;
    const double nr = re * qr - im * qi;
    im = re * qi + im * qr;
    re = nr;
  }
  re *= mval[t];
  im *= mval[t];
  if (odt == 2) ((double2*)out)[o] = make_double2(re, im);
  else ((float2*)out)[o] = make_float2((float)re, (float)im);
}

// The order's prefilter response h[|k|], k <= kPrefilterTaps, computed once
// per order on the host: the product over the B-spline's poles z (Unser's
// values) of (1 - z) / (1 + z) z^|k|, formed on a support four times wider
// than kept so the truncation drops only terms below double rounding
// (|z|^48 < 3e-18 for the largest pole, order 5).
static PrefilterTaps make_prefilter_taps(int order) {
  PrefilterTaps h = {};
  h.h[0] = 1.0;
  if (order < 2) return h;
  std::vector<double> z;
  if (order == 2) z = {std::sqrt(8.0) - 3.0};
  if (order == 3) z = {std::sqrt(3.0) - 2.0};
  if (order == 4)
    z = {std::sqrt(664.0 - std::sqrt(438976.0)) + std::sqrt(304.0) - 19.0,
         std::sqrt(664.0 + std::sqrt(438976.0)) - std::sqrt(304.0) - 19.0};
  if (order == 5)
    z = {std::sqrt(67.5 - std::sqrt(4436.25)) + std::sqrt(26.25) - 6.5,
         std::sqrt(67.5 + std::sqrt(4436.25)) - std::sqrt(26.25) - 6.5};
  constexpr int W = 4 * kPrefilterTaps;
  std::vector<double> cur(2 * W + 1, 0.0), nxt(2 * W + 1), g(2 * W + 1);
  cur[W] = 1.0;
  for (double zp : z) {
    double zk = (1.0 - zp) / (1.0 + zp);
    for (int k = 0; k <= W; ++k, zk *= zp) g[W + k] = g[W - k] = zk;
    for (int i = 0; i <= 2 * W; ++i) {
      double acc = 0.0;
      for (int j = std::max(0, i - W); j <= std::min(2 * W, i + W); ++j) acc += g[W + i - j] * cur[j];
      nxt[i] = acc;
    }
    cur.swap(nxt);
  }
  h.n = kPrefilterTaps;
  for (int k = 0; k <= kPrefilterTaps; ++k) h.h[k] = cur[W + k];
  return h;
}

PrefilterTaps prefilter_taps(int order) {
  static const PrefilterTaps table[6] = {make_prefilter_taps(0), make_prefilter_taps(1), make_prefilter_taps(2),
                                         make_prefilter_taps(3), make_prefilter_taps(4), make_prefilter_taps(5)};
  return table[order < 0 || order > 5 ? 0 : order];
}

hipError_t launch_rescale(const void* img, int dtype, long long rows, long long cols, long long ld, const void* msrc,
                          int mask_mode, int mdt, long long ldm, double eps, double scale, long long orows,
                          long long ocols, int order, int mode, int unitary, void* out, int odt, long long ldo,
                          double* P, double* T, double* raw, double* raw_im, double* mval, double* part, double* sums,
                          hipStream_t s) {
  const bool cplx = dtype == 0 || dtype == 2;
  const int npad = order > 1 && (mode == MFTX_MODE_NEAREST || mode == MFTX_MODE_GRID_CONSTANT) ? 12 : 0;
  const long long P0 = rows + 2 * npad, P1 = cols + 2 * npad;
  const unsigned nbp = blocks_for(P0 * P1);
  const unsigned nbo = blocks_for(orows * ocols);
  const PrefilterTaps taps = prefilter_taps(order);
  for (int comp = 0; comp < (cplx ? 2 : 1); ++comp) {
    hipLaunchKernelGGL(rescale_load_kernel, dim3(nbp), dim3(kBlock), 0, s, img, dtype, comp, rows, cols, ld, npad,
                       mode == MFTX_MODE_GRID_CONSTANT ? 1 : 0, P);
    if (unitary) {  // sum of the image component: the interior of P
      hipError_t e = launch_sum(P + npad * P1 + npad, 3, rows, cols, P1, part, sums + comp, s);
      if (e != hipSuccess) return e;
    }
    if (order > 1) {
      hipLaunchKernelGGL(spline_fir_kernel, dim3(nbp), dim3(kBlock), 0, s, (const double*)P, T, P0, P1, 0,
                         filter_rule(mode), taps);
      hipLaunchKernelGGL(spline_fir_kernel, dim3(nbp), dim3(kBlock), 0, s, (const double*)T, P, P0, P1, 1,
                         filter_rule(mode), taps);
    }
    hipLaunchKernelGGL(resample_kernel, dim3(nbo), dim3(kBlock), 0, s, (const double*)P, P0, P1, npad, order, mode,
                       rows, cols, scale, orows, ocols, msrc, mask_mode, mdt, ldm, eps, comp ? raw_im : raw,
                       comp ? (double*)nullptr : mval, part);
    hipLaunchKernelGGL(finalize_sum_kernel, dim3(1), dim3(kBlock), 0, s, (const double*)part, (long long)nbo,
                       sums + 2 + comp);
  }
  hipLaunchKernelGGL(rescale_finish_kernel, dim3(nbo), dim3(kBlock), 0, s, (const double*)raw,
                     cplx ? (const double*)raw_im : nullptr, (const double*)mval, orows, ocols, unitary,
                     (const double*)sums, out, odt, ldo);
  return hipGetLastError();
}

}  // namespace mftx
