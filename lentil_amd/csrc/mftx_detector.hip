// MI355X (gfx950) kernels of the detector-grid post-processing: integer
// rebinning (util.py:221-259), the magnitude / weight-preserving passes of
// the FFT filters (detector.py:167-225, convolvable.py:4-169) and the cubic
// B-spline resampling of rescale (util.py:262-349).  All HBM-bound; the
// arithmetic is FP64 so results follow the reference's float64 numpy/scipy
// to rounding.
#include "mftx_device.h"

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
// rescale: map_coordinates(order, mode='nearest') on the device
// --------------------------------------------------------------------------

// P (P0 x P1, double) = img (rows x cols) with npad edge-replicated samples
// on every side (numpy.pad mode='edge').
__global__ void edge_pad_kernel(const void* __restrict__ img, int dtype, long long rows, long long cols, long long ld,
                                int npad, double* __restrict__ P) {
  const long long P0 = rows + 2 * npad, P1 = cols + 2 * npad;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P0 * P1) return;
  const long long i = t / P1, j = t % P1;
  const long long r = min(max(i - npad, 0LL), rows - 1), c = min(max(j - npad, 0LL), cols - 1);
  P[t] = load_real(img, dtype, r * ld + c);
}

// Half-sample symmetric (reflect) index of i on an n-sample line.
__device__ __forceinline__ long long fold_index(long long i, long long n) {
  const long long p = 2 * n;
  i %= p;
  if (i < 0) i += p;
  return i >= n ? p - 1 - i : i;
}

// One axis of the cubic B-spline prefilter as its two-sided impulse
// response: dst[i] = sqrt(3) * sum_{|k| <= 32} z^|k| src[fold(i + k)],
// z = sqrt(3) - 2, along axis 0 (columns) or 1 (rows) of a P0 x P1 array.
// z^32 < 6e-19: the truncation is below double rounding.
__global__ void spline_fir_kernel(const double* __restrict__ src, double* __restrict__ dst, long long P0, long long P1,
                                  int axis) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P0 * P1) return;
  const long long i = t / P1, j = t % P1;
  const long long n = axis == 0 ? P0 : P1, pos = axis == 0 ? i : j;
  const long long stride = axis == 0 ? P1 : 1;
  const double* line = src + (axis == 0 ? j : i * P1);
  const double z = 1.7320508075688772 - 2.0;  // sqrt(3) - 2
  double s = line[pos * stride], zk = 1.0;
  for (int k = 1; k <= 32; ++k) {
    zk *= z;
    s += zk * (line[fold_index(pos - k, n) * stride] + line[fold_index(pos + k, n) * stride]);
  }
  dst[t] = 1.7320508075688772 * s;
}

// Resample at the coordinates of util.py:333-336,
//   y = (r - orows / 2) / scale + rows / 2,  x = (c - ocols / 2) / scale + cols / 2,
// from coefficients C (P0 x P1, npad samples of padding) with spline order
// 0, 1 or 3, mode 'nearest' (coordinate and taps clamped to the padded grid).
// The mask (mask_mode 1: img != 0; 2: a real mask array) is sampled with
// order 1 on the unpadded grid and zeroed below eps (util.py:338-339).
// raw = interpolated value, mval = mask value, part = block sums of raw.
__global__ void resample_kernel(const double* __restrict__ C, long long P0, long long P1, int npad, int order,
                                long long rows, long long cols, double scale, long long orows, long long ocols,
                                const void* __restrict__ msrc, int mask_mode, int mdt, long long ldm, double eps,
                                double* __restrict__ raw, double* __restrict__ mval, double* __restrict__ part) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (t < orows * ocols) {
    const long long r = t / ocols, c = t % ocols;
    const double y = ((double)r - (double)orows / 2.0) / scale + (double)rows / 2.0;
    const double x = ((double)c - (double)ocols / 2.0) / scale + (double)cols / 2.0;
    const double cy = fmin(fmax(y + npad, 0.0), (double)(P0 - 1));
    const double cx = fmin(fmax(x + npad, 0.0), (double)(P1 - 1));
    if (order == 3) {
      const double fy = floor(cy), fx = floor(cx);
      const double ty = cy - fy, tx = cx - fx;
      const double wy[4] = {(1 - ty) * (1 - ty) * (1 - ty) / 6, (4 - 6 * ty * ty + 3 * ty * ty * ty) / 6,
                            (1 + 3 * ty + 3 * ty * ty - 3 * ty * ty * ty) / 6, ty * ty * ty / 6};
      const double wx[4] = {(1 - tx) * (1 - tx) * (1 - tx) / 6, (4 - 6 * tx * tx + 3 * tx * tx * tx) / 6,
                            (1 + 3 * tx + 3 * tx * tx - 3 * tx * tx * tx) / 6, tx * tx * tx / 6};
      const long long sy = (long long)fy - 1, sx = (long long)fx - 1;
      for (int a = 0; a < 4; ++a) {
        const long long rr = min(max(sy + a, 0LL), P0 - 1);
        double row = 0.0;
        for (int b = 0; b < 4; ++b) row += wx[b] * C[rr * P1 + min(max(sx + b, 0LL), P1 - 1)];
        v += wy[a] * row;
      }
    } else if (order == 1) {
      const long long fy = (long long)floor(cy), fx = (long long)floor(cx);
      const double ty = cy - fy, tx = cx - fx;
      const long long r1 = min(fy + 1, P0 - 1), c1 = min(fx + 1, P1 - 1);
      v = (1 - ty) * ((1 - tx) * C[fy * P1 + fx] + tx * C[fy * P1 + c1]) +
          ty * ((1 - tx) * C[r1 * P1 + fx] + tx * C[r1 * P1 + c1]);
    } else {
      v = C[(long long)floor(cy + 0.5) * P1 + (long long)floor(cx + 0.5)];
    }
    raw[t] = v;
    // mask: order 1 on the unpadded grid
    const double my = fmin(fmax(y, 0.0), (double)(rows - 1)), mx = fmin(fmax(x, 0.0), (double)(cols - 1));
    const long long fy = (long long)floor(my), fx = (long long)floor(mx);
    const double ty = my - fy, tx = mx - fx;
    const long long r1 = min(fy + 1, rows - 1), c1 = min(fx + 1, cols - 1);
    auto m = [&](long long rr, long long cc) -> double {
      const double s = load_real(msrc, mdt, rr * ldm + cc);
      return mask_mode == 1 ? (s != 0.0 ? 1.0 : 0.0) : s;
    };
    double mv = (1 - ty) * ((1 - tx) * m(fy, fx) + tx * m(fy, c1)) + ty * ((1 - tx) * m(r1, fx) + tx * m(r1, c1));
    if (mv < eps) mv = 0.0;
    mval[t] = mv;
  }
  block_partial(v, part);
}

// out = raw * (unitary ? num / den : 1) * mval   (util.py:344-347)
__global__ void rescale_finish_kernel(const double* __restrict__ raw, const double* __restrict__ mval, long long orows,
                                      long long ocols, int unitary, const double* __restrict__ num,
                                      const double* __restrict__ den, void* __restrict__ out, int odt, long long ldo) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= orows * ocols) return;
  double v = raw[t];
  if (unitary) v *= *num / *den;
  store_real(out, odt, (t / ocols) * ldo + t % ocols, v * mval[t]);
}

hipError_t launch_rescale(const void* img, int dtype, long long rows, long long cols, long long ld, const void* msrc,
                          int mask_mode, int mdt, long long ldm, double eps, double scale, long long orows,
                          long long ocols, int order, int unitary, void* out, int odt, long long ldo, double* P,
                          double* T, double* raw, double* mval, double* part, double* sums, hipStream_t s) {
  const int npad = order > 1 ? 12 : 0;
  const long long P0 = rows + 2 * npad, P1 = cols + 2 * npad;
  const unsigned nbp = blocks_for(P0 * P1);
  hipLaunchKernelGGL(edge_pad_kernel, dim3(nbp), dim3(kBlock), 0, s, img, dtype, rows, cols, ld, npad, P);
  const double* C = P;
  if (order > 1) {
    hipLaunchKernelGGL(spline_fir_kernel, dim3(nbp), dim3(kBlock), 0, s, (const double*)P, T, P0, P1, 0);
    hipLaunchKernelGGL(spline_fir_kernel, dim3(nbp), dim3(kBlock), 0, s, (const double*)T, P, P0, P1, 1);
  }
  if (unitary) {
    hipError_t e = launch_sum(img, dtype, rows, cols, ld, part, sums, s);
    if (e != hipSuccess) return e;
  }
  const unsigned nbo = blocks_for(orows * ocols);
  hipLaunchKernelGGL(resample_kernel, dim3(nbo), dim3(kBlock), 0, s, C, P0, P1, npad, order, rows, cols, scale, orows,
                     ocols, msrc, mask_mode, mdt, ldm, eps, raw, mval, part);
  hipLaunchKernelGGL(finalize_sum_kernel, dim3(1), dim3(kBlock), 0, s, (const double*)part, (long long)nbo, sums + 1);
  hipLaunchKernelGGL(rescale_finish_kernel, dim3(nbo), dim3(kBlock), 0, s, (const double*)raw, (const double*)mval,
                     orows, ocols, unitary, (const double*)sums, (const double*)(sums + 1), out, odt, ldo);
  return hipGetLastError();
}

}  // namespace mftx
