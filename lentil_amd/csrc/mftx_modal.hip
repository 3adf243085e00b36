// MI355X (gfx950) kernels of the modal OPD operations: the least-squares
// piston/tip/tilt fit of Plane.fit_tilt (plane.py:521-611) and the Zernike
// basis of zernike_basis (zernike.py:11-109,183-234).  FP64 throughout.
#include "mftx_device.h"

namespace mftx {

namespace {

constexpr int kBlock = 256;
constexpr int kNormal = 9;  // G00 G01 G02 G11 G12 G22 b0 b1 b2

__device__ __forceinline__ double load_real(const void* p, int dtype, long long i) {
  return dtype == 3 ? ((const double*)p)[i] : (double)((const float*)p)[i];
}

}  // namespace

// Normal equations of the per-segment fit opd ~ t0 a0 + t1 a1 + t2 a2 with
// a = mask * (1, r * dx_r, -c * dx_c), r and c the centred mesh
// (helper.py:7-17).  Grid (pixel blocks, segments); part[(seg * nblk + blk) * 9 + q].
__global__ void ptt_normal_kernel(const void* __restrict__ opd, int odt, long long ld_opd,
                                  const void* __restrict__ masks, int mdt, long long ld_mask, long long mask_stride,
                                  long long rows, long long cols, double dx_r, double dx_c,
                                  double* __restrict__ part) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int seg = blockIdx.y;
  double v[kNormal] = {};
  if (t < rows * cols) {
    const long long i = t / cols, j = t % cols;
    const double m = load_real(masks, mdt, seg * mask_stride + i * ld_mask + j);
    const double a0 = m, a1 = m * ((double)(i - rows / 2) * dx_r), a2 = m * (-(double)(j - cols / 2) * dx_c);
    const double b = load_real(opd, odt, i * ld_opd + j);
    v[0] = a0 * a0;
    v[1] = a0 * a1;
    v[2] = a0 * a2;
    v[3] = a1 * a1;
    v[4] = a1 * a2;
    v[5] = a2 * a2;
    v[6] = a0 * b;
    v[7] = a1 * b;
    v[8] = a2 * b;
  }
  __shared__ double w[kBlock / 64][kNormal];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kNormal; ++q) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) w[wave][q] = x;
  }
  __syncthreads();
  if (threadIdx.x < kNormal) {
    double s = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) s += w[k][threadIdx.x];
    part[((long long)seg * gridDim.x + blockIdx.x) * kNormal + threadIdx.x] = s;
  }
}

// out[seg][q] = sum over blocks (one block per segment, fixed order)
__global__ void ptt_reduce_kernel(const double* __restrict__ part, long long nblk, double* __restrict__ out) {
  const int seg = blockIdx.x;
  __shared__ double w[kBlock / 64][kNormal];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = 0; q < kNormal; ++q) {
    double x = 0.0;
    for (long long b = threadIdx.x; b < nblk; b += blockDim.x) x += part[((long long)seg * nblk + b) * kNormal + q];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) w[wave][q] = x;
  }
  __syncthreads();
  if (threadIdx.x < kNormal) {
    double s = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) s += w[k][threadIdx.x];
    out[seg * kNormal + threadIdx.x] = s;
  }
}

// Tilt removal (plane.py:590-609): with one segment (multi = 0)
//   out = opd - m (r dx_r t1 - c dx_c t2);
// with several, out = sum_s (opd - m_s (r dx_r t1_s - c dx_c t2_s)) m_s.
// tilt[s] = (t1_s, t2_s).
__global__ void remove_tilt_kernel(const void* __restrict__ opd, int odt, long long ld_opd,
                                   const void* __restrict__ masks, int mdt, long long ld_mask, long long mask_stride,
                                   int nseg, int multi, long long rows, long long cols, double dx_r, double dx_c,
                                   const double* __restrict__ tilt, double* __restrict__ out, long long ldo) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * cols) return;
  const long long i = t / cols, j = t % cols;
  const double o = load_real(opd, odt, i * ld_opd + j);
  const double ri = (double)(i - rows / 2) * dx_r, cj = -(double)(j - cols / 2) * dx_c;
  if (!multi) {
    const double m = load_real(masks, mdt, i * ld_mask + j);
    out[i * ldo + j] = o - (m * ri * tilt[0] + m * cj * tilt[1]);
    return;
  }
  double s = 0.0;
  for (int k = 0; k < nseg; ++k) {
    const double m = load_real(masks, mdt, k * mask_stride + i * ld_mask + j);
    s += (o - (m * ri * tilt[2 * k] + m * cj * tilt[2 * k + 1])) * m;
  }
  out[i * ldo + j] = s;
}

// One Zernike mode (zernike.py:64-93) per blockIdx.y at every pixel:
//   m == 0: c R(rho) mask   (n == 0: mask)
//   m != 0: c R(rho) cos(m theta) mask  (sin(m theta) for m < 0)
// with R(rho) = sum_k coef[k] rho^pow[k] accumulated in the order k = 0, 1, ..
// (zernike.py:96-109).  desc[mode] = (n, m, first term, term count) and
// norm[mode] = c.
__global__ void zernike_kernel(const double* __restrict__ rho, const double* __restrict__ theta,
                               const void* __restrict__ mask, int mdt, long long npix, const int* __restrict__ desc,
                               const double* __restrict__ coef, const int* __restrict__ pw,
                               const double* __restrict__ norm, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npix) return;
  const int mode = blockIdx.y;
  const int n = desc[4 * mode], m = desc[4 * mode + 1], k0 = desc[4 * mode + 2], nk = desc[4 * mode + 3];
  const double mk = load_real(mask, mdt, t) != 0.0 ? 1.0 : 0.0;
  double z;
  if (n == 0 && m == 0) {
    z = mk;
  } else {
    const double r = rho[t];
    double R = 0.0;
    for (int k = 0; k < nk; ++k) R += coef[k0 + k] * pow(r, (double)pw[k0 + k]);
    z = norm[mode] * R;
    if (m > 0) z = z * cos((double)m * theta[t]);
    else if (m < 0) z = z * sin((double)m * theta[t]);
    z = z * mk;
  }
  out[(long long)mode * npix + t] = z;
}

hipError_t launch_ptt_normal(const void* opd, int odt, long long ld_opd, const void* masks, int mdt, long long ld_mask,
                             long long mask_stride, int nseg, long long rows, long long cols, double dx_r, double dx_c,
                             double* part, double* out, hipStream_t s) {
  const long long nblk = (rows * cols + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(ptt_normal_kernel, dim3((unsigned)nblk, (unsigned)nseg), dim3(kBlock), 0, s, opd, odt, ld_opd,
                     masks, mdt, ld_mask, mask_stride, rows, cols, dx_r, dx_c, part);
  hipLaunchKernelGGL(ptt_reduce_kernel, dim3((unsigned)nseg), dim3(kBlock), 0, s, (const double*)part, nblk, out);
  return hipGetLastError();
}

hipError_t launch_remove_tilt(const void* opd, int odt, long long ld_opd, const void* masks, int mdt,
                              long long ld_mask, long long mask_stride, int nseg, int multi, long long rows,
                              long long cols, double dx_r, double dx_c, const double* tilt, double* out,
                              long long ldo, hipStream_t s) {
  const long long n = rows * cols;
  hipLaunchKernelGGL(remove_tilt_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, opd, odt,
                     ld_opd, masks, mdt, ld_mask, mask_stride, nseg, multi, rows, cols, dx_r, dx_c, tilt, out, ldo);
  return hipGetLastError();
}

hipError_t launch_zernike(const double* rho, const double* theta, const void* mask, int mdt, long long npix,
                          int nmodes, const int* desc, const double* coef, const int* pw, const double* norm,
                          double* out, hipStream_t s) {
  hipLaunchKernelGGL(zernike_kernel, dim3((unsigned)((npix + kBlock - 1) / kBlock), (unsigned)nmodes), dim3(kBlock), 0,
                     s, rho, theta, mask, mdt, npix, desc, coef, pw, norm, out);
  return hipGetLastError();
}

}  // namespace mftx
