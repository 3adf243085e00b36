// mftx_dft2_noalloc: mftx_dft2 on a fixed device array instead of the
// context's growable workspace, for callers that must not allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// The array `scratch` is file-scope device memory of the library's code
// object: every GPU the library runs on holds its own copy, from the first
// use on that device until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in symbol): the
// device address is looked up once per device with hipGetSymbolAddress,
// with that device current, and that address is what the launches use.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two host threads, or two streams with no ordering between them) race
// on it.  Calls queued on one stream are ordered and safe.
#include <mutex>
#include <string>

#include "../../include/mftx.h"
#include "mftx_internal.h"

namespace mftx {

constexpr size_t kScratchBytes = MFTX_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of this device's `scratch`, looked up on the first call for
// the device (mftx_create makes that call with the device current).
hipError_t scratch_address(int device, void** out) {
  static std::mutex mu;
  static void* ptr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    ptr[device] = p;
  }
  *out = ptr[device];
  return hipSuccess;
}

}  // namespace mftx

extern "C" int mftx_dft2_noalloc(mftx_ctx* ctx, const void* f, int f_dtype, int64_t m, int64_t n, int64_t ldf,
                                 void* F, int64_t M, int64_t N, int64_t ldF, double alpha_r, double alpha_c,
                                 double shift_r, double shift_c, double offset_r, double offset_c, double scale,
                                 int conj, int precision) {
  if (!ctx) return MFTX_EHANDLE;
  const int device = mftx::context_device(ctx);
  hipError_t e = hipSetDevice(device);
  void* ws = nullptr;
  if (e == hipSuccess) e = mftx::scratch_address(device, &ws);
  if (e != hipSuccess)
    return mftx::set_error(ctx, MFTX_EDEVICE,
                           (std::string("scratch address lookup: ") + hipGetErrorString(e)).c_str());
  return mftx::dft2_on_workspace(ctx, f, f_dtype, m, n, ldf, F, M, N, ldF, alpha_r, alpha_c, shift_r, shift_c,
                                 offset_r, offset_c, scale, conj, precision, ws, mftx::kScratchBytes);
}
