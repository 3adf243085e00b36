// C-ABI of the mftx library (declared in include/mftx.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mftx.h"
#include "mftx_internal.h"

using namespace mftx;

namespace {

constexpr int kRingSlots = 16;
constexpr size_t kSlotBytes = 512 * 1024;   // up to ~13k work items per launch

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }
int64_t floor_half(int64_t k) { return k / 2; }  // floor(k/2) for k >= 0

}  // namespace

struct mftx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // descriptor ring: pinned host staging + device slots, reused round-robin
  char* ring_host = nullptr;
  char* ring_dev = nullptr;
  hipEvent_t ring_ev[kRingSlots] = {};
  bool ring_used[kRingSlots] = {};
  int ring_next = 0;
};

struct mftx_plan {
  mftx_ctx* ctx = nullptr;
  int prec = 0;
  int64_t img_rows = 0, img_cols = 0;
  // device workspace and descriptor arrays
  void* ws = nullptr;
  size_t ws_bytes = 0;
  Work* works1 = nullptr;
  Work* works2 = nullptr;
  Problem* probs1 = nullptr;
  Problem* probs2 = nullptr;
  Segment* segs = nullptr;
  double* inv_wl = nullptr;
  int nwork1 = 0, nwork2 = 0;
  int epi2 = EPI_INTENSITY;
  // phasor launches
  struct Phasor {
    mftx_tile tile;
    int nwave;
    int wave_off;          // offset into inv_wl
    void* dst;
    long long wave_stride; // bytes
    long long rows_pad;
    int kpad;
  };
  std::vector<Phasor> phasors;
  // reductions of partial images onto the detector
  struct Reduce {
    float* partial;
    long long gstride;
    int ngroups, rows, cols;
    long long ldp;
    long long r0, c0;
  };
  std::vector<Reduce> reduces;
  // fallback (partially overlapping coherent groups): complex accumulators
  struct Fallback {
    float2* acc;                    // image-sized complex accumulator for the group
    std::vector<std::tuple<const float2*, int, int, int64_t, int64_t, long long>> adds;  // F, M, N, r0, c0, ldF
    double weight;
  };
  std::vector<Fallback> fallbacks;
  mftx_plan_info info = {};
  // graph replay
  int use_graph = 0;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  void* g_image = nullptr;
  int g_dtype = -1, g_accum = -1;
  int64_t g_ld = -1;
};

static int fail(mftx_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

static int hipfail(mftx_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, MFTX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, expr)                                      \
  do {                                                         \
    hipError_t e_ = (expr);                                    \
    if (e_ != hipSuccess) return hipfail((ctx), e_, #expr);    \
  } while (0)

static int ensure_ws(mftx_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->ws_bytes) return MFTX_OK;
  if (ctx->ws) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  size_t nb = std::max(bytes, (size_t)(ctx->ws_bytes * 3 / 2));
  if (hipMalloc(&ctx->ws, nb) != hipSuccess) {
    ctx->ws = nullptr;
    return fail(ctx, MFTX_ENOMEM, "workspace allocation of " + std::to_string(nb) + " bytes failed");
  }
  ctx->ws_bytes = nb;
  return MFTX_OK;
}

// Stage `bytes` of descriptors into a ring slot and queue the upload on the
// context stream; returns the device address of the slot.
static int upload_desc(mftx_ctx* ctx, const void* src, size_t bytes, void** dev) {
  if (bytes > kSlotBytes) return fail(ctx, MFTX_EINVAL, "descriptor block too large");
  const int s = ctx->ring_next;
  ctx->ring_next = (s + 1) % kRingSlots;
  if (ctx->ring_used[s]) HIPCHK(ctx, hipEventSynchronize(ctx->ring_ev[s]));
  char* h = ctx->ring_host + (size_t)s * kSlotBytes;
  char* d = ctx->ring_dev + (size_t)s * kSlotBytes;
  std::memcpy(h, src, bytes);
  HIPCHK(ctx, hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ring_ev[s], ctx->stream));
  ctx->ring_used[s] = true;
  *dev = d;
  return MFTX_OK;
}

extern "C" {

int mftx_version(void) { return 100; }

int mftx_create(int device, mftx_ctx** out) {
  if (!out) return MFTX_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MFTX_EDEVICE;
  if (device < 0 || device >= n) return MFTX_EINVAL;
  mftx_ctx* ctx = new mftx_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipHostMalloc((void**)&ctx->ring_host, kRingSlots * kSlotBytes, hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&ctx->ring_dev, kRingSlots * kSlotBytes) != hipSuccess) {
    delete ctx;
    return MFTX_ENOMEM;
  }
  for (int i = 0; i < kRingSlots; ++i) {
    if (hipEventCreateWithFlags(&ctx->ring_ev[i], hipEventDisableTiming) != hipSuccess) {
      delete ctx;
      return MFTX_EDEVICE;
    }
  }
  *out = ctx;
  return MFTX_OK;
}

int mftx_destroy(mftx_ctx* ctx) {
  if (!ctx) return MFTX_EHANDLE;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (int i = 0; i < kRingSlots; ++i)
    if (ctx->ring_ev[i]) (void)hipEventDestroy(ctx->ring_ev[i]);
  if (ctx->ring_host) (void)hipHostFree(ctx->ring_host);
  if (ctx->ring_dev) (void)hipFree(ctx->ring_dev);
  if (ctx->ws) (void)hipFree(ctx->ws);
  delete ctx;
  return MFTX_OK;
}

const char* mftx_last_error(const mftx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mftx_set_stream(mftx_ctx* ctx, void* stream) {
  if (!ctx) return MFTX_EHANDLE;
  if ((hipStream_t)stream != ctx->stream) HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->stream = (hipStream_t)stream;
  return MFTX_OK;
}

int mftx_synchronize(mftx_ctx* ctx) {
  if (!ctx) return MFTX_EHANDLE;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return MFTX_OK;
}

static bool valid_dtype(int d) { return d >= 0 && d <= 3; }

// Fill a single-problem GEMM descriptor set and launch it on tiles covering
// rows_pad x cols of C.
static int run_single(mftx_ctx* ctx, int prec, int epi, const Problem& p, const Segment& sg,
                      long long rows_tiled) {
  const int tr = (int)(rows_tiled / BM);
  const int tc = (int)((p.cols + BN - 1) / BN);
  const int nw = tr * tc;
  const size_t need = sizeof(Problem) + sizeof(Segment) + sizeof(Work) * (size_t)nw;
  std::vector<char> host(need);
  Problem* hp = (Problem*)host.data();
  Segment* hs = (Segment*)(host.data() + sizeof(Problem));
  Work* hw = (Work*)(host.data() + sizeof(Problem) + sizeof(Segment));
  *hp = p;
  hp->seg_begin = 0;
  hp->seg_count = 1;
  *hs = sg;
  // XCD-aware order: works of one tile row (sharing A) land on one XCD
  std::vector<std::vector<Work>> bins(8);
  for (int r = 0; r < tr; ++r)
    for (int c = 0; c < tc; ++c) bins[r % 8].push_back(Work{0, 1, r, c, nullptr, 0, 0, 0});
  size_t mx = 0;
  for (auto& b : bins) mx = std::max(mx, b.size());
  std::vector<Work> order;
  for (size_t i = 0; i < mx; ++i)
    for (int x = 0; x < 8; ++x) order.push_back(i < bins[x].size() ? bins[x][i] : Work{0, 0, 0, 0, nullptr, 0, 0, 0});
  // trim trailing empty works
  while (!order.empty() && order.back().pend == 0) order.pop_back();
  if (order.size() * sizeof(Work) + sizeof(Problem) + sizeof(Segment) > kSlotBytes) {
    order.clear();
    for (int r = 0; r < tr; ++r)
      for (int c = 0; c < tc; ++c) order.push_back(Work{0, 1, r, c, nullptr, 0, 0, 0});
  }
  host.resize(sizeof(Problem) + sizeof(Segment) + sizeof(Work) * order.size());
  hp = (Problem*)host.data();
  hw = (Work*)(host.data() + sizeof(Problem) + sizeof(Segment));
  std::memcpy(hw, order.data(), sizeof(Work) * order.size());
  void* dev = nullptr;
  int rc = upload_desc(ctx, host.data(), host.size(), &dev);
  if (rc) return rc;
  const Problem* dp = (const Problem*)dev;
  const Segment* ds = (const Segment*)((char*)dev + sizeof(Problem));
  const Work* dw = (const Work*)((char*)dev + sizeof(Problem) + sizeof(Segment));
  HIPCHK(ctx, launch_gemm(prec, epi, dw, (int)order.size(), dp, ds, ctx->stream));
  return MFTX_OK;
}

int mftx_dft2(mftx_ctx* ctx, const void* f, int f_dtype, int64_t m, int64_t n, int64_t ldf, void* F, int64_t M,
              int64_t N, int64_t ldF, double alpha_r, double alpha_c, double shift_r, double shift_c,
              double offset_r, double offset_c, double scale, int conj, int precision) {
  if (!ctx) return MFTX_EHANDLE;
  if (!f || !F) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if (m <= 0 || n <= 0 || M <= 0 || N <= 0) return fail(ctx, MFTX_EINVAL, "shapes must be positive");
  if (ldf < n || ldF < N) return fail(ctx, MFTX_EINVAL, "row stride smaller than row length");
  if (!valid_dtype(f_dtype)) return fail(ctx, MFTX_EINVAL, "unknown input dtype");
  if (precision != MFTX_PREC_FP32 && precision != MFTX_PREC_BF16X3)
    return fail(ctx, MFTX_EINVAL, "unknown precision");
  if (m > (1 << 20) || n > (1 << 20) || M > (1 << 20) || N > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "dimension too large");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const double sgn = conj ? 1.0 : -1.0;
  const int prec = precision;
  // order B: G = f . E2 (stage 1, stored as G^T), F^T = G^T . E1^T (stage 2)
  const int64_t rows1 = round_up(m, BM);
  const int kpad1 = (int)round_up(n, BK);
  const int64_t rows2 = round_up(N, BM);
  const int kpad2 = (int)round_up(m, BK);
  const size_t b1 = packed_bytes(prec, rows1, kpad1);
  const size_t b2 = packed_bytes(prec, rows2, kpad2);
  int rc = ensure_ws(ctx, b1 + b2 + 256);
  if (rc) return rc;
  char* a1 = (char*)ctx->ws;
  char* a2 = a1 + round_up((int64_t)b1, 256);
  HIPCHK(ctx, launch_pack(prec, f, f_dtype, m, n, ldf, a1, rows1, kpad1, ctx->stream));

  Problem p1 = {};
  p1.a = a1;
  p1.lda = packed_lda(prec, kpad1);
  p1.alpha = sgn * alpha_c;
  p1.y0 = -(double)floor_half(N) - shift_c;
  p1.rows = (int)m;
  p1.cols = (int)N;
  p1.kpad = kpad1;
  p1.scale = 1.0f;
  p1.limit = kpad2;
  p1.dst = a2;
  p1.ldd = packed_lda(prec, kpad2);
  Segment s1 = {0, kpad1, -(double)floor_half(n) + offset_c};
  rc = run_single(ctx, prec, EPI_STORE_PACKED_T, p1, s1, rows1);
  if (rc) return rc;

  Problem p2 = {};
  p2.a = a2;
  p2.lda = packed_lda(prec, kpad2);
  p2.alpha = sgn * alpha_r;
  p2.y0 = -(double)floor_half(M) - shift_r;
  p2.rows = (int)N;
  p2.cols = (int)M;
  p2.kpad = kpad2;
  p2.scale = (float)scale;
  p2.dst = F;
  p2.ldd = ldF;
  p2.trans = 1;
  Segment s2 = {0, kpad2, -(double)floor_half(m) + offset_r};
  return run_single(ctx, prec, EPI_STORE_CPLX, p2, s2, rows2);
}

int mftx_phasor(mftx_ctx* ctx, const void* amp, int amp_dtype, int64_t ld_amp, const void* opd, int opd_dtype,
                int64_t ld_opd, int64_t m, int64_t n, double wavelength, void* out) {
  if (!ctx) return MFTX_EHANDLE;
  if (!amp || !opd || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if ((amp_dtype != 1 && amp_dtype != 3) || (opd_dtype != 1 && opd_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "amp/opd must be float32 or float64");
  if (m <= 0 || n <= 0 || ld_amp < n || ld_opd < n) return fail(ctx, MFTX_EINVAL, "bad phasor shape");
  if (!(wavelength > 0)) return fail(ctx, MFTX_EINVAL, "wavelength must be positive");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_phasor_field(amp, amp_dtype, ld_amp, opd, opd_dtype, ld_opd, (int)m, (int)n, 1.0 / wavelength,
                                  out, ctx->stream));
  return MFTX_OK;
}

int mftx_insert(mftx_ctx* ctx, const void* field, int64_t ldf, int64_t fr0, int64_t fc0, int64_t h, int64_t w,
                void* out, int out_dtype, int64_t ldo, int64_t or0, int64_t oc0, int intensity, double weight) {
  if (!ctx) return MFTX_EHANDLE;
  if (!field || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if (h < 0 || w < 0 || fr0 < 0 || fc0 < 0 || or0 < 0 || oc0 < 0) return fail(ctx, MFTX_EINVAL, "negative slice");
  if (intensity && out_dtype != 1 && out_dtype != 3) return fail(ctx, MFTX_EINVAL, "intensity output must be real");
  if (!intensity && out_dtype != 0 && out_dtype != 2) return fail(ctx, MFTX_EINVAL, "field output must be complex");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_insert(field, ldf, (int)fr0, (int)fc0, (int)h, (int)w, out, out_dtype, ldo, (int)or0,
                            (int)oc0, intensity, weight, ctx->stream));
  return MFTX_OK;
}

int mftx_multiply(mftx_ctx* ctx, const void* a, int64_t lda, const void* b, int64_t ldb, void* out, int64_t ldo,
                  int64_t h, int64_t w) {
  if (!ctx) return MFTX_EHANDLE;
  if (!a || !b || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if (h < 0 || w < 0 || lda < w || ldb < w || ldo < w) return fail(ctx, MFTX_EINVAL, "bad multiply shape");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_multiply(a, lda, b, ldb, out, ldo, (int)h, (int)w, ctx->stream));
  return MFTX_OK;
}

int mftx_zero(mftx_ctx* ctx, void* p, int64_t bytes) {
  if (!ctx) return MFTX_EHANDLE;
  if (!p || bytes < 0) return fail(ctx, MFTX_EINVAL, "bad zero request");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_zero(p, (size_t)bytes, ctx->stream));
  return MFTX_OK;
}

// ---------------------------------------------------------------------------
// batched plan
// ---------------------------------------------------------------------------

}  // extern "C"

namespace {

struct ClassKey {
  int wave, M, N;
  double ar, ac, sr, sc;
  int64_t r0, c0;
  bool operator<(const ClassKey& o) const {
    return std::tie(wave, M, N, ar, ac, sr, sc, r0, c0) < std::tie(o.wave, o.M, o.N, o.ar, o.ac, o.sr, o.sc, o.r0, o.c0);
  }
};

// XCD-aware interleave: clusters of works that share the same A rows are
// assigned round-robin to the 8 XCDs; workgroup i runs on XCD i % 8.
std::vector<Work> xcd_order(const std::vector<std::vector<Work>>& clusters) {
  std::vector<std::vector<Work>> bins(8);
  for (size_t c = 0; c < clusters.size(); ++c) {
    auto& b = bins[c % 8];
    b.insert(b.end(), clusters[c].begin(), clusters[c].end());
  }
  size_t mx = 0;
  for (auto& b : bins) mx = std::max(mx, b.size());
  std::vector<Work> out;
  out.reserve(mx * 8);
  for (size_t i = 0; i < mx; ++i)
    for (int x = 0; x < 8; ++x) out.push_back(i < bins[x].size() ? bins[x][i] : Work{0, 0, 0, 0, nullptr, 0, 0, 0});
  while (!out.empty() && out.back().pend == 0) out.pop_back();
  return out;
}

}  // namespace

extern "C" {

int mftx_plan_psf(mftx_ctx* ctx, const mftx_tile* tiles, int32_t ntiles, const double* wavelengths, int32_t nwave,
                  const mftx_problem* problems, int32_t nprob, int64_t image_rows, int64_t image_cols,
                  int32_t precision, int32_t lambda_group, mftx_plan** out) {
  if (!ctx) return MFTX_EHANDLE;
  if (!out || !tiles || !wavelengths || (!problems && nprob > 0)) return fail(ctx, MFTX_EINVAL, "null argument");
  *out = nullptr;
  if (ntiles <= 0 || nwave <= 0 || nprob < 0) return fail(ctx, MFTX_EINVAL, "empty plan");
  if (precision != MFTX_PREC_FP32 && precision != MFTX_PREC_BF16X3) return fail(ctx, MFTX_EINVAL, "unknown precision");
  if (image_rows <= 0 || image_cols <= 0) return fail(ctx, MFTX_EINVAL, "bad image shape");
  if (lambda_group <= 0) lambda_group = 4;
  for (int t = 0; t < ntiles; ++t) {
    const mftx_tile& T = tiles[t];
    if (!T.amp || !T.opd || T.m <= 0 || T.n <= 0 || T.ld_amp < T.n || T.ld_opd < T.n ||
        (T.amp_dtype != 1 && T.amp_dtype != 3) || (T.opd_dtype != 1 && T.opd_dtype != 3))
      return fail(ctx, MFTX_EINVAL, "bad tile " + std::to_string(t));
  }
  for (int w = 0; w < nwave; ++w)
    if (!(wavelengths[w] > 0)) return fail(ctx, MFTX_EINVAL, "wavelengths must be positive");
  for (int i = 0; i < nprob; ++i) {
    const mftx_problem& P = problems[i];
    if (P.tile < 0 || P.tile >= ntiles || P.wave < 0 || P.wave >= nwave || P.M <= 0 || P.N <= 0 ||
        P.M > (1 << 20) || P.N > (1 << 20))
      return fail(ctx, MFTX_EINVAL, "bad problem " + std::to_string(i));
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int prec = precision;
  mftx_plan* plan = new mftx_plan();
  plan->ctx = ctx;
  plan->prec = prec;
  plan->img_rows = image_rows;
  plan->img_cols = image_cols;

  // ---- classes: problems that can share one K-concatenated stage-2 GEMM ----
  std::map<ClassKey, std::vector<int>> classes;
  for (int i = 0; i < nprob; ++i) {
    const mftx_problem& P = problems[i];
    classes[ClassKey{P.wave, P.M, P.N, P.alpha_r, P.alpha_c, P.shift_r, P.shift_c, P.out_r0, P.out_c0}].push_back(i);
  }
  // coherent groups -> classes
  std::map<std::pair<int, int>, std::vector<const std::vector<int>*>> groups;  // (wave, group) -> classes
  for (auto& kv : classes) {
    // a class must not mix coherent groups; split it if it does
    groups[{kv.first.wave, problems[kv.second[0]].group}].push_back(&kv.second);
  }
  for (auto& kv : classes)
    for (int i : kv.second)
      if (problems[i].group != problems[kv.second[0]].group) {
        delete plan;
        return fail(ctx, MFTX_EINVAL, "problems with identical geometry must share a coherent group");
      }
  bool fast = true;
  for (auto& g : groups)
    if (g.second.size() > 1) fast = false;
  plan->info.coherent_fast = fast ? 1 : 0;

  // ---- workspace layout ----------------------------------------------------
  struct Buf { size_t off, bytes; };
  size_t ws_total = 0;
  auto reserve = [&](size_t bytes) { Buf b{ws_total, bytes}; ws_total += (size_t)round_up((int64_t)bytes, 256); return b; };
  // Buffer offsets are carried in pointer fields until the workspace exists;
  // a leading pad keeps every real offset non-zero.
  reserve(256);

  // phasor buffers: per tile, all wavelengths the tile is used with
  std::vector<std::vector<int>> tile_waves(ntiles);
  for (int i = 0; i < nprob; ++i) tile_waves[problems[i].tile].push_back(problems[i].wave);
  std::vector<int> inv_order;        // wavelength index list for inv_wl upload
  std::vector<std::map<int, int>> tile_wave_slot(ntiles);
  std::vector<Buf> tile_buf(ntiles);
  std::vector<int64_t> tile_rows(ntiles), tile_kpad(ntiles);
  std::vector<int> tile_wave_off(ntiles);
  for (int t = 0; t < ntiles; ++t) {
    auto& v = tile_waves[t];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    tile_rows[t] = round_up(tiles[t].m, BM);
    tile_kpad[t] = round_up(tiles[t].n, BK);
    tile_wave_off[t] = (int)inv_order.size();
    for (size_t j = 0; j < v.size(); ++j) {
      tile_wave_slot[t][v[j]] = (int)j;
      inv_order.push_back(v[j]);
    }
    tile_buf[t] = reserve(packed_bytes(prec, tile_rows[t], (int)tile_kpad[t]) * v.size());
  }

  // stage-2 A buffers: one per class, K = sum of member kpads (roundup(m_t, BK))
  struct ClassInfo {
    const std::vector<int>* members;
    ClassKey key;
    int64_t rows2;   // round_up(N, BM)
    int ktot;
    Buf abuf;
    Buf fbuf;        // fallback complex output
    std::vector<int> kbeg;
  };
  std::vector<ClassInfo> cinfo;
  for (auto& kv : classes) {
    ClassInfo ci;
    ci.members = &kv.second;
    ci.key = kv.first;
    ci.rows2 = round_up(kv.first.N, BM);
    int k = 0;
    for (int i : kv.second) {
      ci.kbeg.push_back(k);
      k += (int)round_up(tiles[problems[i].tile].m, BK);
    }
    ci.ktot = k;
    ci.abuf = reserve(packed_bytes(prec, ci.rows2, ci.ktot));
    if (!fast) ci.fbuf = reserve((size_t)kv.first.M * kv.first.N * sizeof(float2));
    cinfo.push_back(ci);
  }

  // λ-groups for the intensity epilogue: classes with identical output
  // geometry (M, N, r0, c0) in chunks of lambda_group
  std::map<std::tuple<int, int, int64_t, int64_t>, std::vector<int>> geo;  // -> class indices
  if (fast)
    for (size_t c = 0; c < cinfo.size(); ++c)
      geo[{cinfo[c].key.M, cinfo[c].key.N, cinfo[c].key.r0, cinfo[c].key.c0}].push_back((int)c);
  struct GeoInfo { std::vector<std::vector<int>> lgroups; Buf partial; int64_t prow, pcol; };
  std::vector<GeoInfo> ginfo;
  for (auto& kv : geo) {
    GeoInfo gi;
    auto& cl = kv.second;
    for (size_t s = 0; s < cl.size(); s += lambda_group)
      gi.lgroups.emplace_back(cl.begin() + s, cl.begin() + std::min(cl.size(), s + (size_t)lambda_group));
    gi.prow = round_up(std::get<0>(kv.first), BN);  // partial[u][v]: u over M (C cols)
    gi.pcol = round_up(std::get<1>(kv.first), BM);  // v over N (C rows)
    gi.partial = reserve(sizeof(float) * gi.prow * gi.pcol * gi.lgroups.size());
    ginfo.push_back(gi);
  }
  std::vector<Buf> fb_acc;
  std::map<std::pair<int, int>, int> group_fb;
  if (!fast) {
    for (auto& g : groups) {
      group_fb[g.first] = (int)fb_acc.size();
      fb_acc.push_back(reserve(sizeof(float2) * image_rows * image_cols));
    }
  }

  // descriptor arrays (sizes known after building them on the host)
  std::vector<Problem> p1, p2;
  std::vector<Segment> segs;
  std::vector<std::vector<Work>> clus1, clus2;

  // stage 1: one problem per (tile, λ) member, output into its class A buffer
  // stage 2: one problem per class
  for (size_t c = 0; c < cinfo.size(); ++c) {
    ClassInfo& ci = cinfo[c];
    const ClassKey& K = ci.key;
    Problem q2 = {};
    q2.lda = packed_lda(prec, ci.ktot);
    q2.alpha = -K.ar;
    q2.y0 = -(double)floor_half(K.M) - K.sr;
    q2.rows = K.N;
    q2.cols = K.M;
    q2.kpad = ci.ktot;
    q2.seg_begin = (int)segs.size();
    q2.seg_count = (int)ci.members->size();
    q2.trans = 1;
    for (size_t j = 0; j < ci.members->size(); ++j) {
      const mftx_problem& P = problems[(*ci.members)[j]];
      const mftx_tile& T = tiles[P.tile];
      const int kp2 = (int)round_up(T.m, BK);
      segs.push_back(Segment{ci.kbeg[j], ci.kbeg[j] + kp2, -(double)floor_half(T.m) + P.offset_r});
      Problem q1 = {};
      const int slot = tile_wave_slot[P.tile][P.wave];
      q1.a = (void*)(tile_buf[P.tile].off + (size_t)slot * packed_bytes(prec, tile_rows[P.tile], (int)tile_kpad[P.tile]));
      q1.lda = packed_lda(prec, (int)tile_kpad[P.tile]);
      q1.alpha = -P.alpha_c;
      q1.y0 = -(double)floor_half(P.N) - P.shift_c;
      q1.rows = (int)T.m;
      q1.cols = P.N;
      q1.kpad = (int)tile_kpad[P.tile];
      q1.seg_begin = -1;  // patched below (own segment)
      q1.seg_count = 1;
      q1.scale = 1.0f;
      q1.limit = kp2;
      q1.dst = (void*)(ci.abuf.off + (size_t)ci.kbeg[j] * (prec == PREC_FP32 ? sizeof(float2) : 4 * sizeof(__bf16)));
      q1.ldd = packed_lda(prec, ci.ktot);
      const int pi1 = (int)p1.size();
      p1.push_back(q1);
      // stage-1 segment
      // (appended after all class segments; index fixed up below)
      std::vector<Work> cl;
      const int trs = (int)(tile_rows[P.tile] / BM), tcs = (int)((P.N + BN - 1) / BN);
      for (int r = 0; r < trs; ++r) {
        std::vector<Work> row;
        for (int cc = 0; cc < tcs; ++cc) row.push_back(Work{pi1, pi1 + 1, r, cc, nullptr, 0, 0, 0});
        clus1.push_back(row);
      }
    }
    const double unit = std::sqrt(std::fabs(K.ar * K.ac));
    q2.scale = (float)(unit * unit * problems[(*ci.members)[0]].weight);
    if (!fast) {
      q2.scale = (float)unit;
      q2.dst = (void*)ci.fbuf.off;
      q2.ldd = K.N;  // F[u][v], u < M
    }
    p2.push_back(q2);
  }
  // stage-1 segments: x_s = s - floor(n/2) + offset_c
  for (size_t i = 0, c = 0; c < cinfo.size(); ++c)
    for (size_t j = 0; j < cinfo[c].members->size(); ++j, ++i) {
      const mftx_problem& P = problems[(*cinfo[c].members)[j]];
      const mftx_tile& T = tiles[P.tile];
      p1[i].seg_begin = (int)segs.size();
      segs.push_back(Segment{0, (int)tile_kpad[P.tile], -(double)floor_half(T.n) + P.offset_c});
    }
  // stage-2 works
  if (fast) {
    for (auto& gi : ginfo) {
      for (size_t g = 0; g < gi.lgroups.size(); ++g) {
        const auto& lg = gi.lgroups[g];
        const ClassInfo& c0 = cinfo[lg[0]];
        // problems of a λ-group must be contiguous in p2: they are (class order)
        // only if lg is a contiguous run; build explicit contiguous copies
        const int pb = (int)p2.size();
        for (int c : lg) p2.push_back(p2[c]);
        const int pe = (int)p2.size();
        const int trs = (int)(c0.rows2 / BM), tcs = (int)((c0.key.M + BN - 1) / BN);
        for (int r = 0; r < trs; ++r) {
          std::vector<Work> row;
          for (int cc = 0; cc < tcs; ++cc)
            row.push_back(Work{pb, pe, r, cc, (float*)(gi.partial.off + sizeof(float) * gi.prow * gi.pcol * g),
                               (long long)gi.pcol, c0.key.M, c0.key.N});
          clus2.push_back(row);
        }
      }
    }
  } else {
    plan->epi2 = EPI_STORE_CPLX;
    for (size_t c = 0; c < cinfo.size(); ++c) {
      const int trs = (int)(cinfo[c].rows2 / BM), tcs = (int)((cinfo[c].key.M + BN - 1) / BN);
      for (int r = 0; r < trs; ++r) {
        std::vector<Work> row;
        for (int cc = 0; cc < tcs; ++cc) row.push_back(Work{(int)c, (int)c + 1, r, cc, nullptr, 0, 0, 0});
        clus2.push_back(row);
      }
    }
  }
  std::vector<Work> w1 = xcd_order(clus1), w2 = xcd_order(clus2);

  // descriptor block
  Buf bw1 = reserve(sizeof(Work) * std::max<size_t>(w1.size(), 1));
  Buf bw2 = reserve(sizeof(Work) * std::max<size_t>(w2.size(), 1));
  Buf bp1 = reserve(sizeof(Problem) * std::max<size_t>(p1.size(), 1));
  Buf bp2 = reserve(sizeof(Problem) * std::max<size_t>(p2.size(), 1));
  Buf bsg = reserve(sizeof(Segment) * std::max<size_t>(segs.size(), 1));
  Buf biw = reserve(sizeof(double) * std::max<size_t>(inv_order.size(), 1));

  if (hipMalloc(&plan->ws, ws_total) != hipSuccess) {
    delete plan;
    return fail(ctx, MFTX_ENOMEM, "plan workspace allocation of " + std::to_string(ws_total) + " bytes failed");
  }
  plan->ws_bytes = ws_total;
  char* base = (char*)plan->ws;
  for (size_t i = 0; i < p1.size(); ++i) {
    p1[i].a = base + (size_t)p1[i].a;
    p1[i].dst = base + (size_t)p1[i].dst;
  }
  // p2 holds one entry per class, then (fast path) the λ-group copies; each
  // class owns a distinct seg_begin, which identifies the class of a copy
  for (size_t i = 0; i < p2.size(); ++i) {
    size_t c = 0;
    while (c < cinfo.size() && p2[i].seg_begin != p2[c].seg_begin) ++c;
    p2[i].a = base + cinfo[c].abuf.off;
    if (!fast) p2[i].dst = base + cinfo[c].fbuf.off;
  }
  for (auto& w : w2)
    if (w.partial) w.partial = (float*)(base + (size_t)w.partial);

  std::vector<double> iw(inv_order.size());
  for (size_t i = 0; i < inv_order.size(); ++i) iw[i] = 1.0 / wavelengths[inv_order[i]];

  plan->works1 = (Work*)(base + bw1.off);
  plan->works2 = (Work*)(base + bw2.off);
  plan->probs1 = (Problem*)(base + bp1.off);
  plan->probs2 = (Problem*)(base + bp2.off);
  plan->segs = (Segment*)(base + bsg.off);
  plan->inv_wl = (double*)(base + biw.off);
  plan->nwork1 = (int)w1.size();
  plan->nwork2 = (int)w2.size();
  hipError_t e = hipSuccess;
  if (!w1.empty()) e = hipMemcpy(plan->works1, w1.data(), sizeof(Work) * w1.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !w2.empty()) e = hipMemcpy(plan->works2, w2.data(), sizeof(Work) * w2.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !p1.empty()) e = hipMemcpy(plan->probs1, p1.data(), sizeof(Problem) * p1.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !p2.empty()) e = hipMemcpy(plan->probs2, p2.data(), sizeof(Problem) * p2.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !segs.empty()) e = hipMemcpy(plan->segs, segs.data(), sizeof(Segment) * segs.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !iw.empty()) e = hipMemcpy(plan->inv_wl, iw.data(), sizeof(double) * iw.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(plan->ws);
    delete plan;
    return hipfail(ctx, e, "plan descriptor upload");
  }

  for (int t = 0; t < ntiles; ++t) {
    if (tile_waves[t].empty()) continue;
    mftx_plan::Phasor ph;
    ph.tile = tiles[t];
    ph.nwave = (int)tile_waves[t].size();
    ph.wave_off = tile_wave_off[t];
    ph.dst = base + tile_buf[t].off;
    ph.wave_stride = (long long)packed_bytes(prec, tile_rows[t], (int)tile_kpad[t]);
    ph.rows_pad = tile_rows[t];
    ph.kpad = (int)tile_kpad[t];
    plan->phasors.push_back(ph);
  }
  if (fast) {
    for (auto& gi : ginfo) {
      const ClassInfo& c0 = cinfo[gi.lgroups[0][0]];
      plan->reduces.push_back(mftx_plan::Reduce{(float*)(base + gi.partial.off), (long long)(gi.prow * gi.pcol),
                                                (int)gi.lgroups.size(), c0.key.M, c0.key.N, (long long)gi.pcol,
                                                c0.key.r0, c0.key.c0});
    }
  } else {
    for (auto& g : groups) {
      mftx_plan::Fallback fb;
      fb.acc = (float2*)(base + fb_acc[group_fb[g.first]].off);
      fb.weight = problems[(*g.second[0])[0]].weight;
      for (auto* mem : g.second) {
        size_t c = 0;
        for (; c < cinfo.size(); ++c)
          if (cinfo[c].members == mem) break;
        const ClassKey& K = cinfo[c].key;
        fb.adds.emplace_back((const float2*)(base + cinfo[c].fbuf.off), K.M, K.N, K.r0, K.c0, (long long)K.N);
      }
      plan->fallbacks.push_back(fb);
    }
  }

  // ---- info ------------------------------------------------------------------
  double fl = 0, g1 = 0, g2 = 0;
  for (int i = 0; i < nprob; ++i) {
    const mftx_problem& P = problems[i];
    const double m = (double)tiles[P.tile].m, n = (double)tiles[P.tile].n, M = P.M, N = P.N;
    fl += 8.0 * std::min(M * m * n + M * n * N, m * n * N + M * m * N);
    g1 += 8.0 * m * n * N;
    g2 += 8.0 * M * m * N;
  }
  plan->info.flops = fl;
  plan->info.gemm_flops[0] = g1;
  plan->info.gemm_flops[1] = g2;
  double bytes = 0;
  for (auto& ph : plan->phasors) bytes += (double)ph.wave_stride * ph.nwave * 2;  // write + read
  for (auto& ci : cinfo) bytes += 2.0 * packed_bytes(prec, ci.rows2, ci.ktot);
  bytes += 4.0 * image_rows * image_cols;
  plan->info.bytes_hbm_min = bytes;
  plan->info.n_stage1 = (int)p1.size();
  plan->info.n_stage2 = (int)cinfo.size();
  plan->info.n_work1 = plan->nwork1;
  plan->info.n_work2 = plan->nwork2;
  plan->info.n_launch = (int)plan->phasors.size() + 2 + (int)plan->reduces.size();
  *out = plan;
  return MFTX_OK;
}

static int plan_launch(mftx_plan* plan, void* image, int dtype, int64_t ldi, int accumulate) {
  mftx_ctx* ctx = plan->ctx;
  hipStream_t s = ctx->stream;
  for (auto& ph : plan->phasors)
    HIPCHK(ctx, launch_phasor(plan->prec, ph.tile.amp, ph.tile.amp_dtype, ph.tile.ld_amp, ph.tile.opd,
                              ph.tile.opd_dtype, ph.tile.ld_opd, (int)ph.tile.m, (int)ph.tile.n,
                              plan->inv_wl + ph.wave_off, ph.nwave, ph.dst, ph.wave_stride, ph.rows_pad, ph.kpad, s));
  HIPCHK(ctx, launch_gemm(plan->prec, EPI_STORE_PACKED_T, plan->works1, plan->nwork1, plan->probs1, plan->segs, s));
  HIPCHK(ctx, launch_gemm(plan->prec, plan->epi2, plan->works2, plan->nwork2, plan->probs2, plan->segs, s));
  if (plan->epi2 == EPI_INTENSITY) {
    bool first = true;
    for (auto& r : plan->reduces) {
      // overwrite only on the first reduce that covers the whole image; later
      // reduces (other output geometries) always accumulate
      int acc = accumulate;
      if (!accumulate && !first) acc = 1;
      if (!accumulate && first) {
        if (!(r.r0 == 0 && r.c0 == 0 && r.rows == plan->img_rows && r.cols == plan->img_cols)) {
          HIPCHK(ctx, launch_zero(image, (size_t)(dtype == 3 ? 8 : 4) * plan->img_rows * ldi, s));
          acc = 1;
        }
      }
      first = false;
      HIPCHK(ctx, launch_reduce_partials(r.partial, r.gstride, r.ngroups, r.rows, r.cols, r.ldp, image, dtype, ldi,
                                         (int)plan->img_rows, (int)plan->img_cols, (int)r.r0, (int)r.c0, acc, s));
    }
    if (!accumulate && plan->reduces.empty())
      HIPCHK(ctx, launch_zero(image, (size_t)(dtype == 3 ? 8 : 4) * plan->img_rows * ldi, s));
  } else {
    if (!accumulate) HIPCHK(ctx, launch_zero(image, (size_t)(dtype == 3 ? 8 : 4) * plan->img_rows * ldi, s));
    for (auto& fb : plan->fallbacks) {
      HIPCHK(ctx, launch_zero(fb.acc, sizeof(float2) * plan->img_rows * plan->img_cols, s));
      for (auto& a : fb.adds) {
        const float2* F = std::get<0>(a);
        const int M = std::get<1>(a), N = std::get<2>(a);
        const int64_t r0 = std::get<3>(a), c0 = std::get<4>(a);
        const long long ldF = std::get<5>(a);
        // clip F's placement to the image
        const int64_t or0 = std::max<int64_t>(r0, 0), oc0 = std::max<int64_t>(c0, 0);
        const int64_t or1 = std::min<int64_t>(r0 + M, plan->img_rows), oc1 = std::min<int64_t>(c0 + N, plan->img_cols);
        if (or1 <= or0 || oc1 <= oc0) continue;
        HIPCHK(ctx, launch_insert(F, ldF, (int)(or0 - r0), (int)(oc0 - c0), (int)(or1 - or0), (int)(oc1 - oc0),
                                  fb.acc, 0, plan->img_cols, (int)or0, (int)oc0, 0, 1.0, s));
      }
      HIPCHK(ctx, launch_insert(fb.acc, plan->img_cols, 0, 0, (int)plan->img_rows, (int)plan->img_cols, image, dtype,
                                ldi, 0, 0, 1, fb.weight, s));
    }
  }
  return MFTX_OK;
}

int mftx_plan_execute(mftx_plan* plan, void* image, int32_t image_dtype, int64_t ldi, int32_t accumulate) {
  if (!plan || !plan->ctx) return MFTX_EHANDLE;
  mftx_ctx* ctx = plan->ctx;
  if (!image) return fail(ctx, MFTX_EINVAL, "null image");
  if (image_dtype != 1 && image_dtype != 3) return fail(ctx, MFTX_EINVAL, "image must be float32 or float64");
  if (ldi < plan->img_cols) return fail(ctx, MFTX_EINVAL, "image row stride too small");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (!plan->use_graph) return plan_launch(plan, image, image_dtype, ldi, accumulate);
  if (!plan->gexec || plan->g_image != image || plan->g_dtype != image_dtype || plan->g_ld != ldi ||
      plan->g_accum != accumulate) {
    if (plan->gexec) {
      (void)hipGraphExecDestroy(plan->gexec);
      plan->gexec = nullptr;
    }
    if (plan->graph) {
      (void)hipGraphDestroy(plan->graph);
      plan->graph = nullptr;
    }
    hipStream_t cap = nullptr;
    HIPCHK(ctx, hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipStream_t saved = ctx->stream;
    ctx->stream = cap;
    HIPCHK(ctx, hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    int rc = plan_launch(plan, image, image_dtype, ldi, accumulate);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &g);
    ctx->stream = saved;
    (void)hipStreamDestroy(cap);
    if (rc) return rc;
    if (e != hipSuccess) return hipfail(ctx, e, "hipStreamEndCapture");
    plan->graph = g;
    HIPCHK(ctx, hipGraphInstantiate(&plan->gexec, g, nullptr, nullptr, 0));
    plan->g_image = image;
    plan->g_dtype = image_dtype;
    plan->g_ld = ldi;
    plan->g_accum = accumulate;
  }
  HIPCHK(ctx, hipGraphLaunch(plan->gexec, ctx->stream));
  return MFTX_OK;
}

int mftx_plan_profile(mftx_plan* plan, void* image, int32_t image_dtype, int64_t ldi, int32_t repeats,
                      double* ms) {
  if (!plan || !plan->ctx) return MFTX_EHANDLE;
  mftx_ctx* ctx = plan->ctx;
  if (!image || !ms || repeats <= 0) return fail(ctx, MFTX_EINVAL, "bad profile request");
  if (image_dtype != 1 && image_dtype != 3) return fail(ctx, MFTX_EINVAL, "image must be float32 or float64");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  hipEvent_t ev[5];
  for (auto& e : ev) HIPCHK(ctx, hipEventCreate(&e));
  double acc[4] = {0, 0, 0, 0};
  int rc = MFTX_OK;
  for (int r = 0; r < repeats && rc == MFTX_OK; ++r) {
    hipError_t e = hipEventRecord(ev[0], s);
    for (auto& ph : plan->phasors)
      if (e == hipSuccess)
        e = launch_phasor(plan->prec, ph.tile.amp, ph.tile.amp_dtype, ph.tile.ld_amp, ph.tile.opd, ph.tile.opd_dtype,
                          ph.tile.ld_opd, (int)ph.tile.m, (int)ph.tile.n, plan->inv_wl + ph.wave_off, ph.nwave,
                          ph.dst, ph.wave_stride, ph.rows_pad, ph.kpad, s);
    if (e == hipSuccess) e = hipEventRecord(ev[1], s);
    if (e == hipSuccess) e = launch_gemm(plan->prec, EPI_STORE_PACKED_T, plan->works1, plan->nwork1, plan->probs1, plan->segs, s);
    if (e == hipSuccess) e = hipEventRecord(ev[2], s);
    if (e == hipSuccess) e = launch_gemm(plan->prec, plan->epi2, plan->works2, plan->nwork2, plan->probs2, plan->segs, s);
    if (e == hipSuccess) e = hipEventRecord(ev[3], s);
    if (e != hipSuccess) {
      rc = hipfail(ctx, e, "profile launch");
      break;
    }
    // the detector reduction (or the coherent fallback) via the regular path
    // would re-run every stage; time only its own launches
    if (plan->epi2 == EPI_INTENSITY) {
      for (auto& red : plan->reduces) {
        e = launch_reduce_partials(red.partial, red.gstride, red.ngroups, red.rows, red.cols, red.ldp, image,
                                   image_dtype, ldi, (int)plan->img_rows, (int)plan->img_cols, (int)red.r0,
                                   (int)red.c0, 0, s);
        if (e != hipSuccess) break;
      }
    }
    if (e == hipSuccess) e = hipEventRecord(ev[4], s);
    if (e == hipSuccess) e = hipEventSynchronize(ev[4]);
    if (e != hipSuccess) {
      rc = hipfail(ctx, e, "profile run");
      break;
    }
    for (int k = 0; k < 4; ++k) {
      float t = 0;
      hipEventElapsedTime(&t, ev[k], ev[k + 1]);
      acc[k] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  for (int k = 0; k < 4; ++k) ms[k] = acc[k] / repeats;
  return MFTX_OK;
}

int mftx_plan_set_graph(mftx_plan* plan, int32_t use_graph) {
  if (!plan) return MFTX_EHANDLE;
  plan->use_graph = use_graph ? 1 : 0;
  return MFTX_OK;
}

int mftx_plan_get_info(const mftx_plan* plan, mftx_plan_info* info) {
  if (!plan || !info) return MFTX_EHANDLE;
  *info = plan->info;
  return MFTX_OK;
}

int mftx_plan_destroy(mftx_plan* plan) {
  if (!plan) return MFTX_EHANDLE;
  if (plan->ctx) {
    (void)hipSetDevice(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);
  }
  if (plan->gexec) (void)hipGraphExecDestroy(plan->gexec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  if (plan->ws) (void)hipFree(plan->ws);
  delete plan;
  return MFTX_OK;
}

}  // extern "C"
