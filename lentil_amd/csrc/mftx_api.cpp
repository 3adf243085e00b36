// C-ABI of the mftx library (declared in include/mftx.h).
//
// An MFT F = c * E1 . f . E2 (lentil fourier.py:97) runs as two launches of
// the complex GEMM C = A . B (A packed data, B generated twiddles):
//   order B (row-first):    G   = f . E2        -> stored as G^T
//                           F^T = G^T . E1^T    -> F or |F|^2
//   order A (column-first): T^T = f^T . E1^T    -> stored as T
//                           F   = T . E2        -> F or |F|^2
// Order B costs m*n*N + M*m*N complex MACs, order A M*m*n + M*n*N; the
// cheaper one is used.  Several Fields summed coherently share one output
// grid, so their stage-2 GEMMs concatenate along K (one GEMM per coherent
// group, each K segment with its own twiddle coordinate origin).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mftx.h"
#include "mftx_internal.h"

using namespace mftx;

namespace {

constexpr int kRingSlots = 16;
constexpr size_t kSlotBytes = 512 * 1024;  // up to ~13k work items per launch

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }
double centre(int64_t k) { return (double)(k / 2); }  // floor(k/2), k >= 0 (fourier.py:117-120)

// MACs of the two contraction orders for an m x n -> M x N transform
double cost_row_first(double m, double n, double M, double N) { return m * n * N + M * m * N; }
double cost_col_first(double m, double n, double M, double N) { return M * m * n + M * n * N; }

Work empty_work() { return Work{0, 0, 0, 0, nullptr, 0, 0, 0}; }

// XCD-aware order: clusters of works that share A rows are assigned round-
// robin to the 8 XCDs; workgroup i of a launch runs on XCD i % 8, so a
// cluster's works land on one XCD and share its L2.
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
    for (int x = 0; x < 8; ++x) out.push_back(i < bins[x].size() ? bins[x][i] : empty_work());
  while (!out.empty() && out.back().pend == 0) out.pop_back();
  return out;
}

// Works covering rows_tiled C rows x `slots` slots for problems [pb, pe) with
// bn-column workgroups (bn / 2 slots each), one cluster per tile row.
void add_tiles(std::vector<std::vector<Work>>& clusters, int pb, int pe, int64_t rows_tiled, int64_t slots, int bn,
               float* partial = nullptr, long long ldp = 0, int prows = 0, int pcols = 0) {
  const int bs = slots_per_tile(bn);
  const int tr = (int)(rows_tiled / BM), tc = (int)((slots + bs - 1) / bs);
  for (int r = 0; r < tr; ++r) {
    std::vector<Work> row;
    for (int c = 0; c < tc; ++c) row.push_back(Work{pb, pe, r, c, partial, ldp, prows, pcols});
    clusters.push_back(row);
  }
}

// Output grid of one MFT stage with N columns, signed twiddle scale alpha and
// output shift: the slots and pairing centre the GEMM runs on (see Problem)
// and the fold of its A operand.  Outputs are paired about the PSF core
// j = N/2 + shift, snapped to the half-pixel grid, so the two outputs of a
// slot have comparable magnitude and neither inherits the rounding of a much
// brighter partner; when the core sits within a pixel of the grid centre the
// centre itself is used (fewest slots), and when pairing would save less than
// a quarter of the work (core far from the centre or outside the grid) every
// slot computes one column directly (c = 0).
struct OutGrid {
  int slots, cp, cm;
  double y0, fold;
};

OutGrid make_grid(double alpha, double shift, int N) {
  const int h = N / 2;
  const double core = h + shift;
  const double mid = 0.5 * (N - 1);
  double c;
  if (std::fabs(core - mid) <= 1.0) {
    c = mid;
  } else {
    c = std::nearbyint(2.0 * core) / 2.0;
    if (!(c >= 0.0 && c <= N - 1) || std::max(N - std::ceil(c), std::floor(c) + 1) > 0.75 * N) c = 0.0;
  }
  OutGrid g;
  g.cm = (int)std::floor(c);
  g.cp = (int)std::ceil(c);
  g.y0 = c - g.cm;
  g.slots = std::max(N - g.cp, g.cm + 1);
  g.fold = -alpha * (shift + h - c);
  return g;
}

void push_chunks(std::vector<double>& cx, int kpad, double x0) {
  for (int c = 0; c < kpad / BK; ++c) cx.push_back(x0 + (double)(c * BK));
}

}  // namespace

struct mftx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  void* aux = nullptr;  // intermediates of the detector operators (ws holds their MFT operands)
  size_t aux_bytes = 0;
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
  int bn = BN_DEFAULT;  // C columns per GEMM workgroup
  int64_t img_rows = 0, img_cols = 0;
  void* ws = nullptr;  // workspace: operands, intermediates, partial images, descriptors
  Work* works1 = nullptr;
  Work* works2 = nullptr;
  Problem* probs1 = nullptr;
  Problem* probs2 = nullptr;
  double* chunk_x = nullptr;
  PhasorWave* waves = nullptr;  // per phasor wavelength slot: 1/λ and the stage-1 fold
  int nwork1 = 0, nwork2 = 0;
  // stage 2 of the classes whose coherent group has several output grids:
  // their complex fields, summed by the fallbacks below (every other class
  // runs the intensity launch above)
  Work* works2c = nullptr;
  Problem* probs2c = nullptr;
  int nwork2c = 0;
  PhasorOp* phasor_ops = nullptr;  // device descriptors of the grouped phasor launch
  int n_phasor_ops = 0;
  long long phasor_threads = 0;
  struct Reduce {
    float* partial;
    long long gstride, ldp;
    int ngroups, rows, cols;
    long long r0, c0;
  };
  std::vector<Reduce> reduces;
  // coherent groups whose Fields overlap only partially: complex outputs
  // summed into an image-sized accumulator, then |.|^2 (field.py:331-346)
  struct Add {
    const float2* F;
    int M, N;
    int64_t r0, c0;
  };
  struct Fallback {
    float2* acc;            // the plan's one accumulator, shared by the groups in launch order
    int64_t br0, bc0;       // the group's union box on the image (clipped)
    int64_t brows, bcols;
    std::vector<Add> adds;
    double weight;
  };
  std::vector<Fallback> fallbacks;
  mftx_plan_info info = {};
  int use_graph = 0;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  void* g_image = nullptr;
  int g_dtype = -1, g_accum = -1;
  int64_t g_ld = -1;
  // complex fields of every class from the stage-1 operands already in the
  // workspace (mftx_plan_set_fields / mftx_plan_execute_fields)
  Work* works_f = nullptr;
  Problem* probs_f = nullptr;
  int nwork_f = 0;
  std::vector<Problem> pf;       // host copy, pointers filled by set_fields
  std::vector<int> class_first;  // first member (problem index) of each class
  int fields_ready = 0;
};

static int fail(mftx_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int mftx::set_error(mftx_ctx* ctx, int code, const char* msg) { return fail(ctx, code, msg); }
int mftx::context_device(const mftx_ctx* ctx) { return ctx->device; }

static int hipfail(mftx_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, MFTX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Host-side exceptions (allocation failures of the plan builders' containers)
// end at the C boundary as status codes.
template <class F> static int guarded(mftx_ctx* ctx, F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(ctx, MFTX_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(ctx, MFTX_EDEVICE, e.what());
  } catch (...) {
    return fail(ctx, MFTX_EDEVICE, "unexpected host exception");
  }
}

// Device allocation whose failure is an ordinary status: the runtime's
// pending error is cleared, so the next launch check (hipGetLastError after a
// kernel launch) does not report this allocation's failure as its own.
static bool try_malloc(void** p, size_t bytes) {
  if (hipMalloc(p, bytes) == hipSuccess) return true;
  *p = nullptr;
  (void)hipGetLastError();
  return false;
}

#define HIPCHK(ctx, expr)                                   \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return hipfail((ctx), e_, #expr); \
  } while (0)

// Grow the context workspace to at least `bytes`, by at least half its old
// size, so a run of slowly growing calls reallocates (and synchronises) only
// a few times.
static int ensure_ws(mftx_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->ws_bytes) return MFTX_OK;
  const size_t old = ctx->ws_bytes;
  if (ctx->ws) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  const size_t nb = std::max(bytes, old + old / 2);
  if (!try_malloc(&ctx->ws, nb)) {
    ctx->ws = nullptr;
    return fail(ctx, MFTX_ENOMEM, "workspace allocation of " + std::to_string(nb) + " bytes failed");
  }
  ctx->ws_bytes = nb;
  return MFTX_OK;
}

// Take the next ring slot (waiting until its previous upload has been
// consumed); *host is its pinned staging area, *dev its device address.
static int reserve_desc(mftx_ctx* ctx, int* slot, char** host, char** dev) {
  const int s = ctx->ring_next;
  ctx->ring_next = (s + 1) % kRingSlots;
  if (ctx->ring_used[s]) HIPCHK(ctx, hipEventSynchronize(ctx->ring_ev[s]));
  ctx->ring_used[s] = false;
  *slot = s;
  *host = ctx->ring_host + (size_t)s * kSlotBytes;
  *dev = ctx->ring_dev + (size_t)s * kSlotBytes;
  return MFTX_OK;
}

// Queue the upload of the first `bytes` of a reserved slot on the context stream.
static int commit_desc(mftx_ctx* ctx, int s, size_t bytes) {
  char* h = ctx->ring_host + (size_t)s * kSlotBytes;
  char* d = ctx->ring_dev + (size_t)s * kSlotBytes;
  HIPCHK(ctx, hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ring_ev[s], ctx->stream));
  ctx->ring_used[s] = true;
  return MFTX_OK;
}

static int ensure_aux(mftx_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->aux_bytes) return MFTX_OK;
  if (ctx->aux) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(ctx->aux));
    ctx->aux = nullptr;
    ctx->aux_bytes = 0;
  }
  if (!try_malloc(&ctx->aux, bytes)) {
    ctx->aux = nullptr;
    return fail(ctx, MFTX_ENOMEM, "workspace allocation of " + std::to_string(bytes) + " bytes failed");
  }
  ctx->aux_bytes = bytes;
  return MFTX_OK;
}

// Stage `bytes` of descriptors into a ring slot and queue the upload on the
// context stream; returns the device address of the slot.
static int upload_desc(mftx_ctx* ctx, const void* src, size_t bytes, void** dev) {
  if (bytes > kSlotBytes) return fail(ctx, MFTX_EINVAL, "descriptor block too large (output too big for one call)");
  int s;
  char *h, *d;
  int rc = reserve_desc(ctx, &s, &h, &d);
  if (rc) return rc;
  std::memcpy(h, src, bytes);
  rc = commit_desc(ctx, s, bytes);
  if (rc) return rc;
  *dev = d;
  return MFTX_OK;
}

// One GEMM launch of a single problem covering rows_tiled x p.cols of C with
// bn-column workgroups (p.kpad a multiple of k_granule(bn)).
static int run_single(mftx_ctx* ctx, int prec, int epi, int bn, Problem p, double x0, int64_t rows_tiled) {
  std::vector<std::vector<Work>> clusters;
  add_tiles(clusters, 0, 1, rows_tiled, p.cols, bn);
  const std::vector<Work> works = xcd_order(clusters);
  std::vector<double> cx;
  push_chunks(cx, p.kpad, x0);
  p.chunk_begin = 0;
  const size_t off_cx = round_up(sizeof(Problem), 16);
  const size_t off_w = round_up(off_cx + sizeof(double) * cx.size(), 16);
  std::vector<char> host(off_w + sizeof(Work) * works.size());
  std::memcpy(host.data(), &p, sizeof(Problem));
  std::memcpy(host.data() + off_cx, cx.data(), sizeof(double) * cx.size());
  std::memcpy(host.data() + off_w, works.data(), sizeof(Work) * works.size());
  void* dev = nullptr;
  int rc = upload_desc(ctx, host.data(), host.size(), &dev);
  if (rc) return rc;
  char* d = (char*)dev;
  HIPCHK(ctx, launch_gemm(prec, epi, bn, (const Work*)(d + off_w), (int)works.size(), (const Problem*)d,
                          (const double*)(d + off_cx), ctx->stream));
  return MFTX_OK;
}

// Problem ranges [b, e) for work items: runs of g consecutive problems with
// one tiling (rows tiled, output slots) and K length share a work item, so
// the GEMM's staging pipeline runs across their K loops.  g is at most 8 and
// keeps at least one work item per CU.  Grouping is used only when every run
// has the same length g: a launch that mixed long and short work items would
// finish on its long tail.
static std::vector<std::pair<int, int>> group_runs(const std::vector<Problem>& ps, const std::vector<int64_t>& rows,
                                                   int bn) {
  const int n = (int)ps.size();
  int64_t single = 0;
  for (int i = 0; i < n; ++i) single += (rows[i] / BM) * ((ps[i].cols + slots_per_tile(bn) - 1) / slots_per_tile(bn));
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(8, single / 256));
  std::vector<std::pair<int, int>> runs;
  bool uniform = true;
  for (int i = 0; i < n;) {
    int e = i + 1;
    while (e < n && e - i < g && rows[e] == rows[i] && ps[e].cols == ps[i].cols && ps[e].kpad == ps[i].kpad &&
           ps[e].rows == ps[i].rows)
      ++e;
    uniform = uniform && e - i == g;
    runs.emplace_back(i, e);
    i = e;
  }
  if (!uniform) {
    runs.clear();
    for (int i = 0; i < n; ++i) runs.emplace_back(i, i + 1);
  }
  return runs;
}

// Geometry of one entry of mftx_dft2_many (the quantities mftx_dft2 derives).
namespace {
struct ManyGeo {
  bool col_first;
  int64_t r1, c1, rows1, rows2;
  double alpha1, alpha2, x01, x02;
  OutGrid g1, g2;
  int kpad1, kpad2;
  size_t off1, off2;  // workspace offsets of the packed stage-1 and stage-2 operands
};
}  // namespace

// The launches of a list of MFTs: all entries in one pack launch per input
// kind and one GEMM launch per stage.  many_build derives the geometry and the
// descriptors (workspace and descriptor addresses as offsets); many_place
// rebases them onto a workspace and a device descriptor block and writes the
// host copy of that block; many_launch issues the launches.
struct ManyBatch {
  int bn = 64;
  size_t ws = 0, bytes = 0;  // workspace and descriptor block sizes
  size_t o_wv = 0, o_ph = 0, o_pk = 0, o_p1 = 0, o_p2 = 0, o_cx = 0, o_w1 = 0, o_w2 = 0;
  std::vector<ManyGeo> geo;
  std::vector<PhasorWave> waves;
  std::vector<PhasorOp> phops;
  std::vector<PackOp> packs;
  std::vector<int> wave_of;  // phasor wave of entry i, or -(pack index + 1)
  std::vector<Problem> p1, p2;
  std::vector<double> cx;
  std::vector<Work> w1, w2;
  long long ph_total = 0, pk_total = 0;
};

static void many_build(const mftx_mft* mfts, int32_t count, int prec, ManyBatch& mb) {
  std::vector<ManyGeo>& geo = mb.geo;
  geo.assign(count, ManyGeo{});
  int64_t wide1 = 0, wide2 = 0;
  for (int32_t i = 0; i < count; ++i) {
    const mftx_mft& q = mfts[i];
    ManyGeo& g = geo[i];
    g.col_first = cost_col_first(q.m, q.n, q.M, q.N) < cost_row_first(q.m, q.n, q.M, q.N);
    g.r1 = g.col_first ? q.n : q.m;
    g.c1 = g.col_first ? q.M : q.N;
    const int64_t c2 = g.col_first ? q.N : q.M;
    g.rows1 = round_up(g.r1, BM);
    g.rows2 = round_up(g.c1, BM);
    const double sgn = q.conj ? 1.0 : -1.0;
    g.alpha1 = sgn * (g.col_first ? q.alpha_r : q.alpha_c);
    g.alpha2 = sgn * (g.col_first ? q.alpha_c : q.alpha_r);
    g.g1 = make_grid(g.alpha1, g.col_first ? q.shift_r : q.shift_c, (int)g.c1);
    g.g2 = make_grid(g.alpha2, g.col_first ? q.shift_c : q.shift_r, (int)c2);
    const double xr0 = -centre(q.m) + q.offset_r, xc0 = -centre(q.n) + q.offset_c;
    g.x01 = g.col_first ? xr0 : xc0;
    g.x02 = g.col_first ? xc0 : xr0;
    wide1 += (g.rows1 / BM) * ((g.g1.slots + 63) / 64);
    wide2 += (g.rows2 / BM) * ((g.g2.slots + 63) / 64);
  }
  // 128-column tiles when the larger stage still gives every CU a workgroup
  const int bn = mb.bn = std::max(wide1, wide2) >= 256 ? 128 : 64;
  const int kg = k_granule(bn);
  size_t& ws = mb.ws;
  ws = 0;
  for (int32_t i = 0; i < count; ++i) {
    const mftx_mft& q = mfts[i];
    ManyGeo& g = geo[i];
    g.kpad1 = (int)round_up(g.col_first ? q.m : q.n, kg);
    g.kpad2 = (int)round_up(g.r1, kg);
    g.off1 = ws;
    ws += round_up((int64_t)packed_bytes(prec, g.rows1, g.kpad1), 256);
  }
  for (int32_t i = 0; i < count; ++i) {
    geo[i].off2 = ws;
    ws += round_up((int64_t)packed_bytes(prec, geo[i].rows2, geo[i].kpad2), 256);
  }

  // ---- descriptors -------------------------------------------------------------
  std::vector<PhasorWave>& waves = mb.waves;
  std::vector<PhasorOp>& phops = mb.phops;
  std::vector<PackOp>& packs = mb.packs;
  long long& ph_total = mb.ph_total;
  long long& pk_total = mb.pk_total;
  std::vector<int>& wave_of = mb.wave_of;
  for (int32_t i = 0; i < count; ++i) {
    const mftx_mft& q = mfts[i];
    const ManyGeo& g = geo[i];
    if (q.f) {
      PackOp op = {};
      op.src = q.f;
      op.dtype = q.f_dtype;
      op.trans = g.col_first ? 1 : 0;
      op.m = q.m;
      op.n = q.n;
      op.lds = q.ldf;
      op.fold = g.g1.fold;
      op.x0 = g.x01;
      op.fold2 = g.g2.fold;
      op.x02 = g.x02;
      op.dst = nullptr;  // rebased below
      op.rows_pad = g.rows1;
      op.kpad = g.kpad1;
      op.begin = group_begin(pk_total);
      pk_total = op.begin + g.rows1 * (g.kpad1 / 2);  // one thread per (row, pair of K values)
      packs.push_back(op);
      wave_of.push_back(-(int)packs.size());  // -(pack index + 1)
      continue;
    }
    // consecutive phasor entries of one pupil and geometry share an op
    // (their operands are one fixed stride apart in the workspace)
    const size_t stride = round_up((int64_t)packed_bytes(prec, g.rows1, g.kpad1), 256);
    bool joined = false;
    if (i > 0 && !mfts[i - 1].f && !phops.empty()) {
      const mftx_mft& p = mfts[i - 1];
      const ManyGeo& h = geo[i - 1];
      PhasorOp& last = phops.back();
      joined = p.amp == q.amp && p.opd == q.opd && p.amp_dtype == q.amp_dtype && p.opd_dtype == q.opd_dtype &&
               p.ld_amp == q.ld_amp && p.ld_opd == q.ld_opd && p.m == q.m && p.n == q.n &&
               h.col_first == g.col_first && h.kpad1 == g.kpad1 && h.rows1 == g.rows1 &&
               (long long)stride == last.wave_stride;
    }
    if (!joined) {
      if (!phops.empty())
        ph_total = phops.back().begin + phasor_threads(phops.back().rows_pad, phops.back().kpad, phops.back().nwave);
      PhasorOp op = {};
      op.amp = q.amp;
      op.opd = q.opd;
      op.amp_dtype = q.amp_dtype;
      op.opd_dtype = q.opd_dtype;
      op.ld_amp = q.ld_amp;
      op.ld_opd = q.ld_opd;
      op.m = (int)q.m;
      op.n = (int)q.n;
      op.trans = g.col_first ? 1 : 0;
      op.nwave = 0;
      op.waves = (const PhasorWave*)(uintptr_t)waves.size();  // index, rebased below
      op.dst = (void*)(uintptr_t)g.off1;                       // offset, rebased below
      op.wave_stride = (long long)stride;
      op.rows_pad = g.rows1;
      op.kpad = g.kpad1;
      op.begin = group_begin(ph_total);
      phops.push_back(op);
    }
    phops.back().nwave += 1;
    waves.push_back(PhasorWave{q.inv_wl, g.g1.fold, g.x01, g.g2.fold, g.x02});
    wave_of.push_back((int)waves.size() - 1);
  }
  if (!phops.empty())
    ph_total = phops.back().begin + phasor_threads(phops.back().rows_pad, phops.back().kpad, phops.back().nwave);

  std::vector<Problem>& p1 = mb.p1;
  std::vector<Problem>& p2 = mb.p2;
  p1.assign(count, Problem{});
  p2.assign(count, Problem{});
  std::vector<double>& cx = mb.cx;
  std::vector<std::vector<Work>> clus1, clus2;
  for (int32_t i = 0; i < count; ++i) {
    const mftx_mft& q = mfts[i];
    const ManyGeo& g = geo[i];
    Problem& a = p1[i];
    a.lda = packed_lda(prec, g.kpad1);
    a.alpha = g.alpha1;
    a.y0 = g.g1.y0;
    a.rows = (int)g.r1;
    a.cols = g.g1.slots;
    a.ncols = (int)g.c1;
    a.cp = g.g1.cp;
    a.cm = g.g1.cm;
    a.kpad = g.kpad1;
    a.chunk_begin = (int)cx.size();
    push_chunks(cx, g.kpad1, g.x01);
    a.scale = 1.0f;
    a.limit = g.kpad2;
    a.ldd = packed_lda(prec, g.kpad2);

    Problem& b = p2[i];
    b.lda = packed_lda(prec, g.kpad2);
    b.alpha = g.alpha2;
    b.y0 = g.g2.y0;
    b.rows = (int)g.c1;
    b.cols = g.g2.slots;
    b.ncols = (int)(g.col_first ? q.N : q.M);
    b.cp = g.g2.cp;
    b.cm = g.g2.cm;
    b.kpad = g.kpad2;
    b.chunk_begin = (int)cx.size();
    push_chunks(cx, g.kpad2, g.x02);
    b.scale = (float)q.scale;
    b.dst = q.F;
    b.ldd = q.ldF;
    b.trans = g.col_first ? 0 : 1;
    b.wmap = q.w;
    b.ldw = q.ldw;
    b.wdtype = q.w_dtype;
  }
  // works: runs of consecutive entries with one geometry share a work item
  auto group_works = [&](std::vector<std::vector<Work>>& clus, const std::vector<Problem>& ps, bool first) {
    std::vector<int64_t> rows(count);
    for (int32_t i = 0; i < count; ++i) rows[i] = first ? geo[i].rows1 : geo[i].rows2;
    for (const auto& r : group_runs(ps, rows, bn)) add_tiles(clus, r.first, r.second, rows[r.first], ps[r.first].cols, bn);
  };
  group_works(clus1, p1, true);
  group_works(clus2, p2, false);
  mb.w1 = xcd_order(clus1);
  mb.w2 = xcd_order(clus2);

  // one descriptor block: waves, phasor ops, pack ops, problems, chunk origins, works
  size_t off = 0;
  auto take = [&](size_t b) { const size_t o = off; off = round_up((int64_t)(off + b), 16); return o; };
  mb.o_wv = take(sizeof(PhasorWave) * waves.size());
  mb.o_ph = take(sizeof(PhasorOp) * phops.size());
  mb.o_pk = take(sizeof(PackOp) * packs.size());
  mb.o_p1 = take(sizeof(Problem) * p1.size());
  mb.o_p2 = take(sizeof(Problem) * p2.size());
  mb.o_cx = take(sizeof(double) * cx.size());
  mb.o_w1 = take(sizeof(Work) * mb.w1.size());
  mb.o_w2 = take(sizeof(Work) * mb.w2.size());
  mb.bytes = off;
}

// Rebase the batch's descriptors onto workspace `base` and the device
// descriptor block `d`, and write the block's contents to host memory `h`.
static void many_place(ManyBatch& mb, char* base, char* d, char* h) {
  for (auto& op : mb.phops) {
    op.waves = (const PhasorWave*)(d + mb.o_wv) + (uintptr_t)op.waves;
    op.dst = base + (uintptr_t)op.dst;
  }
  for (size_t i = 0; i < mb.geo.size(); ++i) {
    mb.p1[i].a = base + mb.geo[i].off1;
    mb.p1[i].dst = base + mb.geo[i].off2;
    mb.p2[i].a = base + mb.geo[i].off2;
    if (mb.wave_of[i] < 0) mb.packs[-mb.wave_of[i] - 1].dst = base + mb.geo[i].off1;
  }
  std::memcpy(h + mb.o_wv, mb.waves.data(), sizeof(PhasorWave) * mb.waves.size());
  std::memcpy(h + mb.o_ph, mb.phops.data(), sizeof(PhasorOp) * mb.phops.size());
  std::memcpy(h + mb.o_pk, mb.packs.data(), sizeof(PackOp) * mb.packs.size());
  std::memcpy(h + mb.o_p1, mb.p1.data(), sizeof(Problem) * mb.p1.size());
  std::memcpy(h + mb.o_p2, mb.p2.data(), sizeof(Problem) * mb.p2.size());
  std::memcpy(h + mb.o_cx, mb.cx.data(), sizeof(double) * mb.cx.size());
  std::memcpy(h + mb.o_w1, mb.w1.data(), sizeof(Work) * mb.w1.size());
  std::memcpy(h + mb.o_w2, mb.w2.data(), sizeof(Work) * mb.w2.size());
}

// The launches of a placed batch (descriptors at device address d).
static hipError_t many_launch(const ManyBatch& mb, int prec, const char* d, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (!mb.phops.empty())
    e = launch_phasor_group(prec, (const PhasorOp*)(d + mb.o_ph), (int)mb.phops.size(), mb.ph_total, s);
  if (e == hipSuccess && !mb.packs.empty())
    e = launch_pack_group(prec, (const PackOp*)(d + mb.o_pk), (int)mb.packs.size(), mb.pk_total, s);
  if (e == hipSuccess)
    e = launch_gemm(prec, EPI_STORE_PACKED_T, mb.bn, (const Work*)(d + mb.o_w1), (int)mb.w1.size(),
                    (const Problem*)(d + mb.o_p1), (const double*)(d + mb.o_cx), s);
  if (e == hipSuccess)
    e = launch_gemm(prec, EPI_STORE_CPLX, mb.bn, (const Work*)(d + mb.o_w2), (int)mb.w2.size(),
                    (const Problem*)(d + mb.o_p2), (const double*)(d + mb.o_cx), s);
  return e;
}

// All entries [0, count) on the context stream with the context workspace and
// one descriptor ring slot; a batch whose descriptors exceed a slot is split.
static int dft2_many_run(mftx_ctx* ctx, const mftx_mft* mfts, int32_t count, int prec) {
  ManyBatch mb;
  many_build(mfts, count, prec, mb);
  if (mb.bytes > kSlotBytes) {
    if (count == 1) return fail(ctx, MFTX_EINVAL, "descriptor block too large (output too big for one call)");
    const int32_t h = count / 2;
    const int rc = dft2_many_run(ctx, mfts, h, prec);
    return rc ? rc : dft2_many_run(ctx, mfts + h, count - h, prec);
  }
  int rc = ensure_ws(ctx, mb.ws);
  if (rc) return rc;
  // the slot is reserved first: the descriptors hold device addresses inside it
  int slot;
  char *h, *d;
  rc = reserve_desc(ctx, &slot, &h, &d);
  if (rc) return rc;
  many_place(mb, (char*)ctx->ws, d, h);
  rc = commit_desc(ctx, slot, mb.bytes);
  if (rc) return rc;
  HIPCHK(ctx, many_launch(mb, prec, d, ctx->stream));
  return MFTX_OK;
}

// A persistent list of MFTs (mftx_mft_plan_*): plan-owned workspace and
// descriptors, built once.
struct mftx_mft_plan {
  mftx_ctx* ctx = nullptr;
  int prec = 0;
  ManyBatch mb;
  void* mem = nullptr;   // operand workspace, then the descriptor block
  char* desc = nullptr;
  ~mftx_mft_plan() {
    if (mem) (void)hipFree(mem);
  }
};

// Argument checks shared by mftx_dft2_many and mftx_mft_plan_create.
static int check_mfts(mftx_ctx* ctx, const mftx_mft* mfts, int32_t count, int32_t precision) {
  if (count < 0 || (count > 0 && !mfts)) return fail(ctx, MFTX_EINVAL, "bad MFT list");
  if (precision != MFTX_PREC_FP32 && precision != MFTX_PREC_BF16X3) return fail(ctx, MFTX_EINVAL, "unknown precision");
  for (int32_t i = 0; i < count; ++i) {
    const mftx_mft& q = mfts[i];
    const std::string at = "MFT " + std::to_string(i) + ": ";
    if (!q.F) return fail(ctx, MFTX_EINVAL, at + "null output pointer");
    if (q.m <= 0 || q.n <= 0 || q.M <= 0 || q.N <= 0) return fail(ctx, MFTX_EINVAL, at + "shapes must be positive");
    if (q.m > (1 << 20) || q.n > (1 << 20) || q.M > (1 << 20) || q.N > (1 << 20))
      return fail(ctx, MFTX_EINVAL, at + "dimension too large");
    if (q.ldF < q.N) return fail(ctx, MFTX_EINVAL, at + "output row stride smaller than row length");
    if (q.f) {
      if (q.f_dtype < 0 || q.f_dtype > 3) return fail(ctx, MFTX_EINVAL, at + "unknown input dtype");
      if (q.ldf < q.n) return fail(ctx, MFTX_EINVAL, at + "input row stride smaller than row length");
    } else {
      if (!q.amp || !q.opd) return fail(ctx, MFTX_EINVAL, at + "phasor input needs amp and opd");
      if ((q.amp_dtype != 1 && q.amp_dtype != 3) || (q.opd_dtype != 1 && q.opd_dtype != 3))
        return fail(ctx, MFTX_EINVAL, at + "amp/opd must be float32 or float64");
      if (q.ld_amp < q.n || q.ld_opd < q.n) return fail(ctx, MFTX_EINVAL, at + "phasor row stride too small");
      if (!(q.inv_wl > 0) || !std::isfinite(q.inv_wl)) return fail(ctx, MFTX_EINVAL, at + "inv_wl must be positive");
    }
    if (q.w) {
      if (q.w_dtype != 1 && q.w_dtype != 3) return fail(ctx, MFTX_EINVAL, at + "weight must be float32 or float64");
      if (q.ldw < q.N) return fail(ctx, MFTX_EINVAL, at + "weight row stride smaller than row length");
    }
  }
  return MFTX_OK;
}

extern "C" {

int mftx_version(void) { return 10100; }

int mftx_create(int device, mftx_ctx** out) {
  return guarded(nullptr, [&]() -> int {
  if (!out) return MFTX_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MFTX_EDEVICE;
  if (device < 0 || device >= n) return MFTX_EINVAL;
  mftx_ctx* ctx = new mftx_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipHostMalloc((void**)&ctx->ring_host, kRingSlots * kSlotBytes, hipHostMallocDefault) != hipSuccess ||
      !try_malloc((void**)&ctx->ring_dev, kRingSlots * kSlotBytes)) {
    mftx_destroy(ctx);
    return MFTX_ENOMEM;
  }
  for (int i = 0; i < kRingSlots; ++i) {
    if (hipEventCreateWithFlags(&ctx->ring_ev[i], hipEventDisableTiming) != hipSuccess) {
      mftx_destroy(ctx);
      return MFTX_EDEVICE;
    }
  }
  // the no-allocation entry point's scratch address on this device
  void* scratch = nullptr;
  if (scratch_address(device, &scratch) != hipSuccess) {
    mftx_destroy(ctx);
    return MFTX_EDEVICE;
  }

  *out = ctx;
  return MFTX_OK;
  });
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
  if (ctx->aux) (void)hipFree(ctx->aux);
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

int mftx_dft2(mftx_ctx* ctx, const void* f, int f_dtype, int64_t m, int64_t n, int64_t ldf, void* F, int64_t M,
              int64_t N, int64_t ldF, double alpha_r, double alpha_c, double shift_r, double shift_c,
              double offset_r, double offset_c, double scale, int conj, int precision) {
  return mftx::dft2_on_workspace(ctx, f, f_dtype, m, n, ldf, F, M, N, ldF, alpha_r, alpha_c, shift_r, shift_c,
                                 offset_r, offset_c, scale, conj, precision, nullptr, 0);
}

}  // extern "C"

// mftx_dft2 on the caller's workspace `ws` of `cap` bytes, or on the
// context's growable workspace when ws is null.  The operands are fully
// written before they are read (the packed input with its zero padding,
// then stage 1's output for every K the second stage reads), so the
// workspace needs no clearing; its padding rows beyond the output only feed
// output rows that are never stored.
int mftx::dft2_on_workspace(mftx_ctx* ctx, const void* f, int f_dtype, int64_t m, int64_t n, int64_t ldf, void* F,
                            int64_t M, int64_t N, int64_t ldF, double alpha_r, double alpha_c, double shift_r,
                            double shift_c, double offset_r, double offset_c, double scale, int conj, int precision,
                            void* ws, size_t cap) {
  if (!ctx) return MFTX_EHANDLE;
  if (!f || !F) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if (m <= 0 || n <= 0 || M <= 0 || N <= 0) return fail(ctx, MFTX_EINVAL, "shapes must be positive");
  if (ldf < n || ldF < N) return fail(ctx, MFTX_EINVAL, "row stride smaller than row length");
  if (f_dtype < 0 || f_dtype > 3) return fail(ctx, MFTX_EINVAL, "unknown input dtype");
  if (precision != MFTX_PREC_FP32 && precision != MFTX_PREC_BF16X3) return fail(ctx, MFTX_EINVAL, "unknown precision");
  if (m > (1 << 20) || n > (1 << 20) || M > (1 << 20) || N > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "dimension too large");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int prec = precision;
  const bool col_first = cost_col_first(m, n, M, N) < cost_row_first(m, n, M, N);
  // stage-1 operand: f (row-first) or f^T (column-first)
  const int64_t r1 = col_first ? n : m, k1 = col_first ? m : n;  // rows and K of stage 1
  const int64_t c1 = col_first ? M : N;                           // stage-1 output columns
  const int64_t c2 = col_first ? N : M;                           // stage-2 output columns
  const int64_t rows1 = round_up(r1, BM);
  const int64_t rows2 = round_up(c1, BM);
  const double sgn = conj ? 1.0 : -1.0;
  const double alpha1 = sgn * (col_first ? alpha_r : alpha_c), alpha2 = sgn * (col_first ? alpha_c : alpha_r);
  const OutGrid g1 = make_grid(alpha1, col_first ? shift_r : shift_c, (int)c1);
  const OutGrid g2 = make_grid(alpha2, col_first ? shift_c : shift_r, (int)c2);
  const int s1 = g1.slots, s2 = g2.slots;
  // 128-column tiles when the larger stage still gives every CU a workgroup
  const int64_t wide = std::max((rows1 / BM) * ((s1 + 63) / 64), (rows2 / BM) * ((s2 + 63) / 64));
  const int bn = wide >= 256 ? 128 : 64;
  const int kg = k_granule(bn);
  const int kpad1 = (int)round_up(k1, kg);
  const int kpad2 = (int)round_up(r1, kg);
  const size_t b1 = packed_bytes(prec, rows1, kpad1);
  const size_t b2 = packed_bytes(prec, rows2, kpad2);
  const size_t need = (size_t)round_up((int64_t)b1, 256) + b2;
  int rc = MFTX_OK;
  if (!ws) {
    rc = ensure_ws(ctx, need);
    if (rc) return rc;
    ws = ctx->ws;
  } else if (need > cap) {
    return fail(ctx, MFTX_EINVAL, "transform needs " + std::to_string(need) + " bytes of workspace, more than the " +
                                      std::to_string(cap) + " available");
  }
  char* a1 = (char*)ws;
  char* a2 = a1 + round_up((int64_t)b1, 256);
  // twiddle coordinates: rows x = R + or against U - sr; columns x = S + oc
  // against V - sc; each stage's output shift is folded into its A operand
  const double xr0 = -centre(m) + offset_r, xc0 = -centre(n) + offset_c;
  const double fold1 = g1.fold, fold2 = g2.fold;
  const double x01 = col_first ? xr0 : xc0, x02 = col_first ? xc0 : xr0;
  // the operand carries both folds: stage 1's along K, stage 2's along its
  // rows (stage 1's output rows, stage 2's K)
  HIPCHK(ctx, launch_pack(prec, f, f_dtype, m, n, ldf, col_first ? 1 : 0, fold1, x01, fold2, x02, a1, rows1,
                          kpad1, ctx->stream));

  Problem p1 = {};
  p1.a = a1;
  p1.lda = packed_lda(prec, kpad1);
  p1.alpha = alpha1;
  p1.y0 = g1.y0;
  p1.rows = (int)r1;
  p1.cols = s1;
  p1.ncols = (int)c1;
  p1.cp = g1.cp;
  p1.cm = g1.cm;
  p1.kpad = kpad1;
  p1.scale = 1.0f;
  p1.limit = kpad2;
  p1.dst = a2;
  p1.ldd = packed_lda(prec, kpad2);
  rc = run_single(ctx, prec, EPI_STORE_PACKED_T, bn, p1, x01, rows1);
  if (rc) return rc;

  Problem p2 = {};
  p2.a = a2;
  p2.lda = packed_lda(prec, kpad2);
  p2.alpha = alpha2;
  p2.y0 = g2.y0;
  p2.rows = (int)c1;
  p2.cols = s2;
  p2.ncols = (int)c2;
  p2.cp = g2.cp;
  p2.cm = g2.cm;
  p2.kpad = kpad2;
  p2.scale = (float)scale;
  p2.dst = F;
  p2.ldd = ldF;
  p2.trans = col_first ? 0 : 1;  // C is F (column-first) or F^T (row-first)
  return run_single(ctx, prec, EPI_STORE_CPLX, bn, p2, x02, rows2);
}

extern "C" {

int mftx_phasor(mftx_ctx* ctx, const void* amp, int amp_dtype, int64_t ld_amp, const void* opd, int opd_dtype,
                int64_t ld_opd, int64_t m, int64_t n, double wavelength, void* out) {
  if (!ctx) return MFTX_EHANDLE;
  if (!amp || !opd || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if ((amp_dtype != 1 && amp_dtype != 3) || (opd_dtype != 1 && opd_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "amp/opd must be float32 or float64");
  if (m <= 0 || n <= 0 || ld_amp < n || ld_opd < n || m > (1 << 20) || n > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "bad phasor shape");
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
  if (h > (1 << 20) || w > (1 << 20) || fr0 > (1 << 30) || fc0 > (1 << 30) || or0 > (1 << 30) || oc0 > (1 << 30))
    return fail(ctx, MFTX_EINVAL, "slice too large");
  if (h > 0 && w > 0 && (ldf < fc0 + w || ldo < oc0 + w))
    return fail(ctx, MFTX_EINVAL, "slice wider than its row stride");
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
  if (h < 0 || w < 0 || lda < w || ldb < w || ldo < w || h > (1 << 20) || w > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "bad multiply shape");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_multiply(a, lda, b, ldb, out, ldo, (int)h, (int)w, ctx->stream));
  return MFTX_OK;
}

int mftx_weight_field(mftx_ctx* ctx, const void* F, int64_t ldF, const void* w, int w_dtype, int64_t ldw,
                      double scale, void* out, int64_t ldo, int64_t h, int64_t wd) {
  if (!ctx) return MFTX_EHANDLE;
  if (!F || !w || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if (w_dtype != 1 && w_dtype != 3) return fail(ctx, MFTX_EINVAL, "weight must be float32 or float64");
  if (h < 0 || wd < 0 || ldF < wd || ldw < wd || ldo < wd || h > (1 << 20) || wd > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "bad weight shape");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_weight_field(F, ldF, w, w_dtype, ldw, scale, out, ldo, (int)h, (int)wd, ctx->stream));
  return MFTX_OK;
}

int mftx_phase_grad(mftx_ctx* ctx, const void* amp, int amp_dtype, int64_t ld_amp, const void* opd, int opd_dtype,
                    int64_t ld_opd, int64_t m, int64_t n, double wavelength, const void* D, int64_t ldD,
                    double* grad_opd, int64_t ld_gopd, double* grad_amp, int64_t ld_gamp) {
  if (!ctx) return MFTX_EHANDLE;
  if (!amp || !opd || !D || !grad_opd) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if ((amp_dtype != 1 && amp_dtype != 3) || (opd_dtype != 1 && opd_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "amp/opd must be float32 or float64");
  if (m <= 0 || n <= 0 || ld_amp < n || ld_opd < n || ldD < n || ld_gopd < n || (grad_amp && ld_gamp < n) ||
      m > (1 << 20) || n > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "bad gradient shape");
  if (!(wavelength > 0)) return fail(ctx, MFTX_EINVAL, "wavelength must be positive");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_phase_grad(amp, amp_dtype, ld_amp, opd, opd_dtype, ld_opd, (int)m, (int)n, nullptr,
                                1.0 / wavelength, 1, D, ldD, 0, grad_opd, ld_gopd, grad_amp, ld_gamp, ctx->stream));
  return MFTX_OK;
}

int mftx_phase_grad_many(mftx_ctx* ctx, const void* amp, int amp_dtype, int64_t ld_amp, const void* opd,
                         int opd_dtype, int64_t ld_opd, int64_t m, int64_t n, const double* wavelengths,
                         int32_t nwave, const void* D, int64_t ldD, int64_t strideD, double* grad_opd,
                         int64_t ld_gopd, double* grad_amp, int64_t ld_gamp) {
  if (!ctx) return MFTX_EHANDLE;
  if (!amp || !opd || !D || !grad_opd || (nwave > 0 && !wavelengths))
    return fail(ctx, MFTX_EINVAL, "null array pointer");
  if ((amp_dtype != 1 && amp_dtype != 3) || (opd_dtype != 1 && opd_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "amp/opd must be float32 or float64");
  if (m <= 0 || n <= 0 || ld_amp < n || ld_opd < n || ldD < n || ld_gopd < n || (grad_amp && ld_gamp < n) ||
      m > (1 << 20) || n > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "bad gradient shape");
  if (nwave < 0 || (nwave > 1 && strideD < m * ldD)) return fail(ctx, MFTX_EINVAL, "bad wavelength count or stride");
  if ((size_t)nwave * sizeof(double) > kSlotBytes) return fail(ctx, MFTX_EINVAL, "too many wavelengths");
  std::vector<double> inv(nwave);
  for (int32_t w = 0; w < nwave; ++w) {
    if (!(wavelengths[w] > 0) || !std::isfinite(wavelengths[w]))
      return fail(ctx, MFTX_EINVAL, "wavelengths must be positive");
    inv[w] = 1.0 / wavelengths[w];
  }
  if (nwave == 0) return MFTX_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  void* dev = nullptr;
  const int rc = upload_desc(ctx, inv.data(), sizeof(double) * inv.size(), &dev);
  if (rc) return rc;
  HIPCHK(ctx, launch_phase_grad(amp, amp_dtype, ld_amp, opd, opd_dtype, ld_opd, (int)m, (int)n, (const double*)dev,
                                0.0, nwave, D, ldD, strideD, grad_opd, ld_gopd, grad_amp, ld_gamp, ctx->stream));
  return MFTX_OK;
}

int mftx_dft2_many(mftx_ctx* ctx, const mftx_mft* mfts, int32_t count, int32_t precision) {
  return guarded(ctx, [&]() -> int {
  if (!ctx) return MFTX_EHANDLE;
  const int rc = check_mfts(ctx, mfts, count, precision);
  if (rc) return rc;
  if (count == 0) return MFTX_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return dft2_many_run(ctx, mfts, count, precision);
  });
}

int mftx_mft_plan_create(mftx_ctx* ctx, const mftx_mft* mfts, int32_t count, int32_t precision,
                         mftx_mft_plan** out) {
  return guarded(ctx, [&]() -> int {
  if (!ctx) return MFTX_EHANDLE;
  if (!out) return fail(ctx, MFTX_EINVAL, "null argument");
  *out = nullptr;
  if (count <= 0) return fail(ctx, MFTX_EINVAL, "empty MFT list");
  int rc = check_mfts(ctx, mfts, count, precision);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<mftx_mft_plan> plan(new mftx_mft_plan());  // frees its device block on every exit
  plan->ctx = ctx;
  plan->prec = precision;
  many_build(mfts, count, precision, plan->mb);
  const size_t wsb = (size_t)round_up((int64_t)plan->mb.ws, 256);
  if (!try_malloc(&plan->mem, wsb + plan->mb.bytes)) {
    plan->mem = nullptr;
    return fail(ctx, MFTX_ENOMEM, "MFT plan allocation failed");
  }
  plan->desc = (char*)plan->mem + wsb;
  std::vector<char> host(plan->mb.bytes);
  many_place(plan->mb, (char*)plan->mem, plan->desc, host.data());
  if (hipMemcpy(plan->desc, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(ctx, MFTX_EDEVICE, "MFT plan upload failed");
  *out = plan.release();
  return MFTX_OK;
  });
}

int mftx_mft_plan_execute(mftx_mft_plan* plan) {
  if (!plan) return MFTX_EHANDLE;
  mftx_ctx* ctx = plan->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, many_launch(plan->mb, plan->prec, plan->desc, ctx->stream));
  return MFTX_OK;
}

int mftx_mft_plan_destroy(mftx_mft_plan* plan) {
  if (!plan) return MFTX_EHANDLE;
  (void)hipSetDevice(plan->ctx->device);
  (void)hipStreamSynchronize(plan->ctx->stream);
  delete plan;
  return MFTX_OK;
}

// ---------------------------------------------------------------------------
// detector-grid post-processing
// ---------------------------------------------------------------------------

int mftx_rebin(mftx_ctx* ctx, const void* img, int dtype, int64_t nimg, int64_t rows, int64_t cols, int64_t ld,
               int64_t factor, void* out, int out_dtype) {
  if (!ctx) return MFTX_EHANDLE;
  if (!img || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if ((dtype != 1 && dtype != 3) || (out_dtype != 1 && out_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "rebin is defined for float32/float64 data");
  if (nimg <= 0 || rows <= 0 || cols <= 0 || ld < cols || factor <= 0 || factor > (1 << 16))
    return fail(ctx, MFTX_EINVAL, "bad rebin shape");
  // util.rebin reshapes into factor x factor blocks, which needs whole blocks
  if (rows % factor || cols % factor) return fail(ctx, MFTX_EINVAL, "rebin factor must divide the image shape");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_rebin(img, dtype, nimg, rows, cols, ld, (int)factor, out, out_dtype, ctx->stream));
  return MFTX_OK;
}

// wrap v modulo n into (-n/2, n/2]
static double wrap_coord(int64_t v, int64_t n) {
  int64_t w = ((v % n) + n) % n;
  if (2 * w > n) w -= n;
  return (double)w;
}

int mftx_fft_filter(mftx_ctx* ctx, const void* img, int dtype, int64_t rows, int64_t cols, int64_t ld,
                    const void* kernel, int kernel_dtype, int64_t ldk, int normalize, void* out, int out_dtype,
                    int64_t ldo, int precision) {
  if (!ctx) return MFTX_EHANDLE;
  if (!img || !kernel || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if ((dtype != 1 && dtype != 3) || (kernel_dtype != 1 && kernel_dtype != 3) || (out_dtype != 1 && out_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "image, kernel and output must be float32 or float64");
  if (rows <= 0 || cols <= 0 || ld < cols || ldk < cols || ldo < cols || rows > (1 << 20) || cols > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "bad filter shape");
  if (precision != MFTX_PREC_FP32 && precision != MFTX_PREC_BF16X3) return fail(ctx, MFTX_EINVAL, "unknown precision");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int64_t n = rows * cols;
  const int64_t nb = (n + 255) / 256;
  const size_t fbytes = (size_t)round_up(n * (int64_t)sizeof(float2), 256);
  int rc = ensure_aux(ctx, 2 * fbytes + sizeof(double) * (size_t)(2 * nb + 2));
  if (rc) return rc;
  char* a = (char*)ctx->aux;
  float2* F = (float2*)a;
  float2* Y = (float2*)(a + fbytes);
  double* part = (double*)(a + 2 * fbytes);
  double* sums = part + 2 * nb;
  // fft2 then ifft2 as MFTs with alpha = 1/N: input coordinate r and output
  // coordinate u (uncentred) are r - N//2 + N//2 and (u - N//2) + N//2,
  // offsets taken modulo N
  mftx_mft e = {};
  e.f = img;
  e.f_dtype = dtype;
  e.m = rows;
  e.n = cols;
  e.ldf = ld;
  e.F = F;
  e.M = rows;
  e.N = cols;
  e.ldF = cols;
  e.alpha_r = 1.0 / (double)rows;
  e.alpha_c = 1.0 / (double)cols;
  e.offset_r = wrap_coord(rows / 2, rows);
  e.offset_c = wrap_coord(cols / 2, cols);
  e.shift_r = wrap_coord(-(rows / 2), rows);
  e.shift_c = wrap_coord(-(cols / 2), cols);
  e.scale = 1.0;
  e.w = kernel;  // the MTF multiplies the spectrum in the store
  e.w_dtype = kernel_dtype;
  e.ldw = ldk;
  rc = dft2_many_run(ctx, &e, 1, precision);
  if (rc) return rc;
  mftx_mft b = e;
  b.f = F;
  b.f_dtype = 0;
  b.ldf = cols;
  b.F = Y;
  b.conj = 1;
  b.scale = 1.0 / ((double)rows * (double)cols);
  b.w = nullptr;
  rc = dft2_many_run(ctx, &b, 1, precision);
  if (rc) return rc;
  HIPCHK(ctx, launch_magnitude(Y, cols, rows, cols, out, out_dtype, ldo, part, sums + 1, ctx->stream));
  if (normalize) {
    HIPCHK(ctx, launch_sum(img, dtype, rows, cols, ld, part + nb, sums, ctx->stream));
    HIPCHK(ctx, launch_scale_ratio(out, out_dtype, rows, cols, ldo, sums, sums + 1, ctx->stream));
  }
  return MFTX_OK;
}

int mftx_rescale(mftx_ctx* ctx, const void* img, int dtype, int64_t rows, int64_t cols, int64_t ld, const void* mask,
                 int mask_dtype, int64_t ldm, double scale, int64_t out_rows, int64_t out_cols, int order, int mode,
                 int unitary, void* out, int out_dtype, int64_t ldo) {
  if (!ctx) return MFTX_EHANDLE;
  if (!img || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  const bool cplx = dtype == 0 || dtype == 2;
  if (dtype < 0 || dtype > 3 || (mask && mask_dtype != 1 && mask_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "image must be float32, float64, complex64 or complex128 and a mask real");
  if (cplx ? (out_dtype != 0 && out_dtype != 2) : (out_dtype != 1 && out_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "the output must be real for a real image and complex for a complex one");
  if (order < 0 || order > 5) return fail(ctx, MFTX_EINVAL, "spline order must be in 0..5");
  if (mode < MFTX_MODE_NEAREST || mode > MFTX_MODE_GRID_CONSTANT) return fail(ctx, MFTX_EINVAL, "unknown boundary mode");
  if (rows <= 0 || cols <= 0 || ld < cols || out_rows <= 0 || out_cols <= 0 || ldo < out_cols ||
      (mask && ldm < cols) || rows > (1 << 20) || cols > (1 << 20) || out_rows > (1 << 20) || out_cols > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "bad rescale shape");
  if (!(scale > 0) || !std::isfinite(scale)) return fail(ctx, MFTX_EINVAL, "scale must be positive");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int npad = 12;  // the most any mode pads
  const int64_t np_ = (rows + 2 * npad) * (cols + 2 * npad), no = out_rows * out_cols;
  const int64_t nb = (std::max(np_, no) + 255) / 256;
  const size_t b_pad = (size_t)round_up(np_ * 8, 256), b_out = (size_t)round_up(no * 8, 256);
  int rc = ensure_aux(ctx, 2 * b_pad + 3 * b_out + sizeof(double) * (size_t)(nb + 4));
  if (rc) return rc;
  char* a = (char*)ctx->aux;
  double* P = (double*)a;
  double* T = (double*)(a + b_pad);
  double* raw = (double*)(a + 2 * b_pad);
  double* raw_im = (double*)(a + 2 * b_pad + b_out);
  double* mval = (double*)(a + 2 * b_pad + 2 * b_out);
  double* part = (double*)(a + 2 * b_pad + 3 * b_out);
  double* sums = part + nb;
  // the mask's dtype sets the threshold (util.py:338-339, numpy finfo eps of
  // the mask, or of the image's real dtype when the mask is derived from it)
  const int mdt = mask ? mask_dtype : dtype;
  const bool single = mdt == 0 || mdt == 1;
  const double eps = single ? 1.1920928955078125e-07 : 2.220446049250313e-16;
  HIPCHK(ctx, launch_rescale(img, dtype, rows, cols, ld, mask ? mask : img, mask ? 2 : 1, mdt, mask ? ldm : ld, eps,
                             scale, out_rows, out_cols, order, mode, unitary, out, out_dtype, ldo, P, T, raw, raw_im,
                             mval, part, sums, ctx->stream));
  return MFTX_OK;
}

// ---------------------------------------------------------------------------
// modal OPD operations
// ---------------------------------------------------------------------------

int mftx_ptt_fit(mftx_ctx* ctx, const void* opd, int opd_dtype, int64_t ld_opd, const void* masks, int mask_dtype,
                 int64_t ld_mask, int64_t mask_stride, int32_t nseg, int64_t rows, int64_t cols, double dx_r,
                 double dx_c, double* normal) {
  if (!ctx) return MFTX_EHANDLE;
  if (!opd || !masks || !normal) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if ((opd_dtype != 1 && opd_dtype != 3) || (mask_dtype != 1 && mask_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "opd and masks must be float32 or float64");
  if (rows <= 0 || cols <= 0 || ld_opd < cols || ld_mask < cols || nseg <= 0 || nseg > 65535 ||
      (nseg > 1 && mask_stride < rows * ld_mask))
    return fail(ctx, MFTX_EINVAL, "bad tilt-fit shape");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int64_t nblk = (rows * cols + 255) / 256;
  int rc = ensure_aux(ctx, sizeof(double) * (size_t)(nseg * nblk * 9 + nseg * 9));
  if (rc) return rc;
  double* part = (double*)ctx->aux;
  double* out = part + nseg * nblk * 9;
  HIPCHK(ctx, launch_ptt_normal(opd, opd_dtype, ld_opd, masks, mask_dtype, ld_mask, mask_stride, nseg, rows, cols,
                                dx_r, dx_c, part, out, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(normal, out, sizeof(double) * 9 * nseg, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return MFTX_OK;
}

int mftx_remove_tilt(mftx_ctx* ctx, const void* opd, int opd_dtype, int64_t ld_opd, const void* masks, int mask_dtype,
                     int64_t ld_mask, int64_t mask_stride, int32_t nseg, int32_t multi, int64_t rows, int64_t cols,
                     double dx_r, double dx_c, const double* tilt, double* out, int64_t ldo) {
  if (!ctx) return MFTX_EHANDLE;
  if (!opd || !masks || !tilt || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if ((opd_dtype != 1 && opd_dtype != 3) || (mask_dtype != 1 && mask_dtype != 3))
    return fail(ctx, MFTX_EINVAL, "opd and masks must be float32 or float64");
  if (rows <= 0 || cols <= 0 || ld_opd < cols || ld_mask < cols || ldo < cols || nseg <= 0 ||
      (!multi && nseg != 1) || (size_t)nseg * 2 * sizeof(double) > kSlotBytes ||
      (nseg > 1 && mask_stride < rows * ld_mask))
    return fail(ctx, MFTX_EINVAL, "bad tilt-removal shape");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  void* dev = nullptr;
  const int rc = upload_desc(ctx, tilt, sizeof(double) * 2 * nseg, &dev);
  if (rc) return rc;
  HIPCHK(ctx, launch_remove_tilt(opd, opd_dtype, ld_opd, masks, mask_dtype, ld_mask, mask_stride, nseg, multi, rows,
                                 cols, dx_r, dx_c, (const double*)dev, out, ldo, ctx->stream));
  return MFTX_OK;
}

int mftx_zernike_basis(mftx_ctx* ctx, const double* rho, const double* theta, const void* mask, int mask_dtype,
                       int64_t npix, const int32_t* noll_nm, int32_t nmodes, int normalize, double* out) {
  return guarded(ctx, [&]() -> int {
  if (!ctx) return MFTX_EHANDLE;
  if (!rho || !theta || !mask || !noll_nm || !out) return fail(ctx, MFTX_EINVAL, "null array pointer");
  if (mask_dtype != 1 && mask_dtype != 3) return fail(ctx, MFTX_EINVAL, "mask must be float32 or float64");
  if (npix <= 0 || nmodes <= 0 || nmodes > 65535) return fail(ctx, MFTX_EINVAL, "bad basis shape");
  // radial coefficients (-1)^k (n-k)! / (k! ((n+m)/2-k)! ((n-m)/2-k)!) (zernike.py:104-107),
  // integers, exact in double while below 2^53
  std::vector<int> desc, pw;
  std::vector<double> coef, norm;
  for (int32_t i = 0; i < nmodes; ++i) {
    const int n = noll_nm[2 * i], m = noll_nm[2 * i + 1];
    const int am = std::abs(m);
    if (n < 0 || am > n || n > 60) return fail(ctx, MFTX_EINVAL, "bad Zernike (n, m)");
    desc.push_back(n);
    desc.push_back(m);
    desc.push_back((int)coef.size());
    int nk = 0;
    if ((n - am) % 2 == 0) {
      for (int k = 0; k <= (n - am) / 2; ++k) {
        // exact rational in long double, then rounded once (an integer)
        long double c = 1.0L;
        for (int q = 2; q <= n - k; ++q) c *= q;
        for (int q = 2; q <= k; ++q) c /= q;
        for (int q = 2; q <= (n + am) / 2 - k; ++q) c /= q;
        for (int q = 2; q <= (n - am) / 2 - k; ++q) c /= q;
        coef.push_back((double)((k % 2 ? -1.0L : 1.0L) * std::nearbyint((double)c)));
        pw.push_back(n - 2 * k);
        ++nk;
      }
    }
    desc.push_back(nk);
    norm.push_back(normalize ? (m == 0 ? std::sqrt((double)(n + 1)) : std::sqrt(2.0) * std::sqrt((double)(n + 1)))
                             : 1.0);
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t o_desc = 0, o_norm = round_up(sizeof(int) * desc.size(), 16);
  const size_t o_coef = round_up(o_norm + sizeof(double) * norm.size(), 16);
  const size_t o_pw = round_up(o_coef + sizeof(double) * std::max<size_t>(coef.size(), 1), 16);
  const size_t bytes = o_pw + sizeof(int) * std::max<size_t>(pw.size(), 1);
  if (bytes > kSlotBytes) return fail(ctx, MFTX_EINVAL, "too many Zernike modes for one call");
  std::vector<char> host(bytes);
  std::memcpy(host.data() + o_desc, desc.data(), sizeof(int) * desc.size());
  std::memcpy(host.data() + o_norm, norm.data(), sizeof(double) * norm.size());
  if (!coef.empty()) std::memcpy(host.data() + o_coef, coef.data(), sizeof(double) * coef.size());
  if (!pw.empty()) std::memcpy(host.data() + o_pw, pw.data(), sizeof(int) * pw.size());
  void* dev = nullptr;
  const int rc = upload_desc(ctx, host.data(), bytes, &dev);
  if (rc) return rc;
  char* d = (char*)dev;
  HIPCHK(ctx, launch_zernike(rho, theta, mask, mask_dtype, npix, nmodes, (const int*)(d + o_desc),
                             (const double*)(d + o_coef), (const int*)(d + o_pw), (const double*)(d + o_norm), out,
                             ctx->stream));
  return MFTX_OK;
  });
}

int mftx_zero(mftx_ctx* ctx, void* p, int64_t bytes) {
  if (!ctx) return MFTX_EHANDLE;
  if (!p || bytes < 0) return fail(ctx, MFTX_EINVAL, "bad zero request");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_zero(p, (size_t)bytes, ctx->stream));
  return MFTX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched plan
// ---------------------------------------------------------------------------

namespace {

// Problems with equal keys have one output grid and one stage-1 twiddle set
// (per member tile), so their stage-2 GEMMs concatenate along K.
struct ClassKey {
  int wave, M, N;
  double ar, ac, sr, sc;
  int64_t r0, c0;
  bool operator<(const ClassKey& o) const {
    return std::tie(wave, M, N, ar, ac, sr, sc, r0, c0) < std::tie(o.wave, o.M, o.N, o.ar, o.ac, o.sr, o.sc, o.r0, o.c0);
  }
};

struct Layout {
  size_t total = 256;  // leading pad: no real buffer sits at offset 0
  size_t take(size_t bytes) {
    const size_t off = total;
    total += (size_t)round_up((int64_t)bytes, 256);
    return off;
  }
};

struct ClassInfo {
  ClassKey key;
  std::vector<int> members;  // problem indices
  int group;
  bool col_first;
  int64_t rows2;      // stage-2 A rows (tile-padded): N (row-first) or M (column-first)
  int ktot;           // stage-2 K
  std::vector<int> kbeg;
  size_t abuf = 0;    // stage-2 A buffer
  size_t fbuf = 0;    // complex F (partial-overlap groups)
  int p2 = -1;        // index of its stage-2 problem
};

}  // namespace

extern "C" {

int mftx_plan_psf(mftx_ctx* ctx, const mftx_tile* tiles, int32_t ntiles, const double* wavelengths, int32_t nwave,
                  const mftx_problem* problems, int32_t nprob, int64_t image_rows, int64_t image_cols,
                  int32_t precision, int32_t lambda_group, int32_t tile_n, mftx_plan** out) {
  return guarded(ctx, [&]() -> int {
  if (!ctx) return MFTX_EHANDLE;
  if (!out || !tiles || !wavelengths || (!problems && nprob > 0)) return fail(ctx, MFTX_EINVAL, "null argument");
  *out = nullptr;
  if (ntiles <= 0 || nwave <= 0 || nprob < 0) return fail(ctx, MFTX_EINVAL, "empty plan");
  if (precision != MFTX_PREC_FP32 && precision != MFTX_PREC_BF16X3) return fail(ctx, MFTX_EINVAL, "unknown precision");
  if (image_rows <= 0 || image_cols <= 0 || image_rows > (1 << 20) || image_cols > (1 << 20))
    return fail(ctx, MFTX_EINVAL, "bad image shape");
  if (lambda_group <= 0) lambda_group = 8;
  if (tile_n == 0) tile_n = BN_DEFAULT;
  if (!valid_bn(tile_n)) return fail(ctx, MFTX_EINVAL, "tile_n must be 0, 64 or 128");
  const int kg = k_granule(tile_n);
  for (int t = 0; t < ntiles; ++t) {
    const mftx_tile& T = tiles[t];
    if (!T.amp || !T.opd || T.m <= 0 || T.n <= 0 || T.m > (1 << 20) || T.n > (1 << 20) || T.ld_amp < T.n ||
        T.ld_opd < T.n || (T.amp_dtype != 1 && T.amp_dtype != 3) || (T.opd_dtype != 1 && T.opd_dtype != 3))
      return fail(ctx, MFTX_EINVAL, "bad tile " + std::to_string(t));
  }
  for (int w = 0; w < nwave; ++w)
    if (!(wavelengths[w] > 0)) return fail(ctx, MFTX_EINVAL, "wavelengths must be positive");
  for (int i = 0; i < nprob; ++i) {
    const mftx_problem& P = problems[i];
    if (P.tile < 0 || P.tile >= ntiles || P.wave < 0 || P.wave >= nwave || P.M <= 0 || P.N <= 0 ||
        P.M > (1 << 20) || P.N > (1 << 20) || std::llabs(P.out_r0) > (1 << 30) || std::llabs(P.out_c0) > (1 << 30) ||
        !std::isfinite(P.alpha_r) || !std::isfinite(P.alpha_c) ||
        !std::isfinite(P.shift_r) || !std::isfinite(P.shift_c) || !std::isfinite(P.offset_r) ||
        !std::isfinite(P.offset_c) || !std::isfinite(P.weight))
      return fail(ctx, MFTX_EINVAL, "bad problem " + std::to_string(i));
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int prec = precision;

  // ---- classes and coherent groups ---------------------------------------------
  std::map<ClassKey, int> class_of;
  std::vector<ClassInfo> cls;
  for (int i = 0; i < nprob; ++i) {
    const mftx_problem& P = problems[i];
    const ClassKey k{P.wave, P.M, P.N, P.alpha_r, P.alpha_c, P.shift_r, P.shift_c, P.out_r0, P.out_c0};
    auto it = class_of.find(k);
    if (it == class_of.end()) {
      it = class_of.emplace(k, (int)cls.size()).first;
      ClassInfo c;
      c.key = k;
      c.group = P.group;
      cls.push_back(c);
    }
    if (cls[it->second].group != P.group)
      return fail(ctx, MFTX_EINVAL, "problems with identical output geometry must share a coherent group");
    cls[it->second].members.push_back(i);
  }
  std::map<std::pair<int, int>, std::vector<int>> groups;  // (wave, group) -> classes
  for (size_t c = 0; c < cls.size(); ++c) groups[{cls[c].key.wave, cls[c].group}].push_back((int)c);
  // A group with one class runs the intensity path (|F|^2 summed in
  // registers); the classes of a group with several are summed as complex
  // fields first (the fallback), and only they.
  bool fast = true;  // every group has one class
  std::vector<char> cfast(cls.size(), 1);
  for (auto& g : groups)
    if (g.second.size() > 1) {
      fast = false;
      for (int c : g.second) cfast[c] = 0;
    }

  // contraction order per class (all members share M, N)
  for (auto& c : cls) {
    double row_first = 0, col_first = 0;
    for (int i : c.members) {
      const mftx_tile& T = tiles[problems[i].tile];
      row_first += cost_row_first(T.m, T.n, c.key.M, c.key.N);
      col_first += cost_col_first(T.m, T.n, c.key.M, c.key.N);
    }
    c.col_first = col_first < row_first;
  }

  // Each problem's stage-1 operand is its tile's phasor at its wavelength,
  // carrying the folds of its class's two output grids (stage 1's along K,
  // stage 2's along rows; the grids follow the shift and size, x0 the Field
  // offset).  Problems of one tile and orientation share an operand slot
  // when wavelength and folds agree, and get one each when they do not (the
  // same tile placed twice at one wavelength).
  using Slot = std::tuple<int, double, double, double, double>;  // wave, fold1, x01, fold2, x02
  std::vector<Slot> pslot(nprob);
  for (auto& c : cls) {
    const ClassKey& K = c.key;
    const bool cf = c.col_first;
    for (int i : c.members) {
      const mftx_problem& P = problems[i];
      const mftx_tile& T = tiles[P.tile];
      const OutGrid g1 = make_grid(-(cf ? P.alpha_r : P.alpha_c), cf ? K.sr : K.sc, cf ? K.M : K.N);
      const OutGrid g2 = make_grid(-(cf ? K.ac : K.ar), cf ? K.sc : K.sr, cf ? K.N : K.M);
      pslot[i] = Slot{P.wave, g1.fold, cf ? -centre(T.m) + P.offset_r : -centre(T.n) + P.offset_c, g2.fold,
                      cf ? -centre(T.n) + P.offset_c : -centre(T.m) + P.offset_r};
    }
  }

  // ---- workspace layout ----------------------------------------------------------
  Layout lay;
  // phasor operands per (tile, orientation): every slot the pair is used with
  std::map<std::pair<int, int>, std::vector<Slot>> tile_slots;
  for (auto& c : cls)
    for (int i : c.members) tile_slots[{problems[i].tile, c.col_first ? 1 : 0}].push_back(pslot[i]);
  struct TileOp { int64_t rows; int kpad; size_t off; std::map<Slot, int> slot; int wave_off; int nslot; };
  std::map<std::pair<int, int>, TileOp> tops;
  std::vector<Slot> wave_slots;  // every phasor slot, grouped by (tile, orientation)
  for (auto& kv : tile_slots) {
    auto v = kv.second;
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    const mftx_tile& T = tiles[kv.first.first];
    const bool tr = kv.first.second == 1;
    TileOp op;
    op.rows = round_up(tr ? T.n : T.m, BM);
    op.kpad = (int)round_up(tr ? T.m : T.n, kg);
    op.wave_off = (int)wave_slots.size();
    op.nslot = (int)v.size();
    for (size_t j = 0; j < v.size(); ++j) {
      op.slot[v[j]] = (int)j;
      wave_slots.push_back(v[j]);
    }
    op.off = lay.take(packed_bytes(prec, op.rows, op.kpad) * v.size());
    tops[kv.first] = op;
  }
  // stage-2 operands per class
  for (auto& c : cls) {
    c.rows2 = round_up(c.col_first ? c.key.M : c.key.N, BM);
    int k = 0;
    for (int i : c.members) {
      const mftx_tile& T = tiles[problems[i].tile];
      c.kbeg.push_back(k);
      k += (int)round_up(c.col_first ? T.n : T.m, kg);
    }
    c.ktot = k;
    c.abuf = lay.take(packed_bytes(prec, c.rows2, c.ktot));
    if (!cfast[&c - cls.data()]) c.fbuf = lay.take(sizeof(float2) * (size_t)c.key.M * c.key.N);
  }

  // ---- descriptors ---------------------------------------------------------------
  std::vector<Problem> p1, p2;
  std::vector<double> cx;
  std::vector<std::vector<Work>> clus1, clus2;
  struct Patch { int idx; size_t a, dst; };  // workspace offsets to rebase after allocation
  std::vector<Patch> patch1, patch2;
  std::vector<int64_t> s1rows;  // tiled rows of each stage-1 problem

  for (auto& c : cls) {
    const ClassKey& K = c.key;
    const bool cf = c.col_first;
    // stage 1, one problem per member
    for (size_t j = 0; j < c.members.size(); ++j) {
      const mftx_problem& P = problems[c.members[j]];
      const mftx_tile& T = tiles[P.tile];
      const TileOp& op = tops[{P.tile, cf ? 1 : 0}];
      const int kp2 = (int)round_up(cf ? T.n : T.m, kg);
      const int n1 = cf ? K.M : K.N;  // output columns of stage 1
      const Slot& sl = pslot[c.members[j]];
      Problem q = {};
      q.lda = packed_lda(prec, op.kpad);
      q.alpha = -(cf ? P.alpha_r : P.alpha_c);
      const OutGrid g1 = make_grid(q.alpha, cf ? K.sr : K.sc, n1);
      q.y0 = g1.y0;
      q.rows = (int)(cf ? T.n : T.m);
      q.cols = g1.slots;
      q.ncols = n1;
      q.cp = g1.cp;
      q.cm = g1.cm;
      q.kpad = op.kpad;
      q.chunk_begin = (int)cx.size();
      push_chunks(cx, op.kpad, std::get<2>(sl));  // x01
      q.scale = 1.0f;
      q.limit = kp2;
      q.ldd = packed_lda(prec, c.ktot);
      // both folds are in the operand
      const size_t elem = prec == PREC_FP32 ? sizeof(float2) : 4 * sizeof(uint16_t);
      const int slot = op.slot.at(sl);
      patch1.push_back(Patch{(int)p1.size(), op.off + (size_t)slot * packed_bytes(prec, op.rows, op.kpad),
                             c.abuf + (size_t)c.kbeg[j] * elem});
      s1rows.push_back(op.rows);
      p1.push_back(q);
    }
    // stage 2, one problem per class (K concatenated over members)
    Problem q = {};
    const int n2 = cf ? K.N : K.M;
    q.lda = packed_lda(prec, c.ktot);
    q.alpha = -(cf ? K.ac : K.ar);
    const OutGrid g2 = make_grid(q.alpha, cf ? K.sc : K.sr, n2);
    q.y0 = g2.y0;
    q.rows = cf ? K.M : K.N;
    q.cols = g2.slots;
    q.ncols = n2;
    q.cp = g2.cp;
    q.cm = g2.cm;
    q.kpad = c.ktot;
    q.chunk_begin = (int)cx.size();
    for (size_t j = 0; j < c.members.size(); ++j) {
      const mftx_problem& P = problems[c.members[j]];
      const mftx_tile& T = tiles[P.tile];
      push_chunks(cx, (int)round_up(cf ? T.n : T.m, kg), cf ? -centre(T.n) + P.offset_c : -centre(T.m) + P.offset_r);
    }
    q.trans = cf ? 0 : 1;
    const double unit = std::sqrt(std::fabs(K.ar * K.ac));  // unitary scale (fourier.py:100-101)
    const bool cl_fast = cfast[&c - cls.data()];
    if (cl_fast) {
      q.scale = (float)(unit * unit * problems[c.members[0]].weight);
    } else {
      q.scale = (float)unit;
      q.ldd = K.N;  // F[u][v]
    }
    c.p2 = (int)p2.size();
    patch2.push_back(Patch{(int)p2.size(), c.abuf, cl_fast ? 0 : c.fbuf});
    p2.push_back(q);
  }

  // stage-1 works: runs of g1 consecutive problems with one tiling and K
  // length (the wavelengths of a tile) share a work item, so the staging
  // pipeline runs on across their K loops instead of refilling for every
  // short K loop; g1 keeps at least one work item per CU
  {
    std::vector<int64_t> rows(s1rows.begin(), s1rows.end());
    for (const auto& r : group_runs(p1, rows, tile_n)) add_tiles(clus1, r.first, r.second, s1rows[r.first],
                                                                 p1[r.first].cols, tile_n);
  }
  // owned until returned: every failure or exception path frees the plan and its workspace
  std::unique_ptr<mftx_plan, int (*)(mftx_plan*)> owned(new mftx_plan(), mftx_plan_destroy);
  mftx_plan* plan = owned.get();
  plan->ctx = ctx;
  plan->prec = prec;
  plan->bn = tile_n;
  plan->img_rows = image_rows;
  plan->img_cols = image_cols;

  // stage-2 works: λ-groups of classes with one output window and orientation
  // accumulate |F|^2 in registers (the intensity path); a class of a
  // several-class group gets a complex-store work set of its own.
  // A work item maps every problem's slots through its first problem's output
  // grid (store_partial), so classes share a λ-group only when their stage-2
  // grids agree too: the pairing columns (cp, cm) and the slot count, which
  // follow from the sub-pixel shift (DispersiveTilt moves it per wavelength
  // while the integer placement stays).  Partial images are indexed by output
  // pixel, so every λ-group of one window, whatever its grid, is summed by
  // that window's single reduction.
  struct GeoInfo { std::vector<std::vector<int>> lgroups; size_t partial; int64_t prow, pcol; };
  std::vector<GeoInfo> geos;
  {
    using Grid = std::tuple<int, int, int>;  // stage-2 (cp, cm, slots)
    std::map<std::tuple<int, int, int64_t, int64_t, int>, std::map<Grid, std::vector<int>>> geo;
    for (size_t c = 0; c < cls.size(); ++c) {
      if (!cfast[c]) continue;
      const Problem& q = p2[cls[c].p2];
      geo[{cls[c].key.M, cls[c].key.N, cls[c].key.r0, cls[c].key.c0, cls[c].col_first ? 1 : 0}][{q.cp, q.cm, q.cols}]
          .push_back((int)c);
    }
    for (auto& kv : geo) {
      GeoInfo gi;
      for (auto& sub : kv.second) {
        const auto& cl = sub.second;
        for (size_t s = 0; s < cl.size(); s += lambda_group)
          gi.lgroups.emplace_back(cl.begin() + s, cl.begin() + std::min(cl.size(), s + (size_t)lambda_group));
      }
      gi.prow = round_up(std::get<0>(kv.first), BM);  // partial[u][v]
      gi.pcol = round_up(std::get<1>(kv.first), BM);
      gi.partial = lay.take(sizeof(float) * gi.prow * gi.pcol * gi.lgroups.size());
      geos.push_back(gi);
    }
  }
  // stage-2 problem lists in launch order (λ-groups contiguous), intensity
  // and complex-store
  std::vector<Problem> p2_launch, p2c_launch;
  std::vector<Patch> patch2_launch, patch2c_launch;
  std::vector<std::vector<Work>> clus2c;
  {
    for (auto& gi : geos) {
      for (size_t g = 0; g < gi.lgroups.size(); ++g) {
        const int pb = (int)p2_launch.size();
        for (int c : gi.lgroups[g]) {
          p2_launch.push_back(p2[cls[c].p2]);
          patch2_launch.push_back(Patch{(int)p2_launch.size() - 1, patch2[cls[c].p2].a, 0});
        }
        const ClassInfo& c0 = cls[gi.lgroups[g][0]];
        const size_t poff = gi.partial + sizeof(float) * gi.prow * gi.pcol * g;
        // the partial image offset rides in the pointer field until the workspace exists
        add_tiles(clus2, pb, (int)p2_launch.size(), c0.rows2, p2[c0.p2].cols, tile_n,
                  (float*)(uintptr_t)poff, (long long)gi.pcol, c0.key.M, c0.key.N);
      }
    }
    for (size_t ci = 0; ci < cls.size(); ++ci) {
      if (cfast[ci]) continue;
      const ClassInfo& c = cls[ci];
      const int pb = (int)p2c_launch.size();
      p2c_launch.push_back(p2[c.p2]);
      patch2c_launch.push_back(Patch{pb, patch2[c.p2].a, patch2[c.p2].dst});
      add_tiles(clus2c, pb, pb + 1, c.rows2, p2[c.p2].cols, tile_n);
    }
  }
  std::vector<Work> w1 = xcd_order(clus1), w2 = xcd_order(clus2), w2c = xcd_order(clus2c);
  // the complex-field set: each class's stage-2 problem, one work set per class
  std::vector<std::vector<Work>> clusf;
  for (size_t c = 0; c < cls.size(); ++c) {
    Problem q = p2[cls[c].p2];
    q.dst = nullptr;
    q.wmap = nullptr;
    plan->pf.push_back(q);
    plan->class_first.push_back(cls[c].members[0]);
    add_tiles(clusf, (int)c, (int)c + 1, cls[c].rows2, q.cols, tile_n);
  }
  std::vector<Work> wf = xcd_order(clusf);
  // executed MFMA FLOPs of the two launches: per workgroup and problem, BM rows
  // x slots x kpad x 4 real products (re/im x cos/sin) x 2, times the passes
  const double passes = prec == PREC_BF16X3 ? 3.0 : 1.0;
  const double per_slot_k = 8.0 * BM * slots_per_tile(tile_n) * passes;
  double mf1 = 0, mf2 = 0;
  for (auto& w : w1)
    for (int i = w.pbeg; i < w.pend; ++i) mf1 += per_slot_k * p1[i].kpad;
  for (auto& w : w2)
    for (int i = w.pbeg; i < w.pend; ++i) mf2 += per_slot_k * p2_launch[i].kpad;
  for (auto& w : w2c)
    for (int i = w.pbeg; i < w.pend; ++i) mf2 += per_slot_k * p2c_launch[i].kpad;

  const size_t o_w1 = lay.take(sizeof(Work) * std::max<size_t>(w1.size(), 1));
  const size_t o_w2 = lay.take(sizeof(Work) * std::max<size_t>(w2.size(), 1));
  const size_t o_p1 = lay.take(sizeof(Problem) * std::max<size_t>(p1.size(), 1));
  const size_t o_p2 = lay.take(sizeof(Problem) * std::max<size_t>(p2_launch.size(), 1));
  const size_t o_w2c = lay.take(sizeof(Work) * std::max<size_t>(w2c.size(), 1));
  const size_t o_p2c = lay.take(sizeof(Problem) * std::max<size_t>(p2c_launch.size(), 1));
  const size_t o_cx = lay.take(sizeof(double) * std::max<size_t>(cx.size(), 1));
  const size_t o_iw = lay.take(sizeof(PhasorWave) * std::max<size_t>(wave_slots.size(), 1));
  const size_t o_ops = lay.take(sizeof(PhasorOp) * std::max<size_t>(tops.size(), 1));
  const size_t o_wf = lay.take(sizeof(Work) * std::max<size_t>(wf.size(), 1));
  const size_t o_pf = lay.take(sizeof(Problem) * std::max<size_t>(plan->pf.size(), 1));
  // partial-overlap groups: each group's union box on the image (its Fields'
  // placements clipped to the image, field.py:282-296); one accumulator of
  // the largest box serves every group in turn (same stream, in order)
  std::vector<std::array<int64_t, 4>> fb_box;  // r0, c0, rows, cols per several-class group (map order)
  size_t fb_acc = 0;
  if (!fast) {
    int64_t most = 0;
    for (auto& kv : groups) {
      if (kv.second.size() < 2) continue;
      int64_t r0 = image_rows, c0 = image_cols, r1 = 0, c1 = 0;
      for (int c : kv.second) {
        const ClassKey& K = cls[c].key;
        r0 = std::min(r0, std::max<int64_t>(K.r0, 0));
        c0 = std::min(c0, std::max<int64_t>(K.c0, 0));
        r1 = std::max(r1, std::min<int64_t>(K.r0 + K.M, image_rows));
        c1 = std::max(c1, std::min<int64_t>(K.c0 + K.N, image_cols));
      }
      const int64_t rows = std::max<int64_t>(r1 - r0, 0), cols = std::max<int64_t>(c1 - c0, 0);
      fb_box.push_back({r0, c0, rows, cols});
      most = std::max(most, rows * cols);
    }
    fb_acc = lay.take(sizeof(float2) * (size_t)std::max<int64_t>(most, 1));
  }

  if (!try_malloc(&plan->ws, lay.total)) {
    plan->ws = nullptr;
    return fail(ctx, MFTX_ENOMEM, "plan workspace allocation of " + std::to_string(lay.total) + " bytes failed");
  }
  char* base = (char*)plan->ws;
  for (auto& pt : patch1) {
    p1[pt.idx].a = base + pt.a;
    p1[pt.idx].dst = base + pt.dst;
  }
  for (auto& pt : patch2_launch) {
    p2_launch[pt.idx].a = base + pt.a;
    p2_launch[pt.idx].dst = pt.dst ? (void*)(base + pt.dst) : nullptr;
  }
  for (auto& pt : patch2c_launch) {
    p2c_launch[pt.idx].a = base + pt.a;
    p2c_launch[pt.idx].dst = base + pt.dst;
  }
  for (auto& w : w2)
    if (w.partial) w.partial = (float*)(base + (size_t)(uintptr_t)w.partial);
  for (size_t c = 0; c < cls.size(); ++c) plan->pf[c].a = base + patch2[cls[c].p2].a;
  std::vector<PhasorWave> iw(wave_slots.size());
  for (size_t i = 0; i < wave_slots.size(); ++i) {
    const Slot& sl = wave_slots[i];
    iw[i] = PhasorWave{1.0 / wavelengths[std::get<0>(sl)], std::get<1>(sl), std::get<2>(sl), std::get<3>(sl),
                       std::get<4>(sl)};
  }

  plan->works1 = (Work*)(base + o_w1);
  plan->works2 = (Work*)(base + o_w2);
  plan->probs1 = (Problem*)(base + o_p1);
  plan->probs2 = (Problem*)(base + o_p2);
  plan->chunk_x = (double*)(base + o_cx);
  plan->waves = (PhasorWave*)(base + o_iw);
  plan->nwork1 = (int)w1.size();
  plan->nwork2 = (int)w2.size();
  plan->works2c = (Work*)(base + o_w2c);
  plan->probs2c = (Problem*)(base + o_p2c);
  plan->nwork2c = (int)w2c.size();
  plan->works_f = (Work*)(base + o_wf);
  plan->probs_f = (Problem*)(base + o_pf);
  plan->nwork_f = (int)wf.size();
  hipError_t e = hipSuccess;
  auto up = [&](void* d, const void* s, size_t n) {
    if (e == hipSuccess && n) e = hipMemcpy(d, s, n, hipMemcpyHostToDevice);
  };
  up(plan->works1, w1.data(), sizeof(Work) * w1.size());
  up(plan->works2, w2.data(), sizeof(Work) * w2.size());
  up(plan->probs1, p1.data(), sizeof(Problem) * p1.size());
  up(plan->probs2, p2_launch.data(), sizeof(Problem) * p2_launch.size());
  up(plan->works2c, w2c.data(), sizeof(Work) * w2c.size());
  up(plan->probs2c, p2c_launch.data(), sizeof(Problem) * p2c_launch.size());
  up(plan->chunk_x, cx.data(), sizeof(double) * cx.size());
  up(plan->waves, iw.data(), sizeof(PhasorWave) * iw.size());
  up(plan->works_f, wf.data(), sizeof(Work) * wf.size());
  if (e != hipSuccess) {
    return hipfail(ctx, e, "plan descriptor upload");
  }

  std::vector<PhasorOp> ops;
  long long pthreads = 0;
  for (auto& kv : tops) {
    const TileOp& op = kv.second;
    const mftx_tile& T = tiles[kv.first.first];
    PhasorOp ph = {};
    ph.amp = T.amp;
    ph.opd = T.opd;
    ph.amp_dtype = T.amp_dtype;
    ph.opd_dtype = T.opd_dtype;
    ph.ld_amp = T.ld_amp;
    ph.ld_opd = T.ld_opd;
    ph.m = (int)T.m;
    ph.n = (int)T.n;
    ph.trans = kv.first.second;
    ph.nwave = op.nslot;
    ph.waves = plan->waves + op.wave_off;
    ph.dst = base + op.off;
    ph.wave_stride = (long long)packed_bytes(prec, op.rows, op.kpad);
    ph.rows_pad = op.rows;
    ph.kpad = op.kpad;
    ph.begin = group_begin(pthreads);
    pthreads = ph.begin + phasor_threads(ph.rows_pad, ph.kpad, ph.nwave);
    ops.push_back(ph);
  }
  plan->phasor_ops = (PhasorOp*)(base + o_ops);
  plan->n_phasor_ops = (int)ops.size();
  plan->phasor_threads = pthreads;
  e = hipMemcpy(plan->phasor_ops, ops.data(), sizeof(PhasorOp) * ops.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    return hipfail(ctx, e, "plan descriptor upload");
  }
  for (auto& gi : geos) {
    const ClassInfo& c0 = cls[gi.lgroups[0][0]];
    plan->reduces.push_back(mftx_plan::Reduce{(float*)(base + gi.partial), (long long)(gi.prow * gi.pcol),
                                              (long long)gi.pcol, (int)gi.lgroups.size(), c0.key.M, c0.key.N,
                                              c0.key.r0, c0.key.c0});
  }
  {
    size_t g = 0;
    for (auto& kv : groups) {
      if (kv.second.size() < 2) continue;
      mftx_plan::Fallback fb;
      fb.acc = (float2*)(base + fb_acc);
      fb.br0 = fb_box[g][0];
      fb.bc0 = fb_box[g][1];
      fb.brows = fb_box[g][2];
      fb.bcols = fb_box[g][3];
      ++g;
      fb.weight = problems[cls[kv.second[0]].members[0]].weight;
      for (int c : kv.second) {
        const ClassKey& K = cls[c].key;
        fb.adds.push_back(mftx_plan::Add{(const float2*)(base + cls[c].fbuf), K.M, K.N, K.r0, K.c0});
      }
      plan->fallbacks.push_back(fb);
    }
  }

  // ---- info: algorithmic FLOPs are 8 per complex MAC of the cheaper order -------
  double fl = 0, g1 = 0, g2 = 0;
  for (auto& c : cls)
    for (int i : c.members) {
      const mftx_problem& P = problems[i];
      const double m = (double)tiles[P.tile].m, n = (double)tiles[P.tile].n, M = P.M, N = P.N;
      fl += 8.0 * std::min(cost_row_first(m, n, M, N), cost_col_first(m, n, M, N));
      g1 += 8.0 * (c.col_first ? M * m * n : m * n * N);
      g2 += 8.0 * (c.col_first ? M * n * N : M * m * N);
    }
  plan->info.flops = fl;
  plan->info.gemm_flops[0] = g1;
  plan->info.gemm_flops[1] = g2;
  plan->info.mfma_flops[0] = mf1;
  plan->info.mfma_flops[1] = mf2;
  double bytes = 0;
  for (auto& ph : ops) bytes += 2.0 * (double)ph.wave_stride * ph.nwave;  // write + read
  for (auto& c : cls) bytes += 2.0 * packed_bytes(prec, c.rows2, c.ktot);
  bytes += 4.0 * image_rows * image_cols;
  plan->info.bytes_hbm_min = bytes;
  double bph = 0, bre = 0;
  for (auto& ph : ops)
    bph += (double)ph.m * ph.n * ((ph.amp_dtype == MFTX_DTYPE_F64 ? 8.0 : 4.0) + (ph.opd_dtype == MFTX_DTYPE_F64 ? 8.0 : 4.0)) +
           (double)ph.wave_stride * ph.nwave;
  for (auto& r : plan->reduces) {
    const double vr = (double)std::max<int64_t>(0, std::min<int64_t>(r.r0 + r.rows, image_rows) - std::max<int64_t>(r.r0, 0));
    const double vc = (double)std::max<int64_t>(0, std::min<int64_t>(r.c0 + r.cols, image_cols) - std::max<int64_t>(r.c0, 0));
    bre += 4.0 * r.ngroups * vr * vc + 4.0 * vr * vc;
  }
  plan->info.bytes_phasor = bph;
  plan->info.bytes_reduce = bre;
  plan->info.coherent_fast = fast ? 1 : 0;
  plan->info.n_stage1 = (int)p1.size();
  plan->info.n_stage2 = (int)cls.size();
  plan->info.n_work1 = plan->nwork1;
  plan->info.n_work2 = plan->nwork2 + plan->nwork2c;
  int nl = (ops.empty() ? 0 : 1) + 1 + (w2.empty() ? 0 : 1) + (w2c.empty() ? 0 : 1) + (int)plan->reduces.size();
  for (auto& fb : plan->fallbacks) nl += 2 + (int)fb.adds.size();  // clear, adds, |.|^2 into the image
  plan->info.n_launch = nl;
  *out = owned.release();
  return MFTX_OK;
  });
}

}  // extern "C"

static hipError_t launch_phasors(const mftx_plan* plan, hipStream_t s) {
  return launch_phasor_group(plan->prec, plan->phasor_ops, plan->n_phasor_ops, plan->phasor_threads, s);
}

static hipError_t launch_reduces(const mftx_plan* plan, void* image, int dtype, int64_t ldi, int accumulate,
                                 hipStream_t s) {
  const size_t esz = dtype == 3 ? 8 : 4;
  bool first = true;
  for (auto& r : plan->reduces) {
    int acc = accumulate;
    if (!accumulate && first && !(r.r0 == 0 && r.c0 == 0 && r.rows == plan->img_rows && r.cols == plan->img_cols)) {
      hipError_t e = launch_zero_rows(image, (int)esz, plan->img_rows, plan->img_cols, ldi, s);
      if (e != hipSuccess) return e;
      acc = 1;
    } else if (!accumulate && !first) {
      acc = 1;
    }
    first = false;
    hipError_t e = launch_reduce_partials(r.partial, r.gstride, r.ngroups, r.rows, r.cols, r.ldp, image, dtype, ldi,
                                          (int)plan->img_rows, (int)plan->img_cols, (int)r.r0, (int)r.c0, acc, s);
    if (e != hipSuccess) return e;
  }
  if (!accumulate && plan->reduces.empty()) return launch_zero_rows(image, (int)esz, plan->img_rows, plan->img_cols, ldi, s);
  return hipSuccess;
}

static hipError_t launch_fallbacks(const mftx_plan* plan, void* image, int dtype, int64_t ldi, int accumulate,
                                   hipStream_t s) {
  const size_t esz = dtype == 3 ? 8 : 4;
  hipError_t e = hipSuccess;
  if (!accumulate) e = launch_zero_rows(image, (int)esz, plan->img_rows, plan->img_cols, ldi, s);
  for (auto& fb : plan->fallbacks) {
    if (fb.brows <= 0 || fb.bcols <= 0) continue;  // every Field of the group lies off the image
    if (e == hipSuccess) e = launch_zero(fb.acc, sizeof(float2) * fb.brows * fb.bcols, s);
    for (auto& a : fb.adds) {
      // clip F's placement to the image (field.py:282-296); acc holds the group's union box
      const int64_t or0 = std::max<int64_t>(a.r0, 0), oc0 = std::max<int64_t>(a.c0, 0);
      const int64_t or1 = std::min<int64_t>(a.r0 + a.M, plan->img_rows), oc1 = std::min<int64_t>(a.c0 + a.N, plan->img_cols);
      if (or1 <= or0 || oc1 <= oc0 || e != hipSuccess) continue;
      e = launch_insert(a.F, a.N, (int)(or0 - a.r0), (int)(oc0 - a.c0), (int)(or1 - or0), (int)(oc1 - oc0), fb.acc, 0,
                        fb.bcols, (int)(or0 - fb.br0), (int)(oc0 - fb.bc0), 0, 1.0, s);
    }
    if (e == hipSuccess)
      e = launch_insert(fb.acc, fb.bcols, 0, 0, (int)fb.brows, (int)fb.bcols, image, dtype, ldi, (int)fb.br0,
                        (int)fb.bc0, 1, fb.weight, s);
  }
  return e;
}

// Stage 2: the intensity launch (classes alone in their coherent group) and
// the complex-store launch (the classes of several-class groups).
static hipError_t launch_stage2(const mftx_plan* plan, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (plan->nwork2)
    e = launch_gemm(plan->prec, EPI_INTENSITY, plan->bn, plan->works2, plan->nwork2, plan->probs2, plan->chunk_x, s);
  if (e == hipSuccess && plan->nwork2c)
    e = launch_gemm(plan->prec, EPI_STORE_CPLX, plan->bn, plan->works2c, plan->nwork2c, plan->probs2c, plan->chunk_x,
                    s);
  return e;
}

// The image: the reductions write (or clear and add) it, then the
// several-class groups add their |sum of fields|^2.
static hipError_t launch_finish(const mftx_plan* plan, void* image, int dtype, int64_t ldi, int accumulate,
                                hipStream_t s) {
  if (plan->fallbacks.empty()) return launch_reduces(plan, image, dtype, ldi, accumulate, s);
  hipError_t e = hipSuccess;
  if (!plan->reduces.empty()) {
    e = launch_reduces(plan, image, dtype, ldi, accumulate, s);
    accumulate = 1;
  }
  if (e == hipSuccess) e = launch_fallbacks(plan, image, dtype, ldi, accumulate, s);
  return e;
}

static int plan_launch(mftx_plan* plan, void* image, int dtype, int64_t ldi, int accumulate) {
  mftx_ctx* ctx = plan->ctx;
  hipStream_t s = ctx->stream;
  HIPCHK(ctx, launch_phasors(plan, s));
  HIPCHK(ctx, launch_gemm(plan->prec, EPI_STORE_PACKED_T, plan->bn, plan->works1, plan->nwork1, plan->probs1, plan->chunk_x, s));
  HIPCHK(ctx, launch_stage2(plan, s));
  HIPCHK(ctx, launch_finish(plan, image, dtype, ldi, accumulate, s));
  return MFTX_OK;
}

extern "C" {

int mftx_plan_execute(mftx_plan* plan, void* image, int32_t image_dtype, int64_t ldi, int32_t accumulate) {
  return guarded(plan ? plan->ctx : nullptr, [&]() -> int {
  if (!plan || !plan->ctx) return MFTX_EHANDLE;
  mftx_ctx* ctx = plan->ctx;
  if (!image) return fail(ctx, MFTX_EINVAL, "null image");
  if (image_dtype != 1 && image_dtype != 3) return fail(ctx, MFTX_EINVAL, "image must be float32 or float64");
  if (ldi < plan->img_cols) return fail(ctx, MFTX_EINVAL, "image row stride too small");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (!plan->use_graph) return plan_launch(plan, image, image_dtype, ldi, accumulate);
  if (!plan->gexec || plan->g_image != image || plan->g_dtype != image_dtype || plan->g_ld != ldi ||
      plan->g_accum != accumulate) {
    if (plan->gexec) (void)hipGraphExecDestroy(plan->gexec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    plan->gexec = nullptr;
    plan->graph = nullptr;
    hipStream_t cap = nullptr;
    HIPCHK(ctx, hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipStream_t saved = ctx->stream;
    ctx->stream = cap;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    int rc = e == hipSuccess ? plan_launch(plan, image, image_dtype, ldi, accumulate) : MFTX_OK;
    hipGraph_t g = nullptr;
    if (e == hipSuccess) e = hipStreamEndCapture(cap, &g);
    ctx->stream = saved;
    (void)hipStreamDestroy(cap);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hipfail(ctx, e, "graph capture");
    plan->graph = g;
    HIPCHK(ctx, hipGraphInstantiate(&plan->gexec, g, nullptr, nullptr, 0));
    plan->g_image = image;
    plan->g_dtype = image_dtype;
    plan->g_ld = ldi;
    plan->g_accum = accumulate;
  }
  HIPCHK(ctx, hipGraphLaunch(plan->gexec, ctx->stream));
  return MFTX_OK;
  });
}

// One execute's launch groups with a timing event before, between and after
// them (ev[0..4]) on stream s.
// The plan's launch groups in execute order (phasors, stage-1 GEMM, stage-2
// GEMM, reduction or fallback), the first `upto` of them; with `ev`, an event
// is recorded before the first and after each.
static hipError_t profile_launches(mftx_plan* plan, void* image, int dtype, int64_t ldi, hipStream_t s, int upto,
                                   const hipEvent_t* ev) {
  hipError_t e = ev ? hipEventRecord(ev[0], s) : hipSuccess;
  for (int k = 0; k < upto && e == hipSuccess; ++k) {
    if (k == 0) e = launch_phasors(plan, s);
    if (k == 1)
      e = launch_gemm(plan->prec, EPI_STORE_PACKED_T, plan->bn, plan->works1, plan->nwork1, plan->probs1, plan->chunk_x,
                      s);
    if (k == 2) e = launch_stage2(plan, s);
    if (k == 3) e = launch_finish(plan, image, dtype, ldi, 0, s);
    if (e == hipSuccess && ev) e = hipEventRecord(ev[k + 1], s);
  }
  return e;
}

// Graph mode of mftx_plan_profile: one graph per prefix of the launch groups
// (the first k of them, k = 1..4), each replayed `repeats` times back to back
// between two stream events, twice; group k's time is the difference of the
// mean replay times of prefixes k + 1 and k, so every launch runs as in a
// replayed execute, after its producer, with no event between launches.
// Returns the runtime's error when it cannot capture.
static hipError_t profile_graphs(mftx_plan* plan, void* image, int dtype, int64_t ldi, int repeats, double* ms) {
  hipStream_t s = plan->ctx->stream;
  hipStream_t cap = nullptr;
  hipGraphExec_t x[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  for (int k = 0; k < 4 && e == hipSuccess; ++k) {
    hipGraph_t g = nullptr;
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t l = profile_launches(plan, image, dtype, ldi, cap, k + 1, nullptr);
      e = hipStreamEndCapture(cap, &g);
      if (e == hipSuccess) e = l;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&x[k], g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
  }
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  // two full replays first: caches and clocks as in a run of executes
  for (int r = 0; r < 2 && e == hipSuccess; ++r) e = hipGraphLaunch(x[3], s);
  // prefixes timed in the order 1 2 3 4 4 3 2 1: a slow drift of the clock
  // over the measurement adds the same to both timings of every prefix
  double t[4] = {0, 0, 0, 0};
  for (int i = 0; i < 8 && e == hipSuccess; ++i) {
    const int k = i < 4 ? i : 7 - i;
    e = hipEventRecord(a, s);
    for (int r = 0; r < repeats && e == hipSuccess; ++r) e = hipGraphLaunch(x[k], s);
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float f = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, a, b);
    t[k] += 0.5 * (double)f / repeats;
  }
  for (auto& g : x)
    if (g) (void)hipGraphExecDestroy(g);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (cap) (void)hipStreamDestroy(cap);
  if (e != hipSuccess) return e;
  ms[0] = t[0];
  for (int k = 1; k < 4; ++k) ms[k] = t[k] - t[k - 1];
  return hipSuccess;
}

int mftx_plan_profile(mftx_plan* plan, void* image, int32_t image_dtype, int64_t ldi, int32_t repeats, double* ms) {
  if (!plan || !plan->ctx) return MFTX_EHANDLE;
  mftx_ctx* ctx = plan->ctx;
  if (!image || !ms || repeats <= 0 || repeats > 4096) return fail(ctx, MFTX_EINVAL, "bad profile request");
  if (image_dtype != 1 && image_dtype != 3) return fail(ctx, MFTX_EINVAL, "image must be float32 or float64");
  if (ldi < plan->img_cols) return fail(ctx, MFTX_EINVAL, "image row stride too small");
  return guarded(ctx, [&]() -> int {
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (plan->use_graph) {
    double g[4];
    if (profile_graphs(plan, image, image_dtype, ldi, repeats, g) == hipSuccess) {
      for (int k = 0; k < 4; ++k) ms[k] = g[k];
      ms[4] = 1.0;
      return MFTX_OK;
    }
    (void)hipGetLastError();  // the runtime refused a capture: time eagerly instead
  }
  // eager: every repeat back to back on the context stream with events
  // between the launch groups, one wait at the end
  hipStream_t s = ctx->stream;
  std::vector<hipEvent_t> ev(5 * (size_t)repeats, nullptr);
  hipError_t e = hipSuccess;
  for (auto& x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  for (int r = 0; r < repeats && e == hipSuccess; ++r)
    e = profile_launches(plan, image, image_dtype, ldi, s, 4, &ev[5 * (size_t)r]);
  if (e == hipSuccess) e = hipEventSynchronize(ev.back());
  double acc[4] = {0, 0, 0, 0};
  for (int r = 0; r < repeats && e == hipSuccess; ++r)
    for (int k = 0; k < 4 && e == hipSuccess; ++k) {
      float t = 0;
      e = hipEventElapsedTime(&t, ev[5 * (size_t)r + k], ev[5 * (size_t)r + k + 1]);
      acc[k] += t;
    }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  if (e != hipSuccess) return hipfail(ctx, e, "plan profile");
  for (int k = 0; k < 4; ++k) ms[k] = acc[k] / repeats;
  ms[4] = 0.0;
  return MFTX_OK;
  });
}

int mftx_plan_set_fields(mftx_plan* plan, void* const* dst, int64_t ldd, const void* const* wmap, int32_t wmap_dtype,
                         int64_t ldw, const double* scale) {
  if (!plan || !plan->ctx) return MFTX_EHANDLE;
  mftx_ctx* ctx = plan->ctx;
  return guarded(ctx, [&]() -> int {
    if (!dst || !scale) return fail(ctx, MFTX_EINVAL, "null argument");
    if (wmap && wmap_dtype != 1 && wmap_dtype != 3) return fail(ctx, MFTX_EINVAL, "weight map must be float32 or float64");
    // every argument is checked before any descriptor changes
    for (size_t c = 0; c < plan->pf.size(); ++c) {
      const Problem& q = plan->pf[c];
      const int i = plan->class_first[c];
      const int64_t N = q.trans ? q.rows : q.ncols;  // the class's F is M x N, row stride ldd
      if (!dst[i]) return fail(ctx, MFTX_EINVAL, "null field pointer for problem " + std::to_string(i));
      if (ldd < N || (wmap && wmap[i] && ldw < N)) return fail(ctx, MFTX_EINVAL, "field row stride too small");
      if (!std::isfinite(scale[i])) return fail(ctx, MFTX_EINVAL, "non-finite field scale");
    }
    for (size_t c = 0; c < plan->pf.size(); ++c) {
      Problem& q = plan->pf[c];
      const int i = plan->class_first[c];
      q.dst = dst[i];
      q.ldd = ldd;
      q.wmap = wmap ? wmap[i] : nullptr;
      q.wdtype = wmap_dtype;
      q.ldw = ldw;
      q.scale = (float)scale[i];
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (!plan->pf.empty())
      HIPCHK(ctx, hipMemcpy(plan->probs_f, plan->pf.data(), sizeof(Problem) * plan->pf.size(), hipMemcpyHostToDevice));
    plan->fields_ready = 1;
    return MFTX_OK;
  });
}

int mftx_plan_execute_fields(mftx_plan* plan) {
  if (!plan || !plan->ctx) return MFTX_EHANDLE;
  mftx_ctx* ctx = plan->ctx;
  if (!plan->fields_ready) return fail(ctx, MFTX_EINVAL, "mftx_plan_set_fields has not been called");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, launch_gemm(plan->prec, EPI_STORE_CPLX, plan->bn, plan->works_f, plan->nwork_f, plan->probs_f,
                          plan->chunk_x, ctx->stream));
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
