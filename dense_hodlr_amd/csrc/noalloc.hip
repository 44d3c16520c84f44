/*
 * The HODLR operator product Y = A_H X without any device allocation, for
 * callers that cannot allocate (code under stream capture, serving loops,
 * harnesses that forbid hipMalloc): every temporary of the product (the
 * descriptor arrays of the grouped GEMM launches, the split-K partials and
 * the projections T = V^T X of a level) lives in one file-scope device array
 * of NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.
 *
 * The array's name is a host-side stand-in in host code; its device address
 * is obtained once per device with hipGetSymbolAddress (with that device
 * current) and only that address is handed to copies and kernels.  Calls on
 * one GPU share the array: calls that overlap on one GPU (from several
 * threads or handles) race and must be serialized by the caller.
 */
#include <mutex>

#include "engine.h"

namespace hodlr {

static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

/* Device address of this GPU's scratch array (looked up on first use). */
char* scratch_device_ptr() {
  int dev = 0;
  HODLR_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HODLR_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  return g_scratch_ptr[dev];
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

void Engine::noalloc_setup() { (void)scratch_device_ptr(); }

/* Y = A_H X with the launches and arithmetic of matvec_device (same tiles,
 * same split-K plan, so the result is bit-identical), every temporary placed
 * in the scratch array: a level's projections at its start, then the
 * split-K partials, then the descriptors of one launch; launches on the
 * engine's stream reuse the regions in stream order (the descriptor copy of
 * a launch is queued after the kernels of the previous one). */
void Engine::matvec_noalloc(const double* dX, int ldx, int nrhs, double* dY, int ldy) {
  if (nrhs <= 0) return;
  if (!compressed_)
    throw std::logic_error("the allocation-free product needs a compressed operator "
                           "(call hodlr_compress first)");
  const int ld = ntot_;
  const bool mfma = tune_.gemm_mfma != 0;
  char* const base = scratch_device_ptr();
  // projections T = V^T X of the largest level
  size_t proj_max = 0;
  for (int d = 0; d < nlev_; d++) {
    size_t need = 0;
    for (int id : level_nodes_[d]) need += (size_t)(nodes_[id].kt + nodes_[id].kb) * nrhs;
    proj_max = std::max(proj_max, need);
  }
  const size_t proj_bytes = align256(proj_max * sizeof(double));
  auto too_large = [&](size_t need) {
    char msg[200];
    std::snprintf(msg, sizeof msg,
                  "the allocation-free product of %d right-hand sides needs %zu bytes of scratch; "
                  "the maximum is %zu (use fewer right-hand sides per call)",
                  nrhs, need, (size_t)NOALLOC_SCRATCH_BYTES);
    return std::invalid_argument(msg);
  };
  if (proj_bytes > NOALLOC_SCRATCH_BYTES) throw too_large(proj_bytes);
  double* const proj = reinterpret_cast<double*>(base);
  struct Launch {
    std::vector<GemmDesc> g;
    std::vector<int> ids;
    int ntiles = 0, bm = 64;
    size_t ws = 0;
  };
  std::vector<Launch> launches;
  auto add = [&](std::vector<GemmDesc>&& g) {
    Launch L;
    L.g = std::move(g);
    L.bm = tune_.gemm_bm ? tune_.gemm_bm : gemm_tile_m(L.g, mfma);
    L.ws = gemm_plan(L.g, L.ids, L.ntiles, tune_.gemm_target, L.bm, mfma);
    launches.push_back(std::move(L));
  };
  {  // leaves: Y = L X
    std::vector<GemmDesc> g;
    for (auto& lf : leaves_)
      g.push_back({lf.s, nrhs, lf.s, 0, lf.s, ldx, ldy, 1, 0, 0, 1.0, 0.0,
                   Lorig_.p + lf.mat_off, dX + lf.row0, dY + lf.row0});
    add(std::move(g));
  }
  for (int d = 0; d < nlev_; d++) {  // per level: T = V^T X, then Y += U T
    if (level_nodes_[d].empty()) continue;
    std::vector<GemmDesc> g1, g2;
    size_t o = 0;
    for (int id : level_nodes_[d]) {
      const TreeNodeH& nd = nodes_[id];
      const int kt = nd.kt, kb = nd.kb;
      const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
      const int top = nd.row0, bot = nd.row0 + st;
      double* tt = proj + o;
      double* tb = tt + (size_t)kt * nrhs;
      o += (size_t)(kt + kb) * nrhs;
      g1.push_back({kt, nrhs, sb, 1, ld, ldx, kt, 1, 0, 0, 1.0, 0.0, V_[d].p + bot, dX + bot, tt});
      g1.push_back({kb, nrhs, st, 1, ld, ldx, kb, 1, 0, 0, 1.0, 0.0, V_[d].p + top, dX + top, tb});
      g2.push_back({st, nrhs, kt, 0, ld, kt, ldy, 1, 0, 0, 1.0, 1.0, U_[d].p + top, tt, dY + top});
      g2.push_back({sb, nrhs, kb, 0, ld, kb, ldy, 1, 0, 0, 1.0, 1.0, U_[d].p + bot, tb, dY + bot});
    }
    add(std::move(g1));
    add(std::move(g2));
  }
  size_t ws_max = 0, desc_max = 0;
  for (auto& L : launches) {
    ws_max = std::max(ws_max, L.ws);
    desc_max = std::max(desc_max, align256(L.g.size() * sizeof(GemmDesc)) +
                                      align256(L.ids.size() * sizeof(int)));
  }
  const size_t ws_bytes = align256(ws_max * sizeof(double));
  const size_t total = proj_bytes + ws_bytes + desc_max;
  if (total > NOALLOC_SCRATCH_BYTES) throw too_large(total);
  double* const wsp = reinterpret_cast<double*>(base + proj_bytes);
  char* const dregion = base + proj_bytes + ws_bytes;
  for (auto& L : launches) {
    if (L.g.empty() || L.ntiles == 0) continue;
    const size_t gb = L.g.size() * sizeof(GemmDesc);
    HODLR_HIP_CHECK(hipMemcpyAsync(dregion, L.g.data(), gb, hipMemcpyHostToDevice, stream_));
    int* ids = nullptr;
    if (!L.ids.empty()) {
      ids = reinterpret_cast<int*>(dregion + align256(gb));
      HODLR_HIP_CHECK(hipMemcpyAsync(ids, L.ids.data(), L.ids.size() * sizeof(int),
                                     hipMemcpyHostToDevice, stream_));
    }
    launch_gemm<double>(reinterpret_cast<const GemmDesc*>(dregion), (int)L.g.size(), L.ntiles,
                        wsp, ids, (int)L.ids.size(), L.bm, stream_, mfma);
  }
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));  // (the host descriptor vectors end here)
}

}  // namespace hodlr
