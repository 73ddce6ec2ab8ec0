// Host side of the RPD engine: implements include/rpd_capi.h.
//
// Replaces the device orchestration of compute_clipped_voro_diagram_GPU
// (src/rpd3d/voronoi.cu:455-795): instead of per-call cudaMalloc/free, dense
// n_site x n_vert / n_site x n_tet matrices and a D2H of one 3.4 KB record
// per (tet, slot), the context keeps device buffers across calls, gathers
// candidates sparsely on the device and compacts only emitted cells.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "rpd_capi.h"
#include "rpd_device.h"
#include "rpd_prep.h"
#include "../common/rpd_classgrid.h"
#include "rpd_kernels.hip"
#include "rpd_learn.hip"
#include "rpd_nbr.hip"
#include "rpd_cells.hip"

// Device scratch of rpd_run_noalloc, the entry point for callers that
// cannot allocate: every buffer of one upload + plain run + compaction is
// carved from it.  Each device holds its own copy of the array, and host
// code reaches it only through the address hipGetSymbolAddress returns for
// that device (taken once per device when a context is created, kept in
// g_scratch_dev); the array's name is not a device address on the host.
constexpr size_t kScratchBytes = (size_t)RPD_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

using namespace rpd;

struct Buf;

// Bump allocator over the scratch for the duration of one rpd_run_noalloc
// call: Bufs that ask for memory get a block of it; their own blocks are
// remembered and put back when the call ends.  Nothing is freed or
// allocated while it is active.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0, high = 0;
  std::vector<std::pair<Buf*, std::pair<void*, size_t>>> saved;
  bool has(const Buf* b) const {
    for (const auto& x : saved)
      if (x.first == b) return true;
    return false;
  }
  void touch(Buf* b);
  bool exhausted = false;
  void* take(size_t bytes) {
    const size_t b = (std::max<size_t>(bytes, 256) + 255) & ~(size_t)255;
    if (used + b > cap) {
      exhausted = true;
      return nullptr;
    }
    void* q = base + used;
    used += b;
    high = std::max(high, used);
    return q;
  }
  void restore();
};
thread_local Arena* t_arena = nullptr;
// set while a stream capture records the pipeline: nothing may allocate,
// free or wait on the host then
thread_local bool t_capture = false;
// bumped whenever a buffer moves (a captured graph holds addresses)
std::atomic<uint64_t> g_buf_generation{0};

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
  Buf() = default;
  Buf(const Buf&) = delete;  // owns its block
  Buf& operator=(const Buf&) = delete;
  void swap(Buf& o) {
    if (t_arena) {  // only scratch blocks change hands inside an arena
      for (Buf* b : {this, &o})
        if (!t_arena->has(b)) {
          t_arena->touch(b);
          b->p = nullptr;
          b->cap = 0;
        }
    }
    std::swap(p, o.p);
    std::swap(cap, o.cap);
  }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  // in an arena: a scratch block (keeping the first `keep` bytes of the
  // current one, copied on stream s), never the Buf's own block
  hipError_t arena_need(size_t bytes, size_t keep, hipStream_t s) {
    Arena& A = *t_arena;
    const bool mine = A.has(this);
    if (mine && bytes <= cap && p) return hipSuccess;
    void* q = A.take(bytes);
    if (!q) return hipErrorOutOfMemory;
    if (mine && p && keep) {
      hipError_t e = hipMemcpyAsync(q, p, std::min(keep, cap),
                                    hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
    }
    A.touch(this);
    p = q;
    cap = (std::max<size_t>(bytes, 256) + 255) & ~(size_t)255;
    return hipSuccess;
  }
  hipError_t need(size_t bytes) {
    if (t_arena) return arena_need(bytes, 0, nullptr);
    if (bytes <= cap && p) return hipSuccess;
    if (t_capture) return hipErrorStreamCaptureUnsupported;
    g_buf_generation.fetch_add(1);
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t b = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) cap = b;
    return e;
  }
  // grow to `bytes` keeping the first `keep` bytes (stream-ordered copy;
  // the old block is freed after the stream has drained)
  hipError_t need_keep(size_t bytes, size_t keep, hipStream_t s) {
    if (t_arena) return arena_need(bytes, keep, s);
    if (bytes <= cap && p) return hipSuccess;
    if (t_capture) return hipErrorStreamCaptureUnsupported;
    g_buf_generation.fetch_add(1);
    size_t b = std::max<size_t>(bytes + bytes / 4, 256);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, b);
    if (e != hipSuccess) return e;
    if (p && keep) {
      e = hipMemcpyAsync(q, p, std::min(keep, cap), hipMemcpyDeviceToDevice,
                         s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        (void)hipFree(q);
        return e;
      }
    }
    if (p) (void)hipFree(p);
    p = q;
    cap = b;
    return hipSuccess;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

void Arena::touch(Buf* b) {
  if (!has(b)) saved.push_back({b, {b->p, b->cap}});
}
void Arena::restore() {
  for (auto& x : saved) {
    x.first->p = x.second.first;
    x.first->cap = x.second.second;
  }
  saved.clear();
}

struct TierBufs {
  Buf hdr, words, geom, areas;
  SlotView view(int V, int P, int E) const {
    SlotView s;
    s.stride_words = slot_stride_words(V, P, E);
    s.stride_areas = P;
    s.hdr = hdr.as<uint32_t>();
    s.words = words.as<uint32_t>();
    s.geom = geom.as<double>();
    s.areas = areas.as<double>();
    return s;
  }
  hipError_t need(int64_t n, int V, int P, int E) {
    hipError_t e;
    n = std::max<int64_t>(n, 1);
    if ((e = hdr.need(n * 4)) != hipSuccess) return e;
    if ((e = words.need(n * slot_stride_words(V, P, E) * 4)) != hipSuccess)
      return e;
    if ((e = geom.need(n * 4 * 8)) != hipSuccess) return e;
    return areas.need(n * P * 8);
  }
  hipError_t need_keep(int64_t n, int64_t keep, int V, int P, int E,
                       hipStream_t s) {
    hipError_t e;
    n = std::max<int64_t>(n, 1);
    if ((e = hdr.need_keep(n * 4, keep * 4, s)) != hipSuccess) return e;
    const int64_t sw = slot_stride_words(V, P, E);
    if ((e = words.need_keep(n * sw * 4, keep * sw * 4, s)) != hipSuccess)
      return e;
    if ((e = geom.need_keep(n * 32, keep * 32, s)) != hipSuccess) return e;
    return areas.need_keep(n * P * 8, keep * P * 8, s);
  }
};

// sizes of one plain run: sphere-bin registrations, candidates, middle /
// full / exact re-run entries, output stream
struct RunCaps {
  int64_t n_reg = 0, n_cand = 0, n_mid = 0, n_full = 0, n_exact = 0;
  int64_t cells = 0, words = 0, areas = 0;
  // learned path: closure rounds run; the largest miss list, unique pairs
  // and new candidates of a round; the merged pairs, tets re-run on the
  // full relation, their candidates and the output candidates; the
  // re-run tiers of the closure and fallback clips (largest launch)
  int64_t rounds = 0, miss = 0, uniq = 0, added = 0, ext = 0;
  int64_t flagged = 0, fb = 0, incl = 0;
  int64_t c_mid = 0, c_full = 0, c_exact = 0;
  // per round: the miss list read at round k (k = the closure round, up to
  // the last check), the round's new candidates, the merged pairs after it
  int64_t miss_r[8] = {}, added_r[8] = {}, ext_r[8] = {};
};

}  // namespace

struct rpd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // resident batch of instances (one instance = a batch of one), stored
  // back to back: vertex, tet and site ids are batch-wide
  struct Inst {
    int32_t n_vert = 0, n_tet = 0, n_site = 0;
    int64_t vert_off = 0, tet_off = 0, site_off = 0;
    double dom[6] = {0, 0, 0, 1, 1, 1};
    double rho = 0;
  };
  std::vector<Inst> inst;
  int32_t n_vert = 0, n_tet = 0, n_site = 0;
  bool has_mesh = false, has_sites = false;
  Buf vert, tet_v, tet_fid, tet_meta, tet_inst, vert_inst, inst_dev;
  Buf site, flags, knn_off, knn_ids, knn_near, knn_site_near, site_inst;
  Buf knn_bis;
  Buf knn_bisf;
  // pipeline
  Buf bbox, reg_n, reg_off, keys_a, keys_b, bin_start, bin_ids;
  Buf tet_count, tet_off, tet_pass, cand, rel_ids;
  Buf counters;  // int32[8]: fallback, n_rerun_large, n_rerun_exact, ...,
                 // [6] n_rerun_full (after the middle tier)
  Buf rerun_large, rerun_full, rerun_exact, loc, sizes, offs, drops, stats64,
      work, rel_work;
  Buf scan_tmp;
  Buf trip_k, trip_k2, trip_v, trip_v2;  // step-count order of a clip range
  TierBufs small, large, exact, mid;
  Buf out_cells, out_words, out_geom, out_areas;
  bool last_geometry = false;  // the last run computed cell geometry
  bool has_results = false;    // the last run's cells belong to the
                               // resident instance and sites
  int64_t n_knn = 0;
  // tet subset of the partial RPD (one byte per tet, 1 = clipped)
  Buf tet_sel;
  bool has_subset = false;
  // per-site centroids of the last run's cells
  Buf cs_keys_a, cs_keys_b, cs_vals_a, cs_vals_b, cs_start, cs_out;
  // power-cell adjacency of the last run's cells
  Buf adj_act, adj_euler, adj_cnt, adj_off, adj_nbr, adj_cpar, adj_err;
  Buf tf_ka, tf_kb, tf_va, tf_vb;
  Buf fac_ka, fac_kb, fac_va, fac_vb, fac_par, fac_eul, fac_flag, fac_pos,
      fac_start, fac_runs, fac_reul, fac_rec;
  Buf edg_ka, edg_kb, edg_va, edg_vb, edg_par, edg_flag, edg_pos, edg_start,
      edg_runs, edg_rec;
  Buf site_cells, site_cc, site_eul, adj_soff, adj_eul_s, adj_root_s;
  rpd_adj_sizes adj{};
  bool adj_ready = false;
  RunCaps seen{};  // sizes of the last eager plain run
  // the captured plain pipeline of rpd_graph_build, its capacities and the
  // state it was captured against
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  RunCaps graph_cap{};
  rpd_run_cfg graph_cfg{};
  double graph_headroom = 0;
  std::vector<int64_t> graph_sig;
  // learned proposals: features, model (fp32 master, Adam state, bf16
  // copies), certificate and fallback
  Buf knn, feat, sel, omask, vowner, sg_keys_a, sg_keys_b, sg_start, sg_ids;
  int32_t feat_n = 0;
  Buf mlp_master, mlp_m1, mlp_m2, mlp_w1t, mlp_w2t, mlp_w3t, mlp_w2, mlp_w3,
      mlp_bias;
  bool mlp_ready = false;
  int64_t adam_step = 0;
  Buf lab_y, lab_m;
  int32_t lab_n = 0;
  Buf tr_idx, tr_xb, tr_h1, tr_h2, tr_dz1, tr_dz2, tr_dz3, tr_loss, tr_p1,
      tr_p2, tr_p3, tr_pb, tr_grad, tr_x, tr_y, tr_mk;
  Buf flag, has_cell, flag32, flag_pos, flag_list, brute_n, brute_off,
      vals_a, vals_b, perm, miss_a, miss_b, umark, upos, ukeys, upass, ppos,
      mem_a, mem_b, reopened, out_off, out_cnt, out_ids, reop64;
  // a captured learned run: device counts of its ranges, and where its
  // final offsets land (eager learned runs swap the offset buffers)
  Buf gcnt;
  void* graph_tet_off = nullptr;
  rpd_stats st;
  int tuned_G = 0, tuned_caps = 0;  // rpd_tune's choice (0: not tuned)
  // ev[0..15]: pipeline phases; ev[16 + 2k], ev[17 + 2k]: around the k-th
  // small-tier clip launch of a run (up to kSmallEv)
  hipEvent_t ev[48];
  bool events = false;
  int n_small_ev = 0;
  // host waits: an event polled by wait_stream and a pinned block that
  // receives the counts read back between pipeline phases
  hipEvent_t ev_sync = nullptr;
  char* pin = nullptr;
  // pinned staging of the per-run instance table; ev_inst marks the end of
  // its last copy to the device
  void* pin_inst = nullptr;
  size_t pin_inst_cap = 0;
  hipEvent_t ev_inst = nullptr;
  bool inst_copy_recorded = false;
  int32_t n_reads = 0;  // read_back waits since the run began
  // this device's copy of the no-alloc scratch (nullptr if its address
  // could not be taken)
  void* scratch_dev = nullptr;
};

namespace {

// the scratch's address on each device, taken once per device
std::mutex g_scratch_mu;
void* g_scratch_dev[64] = {};
bool g_scratch_tried[64] = {};

// per-run instance table staging allocated with the context, so that a
// single-instance run never allocates pinned memory
constexpr size_t kPinInstBytes = 64 << 10;

int fail(rpd_ctx* c, const std::string& m, int code = -1) {
  if (c) c->err = m;
  return code;
}

#define HIP_TRY(expr)                                                  \
  do {                                                                 \
    hipError_t _e = (expr);                                            \
    if (_e != hipSuccess)                                              \
      return fail(ctx, std::string(#expr) + ": " + hipGetErrorString(_e), \
                  -2);                                                 \
  } while (0)

int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Waits until the stream has drained by polling an event.  The pipeline
// reads counts back between its phases to size the next launches; a
// blocking wait adds its wake-up latency to each of these round trips,
// polling returns as soon as the event has completed.
hipError_t wait_stream(rpd_ctx* ctx, hipStream_t s) {
  hipError_t e = hipEventRecord(ctx->ev_sync, s);
  if (e != hipSuccess) return e;
  while ((e = hipEventQuery(ctx->ev_sync)) == hipErrorNotReady) {
  }
  return e;
}

// Small device-to-host reads (counts, totals): copied into the context's
// pinned block, the stream waited on, then handed to the host variables.
constexpr size_t kPinBytes = 8192;
struct D2H {
  void* dst;
  const void* src;
  size_t n;
};
hipError_t read_back(rpd_ctx* ctx, hipStream_t s,
                     const std::vector<D2H>& reads) {
  if (t_capture) return hipErrorStreamCaptureUnsupported;
  ++ctx->n_reads;
  size_t off = 0;
  for (const D2H& r : reads) {
    if (off + r.n > kPinBytes) return hipErrorInvalidValue;
    hipError_t e = hipMemcpyAsync(ctx->pin + off, r.src, r.n,
                                  hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    off += (r.n + 15) & ~(size_t)15;
  }
  hipError_t e = wait_stream(ctx, s);
  if (e != hipSuccess) return e;
  off = 0;
  for (const D2H& r : reads) {
    std::memcpy(r.dst, ctx->pin + off, r.n);
    off += (r.n + 15) & ~(size_t)15;
  }
  return hipSuccess;
}
hipError_t read_back(rpd_ctx* ctx, hipStream_t s,
                     std::initializer_list<D2H> reads) {
  return read_back(ctx, s, std::vector<D2H>(reads));
}

template <class T>
hipError_t exclusive_sum(rpd_ctx* ctx, const T* in, T* out, int64_t n) {
  size_t tmp = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out,
                                                  (int)n, ctx->stream);
  if (e != hipSuccess) return e;
  if ((e = ctx->scan_tmp.need(tmp)) != hipSuccess) return e;
  return hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, tmp, in, out,
                                          (int)n, ctx->stream);
}

// zero n int32 words with a kernel (a memset node in a captured graph is
// avoided)
hipError_t zero32(hipStream_t s, void* p, int64_t n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill32, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                     static_cast<int32_t*>(p), n, 0);
  return hipGetLastError();
}

}  // namespace

extern "C" {

int rpd_ctx_create(int32_t device, rpd_ctx** out) {
  if (!out) return -1;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return -3;
  if (device < 0 || device >= n) return -1;
  if (hipSetDevice(device) != hipSuccess) return -3;
  auto* ctx = new rpd_ctx();
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) !=
      hipSuccess) {
    delete ctx;
    return -3;
  }
  for (int i = 0; i < 48; ++i)
    if (hipEventCreate(&ctx->ev[i]) != hipSuccess) {
      for (int k = 0; k < i; ++k) (void)hipEventDestroy(ctx->ev[k]);
      (void)hipStreamDestroy(ctx->stream);
      delete ctx;
      return -3;
    }
  ctx->events = true;
  if (hipEventCreateWithFlags(&ctx->ev_sync, hipEventDisableTiming) !=
          hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_inst, hipEventDisableTiming) !=
          hipSuccess ||
      hipHostMalloc((void**)&ctx->pin, kPinBytes, hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc(&ctx->pin_inst, kPinInstBytes, hipHostMallocDefault) !=
          hipSuccess) {
    rpd_ctx_destroy(ctx);
    return -3;
  }
  ctx->pin_inst_cap = kPinInstBytes;
  if (device < 64) {  // the scratch's address on this device, once
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_tried[device]) {
      g_scratch_tried[device] = true;
      void* ptr = nullptr;
      if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess)
        g_scratch_dev[device] = ptr;
    }
    ctx->scratch_dev = g_scratch_dev[device];
  }
  std::memset(&ctx->st, 0, sizeof(ctx->st));
  *out = ctx;
  return 0;
}

void rpd_ctx_destroy(rpd_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->events)
    for (int i = 0; i < 48; ++i) (void)hipEventDestroy(ctx->ev[i]);
  if (ctx->ev_sync) (void)hipEventDestroy(ctx->ev_sync);
  if (ctx->pin) (void)hipHostFree(ctx->pin);
  if (ctx->pin_inst) (void)hipHostFree(ctx->pin_inst);
  if (ctx->ev_inst) (void)hipEventDestroy(ctx->ev_inst);
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int32_t rpd_current_device(void) {
  int d = -1;
  return hipGetDevice(&d) == hipSuccess ? (int32_t)d : -1;
}

const char* rpd_last_error(const rpd_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

void* rpd_ctx_stream(rpd_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int rpd_update_sites_batch(rpd_ctx* ctx, int32_t n_inst,
                           const rpd_sites_desc* sd) {
  if (!ctx) return -1;
  if (!ctx->has_mesh) return fail(ctx, "rpd_update_sites: no resident mesh");
  if (n_inst != (int32_t)ctx->inst.size() || (n_inst > 0 && !sd))
    return fail(ctx, "rpd_update_sites: one site set per resident instance");
  HIP_TRY(hipSetDevice(ctx->device));
  int64_t ns_tot = 0, nk_tot = 0;
  for (int32_t b = 0; b < n_inst; ++b) {
    const rpd_sites_desc& d = sd[b];
    if (d.n_site < 0 ||
        (d.n_site > 0 && (!d.sites || !d.flags || !d.knn_off)))
      return fail(ctx, "rpd_update_sites: bad arguments");
    const int64_t n_knn = d.n_site ? d.knn_off[d.n_site] : 0;
    if (d.n_site && d.knn_off[0] != 0)
      return fail(ctx, "rpd_update_sites: knn_off[0] must be 0");
    for (int32_t i = 0; i < d.n_site; ++i)
      if (d.knn_off[i + 1] < d.knn_off[i])
        return fail(ctx, "rpd_update_sites: knn_off not monotone");
    if (n_knn && !d.knn_ids)
      return fail(ctx, "rpd_update_sites: missing knn_ids");
    for (int64_t q = 0; q < n_knn; ++q)
      if (d.knn_ids[q] < 0 || d.knn_ids[q] >= d.n_site)
        return fail(ctx, "rpd_update_sites: neighbour id out of range");
    ns_tot += d.n_site;
    nk_tot += n_knn;
  }
  if (ns_tot >= (1LL << 31) || nk_tot >= (1LL << 31))
    return fail(ctx, "rpd_update_sites: batch too large");
  std::vector<double> site(4 * std::max<int64_t>(ns_tot, 1));
  std::vector<uint32_t> flags(std::max<int64_t>(ns_tot, 1));
  std::vector<int64_t> koff(ns_tot + 1);
  std::vector<int32_t> kids(std::max<int64_t>(nk_tot, 1));
  std::vector<int32_t> sinst(std::max<int64_t>(ns_tot, 1));
  int64_t so = 0, ko = 0;
  int32_t max_deg = 0;
  for (int32_t b = 0; b < n_inst; ++b) {
    const rpd_sites_desc& d = sd[b];
    ctx->inst[b].n_site = d.n_site;
    ctx->inst[b].site_off = so;
    if (d.n_site) std::memcpy(&site[4 * so], d.sites, (size_t)d.n_site * 32);
    for (int32_t i = 0; i < d.n_site; ++i) {
      flags[so + i] = d.flags[i];
      sinst[so + i] = b;
      koff[so + i] = ko + d.knn_off[i];
      max_deg = std::max<int32_t>(max_deg,
                                  (int32_t)(d.knn_off[i + 1] - d.knn_off[i]));
    }
    const int64_t n_knn = d.n_site ? d.knn_off[d.n_site] : 0;
    for (int64_t q = 0; q < n_knn; ++q)
      kids[ko + q] = (int32_t)(d.knn_ids[q] + so);
    so += d.n_site;
    ko += n_knn;
  }
  koff[ns_tot] = ko;
  ctx->n_site = (int32_t)ns_tot;
  ctx->n_knn = nk_tot;
  HIP_TRY(ctx->site.need(std::max<int64_t>(ns_tot, 1) * 32));
  HIP_TRY(ctx->flags.need(std::max<int64_t>(ns_tot, 1) * 4));
  HIP_TRY(ctx->site_inst.need(std::max<int64_t>(ns_tot, 1) * 4));
  HIP_TRY(ctx->knn_off.need((ns_tot + 1) * 8));
  HIP_TRY(ctx->knn_ids.need(std::max<int64_t>(nk_tot, 1) * 4));
  HIP_TRY(ctx->knn_near.need(std::max<int64_t>(nk_tot, 1) * 4));
  HIP_TRY(ctx->knn_site_near.need(std::max<int64_t>(nk_tot, 1) * 32));
  if (ns_tot) {
    HIP_TRY(hipMemcpyAsync(ctx->site.p, site.data(), (size_t)ns_tot * 32,
                           hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->flags.p, flags.data(), (size_t)ns_tot * 4,
                           hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->site_inst.p, sinst.data(),
                           (size_t)ns_tot * 4, hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->knn_off.p, koff.data(),
                           (size_t)(ns_tot + 1) * 8, hipMemcpyHostToDevice,
                           ctx->stream));
    if (nk_tot)
      HIP_TRY(hipMemcpyAsync(ctx->knn_ids.p, kids.data(), (size_t)nk_tot * 4,
                             hipMemcpyHostToDevice, ctx->stream));
  }
  ctx->st.site_k = max_deg;
  HIP_TRY(wait_stream(ctx, ctx->stream));
  ctx->has_sites = true;
  ctx->last_geometry = false;  // the last run's cells belong to other sites
  ctx->has_results = false;
  ctx->adj_ready = false;
  return 0;
}

int rpd_update_sites(rpd_ctx* ctx, int32_t n_site, const double* sites,
                     const uint32_t* flags, const int64_t* knn_off,
                     const int32_t* knn_ids) {
  if (!ctx) return -1;
  if (ctx->inst.size() != 1)
    return fail(ctx, "rpd_update_sites: the resident batch has more than "
                     "one instance; use rpd_update_sites_batch");
  const rpd_sites_desc d{n_site, sites, flags, knn_off, knn_ids};
  return rpd_update_sites_batch(ctx, 1, &d);
}

int rpd_upload_batch(rpd_ctx* ctx, int32_t n_inst,
                     const rpd_instance_desc* descs) {
  if (!ctx || n_inst <= 0 || !descs) return -1;
  HIP_TRY(hipSetDevice(ctx->device));
  int64_t nv_tot = 0, nt_tot = 0;
  for (int32_t b = 0; b < n_inst; ++b) {
    const rpd_instance_desc* d = descs + b;
    if (d->n_vert <= 0 || d->n_tet <= 0 || !d->verts || !d->tets ||
        !d->v_adj || !d->e_adj || !d->f_adj || !d->f_id)
      return fail(ctx, "rpd_upload: missing mesh arrays");
    for (int64_t q = 0; q < 4LL * d->n_tet; ++q)
      if (d->tets[q] < 0 || d->tets[q] >= d->n_vert)
        return fail(ctx, "rpd_upload: tet vertex id out of range");
    nv_tot += d->n_vert;
    nt_tot += d->n_tet;
  }
  if (nv_tot >= (1LL << 31) || nt_tot >= (1LL << 31))
    return fail(ctx, "rpd_upload: batch too large");
  std::vector<double> v4(4 * nv_tot);
  std::vector<int32_t> tv_all(4 * nt_tot), fid_all(4 * nt_tot),
      tinst(nt_tot), vinst(nv_tot);
  std::vector<TetMeta> meta(nt_tot);
  std::vector<rpd_ctx::Inst> inst(n_inst);
  int64_t vo = 0, to = 0;
  for (int32_t b = 0; b < n_inst; ++b) {
    const rpd_instance_desc* d = descs + b;
    const int32_t nv = d->n_vert, nt = d->n_tet;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < nv; ++i) {
      for (int a = 0; a < 3; ++a) {
        double x = d->verts[3 * i + a];
        if (!std::isfinite(x))
          return fail(ctx, "rpd_upload: non-finite vertex");
        v4[4 * (vo + i) + a] = x;
        lo[a] = std::min(lo[a], x);
        hi[a] = std::max(hi[a], x);
      }
      v4[4 * (vo + i) + 3] = 0.0;
      vinst[vo + i] = b;
    }
    double rho = 0;
    for (int64_t t = 0; t < nt; ++t) {
      const int32_t* tv = d->tets + 4 * t;
      uint32_t f = 0, va = 0, e0 = 0, e1 = 0;
      for (int k = 0; k < 4; ++k) {
        f |= ((uint32_t)d->f_adj[4 * t + k] & 255u) << (8 * k);
        va |= ((uint32_t)d->v_adj[tv[k]] & 255u) << (8 * k);
        tv_all[4 * (to + t) + k] = (int32_t)(tv[k] + vo);
        fid_all[4 * (to + t) + k] = d->f_id[4 * t + k];
      }
      for (int k = 0; k < 4; ++k)
        e0 |= ((uint32_t)d->e_adj[6 * t + k] & 255u) << (8 * k);
      for (int k = 0; k < 2; ++k)
        e1 |= ((uint32_t)d->e_adj[6 * t + 4 + k] & 255u) << (8 * k);
      meta[to + t] = TetMeta{f, va, e0, e1};
      tinst[to + t] = b;
      double c[3];
      for (int a = 0; a < 3; ++a)
        c[a] = (((d->verts[3 * tv[0] + a] + d->verts[3 * tv[1] + a]) +
                 d->verts[3 * tv[2] + a]) +
                d->verts[3 * tv[3] + a]) *
               0.25;
      for (int k = 0; k < 4; ++k) {
        double s2 = 0;
        for (int a = 0; a < 3; ++a) {
          double x = d->verts[3 * tv[k] + a] - c[a];
          s2 += x * x;
        }
        rho = std::max(rho, std::sqrt(s2));
      }
    }
    double ext =
        std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    rpd_ctx::Inst& I = inst[b];
    I.n_vert = nv;
    I.n_tet = nt;
    I.vert_off = vo;
    I.tet_off = to;
    I.rho = rho * (1 + 1e-9) + 1e-9 * ext;
    for (int a = 0; a < 3; ++a) {
      I.dom[a] = lo[a];
      I.dom[3 + a] = hi[a];
    }
    vo += nv;
    to += nt;
  }
  ctx->inst = inst;
  ctx->n_vert = (int32_t)nv_tot;
  ctx->n_tet = (int32_t)nt_tot;
  ctx->has_sites = false;
  ctx->has_subset = false;
  ctx->has_results = false;
  ctx->adj_ready = false;
  HIP_TRY(ctx->vert.need((size_t)nv_tot * 32));
  HIP_TRY(ctx->tet_v.need((size_t)nt_tot * 16));
  HIP_TRY(ctx->tet_fid.need((size_t)nt_tot * 16));
  HIP_TRY(ctx->tet_meta.need((size_t)nt_tot * 16));
  HIP_TRY(ctx->tet_inst.need((size_t)nt_tot * 4));
  HIP_TRY(ctx->vert_inst.need((size_t)nv_tot * 4));
  HIP_TRY(hipMemcpyAsync(ctx->vert.p, v4.data(), (size_t)nv_tot * 32,
                         hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(ctx->tet_v.p, tv_all.data(), (size_t)nt_tot * 16,
                         hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(ctx->tet_fid.p, fid_all.data(), (size_t)nt_tot * 16,
                         hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(ctx->tet_meta.p, meta.data(), (size_t)nt_tot * 16,
                         hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(ctx->tet_inst.p, tinst.data(), (size_t)nt_tot * 4,
                         hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(ctx->vert_inst.p, vinst.data(), (size_t)nv_tot * 4,
                         hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(wait_stream(ctx, ctx->stream));
  ctx->has_mesh = true;
  bool all_sites = true;
  for (int32_t b = 0; b < n_inst; ++b)
    all_sites = all_sites && (descs[b].sites || descs[b].n_site == 0);
  if (!all_sites) return 0;
  std::vector<rpd_sites_desc> sd(n_inst);
  for (int32_t b = 0; b < n_inst; ++b)
    sd[b] = rpd_sites_desc{descs[b].n_site, descs[b].sites, descs[b].flags,
                           descs[b].knn_off, descs[b].knn_ids};
  return rpd_update_sites_batch(ctx, n_inst, sd.data());
}

int rpd_upload(rpd_ctx* ctx, const rpd_instance_desc* d) {
  if (!ctx || !d) return -1;
  return rpd_upload_batch(ctx, 1, d);
}

int rpd_batch_offsets(const rpd_ctx* ctx, int64_t* tet_off,
                      int64_t* site_off) {
  if (!ctx) return -1;
  const size_t n = ctx->inst.size();
  for (size_t b = 0; b <= n; ++b) {
    if (tet_off)
      tet_off[b] = b < n ? ctx->inst[b].tet_off : (int64_t)ctx->n_tet;
    if (site_off)
      site_off[b] = b < n ? ctx->inst[b].site_off : (int64_t)ctx->n_site;
  }
  return 0;
}

int32_t rpd_batch_size(const rpd_ctx* ctx) {
  return ctx ? (int32_t)ctx->inst.size() : 0;
}

}  // extern "C"

namespace {
// capacities of the small tier: 0 = SmallCaps, 1 = WideCaps, 2 = TightCaps,
// 3 = MidCaps (2 and 3 with 4, 8 or 16 cells per wavefront)
struct CapsRT {
  int V, P, E;
};
template <class C>
CapsRT caps_rt() {
  return CapsRT{C::V, C::P, C::E};
}
CapsRT small_caps(int id) {
  switch (id) {
    case 1: return caps_rt<WideCaps>();
    case 2: return caps_rt<TightCaps>();
    case 3: return caps_rt<MidCaps>();
    default: return caps_rt<SmallCaps>();
  }
}

static int launch_clip(rpd_ctx* ctx, int tier, int G, ClipArgs& a,
                       int caps_id = 0) {
  if (a.n_items <= 0) return 0;
  const int blocks = ceil_div(a.n_items, G);
#define RPD_LAUNCH(GG, CAPS, EX)                                           \
  do {                                                                    \
    if (a.cert.tet_off)                                                   \
      hipLaunchKernelGGL((k_clip<GG, CAPS, EX, true>), dim3(blocks),      \
                         dim3(64), 0, ctx->stream, a);                    \
    else                                                                  \
      hipLaunchKernelGGL((k_clip<GG, CAPS, EX, false>), dim3(blocks),     \
                         dim3(64), 0, ctx->stream, a);                    \
  } while (0)
  if (tier == 0) {
    if (caps_id == 2 || caps_id == 3) {
      if (caps_id == 2) {
        switch (G) {
          case 4: RPD_LAUNCH(4, TightCaps, false); break;
          case 8: RPD_LAUNCH(8, TightCaps, false); break;
          default: RPD_LAUNCH(16, TightCaps, false); break;
        }
      } else {
        switch (G) {
          case 4: RPD_LAUNCH(4, MidCaps, false); break;
          case 8: RPD_LAUNCH(8, MidCaps, false); break;
          default: RPD_LAUNCH(16, MidCaps, false); break;
        }
      }
    } else if (caps_id == 1) {
      switch (G) {
        case 1: RPD_LAUNCH(1, WideCaps, false); break;
        case 2: RPD_LAUNCH(2, WideCaps, false); break;
        case 4: RPD_LAUNCH(4, WideCaps, false); break;
        case 8: RPD_LAUNCH(8, WideCaps, false); break;
        default: RPD_LAUNCH(16, WideCaps, false); break;
      }
    } else {
      switch (G) {
        case 1: RPD_LAUNCH(1, SmallCaps, false); break;
        case 2: RPD_LAUNCH(2, SmallCaps, false); break;
        case 4: RPD_LAUNCH(4, SmallCaps, false); break;
        case 8: RPD_LAUNCH(8, SmallCaps, false); break;
        default: RPD_LAUNCH(16, SmallCaps, false); break;
      }
    }
  } else if (tier == 1) {
    switch (G) {
      case 1: RPD_LAUNCH(1, LargeCaps, false); break;
      case 2: RPD_LAUNCH(2, LargeCaps, false); break;
      default: RPD_LAUNCH(4, LargeCaps, false); break;
    }
  } else {
    RPD_LAUNCH(1, LargeCaps, true);
  }
#undef RPD_LAUNCH
  HIP_TRY(hipGetLastError());
  return 0;
}

// ----------------------------------------------------------- run stages
struct RunState {
  rpd_run_cfg cfg;
  CapsRT sc;
  int G = 8;
  hipStream_t s = nullptr;
  int32_t nt = 0, ns = 0;
  DevMesh M;
  DevSites S;
  int32_t* cnt = nullptr;  // int32[8]: bbox fallbacks, re-run counters
  const InstInfo* inst_d = nullptr;
  const int32_t* site_inst_d = nullptr;
  const int32_t* tet_inst_d = nullptr;
  int64_t n_bins = 0, n_sg = 0;
  int32_t r_large = 0, r_exact = 0;  // re-run entries clipped so far
  int32_t r_full = 0;  // entries of the full-capacity list after a middle
                       // tier (tight small tiers only)
  CertArgs cert{};  // learned path: certificate evaluated in the clip
  // Captured pipeline (rpd_graph_build): no host reads inside the run;
  // launches and buffers are sized by cap, the counts stay on the device
  // and the host checks them against cap after the graph has run.
  bool graph = false;
  RunCaps cap{};
  // counts of an eager run (the capacities of a later capture)
  RunCaps seen{};
  int n_clip_calls = 0;  // clip_range calls of this run so far
};

MlpDev mlp_dev(rpd_ctx* ctx) {
  const float* b = ctx->mlp_bias.as<float>();
  return MlpDev{ctx->mlp_w1t.as<uint16_t>(), ctx->mlp_w2t.as<uint16_t>(),
                ctx->mlp_w3t.as<uint16_t>(), ctx->mlp_w2.as<uint16_t>(),
                ctx->mlp_w3.as<uint16_t>(), b, b + kHid, b + 2 * kHid};
}

int upload_inst_info(rpd_ctx* ctx, RunState& R);

// stream, sizes, device views and per-instance grids of the resident batch
int init_state(rpd_ctx* ctx, RunState& R) {
  R.s = ctx->stream;
  R.nt = ctx->n_tet;
  R.ns = ctx->n_site;
  R.M = DevMesh{ctx->n_vert, R.nt, ctx->vert.as<double4>(),
                ctx->tet_v.as<int4>(), ctx->tet_fid.as<int4>(),
                ctx->tet_meta.as<TetMeta>()};
  R.S = DevSites{R.ns, ctx->site.as<double4>(), ctx->flags.as<uint32_t>(),
                 ctx->knn_off.as<int64_t>(), ctx->knn_ids.as<int32_t>(),
                 ctx->knn_near.as<int32_t>(),
                 ctx->knn_site_near.as<double4>(), nullptr};
  HIP_TRY(ctx->counters.need(8 * 4));
  R.cnt = ctx->counters.as<int32_t>();
  // the bisector of every neighbour-list entry, read by the plane masks
  // and the clip instead of two dependent loads and the arithmetic
  HIP_TRY(ctx->knn_bis.need((size_t)std::max<int64_t>(ctx->n_knn, 1) * 32));
  HIP_TRY(ctx->knn_bisf.need((size_t)std::max<int64_t>(ctx->n_knn, 1) * 16));
  R.S.knn_bis = ctx->knn_bis.as<double4>();
  R.S.knn_bisf = ctx->knn_bisf.as<float4>();
  if (R.ns > 0)
    hipLaunchKernelGGL(k_knn_bis, dim3(R.ns), dim3(64), 0, R.s, R.S,
                       ctx->knn_bis.as<double4>(), ctx->knn_bisf.as<float4>());
  HIP_TRY(hipGetLastError());
  return upload_inst_info(ctx, R);
}

// per instance: candidate bin grid (about n_tet/6 bins) and sphere-centre
// grid (about 2 spheres per cell), numbered batch-wide
int upload_inst_info(rpd_ctx* ctx, RunState& R) {
  std::vector<InstInfo> ii(ctx->inst.size());
  int64_t n_bins = 0, n_sg = 0;
  for (size_t b = 0; b < ctx->inst.size(); ++b) {
    const rpd_ctx::Inst& I = ctx->inst[b];
    InstInfo& o = ii[b];
    double ext[3], vol = 1;
    for (int a = 0; a < 3; ++a) {
      ext[a] = std::max(I.dom[3 + a] - I.dom[a], 1e-300);
      vol *= ext[a];
    }
    auto grid = [&](double per_cell_units, BinGrid& g) {
      const double bs = std::cbrt(vol / std::max(1.0, per_cell_units));
      int64_t nb = 1;
      for (int a = 0; a < 3; ++a) {
        g.org[a] = I.dom[a];
        g.dim[a] = (int32_t)std::min<double>(
            512, std::max(1.0, std::ceil(ext[a] / bs)));
        nb *= g.dim[a];
      }
      g.inv_bs = 1.0 / bs;
      g.pad = 0;
      return nb;
    };
    const int64_t nb = grid(I.n_tet / 6.0, o.g);
    o.bin_base = n_bins;
    n_bins += nb;
    const int64_t ng = grid(I.n_site / 2.0, o.sg);
    o.sg_base = n_sg;
    n_sg += ng;
    for (int a = 0; a < 6; ++a) o.dom[a] = I.dom[a];
    o.rho = I.rho;
    o.site_off = (int32_t)I.site_off;
    o.n_site = I.n_site;
  }
  if (n_bins >= (1LL << 31) || n_sg >= (1LL << 31))
    return fail(ctx, "rpd_run: too many grid cells");
  const size_t bytes = ii.size() * sizeof(InstInfo);
  HIP_TRY(ctx->inst_dev.need(bytes));
  // the previous copy out of the pinned block must have completed before
  // the block is rewritten (normally long done: runs end with a wait)
  if (ctx->inst_copy_recorded && !t_capture) {
    hipError_t e;
    while ((e = hipEventQuery(ctx->ev_inst)) == hipErrorNotReady) {
    }
    HIP_TRY(e);
  }
  if (bytes > ctx->pin_inst_cap && t_arena)
    return fail(ctx, "rpd_run_noalloc: instance table over the pinned "
                     "staging block");
  if (bytes > ctx->pin_inst_cap) {
    if (ctx->pin_inst) (void)hipHostFree(ctx->pin_inst);
    ctx->pin_inst = nullptr;
    ctx->pin_inst_cap = 0;
    HIP_TRY(hipHostMalloc(&ctx->pin_inst, bytes, hipHostMallocDefault));
    ctx->pin_inst_cap = bytes;
  }
  std::memcpy(ctx->pin_inst, ii.data(), bytes);
  HIP_TRY(hipMemcpyAsync(ctx->inst_dev.p, ctx->pin_inst, bytes,
                         hipMemcpyHostToDevice, R.s));
  if (!R.graph) {
    HIP_TRY(hipEventRecord(ctx->ev_inst, R.s));
    ctx->inst_copy_recorded = true;
  }
  R.n_bins = n_bins;
  R.n_sg = n_sg;
  R.inst_d = ctx->inst_dev.as<InstInfo>();
  R.site_inst_d = ctx->site_inst.as<int32_t>();
  R.tet_inst_d = ctx->tet_inst.as<int32_t>();
  return 0;
}

int sort_keys(rpd_ctx* ctx, hipStream_t s, Buf& a, Buf& b, int64_t n,
              int end_bit) {
  if (n <= 0) return 0;
  size_t tmp = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortKeys(
      nullptr, tmp, a.as<uint64_t>(), b.as<uint64_t>(), (int)n, 0, end_bit,
      s));
  HIP_TRY(ctx->scan_tmp.need(tmp));
  HIP_TRY(hipcub::DeviceRadixSort::SortKeys(ctx->scan_tmp.p, tmp,
                                            a.as<uint64_t>(), b.as<uint64_t>(),
                                            (int)n, 0, end_bit, s));
  return 0;
}

// plain path: site boxes, binning, reference relation over the bins
int plain_candidates(rpd_ctx* ctx, RunState& R, int64_t* n_cand_out) {
  hipStream_t s = R.s;
  const int32_t nt = R.nt, ns = R.ns;
  const int64_t n_bins = R.n_bins;
  HIP_TRY(ctx->bbox.need((size_t)std::max(ns, 1) * 48));
  if (ns > 0)
    hipLaunchKernelGGL(k_site_bbox, dim3(ns), dim3(64), 0, s, R.S, R.inst_d,
                       R.site_inst_d, ctx->bbox.as<double>(),
                       ctx->knn_near.as<int32_t>(),
                       ctx->knn_site_near.as<double4>(), R.cnt + 0);
  HIP_TRY(hipGetLastError());
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[10], s));
  // sphere -> bin registrations: count, expand, radix sort by
  // (bin, sphere), bin starts by binary search
  HIP_TRY(ctx->reg_n.need(((size_t)ns + 1) * 8));
  HIP_TRY(ctx->reg_off.need(((size_t)ns + 1) * 8));
  hipLaunchKernelGGL(k_bin_nreg, dim3(ceil_div(ns + 1, 256)), dim3(256), 0, s,
                     ns, ctx->bbox.as<double>(), R.inst_d, R.site_inst_d,
                     ctx->reg_n.as<int64_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(exclusive_sum<int64_t>(ctx, ctx->reg_n.as<int64_t>(),
                                 ctx->reg_off.as<int64_t>(), ns + 1));
  int64_t n_reg = 0;
  if (R.graph)
    n_reg = R.cap.n_reg;
  else
    HIP_TRY(read_back(ctx, s, {{&n_reg, ctx->reg_off.as<int64_t>() + ns, 8}}));
  if (n_reg >= (1LL << 31))
    return fail(ctx, "rpd_run: too many sphere-bin registrations");
  ctx->st.n_bin_registrations = n_reg;
  R.seen.n_reg = n_reg;
  HIP_TRY(ctx->keys_a.need((size_t)std::max<int64_t>(n_reg, 1) * 8));
  HIP_TRY(ctx->keys_b.need((size_t)std::max<int64_t>(n_reg, 1) * 8));
  HIP_TRY(ctx->bin_start.need((n_bins + 1) * 4));
  HIP_TRY(ctx->bin_ids.need((size_t)std::max<int64_t>(n_reg, 1) * 4));
  if (n_reg > 0) {
    int end_bit = 32;
    while ((1LL << (end_bit - 32)) < n_bins + 1) ++end_bit;
    // (a captured run pads its registrations past the device count with a
    // key above every bin)
    hipLaunchKernelGGL(k_bin_expand, dim3(ceil_div(n_reg, 256)), dim3(256), 0,
                       s, ns, n_reg, ctx->bbox.as<double>(), R.inst_d,
                       R.site_inst_d, ctx->reg_off.as<int64_t>(),
                       ctx->keys_a.as<uint64_t>(),
                       R.graph ? ctx->reg_off.as<int64_t>() + ns : nullptr,
                       end_bit >= 64 ? ~0ull : (1ull << end_bit) - 1);
    HIP_TRY(hipGetLastError());
    if (int rc = sort_keys(ctx, s, ctx->keys_a, ctx->keys_b, n_reg, end_bit))
      return rc;
  }
  hipLaunchKernelGGL(k_bin_starts,
                     dim3(ceil_div(std::max<int64_t>(n_bins + 1, n_reg), 256)),
                     dim3(256), 0, s, n_bins, n_reg, ctx->keys_b.as<uint64_t>(),
                     ctx->bin_start.as<int32_t>(), ctx->bin_ids.as<int32_t>(),
                     R.graph ? ctx->reg_off.as<int64_t>() + ns : nullptr);
  HIP_TRY(hipGetLastError());
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[11], s));
  HIP_TRY(ctx->tet_count.need(((size_t)nt + 1) * 8));
  HIP_TRY(ctx->tet_off.need(((size_t)nt + 1) * 8));
  HIP_TRY(ctx->tet_pass.need(((size_t)nt + 1) * 2 * 8));  // 2 words per tet
  HIP_TRY(zero32(s, ctx->tet_count.as<int64_t>() + nt, 2));
  hipLaunchKernelGGL(k_tet_count, dim3(ceil_div(nt, 256)), dim3(256), 0, s,
                     R.M, R.S, R.inst_d, R.tet_inst_d,
                     ctx->bin_start.as<int32_t>(), ctx->bin_ids.as<int32_t>(),
                     ctx->tet_count.as<int64_t>(), ctx->tet_pass.as<uint64_t>(),
                     R.cfg.count_work
                         ? ctx->rel_work.as<unsigned long long>()
                         : nullptr,
                     ctx->has_subset ? ctx->tet_sel.as<uint8_t>() : nullptr);
  HIP_TRY(hipGetLastError());
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[12], s));
  HIP_TRY(exclusive_sum<int64_t>(ctx, ctx->tet_count.as<int64_t>(),
                                 ctx->tet_off.as<int64_t>(), nt + 1));
  int64_t n_cand = 0;
  if (R.graph)
    n_cand = R.cap.n_cand;
  else
    HIP_TRY(
        read_back(ctx, s, {{&n_cand, ctx->tet_off.as<int64_t>() + nt, 8}}));
  R.seen.n_cand = n_cand;
  HIP_TRY(ctx->cand.need((size_t)std::max<int64_t>(n_cand, 1) * sizeof(Cand)));
  HIP_TRY(ctx->rel_ids.need((size_t)std::max<int64_t>(n_cand, 1) * 4));
  hipLaunchKernelGGL(k_tet_fill, dim3(ceil_div(nt, 256)), dim3(256), 0, s, R.M,
                     R.S, R.inst_d, R.tet_inst_d, ctx->bin_start.as<int32_t>(),
                     ctx->bin_ids.as<int32_t>(), ctx->tet_off.as<int64_t>(),
                     ctx->tet_pass.as<uint64_t>(), ctx->cand.as<Cand>(),
                     ctx->rel_ids.as<int32_t>(),
                     ctx->has_subset ? ctx->tet_sel.as<uint8_t>() : nullptr,
                     n_cand, R.cnt + 7);
  HIP_TRY(hipGetLastError());
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[13], s));
  *n_cand_out = n_cand;
  return 0;
}

// neighbour spheres inline in list order (descent and relation read them)
int site_inline(rpd_ctx* ctx, RunState& R) {
  // (one wavefront per site: the lists' lengths are read on the device, so
  // a captured run follows sphere updates that change them)
  if (R.ns > 0)
    hipLaunchKernelGGL(k_site_inline, dim3(R.ns), dim3(64), 0, R.s, R.S,
                       ctx->knn_near.as<int32_t>(),
                       ctx->knn_site_near.as<double4>());
  HIP_TRY(hipGetLastError());
  return 0;
}

// owner descent, 32 proposals and features of every tet
int learned_features(rpd_ctx* ctx, RunState& R) {
  hipStream_t s = R.s;
  const int32_t nt = R.nt, ns = R.ns;
  if (int rc = site_inline(ctx, R)) return rc;
  HIP_TRY(ctx->sg_keys_a.need((size_t)std::max(ns, 1) * 8));
  HIP_TRY(ctx->sg_keys_b.need((size_t)std::max(ns, 1) * 8));
  HIP_TRY(ctx->sg_start.need((size_t)(R.n_sg + 1) * 4));
  HIP_TRY(ctx->sg_ids.need((size_t)std::max(ns, 1) * 4));
  if (ns > 0) {
    hipLaunchKernelGGL(k_sg_keys, dim3(ceil_div(ns, 256)), dim3(256), 0, s,
                       R.S, R.inst_d, R.site_inst_d,
                       ctx->sg_keys_a.as<uint64_t>());
    HIP_TRY(hipGetLastError());
    int end_bit = 32;
    while ((1LL << (end_bit - 32)) < R.n_sg + 1) ++end_bit;
    if (int rc = sort_keys(ctx, s, ctx->sg_keys_a, ctx->sg_keys_b, ns,
                           end_bit))
      return rc;
  }
  hipLaunchKernelGGL(k_bin_starts,
                     dim3(ceil_div(std::max<int64_t>(R.n_sg + 1, ns), 256)),
                     dim3(256), 0, s, R.n_sg, (int64_t)ns,
                     ctx->sg_keys_b.as<uint64_t>(), ctx->sg_start.as<int32_t>(),
                     ctx->sg_ids.as<int32_t>(), nullptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(ctx->knn.need((size_t)std::max(nt, 1) * kOut * 4));
  HIP_TRY(ctx->feat.need((size_t)std::max(nt, 1) * kIn * 2));
  HIP_TRY(ctx->omask.need((size_t)std::max(nt, 1) * 4));
  HIP_TRY(ctx->vowner.need((size_t)std::max(ctx->n_vert, 1) * 4));
  hipLaunchKernelGGL(k_vertex_owner, dim3(ceil_div(ctx->n_vert, 256)),
                     dim3(256), 0, s, R.M, R.S, R.inst_d,
                     ctx->vert_inst.as<int32_t>(), ctx->sg_start.as<int32_t>(),
                     ctx->sg_ids.as<int32_t>(), ctx->vowner.as<int32_t>());
  HIP_TRY(hipGetLastError());
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[14], s));
  // 32-bit proposal entries hold instance-local ids of 16 bits
  bool wide = false;
  for (const rpd_ctx::Inst& I : ctx->inst) wide = wide || I.n_site > 65536;
  auto* feat_kernel = wide ? k_knn_features_wide : k_knn_features;
  hipLaunchKernelGGL(feat_kernel, dim3(ceil_div(nt, 64)), dim3(64), 0, s,
                     R.M, R.S, R.inst_d, R.tet_inst_d,
                     ctx->vowner.as<int32_t>(), ctx->knn.as<int32_t>(),
                     ctx->feat.as<uint16_t>(), ctx->omask.as<uint32_t>());
  HIP_TRY(hipGetLastError());
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[15], s));
  ctx->feat_n = nt;
  return 0;
}

int learned_candidates(rpd_ctx* ctx, RunState& R, int64_t* n_cand_out) {
  hipStream_t s = R.s;
  const int32_t nt = R.nt;
  if (R.cfg.learned == 1 && !ctx->mlp_ready)
    return fail(ctx, "rpd_run: learned path without a model (rpd_mlp_set "
                     "or rpd_mlp_train first)");
  if (R.cfg.learned != 1 && R.cfg.learned != 2)
    return fail(ctx, "rpd_run: learned must be 0, 1 or 2");
  if (int rc = learned_features(ctx, R)) return rc;
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[10], s));
  HIP_TRY(ctx->sel.need((size_t)std::max(nt, 1) * 4));
  if (R.cfg.learned == 1) {
    const int64_t tiles = (nt + 15) / 16;
    hipLaunchKernelGGL(
        k_mlp_infer,
        dim3((int)std::max<int64_t>(
            1, std::min<int64_t>((tiles + kInferWaves - 1) / kInferWaves,
                                 1024))),
        dim3(64 * kInferWaves), 0, s, mlp_dev(ctx), ctx->feat.as<uint16_t>(),
        (int64_t)nt, R.cfg.logit_threshold, ctx->sel.as<uint32_t>(),
        (float*)nullptr);
  } else {
    hipLaunchKernelGGL(k_sel_topk, dim3(ceil_div(nt, 256)), dim3(256), 0, s,
                       nt, R.cfg.proposal_k, ctx->sel.as<uint32_t>());
  }
  HIP_TRY(hipGetLastError());
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[11], s));
  HIP_TRY(ctx->tet_count.need(((size_t)nt + 1) * 8));
  HIP_TRY(ctx->tet_off.need(((size_t)nt + 1) * 8));
  HIP_TRY(zero32(s, ctx->tet_count.as<int64_t>() + nt, 2));
  hipLaunchKernelGGL(k_learn_count, dim3(ceil_div(nt, 64)), dim3(64), 0, s, nt,
                     R.M, R.S, ctx->knn.as<int32_t>(), ctx->sel.as<uint32_t>(),
                     ctx->omask.as<uint32_t>(), ctx->tet_count.as<int64_t>());
  HIP_TRY(hipGetLastError());
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[12], s));
  HIP_TRY(exclusive_sum<int64_t>(ctx, ctx->tet_count.as<int64_t>(),
                                 ctx->tet_off.as<int64_t>(), nt + 1));
  // (a captured run: the first-round count stays at tet_off[n_tet], the
  // lists hold the capacity)
  int64_t n_cand = R.cap.n_cand;
  if (!R.graph)
    HIP_TRY(
        read_back(ctx, s, {{&n_cand, ctx->tet_off.as<int64_t>() + nt, 8}}));
  R.seen.n_cand = n_cand;
  HIP_TRY(ctx->cand.need((size_t)std::max<int64_t>(n_cand, 1) * sizeof(Cand)));
  // (a captured run swaps rel_ids with the output ids every run: both hold
  // the larger of the two capacities, so no capture reallocates them)
  const int64_t n_ids =
      R.graph ? std::max(n_cand, R.cap.incl) : std::max<int64_t>(n_cand, 1);
  HIP_TRY(ctx->rel_ids.need((size_t)n_ids * 4));
  hipLaunchKernelGGL(k_learn_fill, dim3(ceil_div(nt, 64)), dim3(64), 0, s, nt,
                     ctx->knn.as<int32_t>(), ctx->sel.as<uint32_t>(),
                     ctx->tet_off.as<int64_t>(), ctx->cand.as<Cand>(),
                     ctx->rel_ids.as<int32_t>(),
                     R.graph ? n_cand : INT64_MAX, R.cnt + 7);
  HIP_TRY(hipGetLastError());
  if (R.graph) {
    hipLaunchKernelGGL(k_clamp_offsets, dim3(ceil_div(nt + 1, 256)),
                       dim3(256), 0, s, ctx->tet_off.as<int64_t>(),
                       (int64_t)nt, n_cand, (const int32_t*)(R.cnt + 7));
    HIP_TRY(hipGetLastError());
  }
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[13], s));
  *n_cand_out = n_cand;
  return 0;
}

ClipArgs clip_args(rpd_ctx* ctx, RunState& R) {
  ClipArgs a;
  a.M = R.M;
  a.S = R.S;
  a.cand = ctx->cand.as<Cand>();
  a.items = nullptr;
  a.item_base = 0;
  a.n_items = 0;
  a.slot_base = 0;
  a.slot_by_cell = 0;
  a.use_mask = R.cfg.use_active_mask;
  a.force_exact = R.cfg.force_exact;
  a.geometry = R.cfg.compute_geometry;
  a.rerun_large = ctx->rerun_large.as<int32_t>();
  a.n_rerun_large = R.cnt + 1;
  a.rerun_exact = ctx->rerun_exact.as<int32_t>();
  a.n_rerun_exact = R.cnt + 2;
  a.work = R.cfg.count_work ? ctx->work.as<unsigned long long>() : nullptr;
  a.cert = R.cert;
  a.n_items_d64 = nullptr;
  a.n_items_d32 = nullptr;
  a.n_sub = 0;
  a.cap_large = 0x7fffffff;
  a.cap_exact = 0x7fffffff;
  a.overflow = R.cnt + 7;
  return a;
}

// event pairs around the small-tier clip launches of one run
constexpr int kSmallEv = 16;

// masks, small tier, full-capacity and exact re-runs of candidates
// [c0, c1); buffers grow keeping earlier rounds; loc of the range is set
// (below this many cells every wavefront of the launch is resident at once,
// so the order cannot shorten it and the sort would only add latency)
constexpr int64_t kStepOrderMin = 65536;

// (a captured run: the range [c0, c1) is a capacity; its cells are the
// first *rel_dev, ending at *abs_dev = c0 + *rel_dev, both on the device --
// the first launch of a run, from 0, defaults to tet_off[n_tet])
int clip_range(rpd_ctx* ctx, RunState& R, int64_t c0, int64_t c1,
               bool events, const int64_t* rel_dev = nullptr,
               const int64_t* abs_dev = nullptr) {
  hipStream_t s = R.s;
  const int64_t n = c1 - c0;
  const CapsRT sc = R.sc;
  const bool first_call = R.n_clip_calls++ == 0;
  // a captured run's re-run tiers are sized by the capacities of this
  // launch (the first, or the largest of the later ones)
  const int64_t cap_mid = first_call ? R.cap.n_mid : R.cap.c_mid;
  const int64_t cap_full = first_call ? R.cap.n_full : R.cap.c_full;
  const int64_t cap_exact = first_call ? R.cap.n_exact : R.cap.c_exact;
  // (a captured run's counters start each launch at R.r_large, R.r_exact,
  // R.r_full: zero for the first, set by the previous launch's check)
  // the small tier takes its cells in decreasing order of clip steps (a
  // stable sort of the step counts), so that the cells of one wavefront
  // need about as many steps as each other
  const bool by_steps = R.cfg.use_active_mask && R.cfg.small_tier &&
                        n >= kStepOrderMin;
  if (by_steps) {
    HIP_TRY(ctx->trip_k.need((size_t)n));
    HIP_TRY(ctx->trip_k2.need((size_t)n));
    HIP_TRY(ctx->trip_v.need((size_t)n * 4));
    HIP_TRY(ctx->trip_v2.need((size_t)n * 4));
  }
  // (a captured run: the candidate count is tet_off[n_tet] on the device)
  const int64_t* n_dev =
      R.graph ? (abs_dev ? abs_dev : ctx->tet_off.as<int64_t>() + R.nt)
              : nullptr;
  const int64_t* n_rel = R.graph ? (rel_dev ? rel_dev : n_dev) : nullptr;
  if (R.cfg.use_active_mask && n > 0)
    hipLaunchKernelGGL(k_cand_mask, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                       R.M, R.S, n, ctx->cand.as<Cand>() + c0,
                       by_steps ? ctx->trip_k.as<uint8_t>() : nullptr,
                       by_steps ? ctx->trip_v.as<int32_t>() : nullptr, c0,
                       n_dev);
  HIP_TRY(hipGetLastError());
  if (by_steps) {
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairsDescending(
        nullptr, tmp, ctx->trip_k.as<uint8_t>(), ctx->trip_k2.as<uint8_t>(),
        ctx->trip_v.as<int32_t>(), ctx->trip_v2.as<int32_t>(), (int)n, 0, 8,
        s));
    HIP_TRY(ctx->scan_tmp.need(tmp));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairsDescending(
        ctx->scan_tmp.p, tmp, ctx->trip_k.as<uint8_t>(),
        ctx->trip_k2.as<uint8_t>(), ctx->trip_v.as<int32_t>(),
        ctx->trip_v2.as<int32_t>(), (int)n, 0, 8, s));
  }
  if (events) HIP_TRY(hipEventRecord(ctx->ev[1], s));
  // the re-run lists: room for every cell of the range after the entries
  // of earlier launches
  const int64_t list_cap =
      std::max<int64_t>({c1, (int64_t)R.r_large + n, (int64_t)R.r_exact + n, 1});
  HIP_TRY(ctx->rerun_large.need_keep((size_t)list_cap * 4,
                                     (size_t)R.r_large * 4, s));
  HIP_TRY(ctx->rerun_exact.need_keep((size_t)list_cap * 4,
                                     (size_t)R.r_exact * 4, s));
  HIP_TRY(ctx->loc.need_keep((size_t)std::max<int64_t>(c1, 1) * 4,
                             (size_t)c0 * 4, s));
  HIP_TRY(zero32(s, ctx->loc.as<int32_t>() + c0, n));
  if (R.cfg.small_tier) {
    HIP_TRY(ctx->small.need_keep(c1, c0, sc.V, sc.P, sc.E, s));
  } else {
    HIP_TRY(ctx->small.need_keep(1, 0, sc.V, sc.P, sc.E, s));
  }
  ClipArgs a = clip_args(ctx, R);
  // the re-run lists hold up to list_cap entries (need_keep above)
  a.cap_large = (int32_t)std::min<int64_t>(list_cap, 0x7fffffff);
  a.cap_exact = a.cap_large;
  if (R.cfg.small_tier) {
    a.item_base = c0;
    a.slot_base = c0;
    a.n_items = n;
    a.n_items_d64 = n_rel;
    if (by_steps) {  // slot = cell index, as with the identity order
      a.items = ctx->trip_v2.as<int32_t>();
      a.slot_by_cell = 1;
    }
    a.out = ctx->small.view(sc.V, sc.P, sc.E);
    // the tier's shape is the same for every launch of a run: the FP64
    // geometry sums over the lanes of a cell, so its rounding depends on
    // the cells per wavefront
    const int k_ev = ctx->n_small_ev;
    const bool ev = !R.graph && n > 0 && k_ev < kSmallEv;
    if (ev) HIP_TRY(hipEventRecord(ctx->ev[16 + 2 * k_ev], s));
    if (int rc = launch_clip(ctx, 0, R.G, a, R.cfg.small_caps)) return rc;
    if (ev) {
      HIP_TRY(hipEventRecord(ctx->ev[17 + 2 * k_ev], s));
      ctx->n_small_ev = k_ev + 1;
    }
    a.items = nullptr;  // the re-run tiers index their own lists
    a.slot_by_cell = 0;
    a.n_items_d64 = nullptr;
  } else if (n > 0) {
    if (R.graph)
      return fail(ctx, "rpd_graph_build: the small tier is required");
    // every cell goes straight to the full-capacity tier
    std::vector<int32_t> idx(n);
    for (int64_t q = 0; q < n; ++q) idx[q] = (int32_t)(c0 + q);
    HIP_TRY(hipMemcpyAsync(ctx->rerun_large.as<int32_t>() + R.r_large,
                           idx.data(), n * 4, hipMemcpyHostToDevice, s));
    int32_t nf = (int32_t)(R.r_large + n);
    HIP_TRY(hipMemcpyAsync(R.cnt + 1, &nf, 4, hipMemcpyHostToDevice, s));
    HIP_TRY(wait_stream(ctx, s));
  }
  if (events) HIP_TRY(hipEventRecord(ctx->ev[2], s));
  int32_t host_cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (!R.graph) HIP_TRY(read_back(ctx, s, {{host_cnt, R.cnt, 32}}));
  // (a captured run launches its re-run tiers at capacity and reads their
  // counts on the device)
  const bool mid = R.cfg.small_tier && R.sc.V < SmallCaps::V;
  const int32_t n_over =
      R.graph ? (int32_t)(R.r_large + (mid ? cap_mid : cap_full))
              : host_cnt[1];
  // counts of the eager run (first launch; the largest later one)
  auto note = [&](int64_t& first, int64_t& later, int64_t v) {
    if (first_call)
      first = v;
    else
      later = std::max(later, v);
  };
  note(R.seen.n_mid, R.seen.c_mid, n_over - R.r_large);
  // cells over a tight small tier first get the middle tier (SmallCaps, 8
  // per wavefront; tier 3, slot = position in rerun_large); the ones over
  // that, or over any other small tier, get the full-capacity tier (tier 1,
  // slot = position in its list)
  const int32_t* full_list = ctx->rerun_large.as<int32_t>();
  int32_t full_r0 = R.r_large, full_r1 = n_over;
  if (mid) {
    HIP_TRY(ctx->mid.need_keep(n_over, R.r_large, SmallCaps::V, SmallCaps::P,
                               SmallCaps::E, s));
    HIP_TRY(ctx->rerun_full.need_keep(
        (size_t)std::max<int64_t>(R.r_full + (n_over - R.r_large), 1) * 4,
        (size_t)R.r_full * 4, s));
    if (n_over > R.r_large) {
      ClipArgs m = a;
      m.items = ctx->rerun_large.as<int32_t>() + R.r_large;
      m.n_items = n_over - R.r_large;
      m.slot_base = R.r_large;
      m.out = ctx->mid.view(SmallCaps::V, SmallCaps::P, SmallCaps::E);
      m.rerun_large = ctx->rerun_full.as<int32_t>();
      m.n_rerun_large = R.cnt + 6;
      m.cap_large = (int32_t)(R.r_full + (n_over - R.r_large));
      if (R.graph) {
        m.n_items_d32 = R.cnt + 1;
        m.n_sub = R.r_large;
      }
      if (m.work) m.work = ctx->work.as<unsigned long long>() + 256;
      if (int rc = launch_clip(ctx, 0, 8, m, 0)) return rc;
      hipLaunchKernelGGL(k_mark_rerun, dim3(ceil_div(n_over - R.r_large, 256)),
                         dim3(256), 0, s, ctx->rerun_large.as<int32_t>(),
                         R.r_large, n_over, 3, ctx->loc.as<int32_t>(),
                         R.graph ? R.cnt + 1 : nullptr);
      HIP_TRY(hipGetLastError());
      if (!R.graph) HIP_TRY(read_back(ctx, s, {{host_cnt, R.cnt, 32}}));
    }
    full_list = ctx->rerun_full.as<int32_t>();
    full_r0 = R.r_full;
    full_r1 = R.graph ? (int32_t)(R.r_full + cap_full) : host_cnt[6];
  }
  note(R.seen.n_full, R.seen.c_full, full_r1 - full_r0);
  // a captured run reads the full tier's count here
  const int32_t* full_cnt = mid ? R.cnt + 6 : R.cnt + 1;
  HIP_TRY(ctx->large.need_keep(full_r1, full_r0, LargeCaps::V, LargeCaps::P,
                               LargeCaps::E, s));
  if (full_r1 > full_r0) {
    a.items = full_list + full_r0;
    a.n_items = full_r1 - full_r0;
    a.slot_base = full_r0;
    a.out = ctx->large.view(LargeCaps::V, LargeCaps::P, LargeCaps::E);
    if (R.graph) {
      a.n_items_d32 = full_cnt;
      a.n_sub = full_r0;
    }
    if (a.work) a.work = ctx->work.as<unsigned long long>() + 256;
    if (int rc = launch_clip(ctx, 1, std::min(R.G, 4), a)) return rc;
    if (!R.graph) HIP_TRY(read_back(ctx, s, {{host_cnt, R.cnt, 32}}));
  }
  const int32_t n_exact =
      R.graph ? (int32_t)(R.r_exact + cap_exact) : host_cnt[2];
  note(R.seen.n_exact, R.seen.c_exact, n_exact - R.r_exact);
  HIP_TRY(ctx->exact.need_keep(n_exact, R.r_exact, LargeCaps::V, LargeCaps::P,
                               LargeCaps::E, s));
  if (n_exact > R.r_exact) {
    a.items = ctx->rerun_exact.as<int32_t>() + R.r_exact;
    a.n_items = n_exact - R.r_exact;
    a.slot_base = R.r_exact;
    a.out = ctx->exact.view(LargeCaps::V, LargeCaps::P, LargeCaps::E);
    if (R.graph) {
      a.n_items_d32 = R.cnt + 2;
      a.n_sub = R.r_exact;
    }
    if (a.work) a.work = ctx->work.as<unsigned long long>() + 512;
    if (int rc = launch_clip(ctx, 2, 1, a)) return rc;
  }
  if (full_r1 > full_r0)
    hipLaunchKernelGGL(k_mark_rerun, dim3(ceil_div(full_r1 - full_r0, 256)),
                       dim3(256), 0, s, full_list, full_r0, full_r1, 1,
                       ctx->loc.as<int32_t>(), R.graph ? full_cnt : nullptr);
  if (n_exact > R.r_exact)
    hipLaunchKernelGGL(k_mark_rerun, dim3(ceil_div(n_exact - R.r_exact, 256)),
                       dim3(256), 0, s, ctx->rerun_exact.as<int32_t>(),
                       R.r_exact, n_exact, 2, ctx->loc.as<int32_t>(),
                       R.graph ? R.cnt + 2 : nullptr);
  HIP_TRY(hipGetLastError());
  R.r_large = n_over;
  if (mid) R.r_full = full_r1;
  R.r_exact = n_exact;
  if (R.graph) {
    // this launch's re-run counts within what it reserved; a learned run's
    // next launch starts at the next reserved positions (the plain run has
    // one launch, whose counts the host reads)
    const bool next = R.cfg.learned != 0;
    hipLaunchKernelGGL(k_check_rerun_counts, dim3(1), dim3(64), 0, s, R.cnt,
                       n_over, n_exact, mid ? full_r1 : 0x7fffffff,
                       R.cnt + 7, next ? R.r_large : -1, R.r_exact,
                       R.r_full);
    HIP_TRY(hipGetLastError());
  }
  if (events) HIP_TRY(hipEventRecord(ctx->ev[3], s));
  return 0;
}

TierViews tier_views(rpd_ctx* ctx, const RunState& R) {
  TierViews tv;
  tv.t[0] = ctx->small.view(R.sc.V, R.sc.P, R.sc.E);
  tv.t[1] = ctx->large.view(LargeCaps::V, LargeCaps::P, LargeCaps::E);
  tv.t[2] = ctx->exact.view(LargeCaps::V, LargeCaps::P, LargeCaps::E);
  tv.t[3] = ctx->mid.view(SmallCaps::V, SmallCaps::P, SmallCaps::E);
  return tv;
}

// learned path: certificate of the first-round candidates [0, n1),
// closure rounds (missing related facet neighbours added and clipped, new
// cells certified), full-relation re-run of the tets left open, final
// (tet, site) order.  Returns the number of candidates in the output order
// (the first n_incl of perm).
constexpr int kClosureRounds = 6;
static_assert(kClosureRounds + 1 <= 8, "RunCaps holds 8 rounds");

int sort_u64(rpd_ctx* ctx, hipStream_t s, const uint64_t* in, uint64_t* out,
             int64_t n) {
  if (n <= 0) return 0;
  size_t tmp = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, in, out, (int)n, 0,
                                            64, s));
  HIP_TRY(ctx->scan_tmp.need(tmp));
  HIP_TRY(hipcub::DeviceRadixSort::SortKeys(ctx->scan_tmp.p, tmp, in, out,
                                            (int)n, 0, 64, s));
  return 0;
}

// certificate state for a round of n new candidates: a cleared miss list
// of capacity max(4096, 2n) and the current ext keys
int certify_round(rpd_ctx* ctx, RunState& R, int64_t n, int64_t n_ext,
                  const int64_t* cap_n = nullptr, int64_t cap_graph = 0) {
  // (a captured run: the capacity of this miss list)
  const int64_t cap = R.graph ? cap_graph : std::max<int64_t>(4096, 2 * n);
  HIP_TRY(ctx->miss_a.need((size_t)cap * 8));
  HIP_TRY(ctx->miss_b.need((size_t)cap * 8));
  HIP_TRY(zero32(R.s, R.cnt + 4, 1));
  R.cert.ext = ctx->mem_a.as<uint64_t>();
  R.cert.n_ext = n_ext;
  R.cert.miss = ctx->miss_a.as<uint64_t>();
  R.cert.cap = (int32_t)std::min<int64_t>(cap, INT32_MAX);
  R.cert.n_miss = R.cnt + 4;
  R.cert.cap_n = R.graph ? cap_n : nullptr;
  return 0;
}

// before the first round's clip: cleared tet flags, members = the
// first-round sets (CSR tet_off / rel_ids of the learned fill)
int certify_begin(rpd_ctx* ctx, RunState& R, int64_t n1) {
  hipStream_t s = R.s;
  const int32_t nt = R.nt;
  // byte flags per tet, cleared by 32-bit words
  const int64_t w = ((int64_t)nt + 4) / 4;
  HIP_TRY(ctx->flag.need((size_t)w * 4));
  HIP_TRY(ctx->has_cell.need((size_t)w * 4));
  HIP_TRY(ctx->reopened.need((size_t)w * 4));
  HIP_TRY(zero32(s, ctx->flag.p, w));
  HIP_TRY(zero32(s, ctx->has_cell.p, w));
  HIP_TRY(zero32(s, ctx->reopened.p, w));
  R.cert.tet_off = ctx->tet_off.as<int64_t>();
  R.cert.mem_ids = ctx->rel_ids.as<int32_t>();
  R.cert.flag = ctx->flag.as<uint8_t>();
  R.cert.has_cell = ctx->has_cell.as<uint8_t>();
  return certify_round(ctx, R, n1, 0, ctx->tet_off.as<int64_t>() + R.nt,
                       R.cap.miss_r[0]);
}

int certify_captured(rpd_ctx* ctx, RunState& R, int64_t n1,
                     int64_t* n_incl_out);

int certify_and_fallback(rpd_ctx* ctx, RunState& R, int64_t n1,
                         int64_t* n_incl_out) {
  if (R.graph) return certify_captured(ctx, R, n1, n_incl_out);
  hipStream_t s = R.s;
  const int32_t nt = R.nt;
  HIP_TRY(ctx->flag32.need(((size_t)nt + 1) * 4));
  HIP_TRY(ctx->flag_pos.need(((size_t)nt + 1) * 4));
  // the cells of [c_lo, c_hi) were certified while clipped; ext = the
  // sorted keys of every pair added since the first round
  int64_t n_ext = 0;
  int64_t c_lo = 0, c_hi = n1, added = 0;
  for (int round = 0;; ++round) {
    int32_t nm = 0;
    HIP_TRY(read_back(ctx, s, {{&nm, R.cnt + 4, 4}}));
    const int64_t n_miss = std::min<int64_t>(nm, R.cert.cap);
    R.seen.miss = std::max<int64_t>(R.seen.miss, n_miss);
    R.seen.miss_r[round] = n_miss;
    if (n_miss == 0) break;
    if (round == kClosureRounds) {  // still open: full relation
      hipLaunchKernelGGL(k_flag_miss, dim3(ceil_div(n_miss, 256)), dim3(256),
                         0, s, n_miss, ctx->miss_a.as<uint64_t>(),
                         ctx->flag.as<uint8_t>(), (const int32_t*)nullptr);
      HIP_TRY(hipGetLastError());
      break;
    }
    R.seen.rounds = round + 1;
    // unique missing pairs; related ones become candidates, all members
    if (int rc = sort_u64(ctx, s, ctx->miss_a.as<uint64_t>(),
                          ctx->miss_b.as<uint64_t>(), n_miss))
      return rc;
    HIP_TRY(ctx->umark.need(((size_t)n_miss + 1) * 4));
    HIP_TRY(ctx->upos.need(((size_t)n_miss + 1) * 4));
    hipLaunchKernelGGL(k_unique_mark, dim3(ceil_div(n_miss + 1, 256)),
                       dim3(256), 0, s, n_miss, ctx->miss_b.as<uint64_t>(),
                       ctx->umark.as<int32_t>(), (const int32_t*)nullptr);
    HIP_TRY(exclusive_sum<int32_t>(ctx, ctx->umark.as<int32_t>(),
                                   ctx->upos.as<int32_t>(), n_miss + 1));
    // the unique count n_u = upos[n_miss] stays on the device: buffers and
    // launches are sized by n_miss >= n_u, and the pass flags past n_u are
    // zero, so the number of new candidates is ppos[n_miss]
    HIP_TRY(ctx->ukeys.need((size_t)std::max<int64_t>(n_miss, 1) * 8));
    HIP_TRY(ctx->upass.need((size_t)n_miss + 1));
    HIP_TRY(ctx->ppos.need(((size_t)n_miss + 1) * 4));
    hipLaunchKernelGGL(k_closure_split, dim3(ceil_div(n_miss, 256)), dim3(256),
                       0, s, n_miss, ctx->miss_b.as<uint64_t>(),
                       ctx->umark.as<int32_t>(), ctx->upos.as<int32_t>(), R.M,
                       R.S, ctx->ukeys.as<uint64_t>(),
                       ctx->upass.as<uint8_t>());
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_pass32, dim3(ceil_div(n_miss + 1, 256)), dim3(256), 0,
                       s, n_miss, ctx->upos.as<int32_t>() + n_miss,
                       ctx->upass.as<uint8_t>(), ctx->umark.as<int32_t>());
    HIP_TRY(exclusive_sum<int32_t>(ctx, ctx->umark.as<int32_t>(),
                                   ctx->ppos.as<int32_t>(), n_miss + 1));
    int32_t n_u = 0, n_new = 0;
    HIP_TRY(read_back(ctx, s,
                      {{&n_u, ctx->upos.as<int32_t>() + n_miss, 4},
                       {&n_new, ctx->ppos.as<int32_t>() + n_miss, 4}}));
    R.seen.uniq = std::max<int64_t>(R.seen.uniq, n_u);
    R.seen.added = std::max<int64_t>(R.seen.added, n_new);
    R.seen.added_r[round] = n_new;
    HIP_TRY(ctx->cand.need_keep((size_t)(c_hi + n_new) * sizeof(Cand),
                                (size_t)c_hi * sizeof(Cand), s));
    if (n_u > 0)
      hipLaunchKernelGGL(k_closure_fill, dim3(ceil_div(n_u, 256)), dim3(256),
                         0, s, (int64_t)n_u, ctx->ukeys.as<uint64_t>(),
                         ctx->upass.as<uint8_t>(), ctx->ppos.as<int32_t>(),
                         ctx->cand.as<Cand>(), c_hi,
                         ctx->reopened.as<uint8_t>(), (const int32_t*)nullptr,
                         INT64_MAX, R.cnt + 7);
    HIP_TRY(hipGetLastError());
    // ext += the unique pairs (disjoint from it), sorted again
    HIP_TRY(ctx->mem_a.need_keep((size_t)(n_ext + n_u) * 8, (size_t)n_ext * 8,
                                 s));
    HIP_TRY(hipMemcpyAsync(ctx->mem_a.as<uint64_t>() + n_ext, ctx->ukeys.p,
                           (size_t)n_u * 8, hipMemcpyDeviceToDevice, s));
    n_ext += n_u;
    R.seen.ext_r[round] = n_ext;
    HIP_TRY(ctx->mem_b.need((size_t)n_ext * 8));
    if (int rc = sort_u64(ctx, s, ctx->mem_a.as<uint64_t>(),
                          ctx->mem_b.as<uint64_t>(), n_ext))
      return rc;
    ctx->mem_a.swap(ctx->mem_b);
    if (int rc = certify_round(ctx, R, n_new, n_ext)) return rc;
    if (n_new > 0)
      if (int rc = clip_range(ctx, R, c_hi, c_hi + n_new, false)) return rc;
    added += n_new;
    c_lo = c_hi;
    c_hi += n_new;
  }
  (void)c_lo;
  R.seen.ext = n_ext;
  R.cert.tet_off = nullptr;  // the full-relation re-runs are final
  // tets without a cell, with a dropped cell or left open: full relation
  HIP_TRY(ctx->reop64.need(64 * 4));
  HIP_TRY(hipMemsetAsync(ctx->reop64.p, 0, 64 * 4, s));
  hipLaunchKernelGGL(k_flag_final, dim3(ceil_div(nt, 256)), dim3(256), 0, s,
                     nt, ctx->flag.as<uint8_t>(), ctx->has_cell.as<uint8_t>(),
                     ctx->reopened.as<uint8_t>(), ctx->flag32.as<int32_t>(),
                     ctx->reop64.as<int32_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemsetAsync(ctx->flag32.as<int32_t>() + nt, 0, 4, s));
  HIP_TRY(exclusive_sum<int32_t>(ctx, ctx->flag32.as<int32_t>(),
                                 ctx->flag_pos.as<int32_t>(), nt + 1));
  int32_t nf = 0, n_reopened = 0, reop[64];
  HIP_TRY(read_back(ctx, s,
                    {{&nf, ctx->flag_pos.as<int32_t>() + nt, 4},
                     {reop, ctx->reop64.p, sizeof(reop)}}));
  for (int q = 0; q < 64; ++q) n_reopened += reop[q];
  int64_t n_fb = 0;
  const int64_t n_brute = c_hi;
  if (nf > 0) {
    HIP_TRY(ctx->flag_list.need((size_t)nf * 4));
    hipLaunchKernelGGL(k_flag_list, dim3(ceil_div(nt, 256)), dim3(256), 0, s,
                       nt, ctx->flag32.as<int32_t>(),
                       ctx->flag_pos.as<int32_t>(),
                       ctx->flag_list.as<int32_t>(), INT64_MAX, R.cnt + 7);
    HIP_TRY(ctx->brute_n.need(((size_t)nf + 1) * 8));
    HIP_TRY(ctx->brute_off.need(((size_t)nf + 1) * 8));
    HIP_TRY(hipMemsetAsync(ctx->brute_n.as<int64_t>() + nf, 0, 8, s));
    hipLaunchKernelGGL(k_brute, dim3(nf), dim3(64), 0, s, nf,
                       ctx->flag_list.as<int32_t>(), R.M, R.S, R.inst_d,
                       R.tet_inst_d, (const int64_t*)nullptr,
                       ctx->brute_n.as<int64_t>(), (Cand*)nullptr, (int64_t)0,
                       (const int32_t*)nullptr, INT64_MAX);
    HIP_TRY(hipGetLastError());
    HIP_TRY(exclusive_sum<int64_t>(ctx, ctx->brute_n.as<int64_t>(),
                                   ctx->brute_off.as<int64_t>(), nf + 1));
    HIP_TRY(read_back(ctx, s, {{&n_fb, ctx->brute_off.as<int64_t>() + nf, 8}}));
    HIP_TRY(ctx->cand.need_keep((size_t)(c_hi + n_fb) * sizeof(Cand),
                                (size_t)c_hi * sizeof(Cand), s));
    hipLaunchKernelGGL(k_brute, dim3(nf), dim3(64), 0, s, nf,
                       ctx->flag_list.as<int32_t>(), R.M, R.S, R.inst_d,
                       R.tet_inst_d, ctx->brute_off.as<int64_t>(),
                       (int64_t*)nullptr, ctx->cand.as<Cand>(), c_hi,
                       (const int32_t*)nullptr, INT64_MAX);
    HIP_TRY(hipGetLastError());
    if (int rc = clip_range(ctx, R, c_hi, c_hi + n_fb, false)) return rc;
  }
  R.seen.flagged = nf;
  R.seen.fb = n_fb;
  const int64_t n_all = c_hi + n_fb;
  ctx->st.learned[0] = n1;
  ctx->st.learned[1] = n_reopened;
  ctx->st.learned[2] = added + n_fb;
  ctx->st.learned[3] = nf;
  // output order: A (first round, kept order) merged per tet with the
  // sorted added candidates B
  const int64_t nB = n_all - n1;
  HIP_TRY(ctx->keys_a.need((size_t)std::max<int64_t>(nB, 1) * 8));
  HIP_TRY(ctx->keys_b.need((size_t)std::max<int64_t>(nB, 1) * 8));
  HIP_TRY(ctx->vals_a.need((size_t)std::max<int64_t>(nB, 1) * 4));
  HIP_TRY(ctx->vals_b.need((size_t)std::max<int64_t>(nB, 1) * 4));
  if (nB > 0) {
    hipLaunchKernelGGL(k_b_keys, dim3(ceil_div(nB, 256)), dim3(256), 0, s, n1,
                       n_all, n_brute, ctx->cand.as<Cand>(),
                       ctx->flag.as<uint8_t>(), ctx->keys_a.as<uint64_t>(),
                       ctx->vals_a.as<int32_t>());
    HIP_TRY(hipGetLastError());
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
        nullptr, tmp, ctx->keys_a.as<uint64_t>(), ctx->keys_b.as<uint64_t>(),
        ctx->vals_a.as<int32_t>(), ctx->vals_b.as<int32_t>(), (int)nB, 0, 64,
        s));
    HIP_TRY(ctx->scan_tmp.need(tmp));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
        ctx->scan_tmp.p, tmp, ctx->keys_a.as<uint64_t>(),
        ctx->keys_b.as<uint64_t>(), ctx->vals_a.as<int32_t>(),
        ctx->vals_b.as<int32_t>(), (int)nB, 0, 64, s));
  }
  // the excluded B keys (~0) sort after every real key, so searches over
  // all nB keys find what searches over the n_b real ones would
  const int64_t n_b = nB;
  HIP_TRY(ctx->out_off.need(((size_t)nt + 1) * 8));
  HIP_TRY(ctx->out_cnt.need(((size_t)nt + 1) * 8));
  hipLaunchKernelGGL(k_out_counts, dim3(ceil_div(nt + 1, 256)), dim3(256), 0,
                     s, nt, ctx->flag.as<uint8_t>(), ctx->reopened.as<uint8_t>(),
                     ctx->tet_off.as<int64_t>(), ctx->keys_b.as<uint64_t>(),
                     (int64_t)n_b, ctx->out_cnt.as<int64_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(exclusive_sum<int64_t>(ctx, ctx->out_cnt.as<int64_t>(),
                                 ctx->out_off.as<int64_t>(), nt + 1));
  int64_t n_incl = 0;
  HIP_TRY(read_back(ctx, s, {{&n_incl, ctx->out_off.as<int64_t>() + nt, 8}}));
  R.seen.incl = n_incl;
  HIP_TRY(ctx->perm.need((size_t)std::max<int64_t>(n_incl, 1) * 4));
  if (n1 > 0)
    hipLaunchKernelGGL(k_place_a, dim3(ceil_div(n1, 256)), dim3(256), 0, s, n1,
                       ctx->cand.as<Cand>(), ctx->flag.as<uint8_t>(),
                       ctx->reopened.as<uint8_t>(), ctx->tet_off.as<int64_t>(),
                       ctx->keys_b.as<uint64_t>(), (int64_t)n_b,
                       ctx->out_off.as<int64_t>(), ctx->perm.as<int32_t>(),
                       (const int64_t*)nullptr, INT64_MAX);
  if (n_b > 0)
    hipLaunchKernelGGL(k_place_b, dim3(ceil_div(n_b, 256)), dim3(256), 0, s,
                       (int64_t)n_b, ctx->keys_b.as<uint64_t>(),
                       ctx->vals_b.as<int32_t>(), ctx->flag.as<uint8_t>(),
                       ctx->tet_off.as<int64_t>(), ctx->rel_ids.as<int32_t>(),
                       ctx->out_off.as<int64_t>(), ctx->perm.as<int32_t>(),
                       INT64_MAX);
  HIP_TRY(hipGetLastError());
  // relation CSR of the output sets (tet_off, rel_ids of rpd_download_*)
  HIP_TRY(ctx->out_ids.need((size_t)std::max<int64_t>(n_incl, 1) * 4));
  if (n_incl > 0)
    hipLaunchKernelGGL(k_out_ids, dim3(ceil_div(n_incl, 256)), dim3(256), 0, s,
                       n_incl, ctx->cand.as<Cand>(), ctx->perm.as<int32_t>(),
                       ctx->out_ids.as<int32_t>(), (const int64_t*)nullptr);
  HIP_TRY(hipGetLastError());
  ctx->tet_off.swap(ctx->out_off);
  ctx->rel_ids.swap(ctx->out_ids);
  ctx->tet_count.swap(ctx->out_cnt);
  *n_incl_out = n_incl;
  return 0;
}

// The certificate, closure rounds, fallback and output order of a captured
// learned run (rpd_graph_build): the steps of certify_and_fallback with
// every count left on the device.  Lists are sized by the capacities in
// R.cap and padded (~0 keys sort last); closure round r puts its new
// candidates at [n1_cap + r * added, + its count), the fallback after the
// last round; every count past its capacity sets the overflow counter
// (cnt[7]), and misses left after the last captured round do too unless
// that round is the last the eager path runs.  The output order is the
// eager one: the same (tet, site) pairs in the same merge.  Device counts: gcnt[1 + 2 r] / gcnt[2 + 2 r] the count and the
// end of range r (r = rounds: the fallback), gcnt[3 + 2 rounds + r] the
// merged pairs before round r.
int certify_captured(rpd_ctx* ctx, RunState& R, int64_t n1_cap,
                     int64_t* n_incl_out) {
  hipStream_t s = R.s;
  const int32_t nt = R.nt;
  const RunCaps& C = R.cap;
  const int rounds = (int)C.rounds;
  // buffers at the largest round's capacities; each round's lists, sorts
  // and launches at its own
  const int64_t cap_miss = C.miss, cap_ext = C.ext;
  const int64_t cap_nf = C.flagged, cap_fb = C.fb, cap_incl = C.incl;
  int32_t* ovf = R.cnt + 7;
  auto grid = [](int64_t n) { return dim3(ceil_div(std::max<int64_t>(n, 1), 256)); };
  const int64_t n_gc = 4 + 3 * (int64_t)rounds;
  HIP_TRY(ctx->gcnt.need((size_t)n_gc * 8));
  int64_t* gc = ctx->gcnt.as<int64_t>();
  int64_t* ext_n = gc + 3 + 2 * rounds;  // merged pairs before round r
  HIP_TRY(zero32(s, gc, 2 * n_gc));
  HIP_TRY(ctx->flag32.need(((size_t)nt + 1) * 4));
  HIP_TRY(ctx->flag_pos.need(((size_t)nt + 1) * 4));
  HIP_TRY(ctx->umark.need(((size_t)cap_miss + 1) * 4));
  HIP_TRY(ctx->upos.need(((size_t)cap_miss + 1) * 4));
  HIP_TRY(ctx->ppos.need(((size_t)cap_miss + 1) * 4));
  HIP_TRY(ctx->ukeys.need((size_t)std::max<int64_t>(cap_miss, 1) * 8));
  HIP_TRY(ctx->upass.need((size_t)cap_miss + 1));
  HIP_TRY(ctx->mem_a.need((size_t)std::max<int64_t>(cap_ext, 1) * 8));
  HIP_TRY(ctx->mem_b.need((size_t)std::max<int64_t>(cap_ext, 1) * 8));
  // candidate ranges: the first round, closure round r, the fallback
  RegionTable reg{};
  reg.n = rounds;
  reg.base[0] = n1_cap;
  for (int r = 0; r < rounds; ++r) reg.base[r + 1] = reg.base[r] + C.added_r[r];
  const int64_t fb_base = reg.base[rounds];
  const int64_t c_end = fb_base + cap_fb;
  HIP_TRY(ctx->cand.need_keep((size_t)c_end * sizeof(Cand),
                              (size_t)n1_cap * sizeof(Cand), s));
  const int32_t* nm = R.cnt + 4;
  int32_t* n_eff = R.cnt + 5;  // the misses listed
  // candidates of the launch whose misses are listed
  const int64_t* launch_n = ctx->tet_off.as<int64_t>() + nt;
  for (int r = 0; r < rounds; ++r) {
    const int64_t cm = C.miss_r[r], cn = C.added_r[r], ce = C.ext_r[r];
    const int32_t* n_u = ctx->upos.as<int32_t>() + cm;
    const int64_t c_lo = reg.base[r];
    // unique missing pairs of the round (a padded sort of the miss list)
    hipLaunchKernelGGL(k_pad_miss, grid(cm), dim3(256), 0, s,
                       ctx->miss_a.as<uint64_t>(), cm, nm, launch_n, n_eff,
                       ovf);
    if (int rc = sort_u64(ctx, s, ctx->miss_a.as<uint64_t>(),
                          ctx->miss_b.as<uint64_t>(), cm))
      return rc;
    hipLaunchKernelGGL(k_unique_mark, grid(cm + 1), dim3(256), 0, s, cm,
                       ctx->miss_b.as<uint64_t>(), ctx->umark.as<int32_t>(),
                       (const int32_t*)n_eff);
    HIP_TRY(exclusive_sum<int32_t>(ctx, ctx->umark.as<int32_t>(),
                                   ctx->upos.as<int32_t>(), cm + 1));
    hipLaunchKernelGGL(k_closure_split, grid(cm), dim3(256), 0, s, cm,
                       ctx->miss_b.as<uint64_t>(), ctx->umark.as<int32_t>(),
                       ctx->upos.as<int32_t>(), R.M, R.S,
                       ctx->ukeys.as<uint64_t>(), ctx->upass.as<uint8_t>());
    hipLaunchKernelGGL(k_pass32, grid(cm + 1), dim3(256), 0, s, cm, n_u,
                       ctx->upass.as<uint8_t>(), ctx->umark.as<int32_t>());
    HIP_TRY(exclusive_sum<int32_t>(ctx, ctx->umark.as<int32_t>(),
                                   ctx->ppos.as<int32_t>(), cm + 1));
    // the related ones: new candidates at [c_lo, c_lo + count)
    hipLaunchKernelGGL(k_closure_fill, grid(cm), dim3(256), 0, s, cm,
                       ctx->ukeys.as<uint64_t>(), ctx->upass.as<uint8_t>(),
                       ctx->ppos.as<int32_t>(), ctx->cand.as<Cand>(), c_lo,
                       ctx->reopened.as<uint8_t>(), n_u, cn, ovf);
    // the round's count and end; merged pairs += the unique pairs (disjoint
    // from them), padded, sorted again
    hipLaunchKernelGGL(k_round_close, grid(ce), dim3(256), 0, s,
                       ctx->ppos.as<int32_t>() + cm, c_lo, cn, gc + 1 + 2 * r,
                       gc + 2 + 2 * r, ctx->mem_a.as<uint64_t>(), ce,
                       ctx->ukeys.as<uint64_t>(), n_u,
                       (const int64_t*)(ext_n + r), ext_n + r + 1, ovf);
    HIP_TRY(hipGetLastError());
    if (int rc = sort_u64(ctx, s, ctx->mem_a.as<uint64_t>(),
                          ctx->mem_b.as<uint64_t>(), ce))
      return rc;
    ctx->mem_a.swap(ctx->mem_b);
    // the round's cells, certified as they are clipped (the padded pair
    // list: its ~0 keys match no pair)
    if (int rc = certify_round(ctx, R, 0, ce, gc + 1 + 2 * r,
                               C.miss_r[r + 1]))
      return rc;
    if (int rc = clip_range(ctx, R, c_lo, c_lo + cn, false, gc + 1 + 2 * r,
                            gc + 2 + 2 * r))
      return rc;
    launch_n = gc + 1 + 2 * r;
  }
  const int64_t cm_last = C.miss_r[rounds];
  if (rounds == kClosureRounds) {  // still open: full relation (every
    // missed tet, listed or not, the same ones as the eager path)
    hipLaunchKernelGGL(k_pad_miss, grid(cm_last), dim3(256), 0, s,
                       ctx->miss_a.as<uint64_t>(), cm_last, nm, launch_n,
                       n_eff, (int32_t*)nullptr);
    hipLaunchKernelGGL(k_flag_miss, grid(cm_last), dim3(256), 0, s, cm_last,
                       ctx->miss_a.as<uint64_t>(), ctx->flag.as<uint8_t>(),
                       (const int32_t*)n_eff);
  } else {  // misses left: the run needs more rounds than were captured
    hipLaunchKernelGGL(k_flag_if_positive, dim3(1), dim3(64), 0, s, nm, ovf);
  }
  HIP_TRY(hipGetLastError());
  R.cert.tet_off = nullptr;  // the full-relation re-runs are final
  // tets without a cell, with a dropped cell or left open: full relation
  HIP_TRY(ctx->reop64.need(64 * 4));
  HIP_TRY(zero32(s, ctx->reop64.p, 64));
  hipLaunchKernelGGL(k_flag_final, dim3(ceil_div(nt, 256)), dim3(256), 0, s,
                     nt, ctx->flag.as<uint8_t>(), ctx->has_cell.as<uint8_t>(),
                     ctx->reopened.as<uint8_t>(), ctx->flag32.as<int32_t>(),
                     ctx->reop64.as<int32_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(zero32(s, ctx->flag32.as<int32_t>() + nt, 1));
  HIP_TRY(exclusive_sum<int32_t>(ctx, ctx->flag32.as<int32_t>(),
                                 ctx->flag_pos.as<int32_t>(), nt + 1));
  const int32_t* nf = ctx->flag_pos.as<int32_t>() + nt;
  HIP_TRY(ctx->flag_list.need((size_t)std::max<int64_t>(cap_nf, 1) * 4));
  hipLaunchKernelGGL(k_flag_list, dim3(ceil_div(nt, 256)), dim3(256), 0, s, nt,
                     ctx->flag32.as<int32_t>(), ctx->flag_pos.as<int32_t>(),
                     ctx->flag_list.as<int32_t>(), cap_nf, ovf);
  HIP_TRY(ctx->brute_n.need(((size_t)cap_nf + 1) * 8));
  HIP_TRY(ctx->brute_off.need(((size_t)cap_nf + 1) * 8));
  HIP_TRY(zero32(s, ctx->brute_n.p, 2 * (cap_nf + 1)));
  if (cap_nf > 0)
    hipLaunchKernelGGL(k_brute, dim3(cap_nf), dim3(64), 0, s, (int32_t)cap_nf,
                       ctx->flag_list.as<int32_t>(), R.M, R.S, R.inst_d,
                       R.tet_inst_d, (const int64_t*)nullptr,
                       ctx->brute_n.as<int64_t>(), (Cand*)nullptr, (int64_t)0,
                       nf, INT64_MAX);
  HIP_TRY(hipGetLastError());
  HIP_TRY(exclusive_sum<int64_t>(ctx, ctx->brute_n.as<int64_t>(),
                                 ctx->brute_off.as<int64_t>(), cap_nf + 1));
  hipLaunchKernelGGL(k_range_count<int64_t>, dim3(1), dim3(64), 0, s,
                     ctx->brute_off.as<int64_t>() + cap_nf, fb_base, cap_fb,
                     gc + 1 + 2 * rounds, gc + 2 + 2 * rounds, ovf);
  if (cap_nf > 0)
    hipLaunchKernelGGL(k_brute, dim3(cap_nf), dim3(64), 0, s, (int32_t)cap_nf,
                       ctx->flag_list.as<int32_t>(), R.M, R.S, R.inst_d,
                       R.tet_inst_d, ctx->brute_off.as<int64_t>(),
                       (int64_t*)nullptr, ctx->cand.as<Cand>(), fb_base, nf,
                       c_end);
  HIP_TRY(hipGetLastError());
  if (int rc = clip_range(ctx, R, fb_base, c_end, false, gc + 1 + 2 * rounds,
                          gc + 2 + 2 * rounds))
    return rc;
  // output order: the first round's candidates (kept order) merged per tet
  // with the sorted added ones
  const int64_t nB = c_end - n1_cap;
  HIP_TRY(ctx->keys_a.need((size_t)std::max<int64_t>(nB, 1) * 8));
  HIP_TRY(ctx->keys_b.need((size_t)std::max<int64_t>(nB, 1) * 8));
  HIP_TRY(ctx->vals_a.need((size_t)std::max<int64_t>(nB, 1) * 4));
  HIP_TRY(ctx->vals_b.need((size_t)std::max<int64_t>(nB, 1) * 4));
  if (nB > 0) {
    hipLaunchKernelGGL(k_b_keys_ranges, grid(nB), dim3(256), 0, s, reg,
                       c_end, (const int64_t*)(gc + 1), ctx->cand.as<Cand>(),
                       ctx->flag.as<uint8_t>(), ctx->keys_a.as<uint64_t>(),
                       ctx->vals_a.as<int32_t>());
    HIP_TRY(hipGetLastError());
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
        nullptr, tmp, ctx->keys_a.as<uint64_t>(), ctx->keys_b.as<uint64_t>(),
        ctx->vals_a.as<int32_t>(), ctx->vals_b.as<int32_t>(), (int)nB, 0, 64,
        s));
    HIP_TRY(ctx->scan_tmp.need(tmp));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
        ctx->scan_tmp.p, tmp, ctx->keys_a.as<uint64_t>(),
        ctx->keys_b.as<uint64_t>(), ctx->vals_a.as<int32_t>(),
        ctx->vals_b.as<int32_t>(), (int)nB, 0, 64, s));
  }
  HIP_TRY(ctx->out_off.need(((size_t)nt + 1) * 8));
  HIP_TRY(ctx->out_cnt.need(((size_t)nt + 1) * 8));
  hipLaunchKernelGGL(k_out_counts, dim3(ceil_div(nt + 1, 256)), dim3(256), 0,
                     s, nt, ctx->flag.as<uint8_t>(), ctx->reopened.as<uint8_t>(),
                     ctx->tet_off.as<int64_t>(), ctx->keys_b.as<uint64_t>(), nB,
                     ctx->out_cnt.as<int64_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(exclusive_sum<int64_t>(ctx, ctx->out_cnt.as<int64_t>(),
                                 ctx->out_off.as<int64_t>(), nt + 1));
  const int64_t* n_incl = ctx->out_off.as<int64_t>() + nt;
  HIP_TRY(ctx->perm.need((size_t)std::max<int64_t>(cap_incl, 1) * 4));
  // (cleared: after an overflow some positions stay unplaced, and later
  // launches then read candidate 0 there)
  HIP_TRY(zero32(s, ctx->perm.p, std::max<int64_t>(cap_incl, 1)));
  if (n1_cap > 0)
    hipLaunchKernelGGL(k_place_a, grid(n1_cap), dim3(256), 0, s, n1_cap,
                       ctx->cand.as<Cand>(), ctx->flag.as<uint8_t>(),
                       ctx->reopened.as<uint8_t>(), ctx->tet_off.as<int64_t>(),
                       ctx->keys_b.as<uint64_t>(), nB,
                       ctx->out_off.as<int64_t>(), ctx->perm.as<int32_t>(),
                       (const int64_t*)(ctx->tet_off.as<int64_t>() + nt),
                       cap_incl);
  if (nB > 0)
    hipLaunchKernelGGL(k_place_b, grid(nB), dim3(256), 0, s, nB,
                       ctx->keys_b.as<uint64_t>(), ctx->vals_b.as<int32_t>(),
                       ctx->flag.as<uint8_t>(), ctx->tet_off.as<int64_t>(),
                       ctx->rel_ids.as<int32_t>(), ctx->out_off.as<int64_t>(),
                       ctx->perm.as<int32_t>(), cap_incl);
  HIP_TRY(hipGetLastError());
  HIP_TRY(ctx->out_ids.need((size_t)std::max(cap_incl, n1_cap) * 4));
  if (cap_incl > 0)
    hipLaunchKernelGGL(k_out_ids, grid(cap_incl), dim3(256), 0, s, cap_incl,
                       ctx->cand.as<Cand>(), ctx->perm.as<int32_t>(),
                       ctx->out_ids.as<int32_t>(), n_incl);
  HIP_TRY(hipGetLastError());
  ctx->tet_off.swap(ctx->out_off);
  ctx->rel_ids.swap(ctx->out_ids);
  ctx->tet_count.swap(ctx->out_cnt);
  *n_incl_out = cap_incl;
  return 0;
}

// sizes, scan and copy of the emitted cells of candidates perm[0..n)
// (identity if perm is null) into the output stream
int compact(rpd_ctx* ctx, RunState& R, int64_t n, const int32_t* perm,
            Sizes3* tot) {
  hipStream_t s = R.s;
  TierViews tv = tier_views(ctx, R);
  const int64_t nb = std::max<int64_t>(ceil_div(n, kCB), 0);
  HIP_TRY(ctx->sizes.need(((size_t)nb + 1) * sizeof(Sizes3)));
  HIP_TRY(ctx->offs.need(((size_t)nb + 1) * sizeof(Sizes3)));
  HIP_TRY(ctx->drops.need(kDropSlots * 9 * 8));
  HIP_TRY(zero32(s, ctx->drops.p, kDropSlots * 9 * 2));
  HIP_TRY(zero32(s, ctx->sizes.as<Sizes3>() + nb, sizeof(Sizes3) / 4));
  // (a captured run: the cell count is tet_off[n_tet] on the device)
  const int64_t* n_dev =
      R.graph ? ctx->tet_off.as<int64_t>() + R.nt : nullptr;
  if (nb > 0)
    hipLaunchKernelGGL(k_block_sizes, dim3(nb), dim3(kCB), 0, s, n, perm,
                       ctx->loc.as<int32_t>(), tv, ctx->sizes.as<Sizes3>(),
                       ctx->drops.as<int64_t>(), n_dev);
  HIP_TRY(hipGetLastError());
  {
    size_t tmp = 0;
    Sizes3 zero{0, 0, 0};
    HIP_TRY(hipcub::DeviceScan::ExclusiveScan(
        nullptr, tmp, ctx->sizes.as<Sizes3>(), ctx->offs.as<Sizes3>(),
        Sizes3Sum(), zero, (int)(nb + 1), s));
    HIP_TRY(ctx->scan_tmp.need(tmp));
    HIP_TRY(hipcub::DeviceScan::ExclusiveScan(
        ctx->scan_tmp.p, tmp, ctx->sizes.as<Sizes3>(), ctx->offs.as<Sizes3>(),
        Sizes3Sum(), zero, (int)(nb + 1), s));
  }
  if (R.graph)
    *tot = Sizes3{R.cap.cells, R.cap.words, R.cap.areas};
  else
    HIP_TRY(read_back(ctx, s,
                      {{tot, ctx->offs.as<Sizes3>() + nb, sizeof(Sizes3)}}));
  R.seen.cells = tot->cells;
  R.seen.words = tot->words;
  R.seen.areas = tot->areas;
  HIP_TRY(ctx->out_cells.need((size_t)std::max<int64_t>(tot->cells, 1) *
                              sizeof(rpd_cell)));
  HIP_TRY(ctx->out_words.need((size_t)std::max<int64_t>(tot->words, 1) * 4));
  HIP_TRY(ctx->out_geom.need((size_t)std::max<int64_t>(tot->cells, 1) * 32));
  HIP_TRY(ctx->out_areas.need((size_t)std::max<int64_t>(tot->areas, 1) * 8));
  if (nb > 0)
    hipLaunchKernelGGL(k_compact, dim3(nb), dim3(kCB), 0, s, n, perm,
                       ctx->cand.as<Cand>(), ctx->loc.as<int32_t>(), tv,
                       ctx->offs.as<Sizes3>(), ctx->out_cells.as<rpd_cell>(),
                       ctx->out_words.as<uint32_t>(),
                       ctx->out_geom.as<double>(), ctx->out_areas.as<double>(),
                       R.cfg.compute_geometry, n_dev, *tot, R.cnt + 7);
  HIP_TRY(hipGetLastError());
  return 0;
}

// completes and checks a run configuration into R.cfg, R.sc, R.G
int run_config(rpd_ctx* ctx, const rpd_run_cfg* cfg_in, RunState& R) {
  rpd_run_cfg& cfg = R.cfg;
  cfg = rpd_run_cfg{};
  cfg.cells_per_wave = 0;
  cfg.use_active_mask = 1;
  cfg.compute_geometry = 1;
  cfg.small_tier = 1;
  if (cfg_in) cfg = *cfg_in;
  if (cfg.cells_per_wave <= 0) {  // tuned shape, else the default
    cfg.cells_per_wave = ctx->tuned_G > 0 ? ctx->tuned_G : 8;
    if (ctx->tuned_G > 0) cfg.small_caps = ctx->tuned_caps;
  }
  if (cfg.small_caps < 0 || cfg.small_caps > 3)
    return fail(ctx, "rpd_run: small_caps must be 0, 1, 2 or 3");
  R.sc = small_caps(cfg.small_caps);
  R.G = cfg.cells_per_wave;
  if (R.G != 1 && R.G != 2 && R.G != 4 && R.G != 8 && R.G != 16)
    return fail(ctx, "rpd_run: cells_per_wave must be 1, 2, 4, 8 or 16");
  if (cfg.small_caps >= 2 && R.G < 4)
    return fail(ctx, "rpd_run: small_caps 2 and 3 need cells_per_wave >= 4");
  if (cfg.learned && ctx->has_subset)
    return fail(ctx, "rpd_run: a tet subset is for the plain path only "
                     "(learned = 0)");
  return 0;
}

// the device stages of one run: candidates, clip tiers, (learned)
// certificate and fallback, compaction.  A captured run (R.graph) records
// no events and reads nothing back.
int run_stages(rpd_ctx* ctx, RunState& R, int64_t* n_cand_out,
               int64_t* n_out_out, Sizes3* tot) {
  const rpd_run_cfg& cfg = R.cfg;
  HIP_TRY(ctx->counters.need(8 * 4));
  R.s = ctx->stream;
  hipStream_t s = R.s;
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[0], s));
  ctx->n_small_ev = 0;
  ctx->n_reads = 0;
  HIP_TRY(zero32(s, ctx->counters.p, 8));
  if (int rc = init_state(ctx, R)) return rc;
  if (cfg.count_work) {
    HIP_TRY(ctx->rel_work.need(64 * 4 * 8));
    HIP_TRY(hipMemsetAsync(ctx->rel_work.p, 0, 64 * 4 * 8, s));
    HIP_TRY(ctx->work.need(3 * 64 * 4 * 8));
    HIP_TRY(hipMemsetAsync(ctx->work.p, 0, 3 * 64 * 4 * 8, s));
  }

  // ---- 1. candidates
  int64_t n_cand = 0;
  if (cfg.learned) {
    if (int rc = learned_candidates(ctx, R, &n_cand)) return rc;
  } else {
    if (int rc = plain_candidates(ctx, R, &n_cand)) return rc;
  }
  // ---- 2. clipping: small tier, then full capacity, then exact (learned:
  // each cell certified as it is clipped)
  if (cfg.learned)
    if (int rc = certify_begin(ctx, R, n_cand)) return rc;
  if (int rc = clip_range(ctx, R, 0, n_cand, !R.graph)) return rc;
  // ---- 3. (learned) certificate and fallback; 4. compaction
  int64_t n_out = n_cand;
  const int32_t* perm = nullptr;
  if (cfg.learned) {
    if (int rc = certify_and_fallback(ctx, R, n_cand, &n_out)) return rc;
    perm = ctx->perm.as<int32_t>();
  }
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[5], s));
  if (int rc = compact(ctx, R, n_out, perm, tot)) return rc;
  if (!R.graph) HIP_TRY(hipEventRecord(ctx->ev[4], s));
  // tet_k: max candidates of one tet, into stats64[2]
  HIP_TRY(ctx->stats64.need(4 * 8));
  HIP_TRY(zero32(s, ctx->stats64.p, 8));
  {
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceReduce::Max(nullptr, tmp, ctx->tet_count.as<int64_t>(),
                                      ctx->stats64.as<int64_t>() + 2, R.nt, s));
    HIP_TRY(ctx->scan_tmp.need(tmp));
    HIP_TRY(hipcub::DeviceReduce::Max(ctx->scan_tmp.p, tmp,
                                      ctx->tet_count.as<int64_t>(),
                                      ctx->stats64.as<int64_t>() + 2, R.nt, s));
  }
  *n_cand_out = n_cand;
  *n_out_out = n_out;
  return 0;
}

// statistics of a finished eager run (outside the timed phases)
int run_stats(rpd_ctx* ctx, RunState& R, int64_t n_cand, int64_t n_out,
              const Sizes3& tot) {
  const rpd_run_cfg& cfg = R.cfg;
  hipStream_t s = R.s;
  if (n_cand > 0 && cfg.count_work)
    hipLaunchKernelGGL(k_mask_stats, dim3(ceil_div(n_cand, 256)), dim3(256), 0,
                       s, n_cand, ctx->cand.as<Cand>(), R.S,
                       ctx->stats64.as<unsigned long long>());
  int64_t st64[4], drops[kDropSlots * 9];
  int32_t host_cnt[8];
  HIP_TRY(read_back(ctx, s,
                    {{st64, ctx->stats64.p, 32},
                     {drops, ctx->drops.p, sizeof(drops)},
                     {host_cnt, R.cnt, 32}}));
  rpd_stats& st = ctx->st;
  st.n_cells = tot.cells;
  st.n_candidates = n_out;
  st.n_words = tot.words;
  st.n_areas = tot.areas;
  for (int q = 0; q < 9; ++q) {
    st.n_dropped[q] = 0;
    for (int b = 0; b < kDropSlots; ++b) st.n_dropped[q] += drops[9 * b + q];
  }
  st.n_rerun_large = R.r_large;
  st.n_rerun_exact = R.r_exact;
  st.n_active_planes = st64[0];
  st.n_all_planes = st64[1];
  if (cfg.count_work) {
    std::vector<uint64_t> w(3 * 64 * 4);
    HIP_TRY(hipMemcpy(w.data(), ctx->work.p, w.size() * 8,
                      hipMemcpyDeviceToHost));
    for (size_t q = 0; q < w.size(); ++q)
      st.work[q / 256][q % 4] += (int64_t)w[q];
    std::vector<uint64_t> rw(64 * 4);
    HIP_TRY(hipMemcpy(rw.data(), ctx->rel_work.p, rw.size() * 8,
                      hipMemcpyDeviceToHost));
    for (size_t q = 0; q < rw.size(); ++q)
      st.relation_work[q % 4] += (int64_t)rw[q];
    st.relation_work[3] = n_cand;
  }
  float ms = 0;
  auto el = [&](int a, int b) -> double {
    float m = 0;
    return hipEventElapsedTime(&m, ctx->ev[a], ctx->ev[b]) == hipSuccess ? m
                                                                         : -1;
  };
  (void)ms;
  st.ms_total = el(0, 4);
  st.ms_phase[0] = el(0, 1);
  st.ms_phase[1] = el(1, 2);
  st.ms_phase[2] = el(2, 3);
  st.ms_phase[3] = el(5, 4);
  st.ms_phase[4] = el(3, 5);
  // kernel-level: plain: site boxes, binning, relation count, relation
  // fill, masks; learned: features, MLP, proposal count, proposal fill,
  // masks
  st.ms_kernel[0] = el(0, 10);
  st.ms_kernel[1] = el(10, 11);
  st.ms_kernel[2] = el(11, 12);
  st.ms_kernel[3] = el(12, 13);
  st.ms_kernel[4] = el(13, 1);
  st.ms_kernel[5] = st.ms_phase[1];
  st.ms_kernel[6] = st.ms_phase[3];
  st.ms_kernel[7] = st.ms_phase[4];
  st.n_bbox_fallback = host_cnt[0];
  st.tet_k = (int32_t)st64[2];
  st.ms_features_kernel = cfg.learned ? el(14, 15) : 0.0;
  st.n_clip_small = ctx->n_small_ev;
  st.n_host_reads = ctx->n_reads;
  st.ms_clip_small_sum = 0;
  for (int k = 0; k < ctx->n_small_ev; ++k)
    st.ms_clip_small_sum += el(16 + 2 * k, 17 + 2 * k);
  return 0;
}

}  // namespace

extern "C" {

int rpd_run(rpd_ctx* ctx, const rpd_run_cfg* cfg_in) {
  if (!ctx) return -1;
  if (!ctx->has_mesh || !ctx->has_sites)
    return fail(ctx, "rpd_run: no resident instance");
  HIP_TRY(hipSetDevice(ctx->device));
  RunState R;
  if (int rc = run_config(ctx, cfg_in, R)) return rc;
  ctx->last_geometry = false;
  ctx->has_results = false;
  ctx->adj_ready = false;
  const int32_t site_k = ctx->st.site_k;
  std::memset(&ctx->st, 0, sizeof(ctx->st));
  ctx->st.site_k = site_k;
  int64_t n_cand = 0, n_out = 0;
  Sizes3 tot{0, 0, 0};
  if (int rc = run_stages(ctx, R, &n_cand, &n_out, &tot)) return rc;
  if (int rc = run_stats(ctx, R, n_cand, n_out, tot)) return rc;
  ctx->seen = R.seen;
  ctx->last_geometry = R.cfg.compute_geometry != 0;
  ctx->has_results = true;
  return 0;
}

}  // extern "C"

namespace {
// the last run's cells grouped by seed: cs_vals_b = cell indices sorted by
// seed (stable: ascending within a seed), cs_start[s] = first position of
// seed s (cs_start[n_site] = n_cells)
int cells_by_seed(rpd_ctx* ctx) {
  hipStream_t s = ctx->stream;
  const int32_t ns = ctx->n_site;
  const int64_t n = ctx->st.n_cells;
  if (n >= (1LL << 31)) return fail(ctx, "too many cells");
  HIP_TRY(ctx->cs_keys_a.need((size_t)std::max<int64_t>(n, 1) * 4));
  HIP_TRY(ctx->cs_keys_b.need((size_t)std::max<int64_t>(n, 1) * 4));
  HIP_TRY(ctx->cs_vals_a.need((size_t)std::max<int64_t>(n, 1) * 4));
  HIP_TRY(ctx->cs_vals_b.need((size_t)std::max<int64_t>(n, 1) * 4));
  HIP_TRY(ctx->cs_start.need(((size_t)ns + 1) * 8));
  if (n > 0) {
    hipLaunchKernelGGL(k_cell_seed_keys, dim3(ceil_div(n, 256)), dim3(256), 0,
                       s, n, ctx->out_cells.as<rpd_cell>(),
                       ctx->cs_keys_a.as<uint32_t>(),
                       ctx->cs_vals_a.as<int32_t>());
    HIP_TRY(hipGetLastError());
    int end_bit = 1;
    while (end_bit < 32 && (1LL << end_bit) <= (int64_t)ns) ++end_bit;
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
        nullptr, tmp, ctx->cs_keys_a.as<uint32_t>(),
        ctx->cs_keys_b.as<uint32_t>(), ctx->cs_vals_a.as<int32_t>(),
        ctx->cs_vals_b.as<int32_t>(), (int)n, 0, end_bit, s));
    HIP_TRY(ctx->scan_tmp.need(tmp));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
        ctx->scan_tmp.p, tmp, ctx->cs_keys_a.as<uint32_t>(),
        ctx->cs_keys_b.as<uint32_t>(), ctx->cs_vals_a.as<int32_t>(),
        ctx->cs_vals_b.as<int32_t>(), (int)n, 0, end_bit, s));
  }
  hipLaunchKernelGGL(k_seed_starts, dim3(ceil_div((int64_t)ns + 1, 256)),
                     dim3(256), 0, s, ns, n, ctx->cs_keys_b.as<uint32_t>(),
                     ctx->cs_start.as<int64_t>());
  HIP_TRY(hipGetLastError());
  return 0;
}

// stable sort of (u64 key, i32 value) pairs over the low end_bit key bits;
// results in ka/va (the b buffers are the sort's other half)
int sort_pairs64(rpd_ctx* ctx, Buf& ka, Buf& kb, Buf& va, Buf& vb, int64_t n,
                 int end_bit) {
  if (n <= 1) return 0;
  if (n >= (1LL << 31)) return fail(ctx, "too many records to sort");
  hipStream_t s = ctx->stream;
  size_t tmp = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
      nullptr, tmp, ka.as<uint64_t>(), kb.as<uint64_t>(), va.as<int32_t>(),
      vb.as<int32_t>(), (int)n, 0, end_bit, s));
  HIP_TRY(ctx->scan_tmp.need(tmp));
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
      ctx->scan_tmp.p, tmp, ka.as<uint64_t>(), kb.as<uint64_t>(),
      va.as<int32_t>(), vb.as<int32_t>(), (int)n, 0, end_bit, s));
  ka.swap(kb);
  va.swap(vb);
  return 0;
}

}  // namespace

extern "C" {

int rpd_adjacency(rpd_ctx* ctx, rpd_adj_sizes* sizes) {
  if (!ctx) return -1;
  if (!ctx->has_results)
    return fail(ctx, "rpd_adjacency: no run on the resident instance");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  ctx->adj_ready = false;
  const int64_t n = ctx->st.n_cells;
  const int32_t ns = ctx->n_site;
  if (n >= (1LL << 29))
    return fail(ctx, "rpd_adjacency: too many cells (4 per cell must fit "
                     "32 bits)");
  DevMesh M{ctx->n_vert, ctx->n_tet, ctx->vert.as<double4>(),
            ctx->tet_v.as<int4>(), ctx->tet_fid.as<int4>(),
            ctx->tet_meta.as<TetMeta>()};
  DevSites S{ns,
             ctx->site.as<double4>(),
             ctx->flags.as<uint32_t>(),
             ctx->knn_off.as<int64_t>(),
             ctx->knn_ids.as<int32_t>(),
             nullptr,
             nullptr,
             nullptr};
  const size_t n1 = (size_t)std::max<int64_t>(n, 1);
  HIP_TRY(ctx->adj_act.need(n1 * 32));
  HIP_TRY(ctx->adj_euler.need(n1 * 4));
  HIP_TRY(ctx->adj_cnt.need(3 * (n1 + 1) * 8));
  HIP_TRY(ctx->adj_off.need(3 * (n1 + 1) * 8));
  HIP_TRY(ctx->adj_nbr.need(n1 * 16));
  HIP_TRY(ctx->adj_cpar.need(n1 * 4));
  HIP_TRY(ctx->adj_err.need(16));
  HIP_TRY(hipMemsetAsync(ctx->adj_err.p, 0, 16, s));
  {  // first site of each instance
    std::vector<int64_t> so(std::max<size_t>(ctx->inst.size(), 1), 0);
    for (size_t b = 0; b < ctx->inst.size(); ++b) so[b] = ctx->inst[b].site_off;
    HIP_TRY(ctx->adj_soff.need(so.size() * 8));
    HIP_TRY(hipMemcpyAsync(ctx->adj_soff.p, so.data(), so.size() * 8,
                           hipMemcpyHostToDevice, s));
    HIP_TRY(wait_stream(ctx, s));  // so outlives the copy
  }
  int64_t* cnt = ctx->adj_cnt.as<int64_t>();
  int64_t* off = ctx->adj_off.as<int64_t>();
  const size_t stride = n1 + 1;  // three count / offset arrays of n+1
  HIP_TRY(hipMemsetAsync(cnt, 0, 3 * stride * 8, s));
  const rpd_cell* cells = ctx->out_cells.as<rpd_cell>();
  const uint32_t* words = ctx->out_words.as<uint32_t>();
  const int nblk = ceil_div(n, kAdjPerBlock);
  if (n > 0)
    hipLaunchKernelGGL(k_adj_cell, dim3(nblk), dim3(64 * kAdjPerBlock), 0, s,
                       n, cells, words, M, ctx->adj_act.as<uint64_t>(),
                       ctx->adj_euler.as<float>(), cnt, cnt + stride,
                       cnt + 2 * stride);
  HIP_TRY(hipGetLastError());
  for (int q = 0; q < 3; ++q)
    HIP_TRY(exclusive_sum<int64_t>(ctx, cnt + q * stride, off + q * stride,
                                   n + 1));
  int64_t tot[3];
  HIP_TRY(read_back(ctx, s,
                    {{&tot[0], off + n, 8},
                     {&tot[1], off + stride + n, 8},
                     {&tot[2], off + 2 * stride + n, 8}}));
  const int64_t n_tf = tot[0], n_fac = tot[1], n_edg = tot[2];
  if (n_fac >= (1LL << 31) || n_edg >= (1LL << 31) || n_tf >= (1LL << 31))
    return fail(ctx, "rpd_adjacency: too many records");
  auto need = [](Buf& b, int64_t m, size_t w) {
    return b.need((size_t)std::max<int64_t>(m, 1) * w);
  };
  HIP_TRY(need(ctx->tf_ka, n_tf, 8));
  HIP_TRY(need(ctx->tf_kb, n_tf, 8));
  HIP_TRY(need(ctx->tf_va, n_tf, 4));
  HIP_TRY(need(ctx->tf_vb, n_tf, 4));
  HIP_TRY(need(ctx->fac_ka, n_fac, 8));
  HIP_TRY(need(ctx->fac_kb, n_fac, 8));
  HIP_TRY(need(ctx->fac_va, n_fac, 4));
  HIP_TRY(need(ctx->fac_vb, n_fac, 4));
  HIP_TRY(need(ctx->edg_ka, n_edg, 8));
  HIP_TRY(need(ctx->edg_kb, n_edg, 8));
  HIP_TRY(need(ctx->edg_va, n_edg, 4));
  HIP_TRY(need(ctx->edg_vb, n_edg, 4));
  if (n > 0)
    hipLaunchKernelGGL(k_adj_fill, dim3(nblk), dim3(64 * kAdjPerBlock), 0, s,
                       n, cells, words, M, S, ctx->adj_act.as<uint64_t>(), off,
                       off + stride, off + 2 * stride,
                       ctx->tf_ka.as<uint64_t>(), ctx->tf_va.as<int32_t>(),
                       ctx->fac_ka.as<uint64_t>(), ctx->fac_va.as<int32_t>(),
                       ctx->edg_ka.as<uint64_t>(), ctx->edg_va.as<int32_t>(),
                       ctx->adj_err.as<int32_t>());
  HIP_TRY(hipGetLastError());
  // key widths: sites need sb bits
  int sb = 1;
  while (sb < 40 && (1LL << sb) <= (int64_t)std::max(ns, 1)) ++sb;
  if (int rc = sort_pairs64(ctx, ctx->tf_ka, ctx->tf_kb, ctx->tf_va,
                            ctx->tf_vb, n_tf, std::min(64, 32 + sb)))
    return rc;
  if (int rc = sort_pairs64(ctx, ctx->fac_ka, ctx->fac_kb, ctx->fac_va,
                            ctx->fac_vb, n_fac, std::min(64, 32 + sb)))
    return rc;
  if (int rc = sort_pairs64(ctx, ctx->edg_ka, ctx->edg_kb, ctx->edg_va,
                            ctx->edg_vb, n_edg,
                            std::min(64, 2 * kEdgePosBits + sb)))
    return rc;
  // cell neighbours across tet faces, and cell components
  int32_t* nbr = ctx->adj_nbr.as<int32_t>();
  HIP_TRY(hipMemsetAsync(nbr, 0xff, n1 * 16, s));
  int32_t* err = ctx->adj_err.as<int32_t>();
  if (n_tf > 1)
    hipLaunchKernelGGL(k_adj_pairs, dim3(ceil_div(n_tf, 256)), dim3(256), 0,
                       s, n_tf, ctx->tf_ka.as<uint64_t>(),
                       ctx->tf_va.as<int32_t>(), nbr, err + 2, err);
  HIP_TRY(hipGetLastError());
  int32_t* cpar = ctx->adj_cpar.as<int32_t>();
  if (n > 0) {
    hipLaunchKernelGGL(k_uf_init, dim3(ceil_div(n, 256)), dim3(256), 0, s, n,
                       cpar);
    hipLaunchKernelGGL(k_uf_cells, dim3(ceil_div(n, 256)), dim3(256), 0, s, n,
                       nbr, cpar);
    hipLaunchKernelGGL(k_uf_flatten, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                       n, cpar);
  }
  HIP_TRY(hipGetLastError());
  // facet and edge records: components, runs
  struct Kind {
    int64_t m;
    Buf *k, *v, *par, *flag, *pos, *start, *runs, *rec;
    int width;
  } kinds[2] = {{n_fac, &ctx->fac_ka, &ctx->fac_va, &ctx->fac_par,
                 &ctx->fac_flag, &ctx->fac_pos, &ctx->fac_start,
                 &ctx->fac_runs, &ctx->fac_rec, 5},
                {n_edg, &ctx->edg_ka, &ctx->edg_va, &ctx->edg_par,
                 &ctx->edg_flag, &ctx->edg_pos, &ctx->edg_start,
                 &ctx->edg_runs, &ctx->edg_rec, 6}};
  int64_t n_runs[2] = {0, 0};
  for (int kd = 0; kd < 2; ++kd) {
    Kind& K = kinds[kd];
    const int64_t m = K.m;
    HIP_TRY(need(*K.par, m, 4));
    HIP_TRY(need(*K.flag, m + 1, 8));
    HIP_TRY(need(*K.pos, m + 1, 8));
    HIP_TRY(need(*K.rec, m, 8));
    if (m > 0) {
      hipLaunchKernelGGL(k_uf_init, dim3(ceil_div(m, 256)), dim3(256), 0, s,
                         m, K.par->as<int32_t>());
      hipLaunchKernelGGL(k_uf_records, dim3(ceil_div(m, 256)), dim3(256), 0,
                         s, m, K.k->as<uint64_t>(), K.v->as<int32_t>(), nbr,
                         K.par->as<int32_t>());
      hipLaunchKernelGGL(k_uf_flatten, dim3(ceil_div(m, 256)), dim3(256), 0,
                         s, m, K.par->as<int32_t>());
      hipLaunchKernelGGL(k_run_flags, dim3(ceil_div(m, 256)), dim3(256), 0, s,
                         m, K.k->as<uint64_t>(), K.flag->as<int64_t>());
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemsetAsync(K.flag->as<int64_t>() + m, 0, 8, s));
    HIP_TRY(exclusive_sum<int64_t>(ctx, K.flag->as<int64_t>(),
                                   K.pos->as<int64_t>(), m + 1));
    HIP_TRY(read_back(ctx, s, {{&n_runs[kd], K.pos->as<int64_t>() + m, 8}}));
    HIP_TRY(need(*K.start, n_runs[kd], 8));
    HIP_TRY(need(*K.runs, n_runs[kd], 4 * K.width));
    if (kd == 0) {
      HIP_TRY(need(ctx->fac_reul, n_runs[kd], 8));
      HIP_TRY(need(ctx->fac_eul, m, 4));
    }
    if (m > 0) {
      hipLaunchKernelGGL(k_run_starts, dim3(ceil_div(m, 256)), dim3(256), 0,
                         s, m, K.flag->as<int64_t>(), K.pos->as<int64_t>(),
                         K.start->as<int64_t>());
      if (kd == 0)
        hipLaunchKernelGGL(k_fac_euler, dim3(ceil_div(m, 256)), dim3(256), 0,
                           s, m, K.k->as<uint64_t>(), K.v->as<int32_t>(),
                           cells, words, M, ctx->adj_act.as<uint64_t>(),
                           ctx->site_inst.as<int32_t>(),
                           ctx->adj_soff.as<int64_t>(),
                           ctx->fac_eul.as<float>());
      hipLaunchKernelGGL(
          k_adj_runs, dim3(ceil_div(n_runs[kd], 256)), dim3(256), 0, s,
          n_runs[kd], m, kd, K.start->as<int64_t>(), K.k->as<uint64_t>(),
          K.v->as<int32_t>(), K.par->as<int32_t>(),
          kd == 0 ? ctx->fac_eul.as<float>() : nullptr, S,
          K.runs->as<int32_t>(),
          kd == 0 ? ctx->fac_reul.as<double>() : nullptr,
          K.rec->as<int32_t>());
      HIP_TRY(hipGetLastError());
    }
  }
  // per site: cells, components, Euler number
  if (int rc = cells_by_seed(ctx)) return rc;
  const size_t ns1 = (size_t)std::max(ns, 1);
  HIP_TRY(ctx->site_cells.need(ns1 * 4));
  HIP_TRY(ctx->site_cc.need(ns1 * 4));
  HIP_TRY(ctx->site_eul.need(ns1 * 8));
  HIP_TRY(ctx->adj_eul_s.need(n1 * 4));
  HIP_TRY(ctx->adj_root_s.need(n1));
  if (n > 0)
    hipLaunchKernelGGL(k_adj_gather, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                       n, ctx->cs_vals_b.as<int32_t>(), cpar,
                       ctx->adj_euler.as<float>(),
                       ctx->adj_eul_s.as<float>(),
                       ctx->adj_root_s.as<uint8_t>());
  if (ns > 0)
    hipLaunchKernelGGL(k_adj_sites, dim3(ceil_div(ns, 256)), dim3(256), 0, s,
                       ns, ctx->cs_start.as<int64_t>(),
                       ctx->adj_eul_s.as<float>(),
                       ctx->adj_root_s.as<uint8_t>(),
                       ctx->site_cells.as<int32_t>(),
                       ctx->site_cc.as<int32_t>(),
                       ctx->site_eul.as<double>());
  HIP_TRY(hipGetLastError());
  int32_t herr[4];
  HIP_TRY(read_back(ctx, s, {{herr, err, 16}}));
  if (herr[0])
    return fail(ctx, "rpd_adjacency: a bisector's site is missing from the "
                     "seed's neighbour list, or sits beyond position " +
                         std::to_string((1 << kEdgePosBits) - 1));
  if (herr[1])
    return fail(ctx, "rpd_adjacency: a tet face is shared by more than two "
                     "cells of one sphere");
  rpd_adj_sizes& z = ctx->adj;
  z.n_cells = n;
  z.n_sites = ns;
  z.n_cell_pairs = herr[2];
  z.n_facets = n_runs[0];
  z.n_facet_cells = n_fac;
  z.n_edges = n_runs[1];
  z.n_edge_cells = n_edg;
  ctx->adj_ready = true;
  if (sizes) *sizes = z;
  return 0;
}

int rpd_adjacency_download(rpd_ctx* ctx, const rpd_adj_out* o) {
  if (!ctx || !o) return -1;
  if (!ctx->adj_ready)
    return fail(ctx, "rpd_adjacency_download: call rpd_adjacency first");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const rpd_adj_sizes& z = ctx->adj;
  auto get = [&](void* dst, const Buf& src, size_t bytes) -> hipError_t {
    if (!dst || bytes == 0) return hipSuccess;
    return hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, s);
  };
  HIP_TRY(get(o->cell_euler, ctx->adj_euler, (size_t)z.n_cells * 4));
  HIP_TRY(get(o->cell_cc, ctx->adj_cpar, (size_t)z.n_cells * 4));
  HIP_TRY(get(o->cell_nbr, ctx->adj_nbr, (size_t)z.n_cells * 16));
  HIP_TRY(get(o->site_cells, ctx->site_cells, (size_t)z.n_sites * 4));
  HIP_TRY(get(o->site_cc, ctx->site_cc, (size_t)z.n_sites * 4));
  HIP_TRY(get(o->site_euler, ctx->site_eul, (size_t)z.n_sites * 8));
  HIP_TRY(get(o->facets, ctx->fac_runs, (size_t)z.n_facets * 20));
  HIP_TRY(get(o->facet_euler, ctx->fac_reul, (size_t)z.n_facets * 8));
  HIP_TRY(get(o->facet_cells, ctx->fac_rec, (size_t)z.n_facet_cells * 8));
  HIP_TRY(get(o->facet_cell_euler, ctx->fac_eul,
              (size_t)z.n_facet_cells * 4));
  HIP_TRY(get(o->edges, ctx->edg_runs, (size_t)z.n_edges * 24));
  HIP_TRY(get(o->edge_cells, ctx->edg_rec, (size_t)z.n_edge_cells * 8));
  HIP_TRY(wait_stream(ctx, s));
  return 0;
}

}  // extern "C"

namespace {

// what a captured graph depends on besides the data it reads: the batch's
// per-instance tet and site counts, the tet subset switch, the pinned
// instance table and every buffer address (g_buf_generation)
std::vector<int64_t> graph_signature(const rpd_ctx* ctx) {
  std::vector<int64_t> v;
  v.push_back((int64_t)ctx->inst.size());
  for (const auto& I : ctx->inst) {
    v.push_back(I.n_tet);
    v.push_back(I.n_site);
  }
  v.push_back(ctx->has_subset);
  v.push_back((int64_t)(intptr_t)ctx->pin_inst);
  v.push_back((int64_t)g_buf_generation.load());
  return v;
}

void graph_drop(rpd_ctx* ctx) {
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_sig.clear();
}

int64_t grow(int64_t n, double f, int64_t pad) {
  return (int64_t)std::ceil((double)n * f) + pad;
}

}  // namespace

extern "C" {

int rpd_graph_build(rpd_ctx* ctx, const rpd_run_cfg* cfg, double headroom) {
  if (!ctx) return -1;
  if (!ctx->has_mesh || !ctx->has_sites)
    return fail(ctx, "rpd_graph_build: no resident instance");
  if (headroom < 0 || headroom > 10)
    return fail(ctx, "rpd_graph_build: headroom must be in [0, 10]");
  HIP_TRY(hipSetDevice(ctx->device));
  RunState R;
  if (int rc = run_config(ctx, cfg, R)) return rc;
  if (R.cfg.count_work || !R.cfg.small_tier)
    return fail(ctx, "rpd_graph_build: the small tier without work counters "
                     "only (count_work = 0, small_tier = 1)");
  graph_drop(ctx);
  // 1. an eager run: the results, and the counts the capacities follow
  if (int rc = rpd_run(ctx, &R.cfg)) return rc;
  const rpd_stats st_eager = ctx->st;
  const RunCaps& n = ctx->seen;
  const double f = 1.0 + headroom;
  RunCaps cap;
  cap.n_reg = grow(n.n_reg, f, 1024);
  cap.n_cand = grow(n.n_cand, f, 1024);
  cap.n_mid = grow(n.n_mid, f, 256);
  cap.n_full = grow(n.n_full, f, 256);
  cap.n_exact = grow(n.n_exact, f, 256);
  cap.cells = grow(n.cells, f, 1024);
  cap.words = grow(n.words, f, 65536);
  cap.areas = grow(n.areas, f, 16384);
  if (R.cfg.learned) {
    // one closure round more than the eager run needed (up to the eager
    // limit), the merged pairs with room for it: a replay that needs still
    // more is captured again, so that its per-tet site sets are the eager
    // ones too (a tet sent to the full relation keeps related sites whose
    // cells are empty, one closed by the rounds does not)
    cap.rounds = std::min<int64_t>(n.rounds + 1, kClosureRounds);
    // per round (the extra round with the last merged count plus its list)
    cap.miss = cap.added = cap.ext = 0;
    for (int64_t k = 0; k <= cap.rounds; ++k) {
      cap.miss_r[k] = grow(k <= n.rounds ? n.miss_r[k] : 0, f, 4096);
      cap.miss = std::max(cap.miss, cap.miss_r[k]);
    }
    for (int64_t r = 0; r < cap.rounds; ++r) {
      const bool seen_r = r < n.rounds;
      cap.added_r[r] = grow(seen_r ? n.added_r[r] : 0, f, 1024);
      cap.ext_r[r] = seen_r ? grow(n.ext_r[r], f, 4096)
                            : (r > 0 ? cap.ext_r[r - 1] : 4096) +
                                  cap.miss_r[r];
      cap.added += cap.added_r[r];
      cap.ext = std::max(cap.ext, cap.ext_r[r]);
    }
    cap.ext = std::max<int64_t>(cap.ext, 1);
    cap.uniq = cap.miss;
    cap.flagged = grow(n.flagged, f, 256);
    cap.fb = grow(n.fb, f, 4096);
    cap.incl = grow(n.incl, f, 1024);
    cap.c_mid = grow(n.c_mid, f, 256);
    cap.c_full = grow(n.c_full, f, 256);
    cap.c_exact = grow(n.c_exact, f, 256);
  }
  if (cap.n_reg >= (1LL << 31) ||
      cap.n_cand + cap.added + cap.fb >= (1LL << 31) ||
      cap.miss >= (1LL << 31) || cap.incl >= (1LL << 31))
    return fail(ctx, "rpd_graph_build: capacities over 2^31 entries");
  // the neighbour lists and the tables sized by them may lengthen in later
  // sphere updates: room for them too, so that those keep their addresses
  {
    const size_t nk = (size_t)std::max<int64_t>(grow(ctx->n_knn, f, 1024), 1);
    HIP_TRY(ctx->knn_bis.need(nk * 32));
    HIP_TRY(ctx->knn_bisf.need(nk * 16));
    HIP_TRY(ctx->knn_ids.need_keep(nk * 4, (size_t)ctx->n_knn * 4,
                                   ctx->stream));
    HIP_TRY(ctx->knn_near.need_keep(nk * 4, (size_t)ctx->n_knn * 4,
                                    ctx->stream));
    HIP_TRY(ctx->knn_site_near.need_keep(nk * 32, (size_t)ctx->n_knn * 32,
                                         ctx->stream));
  }
  if (R.cfg.learned) {
    // eager learned runs between replays size their first miss lists by
    // their candidate count: room for up to the captured capacity, so that
    // they reallocate nothing the graph uses
    const size_t m = (size_t)std::max<int64_t>(4096, 2 * cap.n_cand) * 8;
    HIP_TRY(ctx->miss_a.need(m));
    HIP_TRY(ctx->miss_b.need(m));
  }
  // 2. the captured stages once without capture: every buffer reaches its
  // capacity here, so the capture allocates nothing
  RunState P = R;
  P.graph = true;
  P.cap = cap;
  int64_t nc = 0, no = 0;
  Sizes3 tot{0, 0, 0};
  if (int rc = run_stages(ctx, P, &nc, &no, &tot)) return rc;
  HIP_TRY(wait_stream(ctx, ctx->stream));
  // 3. capture
  RunState Q = R;
  Q.graph = true;
  Q.cap = cap;
  t_capture = true;
  hipError_t e = hipStreamBeginCapture(ctx->stream,
                                       hipStreamCaptureModeThreadLocal);
  int rc = e == hipSuccess ? run_stages(ctx, Q, &nc, &no, &tot) : -2;
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
  t_capture = false;
  if (e != hipSuccess)
    return fail(ctx, std::string("rpd_graph_build: hipStreamBeginCapture: ") +
                         hipGetErrorString(e), -2);
  if (rc != 0) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e2 != hipSuccess || !g)
    return fail(ctx, std::string("rpd_graph_build: hipStreamEndCapture: ") +
                         hipGetErrorString(e2), -2);
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(ctx, std::string("rpd_graph_build: hipGraphInstantiate: ") +
                         hipGetErrorString(e), -2);
  }
  ctx->graph = g;
  ctx->graph_exec = x;
  ctx->graph_tet_off = ctx->tet_off.p;  // where its final offsets land
  if (R.cfg.learned) {
    // the capture swapped the offset buffers without running: the current
    // results are the pre-pass's, in the other ones (rpd_graph_run swaps
    // back to graph_tet_off before a replay)
    ctx->tet_off.swap(ctx->out_off);
    ctx->rel_ids.swap(ctx->out_ids);
    ctx->tet_count.swap(ctx->out_cnt);
  }
  ctx->graph_cap = cap;
  ctx->graph_cfg = R.cfg;
  ctx->graph_headroom = headroom;
  ctx->graph_sig = graph_signature(ctx);
  ctx->st = st_eager;  // the eager run's results are the current ones
  return 0;
}

int rpd_graph_run(rpd_ctx* ctx, int32_t* replayed) {
  if (!ctx) return -1;
  if (replayed) *replayed = 0;
  if (!ctx->graph_exec)
    return fail(ctx, "rpd_graph_run: no graph (rpd_graph_build first)");
  if (!ctx->has_mesh || !ctx->has_sites)
    return fail(ctx, "rpd_graph_run: no resident instance");
  HIP_TRY(hipSetDevice(ctx->device));
  // the batch or a buffer changed since the capture, or the neighbour
  // lists outgrew the captured bisector table: capture again (its eager run
  // gives this call's results)
  if (graph_signature(ctx) != ctx->graph_sig ||
      (size_t)std::max<int64_t>(ctx->n_knn, 1) * 32 > ctx->knn_bis.cap ||
      (size_t)std::max<int64_t>(ctx->n_knn, 1) * 16 > ctx->knn_bisf.cap) {
    const rpd_run_cfg c = ctx->graph_cfg;
    return rpd_graph_build(ctx, &c, ctx->graph_headroom);
  }
  const bool learned = ctx->graph_cfg.learned != 0;
  if (learned && ctx->tet_off.p != ctx->graph_tet_off) {
    // an eager learned run since the capture swapped the offset buffers
    // the graph writes: swap them back
    ctx->tet_off.swap(ctx->out_off);
    ctx->rel_ids.swap(ctx->out_ids);
    ctx->tet_count.swap(ctx->out_cnt);
    if (ctx->tet_off.p != ctx->graph_tet_off) {
      const rpd_run_cfg c = ctx->graph_cfg;
      return rpd_graph_build(ctx, &c, ctx->graph_headroom);
    }
  }
  ctx->last_geometry = false;
  ctx->has_results = false;
  ctx->adj_ready = false;
  hipStream_t s = ctx->stream;
  HIP_TRY(hipEventRecord(ctx->ev[0], s));
  HIP_TRY(hipGraphLaunch(ctx->graph_exec, s));
  HIP_TRY(hipEventRecord(ctx->ev[4], s));
  const RunCaps& cap = ctx->graph_cap;
  const int64_t nb = ceil_div(learned ? cap.incl : cap.n_cand, kCB);
  int64_t st64[4], drops[kDropSlots * 9], n_cand = 0, n_reg = 0, n1 = 0;
  int32_t cnt[8], reop[64], nf = 0;
  int64_t gc[3 + 2 * kClosureRounds];
  Sizes3 tot{0, 0, 0};
  std::vector<D2H> items = {
      {st64, ctx->stats64.p, 32},
      {drops, ctx->drops.p, sizeof(drops)},
      {cnt, ctx->counters.p, 32},
      {&tot, ctx->offs.as<Sizes3>() + nb, sizeof(Sizes3)},
      {&n_cand, ctx->tet_off.as<int64_t>() + ctx->n_tet, 8}};
  if (learned) {  // first-round count, re-run tets, ranges, listed tets
    items.push_back({&n1, ctx->out_off.as<int64_t>() + ctx->n_tet, 8});
    items.push_back({reop, ctx->reop64.p, sizeof(reop)});
    items.push_back({gc, ctx->gcnt.p, (size_t)(3 + 2 * cap.rounds) * 8});
    items.push_back({&nf, ctx->flag_pos.as<int32_t>() + ctx->n_tet, 4});
  } else {
    items.push_back({&n_reg, ctx->reg_off.as<int64_t>() + ctx->n_site, 8});
  }
  HIP_TRY(read_back(ctx, s, items));
  const bool mid = small_caps(ctx->graph_cfg.small_caps).V < SmallCaps::V;
  // (the re-run tiers of a learned run are checked on the device, launch by
  // launch, into cnt[7])
  const bool tiers_over =
      !learned && ((mid ? (cnt[1] > cap.n_mid || cnt[6] > cap.n_full)
                        : cnt[1] > cap.n_full) ||
                   cnt[2] > cap.n_exact);
  const bool over =
      cnt[7] != 0 || tiers_over || n_reg > cap.n_reg ||
      (learned ? (n1 > cap.n_cand || n_cand > cap.incl)
               : n_cand > cap.n_cand) ||
      tot.cells > cap.cells || tot.words > cap.words || tot.areas > cap.areas;
  if (over) {  // counts beyond the capture's capacities: capture again
    std::string why;
    if (cnt[7] != 0) why += " device-side capacity";
    if (tiers_over) why += " re-run tiers";
    if (learned ? n1 > cap.n_cand : n_cand > cap.n_cand) why += " candidates";
    if (learned && n_cand > cap.incl) why += " output candidates";
    if (n_reg > cap.n_reg) why += " bin registrations";
    if (tot.cells > cap.cells || tot.words > cap.words ||
        tot.areas > cap.areas)
      why += " output stream";
    const rpd_run_cfg c = ctx->graph_cfg;
    const int rc = rpd_graph_build(ctx, &c, ctx->graph_headroom);
    if (rc == 0)  // (not an error: the reason of the new capture)
      ctx->err = "rpd_graph_run: captured again, over capacity:" + why;
    return rc;
  }
  rpd_stats& st = ctx->st;
  const int32_t site_k = st.site_k;
  std::memset(&st, 0, sizeof(st));
  st.site_k = site_k;
  st.n_cells = tot.cells;
  st.n_candidates = n_cand;
  st.n_words = tot.words;
  st.n_areas = tot.areas;
  for (int q = 0; q < 9; ++q)
    for (int b = 0; b < kDropSlots; ++b) st.n_dropped[q] += drops[9 * b + q];
  st.n_rerun_large = cnt[1];
  st.n_rerun_exact = cnt[2];
  st.n_bbox_fallback = cnt[0];
  st.n_bin_registrations = n_reg;
  st.tet_k = (int32_t)st64[2];
  st.n_host_reads = 1;  // the counts, once, after the graph
  if (learned) {
    st.n_rerun_large = st.n_rerun_exact = 0;  // (list positions only)
    int64_t re = 0, added = 0;
    for (int q = 0; q < 64; ++q) re += reop[q];
    for (int64_t r = 0; r <= cap.rounds; ++r) added += gc[1 + 2 * r];
    st.learned[0] = n1;
    st.learned[1] = re;
    st.learned[2] = added;
    st.learned[3] = nf;
  }
  float ms = 0;
  st.ms_total =
      hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[4]) == hipSuccess ? ms : -1;
  ctx->last_geometry = ctx->graph_cfg.compute_geometry != 0;
  ctx->has_results = true;
  if (replayed) *replayed = 1;
  return 0;
}

int rpd_graph_destroy(rpd_ctx* ctx) {
  if (!ctx) return -1;
  graph_drop(ctx);
  return 0;
}

int rpd_set_tet_subset(rpd_ctx* ctx, int32_t n, const int32_t* tet_ids) {
  if (!ctx) return -1;
  if (!ctx->has_mesh) return fail(ctx, "rpd_set_tet_subset: no resident mesh");
  if (n < 0 || (n > 0 && !tet_ids))
    return fail(ctx, "rpd_set_tet_subset: bad arguments");
  if (n == 0) {
    ctx->has_subset = false;
    return 0;
  }
  std::vector<uint8_t> sel(ctx->n_tet, 0);
  for (int32_t i = 0; i < n; ++i) {
    if (tet_ids[i] < 0 || tet_ids[i] >= ctx->n_tet)
      return fail(ctx, "rpd_set_tet_subset: tet id " +
                           std::to_string(tet_ids[i]) + " out of range");
    sel[tet_ids[i]] = 1;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(ctx->tet_sel.need((size_t)ctx->n_tet));
  HIP_TRY(hipMemcpyAsync(ctx->tet_sel.p, sel.data(), sel.size(),
                         hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(wait_stream(ctx, ctx->stream));
  ctx->has_subset = true;
  return 0;
}

int rpd_site_centroids(rpd_ctx* ctx, double* out, const double** dev_out) {
  if (!ctx) return -1;
  if (!ctx->has_sites) return fail(ctx, "rpd_site_centroids: no resident sites");
  if (!ctx->last_geometry)
    return fail(ctx, "rpd_site_centroids: the last rpd_run computed no cell "
                     "geometry (compute_geometry = 0, or no run)");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int32_t ns = ctx->n_site;
  if (int rc = cells_by_seed(ctx)) return rc;
  HIP_TRY(ctx->cs_out.need((size_t)std::max(ns, 1) * 32));
  if (ns > 0)
    hipLaunchKernelGGL(k_site_centroid, dim3(ceil_div(ns, 4)), dim3(256), 0, s,
                       ns, ctx->cs_start.as<int64_t>(),
                       ctx->cs_vals_b.as<int32_t>(), ctx->out_geom.as<double>(),
                       ctx->site.as<double4>(), ctx->cs_out.as<double>());
  HIP_TRY(hipGetLastError());
  if (out && ns > 0)
    HIP_TRY(hipMemcpyAsync(out, ctx->cs_out.p, (size_t)ns * 32,
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(wait_stream(ctx, s));
  if (dev_out) *dev_out = ctx->cs_out.as<double>();
  return 0;
}

// Plain streaming copy of `bytes` (16-byte vector loads and stores), the
// reference kernel of known traffic for calibrating HBM counters.
__global__ __launch_bounds__(256) void k_calib_copy(const uint4* __restrict__ a,
                                                    uint4* __restrict__ b,
                                                    int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

int rpd_calib_copy(rpd_ctx* ctx, int64_t bytes) {
  if (!ctx || bytes <= 0 || (bytes & 15)) return -1;
  HIP_TRY(hipSetDevice(ctx->device));
  void *a = nullptr, *b = nullptr;
  HIP_TRY(hipMalloc(&a, bytes));
  hipError_t e = hipMalloc(&b, bytes);
  if (e != hipSuccess) {
    (void)hipFree(a);
    return fail(ctx, "rpd_calib_copy: allocation failed", -2);
  }
  (void)hipMemsetAsync(a, 1, bytes, ctx->stream);
  hipLaunchKernelGGL(k_calib_copy, dim3(4096), dim3(256), 0, ctx->stream,
                     (const uint4*)a, (uint4*)b, bytes / 16);
  e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(ctx, "rpd_calib_copy failed", -2);
  return 0;
}

int rpd_power_neighbors_gpu(rpd_ctx* ctx, int32_t n, const double* xyzw,
                            const double* box, double* ms_out,
                            rpd_nbr** out) {
  if (!ctx || n < 0 || !out || (n > 0 && !xyzw) || !box) return -1;
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // weight-class grids on the host (O(n)), flattened: per class the cell
  // starts (absolute into ids) and the members
  std::vector<RpdClassGrid> grids;
  rpd_build_class_grids(n, xyzw, box, grids);
  NbGrids G{};
  std::vector<int64_t> starts;
  std::vector<int32_t> ids;
  for (int c = 0; c < kRpdClasses; ++c) {
    const RpdClassGrid& g = grids[c];
    NbGrid& d = G.g[c];
    d.used = !g.ids.empty();
    if (!d.used) continue;
    for (int a = 0; a < 3; ++a) {
      d.org[a] = g.org[a];
      d.dim[a] = g.dim[a];
    }
    d.cs = g.cs;
    d.wmax = g.wmax;
    d.cell_base = (int64_t)starts.size();
    const int64_t id_base = (int64_t)ids.size();
    for (int64_t v : g.start) starts.push_back(id_base + v);
    ids.insert(ids.end(), g.ids.begin(), g.ids.end());
  }
  Buf d_sph, d_box, d_start, d_ids, d_valid, d_cnt, d_lists;
  const int64_t nn = std::max<int32_t>(n, 1);
  HIP_TRY(d_sph.need(nn * 32));
  HIP_TRY(d_box.need(6 * 8));
  HIP_TRY(d_start.need(std::max<size_t>(starts.size(), 1) * 8));
  HIP_TRY(d_ids.need(std::max<size_t>(ids.size(), 1) * 4));
  HIP_TRY(d_valid.need(nn * 4));
  HIP_TRY(d_cnt.need(nn * 4));
  HIP_TRY(d_lists.need(nn * kNbOut * 4));
  HIP_TRY(hipMemcpyAsync(d_sph.p, xyzw, (size_t)n * 32, hipMemcpyHostToDevice,
                         s));
  HIP_TRY(hipMemcpyAsync(d_box.p, box, 6 * 8, hipMemcpyHostToDevice, s));
  if (!starts.empty())
    HIP_TRY(hipMemcpyAsync(d_start.p, starts.data(), starts.size() * 8,
                           hipMemcpyHostToDevice, s));
  if (!ids.empty())
    HIP_TRY(hipMemcpyAsync(d_ids.p, ids.data(), ids.size() * 4,
                           hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(ctx->ev[14], s));
  if (n > 0)
    hipLaunchKernelGGL(k_power_nbrs, dim3(n), dim3(64), 0, s, n,
                       d_sph.as<double4>(), d_box.as<double>(), G,
                       d_start.as<int64_t>(), d_ids.as<int32_t>(),
                       d_valid.as<int32_t>(), d_cnt.as<int32_t>(),
                       d_lists.as<int32_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ctx->ev[15], s));
  std::vector<int32_t> valid(n), cnt(n), lists((size_t)n * kNbOut);
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(valid.data(), d_valid.p, (size_t)n * 4,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(cnt.data(), d_cnt.p, (size_t)n * 4,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(lists.data(), d_lists.p, (size_t)n * kNbOut * 4,
                           hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(wait_stream(ctx, s));
  if (ms_out) {
    float m = 0;
    HIP_TRY(hipEventElapsedTime(&m, ctx->ev[14], ctx->ev[15]));
    *ms_out = m;
  }
  for (int32_t i = 0; i < n; ++i)
    if (valid[i] < 0)
      return fail(ctx, "rpd_power_neighbors_gpu: sphere " + std::to_string(i) +
                           " exceeds the cell capacities");
  if (int rc = rpd_nbr_from_lists(n, valid.data(), cnt.data(), lists.data(),
                                  kNbOut, out))
    return fail(ctx, "rpd_power_neighbors_gpu: bad neighbour lists", rc);
  return 0;
}

int rpd_tune(rpd_ctx* ctx, const rpd_run_cfg* base, int32_t* best_G,
             int32_t* best_caps, double* ms_out /* 12, or NULL */) {
  if (!ctx) return -1;
  rpd_run_cfg cfg{};
  cfg.use_active_mask = 1;
  cfg.compute_geometry = 1;
  cfg.small_tier = 1;
  if (base) cfg = *base;
  static const int kG[3] = {4, 8, 16};
  double best = 1e300;
  int bg = 8, bc = 0;
  for (int c = 0; c < 4; ++c)
    for (int q = 0; q < 3; ++q) {
      cfg.cells_per_wave = kG[q];
      cfg.small_caps = c;
      double t = 1e300;
      for (int rep = 0; rep < 3; ++rep) {  // first run warms the buffers
        if (int rc = rpd_run(ctx, &cfg)) return rc;
        if (rep) t = std::min(t, ctx->st.ms_phase[1] + ctx->st.ms_phase[2]);
      }
      if (ms_out) ms_out[3 * c + q] = t;
      if (t < best) {
        best = t;
        bg = kG[q];
        bc = c;
      }
    }
  ctx->tuned_G = bg;
  ctx->tuned_caps = bc;
  if (best_G) *best_G = bg;
  if (best_caps) *best_caps = bc;
  return 0;
}

int rpd_sync(rpd_ctx* ctx) {
  if (!ctx) return -1;
  HIP_TRY(wait_stream(ctx, ctx->stream));
  return 0;
}

int rpd_get_stats(const rpd_ctx* ctx, rpd_stats* out) {
  if (!ctx || !out) return -1;
  *out = ctx->st;
  return 0;
}

int rpd_download_cells(rpd_ctx* ctx, rpd_cell* cells, uint32_t* words,
                       double* geom, double* areas) {
  if (!ctx) return -1;
  HIP_TRY(hipSetDevice(ctx->device));
  const rpd_stats& st = ctx->st;
  hipStream_t s = ctx->stream;
  if (cells && st.n_cells)
    HIP_TRY(hipMemcpyAsync(cells, ctx->out_cells.p, st.n_cells * sizeof(rpd_cell),
                           hipMemcpyDeviceToHost, s));
  if (words && st.n_words)
    HIP_TRY(hipMemcpyAsync(words, ctx->out_words.p, st.n_words * 4,
                           hipMemcpyDeviceToHost, s));
  if (geom && st.n_cells)
    HIP_TRY(hipMemcpyAsync(geom, ctx->out_geom.p, st.n_cells * 32,
                           hipMemcpyDeviceToHost, s));
  if (areas && st.n_areas)
    HIP_TRY(hipMemcpyAsync(areas, ctx->out_areas.p, st.n_areas * 8,
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(wait_stream(ctx, s));
  return 0;
}

int rpd_download_relation(rpd_ctx* ctx, int64_t* offsets, int32_t* ids) {
  if (!ctx) return -1;
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  if (offsets)
    HIP_TRY(hipMemcpyAsync(offsets, ctx->tet_off.p, ((size_t)ctx->n_tet + 1) * 8,
                           hipMemcpyDeviceToHost, s));
  if (ids && ctx->st.n_candidates)
    HIP_TRY(hipMemcpyAsync(ids, ctx->rel_ids.p, ctx->st.n_candidates * 4,
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(wait_stream(ctx, s));
  return 0;
}

int rpd_run_noalloc(rpd_ctx* ctx, const rpd_instance_desc* desc,
                    const rpd_run_cfg* cfg, const rpd_noalloc_out* out,
                    rpd_stats* st_out) {
  if (!ctx || !desc || !out) return -1;
  if (cfg && cfg->learned != 0)
    return fail(ctx, "rpd_run_noalloc: plain path only (learned = 0)");
  if (cfg && cfg->count_work)
    return fail(ctx, "rpd_run_noalloc: count_work is not available");
  if (!out->cells || !out->words || !out->geom || !out->areas)
    return fail(ctx, "rpd_run_noalloc: missing output buffers");
  if (!ctx->scratch_dev)
    return fail(ctx, "rpd_run_noalloc: the scratch address of this device "
                     "is not available", -3);
  if (t_arena) return fail(ctx, "rpd_run_noalloc: nested call");
  HIP_TRY(hipSetDevice(ctx->device));
  // the context's resident instance and last results are put back when the
  // call returns: it runs entirely in the scratch
  struct Saved {
    std::vector<rpd_ctx::Inst> inst;
    int32_t n_vert, n_tet, n_site;
    int64_t n_knn;
    bool has_mesh, has_sites, has_subset, last_geometry, has_results,
        adj_ready;
    rpd_stats st;
  } sv{ctx->inst,          ctx->n_vert,      ctx->n_tet,
       ctx->n_site,        ctx->n_knn,       ctx->has_mesh,
       ctx->has_sites,     ctx->has_subset,  ctx->last_geometry,
       ctx->has_results,   ctx->adj_ready,   ctx->st};
  Arena A;
  A.base = static_cast<char*>(ctx->scratch_dev);
  A.cap = kScratchBytes;
  struct Guard {
    rpd_ctx* c;
    Arena& a;
    Saved& s;
    ~Guard() {
      // the call's work must be complete before its blocks can be handed
      // to anyone else
      (void)hipStreamSynchronize(c->stream);
      t_arena = nullptr;
      a.restore();
      c->inst = s.inst;
      c->n_vert = s.n_vert;
      c->n_tet = s.n_tet;
      c->n_site = s.n_site;
      c->n_knn = s.n_knn;
      c->has_mesh = s.has_mesh;
      c->has_sites = s.has_sites;
      c->has_subset = s.has_subset;
      c->last_geometry = s.last_geometry;
      c->has_results = s.has_results;
      c->adj_ready = s.adj_ready;
      c->st = s.st;
    }
  } guard{ctx, A, sv};
  t_arena = &A;
  auto scratch_err = [&](int rc) {
    if (A.exhausted)
      return fail(ctx,
                  "rpd_run_noalloc: the instance needs more than the " +
                      std::to_string(kScratchBytes >> 20) +
                      " MiB device scratch (used " +
                      std::to_string(A.used >> 20) + " MiB before the "
                      "request that failed)",
                  -4);
    return rc;
  };
  if (desc->n_tet <= 0 || desc->n_vert <= 0 || desc->n_site <= 0 ||
      !desc->sites)
    return fail(ctx, "rpd_run_noalloc: instance without tets or sites");
  if (int rc = rpd_upload(ctx, desc)) return scratch_err(rc);
  if (int rc = rpd_run(ctx, cfg)) return scratch_err(rc);
  const rpd_stats& st = ctx->st;
  if (st.n_cells > out->cells_cap || st.n_words > out->words_cap ||
      st.n_areas > out->areas_cap)
    return fail(ctx, "rpd_run_noalloc: output buffers too small (cells " +
                         std::to_string(st.n_cells) + ", words " +
                         std::to_string(st.n_words) + ", areas " +
                         std::to_string(st.n_areas) + ")",
                -5);
  hipStream_t s = ctx->stream;
  if (st.n_cells) {
    HIP_TRY(hipMemcpyAsync(out->cells, ctx->out_cells.p,
                           (size_t)st.n_cells * sizeof(rpd_cell),
                           hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(out->geom, ctx->out_geom.p,
                           (size_t)st.n_cells * 32, hipMemcpyDeviceToDevice,
                           s));
  }
  if (st.n_words)
    HIP_TRY(hipMemcpyAsync(out->words, ctx->out_words.p,
                           (size_t)st.n_words * 4, hipMemcpyDeviceToDevice,
                           s));
  if (st.n_areas)
    HIP_TRY(hipMemcpyAsync(out->areas, ctx->out_areas.p,
                           (size_t)st.n_areas * 8, hipMemcpyDeviceToDevice,
                           s));
  HIP_TRY(wait_stream(ctx, s));
  if (st_out) {
    *st_out = st;
    st_out->scratch_bytes = (int64_t)A.high;
  }
  return 0;
}

}  // extern "C"

namespace {

int mlp_alloc(rpd_ctx* ctx) {
  HIP_TRY(ctx->mlp_master.need(kNParam * 4));
  HIP_TRY(ctx->mlp_m1.need(kNParam * 4));
  HIP_TRY(ctx->mlp_m2.need(kNParam * 4));
  HIP_TRY(ctx->mlp_w1t.need(kIn * kHid * 2));
  HIP_TRY(ctx->mlp_w2t.need(kHid * kHid * 2));
  HIP_TRY(ctx->mlp_w3t.need(kHid * kOut * 2));
  HIP_TRY(ctx->mlp_w2.need(kHid * kHid * 2));
  HIP_TRY(ctx->mlp_w3.need(kHid * kOut * 2));
  HIP_TRY(ctx->mlp_bias.need((2 * kHid + kOut) * 4));
  return 0;
}

int mlp_pack(rpd_ctx* ctx) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(ceil_div(kHid * kHid, 256)), dim3(256),
                     0, ctx->stream, ctx->mlp_master.as<float>(),
                     ctx->mlp_w1t.as<uint16_t>(), ctx->mlp_w2t.as<uint16_t>(),
                     ctx->mlp_w3t.as<uint16_t>(), ctx->mlp_w2.as<uint16_t>(),
                     ctx->mlp_w3.as<uint16_t>(), ctx->mlp_bias.as<float>());
  HIP_TRY(hipGetLastError());
  return 0;
}

// one training step (or a gradient) over rows idx[0..n_rows) of X/Y/M;
// n_rows is a multiple of 16.  The per-tile losses go to loss_tiles (device,
// n_rows / 16 floats); nothing waits for the device.
int mlp_step(rpd_ctx* ctx, const uint16_t* X, const float* Y, const float* M,
             const int32_t* idx, int32_t n_rows, int32_t n_valid,
             float pos_w, const AdamArgs& adam, float* grad_out,
             float* loss_tiles) {
  hipStream_t s = ctx->stream;
  const int tiles = n_rows / 16;
  const int tpc = 8;
  const int chunks = (tiles + tpc - 1) / tpc;
  HIP_TRY(ctx->tr_xb.need((size_t)n_rows * kIn * 2));
  HIP_TRY(ctx->tr_h1.need((size_t)n_rows * kHid * 2));
  HIP_TRY(ctx->tr_h2.need((size_t)n_rows * kHid * 2));
  HIP_TRY(ctx->tr_dz1.need((size_t)n_rows * kHid * 2));
  HIP_TRY(ctx->tr_dz2.need((size_t)n_rows * kHid * 2));
  HIP_TRY(ctx->tr_dz3.need((size_t)n_rows * kOut * 2));
  HIP_TRY(ctx->tr_p1.need((size_t)chunks * kIn * kHid * 4));
  HIP_TRY(ctx->tr_p2.need((size_t)chunks * kHid * kHid * 4));
  HIP_TRY(ctx->tr_p3.need((size_t)chunks * kHid * kOut * 4));
  HIP_TRY(ctx->tr_pb.need((size_t)chunks * kBiasN * 4));
  const float inv_count = 1.f / (float)std::max(1, n_valid * kOut);
  hipLaunchKernelGGL(k_mlp_fb, dim3(ceil_div(tiles, 4)), dim3(256), 0, s,
                     mlp_dev(ctx), X, Y, M, idx, n_rows, pos_w, inv_count,
                     ctx->tr_xb.as<uint16_t>(), ctx->tr_h1.as<uint16_t>(),
                     ctx->tr_h2.as<uint16_t>(), ctx->tr_dz1.as<uint16_t>(),
                     ctx->tr_dz2.as<uint16_t>(), ctx->tr_dz3.as<uint16_t>(),
                     loss_tiles);
  HIP_TRY(hipGetLastError());
  float* pb = ctx->tr_pb.as<float>();
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((kIn / 16) * (kHid / 16), chunks),
                     dim3(64), 0, s, ctx->tr_xb.as<uint16_t>(),
                     ctx->tr_dz1.as<uint16_t>(), kIn, kHid, tpc, tiles,
                     ctx->tr_p1.as<float>(), pb, 0);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((kHid / 16) * (kHid / 16), chunks),
                     dim3(64), 0, s, ctx->tr_h1.as<uint16_t>(),
                     ctx->tr_dz2.as<uint16_t>(), kHid, kHid, tpc, tiles,
                     ctx->tr_p2.as<float>(), pb, kHid);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((kHid / 16) * (kOut / 16), chunks),
                     dim3(64), 0, s, ctx->tr_h2.as<uint16_t>(),
                     ctx->tr_dz3.as<uint16_t>(), kHid, kOut, tpc, tiles,
                     ctx->tr_p3.as<float>(), pb, 2 * kHid);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_update, dim3(ceil_div(kNParam, 256)), dim3(256), 0,
                     s, ctx->mlp_master.as<float>(), ctx->mlp_m1.as<float>(),
                     ctx->mlp_m2.as<float>(), ctx->tr_p1.as<float>(),
                     ctx->tr_p2.as<float>(), ctx->tr_p3.as<float>(), pb,
                     chunks, adam, grad_out);
  HIP_TRY(hipGetLastError());
  if (adam.apply)
    if (int rc = mlp_pack(ctx)) return rc;
  return 0;
}

// sum of n device floats (host side, after a wait)
int sum_device(rpd_ctx* ctx, const float* d, int64_t n, double* out) {
  std::vector<float> h(n);
  HIP_TRY(hipMemcpyAsync(h.data(), d, (size_t)n * 4, hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(wait_stream(ctx, ctx->stream));
  double sum = 0;
  for (float v : h) sum += v;
  *out = sum;
  return 0;
}

}  // namespace

extern "C" {

int rpd_mlp_set(rpd_ctx* ctx, const float* params) {
  if (!ctx || !params) return -1;
  HIP_TRY(hipSetDevice(ctx->device));
  if (int rc = mlp_alloc(ctx)) return rc;
  for (int p = 0; p < kNParam; ++p)
    if (!std::isfinite(params[p]))
      return fail(ctx, "rpd_mlp_set: non-finite parameter");
  HIP_TRY(hipMemcpyAsync(ctx->mlp_master.p, params, kNParam * 4,
                         hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemsetAsync(ctx->mlp_m1.p, 0, kNParam * 4, ctx->stream));
  HIP_TRY(hipMemsetAsync(ctx->mlp_m2.p, 0, kNParam * 4, ctx->stream));
  if (int rc = mlp_pack(ctx)) return rc;
  HIP_TRY(wait_stream(ctx, ctx->stream));
  ctx->adam_step = 0;
  ctx->mlp_ready = true;
  return 0;
}

int rpd_mlp_get(rpd_ctx* ctx, float* params) {
  if (!ctx || !params) return -1;
  if (!ctx->mlp_ready) return fail(ctx, "rpd_mlp_get: no model");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(params, ctx->mlp_master.p, kNParam * 4,
                         hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(wait_stream(ctx, ctx->stream));
  return 0;
}

int rpd_mlp_init(rpd_ctx* ctx, uint64_t seed) {
  if (!ctx) return -1;
  std::vector<float> p(kNParam, 0.f);
  auto fill = [&](int off, int in, int out, uint64_t stream) {
    const double lim = std::sqrt(6.0 / (in + out));
    for (int q = 0; q < in * out; ++q)
      p[off + q] =
          (float)((2.0 * rpd_rng_uniform(seed, stream, (uint64_t)q) - 1.0) *
                  lim);
  };
  fill(kOffW1, kIn, kHid, 11);
  fill(kOffW2, kHid, kHid, 12);
  fill(kOffW3, kHid, kOut, 13);
  return rpd_mlp_set(ctx, p.data());
}

int rpd_features(rpd_ctx* ctx, int32_t* knn, uint16_t* X) {
  if (!ctx) return -1;
  if (!ctx->has_mesh || !ctx->has_sites)
    return fail(ctx, "rpd_features: no resident instance");
  HIP_TRY(hipSetDevice(ctx->device));
  RunState R;
  if (int rc = init_state(ctx, R)) return rc;
  if (int rc = learned_features(ctx, R)) return rc;
  const size_t nt = (size_t)ctx->n_tet;
  if (knn)
    HIP_TRY(hipMemcpyAsync(knn, ctx->knn.p, nt * kOut * 4,
                           hipMemcpyDeviceToHost, ctx->stream));
  if (X)
    HIP_TRY(hipMemcpyAsync(X, ctx->feat.p, nt * kIn * 2, hipMemcpyDeviceToHost,
                           ctx->stream));
  HIP_TRY(wait_stream(ctx, ctx->stream));
  return 0;
}

int rpd_labels(rpd_ctx* ctx, float* Y, float* M) {
  if (!ctx) return -1;
  if (ctx->feat_n != ctx->n_tet || ctx->n_tet <= 0)
    return fail(ctx, "rpd_labels: compute rpd_features first");
  HIP_TRY(hipSetDevice(ctx->device));
  const int32_t nt = ctx->n_tet;
  HIP_TRY(ctx->lab_y.need((size_t)nt * kOut * 4));
  HIP_TRY(ctx->lab_m.need((size_t)nt * kOut * 4));
  hipLaunchKernelGGL(k_labels, dim3(ceil_div(nt, 256)), dim3(256), 0,
                     ctx->stream, nt, ctx->out_cells.as<rpd_cell>(),
                     ctx->st.n_cells, ctx->knn.as<int32_t>(),
                     ctx->lab_y.as<float>(), ctx->lab_m.as<float>());
  HIP_TRY(hipGetLastError());
  if (Y)
    HIP_TRY(hipMemcpyAsync(Y, ctx->lab_y.p, (size_t)nt * kOut * 4,
                           hipMemcpyDeviceToHost, ctx->stream));
  if (M)
    HIP_TRY(hipMemcpyAsync(M, ctx->lab_m.p, (size_t)nt * kOut * 4,
                           hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(wait_stream(ctx, ctx->stream));
  ctx->lab_n = nt;
  return 0;
}

int rpd_mlp_train(rpd_ctx* ctx, const rpd_train_cfg* tc, float* loss) {
  if (!ctx || !tc) return -1;
  if (!ctx->mlp_ready) return fail(ctx, "rpd_mlp_train: no model");
  if (ctx->lab_n != ctx->n_tet || ctx->feat_n != ctx->n_tet ||
      ctx->n_tet <= 0)
    return fail(ctx, "rpd_mlp_train: features and labels first");
  if (tc->batch <= 0 || tc->batch % 16 || tc->epochs <= 0)
    return fail(ctx, "rpd_mlp_train: batch must be a positive multiple of "
                     "16, epochs positive");
  HIP_TRY(hipSetDevice(ctx->device));
  const int64_t n = ctx->n_tet, B = tc->batch;
  const int64_t n_pad = (n + B - 1) / B * B;
  std::vector<int32_t> order(n_pad, -1);
  HIP_TRY(ctx->tr_idx.need((size_t)n_pad * 4));
  for (int e = 0; e < tc->epochs; ++e) {
    for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
    for (int64_t i = n - 1; i > 0; --i) {  // Fisher-Yates
      const int64_t j = (int64_t)(rpd_rng_uniform(tc->seed, 1000 + e,
                                                  (uint64_t)i) *
                                  (double)(i + 1));
      std::swap(order[i], order[std::min(j, i)]);
    }
    HIP_TRY(hipMemcpyAsync(ctx->tr_idx.p, order.data(), (size_t)n_pad * 4,
                           hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx->tr_loss.need((size_t)(n_pad / 16) * 4));
    int nb = 0;
    for (int64_t b0 = 0; b0 < n_pad; b0 += B) {
      ++ctx->adam_step;
      AdamArgs a;
      a.lr = tc->lr;
      a.beta1 = 0.9f;
      a.beta2 = 0.999f;
      a.eps = 1e-8f;
      a.bc1 = 1.f - std::pow(0.9f, (float)ctx->adam_step);
      a.bc2 = 1.f - std::pow(0.999f, (float)ctx->adam_step);
      a.apply = 1;
      const int32_t valid = (int32_t)std::min<int64_t>(B, n - b0);
      if (int rc = mlp_step(ctx, ctx->feat.as<uint16_t>(),
                            ctx->lab_y.as<float>(), ctx->lab_m.as<float>(),
                            ctx->tr_idx.as<int32_t>() + b0, (int32_t)B, valid,
                            tc->pos_weight, a, nullptr,
                            ctx->tr_loss.as<float>() + b0 / 16))
        return rc;
      ++nb;
    }
    // mean over the epoch's steps of each step's loss
    double el = 0;
    if (int rc = sum_device(ctx, ctx->tr_loss.as<float>(), n_pad / 16, &el))
      return rc;
    if (loss) loss[e] = (float)(el / std::max(1, nb));
  }
  return 0;
}

int rpd_mlp_grad(rpd_ctx* ctx, const float* params, int32_t n,
                 const uint16_t* X, const float* Y, const float* M,
                 float pos_weight, float* grad, float* loss) {
  if (!ctx || !params || n <= 0 || !X || !Y || !M || !grad || !loss)
    return -1;
  if (int rc = rpd_mlp_set(ctx, params)) return rc;
  const int32_t n_pad = (n + 15) / 16 * 16;
  std::vector<int32_t> idx(n_pad, -1);
  for (int32_t i = 0; i < n; ++i) idx[i] = i;
  HIP_TRY(ctx->tr_x.need((size_t)n * kIn * 2));
  HIP_TRY(ctx->tr_y.need((size_t)n * kOut * 4));
  HIP_TRY(ctx->tr_mk.need((size_t)n * kOut * 4));
  HIP_TRY(ctx->tr_idx.need((size_t)n_pad * 4));
  HIP_TRY(ctx->tr_grad.need(kNParam * 4));
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(ctx->tr_x.p, X, (size_t)n * kIn * 2,
                         hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(ctx->tr_y.p, Y, (size_t)n * kOut * 4,
                         hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(ctx->tr_mk.p, M, (size_t)n * kOut * 4,
                         hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(ctx->tr_idx.p, idx.data(), (size_t)n_pad * 4,
                         hipMemcpyHostToDevice, s));
  AdamArgs a{};
  a.apply = 0;
  HIP_TRY(ctx->tr_loss.need((size_t)(n_pad / 16) * 4));
  if (int rc = mlp_step(ctx, ctx->tr_x.as<uint16_t>(), ctx->tr_y.as<float>(),
                        ctx->tr_mk.as<float>(), ctx->tr_idx.as<int32_t>(),
                        n_pad, n, pos_weight, a, ctx->tr_grad.as<float>(),
                        ctx->tr_loss.as<float>()))
    return rc;
  double lsum = 0;
  if (int rc = sum_device(ctx, ctx->tr_loss.as<float>(), n_pad / 16, &lsum))
    return rc;
  *loss = (float)lsum;
  HIP_TRY(hipMemcpyAsync(grad, ctx->tr_grad.p, kNParam * 4,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(wait_stream(ctx, s));
  return 0;
}

int rpd_mlp_scores(rpd_ctx* ctx, float* logits) {
  if (!ctx || !logits) return -1;
  if (!ctx->mlp_ready || ctx->feat_n != ctx->n_tet || ctx->n_tet <= 0)
    return fail(ctx, "rpd_mlp_scores: model and features first");
  HIP_TRY(hipSetDevice(ctx->device));
  const int32_t nt = ctx->n_tet;
  HIP_TRY(ctx->sel.need((size_t)nt * 4));
  HIP_TRY(ctx->tr_grad.need((size_t)nt * kOut * 4));
  const int64_t tiles = (nt + 15) / 16;
  hipLaunchKernelGGL(k_mlp_infer,
                     dim3((int)std::max<int64_t>(
                         1, std::min<int64_t>(
                                (tiles + kInferWaves - 1) / kInferWaves,
                                1024))),
                     dim3(64 * kInferWaves), 0, ctx->stream, mlp_dev(ctx),
                     ctx->feat.as<uint16_t>(), (int64_t)nt, 0.f,
                     ctx->sel.as<uint32_t>(), ctx->tr_grad.as<float>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(logits, ctx->tr_grad.p, (size_t)nt * kOut * 4,
                         hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(wait_stream(ctx, ctx->stream));
  return 0;
}

}  // extern "C"

