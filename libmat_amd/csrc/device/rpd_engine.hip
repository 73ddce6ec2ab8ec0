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
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rpd_capi.h"
#include "rpd_device.h"
#include "rpd_kernels.hip"

namespace {

using namespace rpd;

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  hipError_t need(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t b = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) cap = b;
    return e;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

struct TierBufs {
  Buf hdr, words, geom, areas;
  SlotView view(int V, int P, int E) const {
    SlotView s;
    s.stride_words = V + P + E;
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
    if ((e = words.need(n * (V + P + E) * 4)) != hipSuccess) return e;
    if ((e = geom.need(n * 4 * 8)) != hipSuccess) return e;
    return areas.need(n * P * 8);
  }
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
  Buf vert, tet_v, tet_fid, tet_meta, tet_inst, inst_dev;
  Buf site, flags, knn_off, knn_ids, knn_near, knn_site_near, site_inst;
  // pipeline
  Buf bbox, reg_n, reg_off, keys_a, keys_b, bin_start, bin_ids;
  Buf tet_count, tet_off, tet_pass, cand, rel_ids;
  Buf counters;  // int32[8]: fallback, n_rerun_large, n_rerun_exact
  Buf rerun_large, rerun_exact, loc, sizes, offs, drops, stats64, work,
      rel_work;
  Buf scan_tmp;
  TierBufs small, large, exact;
  Buf out_cells, out_words, out_geom, out_areas;
  rpd_stats st;
  int tuned_G = 0, tuned_caps = 0;  // rpd_tune's choice (0: not tuned)
  hipEvent_t ev[16];
  bool events = false;
};

namespace {

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
  for (int i = 0; i < 16; ++i)
    if (hipEventCreate(&ctx->ev[i]) != hipSuccess) {
      delete ctx;
      return -3;
    }
  ctx->events = true;
  std::memset(&ctx->st, 0, sizeof(ctx->st));
  *out = ctx;
  return 0;
}

void rpd_ctx_destroy(rpd_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->events)
    for (int i = 0; i < 16; ++i) (void)hipEventDestroy(ctx->ev[i]);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
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
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->has_sites = true;
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
      tinst(nt_tot);
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
  HIP_TRY(ctx->vert.need((size_t)nv_tot * 32));
  HIP_TRY(ctx->tet_v.need((size_t)nt_tot * 16));
  HIP_TRY(ctx->tet_fid.need((size_t)nt_tot * 16));
  HIP_TRY(ctx->tet_meta.need((size_t)nt_tot * 16));
  HIP_TRY(ctx->tet_inst.need((size_t)nt_tot * 4));
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
  HIP_TRY(hipStreamSynchronize(ctx->stream));
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
// capacities of the small tier: 0 = SmallCaps, 1 = WideCaps
struct CapsRT {
  int V, P, E;
};
CapsRT small_caps(int id) {
  return id == 1 ? CapsRT{WideCaps::V, WideCaps::P, WideCaps::E}
                 : CapsRT{SmallCaps::V, SmallCaps::P, SmallCaps::E};
}

static int launch_clip(rpd_ctx* ctx, int tier, int G, ClipArgs& a,
                       int caps_id = 0) {
  if (a.n_items <= 0) return 0;
  const int blocks = ceil_div(a.n_items, G);
#define RPD_LAUNCH(GG, CAPS, EX)                                       \
  hipLaunchKernelGGL((k_clip<GG, CAPS, EX>), dim3(blocks), dim3(64), 0, \
                     ctx->stream, a)
  if (tier == 0) {
    if (caps_id == 1) {
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

}  // namespace

extern "C" {

int rpd_run(rpd_ctx* ctx, const rpd_run_cfg* cfg_in) {
  if (!ctx) return -1;
  if (!ctx->has_mesh || !ctx->has_sites)
    return fail(ctx, "rpd_run: no resident instance");
  HIP_TRY(hipSetDevice(ctx->device));
  rpd_run_cfg cfg{};
  cfg.cells_per_wave = 0;
  cfg.use_active_mask = 1;
  cfg.compute_geometry = 1;
  cfg.force_exact = 0;
  cfg.small_tier = 1;
  if (cfg_in) cfg = *cfg_in;
  if (cfg.cells_per_wave <= 0) {  // tuned shape, else the default
    cfg.cells_per_wave = ctx->tuned_G > 0 ? ctx->tuned_G : 8;
    if (ctx->tuned_G > 0) cfg.small_caps = ctx->tuned_caps;
  }
  if (cfg.small_caps != 0 && cfg.small_caps != 1)
    return fail(ctx, "rpd_run: small_caps must be 0 or 1");
  const CapsRT sc = small_caps(cfg.small_caps);
  const int G = cfg.cells_per_wave;
  if (G != 1 && G != 2 && G != 4 && G != 8 && G != 16)
    return fail(ctx, "rpd_run: cells_per_wave must be 1, 2, 4, 8 or 16");
  hipStream_t s = ctx->stream;
  const int32_t nt = ctx->n_tet, ns = ctx->n_site;
  const int32_t site_k = ctx->st.site_k;
  std::memset(&ctx->st, 0, sizeof(ctx->st));
  ctx->st.site_k = site_k;

  DevMesh M{ctx->n_vert, nt, ctx->vert.as<double4>(), ctx->tet_v.as<int4>(),
            ctx->tet_fid.as<int4>(), ctx->tet_meta.as<TetMeta>()};
  DevSites S{ns, ctx->site.as<double4>(), ctx->flags.as<uint32_t>(),
             ctx->knn_off.as<int64_t>(), ctx->knn_ids.as<int32_t>(),
             ctx->knn_near.as<int32_t>(), ctx->knn_site_near.as<double4>()};

  HIP_TRY(hipEventRecord(ctx->ev[0], s));
  HIP_TRY(ctx->counters.need(8 * 4));
  HIP_TRY(hipMemsetAsync(ctx->counters.p, 0, 8 * 4, s));
  int32_t* cnt = ctx->counters.as<int32_t>();

  if (cfg.count_work) {
    HIP_TRY(ctx->rel_work.need(64 * 4 * 8));
    HIP_TRY(hipMemsetAsync(ctx->rel_work.p, 0, 64 * 4 * 8, s));
  }

  // ---- 1. candidate gathering
  // per instance: a uniform grid of about n_tet/6 bins over its domain,
  // numbered from bin_base in the batch-wide bin space
  int64_t n_bins = 0;
  {
    std::vector<InstInfo> ii(ctx->inst.size());
    for (size_t b = 0; b < ctx->inst.size(); ++b) {
      const rpd_ctx::Inst& I = ctx->inst[b];
      InstInfo& o = ii[b];
      double ext[3], vol = 1;
      for (int a = 0; a < 3; ++a) {
        ext[a] = std::max(I.dom[3 + a] - I.dom[a], 1e-300);
        vol *= ext[a];
      }
      const double bs = std::cbrt(vol / std::max(1.0, I.n_tet / 6.0));
      int64_t nb = 1;
      for (int a = 0; a < 3; ++a) {
        o.g.org[a] = I.dom[a];
        o.g.dim[a] = (int32_t)std::min<double>(
            512, std::max(1.0, std::ceil(ext[a] / bs)));
        nb *= o.g.dim[a];
      }
      o.g.inv_bs = 1.0 / bs;
      o.g.pad = 0;
      o.bin_base = n_bins;
      for (int a = 0; a < 6; ++a) o.dom[a] = I.dom[a];
      o.rho = I.rho;
      n_bins += nb;
    }
    if (n_bins >= (1LL << 31))
      return fail(ctx, "rpd_run: too many candidate bins");
    HIP_TRY(ctx->inst_dev.need(ii.size() * sizeof(InstInfo)));
    HIP_TRY(hipMemcpyAsync(ctx->inst_dev.p, ii.data(),
                           ii.size() * sizeof(InstInfo),
                           hipMemcpyHostToDevice, s));
    // (the host vector must outlive the copy)
    HIP_TRY(hipStreamSynchronize(s));
  }
  const InstInfo* inst_d = ctx->inst_dev.as<InstInfo>();
  const int32_t* site_inst_d = ctx->site_inst.as<int32_t>();
  const int32_t* tet_inst_d = ctx->tet_inst.as<int32_t>();
  {
    HIP_TRY(ctx->bbox.need((size_t)std::max(ns, 1) * 48));
    if (ns > 0)
      hipLaunchKernelGGL(k_site_bbox, dim3(ns), dim3(64), 0, s, S, inst_d,
                         site_inst_d,
                         ctx->bbox.as<double>(), ctx->knn_near.as<int32_t>(),
                         ctx->knn_site_near.as<double4>(), cnt + 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[10], s));
    // sphere -> bin registrations: count, expand, radix sort by
    // (bin, sphere), bin starts by binary search
    HIP_TRY(ctx->reg_n.need(((size_t)ns + 1) * 8));
    HIP_TRY(ctx->reg_off.need(((size_t)ns + 1) * 8));
    hipLaunchKernelGGL(k_bin_nreg, dim3(ceil_div(ns + 1, 256)), dim3(256), 0,
                       s, ns, ctx->bbox.as<double>(), inst_d, site_inst_d,
                       ctx->reg_n.as<int64_t>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(exclusive_sum<int64_t>(ctx, ctx->reg_n.as<int64_t>(),
                                   ctx->reg_off.as<int64_t>(), ns + 1));
    int64_t n_reg = 0;
    HIP_TRY(hipMemcpyAsync(&n_reg, ctx->reg_off.as<int64_t>() + ns, 8,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (n_reg >= (1LL << 31))
      return fail(ctx, "rpd_run: too many sphere-bin registrations");
    ctx->st.n_bin_registrations = n_reg;
    HIP_TRY(ctx->keys_a.need((size_t)std::max<int64_t>(n_reg, 1) * 8));
    HIP_TRY(ctx->keys_b.need((size_t)std::max<int64_t>(n_reg, 1) * 8));
    HIP_TRY(ctx->bin_start.need((n_bins + 1) * 4));
    HIP_TRY(ctx->bin_ids.need((size_t)std::max<int64_t>(n_reg, 1) * 4));
    if (n_reg > 0) {
      hipLaunchKernelGGL(k_bin_expand, dim3(ceil_div(n_reg, 256)), dim3(256),
                         0, s, ns, n_reg, ctx->bbox.as<double>(), inst_d,
                         site_inst_d, ctx->reg_off.as<int64_t>(),
                         ctx->keys_a.as<uint64_t>());
      HIP_TRY(hipGetLastError());
      int end_bit = 32;
      while ((1LL << (end_bit - 32)) < n_bins + 1) ++end_bit;
      size_t tmp = 0;
      HIP_TRY(hipcub::DeviceRadixSort::SortKeys(
          nullptr, tmp, ctx->keys_a.as<uint64_t>(), ctx->keys_b.as<uint64_t>(),
          (int)n_reg, 0, end_bit, s));
      HIP_TRY(ctx->scan_tmp.need(tmp));
      HIP_TRY(hipcub::DeviceRadixSort::SortKeys(
          ctx->scan_tmp.p, tmp, ctx->keys_a.as<uint64_t>(),
          ctx->keys_b.as<uint64_t>(), (int)n_reg, 0, end_bit, s));
    }
    hipLaunchKernelGGL(k_bin_starts,
                       dim3(ceil_div(std::max<int64_t>(n_bins + 1, n_reg), 256)),
                       dim3(256), 0, s, n_bins, n_reg,
                       ctx->keys_b.as<uint64_t>(),
                       ctx->bin_start.as<int32_t>(),
                       ctx->bin_ids.as<int32_t>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[11], s));
    HIP_TRY(ctx->tet_count.need(((size_t)nt + 1) * 8));
    HIP_TRY(ctx->tet_off.need(((size_t)nt + 1) * 8));
    HIP_TRY(ctx->tet_pass.need(((size_t)nt + 1) * 8));
    HIP_TRY(hipMemsetAsync(ctx->tet_count.as<int64_t>() + nt, 0, 8, s));
    hipLaunchKernelGGL(k_tet_count, dim3(ceil_div(nt, 256)), dim3(256), 0, s,
                       M, S, inst_d, tet_inst_d, ctx->bin_start.as<int32_t>(),
                       ctx->bin_ids.as<int32_t>(),
                       ctx->tet_count.as<int64_t>(),
                       ctx->tet_pass.as<uint64_t>(),
                       cfg.count_work ? ctx->rel_work.as<unsigned long long>()
                                      : nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[12], s));
    HIP_TRY(exclusive_sum<int64_t>(ctx, ctx->tet_count.as<int64_t>(),
                                   ctx->tet_off.as<int64_t>(), nt + 1));
  }
  int64_t n_cand = 0;
  HIP_TRY(hipMemcpyAsync(&n_cand, ctx->tet_off.as<int64_t>() + nt, 8,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(ctx->cand.need((size_t)std::max<int64_t>(n_cand, 1) * sizeof(Cand)));
  HIP_TRY(ctx->rel_ids.need((size_t)std::max<int64_t>(n_cand, 1) * 4));
  hipLaunchKernelGGL(k_tet_fill, dim3(ceil_div(nt, 256)), dim3(256), 0, s, M,
                     S, inst_d, tet_inst_d, ctx->bin_start.as<int32_t>(),
                     ctx->bin_ids.as<int32_t>(), ctx->tet_off.as<int64_t>(),
                     ctx->tet_pass.as<uint64_t>(), ctx->cand.as<Cand>(),
                     ctx->rel_ids.as<int32_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ctx->ev[13], s));
  if (cfg.use_active_mask && n_cand > 0)
    hipLaunchKernelGGL(k_cand_mask, dim3(ceil_div(n_cand, 256)), dim3(256), 0,
                       s, M, S, n_cand, ctx->cand.as<Cand>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ctx->ev[1], s));

  // ---- 2. clipping: small tier, then full capacity, then exact
  HIP_TRY(ctx->rerun_large.need((size_t)std::max<int64_t>(n_cand, 1) * 4));
  HIP_TRY(ctx->rerun_exact.need((size_t)std::max<int64_t>(n_cand, 1) * 4));
  ClipArgs a;
  a.M = M;
  a.S = S;
  a.cand = ctx->cand.as<Cand>();
  a.use_mask = cfg.use_active_mask;
  a.force_exact = cfg.force_exact;
  a.geometry = cfg.compute_geometry;
  a.rerun_large = ctx->rerun_large.as<int32_t>();
  a.n_rerun_large = cnt + 1;
  a.rerun_exact = ctx->rerun_exact.as<int32_t>();
  a.n_rerun_exact = cnt + 2;
  a.work = nullptr;
  if (cfg.count_work) {
    HIP_TRY(ctx->work.need(3 * 64 * 4 * 8));
    HIP_TRY(hipMemsetAsync(ctx->work.p, 0, 3 * 64 * 4 * 8, s));
    a.work = ctx->work.as<unsigned long long>();
  }
  const int64_t n_first = n_cand;
  if (cfg.small_tier) {
    HIP_TRY(ctx->small.need(n_first, sc.V, sc.P, sc.E));
    a.items = nullptr;
    a.n_items = n_first;
    a.out = ctx->small.view(sc.V, sc.P, sc.E);
    if (int rc = launch_clip(ctx, 0, G, a, cfg.small_caps)) return rc;
  } else {
    // every cell goes straight to the full-capacity tier through the
    // re-run list (identity)
    std::vector<int32_t> idx(n_first);
    for (int64_t q = 0; q < n_first; ++q) idx[q] = (int32_t)q;
    if (n_first)
      HIP_TRY(hipMemcpyAsync(ctx->rerun_large.p, idx.data(), n_first * 4,
                             hipMemcpyHostToDevice, s));
    int32_t nf = (int32_t)n_first;
    HIP_TRY(hipMemcpyAsync(cnt + 1, &nf, 4, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->small.need(1, sc.V, sc.P, sc.E));
  }
  HIP_TRY(hipEventRecord(ctx->ev[2], s));
  int32_t host_cnt[8];
  HIP_TRY(hipMemcpyAsync(host_cnt, cnt, 32, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const int32_t n_large = host_cnt[1];
  HIP_TRY(ctx->large.need(n_large, LargeCaps::V, LargeCaps::P, LargeCaps::E));
  if (n_large > 0) {
    a.items = ctx->rerun_large.as<int32_t>();
    a.n_items = n_large;
    a.out = ctx->large.view(LargeCaps::V, LargeCaps::P, LargeCaps::E);
    if (a.work) a.work = ctx->work.as<unsigned long long>() + 256;
    if (int rc = launch_clip(ctx, 1, std::min(G, 4), a)) return rc;
    HIP_TRY(hipMemcpyAsync(host_cnt, cnt, 32, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  const int32_t n_exact = host_cnt[2];
  HIP_TRY(ctx->exact.need(n_exact, LargeCaps::V, LargeCaps::P, LargeCaps::E));
  if (n_exact > 0) {
    a.items = ctx->rerun_exact.as<int32_t>();
    a.n_items = n_exact;
    a.out = ctx->exact.view(LargeCaps::V, LargeCaps::P, LargeCaps::E);
    if (a.work) a.work = ctx->work.as<unsigned long long>() + 512;
    if (int rc = launch_clip(ctx, 2, 1, a)) return rc;
  }
  HIP_TRY(hipEventRecord(ctx->ev[3], s));

  // ---- 3. compaction
  HIP_TRY(ctx->loc.need((size_t)std::max<int64_t>(n_cand, 1) * 4));
  HIP_TRY(hipMemsetAsync(ctx->loc.p, 0, (size_t)std::max<int64_t>(n_cand, 1) * 4, s));
  if (n_large > 0)
    hipLaunchKernelGGL(k_mark_rerun, dim3(ceil_div(n_large, 256)), dim3(256),
                       0, s, ctx->rerun_large.as<int32_t>(), n_large, 1,
                       ctx->loc.as<int32_t>());
  if (n_exact > 0)
    hipLaunchKernelGGL(k_mark_rerun, dim3(ceil_div(n_exact, 256)), dim3(256),
                       0, s, ctx->rerun_exact.as<int32_t>(), n_exact, 2,
                       ctx->loc.as<int32_t>());
  TierViews tv;
  tv.t[0] = ctx->small.view(sc.V, sc.P, sc.E);
  tv.t[1] = ctx->large.view(LargeCaps::V, LargeCaps::P, LargeCaps::E);
  tv.t[2] = ctx->exact.view(LargeCaps::V, LargeCaps::P, LargeCaps::E);
  HIP_TRY(ctx->sizes.need(((size_t)n_cand + 1) * sizeof(Sizes3)));
  HIP_TRY(ctx->offs.need(((size_t)n_cand + 1) * sizeof(Sizes3)));
  HIP_TRY(ctx->drops.need(9 * 8));
  HIP_TRY(hipMemsetAsync(ctx->drops.p, 0, 9 * 8, s));
  HIP_TRY(hipMemsetAsync(ctx->sizes.as<Sizes3>() + n_cand, 0, sizeof(Sizes3), s));
  if (n_cand > 0)
    hipLaunchKernelGGL(k_sizes, dim3(ceil_div(n_cand, 256)), dim3(256), 0, s,
                       n_cand, ctx->loc.as<int32_t>(), tv,
                       ctx->sizes.as<Sizes3>(), ctx->drops.as<int64_t>());
  HIP_TRY(hipGetLastError());
  {
    size_t tmp = 0;
    Sizes3 zero{0, 0, 0};
    HIP_TRY(hipcub::DeviceScan::ExclusiveScan(
        nullptr, tmp, ctx->sizes.as<Sizes3>(), ctx->offs.as<Sizes3>(),
        Sizes3Sum(), zero, (int)(n_cand + 1), s));
    HIP_TRY(ctx->scan_tmp.need(tmp));
    HIP_TRY(hipcub::DeviceScan::ExclusiveScan(
        ctx->scan_tmp.p, tmp, ctx->sizes.as<Sizes3>(), ctx->offs.as<Sizes3>(),
        Sizes3Sum(), zero, (int)(n_cand + 1), s));
  }
  Sizes3 tot;
  HIP_TRY(hipMemcpyAsync(&tot, ctx->offs.as<Sizes3>() + n_cand, sizeof(Sizes3),
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(ctx->out_cells.need((size_t)std::max<int64_t>(tot.cells, 1) * sizeof(rpd_cell)));
  HIP_TRY(ctx->out_words.need((size_t)std::max<int64_t>(tot.words, 1) * 4));
  HIP_TRY(ctx->out_geom.need((size_t)std::max<int64_t>(tot.cells, 1) * 32));
  HIP_TRY(ctx->out_areas.need((size_t)std::max<int64_t>(tot.areas, 1) * 8));
  if (n_cand > 0)
    hipLaunchKernelGGL(k_compact, dim3(ceil_div(n_cand * kCompactLanes, 256)),
                       dim3(256), 0, s,
                       n_cand, ctx->cand.as<Cand>(), ctx->loc.as<int32_t>(), tv,
                       ctx->offs.as<Sizes3>(), ctx->out_cells.as<rpd_cell>(),
                       ctx->out_words.as<uint32_t>(),
                       ctx->out_geom.as<double>(), ctx->out_areas.as<double>(),
                       cfg.compute_geometry);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ctx->ev[4], s));

  // ---- statistics (outside the timed phases)
  HIP_TRY(ctx->stats64.need(4 * 8));
  HIP_TRY(hipMemsetAsync(ctx->stats64.p, 0, 4 * 8, s));
  if (n_cand > 0 && cfg.count_work)
    hipLaunchKernelGGL(k_mask_stats, dim3(ceil_div(n_cand, 256)), dim3(256), 0,
                       s, n_cand, ctx->cand.as<Cand>(), S,
                       ctx->stats64.as<unsigned long long>());
  int64_t st64[4], drops[9];
  HIP_TRY(hipMemcpyAsync(st64, ctx->stats64.p, 32, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(drops, ctx->drops.p, 72, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(host_cnt, cnt, 32, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  // tet_k: max candidates of one tet
  {
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceReduce::Max(nullptr, tmp, ctx->tet_count.as<int64_t>(),
                                      ctx->stats64.as<int64_t>() + 2, nt, s));
    HIP_TRY(ctx->scan_tmp.need(tmp));
    HIP_TRY(hipcub::DeviceReduce::Max(ctx->scan_tmp.p, tmp,
                                      ctx->tet_count.as<int64_t>(),
                                      ctx->stats64.as<int64_t>() + 2, nt, s));
    HIP_TRY(hipMemcpyAsync(st64, ctx->stats64.p, 32, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  rpd_stats& st = ctx->st;
  st.n_cells = tot.cells;
  st.n_candidates = n_cand;
  st.n_words = tot.words;
  st.n_areas = tot.areas;
  for (int q = 0; q < 9; ++q) st.n_dropped[q] = drops[q];
  st.n_rerun_large = n_large;
  st.n_rerun_exact = n_exact;
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
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[4]));
  st.ms_total = ms;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  st.ms_phase[0] = ms;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
  st.ms_phase[1] = ms;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
  st.ms_phase[2] = ms;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]));
  st.ms_phase[3] = ms;
  // kernel-level: site boxes, binning, relation count, relation fill, masks
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[10]));
  st.ms_kernel[0] = ms;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[10], ctx->ev[11]));
  st.ms_kernel[1] = ms;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[11], ctx->ev[12]));
  st.ms_kernel[2] = ms;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[12], ctx->ev[13]));
  st.ms_kernel[3] = ms;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[13], ctx->ev[1]));
  st.ms_kernel[4] = ms;
  st.ms_kernel[5] = st.ms_phase[1];
  st.ms_kernel[6] = st.ms_phase[3];
  st.n_bbox_fallback = host_cnt[0];
  st.tet_k = (int32_t)st64[2];
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

int rpd_tune(rpd_ctx* ctx, const rpd_run_cfg* base, int32_t* best_G,
             int32_t* best_caps, double* ms_out /* 6, or NULL */) {
  if (!ctx) return -1;
  rpd_run_cfg cfg{};
  cfg.use_active_mask = 1;
  cfg.compute_geometry = 1;
  cfg.small_tier = 1;
  if (base) cfg = *base;
  static const int kG[3] = {4, 8, 16};
  double best = 1e300;
  int bg = 8, bc = 0;
  for (int c = 0; c < 2; ++c)
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
  HIP_TRY(hipStreamSynchronize(ctx->stream));
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
  HIP_TRY(hipStreamSynchronize(s));
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
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
