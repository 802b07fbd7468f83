// abi.hip -- extern "C" boundary of libgcmf_hip.so (include/gcmf_hip.h):
// contexts, device memory helpers, kernel timers, and the batched join / clip
// entry points.  Library exceptions become negative status codes here.
#include <cstring>
#include <functional>
#include <string>

#include <vector>

#include "../../include/gcmf_hip.h"
#include "../../include/gcmf_host.h"
#include "gcmf_internal.hpp"
#include "gcmf_join.hpp"

namespace {

template <typename F>
int guarded(gcmf_ctx *ctx, F &&f) {
  if (!ctx) return GCMF_ERR_ARG;
  try {
    HIP_CHECK(hipSetDevice(ctx->c.device));
    f();
    return GCMF_OK;
  } catch (const gcmf::Error &e) {
    ctx->c.last_error = e.what;
    return e.code;
  } catch (const std::exception &e) {
    ctx->c.last_error = e.what();
    return GCMF_ERR_ALLOC;
  }
}

void no_graph(gcmf_ctx *ctx) {
  if (ctx->graph)
    throw gcmf::Error{GCMF_ERR_ARG,
                      "a captured graph holds this context's workspace (gcmf_graph_destroy)"};
}

gcmf::JoinInputs to_inputs(const gcmf_dlayer *b, const gcmf_dlayer *o) {
  gcmf::JoinInputs in;
  in.B = b->npoly;
  in.O = o->npoly;
  in.bCoords = b->d_coords;
  in.oCoords = o->d_coords;
  in.bVNum = b->d_vnum;
  in.oVNum = o->d_vnum;
  in.bVPSNum = b->d_vpsnum;
  in.oVPSNum = o->d_vpsnum;
  in.bMBR = b->d_mbr;
  in.oMBR = o->d_mbr;
  in.bMBR2 = b->d_mbr2;
  in.oMBR2 = o->d_mbr2;
  return in;
}

void run_join(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
              gcmf_join_stats *st) {
  if (!base || !over) throw gcmf::Error{GCMF_ERR_ARG, "null layer"};
  if (ctx->c.replay == 0) no_graph(ctx);
  ctx->join_valid = false;
  ctx->jin = to_inputs(base, over);
  ctx->K = gcmf::mbr_filter(ctx->c, base->npoly, over->npoly, base->d_mbr, over->d_mbr,
                            &ctx->counts, &ctx->pairs, &ctx->sv, base->d_vnum, over->d_vnum);
  gcmf::join_refine(ctx->c, ctx->jin, ctx->K, ctx->pairs, ctx->jr);
  ctx->c.flush_timers();
  ctx->join_valid = true;
  if (st) {
    st->candidates = ctx->K;
    st->pip_candidates = ctx->jr.pipNum;
    st->within = ctx->jr.within;
    st->cmf_candidates = ctx->jr.cmfCand;
    st->survivors = ctx->jr.E;
    st->b_edges = ctx->jr.nbEdges;
    st->o_edges = ctx->jr.noEdges;
    st->intersecting = ctx->jr.hits;
    st->joined = ctx->jr.nJoined;
  }
}

void fill_clip_stats(gcmf_ctx *ctx, gcmf_clip_stats *st) {
  if (!st) return;
  std::memset(st, 0, sizeof(*st));
  const long M = ctx->cr.M;
  st->pairs = M;
  st->total_comps = ctx->cr.totalComps;
  st->total_verts = ctx->cr.totalVerts;
  st->edges_p = ctx->p1.edgesP;
  st->edges_q = ctx->p1.edgesQ;
  st->slots_p = ctx->p1.slotsP;
  st->slots_q = ctx->p1.slotsQ;
  st->n_eval = ctx->cr.nEvalP;
  st->n_intersections = ctx->cr.totalIntersections;
  st->n_tested = ctx->cr.nTested;
  st->ok = ctx->cr.nOk;
  st->invalid = ctx->cr.nInvalid;
  st->disjoint = ctx->cr.nDisjoint;
  st->failed = ctx->cr.nFailed;
}

void run_clip(gcmf_ctx *ctx, const gcmf::JoinInputs &in, const int *d_pairs, long M, int flags,
              gcmf_clip_stats *st) {
  if (ctx->c.replay == 0) no_graph(ctx);
  ctx->clip_valid = false;
  gcmf::ClipOptions opt;
  // frame bits: 1 CMBR normalisation, 4 edge scale (unless the unit frame is asked for)
  opt.normalize = (flags & GCMF_CLIP_NORMALIZE) ? ((flags & GCMF_CLIP_UNIT_FRAME) ? 1 : 5) : 0;
  opt.cleanup = (flags & GCMF_CLIP_NO_CLEANUP) ? 0 : 1;
  opt.label_mode = (flags & GCMF_CLIP_SEQ_LABEL) ? 1 : 0;
  opt.route = (flags & GCMF_CLIP_GPU_ROUTE) ? 0 : 1;
  gcmf::clip_pairs(ctx->c, in, d_pairs, M, opt, ctx->cr, &ctx->p1);
  ctx->c.flush_timers();
  ctx->clip_valid = true;
  ctx->clip_nested = false;
  fill_clip_stats(ctx, st);
}

// Runs `run` eagerly, recording what it reads back, then captures it on the
// context's stream into one graph: every read-back is answered from the
// record, and a kernel in the graph sets the context's graph flag when the
// value on the device differs.  run(true) is the eager run.
void capture(gcmf_ctx *ctx, const std::function<void(bool)> &run) {
  no_graph(ctx);
  gcmf::Context &c = ctx->c;
  c.rec.clear();
  c.rec_bytes.clear();
  c.rec_pos = c.rec_bytes_pos = 0;
  c.graph_flag = (int *)c.ws.get("graph_flag", 256);
  c.replay = 1;
  try {
    run(true);
  } catch (...) {
    c.replay = 0;
    throw;
  }
  const bool timing0 = c.timing;
  c.timing = false;
  c.replay = 2;
  c.ws.no_alloc = true;
  auto restore = [&] {
    c.replay = 0;
    c.ws.no_alloc = false;
    c.timing = timing0;
  };
  const hipError_t eb = hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed);
  if (eb != hipSuccess) {
    restore();
    HIP_CHECK(eb);
  }
  try {
    gcmf::fill_i32(c, c.graph_flag, 0, 1);
    gcmf::lb_reset(c);
    run(false);
    if (c.rec_pos != c.rec.size() || c.rec_bytes_pos != c.rec_bytes.size())
      throw gcmf::Error{GCMF_ERR_ARG, "capture: read-back sequence differs from the recorded run"};
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    restore();
    throw;
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c.stream, &g);
  restore();
  HIP_CHECK(e);
  hipGraphExec_t ex = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_CHECK(e2);
  ctx->graph = ex;
}

void copy_out(gcmf::Context &c, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return;
  if (!src) throw gcmf::Error{GCMF_ERR_ARG, "field not available"};
  HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace

extern "C" {

int gcmf_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int gcmf_init(int device, gcmf_ctx **out) {
  if (!out) return GCMF_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GCMF_ERR_NODEV;
  if (device < 0 || device >= n) return GCMF_ERR_ARG;
  gcmf_ctx *ctx = new gcmf_ctx();
  ctx->c.device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->c.stream) != hipSuccess) {
    delete ctx;
    return GCMF_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    ctx->c.cu_count = prop.multiProcessorCount;
  try {
    ctx->c.scratch = gcmf::noalloc_scratch_address();
  } catch (const gcmf::Error &) {
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
    return GCMF_ERR_HIP;
  }
  *out = ctx;
  return GCMF_OK;
}

void gcmf_shutdown(gcmf_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  (void)hipStreamSynchronize(ctx->c.stream);
  if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
  ctx->c.ws.release_all();
  for (auto &p : ctx->c.pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  for (auto e : ctx->c.event_pool) (void)hipEventDestroy(e);
  if (ctx->c.hbuf) (void)hipHostFree(ctx->c.hbuf);
  if (ctx->c.side) {
    (void)hipStreamSynchronize(ctx->c.side);
    (void)hipStreamDestroy(ctx->c.side);
    (void)hipEventDestroy(ctx->c.ev_fork);
    (void)hipEventDestroy(ctx->c.ev_join);
  }
  (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

const char *gcmf_last_error(const gcmf_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : ""; }

int gcmf_malloc(gcmf_ctx *ctx, void **d_ptr, size_t bytes) {
  return guarded(ctx, [&] { HIP_CHECK(hipMalloc(d_ptr, bytes ? bytes : 1)); });
}

int gcmf_free(gcmf_ctx *ctx, void *d_ptr) {
  return guarded(ctx, [&] { HIP_CHECK(hipFree(d_ptr)); });
}

int gcmf_memcpy_h2d(gcmf_ctx *ctx, void *d_dst, const void *src, size_t bytes) {
  return guarded(ctx, [&] {
    HIP_CHECK(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
    HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
  });
}

int gcmf_memcpy_d2h(gcmf_ctx *ctx, void *dst, const void *d_src, size_t bytes) {
  return guarded(ctx, [&] { copy_out(ctx->c, dst, d_src, bytes); });
}

int gcmf_radix_sort_u64(gcmf_ctx *ctx, const uint64_t *d_keys_in, const int *d_vals_in,
                        uint64_t *d_keys_out, int *d_vals_out, long n, int bits) {
  return guarded(ctx, [&] {
    if (n < 0 || bits < 0 || bits > 64 || (n > 0 && (!d_keys_in || !d_vals_in || !d_keys_out ||
                                                      !d_vals_out)))
      throw gcmf::Error{GCMF_ERR_ARG, "bad radix sort arguments"};
    if (n >= 0x7fffffffL) throw gcmf::Error{GCMF_ERR_ARG, "radix sort of 2^31 keys or more"};
    no_graph(ctx);
    gcmf::radix_sort_u64(ctx->c, d_keys_in, d_vals_in, d_keys_out, d_vals_out, n, bits);
    gcmf::check_lookback(ctx->c);
  });
}

int gcmf_radix_sort_u64_noalloc(gcmf_ctx *ctx, const uint64_t *d_keys_in, const int *d_vals_in,
                                uint64_t *d_keys_out, int *d_vals_out, long n, int bits) {
  return guarded(ctx, [&] {
    if (n < 0 || bits < 0 || bits > 64 || (n > 0 && (!d_keys_in || !d_vals_in || !d_keys_out ||
                                                      !d_vals_out)))
      throw gcmf::Error{GCMF_ERR_ARG, "bad radix sort arguments"};
    gcmf::radix_sort_u64_noalloc(ctx->c, d_keys_in, d_vals_in, d_keys_out, d_vals_out, n, bits);
  });
}

int gcmf_scan_i32_i64(gcmf_ctx *ctx, const int *d_in, long long *d_out, long n) {
  return guarded(ctx, [&] {
    if (n < 0 || !d_out || (n > 0 && !d_in)) throw gcmf::Error{GCMF_ERR_ARG, "bad scan arguments"};
    no_graph(ctx);
    gcmf::scan_i32_to_i64(ctx->c, d_in, d_out, n);
    gcmf::check_lookback(ctx->c);
  });
}

int gcmf_compact_flags(gcmf_ctx *ctx, const int *d_flags, long n, int *d_idx, long *count) {
  return guarded(ctx, [&] {
    if (n < 0 || !count || (n > 0 && (!d_flags || !d_idx)))
      throw gcmf::Error{GCMF_ERR_ARG, "bad compaction arguments"};
    if (n >= 0x7fffffffL) throw gcmf::Error{GCMF_ERR_ARG, "compaction of 2^31 flags or more"};
    no_graph(ctx);
    *count = gcmf::compact_flags(ctx->c, d_flags, n, d_idx, "api");
    gcmf::check_lookback(ctx->c);
  });
}

int gcmf_synchronize(gcmf_ctx *ctx) {
  return guarded(ctx, [&] { HIP_CHECK(hipStreamSynchronize(ctx->c.stream)); });
}

int gcmf_mem_info(gcmf_ctx *ctx, size_t *free_bytes, size_t *total_bytes,
                  size_t *workspace_bytes) {
  return guarded(ctx, [&] {
    size_t f = 0, t = 0;
    HIP_CHECK(hipMemGetInfo(&f, &t));
    size_t w = 0;
    for (const auto &kv : ctx->c.ws.slots) w += kv.second.bytes;
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    if (workspace_bytes) *workspace_bytes = w;
  });
}

int gcmf_set_timing(gcmf_ctx *ctx, int on) {
  return guarded(ctx, [&] {
    ctx->c.flush_timers();
    ctx->c.timing = on != 0;
  });
}

int gcmf_set_timing_only(gcmf_ctx *ctx, const char *kernel) {
  return guarded(ctx, [&] {
    ctx->c.flush_timers();
    ctx->c.timing_only = kernel ? kernel : "";
  });
}

int gcmf_timer_reset(gcmf_ctx *ctx) {
  return guarded(ctx, [&] {
    ctx->c.flush_timers();
    ctx->c.timers.clear();
  });
}

int gcmf_timer_get(gcmf_ctx *ctx, const char *kernel, double *total_ms, long *launches) {
  return guarded(ctx, [&] {
    ctx->c.flush_timers();
    auto it = ctx->c.timers.find(kernel ? kernel : "");
    if (total_ms) *total_ms = it == ctx->c.timers.end() ? 0.0 : it->second.total_ms;
    if (launches) *launches = it == ctx->c.timers.end() ? 0 : it->second.launches;
  });
}

int gcmf_join(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
              gcmf_join_stats *stats) {
  return guarded(ctx, [&] { run_join(ctx, base, over, stats); });
}

int gcmf_join_get(gcmf_ctx *ctx, int field, void *dst, long count) {
  return guarded(ctx, [&] {
    if (!ctx->join_valid) throw gcmf::Error{GCMF_ERR_ARG, "no join result on this context"};
    if (count < 0) throw gcmf::Error{GCMF_ERR_ARG, "negative count"};
    const long K = ctx->K, N = ctx->jin.B + ctx->jin.O;
    const void *src = nullptr;
    size_t elem = 4;
    long maxc = 0;
    auto &r = ctx->jr;
    switch (field) {
      case GCMF_J_PAIRS: src = ctx->pairs; maxc = 2 * K; break;
      case GCMF_J_COUNTS: src = ctx->counts; maxc = N; break;
      case GCMF_J_PERM: src = ctx->sv.perm; maxc = 2 * N; break;
      case GCMF_J_RANK: src = ctx->sv.rank; maxc = 2 * N; break;
      case GCMF_J_CMBR: src = r.cmbr; maxc = 4 * K; elem = 8; break;
      case GCMF_J_PIPTYPE: src = r.pipType; maxc = K; elem = 1; break;
      case GCMF_J_PIPFLAG: src = r.pipFlag; maxc = K; break;
      case GCMF_J_JOINFLAG: src = r.joinFlag; maxc = K; break;
      case GCMF_J_BCOUNT: src = r.bCount; maxc = K; break;
      case GCMF_J_OCOUNT: src = r.oCount; maxc = K; break;
      case GCMF_J_BEDGE: src = r.bEdge; maxc = r.nbEdges; break;
      case GCMF_J_OEDGE: src = r.oEdge; maxc = r.noEdges; break;
      case GCMF_J_SEGFLAG: src = r.segFlag; maxc = K; break;
      case GCMF_J_JOINED: src = r.joinedPairs; maxc = 2 * r.nJoined; break;
      default: throw gcmf::Error{GCMF_ERR_ARG, "unknown join field"};
    }
    if (count > maxc) throw gcmf::Error{GCMF_ERR_ARG, "count exceeds field size"};
    copy_out(ctx->c, dst, src, elem * (size_t)count);
  });
}

int gcmf_clip(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
              const int *d_pairs, long M, int flags, gcmf_clip_stats *stats) {
  return guarded(ctx, [&] {
    if (!base || !over || (M > 0 && !d_pairs) || M < 0)
      throw gcmf::Error{GCMF_ERR_ARG, "bad clip arguments"};
    run_clip(ctx, to_inputs(base, over), d_pairs, M, flags, stats);
  });
}

int gcmf_clip_complex(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
                      const int *d_parts, long M, int flags, gcmf_clip_stats *stats) {
  return guarded(ctx, [&] {
    if (!base || !over || (M > 0 && !d_parts) || M < 0)
      throw gcmf::Error{GCMF_ERR_ARG, "bad clip arguments"};
    no_graph(ctx);
    ctx->clip_valid = false;
    gcmf::ClipOptions opt;
    opt.normalize = (flags & GCMF_CLIP_NORMALIZE) ? ((flags & GCMF_CLIP_UNIT_FRAME) ? 1 : 5) : 0;
    opt.cleanup = (flags & GCMF_CLIP_NO_CLEANUP) ? 0 : 1;
    gcmf::clip_complex(ctx->c, to_inputs(base, over), d_parts, M, opt,
                       (flags & GCMF_CLIP_UNION) != 0, ctx->cr);
    ctx->p1 = gcmf::Phase1View();
    ctx->p1.M = M;
    ctx->c.flush_timers();
    ctx->clip_valid = true;
    ctx->clip_nested = true;
    fill_clip_stats(ctx, stats);
  });
}

int gcmf_get_tuning(gcmf_ctx *ctx, gcmf_tuning *out) {
  return guarded(ctx, [&] {
    if (!out) throw gcmf::Error{GCMF_ERR_ARG, "null tuning"};
    out->cull_min = ctx->c.tune.cullMin;
    out->big_nodes = ctx->c.tune.bigNodes;
    out->count_block = ctx->c.tune.countBlock;
    out->lds_max_q = ctx->c.tune.ldsMaxQ;
    out->label_lds = ctx->c.tune.labelLds;
    out->grid_nodes = ctx->c.tune.gridNodes;
    out->grid_clean = ctx->c.tune.gridClean;
    out->strip_min = ctx->c.tune.stripMin;
    out->cmf_group = ctx->c.tune.cmfGroup;
    out->label_steal = ctx->c.tune.labelSteal;
    out->block_cap = ctx->c.tune.blockCap;
    out->count_stage = ctx->c.tune.countStage;
  });
}

int gcmf_set_tuning(gcmf_ctx *ctx, const gcmf_tuning *t) {
  return guarded(ctx, [&] {
    if (!t || t->cull_min < 1 || t->big_nodes < 1 ||
        (t->count_block != 64 && t->count_block != 128 && t->count_block != 256) ||
        t->lds_max_q < 0 || t->lds_max_q > 1024 || t->label_lds < 0 || t->label_lds > 65536 ||
        t->label_lds % 16 != 0 || t->grid_nodes < 1 || t->grid_clean < 49 ||
        t->strip_min < 0 ||
        (t->cmf_group != 0 && t->cmf_group != 16 && t->cmf_group != 64) ||
        (t->label_steal != 0 && t->label_steal != 1) || t->block_cap < 1 ||
        t->block_cap > 65536 || (t->count_stage != 0 && t->count_stage != 1))
      throw gcmf::Error{GCMF_ERR_ARG, "bad tuning"};
    ctx->c.tune.cullMin = t->cull_min;
    ctx->c.tune.bigNodes = t->big_nodes;
    ctx->c.tune.countBlock = t->count_block;
    ctx->c.tune.ldsMaxQ = t->lds_max_q;
    ctx->c.tune.labelLds = t->label_lds;
    ctx->c.tune.gridNodes = t->grid_nodes;
    ctx->c.tune.gridClean = t->grid_clean;
    ctx->c.tune.stripMin = t->strip_min;
    ctx->c.tune.cmfGroup = t->cmf_group;
    ctx->c.tune.labelSteal = t->label_steal;
    ctx->c.tune.blockCap = t->block_cap;
    ctx->c.tune.countStage = t->count_stage;
  });
}

int gcmf_tune_clip(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
                   const int *d_pairs, long M, int flags, gcmf_tuning *out) {
  return guarded(ctx, [&] {
    if (!base || !over || (M > 0 && !d_pairs) || M < 0)
      throw gcmf::Error{GCMF_ERR_ARG, "bad tuning arguments"};
    no_graph(ctx);
    gcmf::Context &c = ctx->c;
    const gcmf::JoinInputs in = to_inputs(base, over);
    const bool timing0 = c.timing;
    const std::string only0 = c.timing_only;
    c.flush_timers();
    const auto saved = c.timers;
    gcmf_clip_stats st;
    // milliseconds of the launches named `kernel` in one clip
    auto measure = [&](const char *kernel) {
      c.flush_timers();
      c.timers.clear();
      c.timing = true;
      c.timing_only = kernel;
      run_clip(ctx, in, d_pairs, M, flags, &st);
      c.flush_timers();
      auto it = c.timers.find(kernel);
      return it == c.timers.end() ? 0.0 : it->second.total_ms;
    };
    gcmf::Tuning best = c.tune;
    measure("gh_count");  // warm-up: workspace growth
    {
      double bt = 1e300;
      for (long v : {128L, 256L, 512L, 1024L, 4096L}) {
        c.tune = best;
        c.tune.cullMin = v;
        const double t = measure("gh_count");
        if (t < bt) bt = t, best.cullMin = v;
      }
    }
    {
      double bt = 1e300;
      for (int v : {64, 128, 256}) {
        c.tune = best;
        c.tune.countBlock = v;
        const double t = measure("gh_count");
        if (t < bt) bt = t, best.countBlock = v;
      }
    }
    {
      double bt = 1e300;
      for (int v : {0, 1024}) {
        c.tune = best;
        c.tune.ldsMaxQ = v;
        const double t = measure("gh_count");
        if (t < bt) bt = t, best.ldsMaxQ = v;
      }
    }
    {
      double bt = 1e300;
      for (int v : {0, 1}) {
        c.tune = best;
        c.tune.countStage = v;
        const double t = measure("gh_count");
        if (t < bt) bt = t, best.countStage = v;
      }
    }
    {
      double bt = 1e300;
      for (long long v : {2048LL, 4096LL, 8192LL, 16384LL, 65536LL}) {
        c.tune = best;
        c.tune.bigNodes = v;
        const double t = measure("gh_label_trace");
        if (t < bt) bt = t, best.bigNodes = v;
      }
    }
    {
      double bt = 1e300;
      for (int v : {0, 16384, 20480, 24576, 32768, 49152, 65536}) {
        c.tune = best;
        c.tune.labelLds = v;
        const double t = measure("gh_label_trace");
        if (t < bt) bt = t, best.labelLds = v;
      }
    }
    {
      double bt = 1e300;
      for (int v : {0, 1}) {
        c.tune = best;
        c.tune.labelSteal = v;
        const double t = measure("gh_label_trace");
        if (t < bt) bt = t, best.labelSteal = v;
      }
    }
    c.tune = best;
    c.timers = saved;
    c.timing = timing0;
    c.timing_only = only0;
    if (out) {
      out->cull_min = best.cullMin;
      out->big_nodes = best.bigNodes;
      out->count_block = best.countBlock;
      out->lds_max_q = best.ldsMaxQ;
      out->label_lds = best.labelLds;
      out->grid_nodes = best.gridNodes;
      out->grid_clean = best.gridClean;
      out->strip_min = best.stripMin;
      out->cmf_group = best.cmfGroup;
      out->label_steal = best.labelSteal;
      out->block_cap = best.blockCap;
      out->count_stage = best.countStage;
    }
  });
}

int gcmf_join_clip(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over, int flags,
                   gcmf_join_stats *jstats, gcmf_clip_stats *cstats) {
  return guarded(ctx, [&] {
    run_join(ctx, base, over, jstats);
    run_clip(ctx, ctx->jin, ctx->jr.joinedPairs, ctx->jr.nJoined, flags, cstats);
  });
}

int gcmf_clip_capture(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
                      const int *d_pairs, long M, int flags, gcmf_clip_stats *stats) {
  return guarded(ctx, [&] {
    if (!base || !over || (M > 0 && !d_pairs) || M < 0)
      throw gcmf::Error{GCMF_ERR_ARG, "bad clip arguments"};
    const gcmf::JoinInputs in = to_inputs(base, over);
    capture(ctx, [&](bool eager) {
      run_clip(ctx, in, d_pairs, M, flags, eager ? stats : nullptr);
    });
  });
}

int gcmf_join_clip_capture(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
                           int flags, gcmf_join_stats *jstats, gcmf_clip_stats *cstats) {
  return guarded(ctx, [&] {
    if (!base || !over) throw gcmf::Error{GCMF_ERR_ARG, "null layer"};
    capture(ctx, [&](bool eager) {
      run_join(ctx, base, over, eager ? jstats : nullptr);
      run_clip(ctx, ctx->jin, ctx->jr.joinedPairs, ctx->jr.nJoined, flags,
               eager ? cstats : nullptr);
    });
  });
}

int gcmf_graph_launch(gcmf_ctx *ctx, int *valid) {
  return guarded(ctx, [&] {
    if (!ctx->graph) throw gcmf::Error{GCMF_ERR_ARG, "no captured graph on this context"};
    gcmf::Context &c = ctx->c;
    HIP_CHECK(hipGraphLaunch(ctx->graph, c.stream));
    if (valid) {
      int f = 1;
      HIP_CHECK(hipMemcpyAsync(&f, c.graph_flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
      HIP_CHECK(hipStreamSynchronize(c.stream));
      *valid = f == 0 ? 1 : 0;
    }
  });
}

int gcmf_graph_destroy(gcmf_ctx *ctx) {
  return guarded(ctx, [&] {
    if (!ctx->graph) return;
    HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    HIP_CHECK(hipGraphExecDestroy(ctx->graph));
    ctx->graph = nullptr;
  });
}

int gcmf_clip_areas(gcmf_ctx *ctx, double *d_out) {
  return guarded(ctx, [&] {
    if (!ctx->clip_valid) throw gcmf::Error{GCMF_ERR_ARG, "no clip result on this context"};
    if (ctx->cr.totalComps > 0 && !d_out) throw gcmf::Error{GCMF_ERR_ARG, "null output"};
    gcmf::comp_areas(ctx->c, ctx->cr, d_out);
    HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
  });
}

int gcmf_clip_write(gcmf_ctx *ctx, const char *path, int format) {
  return guarded(ctx, [&] {
    if (!ctx->clip_valid) throw gcmf::Error{GCMF_ERR_ARG, "no clip result on this context"};
    if (!path || (format != GCMF_WRITE_POLY && format != GCMF_WRITE_WKT))
      throw gcmf::Error{GCMF_ERR_ARG, "bad write arguments"};
    const auto &r = ctx->cr;
    std::vector<double> xy(2 * (size_t)r.totalVerts);
    std::vector<int> cs((size_t)r.totalComps), nc((size_t)r.M), st((size_t)r.M);
    gcmf::Context &c = ctx->c;
    if (r.totalVerts > 0)
      HIP_CHECK(hipMemcpyAsync(xy.data(), r.xy, 16 * (size_t)r.totalVerts,
                               hipMemcpyDeviceToHost, c.stream));
    if (r.totalComps > 0)
      HIP_CHECK(hipMemcpyAsync(cs.data(), r.compSizes, 4 * (size_t)r.totalComps,
                               hipMemcpyDeviceToHost, c.stream));
    if (r.M > 0) {
      HIP_CHECK(hipMemcpyAsync(nc.data(), r.ncomp, 4 * (size_t)r.M, hipMemcpyDeviceToHost,
                               c.stream));
      HIP_CHECK(hipMemcpyAsync(st.data(), r.status, 4 * (size_t)r.M, hipMemcpyDeviceToHost,
                               c.stream));
    }
    HIP_CHECK(hipStreamSynchronize(c.stream));
    const int rc = format == GCMF_WRITE_POLY
                       ? gcmf_result_write_poly(path, xy.data(), cs.data(), r.totalComps)
                       : gcmf_result_write_wkt(path, xy.data(), cs.data(), nc.data(),
                                               st.data(), r.M, ctx->clip_nested ? 1 : 0);
    if (rc != 0) throw gcmf::Error{GCMF_ERR_ARG, std::string("cannot write ") + path};
  });
}

int gcmf_clip_get(gcmf_ctx *ctx, int field, void *dst, long count) {
  return guarded(ctx, [&] {
    if (!ctx->clip_valid) throw gcmf::Error{GCMF_ERR_ARG, "no clip result on this context"};
    if (count < 0) throw gcmf::Error{GCMF_ERR_ARG, "negative count"};
    const auto &r = ctx->cr;
    const auto &p = ctx->p1;
    const long M = r.M;
    const void *src = nullptr;
    size_t elem = 4;
    long maxc = 0;
    switch (field) {
      case GCMF_C_STATUS: src = r.status; maxc = M; break;
      case GCMF_C_SWAPPED: src = r.swapped; maxc = M; break;
      case GCMF_C_NCOMP: src = r.ncomp; maxc = M; break;
      case GCMF_C_NVERTS: src = r.nverts; maxc = M; break;
      case GCMF_C_COMPOFF: src = r.compOff; maxc = M + 1; elem = 8; break;
      case GCMF_C_VERTOFF: src = r.vertOff; maxc = M + 1; elem = 8; break;
      case GCMF_C_COMPSIZES: src = r.compSizes; maxc = r.totalComps; break;
      case GCMF_C_XY: src = r.xy; maxc = 2 * r.totalVerts; elem = 8; break;
      case GCMF_C_LABEL_PATH: src = r.labelPath; maxc = r.labelPath ? M : 0; break;
      case GCMF_C_LABEL_TICKS: src = r.labelTicks; maxc = r.labelTicks ? M : 0; elem = 8; break;
      case GCMF_C_OFFP: src = p.offP; maxc = M + 1; elem = 8; break;
      case GCMF_C_OFFQ: src = p.offQ; maxc = M + 1; elem = 8; break;
      case GCMF_C_PSP1: src = p.psP1; maxc = p.edgesP + 1; elem = 8; break;
      case GCMF_C_PSP2: src = p.psP2; maxc = p.edgesP + 1; elem = 8; break;
      case GCMF_C_PSQ1: src = p.psQ1; maxc = p.edgesQ + 1; elem = 8; break;
      case GCMF_C_PSQ2: src = p.psQ2; maxc = p.edgesQ + 1; elem = 8; break;
      case GCMF_C_BOOLP: src = p.boolP; maxc = p.edgesP; break;
      case GCMF_C_BOOLQ: src = p.boolQ; maxc = p.edgesQ; break;
      case GCMF_C_NMAPQ: src = p.nmapQ; maxc = p.slotsQ; break;
      case GCMF_C_IP: src = p.iP; maxc = 2 * p.slotsP; elem = 8; break;
      case GCMF_C_IQ: src = p.iQ; maxc = 2 * p.slotsQ; elem = 8; break;
      case GCMF_C_OP: src = p.oP; maxc = 2 * p.slotsP; elem = 8; break;
      case GCMF_C_OQ: src = p.oQ; maxc = 2 * p.slotsQ; elem = 8; break;
      case GCMF_C_AP: src = p.aP; maxc = p.slotsP; break;
      case GCMF_C_AQ: src = p.aQ; maxc = p.slotsQ; break;
      case GCMF_C_LP: src = p.lP; maxc = p.slotsP; break;
      case GCMF_C_LQ: src = p.lQ; maxc = p.slotsQ; break;
      case GCMF_C_NBP: src = p.nbP; maxc = p.slotsP; break;
      case GCMF_C_NBQ: src = p.nbQ; maxc = p.slotsQ; break;
      case GCMF_C_NORM: src = p.norm; maxc = 3 * M; elem = 8; break;
      default: throw gcmf::Error{GCMF_ERR_ARG, "unknown clip field"};
    }
    if (count > maxc) throw gcmf::Error{GCMF_ERR_ARG, "count exceeds field size"};
    copy_out(ctx->c, dst, src, elem * (size_t)count);
  });
}

int gcmf_clip_last_stats(gcmf_ctx *ctx, gcmf_clip_stats *stats) {
  return guarded(ctx, [&] {
    if (!stats) throw gcmf::Error{GCMF_ERR_ARG, "null stats"};
    if (!ctx->clip_valid) throw gcmf::Error{GCMF_ERR_ARG, "no clip result on this context"};
    fill_clip_stats(ctx, stats);
  });
}

}  // extern "C"
