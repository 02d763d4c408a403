/*
 * hcs_batch.hip -- batched form: thousands of independent pairs stepped by
 * one graph replay.  No reference counterpart (the reference steps one pair
 * at a time inside mj_collision, mujoco_contact_surfaces_plugin.cpp:92-95).
 *
 * Step = [reset counters (fill kernels, no memset nodes), per-pair pose
 * composition, persistent candidate kernel over a work queue, scan of the
 * per-item survivor counts, clip/force kernel, per-pair reduction].  The
 * survivor buffer grows when a step overflows it (checked in
 * hcs_batch_get, which then re-runs the step).  The autotuner times the
 * launch shapes (persistent blocks per CU, work queues, the tet-query
 * threshold for small rigid meshes, the split target of heavy items) on the
 * current state and keeps the fastest; results never depend on the shape.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hcs_internal.h"

using namespace hcs;

struct hcs_batch {
  Ctx* c = nullptr;
  int npairs = 0;
  std::vector<int32_t> geoms;
  hcs_step_opts opts{};
  Plan plan;
  int qpl = 1, blocks_per_cu = 4, tet_max = 64;
  /* split triangle / soft items whose last measured candidate count exceeds
   * this into sub-items over LBVH subtrees (0: never) */
  int split_target = 2048;
  int queues = 1; /* candidate-kernel work queues: 1, or one per XCD */
  /* work order: 0 items by decreasing measured candidate count; 1 pairs by
   * decreasing total count, each pair's items together (decreasing count),
   * so a pair's tree and records stay in L2 while its items run; 2 items by
   * decreasing shader cycles measured in one step (the cost the tail of
   * the persistent launch depends on) */
  int order_mode = 0;
  bool has_state = false;
  DBuf envs, state, items;
  Pipe pipe;
  BatchTaxels tx; /* flat tactile sensors (hcs_batch_set_sensors) */
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int batch_free_graph(hcs_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  return HCS_OK;
}

static int grid_of(const hcs_batch* b) {
  const int nitems = (int)b->plan.items.size();
  return std::max(1, std::min(nitems, b->c->num_cu * b->blocks_per_cu));
}

static QueryArgs batch_args(hcs_batch* b) {
  QueryArgs a = pipe_args(b->pipe, (const EnvDesc*)b->envs.ptr, (const double*)b->state.ptr,
                          (const Item*)b->items.ptr, (int)b->plan.items.size(), true);
  if (b->tx.n > 0) {
    /* the clip kernel keeps the polygons of the pairs that carry a sensor */
    a.poly_rec = (double*)b->tx.poly_rec.ptr;
    a.poly_key = (int64_t*)b->tx.poly_key.ptr;
    a.poly_cnt = (unsigned long long*)b->tx.poly_cnt.ptr;
    a.poly_cap = b->tx.poly_cap;
    a.poly_env = (const unsigned char*)b->tx.mask.ptr;
  }
  return a;
}

/* Enqueues one step on the context stream. */
static void batch_enqueue(hcs_batch* b, const QueryArgs& a) {
  if (b->tx.n > 0) launch_fill_u32(b->c->stream, (unsigned*)b->tx.poly_cnt.ptr, 0u, 2);
  pipe_enqueue(b->c, b->pipe, a, b->npairs, grid_of(b), clip_grid_for(b->c, b->pipe.cap), 0, nullptr, nullptr,
               false, 0);
  batch_taxels_enqueue(b->c, b->tx, (const double*)b->state.ptr);
}

/* Uploads envs and items of b->plan and computes the static item boxes. */
static int batch_upload(hcs_batch* b) {
  Ctx* c = b->c;
  const size_t nitems = std::max<size_t>(1, b->plan.items.size());
  b->pipe.nqueues = b->queues;
  int rc;
  if ((rc = ensure(c, b->envs, sizeof(EnvDesc) * b->npairs)) ||
      (rc = ensure(c, b->state, sizeof(double) * HCS_STATE * b->npairs)) ||
      (rc = ensure(c, b->items, sizeof(Item) * nitems)) ||
      (rc = pipe_reserve(c, b->pipe, (int)b->plan.items.size(), b->npairs, std::max<int64_t>(b->pipe.cap, 65536), 0)))
    return rc;
  rc = hip_check(c, hipMemcpyAsync(b->envs.ptr, b->plan.envs.data(), sizeof(EnvDesc) * b->npairs,
                                   hipMemcpyHostToDevice, c->stream),
                 "upload envs");
  if (rc) return rc;
  if (!b->plan.items.empty())
    rc = hip_check(c, hipMemcpyAsync(b->items.ptr, b->plan.items.data(), sizeof(Item) * b->plan.items.size(),
                                     hipMemcpyHostToDevice, c->stream),
                   "upload items");
  if (!rc)
    rc = pipe_plan(c, b->pipe, (const EnvDesc*)b->envs.ptr, (const Item*)b->items.ptr, b->plan);
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

/* Splits every triangle / soft item whose candidate count in a measured
 * step exceeds split_target into sub-items over the LBVH frontier below its
 * root, at depth ceil(log2(count / split_target)) <= 6, and repeats on the
 * new plan (at most 4 rounds) while some item still exceeds it.  The
 * frontier's subtrees hold each leaf exactly once, so the candidate set is
 * unchanged. */
/* Per-item candidate counts of the current plan on the current state. */
static int batch_measure(hcs_batch* b, std::vector<int>& nc) {
  Ctx* c = b->c;
  const int n = (int)b->plan.items.size();
  nc.assign(n, 0);
  if (n == 0) return HCS_OK;
  /* counts only: no survivors, no clipping */
  const QueryArgs a = batch_args(b);
  pipe_front(c, b->pipe, a, b->npairs);
  launch_cand(c->stream, a, grid_of(b), CAND_COUNT);
  int rc = hip_check(c, hipGetLastError(), "measure step");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(nc.data(), b->pipe.ncand.ptr, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream),
                   "dl");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  return rc;
}

/* Shader cycles of each item in one eager step on the current state (0 for
 * items the step culls). */
static int batch_measure_cycles(hcs_batch* b, std::vector<unsigned long long>& cyc) {
  Ctx* c = b->c;
  const int n = (int)b->plan.items.size();
  cyc.assign(n, 0);
  if (n == 0) return HCS_OK;
  DBuf d;
  const size_t bytes = sizeof(unsigned long long) * 6 * (size_t)n;
  int rc = ensure(c, d, bytes);
  if (rc) return rc;
  launch_fill_u32(c->stream, (unsigned*)d.ptr, 0u, (int)(bytes / 4));
  QueryArgs a = batch_args(b);
  a.item_stats = (unsigned long long*)d.ptr;
  pipe_front(c, b->pipe, a, b->npairs);
  launch_cand(c->stream, a, grid_of(b), CAND_STEP);
  std::vector<unsigned long long> h(6 * (size_t)n);
  rc = hip_check(c, hipGetLastError(), "cycles step");
  if (!rc) rc = hip_check(c, hipMemcpyAsync(h.data(), d.ptr, bytes, hipMemcpyDeviceToHost, c->stream), "dl");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  (void)hipFree(d.ptr);
  if (rc) return rc;
  for (int i = 0; i < n; ++i) cyc[i] = h[6 * (size_t)i];
  return HCS_OK;
}

/* Work order of the items for the candidate kernel (order_mode): decreasing
 * measured candidate count (longest first); pairs by decreasing total count
 * with each pair's items together by decreasing count; or decreasing
 * measured shader cycles.  Ties by item index. */
static int batch_order(hcs_batch* b) {
  std::vector<int> nc;
  int rc = batch_measure(b, nc);
  if (rc) return rc;
  std::vector<unsigned long long> cyc;
  if (b->order_mode == 2 && (rc = batch_measure_cycles(b, cyc))) return rc;
  std::vector<int>& o = b->plan.order;
  o.resize(nc.size());
  for (size_t i = 0; i < o.size(); ++i) o[i] = (int)i;
  if (b->order_mode == 1) {
    std::vector<int64_t> tot(b->npairs, 0);
    for (size_t i = 0; i < nc.size(); ++i) tot[b->plan.items[i].env] += nc[i];
    std::stable_sort(o.begin(), o.end(), [&](int x, int y) {
      const int ex = b->plan.items[x].env, ey = b->plan.items[y].env;
      if (ex != ey) return tot[ex] != tot[ey] ? tot[ex] > tot[ey] : ex < ey;
      return nc[x] > nc[y];
    });
  } else if (b->order_mode == 2) {
    std::stable_sort(o.begin(), o.end(), [&](int x, int y) { return cyc[x] > cyc[y]; });
  } else {
    std::stable_sort(o.begin(), o.end(), [&](int x, int y) { return nc[x] > nc[y]; });
  }
  return batch_upload(b);
}

static int batch_split(hcs_batch* b) {
  Ctx* c = b->c;
  if (b->split_target <= 0) return HCS_OK;
  for (int round = 0; round < 4; ++round) {
    const int n = (int)b->plan.items.size();
    if (n == 0) return HCS_OK;
    std::vector<int> nc;
    int rc = batch_measure(b, nc);
    if (rc) return rc;
    std::vector<int> req, depth, slot(n, -1);
    for (int i = 0; i < n; ++i) {
      const Item& it = b->plan.items[i];
      const EnvDesc& E = b->plan.envs[it.env];
      if ((E.kind != KIND_TRI && E.kind != KIND_SOFT) || E.ntA <= 1 || it.root < 0 || nc[i] <= b->split_target)
        continue;
      int d = 1;
      while (d < 6 && ((int64_t)b->split_target << d) < nc[i]) ++d;
      slot[i] = (int)req.size();
      req.push_back(i);
      depth.push_back(d);
    }
    if (req.empty()) return HCS_OK;
    std::vector<int32_t> roots;
    if ((rc = split_roots(c, (const EnvDesc*)b->envs.ptr, (const Item*)b->items.ptr, req, depth, roots))) return rc;
    std::vector<Item> out;
    out.reserve(n + 64 * req.size());
    for (int e = 0; e < b->npairs; ++e) {
      EnvDesc& E = b->plan.envs[e];
      const int first = E.item_first, cnt = E.item_count;
      E.item_first = (int)out.size();
      for (int i = first; i < first + cnt; ++i) {
        const Item it = b->plan.items[i];
        if (slot[i] < 0) {
          out.push_back(it);
          continue;
        }
        const int32_t* r = &roots[65 * (size_t)slot[i]];
        for (int k = 0; k < r[0]; ++k) out.push_back(Item{it.env, it.first, it.count, r[1 + k]});
      }
      E.item_count = (int)out.size() - E.item_first;
    }
    b->plan.items.swap(out);
    if ((rc = batch_upload(b))) return rc;
  }
  return HCS_OK;
}

/* (Re)builds the plan for the current shape and uploads it; with a state
 * set, heavy items are split and all items ordered by measured steps. */
static int batch_plan(hcs_batch* b) {
  batch_free_graph(b);
  int rc = build_plan(b->c, b->npairs, b->geoms.data(), &b->opts, b->tet_max, b->plan);
  b->plan.order.clear();
  if (!rc) rc = batch_upload(b);
  if (!rc && b->has_state) rc = batch_split(b);
  if (!rc && b->has_state) rc = batch_order(b);
  return rc;
}

static int batch_capture(hcs_batch* b) {
  Ctx* c = b->c;
  batch_free_graph(b);
  int rc = hip_check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "BeginCapture");
  if (rc) return rc;
  batch_enqueue(b, batch_args(b));
  rc = hip_check(c, hipStreamEndCapture(c->stream, &b->graph), "EndCapture");
  if (rc) return rc;
  return hip_check(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "GraphInstantiate");
}

/* Survivor count (and, with sensors, kept polygons and taxel list entries)
 * of the last step; grows what overflowed and re-runs the step.  Returns
 * HCS_OK or an error. */
static int batch_settle(hcs_batch* b) {
  Ctx* c = b->c;
  for (int attempt = 0; attempt < 4; ++attempt) {
    unsigned long long n = 0;
    int rc = hip_check(c, hipMemcpyAsync(&n, b->pipe.surv_cnt.ptr, 8, hipMemcpyDeviceToHost, c->stream), "dl");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
    if (rc) return rc;
    int64_t npoly = 0, nlist = 0;
    if ((int64_t)n <= b->pipe.cap && (rc = batch_taxels_demand(c, b->tx, npoly, nlist))) return rc;
    const bool over_tx = b->tx.n > 0 && (npoly > b->tx.poly_cap || nlist > b->tx.list_cap);
    if ((int64_t)n <= b->pipe.cap && !over_tx) return HCS_OK;
    if ((int64_t)n > b->pipe.cap)
      rc = pipe_reserve(c, b->pipe, (int)b->plan.items.size(), b->npairs, (int64_t)n + (int64_t)n / 4 + 4096, 0);
    else
      rc = batch_taxels_reserve(c, b->tx, std::max(b->tx.poly_cap, npoly + npoly / 4 + 4096),
                                std::max(b->tx.list_cap, nlist + nlist / 4 + 4096));
    if (!rc) rc = batch_capture(b);
    if (!rc) rc = hcs_batch_step(b);
    if (rc) return rc;
  }
  return set_err(c, HCS_ERR_CAPACITY, "survivor or taxel buffers did not settle");
}

int hcs::batches_refresh(Ctx* c, int handle) {
  for (hcs_batch* b : c->batches) {
    if (std::find(b->geoms.begin(), b->geoms.end(), handle) == b->geoms.end()) continue;
    int rc = batch_plan(b);
    if (!rc) rc = batch_capture(b);
    if (rc) return rc;
  }
  return HCS_OK;
}

extern "C" {

int hcs_batch_create(hcs_ctx* h, int32_t npairs, const int32_t* geoms, const hcs_step_opts* opts,
                     hcs_batch** out) {
  if (!h || !geoms || !out || npairs <= 0) return HCS_ERR_ARG;
  *out = nullptr;
  Ctx* c = &h->c;
  int rc = hip_check(c, hipSetDevice(c->device), "hipSetDevice");
  if (rc) return rc;
  hcs_batch* b = new hcs_batch();
  b->c = c;
  c->batches.insert(b);
  b->npairs = npairs;
  b->geoms.assign(geoms, geoms + 2 * (size_t)npairs);
  if (opts) {
    b->opts = *opts;
  } else {
    b->opts.representation = 0, b->opts.apply_forces = 1, b->opts.dt = 0.001;
  }
  b->opts.keep_polygons = 0;
  rc = batch_plan(b);
  if (!rc) {
    std::vector<double> zero((size_t)HCS_STATE * npairs, 0.0);
    for (int e = 0; e < npairs; ++e) {
      double* s = &zero[(size_t)HCS_STATE * e];
      s[0] = s[4] = s[8] = s[12] = s[16] = s[20] = 1.0;
    }
    rc = hcs_batch_set_state(b, zero.data());
    b->has_state = false;
  }
  if (!rc) rc = batch_capture(b);
  if (rc) {
    hcs_batch_destroy(b);
    return rc;
  }
  *out = b;
  return HCS_OK;
}

int hcs_batch_destroy(hcs_batch* b) {
  if (!b) return HCS_ERR_ARG;
  (void)hipStreamSynchronize(b->c->stream);
  batch_free_graph(b);
  for (DBuf* d : {&b->envs, &b->state, &b->items})
    if (d->ptr) (void)hipFree(d->ptr);
  pipe_free(b->pipe);
  batch_taxels_free(b->tx);
  b->c->batches.erase(b);
  delete b;
  return HCS_OK;
}

int hcs_batch_set_state(hcs_batch* b, const double* state) {
  if (!b || !state) return HCS_ERR_ARG;
  Ctx* c = b->c;
  int rc = hip_check(c, hipMemcpyAsync(b->state.ptr, state, sizeof(double) * HCS_STATE * b->npairs,
                                       hipMemcpyHostToDevice, c->stream),
                     "upload state");
  if (rc) return rc;
  b->has_state = true;
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

int hcs_batch_step(hcs_batch* b) {
  if (!b) return HCS_ERR_ARG;
  Ctx* c = b->c;
  if (b->exec) return hip_check(c, hipGraphLaunch(b->exec, c->stream), "hipGraphLaunch");
  batch_enqueue(b, batch_args(b));
  return hip_check(c, hipGetLastError(), "batch step");
}

int hcs_batch_get(hcs_batch* b, hcs_pair_out* out) {
  if (!b || !out) return HCS_ERR_ARG;
  Ctx* c = b->c;
  int rc = batch_settle(b);
  if (rc) return rc;
  rc = hip_check(c, hipMemcpyAsync(out, b->pipe.out.ptr, sizeof(PairOut) * b->npairs, hipMemcpyDeviceToHost, c->stream),
                     "download");
  if (rc) return rc;
  rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  for (int e = 0; e < b->npairs; ++e)
    if (out[e].status) return set_err(c, (int)out[e].status, "traversal stack overflow or survivor overflow");
  return HCS_OK;
}

int hcs_batch_counts(hcs_batch* b, int64_t* counts) {
  if (!b || !counts) return HCS_ERR_ARG;
  std::vector<hcs_pair_out> o(b->npairs);
  int rc = hcs_batch_get(b, o.data());
  counts[0] = counts[1] = counts[2] = 0;
  for (auto& r : o) counts[0] += r.n_candidates, counts[1] += r.n_polygons, counts[2] += r.n_faces;
  return rc;
}

int hcs_batch_set_sensors(hcs_batch* b, int32_t nsensors, const int32_t* pair, const hcs_taxel_grid* grids) {
  if (!b || nsensors < 0 || (nsensors > 0 && (!pair || !grids))) return HCS_ERR_ARG;
  Ctx* c = b->c;
  int rc = hip_check(c, hipSetDevice(c->device), "hipSetDevice");
  if (rc) return rc;
  (void)hipStreamSynchronize(c->stream);
  batch_free_graph(b);
  if (nsensors == 0) {
    batch_taxels_free(b->tx);
  } else {
    rc = batch_taxels_setup(c, b->tx, b->npairs, b->geoms.data(), nsensors, pair, grids);
    if (!rc)
      rc = batch_taxels_reserve(c, b->tx, std::max<int64_t>(b->tx.poly_cap, 65536),
                                std::max<int64_t>(b->tx.list_cap, 4 * (int64_t)b->tx.ncell + 65536));
    if (rc) {
      const std::string msg = c->err;
      batch_taxels_free(b->tx);
      (void)batch_capture(b);
      return set_err(c, rc, msg);
    }
  }
  return batch_capture(b);
}

int hcs_batch_taxel_count(hcs_batch* b, int64_t* ntaxels) {
  if (!b || !ntaxels) return HCS_ERR_ARG;
  *ntaxels = b->tx.ncell;
  return HCS_OK;
}

int hcs_batch_get_taxels(hcs_batch* b, double* out) {
  if (!b || (b->tx.ncell > 0 && !out)) return HCS_ERR_ARG;
  Ctx* c = b->c;
  if (b->tx.n == 0) return set_err(c, HCS_ERR_STATE, "the batch has no sensors (hcs_batch_set_sensors)");
  int rc = batch_settle(b);
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(out, b->tx.out.ptr, sizeof(double) * b->tx.ncell, hipMemcpyDeviceToHost,
                                     c->stream),
                   "download taxels");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  return rc;
}

int hcs_batch_config(hcs_batch* b, int32_t* cfg) {
  if (!b || !cfg) return HCS_ERR_ARG;
  cfg[0] = b->blocks_per_cu, cfg[1] = b->qpl, cfg[2] = b->tet_max, cfg[3] = b->split_target, cfg[4] = b->queues;
  cfg[5] = b->order_mode;
  return HCS_OK;
}

int hcs_batch_set_config(hcs_batch* b, const int32_t* cfg) {
  if (!b || !cfg || cfg[0] < 1 || cfg[0] > 64 || cfg[1] < 1 || cfg[1] > 16 || cfg[2] < 0 || cfg[3] < 0 ||
      cfg[4] < 1 || cfg[4] > HCS_QUEUES || cfg[5] < 0 || cfg[5] > 2)
    return HCS_ERR_ARG;
  /* an item holds one query per lane of the candidate kernel's wave */
  if (cfg[1] != 1)
    return set_err(b->c, HCS_ERR_ARG, "queries_per_lane must be 1 (items of more than 64 queries are switched off)");
  b->blocks_per_cu = cfg[0], b->qpl = cfg[1], b->tet_max = cfg[2], b->split_target = cfg[3], b->queues = cfg[4];
  b->order_mode = cfg[5];
  int rc = batch_plan(b);
  if (!rc) rc = batch_capture(b);
  return rc;
}

/* The median of 4 reps + 3 back-to-back replays after a warm-up one (a
 * median, since single replays differ by a few percent). */
static int time_steps(hcs_batch* b, int reps, double* ms) {
  int rc = hcs_batch_step(b); /* warm-up */
  if (rc) return rc;
  const int n = 4 * reps + 3;
  std::vector<double> t(n);
  if ((rc = hcs_batch_time_steps(b, n, t.data()))) return rc;
  std::nth_element(t.begin(), t.begin() + n / 2, t.end());
  *ms = t[n / 2];
  return HCS_OK;
}

int hcs_batch_tune(hcs_batch* b, int32_t reps, double* best_ms) {
  if (!b || reps <= 0) return HCS_ERR_ARG;
  /* coordinate descent: (tet-query threshold, queues, blocks per CU) at the
   * current split target and work order, then the split target, then the
   * work order, each at the best so far with the best blocks per CU and its
   * neighbours */
  const int bpc[] = {2, 4, 8, 12};
  const int queues[] = {1, HCS_QUEUES};
  const int tets[] = {0, 64, 256};
  const int splits[] = {0, 512, 1024, 2048, 4096};
  double best = 1e300;
  int bb = b->blocks_per_cu, bq = b->queues, bt = b->tet_max, bs = b->split_target, bo = b->order_mode;
  auto try_plan = [&](const int* ks, int nk) -> int {
    int rc = batch_plan(b);
    if (rc) return rc;
    for (int i = 0; i < nk; ++i) {
      b->blocks_per_cu = ks[i];
      double ms = 0;
      rc = time_steps(b, reps, &ms);
      if (rc) return rc;
      if (ms < best)
        best = ms, bb = ks[i], bq = b->queues, bt = b->tet_max, bs = b->split_target, bo = b->order_mode;
    }
    return HCS_OK;
  };
  for (int t : tets)
    for (int q : queues) {
      b->tet_max = t, b->queues = q;
      int rc = try_plan(bpc, 4);
      if (rc) return rc;
    }
  const int s0 = b->split_target;
  b->tet_max = bt, b->queues = bq;
  std::vector<int> near;
  for (int i = 0; i < 4; ++i)
    if (bpc[i] == bb || (i > 0 && bpc[i - 1] == bb) || (i < 3 && bpc[i + 1] == bb)) near.push_back(bpc[i]);
  if (near.empty()) near.push_back(bb);
  for (int s : splits) {
    if (s == s0) continue;
    b->split_target = s;
    int rc = try_plan(near.data(), (int)near.size());
    if (rc) return rc;
  }
  b->split_target = bs;
  /* work orders by measured counts only: the order by measured cycles
   * depends on the timing of the measuring step, so its plan would differ
   * from the one timed here */
  for (int om = 0; om < 2; ++om) {
    if (om == bo) continue;
    b->order_mode = om;
    int rc = try_plan(near.data(), (int)near.size());
    if (rc) return rc;
  }
  b->blocks_per_cu = bb, b->queues = bq, b->tet_max = bt, b->split_target = bs, b->order_mode = bo;
  int rc = batch_plan(b);
  if (!rc) rc = batch_capture(b);
  if (best_ms) *best_ms = best;
  return rc;
}

int hcs_batch_item_stats(hcs_batch* b, int64_t* out, int64_t capacity, int64_t* nitems) {
  if (!b || !nitems || capacity < 0 || (capacity > 0 && !out)) return HCS_ERR_ARG;
  Ctx* c = b->c;
  const int n = (int)b->plan.items.size();
  *nitems = n;
  if (capacity < n) return capacity == 0 ? HCS_OK : set_err(c, HCS_ERR_CAPACITY, "item stats buffer too small");
  DBuf d;
  int rc = ensure(c, d, sizeof(unsigned long long) * 6 * std::max(1, n));
  if (rc) return rc;
  rc = hip_check(c, hipMemsetAsync(d.ptr, 0, sizeof(unsigned long long) * 6 * std::max(1, n), c->stream),
                 "memset");
  if (rc) return rc;
  QueryArgs a = batch_args(b);
  a.item_stats = (unsigned long long*)d.ptr;
  batch_enqueue(b, a);
  std::vector<unsigned long long> h(6 * (size_t)n);
  rc = hip_check(c, hipGetLastError(), "stats step");
  if (!rc) rc = hip_check(c, hipMemcpyAsync(h.data(), d.ptr, h.size() * 8, hipMemcpyDeviceToHost, c->stream), "dl");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  (void)hipFree(d.ptr);
  if (rc) return rc;
  for (int i = 0; i < n; ++i) {
    const Item& it = b->plan.items[i];
    int64_t* o = out + 10 * (size_t)i;
    const unsigned long long* q = &h[6 * (size_t)i];
    o[0] = it.env, o[1] = b->plan.envs[it.env].kind, o[2] = it.count;
    o[3] = (int64_t)q[0], o[4] = (int64_t)q[1], o[5] = (int64_t)q[2], o[6] = (int64_t)q[3];
    o[7] = b->plan.envs[it.env].nB;
    o[8] = (int64_t)q[4], o[9] = (int64_t)q[5];
  }
  return HCS_OK;
}

int hcs_batch_time_steps(hcs_batch* b, int32_t reps, double* ms) {
  if (!b || !ms || reps <= 0) return HCS_ERR_ARG;
  Ctx* c = b->c;
  int rc = batch_settle(b);
  if (rc) return rc;
  std::vector<hipEvent_t> ev(reps + 1, nullptr);
  for (auto& e : ev)
    if ((rc = hip_check(c, hipEventCreate(&e), "event"))) break;
  /* back-to-back replays, an event between consecutive ones */
  if (!rc) rc = hip_check(c, hipEventRecord(ev[0], c->stream), "record");
  for (int r = 0; r < reps && !rc; ++r) {
    rc = hcs_batch_step(b);
    if (!rc) rc = hip_check(c, hipEventRecord(ev[r + 1], c->stream), "record");
  }
  if (!rc) rc = hip_check(c, hipEventSynchronize(ev[reps]), "sync");
  for (int r = 0; r < reps && !rc; ++r) {
    float t = 0;
    rc = hip_check(c, hipEventElapsedTime(&t, ev[r], ev[r + 1]), "elapsed");
    ms[r] = t;
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

int hcs_batch_kernel_times(hcs_batch* b, int32_t reps, double* ms) {
  if (!b || !ms || reps <= 0) return HCS_ERR_ARG;
  Ctx* c = b->c;
  hipStream_t s = c->stream;
  hipEvent_t ev[5];
  for (auto& e : ev) {
    int rc = hip_check(c, hipEventCreate(&e), "event");
    if (rc) return rc;
  }
  double acc[3] = {0, 0, 0};
  int rc = batch_settle(b);
  for (int r = 0; r < reps && !rc; ++r) {
    QueryArgs a = batch_args(b);
    Pipe& P = b->pipe;
    if (b->tx.n > 0) launch_fill_u32(s, (unsigned*)b->tx.poly_cnt.ptr, 0u, 2);
    pipe_front(c, P, a, b->npairs);
    (void)hipEventRecord(ev[0], s);
    launch_cand(s, a, grid_of(b), CAND_STEP);
    (void)hipEventRecord(ev[1], s);
    launch_clip_reduce(c, P, a, b->npairs, clip_grid_for(c, P.cap), ev[2], ev[4]);
    batch_taxels_enqueue(c, b->tx, (const double*)b->state.ptr);
    (void)hipEventRecord(ev[3], s);
    rc = hip_check(c, hipEventSynchronize(ev[3]), "sync");
    float t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    if (!rc) rc = hip_check(c, hipEventElapsedTime(&t0, ev[0], ev[1]), "elapsed");
    if (!rc) rc = hip_check(c, hipEventElapsedTime(&t1, ev[2], ev[4]), "elapsed");
    if (!rc) rc = hip_check(c, hipEventElapsedTime(&t2, ev[1], ev[2]), "elapsed");
    if (!rc) rc = hip_check(c, hipEventElapsedTime(&t3, ev[4], ev[3]), "elapsed");
    acc[0] += t0, acc[1] += t1, acc[2] += t2 + t3;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 3; ++k) ms[k] = acc[k] / reps;
  return rc;
}

}  // extern "C"
