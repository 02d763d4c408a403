/*
 * hcs_mujoco.cpp -- MuJoCo host glue (include/hcs_mujoco.h): the plugin's
 * collision callback queues cs:: pairs, its passive callback steps them all
 * in one hcs_contact_step and applies one wrench per body.
 *
 * Reference: mujoco_contact_surfaces/src/mujoco_contact_surfaces_plugin.cpp
 *   collision_cb :255-318 (pair classification, soft-first swap, point
 *   contacts when forces are off), passiveCallback :411-483 (poses and
 *   velocities per pair, mj_applyFT on both bodies), getGeomPose :116-126,
 *   getGeomVelocity :107-114.
 */
#include <cstring>
#include <new>
#include <vector>

#include "../../include/hcs_mujoco.h"

struct hcs_mj {
  hcs_ctx* ctx = nullptr;
  int32_t ngeom = 0;
  hcs_step_opts opts{};
  std::vector<int32_t> handle; /* per geom: mesh handle, -1 when not a cs:: geom */
  std::vector<int32_t> kind;   /* per geom: 1 soft, 2 rigid mesh, 3 rigid half-space */
  std::vector<int32_t> queue;  /* (soft geom, other geom) per pair */
};

namespace {

/* getGeomPose (plugin.cpp:116-126): geom_xmat then geom_xpos. */
void geom_pose(const hcs_mj_data* d, int32_t g, double* X) {
  std::memcpy(X, d->geom_xmat + 9 * (size_t)g, 9 * sizeof(double));
  std::memcpy(X + 9, d->geom_xpos + 3 * (size_t)g, 3 * sizeof(double));
}

}  // namespace

extern "C" {

int hcs_mj_create(hcs_ctx* ctx, int32_t ngeom, const hcs_step_opts* opts, hcs_mj** out) {
  if (!ctx || !out || ngeom < 0) return HCS_ERR_ARG;
  hcs_mj* g = new (std::nothrow) hcs_mj();
  if (!g) return HCS_ERR_OOM;
  g->ctx = ctx;
  g->ngeom = ngeom;
  if (opts) {
    g->opts = *opts;
  } else {
    g->opts.representation = 0, g->opts.apply_forces = 1, g->opts.dt = 0.001;
  }
  g->handle.assign(ngeom, -1);
  g->kind.assign(ngeom, 0);
  *out = g;
  return HCS_OK;
}

int hcs_mj_destroy(hcs_mj* g) {
  if (!g) return HCS_ERR_ARG;
  delete g;
  return HCS_OK;
}

int hcs_mj_bind_geom(hcs_mj* g, int32_t geom, int32_t handle) {
  if (!g || geom < 0 || geom >= g->ngeom) return HCS_ERR_ARG;
  int64_t info[4];
  const int rc = hcs_mesh_info(g->ctx, handle, info);
  if (rc) return rc;
  g->handle[geom] = handle;
  g->kind[geom] = (int32_t)info[0];
  return HCS_OK;
}

int hcs_mj_collision(hcs_mj* g, const hcs_mj_data* d, int32_t g1, int32_t g2, int32_t* handled,
                     int32_t* default_contacts) {
  if (!g || !handled || !default_contacts || g1 < 0 || g2 < 0 || g1 >= g->ngeom || g2 >= g->ngeom) return HCS_ERR_ARG;
  (void)d;
  *handled = 0;
  *default_contacts = 1;
  /* not the plugin's: a geom without cs:: properties, or both rigid */
  if (g->handle[g1] < 0 || g->handle[g2] < 0 || (g->kind[g1] != 1 && g->kind[g2] != 1)) return HCS_OK;
  if (g->kind[g1] != 1) std::swap(g1, g2); /* soft first */
  g->queue.push_back(g1);
  g->queue.push_back(g2);
  *handled = 1;
  *default_contacts = g->opts.apply_forces ? 0 : 1;
  return HCS_OK;
}

int hcs_mj_passive(hcs_mj* g, const hcs_mj_data* d, hcs_pair_out* out, int32_t* pairs_out, int32_t capacity,
                   int32_t* npairs, hcs_faces* faces) {
  if (!g || !d || !d->geom_xpos || !d->geom_xmat || capacity < 0 || (capacity > 0 && (!out || !pairs_out)))
    return HCS_ERR_ARG;
  const int n = (int)g->queue.size() / 2;
  if (npairs) *npairs = n;
  if (n == 0) {
    if (faces) faces->count = 0;
    return HCS_OK;
  }
  if (g->opts.apply_forces && (!d->geom_bodyid || !d->apply_ft)) return HCS_ERR_ARG;
  std::vector<hcs_pair_in> in(n);
  for (int k = 0; k < n; ++k) {
    const int32_t ga = g->queue[2 * k], gb = g->queue[2 * k + 1];
    hcs_pair_in& p = in[k];
    std::memset(&p, 0, sizeof p);
    p.geom_a = g->handle[ga];
    p.geom_b = g->handle[gb];
    geom_pose(d, ga, p.X_WA);
    geom_pose(d, gb, p.X_WB);
    if (d->object_velocity) {
      d->object_velocity(d->user, ga, p.V_WA);
      d->object_velocity(d->user, gb, p.V_WB);
    }
  }
  std::vector<hcs_pair_out> res(n);
  const int rc = hcs_contact_step(g->ctx, n, in.data(), &g->opts, res.data(), faces);
  if (rc != HCS_OK && rc != HCS_ERR_CAPACITY) {
    g->queue.clear();
    return rc;
  }
  for (int k = 0; k < n; ++k) {
    const int32_t ga = g->queue[2 * k], gb = g->queue[2 * k + 1];
    if (g->opts.apply_forces) {
      d->apply_ft(d->user, res[k].F_A, res[k].tau_A, in[k].X_WA + 9, d->geom_bodyid[ga]);
      d->apply_ft(d->user, res[k].F_B, res[k].tau_B, in[k].X_WB + 9, d->geom_bodyid[gb]);
    }
    if (k < capacity) {
      out[k] = res[k];
      pairs_out[2 * k] = ga, pairs_out[2 * k + 1] = gb;
    }
  }
  g->queue.clear();
  return rc;
}

}  // extern "C"
