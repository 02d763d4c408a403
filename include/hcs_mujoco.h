/*
 * hcs_mujoco.h -- MuJoCo host glue over libhcs: the plugin's collision
 * callback and passive callback with the contact-surface work done by
 * hcs_contact_step (include/hcs.h).
 *
 * Replaces, in the reference (P/ = mujoco_contact_surfaces/):
 *   MujocoContactSurfacesPlugin::collision_cb   P/src/mujoco_contact_surfaces_plugin.cpp:255-318
 *   MujocoContactSurfacesPlugin::passiveCallback                              :411-523
 *   getGeomPose / getGeomVelocity                                             :107-126
 * MuJoCo's headers are not needed: the caller hands over the mjData /
 * mjModel arrays the plugin reads (geom_xpos, geom_xmat, geom_bodyid) and
 * two callbacks for the MuJoCo calls it makes (mj_objectVelocity,
 * mj_applyFT into qfrc_passive).  One hcs_mj per mjData, used from the
 * physics thread (the plugin keeps one instance per mjData,
 * plugin.cpp:88,225).
 */
#ifndef HCS_MUJOCO_H
#define HCS_MUJOCO_H
#include "hcs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hcs_mj hcs_mj;

/* What one step reads from mjModel / mjData. */
typedef struct hcs_mj_data {
  const double* geom_xpos;    /* d->geom_xpos, ngeom x 3 */
  const double* geom_xmat;    /* d->geom_xmat, ngeom x 9 (row-major) */
  const int32_t* geom_bodyid; /* m->geom_bodyid, ngeom */
  void* user;                 /* passed back to the callbacks (e.g. m and d) */
  /* mj_objectVelocity(m, d, mjOBJ_GEOM, geom, res, 0): res = [w; v] of the
   * geom origin in world axes (getGeomVelocity, plugin.cpp:107-114) */
  void (*object_velocity)(void* user, int32_t geom, double* res);
  /* mj_applyFT(m, d, force, torque, point, body, d->qfrc_passive) */
  void (*apply_ft)(void* user, const double* force, const double* torque, const double* point, int32_t body);
} hcs_mj_data;

/* opts: representation, ApplyContactSurfaceForces (plugin.cpp:599-611),
 * dt, keep_polygons (for the tactile sensors after hcs_mj_passive). */
int hcs_mj_create(hcs_ctx* ctx, int32_t ngeom, const hcs_step_opts* opts, hcs_mj** glue);
int hcs_mj_destroy(hcs_mj* glue);
/* A cs:: geom (parseMujocoCustomFields, plugin.cpp:613-808): MuJoCo geom id
 * -> mesh handle (hcs_add_soft_mesh, hcs_add_rigid_mesh or
 * hcs_add_rigid_plane). */
int hcs_mj_bind_geom(hcs_mj* glue, int32_t geom, int32_t handle);

/* collision_cb (plugin.cpp:255-318) for geoms g1, g2.  *handled = 0 when
 * the pair is not the plugin's (a geom without cs:: properties, or both
 * rigid): the caller runs MuJoCo's default collision function and uses its
 * result.  Otherwise the pair is queued for the next hcs_mj_passive (soft
 * geom first, as plugin.cpp:288-295 swaps) and *default_contacts says
 * whether MuJoCo's point contacts are still wanted: 0 when
 * ApplyContactSurfaceForces is on (the callback returns 0 contacts), 1
 * when it is off (plugin.cpp:277). */
int hcs_mj_collision(hcs_mj* glue, const hcs_mj_data* d, int32_t g1, int32_t g2, int32_t* handled,
                     int32_t* default_contacts);

/* passiveCallback (plugin.cpp:411-483): one contact step over the pairs
 * queued since the last call, then, with forces applied, per pair
 * apply_ft(F_A, tau_A, origin of A, body of A) and the same for B (one call
 * per body instead of one per contact point: mj_applyFT is linear in force,
 * torque and point).  The queue is cleared.  out (capacity entries, may be
 * NULL) receives the per-pair results and pairs_out[2 * k] = (g_a, g_b) the
 * MuJoCo geoms of pair k; *npairs = pairs stepped.  faces (may be NULL) as
 * in hcs_contact_step, for SurfacePlugin::update consumers
 * (P/include/mujoco_contact_surfaces/plugin_utils.h:94). */
int hcs_mj_passive(hcs_mj* glue, const hcs_mj_data* d, hcs_pair_out* out, int32_t* pairs_out, int32_t capacity,
                   int32_t* npairs, hcs_faces* faces);

#ifdef __cplusplus
}
#endif
#endif
