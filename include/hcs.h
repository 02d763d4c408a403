/*
 * hcs.h -- C-ABI of the MI355X hydroelastic contact-surface engine (libhcs).
 *
 * Plain C, no exceptions cross it, every call returns an int status
 * (HCS_OK = 0, negative on error; hcs_last_error() gives the text).  The
 * caller owns host buffers, the context owns device memory.  A context is
 * used from one host thread at a time (the reference keys one plugin
 * instance per mjData, plugin.cpp:88,225).
 *
 * Each entry point names the reference interface it replaces; citations are
 * relative to the reference repository (P/ = mujoco_contact_surfaces/,
 * S/ = mujoco_contact_surface_sensors/, plugin.cpp =
 * P/src/mujoco_contact_surfaces_plugin.cpp).
 *
 * Conventions shared by every entry point:
 *   pose X (12 doubles)  = geom_xmat (row-major 3x3) followed by geom_xpos,
 *                          as read by getGeomPose (plugin.cpp:116-126);
 *   twist V (6 doubles)  = [w; v] of the geom origin in world axes, as
 *                          returned by mj_objectVelocity(..., flg_local = 0)
 *                          in getGeomVelocity (plugin.cpp:107-114);
 *   all arithmetic is FP64.
 */
#ifndef HCS_H
#define HCS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HCS_OK 0
#define HCS_ERR_ARG (-1)      /* invalid argument, handle or size */
#define HCS_ERR_HIP (-2)      /* HIP runtime error (text in hcs_last_error) */
#define HCS_ERR_OOM (-3)      /* device allocation failed */
#define HCS_ERR_CAPACITY (-4) /* an output buffer was too small; counts are still valid */
#define HCS_ERR_STATE (-5)    /* call order violated (e.g. taxels before a contact step) */
#define HCS_ERR_LIMIT (-6)    /* an internal limit was exceeded (e.g. BVH deeper than the traversal stack) */

typedef struct hcs_ctx hcs_ctx;
typedef struct hcs_batch hcs_batch;

/* The cs::<geom> 5-tuple read by parseMujocoCustomFields
 * (plugin.cpp:613-626).  hydroelastic_modulus > 0 is a soft geometry; 0 (or
 * +inf) is rigid, whose modulus is +inf and dissipation 1
 * (P/include/mujoco_contact_surfaces/mujoco_contact_surfaces_plugin.h:192-213). */
typedef struct hcs_props {
  double hydroelastic_modulus;
  double dissipation;
  double resolution_hint;
  double static_friction;
  double dynamic_friction;
} hcs_props;

/* One geometry pair of one step: what collision_cb receives for (g1, g2)
 * (plugin.cpp:255) plus the poses and twists passiveCallback reads
 * (plugin.cpp:435-438). */
typedef struct hcs_pair_in {
  int32_t geom_a, geom_b; /* mesh handles from hcs_add_*_mesh */
  int32_t reserved[2];
  double X_WA[12], X_WB[12];
  double V_WA[6], V_WB[6];
} hcs_pair_in;

/* Per-pair result: the net effect of the per-point mj_applyFT calls of
 * passiveCallback (plugin.cpp:477-482) expressed as one wrench per geometry
 * about its own origin (mj_applyFT is linear in force, torque and point). */
typedef struct hcs_pair_out {
  double F_A[3], tau_A[3];
  double F_B[3], tau_B[3];
  int64_t n_candidates; /* |C|: element pairs with overlapping AABBs */
  int64_t n_polygons;   /* non-empty contact polygons */
  int64_t n_faces;      /* PointCollision records (faces passing the filters of plugin.cpp:345,361-368) */
  int64_t status;       /* 0, or an HCS_ERR_* code for this pair */
} hcs_pair_out;

/* Step options: hydroelastic_contact_representation (plugin.h:251; 0 =
 * kTriangle, the reference default, 1 = kPolygon), cs::ApplyContactSurfaceForces
 * (plugin.cpp:599-611), the time constant hard-coded as 0.001 at
 * plugin.cpp:458, and whether to keep the contact polygons on the device for
 * hcs_taxels. */
typedef struct hcs_step_opts {
  int32_t representation;
  int32_t apply_forces;
  double dt;
  int32_t keep_polygons;
  int32_t reserved;
} hcs_step_opts;

/* Optional host copy of the PointCollision list built by
 * evaluateContactSurface (plugin.cpp:404; struct at
 * P/include/mujoco_contact_surfaces/common_types.h:48-56).
 * rec: capacity x 12 doubles = p[3], n[3], fn0, stiffness, damping, area, p0, g
 * key: capacity x 4 int64   = pair, element of A, element of B, fan index
 * Records come in no particular order; sort by key for a canonical order. */
typedef struct hcs_faces {
  int64_t capacity;
  int64_t count; /* out: faces produced (may exceed capacity -> HCS_ERR_CAPACITY) */
  double* rec;
  int64_t* key;
} hcs_faces;

/* Flat tactile sensor of FlatTactileSensor::load/bvh_update
 * (S/src/flat_tactile_sensor.cpp:127-214, 262-402).  The sensor samples the
 * contact polygons of every pair of the last contact step that involves mesh
 * `geom` (the reference's gc->g1 == id or gc->g2 == id test). */
typedef struct hcs_taxel_grid {
  int32_t geom;                /* mesh handle of the sensor geometry */
  int32_t sampling_resolution; /* S: S x S rays per taxel */
  int32_t window;              /* 0 none, 1 gauss, 2 tukey, 3 square */
  int32_t reserved;
  double sigma;
  double X_W_sensor[12];       /* sensor geom pose */
  double half_size[3];         /* geom_size: xs, ys, zs */
  double resolution;           /* taxel pitch */
} hcs_taxel_grid;

/* TaxelSensor (S/include/mujoco_contact_surface_sensors/taxel_sensor.h,
 * S/src/taxel_sensor.cpp:46-156): taxel positions in the sensor geometry's
 * frame (any mesh-geom offset of taxel_sensor.cpp:108-139 already applied),
 * the rosparam values include_margin, sample_resolution, method and
 * sample_method, and the visualize flag, which the reference's closest
 * method depends on. */
typedef struct hcs_taxel_sensor {
  int32_t geom;          /* mesh handle of the sensor geometry */
  int32_t ntaxels;
  int32_t method;        /* 0 closest, 1 weighted, 2 mean, 3 squared */
  int32_t sample_method; /* 0 default, 1 area_importance */
  int32_t visualize;
  int32_t reserved;
  double include_margin;
  double sample_resolution;
  double X_W_sensor[12]; /* sensor geom pose */
  const double* taxels;  /* ntaxels x 3, host memory */
} hcs_taxel_sensor;

/* CurvedSensor (S/src/curved_sensor.cpp:28-380): the surface samples of the
 * sensor geometry (points and normals in its frame) and, per taxel, the
 * samples it averages with their weights (CSR: taxel_ptr[ntaxels + 1],
 * sample_idx / weights[taxel_ptr[ntaxels]]), as the sensor's load step
 * builds them. */
typedef struct hcs_curved_sensor {
  int32_t geom;             /* mesh handle of the sensor geometry */
  int32_t ntaxels;
  int32_t nsamples;
  int32_t reserved;
  double include_margin;
  double X_W_sensor[12];    /* sensor geom pose */
  const double* points;     /* nsamples x 3, host memory */
  const double* normals;    /* nsamples x 3 */
  const int32_t* taxel_ptr; /* ntaxels + 1 */
  const int32_t* sample_idx;
  const double* weights;
} hcs_curved_sensor;

/* ---- lifetime (plugin constructor/load/destructor, plugin.cpp:191-245) */
int hcs_create(hcs_ctx** ctx, int device_ordinal);
/* Also destroys the context's remaining batches; their handles become
 * invalid. */
int hcs_destroy(hcs_ctx* ctx);
const char* hcs_last_error(const hcs_ctx* ctx); /* NULL ctx: last create error */
int hcs_synchronize(hcs_ctx* ctx);
/* The hipStream_t every launch of this context uses (default priority). */
void* hcs_stream(hcs_ctx* ctx);

/* ---- geometry (soft and rigid branches of parseMujocoCustomFields,
 * plugin.cpp:613-808).  Soft: vertices nv x 3 in the geom frame, positively
 * oriented tets nt x 4, per-vertex pressure (the linear field of
 * VolumeMeshFieldLinear).  Builds the tet records and the static LBVH on the
 * device.  Rigid: vertices nv x 3 and triangles ntri x 3 (outward
 * counter-clockwise).  geom_id orders M/N like Drake's GeometryId
 * (plugin.cpp:307-311): the geometry with the smaller id is M. */
int hcs_add_soft_mesh(hcs_ctx* ctx, const double* V, int32_t nv, const int32_t* T, int32_t nt,
                      const double* pressure, const hcs_props* props, int32_t geom_id, int32_t* handle);
int hcs_add_rigid_mesh(hcs_ctx* ctx, const double* V, int32_t nv, const int32_t* F, int32_t ntri,
                       const hcs_props* props, int32_t geom_id, int32_t* handle);
/* A MuJoCo plane geom: the rigid half-space z <= 0 of the geometry's frame,
 * for the soft-plane branch of collision_cb (plugin.cpp:296-299, Drake's
 * ComputeContactSurfaceFromSoftVolumeRigidHalfSpace).  Pairs (soft, plane)
 * slice the soft tets by the plane z = 0; face keys are (pair, tet, 0, k). */
int hcs_add_rigid_plane(hcs_ctx* ctx, const hcs_props* props, int32_t geom_id, int32_t* handle);
/* onGeomChanged (plugin.cpp:828-975): new geometry for an existing soft or
 * rigid mesh handle (the geom was resized or re-meshed), keeping its cs::
 * properties and geom_id.  Unlike the reference, which leaves the new
 * pressure field out of the geometry it installs (plugin.cpp:842-845), the
 * mesh, field and LBVH are replaced together once the new ones are built;
 * on error the handle keeps its old geometry.  Batches that step the
 * handle are planned again. */
int hcs_replace_soft_mesh(hcs_ctx* ctx, int32_t handle, const double* V, int32_t nv, const int32_t* T, int32_t nt,
                          const double* pressure);
int hcs_replace_rigid_mesh(hcs_ctx* ctx, int32_t handle, const double* V, int32_t nv, const int32_t* F,
                           int32_t ntri);
/* info = [kind (1 soft, 2 rigid mesh, 3 rigid half-space), nv, elements,
 * bvh height] */
int hcs_mesh_info(hcs_ctx* ctx, int32_t handle, int64_t* info);
/* Device tet records of a soft mesh, nt x 24 doubles (unit outward face
 * normals [12], plane offsets [4], pressure gradient [3], constant [1],
 * normalised gradient [3], unused [1]); for parity checks. */
int hcs_mesh_tet_records(hcs_ctx* ctx, int32_t handle, double* out);

/* ---- per-step contact (collision_cb + Drake query + evaluateContactSurface
 * + the passiveCallback force loop; plugin.cpp:255-483).  Soft-rigid and
 * soft-soft pairs; rigid-rigid pairs are rejected (the reference falls back
 * to MuJoCo there, plugin.cpp:271).  faces may be NULL. */
int hcs_contact_step(hcs_ctx* ctx, int32_t npairs, const hcs_pair_in* in, const hcs_step_opts* opts,
                     hcs_pair_out* out, hcs_faces* faces);

/* hcs_contact_step for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc); no counterpart in the reference.  Wrenches and counts only
 * (no faces; keep_polygons must be 0).  Every device buffer of the step is
 * a slice of one static 64 MiB array per device, so the call allocates and
 * frees nothing; pairs, work items and survivors must fit it together (at
 * most about 690k survivors of the exact rejections) or the call returns
 * HCS_ERR_CAPACITY.  The array is shared by every call on one device:
 * calls that overlap on one device race. */
int hcs_contact_step_noalloc(hcs_ctx* ctx, int32_t npairs, const hcs_pair_in* in, const hcs_step_opts* opts,
                             hcs_pair_out* out);

/* Canonical candidate keys of one pair for parity checks:
 * key = elemA * n_elements(B) + elemB, sorted ascending. */
int hcs_candidates(hcs_ctx* ctx, const hcs_pair_in* pair, int64_t* keys, int64_t capacity, int64_t* count);

/* ---- tactile sensors (FlatTactileSensor::bvh_update), on the polygons of
 * the last hcs_contact_step run with keep_polygons = 1.  out holds, per
 * sensor, cx * cy values at [x + cx * y] with cx = floor(2 xs / res + 0.1)
 * (S/src/flat_tactile_sensor.cpp:196-197). */
int hcs_taxel_dims(const hcs_taxel_grid* g, int32_t* cx, int32_t* cy);
int hcs_taxels(hcs_ctx* ctx, int32_t nsensors, const hcs_taxel_grid* g, double* out);
/* TaxelSensor::internal_update (S/src/taxel_sensor.cpp:158-478) on the
 * polygons of the last hcs_contact_step run with keep_polygons = 1.
 * values (ntaxels, in/out): taxels no sample reaches keep their value, as
 * the reference's message does; *nsamples (optional) = surface samples. */
int hcs_taxel_sensor_update(hcs_ctx* ctx, const hcs_taxel_sensor* sensor, double* values, int64_t* nsamples);
/* CurvedSensor::internal_update (S/src/curved_sensor.cpp:388-481) on the
 * polygons of the last hcs_contact_step run with keep_polygons = 1: one ray
 * per surface sample along its inward normal; values (ntaxels) all written. */
int hcs_curved_sensor_update(hcs_ctx* ctx, const hcs_curved_sensor* sensor, double* values);

/* ---- batched form (no reference counterpart: the reference runs one pair
 * at a time inside mj_collision).  Thousands of independent pairs, one
 * persistent work-queue launch per step, captured in a hipGraph. */
int hcs_batch_create(hcs_ctx* ctx, int32_t npairs, const int32_t* geoms /* npairs x 2 */,
                     const hcs_step_opts* opts, hcs_batch** batch);
int hcs_batch_destroy(hcs_batch* batch);
/* state: npairs x 36 doubles = X_WA, X_WB, V_WA, V_WB (host, copied to the device) */
int hcs_batch_set_state(hcs_batch* batch, const double* state);
/* One contact step for every pair (graph replay; asynchronous on hcs_stream). */
int hcs_batch_step(hcs_batch* batch);
int hcs_batch_get(hcs_batch* batch, hcs_pair_out* out);
/* Autotuner: times the launch shapes on the current state, keeps the best.
 * Re-planning (here and in hcs_batch_set_config) measures one step on the
 * current state and splits heavy work items over subtrees; results never
 * depend on the plan beyond the order of floating-point sums. */
int hcs_batch_tune(hcs_batch* batch, int32_t reps, double* best_ms);
/* Launch configuration: cfg = [persistent blocks per CU, queries per lane
 * (must be 1: larger values are rejected), largest rigid mesh handled in
 * tet-query mode, candidate count above which
 * a work item is split over LBVH subtrees (0: never), work queues of the
 * candidate kernel (1, or up to 8: one per XCD), work order of the items
 * (0: decreasing candidate count; 1: pairs by decreasing total count, each
 * pair's items together; 2: decreasing shader cycles of one measured
 * step)]; 6 int32. */
int hcs_batch_config(hcs_batch* batch, int32_t* cfg);
int hcs_batch_set_config(hcs_batch* batch, const int32_t* cfg);
/* Per-replay step times: reps graph replays back to back on hcs_stream with
 * a HIP event between consecutive ones; ms[r] = replay r in milliseconds. */
int hcs_batch_time_steps(hcs_batch* batch, int32_t reps, double* ms);
/* Eager per-kernel timing with HIP events on hcs_stream, averaged over
 * reps: ms[0] candidate kernel (traversal, cheap tests, compaction),
 * ms[1] clip/force kernel, ms[2] scan + per-pair reduction. */
int hcs_batch_kernel_times(hcs_batch* batch, int32_t reps, double* ms);
/* Per-item profile of one eager step: out (capacity x 10) = [pair, kind
 * (0 triangle queries, 1 tet queries, 2 soft-soft), queries, shader cycles,
 * production steps, candidates, cheap-test survivors, n elements of B,
 * wall-clock start, wall-clock end (constant-rate device counter; 0 for
 * items culled in that step)]; *nitems = number of items (capacity 0
 * queries it). */
int hcs_batch_item_stats(hcs_batch* batch, int64_t* out, int64_t capacity, int64_t* nitems);
/* Work-unit counts of the last step: [candidates, polygons, faces]. */
int hcs_batch_counts(hcs_batch* batch, int64_t* counts);
/* Flat tactile sensors inside the batch step (FlatTactileSensor::bvh_update,
 * S/src/flat_tactile_sensor.cpp:262-402, for every environment in the same
 * graph replay).  Sensor s belongs to pair pair[s]: as each environment is
 * its own world, it samples only that pair's contact polygons, and its pose
 * at every step is the pose in the batch state of the pair's geometry whose
 * mesh handle is grids[s].geom (B when both are; X_W_sensor is not read).
 * Sampling, windows and taxel order are those of hcs_taxels.  nsensors = 0
 * removes the sensors. */
int hcs_batch_set_sensors(hcs_batch* batch, int32_t nsensors, const int32_t* pair, const hcs_taxel_grid* grids);
/* Taxels of all sensors: sum over sensors of cx * cy. */
int hcs_batch_taxel_count(hcs_batch* batch, int64_t* ntaxels);
/* Taxels of the last step, sensor after sensor, index x + cx * y within a
 * sensor (grows the polygon and list buffers and re-runs the step when the
 * step overflowed them). */
int hcs_batch_get_taxels(hcs_batch* batch, double* out);

/* ---- differentiation (no reference counterpart).  Tangent-linear
 * recomputation of one step at fixed clipping topology:
 * J (npairs x 12 x 7): d(F_A, tau_A, F_B, tau_B) / d(xi), xi = [dtheta (3),
 * dt (3) of the pose of geometry B, applied as X_WB <- (exp([dtheta]) R_WB,
 * t_WB + dt), and dE of geometry A's modulus (pressures scale with E)].
 * Soft-rigid pairs take A soft and B rigid. */
int hcs_contact_jacobian(hcs_ctx* ctx, int32_t npairs, const hcs_pair_in* in, const hcs_step_opts* opts,
                         double* J);
/* Vector-Jacobian product: dL_dxi (npairs x 7) = J^T dL_dout (npairs x 12). */
int hcs_contact_backward(hcs_ctx* ctx, int32_t npairs, const hcs_pair_in* in, const hcs_step_opts* opts,
                         const double* dL_dout, double* dL_dxi);
/* Taxel Jacobian for one sensor attached to geometry B of pair `pair`
 * (the sensor pose moves with B): Jt (cx*cy x 7). */
int hcs_taxel_jacobian(hcs_ctx* ctx, const hcs_pair_in* pair, const hcs_step_opts* opts, const hcs_taxel_grid* g,
                       double* Jt);

/* ---- synthetic geometry (host side; stands in for the Drake mesh and
 * pressure-field generators called at plugin.cpp:648-796).
 * Soft bodies: icosphere of level L, K radial layers around one centre
 * vertex; pressure E (1 - layer/K), E at the centre.
 * shape 0 sphere (a = radius), 1 ellipsoid (a, b, c semi-axes),
 * 2 capsule (a = radius, b = half length of the straight part). */
int hcs_gen_soft_size(int32_t L, int32_t K, int32_t* nv, int32_t* nt);
int hcs_gen_soft(int32_t L, int32_t K, int32_t shape, double a, double b, double c, double E, double* V,
                 int32_t* T, double* p);
/* Rigid surfaces: kind 0 box (a, b, c half sizes, m x m squares per face,
 * 12 m^2 triangles), 1 plane grid z = 0 of side a with m x m squares
 * (2 m^2 triangles, fixed diagonal), 2 icosphere of radius a, level m. */
int hcs_gen_rigid_size(int32_t kind, int32_t m, int32_t* nv, int32_t* ntri);
int hcs_gen_rigid(int32_t kind, int32_t m, double a, double b, double c, double* V, int32_t* F);
/* Meshes of the MuJoCo geom types the plugin meshes for its cs:: geoms
 * (parseMujocoCustomFields and onGeomChanged, plugin.cpp:634-742,
 * 828-915), in place of the Drake generators called there.  type is the
 * mjtGeom value; size is geom_size (sphere radius; ellipsoid semi-axes;
 * cylinder radius and half length; box half sizes); resolution_hint is the
 * cs:: value.  soft = 1: a tet mesh (elements nelem x 4) with pressure 0 on
 * the surface and E at the deepest vertex, linear in every tet (sphere and
 * ellipsoid: a single interior vertex; cylinder: cells around the medial
 * axis; box: a grid of spacing <= hint when hint > 0, else cells around the
 * medial axis); soft = 0: the closed surface (elements nelem x 3).  Plane,
 * capsule and height field are not meshed (HCS_ERR_ARG), as in the
 * reference; a soft cylinder needs hint > 0. */
#define HCS_GEOM_PLANE 0
#define HCS_GEOM_SPHERE 2
#define HCS_GEOM_CAPSULE 3
#define HCS_GEOM_ELLIPSOID 4
#define HCS_GEOM_CYLINDER 5
#define HCS_GEOM_BOX 6
#define HCS_GEOM_MESH 7
int hcs_gen_geom_size(int32_t type, int32_t soft, const double* size, double resolution_hint, int32_t* nv,
                      int32_t* nelem);
int hcs_gen_geom(int32_t type, int32_t soft, const double* size, double resolution_hint, double E, double* V,
                 int32_t* elements, double* pressure);
/* Soft body from a closed surface: one tet per triangle fanned to the
 * enclosed-volume centroid, pressure 0 on the surface and E at the centroid
 * (plugin.cpp:743-793 with CalcCentroidOfEnclosedVolume, plugin.cpp:161-187).
 * V_out: (nv + 1) x 3, T_out: ntri x 4, p_out: nv + 1. */
int hcs_gen_soft_from_surface(const double* V, int32_t nv, const int32_t* F, int32_t ntri, double E, double* V_out,
                              int32_t* T_out, double* p_out);

#ifdef __cplusplus
}
#endif
#endif
