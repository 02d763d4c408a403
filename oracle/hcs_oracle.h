/*
 * hcs_oracle.h -- CPU FP64 restatement of the hydroelastic contact-surface path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product library (libhcs) never links or calls it.
 *
 * The reference path cannot be built here (it needs Drake v1.8.0, MuJoCo 2.x
 * and ROS1), so this file restates, from the reference sources, what
 *   mujoco_contact_surfaces/src/mujoco_contact_surfaces_plugin.cpp
 *   mujoco_contact_surface_sensors/src/flat_tactile_sensor.cpp
 *   mujoco_contact_surface_sensors/src/bvh.cpp
 * compute per step, plus the Drake geometry those files call.  Each function
 * in hcs_oracle.cpp cites the reference lines it follows.
 *
 * Arithmetic contract (shared with the HIP engine, stated here so the two
 * implementations can be compared bit for bit where the algorithm allows):
 *   - FP64 everywhere, no floating-point contraction (-ffp-contract=off);
 *   - 3-term sums and dot products evaluated as ((a0*b0 + a1*b1) + a2*b2);
 *   - a rigid transform X = [R row-major (9 doubles), t (3 doubles)], applied
 *     as x'_i = ((R_i0*x0 + R_i1*x1) + R_i2*x2) + t_i.
 */
#ifndef HCS_ORACLE_H
#define HCS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-tet record, ORC_TREC doubles:
 *  [0..11]  outward unit normals of faces {1,2,3},{0,3,2},{0,1,3},{0,2,1}
 *  [12..15] plane offsets d_k (signed distance = n_k.x + d_k, inside <= 0)
 *  [16..18] pressure gradient g, [19] c, so that p(x) = c + g.x
 *  [20..22] g normalised (g itself when |g| = 0)
 *  [23..25] AABB lo, [26..28] AABB hi, [29..31] unused */
#define ORC_TREC 32
/* Face record (one PointCollision), ORC_FREC doubles:
 *  p[3] (quadrature point, W), n[3] (out of N into M, W), fn0, stiffness,
 *  damping, area, p0, g */
#define ORC_FREC 12
/* Polygon record, ORC_PREC doubles:
 *  [0] vertex count nv (<= 8), [1..24] vertices (W), [25..32] vertex
 *  pressures, [33..35] centroid (W), [36] centroid pressure,
 *  [37..39] polygon normal (W, out of the rigid/N body, into the soft/M body) */
#define ORC_PREC 40

/* Pair parameters for soft-rigid, ORC_PARAMS doubles:
 *  [0] E_soft, [1] d_soft, [2] mu_dynamic_soft, [3] mu_dynamic_rigid,
 *  [4] dt (0.001 in the reference), [5] apply forces (0/1),
 *  [6] representation (0 = kTriangle, 1 = kPolygon), [7] soft is M (0/1) */
#define ORC_PARAMS 8

void orc_tet_precompute(const double* V, int nv, const int32_t* T, int nt, const double* p,
                        double* rec);
void orc_tri_normals(const double* V, int nv, const int32_t* F, int ntri, double* nrm);
void orc_compose_SR(const double* X_WS, const double* X_WR, double* X_SR);

/* Canonical candidate set: every (tet, tri) whose closed AABBs overlap in the
 * soft frame.  keys[i] = tet * ntri + tri, sorted ascending.  Returns |C|
 * (keys written up to cap).  brute = 1 uses the O(nt*ntri) loop, else a CPU
 * BVH that must give the same set. */
int64_t orc_candidates(const double* VS, int nvs, const int32_t* T, int nt,
                       const double* VR, int nvr, const int32_t* F, int ntri,
                       const double* X_WS, const double* X_WR, int brute,
                       int64_t* keys, int64_t cap);

/* Clip one triangle (3x3, soft frame) by one tet record.  Returns the vertex
 * count after near-duplicate removal (0 when < 3), vertices in out[8][3]. */
int orc_clip_tri_tet(const double* tri, const double* rec, double* out);

/* Combined material parameters. */
double orc_combined_modulus(double E_A, double E_B);
double orc_combined_dissipation(double E_A, double d_A, double E_B, double d_B);
double orc_combined_friction(double mu_A, double mu_B);

/* Force of one point collision on body A (the M side), plugin.cpp:435-475.
 *  pc = [p(3), n(3), fn0, k, damping], oA/oB = geom origins,
 *  VA/VB = [w(3), v(3)] at the geom origins.  f_out = force on A, returns fn. */
double orc_point_force(const double* pc, const double* oA, const double* VA, const double* oB,
                       const double* VB, double mu, double dt, int apply, double* f_out);

/* Soft-rigid contact step for one pair.
 *  pose = [X_WS(12), X_WR(12), V_WS(6), V_WR(6)]
 *  wrench (12) = F_S, tau_S about o_S, F_R, tau_R about o_R
 *  counts (3)  = |C|, polygons, faces (point collisions)
 *  faces/face_keys (tet, tri, k) and polys/poly_keys (tet, tri) optional. */
int orc_soft_rigid(const double* VS, int nvs, const int32_t* T, int nt, const double* pS,
                   const double* VR, int nvr, const int32_t* F, int ntri,
                   const double* pose, const double* params, int brute,
                   double* wrench, int64_t* counts,
                   double* faces, int64_t* face_keys, int64_t face_cap,
                   double* polys, int64_t* poly_keys, int64_t poly_cap);

/* The same split into load time (tet records + BVH, once per mesh) and
 * per step, for timing the per-step work alone. */
void* orc_soft_prepare(const double* VS, int nvs, const int32_t* T, int nt, const double* pS);
void orc_soft_free(void* prep);
int orc_soft_rigid_step(void* prep, const double* VR, int nvr, const int32_t* F, int ntri, const double* pose,
                        const double* params, double* wrench, int64_t* counts);

/* Flat tactile sensor sampling (flat_tactile_sensor.cpp:262-402), FP64.
 *  polys: npoly ORC_PREC records in canonical order.
 *  sensor = [X_W_sensor(12), xs, ys, zs, res]; window 0 none, 1 gauss,
 *  2 tukey, 3 square.  out[x + cx*y]. */
int orc_taxels(const double* polys, int64_t npoly, const double* sensor, int S, int window,
               double sigma, int cx, int cy, double* out);

/* TaxelSensor update over kTriangle fan triangles (12 doubles each: 3
 * vertices in W, 3 vertex pressures) with their surface index; method 0
 * closest, 1 weighted, 2 mean, 3 squared; sample_method 0 default, 1 area
 * importance; values in/out.  Returns the number of samples. */
int64_t orc_taxel_sensor(const double* tris, const int32_t* surf, int64_t ntri, const double* X_W_sensor,
                         const double* taxels, int ntaxels, double margin, double sample_resolution, int method,
                         int sample_method, int visualize, double* values);
void orc_minstd_uniform(int n, double* out);
/* CurvedSensor update over fan triangles; taxel lists in CSR form (ptr,
 * idx, w); every value written.  Returns the number of rays that hit. */
int orc_curved_sensor(const double* tris, int64_t ntri, const double* X_W_sensor, const double* pts,
                      const double* nrms, int m, const int32_t* ptr, const int32_t* idx, const double* w,
                      int ntaxels, double margin, double* values);

/* Soft-soft contact step (tet-tet equal-pressure planes).
 *  pose = [X_WA(12), X_WB(12), V_WA(6), V_WB(6)]
 *  params = [E_A, d_A, mu_A, E_B, d_B, mu_B, dt, apply, repr, A is M]
 *  face_keys (tetA, tetB, k), poly_keys (tetA, tetB). */
#define ORC_SS_PARAMS 10
int orc_soft_soft(const double* VA, int nva, const int32_t* TA, int nta, const double* pA,
                  const double* VB, int nvb, const int32_t* TB, int ntb, const double* pB,
                  const double* pose, const double* params, int brute,
                  double* wrench, int64_t* counts,
                  double* faces, int64_t* face_keys, int64_t face_cap,
                  double* polys, int64_t* poly_keys, int64_t poly_cap);

/* Soft body vs the rigid half-space z <= 0 of frame B (a MuJoCo plane
 * geom): params = [E_A, d_A, mu_A, mu_B, dt, apply, representation,
 * a_is_M]; outputs as orc_soft_rigid, keys (tet, 0). */
int orc_soft_plane(const double* VS, int nvs, const int32_t* T, int nt, const double* pS, const double* pose,
                   const double* params, double* wrench, int64_t* counts, double* faces, int64_t* face_keys,
                   int64_t face_cap, double* polys, int64_t* poly_keys, int64_t poly_cap);

#ifdef __cplusplus
}
#endif
#endif
