// Drives the MuJoCo host glue (include/hcs_mujoco.h) the way the plugin's
// callbacks are driven inside mj_step: collision_cb for every candidate geom
// pair, then passiveCallback, with a stand-in for the mjModel / mjData arrays
// the plugin reads and for mj_objectVelocity / mj_applyFT (test
// infrastructure; MuJoCo itself is not in this image).  Prints one JSON
// object: the scene, the handled flags, the per-pair results and, per body,
// the force and the torque about the world origin that the mj_applyFT calls
// added up to.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hcs.h"
#include "hcs_mujoco.h"

namespace {

struct World {
  std::vector<double> xpos, xmat, twist;  // per geom
  std::vector<int32_t> body;              // per geom
  std::vector<double> acc;                // per body: F(3), torque about the world origin(3)
  int calls = 0;
};

void velocity(void* u, int32_t g, double* res) {
  World* w = (World*)u;
  for (int k = 0; k < 6; ++k) res[k] = w->twist[6 * g + k];
}

void apply_ft(void* u, const double* f, const double* t, const double* p, int32_t b) {
  World* w = (World*)u;
  double* a = &w->acc[6 * b];
  const double pxf[3] = {p[1] * f[2] - p[2] * f[1], p[2] * f[0] - p[0] * f[2], p[0] * f[1] - p[1] * f[0]};
  for (int k = 0; k < 3; ++k) a[k] += f[k], a[3 + k] += t[k] + pxf[k];
  ++w->calls;
}

void rot(double ax, double ay, double az, double ang, double* R) {
  const double n = std::sqrt(ax * ax + ay * ay + az * az);
  ax /= n, ay /= n, az /= n;
  const double c = std::cos(ang), s = std::sin(ang), C = 1 - c;
  const double M[9] = {c + ax * ax * C,      ax * ay * C - az * s, ax * az * C + ay * s,
                       ay * ax * C + az * s, c + ay * ay * C,      ay * az * C - ax * s,
                       az * ax * C - ay * s, az * ay * C + ax * s, c + az * az * C};
  for (int k = 0; k < 9; ++k) R[k] = M[k];
}

int check(int rc, const char* what, hcs_ctx* ctx) {
  if (rc != HCS_OK) {
    std::fprintf(stderr, "%s failed: %d %s\n", what, rc, hcs_last_error(ctx));
    std::exit(1);
  }
  return rc;
}

int mesh(hcs_ctx* ctx, int type, int soft, const double* size, double hint, double E, const hcs_props& pr, int gid) {
  int32_t nv, ne, h;
  check(hcs_gen_geom_size(type, soft, size, hint, &nv, &ne), "gen size", ctx);
  std::vector<double> V(3 * (size_t)nv), p(nv);
  std::vector<int32_t> El((soft ? 4 : 3) * (size_t)ne);
  check(hcs_gen_geom(type, soft, size, hint, E, V.data(), El.data(), p.data()), "gen", ctx);
  if (soft) check(hcs_add_soft_mesh(ctx, V.data(), nv, El.data(), ne, p.data(), &pr, gid, &h), "add soft", ctx);
  else check(hcs_add_rigid_mesh(ctx, V.data(), nv, El.data(), ne, &pr, gid, &h), "add rigid", ctx);
  return h;
}

void arr(const char* name, const double* a, int n, bool comma = true) {
  std::printf("\"%s\": [", name);
  for (int i = 0; i < n; ++i) std::printf("%s%.17g", i ? ", " : "", a[i]);
  std::printf("]%s", comma ? ", " : "");
}

}  // namespace

int main(int argc, char** argv) {
  const int apply = argc > 1 ? std::atoi(argv[1]) : 1;
  hcs_ctx* ctx = nullptr;
  if (hcs_create(&ctx, 0) != HCS_OK) {
    std::fprintf(stderr, "hcs_create: %s\n", hcs_last_error(nullptr));
    return 1;
  }
  // geoms: 0 soft sphere (body 1), 1 rigid box (body 2), 2 plane (world body
  // 0), 3 a geom without cs:: properties (body 3), 4 soft box around its
  // medial axis (body 4)
  const int ngeom = 5, nbody = 5;
  const hcs_props soft1{1e5, 2.0, 0.01, 0.4, 0.35}, soft2{5e4, 1.0, 0.0, 0.3, 0.25};
  const hcs_props rigid{0.0, 1.0, 0.05, 0.5, 0.45};
  const double s0[3] = {0.05, 0, 0}, s1[3] = {0.2, 0.2, 0.1}, s4[3] = {0.03, 0.02, 0.02};
  const int h0 = mesh(ctx, HCS_GEOM_SPHERE, 1, s0, 0.01, soft1.hydroelastic_modulus, soft1, 0);
  const int h1 = mesh(ctx, HCS_GEOM_BOX, 0, s1, 0.05, 0.0, rigid, 1);
  int32_t h2;
  check(hcs_add_rigid_plane(ctx, &rigid, 2, &h2), "plane", ctx);
  const int h4 = mesh(ctx, HCS_GEOM_BOX, 1, s4, 0.0, soft2.hydroelastic_modulus, soft2, 4);
  World w;
  w.xpos.assign(3 * ngeom, 0.0), w.xmat.assign(9 * ngeom, 0.0), w.twist.assign(6 * ngeom, 0.0);
  w.body = {1, 2, 0, 3, 4};
  w.acc.assign(6 * nbody, 0.0);
  const double pos[ngeom][3] = {{0.01, -0.02, 0.046}, {0, 0, -0.1}, {0, 0, -0.3}, {0.5, 0.5, 0.5}, {0.02, 0.0, 0.1135}};
  const double ax[ngeom][4] = {{1, 2, 3, 0.05}, {0, 0, 1, 0.0}, {1, 0, 0, 0.0}, {0, 1, 0, 0.0}, {0.2, 1, 0.1, 0.12}};
  const double tw[ngeom][6] = {{0.3, -0.1, 0.2, 0.01, 0.02, -0.03},
                               {0.0, 0.0, 0.05, 0.0, 0.0, 0.001},
                               {0, 0, 0, 0, 0, 0},
                               {0, 0, 0, 0, 0, 0},
                               {-0.2, 0.1, 0.0, -0.01, 0.0, -0.02}};
  for (int g = 0; g < ngeom; ++g) {
    for (int k = 0; k < 3; ++k) w.xpos[3 * g + k] = pos[g][k];
    rot(ax[g][0], ax[g][1], ax[g][2], ax[g][3], &w.xmat[9 * g]);
    for (int k = 0; k < 6; ++k) w.twist[6 * g + k] = tw[g][k];
  }
  // the plane geom: a ceiling at z = 0.13 (its frame turned upside down, so
  // the rigid half-space z <= 0 of that frame is z >= 0.13), 5 mm into the
  // top of the soft box
  w.xpos[3 * 2 + 2] = 0.13;
  rot(1, 0, 0, 3.14159265358979323846, &w.xmat[9 * 2]);
  hcs_step_opts opts{0, apply, 0.001, 0, 0};
  hcs_mj* glue = nullptr;
  check(hcs_mj_create(ctx, ngeom, &opts, &glue), "mj_create", ctx);
  check(hcs_mj_bind_geom(glue, 0, h0), "bind", ctx);
  check(hcs_mj_bind_geom(glue, 1, h1), "bind", ctx);
  check(hcs_mj_bind_geom(glue, 2, h2), "bind", ctx);
  check(hcs_mj_bind_geom(glue, 4, h4), "bind", ctx);
  hcs_mj_data d{w.xpos.data(), w.xmat.data(), w.body.data(), &w, velocity, apply_ft};
  // mj_collision's broadphase hands these pairs to collision_cb
  const int cand[6][2] = {{1, 0}, {2, 4}, {1, 2}, {3, 0}, {0, 4}, {1, 4}};
  int handled[6], dflt[6];
  for (int k = 0; k < 6; ++k) check(hcs_mj_collision(glue, &d, cand[k][0], cand[k][1], &handled[k], &dflt[k]), "coll", ctx);
  hcs_pair_out out[8];
  int32_t pairs[16], n = 0;
  check(hcs_mj_passive(glue, &d, out, pairs, 8, &n, nullptr), "passive", ctx);
  int32_t n2 = -1;
  check(hcs_mj_passive(glue, &d, out + 0, pairs, 0, &n2, nullptr), "passive (empty queue)", ctx);
  std::printf("{\"apply\": %d, \"ngeom\": %d, ", apply, ngeom);
  arr("xpos", w.xpos.data(), 3 * ngeom);
  arr("xmat", w.xmat.data(), 9 * ngeom);
  arr("twist", w.twist.data(), 6 * ngeom);
  std::printf("\"body\": [1, 2, 0, 3, 4], \"candidates\": [[1, 0], [2, 4], [1, 2], [3, 0], [0, 4], [1, 4]], ");
  std::printf("\"handled\": [");
  for (int k = 0; k < 6; ++k) std::printf("%s%d", k ? ", " : "", handled[k]);
  std::printf("], \"default_contacts\": [");
  for (int k = 0; k < 6; ++k) std::printf("%s%d", k ? ", " : "", dflt[k]);
  std::printf("], \"npairs\": %d, \"npairs_after\": %d, \"apply_calls\": %d, \"pairs\": [", n, n2, w.calls);
  for (int k = 0; k < n; ++k) {
    std::printf("%s{\"geoms\": [%d, %d], ", k ? ", " : "", pairs[2 * k], pairs[2 * k + 1]);
    double W[12];
    for (int c = 0; c < 3; ++c)
      W[c] = out[k].F_A[c], W[3 + c] = out[k].tau_A[c], W[6 + c] = out[k].F_B[c], W[9 + c] = out[k].tau_B[c];
    arr("wrench", W, 12);
    std::printf("\"counts\": [%lld, %lld, %lld], \"status\": %lld}", (long long)out[k].n_candidates,
                (long long)out[k].n_polygons, (long long)out[k].n_faces, (long long)out[k].status);
  }
  std::printf("], ");
  arr("body_wrench", w.acc.data(), 6 * nbody, false);
  std::printf("}\n");
  hcs_mj_destroy(glue);
  hcs_destroy(ctx);
  return 0;
}
