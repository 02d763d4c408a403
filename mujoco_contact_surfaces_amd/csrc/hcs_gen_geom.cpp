/*
 * hcs_gen_geom.cpp -- hydroelastic meshes for the MuJoCo geom types the
 * plugin meshes while parsing the cs:: fields and again in onGeomChanged
 * (reference mujoco_contact_surfaces/src/mujoco_contact_surfaces_plugin.cpp:
 * 634-742 and 828-915), in place of the Drake generators it calls there:
 *
 *   sphere     MakeSphereVolumeMesh(kSingleInteriorVertex) +
 *              MakeSpherePressureField; MakeSphereSurfaceMesh      (:648-669)
 *   ellipsoid  MakeEllipsoidVolumeMesh(kSingleInteriorVertex) +
 *              MakeEllipsoidPressureField; MakeEllipsoidSurfaceMesh (:670-693)
 *   cylinder   MakeCylinderVolumeMeshWithMa + MakeCylinderPressureField;
 *              MakeCylinderSurfaceMesh                              (:694-716)
 *   box        MakeBoxVolumeMesh (hint > 0) or MakeBoxVolumeMeshWithMa
 *              (hint <= 0) + MakeBoxPressureField; MakeBoxSurfaceMesh (:717-742)
 *
 * Drake is not in the reference tree, so these are this library's own
 * meshes with the same structure: the pressure is E times the distance to
 * the boundary over its largest value (0 on the surface, E at the deepest
 * vertex), linear in every tet.
 *   - sphere / ellipsoid: an octahedron refined until its edges are no
 *     longer than the hint, projected onto the sphere (scaled to the
 *     ellipsoid), each boundary triangle fanned to the single interior
 *     vertex at the centre;
 *   - box with the medial axis: the box is the union of six frusta, each
 *     between a face and the matching face of the core box [-c, c] with
 *     c_i = h_i - min(h) (the set of points deepest inside); in its frustum
 *     the distance to the boundary is the distance to that face, a linear
 *     function, so the pressure is exact everywhere;
 *   - box on a grid: cells of spacing <= hint, six tets per cell;
 *   - cylinder with the medial axis: the axis segment |z| <= h - r (or the
 *     disk of radius r - h at z = 0 for a flat cylinder) splits the body
 *     into side and cap cells over n sectors of 2 pi / n.
 * Every convex cell is split into tets by a cone from its centroid over its
 * faces, each face fanned from its lowest-numbered vertex, so two cells
 * sharing a face split it alike and the mesh is conforming.  Rigid geoms
 * get the closed boundary surfaces (outward counter-clockwise).
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/hcs.h"

namespace {

const double kPi = 3.14159265358979323846;

struct Out {
  std::vector<double> V;  /* nv x 3 */
  std::vector<int32_t> E; /* nt x 4 (volume) or ntri x 3 (surface) */
  std::vector<double> p;  /* per vertex (volume) */
};

/* Vertices are merged on exact coordinates, so degenerate cells (a core
 * box or medial disk of zero extent) collapse onto shared vertices. */
struct Builder {
  Out o;
  std::map<std::tuple<double, double, double>, int> index;
  int vertex(double x, double y, double z, double pressure) {
    auto key = std::make_tuple(x + 0.0, y + 0.0, z + 0.0);
    auto it = index.find(key);
    if (it != index.end()) return it->second;
    const int id = (int)o.p.size();
    o.V.insert(o.V.end(), {x, y, z});
    o.p.push_back(pressure);
    index.emplace(key, id);
    return id;
  }
  const double* X(int i) const { return &o.V[3 * (size_t)i]; }
  double vol6(int a, int b, int c, int d) const {
    const double *A = X(a), *B = X(b), *C = X(c), *D = X(d);
    const double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    const double e3[3] = {D[0] - A[0], D[1] - A[1], D[2] - A[2]};
    return e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
           e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
  }
  /* A positively oriented tet; nothing when it has no volume. */
  void tet(int a, int b, int c, int d, double scale) {
    const double v = vol6(a, b, c, d);
    if (!(std::fabs(v) > 1e-12 * scale * scale * scale)) return;
    if (v < 0) std::swap(c, d);
    o.E.insert(o.E.end(), {a, b, c, d});
  }
  /* A convex cell given by its faces (vertex cycles): a cone from its
   * centroid (pressure phi at the centroid) over every face, each face fanned
   * from its lowest-numbered distinct vertex. */
  template <class Phi>
  void cell(const std::vector<std::vector<int>>& faces, Phi phi, double scale) {
    std::vector<int> all;
    for (auto& f : faces) all.insert(all.end(), f.begin(), f.end());
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    if (all.size() < 4) return;
    double c[3] = {0, 0, 0};
    for (int i : all)
      for (int k = 0; k < 3; ++k) c[k] += X(i)[k] / all.size();
    const int ctr = vertex(c[0], c[1], c[2], phi(c));
    for (auto f : faces) {
      /* drop repeated vertices of a collapsed edge */
      std::vector<int> g;
      for (size_t i = 0; i < f.size(); ++i)
        if (f[i] != f[(i + 1) % f.size()]) g.push_back(f[i]);
      if (g.size() < 3) continue;
      const size_t m = std::min_element(g.begin(), g.end()) - g.begin();
      for (size_t k = 1; k + 1 < g.size(); ++k)
        tet(ctr, g[m], g[(m + k) % g.size()], g[(m + k + 1) % g.size()], scale);
    }
  }
  void tri(int a, int b, int c) { o.E.insert(o.E.end(), {a, b, c}); }
};

/* Octahedron refined L times, projected to the unit sphere; faces ccw from
 * outside. */
void unit_octasphere(int L, std::vector<double>& v, std::vector<int>& f) {
  v = {1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};
  f = {0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0, 4, 2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3, 5};
  for (int l = 0; l < L; ++l) {
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int a, int b) {
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      double x = v[3 * a] + v[3 * b], y = v[3 * a + 1] + v[3 * b + 1], z = v[3 * a + 2] + v[3 * b + 2];
      const double n = std::sqrt(x * x + y * y + z * z);
      const int id = (int)v.size() / 3;
      v.insert(v.end(), {x / n, y / n, z / n});
      mid[key] = id;
      return id;
    };
    std::vector<int> nf;
    for (size_t i = 0; i < f.size(); i += 3) {
      const int a = f[i], b = f[i + 1], c = f[i + 2];
      const int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
    }
    f.swap(nf);
  }
}

/* The coarsest refinement (level <= 7) whose longest edge, scaled by
 * `radius`, is no longer than the hint. */
void octa_within(double radius, double hint, std::vector<double>& u, std::vector<int>& f) {
  for (int L = 0;; ++L) {
    unit_octasphere(L, u, f);
    double emax = 0.0;
    for (size_t i = 0; i < f.size(); i += 3)
      for (int k = 0; k < 3; ++k) {
        const double* a = &u[3 * (size_t)f[i + k]];
        const double* b = &u[3 * (size_t)f[i + (k + 1) % 3]];
        emax = std::max(emax, std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                                        (a[2] - b[2]) * (a[2] - b[2])));
      }
    if (radius * emax <= hint || L == 7) return;
  }
}

int ellipsoid(const double* s, double hint, double E, bool soft, Out& out) {
  const double a = s[0], b = s[1], c = s[2];
  if (!(a > 0 && b > 0 && c > 0 && hint > 0)) return HCS_ERR_ARG;
  std::vector<double> u;
  std::vector<int> f;
  octa_within(std::max(a, std::max(b, c)), hint, u, f);
  out = Out();
  const int off = soft ? 1 : 0;
  if (soft) out.V = {0, 0, 0}, out.p = {E};
  for (size_t i = 0; i < u.size(); i += 3) {
    out.V.insert(out.V.end(), {a * u[i], b * u[i + 1], c * u[i + 2]});
    if (soft) out.p.push_back(0.0);
  }
  for (size_t i = 0; i < f.size(); i += 3) {
    if (soft) out.E.insert(out.E.end(), {0, f[i] + off, f[i + 1] + off, f[i + 2] + off});
    else out.E.insert(out.E.end(), {f[i], f[i + 1], f[i + 2]});
  }
  return HCS_OK;
}

int box_grid(const double* h, double hint, double E, Out& out) {
  if (!(h[0] > 0 && h[1] > 0 && h[2] > 0 && hint > 0)) return HCS_ERR_ARG;
  int n[3];
  for (int k = 0; k < 3; ++k) {
    const double cells = std::ceil(2 * h[k] / hint);
    if (cells > 512) return HCS_ERR_ARG;
    n[k] = std::max(1, (int)cells);
  }
  if ((int64_t)n[0] * n[1] * n[2] > (1 << 22)) return HCS_ERR_ARG;
  const double hmin = std::min(h[0], std::min(h[1], h[2]));
  Builder B;
  std::vector<int> id((size_t)(n[0] + 1) * (n[1] + 1) * (n[2] + 1));
  auto at = [&](int i, int j, int k) -> int& { return id[((size_t)i * (n[1] + 1) + j) * (n[2] + 1) + k]; };
  for (int i = 0; i <= n[0]; ++i)
    for (int j = 0; j <= n[1]; ++j)
      for (int k = 0; k <= n[2]; ++k) {
        const double x = -h[0] + 2 * h[0] * i / n[0], y = -h[1] + 2 * h[1] * j / n[1], z = -h[2] + 2 * h[2] * k / n[2];
        const double phi = std::min(h[0] - std::fabs(x), std::min(h[1] - std::fabs(y), h[2] - std::fabs(z)));
        at(i, j, k) = B.vertex(x, y, z, E * std::max(0.0, phi) / hmin);
      }
  /* six tets per cell along its main diagonal (the same split in every
   * cell, so shared faces agree) */
  const int kuhn[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
  for (int i = 0; i < n[0]; ++i)
    for (int j = 0; j < n[1]; ++j)
      for (int k = 0; k < n[2]; ++k) {
        int c[8];
        for (int b = 0; b < 8; ++b) c[b] = at(i + (b >> 2 & 1), j + (b >> 1 & 1), k + (b & 1));
        for (auto& t : kuhn) B.tet(c[t[0]], c[t[1]], c[t[2]], c[t[3]], hmin);
      }
  out = B.o;
  return HCS_OK;
}

int box_medial(const double* h, double E, Out& out) {
  if (!(h[0] > 0 && h[1] > 0 && h[2] > 0)) return HCS_ERR_ARG;
  const double m = std::min(h[0], std::min(h[1], h[2]));
  const double c[3] = {h[0] - m, h[1] - m, h[2] - m};
  Builder B;
  /* box corners (pressure 0) and core corners (pressure E) */
  int corner[8], core[8];
  for (int b = 0; b < 8; ++b) {
    const double s[3] = {(b >> 2 & 1) ? 1.0 : -1.0, (b >> 1 & 1) ? 1.0 : -1.0, (b & 1) ? 1.0 : -1.0};
    corner[b] = B.vertex(s[0] * h[0], s[1] * h[1], s[2] * h[2], 0.0);
  }
  for (int b = 0; b < 8; ++b) {
    const double s[3] = {(b >> 2 & 1) ? 1.0 : -1.0, (b >> 1 & 1) ? 1.0 : -1.0, (b & 1) ? 1.0 : -1.0};
    core[b] = B.vertex(s[0] * c[0], s[1] * c[1], s[2] * c[2], E);
  }
  /* frustum of face (axis i, side s): its 4 box corners and the 4 core
   * corners with the same sign on axis i; the cycle (j, l) = (+,+), (-,+),
   * (-,-), (+,-) over the two other axes */
  for (int i = 0; i < 3; ++i)
    for (int sgn = 0; sgn < 2; ++sgn) {
      const int j = (i + 1) % 3, l = (i + 2) % 3;
      auto bit = [&](int si, int sj, int sl) { return (si << (2 - i)) | (sj << (2 - j)) | (sl << (2 - l)); };
      const int cyc[4][2] = {{1, 1}, {0, 1}, {0, 0}, {1, 0}};
      std::vector<int> F, C;
      for (auto& q : cyc) F.push_back(corner[bit(sgn, q[0], q[1])]), C.push_back(core[bit(sgn, q[0], q[1])]);
      std::vector<std::vector<int>> faces{F, C};
      for (int e = 0; e < 4; ++e) faces.push_back({F[e], F[(e + 1) % 4], C[(e + 1) % 4], C[e]});
      const double sg = sgn ? 1.0 : -1.0;
      B.cell(faces, [&](const double* x) { return E * (h[i] - sg * x[i]) / m; }, m);
    }
  out = B.o;
  return HCS_OK;
}

int cylinder_medial(const double* s, double hint, double E, Out& out) {
  const double r = s[0], h = s[1];
  if (!(r > 0 && h > 0 && hint > 0)) return HCS_ERR_ARG;
  const double sectors = std::ceil(2 * kPi * r / hint);
  if (sectors > 4096) return HCS_ERR_ARG;
  const int n = std::max(3, (int)sectors);
  const double pmax = std::min(r, h);
  auto phi = [&](const double* x) {
    const double rho = std::sqrt(x[0] * x[0] + x[1] * x[1]);
    return E * std::max(0.0, std::min(r - rho, h - std::fabs(x[2]))) / pmax;
  };
  Builder B;
  auto ring = [&](double rho, double z) {
    std::vector<int> v(n);
    for (int k = 0; k < n; ++k) {
      const double a = 2 * kPi * k / n, x[3] = {rho * std::cos(a), rho * std::sin(a), z};
      v[k] = B.vertex(x[0], x[1], x[2], rho >= r ? 0.0 : phi(x));
    }
    return v;
  };
  const double scale = pmax;
  auto P = [&](const double* x) { return phi(x); };
  const int ctop = B.vertex(0, 0, h, 0.0), cbot = B.vertex(0, 0, -h, 0.0);
  if (h > r) {
    /* axis segment |z| <= h - r in nz layers, lateral rings at the matching
     * heights stretched to |z| <= h */
    const int nz = std::max(1, (int)std::min(4096.0, std::ceil(2 * (h - r) / hint)));
    std::vector<int> axis(nz + 1);
    std::vector<std::vector<int>> lat(nz + 1);
    for (int q = 0; q <= nz; ++q) {
      const double za = -(h - r) + 2 * (h - r) * q / nz;
      const double x[3] = {0, 0, za};
      axis[q] = B.vertex(0, 0, za, phi(x));
      lat[q] = ring(r, -h + 2 * h * q / nz);
    }
    for (int k = 0; k < n; ++k) {
      const int k1 = (k + 1) % n;
      for (int q = 0; q < nz; ++q) {
        const int a0 = axis[q], a1 = axis[q + 1];
        const int l0 = lat[q][k], l1 = lat[q][k1], u0 = lat[q + 1][k], u1 = lat[q + 1][k1];
        B.cell({{a0, l0, l1}, {a1, u0, u1}, {a0, a1, u0, l0}, {a0, a1, u1, l1}, {l0, l1, u1, u0}}, P, scale);
      }
      B.tet(ctop, lat[nz][k], lat[nz][k1], axis[nz], scale);
      B.tet(cbot, lat[0][k], lat[0][k1], axis[0], scale);
    }
  } else {
    /* medial disk of radius r - h at z = 0 */
    const int cmid = B.vertex(0, 0, 0, E * std::min(r, h) / pmax);
    const std::vector<int> T = ring(r, h), Bt = ring(r, -h), M = ring(r - h, 0.0);
    for (int k = 0; k < n; ++k) {
      const int k1 = (k + 1) % n;
      B.cell({{cmid, M[k], M[k1]}, {ctop, T[k], T[k1]}, {cmid, ctop, T[k], M[k]}, {cmid, ctop, T[k1], M[k1]},
              {M[k], M[k1], T[k1], T[k]}},
             P, scale);
      B.cell({{cmid, M[k], M[k1]}, {cbot, Bt[k], Bt[k1]}, {cmid, cbot, Bt[k], M[k]}, {cmid, cbot, Bt[k1], M[k1]},
              {M[k], M[k1], Bt[k1], Bt[k]}},
             P, scale);
      B.cell({{M[k], T[k], Bt[k]}, {M[k1], T[k1], Bt[k1]}, {M[k], M[k1], T[k1], T[k]}, {M[k], M[k1], Bt[k1], Bt[k]},
              {T[k], T[k1], Bt[k1], Bt[k]}},
             P, scale);
    }
  }
  out = B.o;
  return HCS_OK;
}

int cylinder_surface(const double* s, double hint, Out& out) {
  const double r = s[0], h = s[1];
  if (!(r > 0 && h > 0)) return HCS_ERR_ARG;
  const double sectors = hint > 0 ? std::ceil(2 * kPi * r / hint) : 32.0;
  const double layers = hint > 0 ? std::ceil(2 * h / hint) : 1.0;
  if (sectors > 4096 || layers > 4096) return HCS_ERR_ARG;
  const int n = std::max(3, (int)sectors), nz = std::max(1, (int)layers);
  out = Out();
  auto push = [&](double x, double y, double z) {
    out.V.insert(out.V.end(), {x, y, z});
    return (int)out.V.size() / 3 - 1;
  };
  const int ct = push(0, 0, h), cb = push(0, 0, -h);
  std::vector<std::vector<int>> R(nz + 1, std::vector<int>(n));
  for (int q = 0; q <= nz; ++q)
    for (int k = 0; k < n; ++k) {
      const double a = 2 * kPi * k / n;
      R[q][k] = push(r * std::cos(a), r * std::sin(a), -h + 2 * h * q / nz);
    }
  for (int k = 0; k < n; ++k) {
    const int k1 = (k + 1) % n;
    out.E.insert(out.E.end(), {ct, R[nz][k], R[nz][k1]});
    out.E.insert(out.E.end(), {cb, R[0][k1], R[0][k]});
    for (int q = 0; q < nz; ++q) {
      out.E.insert(out.E.end(), {R[q][k], R[q][k1], R[q + 1][k1]});
      out.E.insert(out.E.end(), {R[q][k], R[q + 1][k1], R[q + 1][k]});
    }
  }
  return HCS_OK;
}

int box_surface(const double* h, double hint, Out& out) {
  if (!(h[0] > 0 && h[1] > 0 && h[2] > 0)) return HCS_ERR_ARG;
  const double cells = hint > 0 ? std::ceil(2 * std::max(h[0], std::max(h[1], h[2])) / hint) : 1.0;
  if (cells > 1024) return HCS_ERR_ARG;
  const int m = std::max(1, (int)cells);
  int32_t nv = 0, nt = 0;
  if (hcs_gen_rigid_size(0, m, &nv, &nt) != HCS_OK) return HCS_ERR_ARG;
  out = Out();
  out.V.resize(3 * (size_t)nv);
  out.E.resize(3 * (size_t)nt);
  return hcs_gen_rigid(0, m, h[0], h[1], h[2], out.V.data(), out.E.data());
}

int generate(int32_t type, int32_t soft, const double* size, double hint, double E, Out& out) {
  if (!size) return HCS_ERR_ARG;
  if (soft && !(E > 0 && std::isfinite(E))) return HCS_ERR_ARG;
  switch (type) {
    case HCS_GEOM_SPHERE: {
      const double s[3] = {size[0], size[0], size[0]};
      return ellipsoid(s, hint, E, soft != 0, out);
    }
    case HCS_GEOM_ELLIPSOID:
      return ellipsoid(size, hint, E, soft != 0, out);
    case HCS_GEOM_CYLINDER:
      return soft ? cylinder_medial(size, hint, E, out) : cylinder_surface(size, hint, out);
    case HCS_GEOM_BOX:
      if (!soft) return box_surface(size, hint, out);
      return hint > 0 ? box_grid(size, hint, E, out) : box_medial(size, E, out);
    default:
      /* plane, capsule and height field are not meshed (plugin.cpp:634-647) */
      return HCS_ERR_ARG;
  }
}

}  // namespace

extern "C" {

int hcs_gen_geom_size(int32_t type, int32_t soft, const double* size, double resolution_hint, int32_t* nv,
                      int32_t* nelem) {
  Out o;
  const int rc = generate(type, soft, size, resolution_hint, 1.0, o);
  if (rc) return rc;
  if (nv) *nv = (int32_t)(o.V.size() / 3);
  if (nelem) *nelem = (int32_t)(o.E.size() / (soft ? 4 : 3));
  return HCS_OK;
}

int hcs_gen_geom(int32_t type, int32_t soft, const double* size, double resolution_hint, double E, double* V,
                 int32_t* elements, double* pressure) {
  if (!V || !elements || (soft && !pressure)) return HCS_ERR_ARG;
  Out o;
  const int rc = generate(type, soft, size, resolution_hint, E, o);
  if (rc) return rc;
  std::copy(o.V.begin(), o.V.end(), V);
  std::copy(o.E.begin(), o.E.end(), elements);
  if (soft) std::copy(o.p.begin(), o.p.end(), pressure);
  return HCS_OK;
}

}  // extern "C"
