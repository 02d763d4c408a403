/*
 * hcs_gen.cpp -- synthetic hydroelastic geometry (host side, load time).
 *
 * Stands in for the Drake generators the plugin calls while parsing the
 * cs:: fields (plugin.cpp:648-742: Make<Shape>VolumeMesh,
 * Make<Shape>PressureField, Make<Shape>SurfaceMesh) and restates the
 * mesh-geom branch (plugin.cpp:743-793): one tet per surface triangle fanned
 * to the enclosed-volume centroid (CalcCentroidOfEnclosedVolume,
 * plugin.cpp:161-187), pressure 0 on the surface and E at the centroid.
 *
 * Soft bodies here are an icosphere of level L with K radial layers of
 * vertices around one centre vertex: the centre pyramids give 20*4^L tets and
 * each of the K-1 prism layers 3*20*4^L more.  Prisms are split with the
 * sorted-vertex rule (diagonal from the lower-indexed bottom vertex's top
 * copy), so neighbouring prisms agree on every shared quad and the mesh is
 * conforming.  Pressure is E (1 - j/K) on layer j, E at the centre.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/hcs.h"

namespace {

struct Ico {
  std::vector<double> v;  // unit vectors
  std::vector<int> f;     // ccw from outside
};

Ico icosphere_build(int L);

/* Subdivided icospheres are cached per level (generation is load time and
 * the batched workload asks for the same few levels thousands of times). */
const Ico& icosphere_cached(int L) {
  static std::mutex mu;
  static std::map<int, Ico> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(L);
  if (it == cache.end()) it = cache.emplace(L, icosphere_build(L)).first;
  return it->second;
}

Ico icosphere(int L) { return icosphere_cached(L); }

Ico icosphere_build(int L) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                              {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  const int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9},  {5, 11, 4},
                            {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6},  {3, 6, 8},
                            {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
  Ico m;
  for (auto& b : base) {
    double n = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    m.v.insert(m.v.end(), {b[0] / n, b[1] / n, b[2] / n});
  }
  for (auto& f : faces) m.f.insert(m.f.end(), {f[0], f[1], f[2]});
  for (int l = 0; l < L; ++l) {
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int a, int b) {
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      double x = m.v[3 * a] + m.v[3 * b], y = m.v[3 * a + 1] + m.v[3 * b + 1], z = m.v[3 * a + 2] + m.v[3 * b + 2];
      double n = std::sqrt(x * x + y * y + z * z);
      int id = (int)m.v.size() / 3;
      m.v.insert(m.v.end(), {x / n, y / n, z / n});
      mid[key] = id;
      return id;
    };
    std::vector<int> nf;
    nf.reserve(m.f.size() * 4);
    for (size_t i = 0; i < m.f.size(); i += 3) {
      int a = m.f[i], b = m.f[i + 1], c = m.f[i + 2];
      int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
    }
    m.f.swap(nf);
  }
  return m;
}

inline double signed_volume6(const double* a, const double* b, const double* c, const double* d) {
  double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double e3[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
  return e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
         e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
}

/* Maps a point r*u of layer fraction s = j/K (u unit) onto the shape. */
void place(int shape, double a, double b, double c, const double* u, double s, double* x) {
  if (shape == 1) {
    x[0] = a * s * u[0], x[1] = b * s * u[1], x[2] = c * s * u[2];
  } else if (shape == 2) {
    double k = std::max(-1.0, std::min(1.0, 4.0 * u[2]));
    x[0] = a * s * u[0], x[1] = a * s * u[1], x[2] = a * s * u[2] + b * s * k;
  } else {
    x[0] = a * s * u[0], x[1] = a * s * u[1], x[2] = a * s * u[2];
  }
}

}  // namespace

extern "C" {

int hcs_gen_soft_size(int32_t L, int32_t K, int32_t* nv, int32_t* nt) {
  if (L < 0 || L > 7 || K < 1 || K > 1024) return HCS_ERR_ARG;
  int64_t F = 20;
  for (int i = 0; i < L; ++i) F *= 4;
  int64_t V = F / 2 + 2;
  int64_t v = 1 + V * K, t = F * (1 + 3 * (int64_t)(K - 1));
  if (v > INT32_MAX || t > INT32_MAX) return HCS_ERR_ARG;
  if (nv) *nv = (int32_t)v;
  if (nt) *nt = (int32_t)t;
  return HCS_OK;
}

int hcs_gen_soft(int32_t L, int32_t K, int32_t shape, double a, double b, double c, double E, double* V, int32_t* T,
                 double* p) {
  int32_t nv, nt;
  if (hcs_gen_soft_size(L, K, &nv, &nt) != HCS_OK || shape < 0 || shape > 2 || !V || !T || !p) return HCS_ERR_ARG;
  Ico m = icosphere(L);
  const int NV = (int)m.v.size() / 3;
  /* vertex 0 = centre; layer j (1..K) vertex i -> 1 + (j-1)*NV + i */
  V[0] = V[1] = V[2] = 0;
  p[0] = E;
  for (int j = 1; j <= K; ++j) {
    double s = (double)j / K;
    for (int i = 0; i < NV; ++i) {
      int id = 1 + (j - 1) * NV + i;
      place(shape, a, b, c, &m.v[3 * i], s, V + 3 * id);
      p[id] = E * (1.0 - s);
    }
  }
  auto vid = [&](int j, int i) { return j == 0 ? 0 : 1 + (j - 1) * NV + i; };
  int64_t k = 0;
  auto push = [&](int q0, int q1, int q2, int q3) {
    if (signed_volume6(V + 3 * q0, V + 3 * q1, V + 3 * q2, V + 3 * q3) < 0) std::swap(q2, q3);
    T[4 * k] = q0, T[4 * k + 1] = q1, T[4 * k + 2] = q2, T[4 * k + 3] = q3;
    ++k;
  };
  const int NF = (int)m.f.size() / 3;
  for (int f = 0; f < NF; ++f) push(0, vid(1, m.f[3 * f]), vid(1, m.f[3 * f + 1]), vid(1, m.f[3 * f + 2]));
  for (int j = 1; j < K; ++j)
    for (int f = 0; f < NF; ++f) {
      int s[3] = {m.f[3 * f], m.f[3 * f + 1], m.f[3 * f + 2]};
      std::sort(s, s + 3);
      int a0 = vid(j, s[0]), b0 = vid(j, s[1]), c0 = vid(j, s[2]);
      int a1 = vid(j + 1, s[0]), b1 = vid(j + 1, s[1]), c1 = vid(j + 1, s[2]);
      push(a0, b0, c0, a1);
      push(b0, c0, a1, b1);
      push(c0, a1, b1, c1);
    }
  return k == nt ? HCS_OK : HCS_ERR_ARG;
}

int hcs_gen_rigid_size(int32_t kind, int32_t m, int32_t* nv, int32_t* ntri) {
  if (m < 0 || m > 4096) return HCS_ERR_ARG;
  int64_t v, t;
  if (kind == 0) {
    if (m < 1) return HCS_ERR_ARG;
    v = 6 * (int64_t)(m + 1) * (m + 1);
    t = 12 * (int64_t)m * m;
  } else if (kind == 1) {
    if (m < 1) return HCS_ERR_ARG;
    v = (int64_t)(m + 1) * (m + 1);
    t = 2 * (int64_t)m * m;
  } else if (kind == 2) {
    if (m > 7) return HCS_ERR_ARG;
    t = 20;
    for (int i = 0; i < m; ++i) t *= 4;
    v = t / 2 + 2;
  } else {
    return HCS_ERR_ARG;
  }
  if (v > INT32_MAX || t > INT32_MAX) return HCS_ERR_ARG;
  if (nv) *nv = (int32_t)v;
  if (ntri) *ntri = (int32_t)t;
  return HCS_OK;
}

int hcs_gen_rigid(int32_t kind, int32_t m, double a, double b, double c, double* V, int32_t* F) {
  int32_t nv, nt;
  if (hcs_gen_rigid_size(kind, m, &nv, &nt) != HCS_OK || !V || !F) return HCS_ERR_ARG;
  if (kind == 2) {
    Ico s = icosphere(m);
    for (size_t i = 0; i < s.v.size(); ++i) V[i] = a * s.v[i];
    for (size_t i = 0; i < s.f.size(); ++i) F[i] = s.f[i];
    return HCS_OK;
  }
  if (kind == 1) {
    for (int j = 0; j <= m; ++j)
      for (int i = 0; i <= m; ++i) {
        int id = j * (m + 1) + i;
        V[3 * id] = -a / 2 + a * i / m;
        V[3 * id + 1] = -a / 2 + a * j / m;
        V[3 * id + 2] = 0;
      }
    int64_t k = 0;
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        int v00 = j * (m + 1) + i, v10 = v00 + 1, v01 = v00 + m + 1, v11 = v01 + 1;
        F[3 * k] = v00, F[3 * k + 1] = v10, F[3 * k + 2] = v11, ++k;
        F[3 * k] = v00, F[3 * k + 1] = v11, F[3 * k + 2] = v01, ++k;
      }
    return HCS_OK;
  }
  /* box: 6 faces, each an (m+1)^2 vertex grid spanned by (u, v) with u x v
   * along the outward normal. */
  const double h[3] = {a, b, c};
  int64_t vi = 0, k = 0;
  for (int axis = 0; axis < 3; ++axis)
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      int ua = (axis + 1) % 3, va = (axis + 2) % 3;
      if (sgn < 0) std::swap(ua, va);
      int64_t base = vi;
      for (int j = 0; j <= m; ++j)
        for (int i = 0; i <= m; ++i) {
          double x[3];
          x[axis] = sgn * h[axis];
          x[ua] = -h[ua] + 2 * h[ua] * i / m;
          x[va] = -h[va] + 2 * h[va] * j / m;
          V[3 * vi] = x[0], V[3 * vi + 1] = x[1], V[3 * vi + 2] = x[2];
          ++vi;
        }
      for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
          int v00 = (int)base + j * (m + 1) + i, v10 = v00 + 1, v01 = v00 + m + 1, v11 = v01 + 1;
          F[3 * k] = v00, F[3 * k + 1] = v10, F[3 * k + 2] = v11, ++k;
          F[3 * k] = v00, F[3 * k + 1] = v11, F[3 * k + 2] = v01, ++k;
        }
    }
  return HCS_OK;
}

int hcs_gen_soft_from_surface(const double* V, int32_t nv, const int32_t* F, int32_t ntri, double E, double* V_out,
                              int32_t* T_out, double* p_out) {
  if (!V || !F || nv < 3 || ntri < 1 || !V_out || !T_out || !p_out) return HCS_ERR_ARG;
  /* CalcCentroidOfEnclosedVolume (plugin.cpp:161-187). */
  double six_total_volume = 0, cen[3] = {0, 0, 0};
  for (int f = 0; f < ntri; ++f) {
    const double *p = V + 3 * F[3 * f], *q = V + 3 * F[3 * f + 1], *r = V + 3 * F[3 * f + 2];
    double pq[3] = {p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]};
    double six_volume = pq[0] * r[0] + pq[1] * r[1] + pq[2] * r[2];
    six_total_volume += six_volume;
    for (int c = 0; c < 3; ++c) cen[c] += six_volume * (p[c] + q[c] + r[c]);
  }
  if (six_total_volume == 0) return HCS_ERR_ARG;
  for (int i = 0; i < 3 * nv; ++i) V_out[i] = V[i];
  for (int c = 0; c < 3; ++c) V_out[3 * nv + c] = cen[c] / (4 * six_total_volume);
  for (int i = 0; i < nv; ++i) p_out[i] = 0;
  p_out[nv] = E;
  for (int f = 0; f < ntri; ++f) {
    T_out[4 * f] = nv;
    T_out[4 * f + 1] = F[3 * f];
    T_out[4 * f + 2] = F[3 * f + 1];
    T_out[4 * f + 3] = F[3 * f + 2];
  }
  return HCS_OK;
}

}  // extern "C"
