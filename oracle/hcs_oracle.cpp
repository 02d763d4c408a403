/*
 * hcs_oracle.cpp -- CPU FP64 restatement of the hydroelastic contact path.
 *
 * TEST INFRASTRUCTURE ONLY (see hcs_oracle.h).  Built with
 * g++ -O2 -ffp-contract=off; never linked into the product library.
 *
 * Reference citations are relative to the reference repository root:
 *   P/ = mujoco_contact_surfaces/, S/ = mujoco_contact_surface_sensors/,
 *   plugin.cpp = P/src/mujoco_contact_surfaces_plugin.cpp.
 * Where the reference delegates to Drake v1.8.0 (not vendored, not buildable
 * here), the Drake semantics restated below are marked "Drake semantics";
 * they are recalled from Drake's published source and are parity-unpinned
 * against Drake itself (only against closed forms and invariants).
 */
#include "hcs_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace {

const double kInf = std::numeric_limits<double>::infinity();
/* Drake semantics: the tet-tri and tet-tet gradient filters accept a face
 * when cos(angle) > cos(5*pi/8). */
const double kCosAlpha = -0.38268343236508978;
/* Drake semantics: vertices closer than 1e-14 are near-duplicates. */
const double kEps2 = 1e-14 * 1e-14;
/* Tet face vertex table; with positively oriented tets the cross product
 * (B-A)x(C-A) of each face points out of the tet. */
const int kFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

inline double dot3(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
inline void sub3(const double* a, const double* b, double* c) {
  c[0] = a[0] - b[0];
  c[1] = a[1] - b[1];
  c[2] = a[2] - b[2];
}
/* x' = R x + t with R row-major. */
inline void xform(const double* X, const double* x, double* y) {
  for (int i = 0; i < 3; ++i) y[i] = ((X[3 * i] * x[0] + X[3 * i + 1] * x[1]) + X[3 * i + 2] * x[2]) + X[9 + i];
}
inline void rot(const double* X, const double* x, double* y) {
  for (int i = 0; i < 3; ++i) y[i] = (X[3 * i] * x[0] + X[3 * i + 1] * x[1]) + X[3 * i + 2] * x[2];
}

/* Neumaier-compensated accumulator for the wrench reference sums. */
struct KSum {
  double s = 0, c = 0;
  void add(double x) {
    double t = s + x;
    if (std::fabs(s) >= std::fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  double get() const { return s + c; }
};

struct Wrench {
  KSum F[3], tA[3], tB[3];
};

/* ---------------------------------------------------------------- meshes */

void tet_record(const double* V, const int32_t* tet, const double* p, double* r) {
  const double* v[4] = {V + 3 * tet[0], V + 3 * tet[1], V + 3 * tet[2], V + 3 * tet[3]};
  for (int k = 0; k < 4; ++k) {
    const double* A = v[kFace[k][0]];
    const double* B = v[kFace[k][1]];
    const double* C = v[kFace[k][2]];
    double e1[3], e2[3], n[3];
    sub3(B, A, e1);
    sub3(C, A, e2);
    cross3(e1, e2, n);
    double len = std::sqrt(dot3(n, n));
    double nh[3] = {n[0] / len, n[1] / len, n[2] / len};
    r[3 * k + 0] = nh[0];
    r[3 * k + 1] = nh[1];
    r[3 * k + 2] = nh[2];
    r[12 + k] = -dot3(nh, A);
  }
  /* Linear field through the 4 vertex values: g solves
   * (v_i - v_0).g = p_i - p_0 (reciprocal-basis form), c = p_0 - g.v_0. */
  double e1[3], e2[3], e3[3];
  sub3(v[1], v[0], e1);
  sub3(v[2], v[0], e2);
  sub3(v[3], v[0], e3);
  double c23[3], c31[3], c12[3];
  cross3(e2, e3, c23);
  cross3(e3, e1, c31);
  cross3(e1, e2, c12);
  double det = dot3(e1, c23);
  double dp1 = p[tet[1]] - p[tet[0]], dp2 = p[tet[2]] - p[tet[0]], dp3 = p[tet[3]] - p[tet[0]];
  double g[3];
  for (int i = 0; i < 3; ++i) g[i] = ((dp1 * c23[i] + dp2 * c31[i]) + dp3 * c12[i]) / det;
  r[16] = g[0];
  r[17] = g[1];
  r[18] = g[2];
  r[19] = p[tet[0]] - dot3(g, v[0]);
  double gg = dot3(g, g);
  if (gg > 0) {
    double gl = std::sqrt(gg);
    r[20] = g[0] / gl;
    r[21] = g[1] / gl;
    r[22] = g[2] / gl;
  } else {
    r[20] = g[0];
    r[21] = g[1];
    r[22] = g[2];
  }
  for (int i = 0; i < 3; ++i) {
    double lo = v[0][i], hi = v[0][i];
    for (int j = 1; j < 4; ++j) {
      lo = std::min(lo, v[j][i]);
      hi = std::max(hi, v[j][i]);
    }
    r[23 + i] = lo;
    r[26 + i] = hi;
  }
  r[29] = r[30] = r[31] = 0;
}

inline double field(const double* rec, const double* x) { return rec[19] + dot3(rec + 16, x); }

/* ------------------------------------------------------------ clipping */

/* Drake semantics (ClipPolygonByHalfSpace): walk current/previous vertex
 * pairs starting with (v0, v_last); a vertex is inside when its signed
 * distance is <= 0; the crossing point is wa*cur + (1-wa)*prev with
 * wa = s_prev / (s_prev - s_cur). */
int clip_halfspace(const double (*in)[3], int n, const double* nh, double off, double (*out)[3], int cap) {
  double s[9];
  for (int i = 0; i < n; ++i) s[i] = dot3(nh, in[i]) + off;
  int m = 0;
  for (int i = 0; i < n; ++i) {
    int ip = (i == 0) ? n - 1 : i - 1;
    bool cin = s[i] <= 0, pin = s[ip] <= 0;
    if (cin != pin && m < cap) {
      double wa = s[ip] / (s[ip] - s[i]);
      double wb = 1.0 - wa;
      for (int c = 0; c < 3; ++c) out[m][c] = wa * in[i][c] + wb * in[ip][c];
      ++m;
    }
    if (cin && m < cap) {
      for (int c = 0; c < 3; ++c) out[m][c] = in[i][c];
      ++m;
    }
  }
  return m;
}

/* Drake semantics (RemoveNearlyDuplicateVertices): drop a vertex within
 * 1e-14 of the previously kept one, then the last if it is within 1e-14 of
 * the first; fewer than 3 vertices is no polygon. */
int dedupe(double (*poly)[3], int n) {
  if (n <= 0) return 0;
  auto near = [](const double* a, const double* b) {
    double d[3];
    sub3(a, b, d);
    return dot3(d, d) < kEps2;
  };
  int m = 1;
  for (int i = 1; i < n; ++i) {
    if (!near(poly[m - 1], poly[i])) {
      if (m != i)
        for (int c = 0; c < 3; ++c) poly[m][c] = poly[i][c];
      ++m;
    }
  }
  if (m >= 3 && near(poly[m - 1], poly[0])) --m;
  return m >= 3 ? m : 0;
}

/* Clip a convex polygon of at most n0 (3 or 4) vertices by the 4 tet half-spaces
 * (unit normals + offsets); stage k keeps at most n0 + k + 1 vertices (the
 * count a convex polygon can reach; rounding surplus is dropped).  Result
 * after dedupe. */
int clip_by_planes(double (*poly)[3], int n, const double* nrm, const double* off, int n0) {
  double buf[9][3];
  double(*a)[3] = poly;
  double(*b)[3] = buf;
  for (int k = 0; k < 4 && n > 0; ++k) {
    n = clip_halfspace(a, n, nrm + 3 * k, off[k], b, n0 + k + 1);
    std::swap(a, b);
  }
  if (a != poly)
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) poly[i][c] = a[i][c];
  return dedupe(poly, n);
}

/* Drake semantics (CalcPolygonCentroid): area-weighted centroid of the fan
 * from v0, weights (v_i-v0)x(v_{i+1}-v0).n; the vertex mean when the
 * weights sum to zero.  Returns the weight sum (twice the area). */
double polygon_centroid(const double (*v)[3], int n, const double* nrm, double* C) {
  if (n == 3) {
    for (int c = 0; c < 3; ++c) C[c] = ((v[0][c] + v[1][c]) + v[2][c]) / 3.0;
    double e1[3], e2[3], cr[3];
    sub3(v[1], v[0], e1);
    sub3(v[2], v[0], e2);
    cross3(e1, e2, cr);
    return dot3(cr, nrm);
  }
  double W = 0, S[3] = {0, 0, 0};
  for (int i = 1; i + 1 < n; ++i) {
    double e1[3], e2[3], cr[3];
    sub3(v[i], v[0], e1);
    sub3(v[i + 1], v[0], e2);
    cross3(e1, e2, cr);
    double w = dot3(cr, nrm);
    W += w;
    for (int c = 0; c < 3; ++c) S[c] += w * ((v[0][c] + v[i][c]) + v[i + 1][c]);
  }
  if (W == 0) {
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int i = 0; i < n; ++i) s += v[i][c];
      C[c] = s / n;
    }
  } else {
    for (int c = 0; c < 3; ++c) C[c] = S[c] / (3.0 * W);
  }
  return W;
}

/* ------------------------------------------------------------ force law */

struct PairCtx {
  const double *X_WA, *X_WB, *V_WA, *V_WB;
  double mu, dstar, dt;
  int apply, repr, a_is_M;
};

/* plugin.cpp:435-475 for one point collision; f_out is the force on A. */
double point_force(const double* p, const double* n, double fn0, double k, double damping, const double* oA,
                   const double* VA, const double* oB, const double* VB, double mu, double dt, int apply,
                   double* f_out) {
  const double stiction_tolerance = 1.0e-4;
  const double relative_tolerance = 1.0e-2;
  double rA[3], rB[3], wA[3], wB[3], vA[3], vB[3], v[3];
  sub3(p, oA, rA);
  sub3(p, oB, rB);
  cross3(VA, rA, wA);
  cross3(VB, rB, wB);
  for (int c = 0; c < 3; ++c) {
    vA[c] = VA[3 + c] + wA[c];
    vB[c] = VB[3 + c] + wB[c];
    v[c] = vA[c] - vB[c];
  }
  double vn = dot3(v, n);
  double fn = std::max(0., 1. - damping * vn) * (fn0 - dt * k * vn);
  if (!apply) {
    f_out[0] = f_out[1] = f_out[2] = 0;
    return fn;
  }
  double vt[3];
  for (int c = 0; c < 3; ++c) vt[c] = v[c] - n[c] * vn;
  double eps = stiction_tolerance * relative_tolerance;
  eps = eps * eps;
  double v_slip = std::sqrt(dot3(vt, vt) + eps);
  double mu_reg = mu;
  double s = v_slip / stiction_tolerance;
  if (s < 1) mu_reg = mu * s * (2.0 - s);
  for (int c = 0; c < 3; ++c) {
    double that = vt[c] / v_slip;
    double f_slip = -mu_reg * that * fn;
    f_out[c] = f_slip + fn * n[c];
  }
  return fn;
}

/* plugin.cpp:320-409 (evaluateContactSurface) for one face followed by the
 * force law; accumulates the wrench; returns 1 when a point collision is
 * emitted.  nW is the polygon normal out of B into A; gAW, gBW the field
 * gradients in W (gBW unused when B is rigid). */
int eval_face(const PairCtx& pc, double area, const double* cW, double p0, const double* nW, const double* gAW,
              const double* gBW, bool b_soft, Wrench& w, double* frec) {
  if (!(area > 1.0e-14)) return 0;
  /* The face normal points out of N into M: n = nW when A is M. */
  double n[3];
  const double *gM, *gN;
  bool M_soft, N_soft;
  if (pc.a_is_M) {
    n[0] = nW[0], n[1] = nW[1], n[2] = nW[2];
    gM = gAW, gN = gBW, M_soft = true, N_soft = b_soft;
  } else {
    n[0] = -nW[0], n[1] = -nW[1], n[2] = -nW[2];
    gM = gBW, gN = gAW, M_soft = b_soft, N_soft = true;
  }
  const double gMv = M_soft ? dot3(gM, n) : kInf;
  const double gNv = N_soft ? -dot3(gN, n) : kInf;
  const double kGradientEpsilon = 1.0e-14;
  if (gMv < kGradientEpsilon || gNv < kGradientEpsilon) return 0;
  const double g = 1.0 / (1.0 / gMv + 1.0 / gNv);
  const double fn0 = area * p0;
  const double k = area * g;
  const double* oM = pc.a_is_M ? pc.X_WA + 9 : pc.X_WB + 9;
  const double* oN = pc.a_is_M ? pc.X_WB + 9 : pc.X_WA + 9;
  const double* VM = pc.a_is_M ? pc.V_WA : pc.V_WB;
  const double* VN = pc.a_is_M ? pc.V_WB : pc.V_WA;
  double f[3];
  point_force(cW, n, fn0, k, pc.dstar, oM, VM, oN, VN, pc.mu, pc.dt, pc.apply, f);
  /* +f on M, -f on N (plugin.cpp:477-482); reported on A and B. */
  double fA[3], fB[3];
  for (int c = 0; c < 3; ++c) {
    fA[c] = pc.a_is_M ? f[c] : -f[c];
    fB[c] = -fA[c];
  }
  double rA[3], rB[3], tA[3], tB[3];
  sub3(cW, pc.X_WA + 9, rA);
  sub3(cW, pc.X_WB + 9, rB);
  cross3(rA, fA, tA);
  cross3(rB, fB, tB);
  for (int c = 0; c < 3; ++c) {
    w.F[c].add(fA[c]);
    w.tA[c].add(tA[c]);
    w.tB[c].add(tB[c]);
  }
  if (frec) {
    for (int c = 0; c < 3; ++c) {
      frec[c] = cW[c];
      frec[3 + c] = n[c];
    }
    frec[6] = fn0;
    frec[7] = k;
    frec[8] = pc.dstar;
    frec[9] = area;
    frec[10] = p0;
    frec[11] = g;
  }
  return 1;
}

struct Outputs {
  double* faces;
  int64_t* face_keys;
  int64_t face_cap;
  double* polys;
  int64_t* poly_keys;
  int64_t poly_cap;
  int64_t npoly = 0, nface = 0;
};

/* One clipped polygon (soft/A frame) -> faces (kTriangle fan around the
 * centroid, or one kPolygon face) -> point collisions -> wrench.
 * Drake semantics (TriMeshBuilder / PolyMeshBuilder): the kTriangle fan is
 * (v_k, v_{k+1}, C) for k = 0..n-1 with the field evaluated at C; a fan
 * face's value at its centroid is Evaluate(f, (1/3,1/3,1/3)). */
void emit_polygon(const PairCtx& pc, const double (*v)[3], int n, const double* nA, const double* recA,
                  const double* gBW, bool b_soft, int64_t ka, int64_t kb, Wrench& w, Outputs& o) {
  double C[3];
  double W2 = polygon_centroid(v, n, nA, C);
  double pv[8];
  for (int i = 0; i < n; ++i) pv[i] = field(recA, v[i]);
  double pC = field(recA, C);
  double nW[3], gAW[3];
  rot(pc.X_WA, nA, nW);
  rot(pc.X_WA, recA + 16, gAW);
  if (o.polys && o.npoly < o.poly_cap) {
    double* P = o.polys + o.npoly * ORC_PREC;
    std::memset(P, 0, sizeof(double) * ORC_PREC);
    P[0] = n;
    for (int i = 0; i < n; ++i) {
      xform(pc.X_WA, v[i], P + 1 + 3 * i);
      P[25 + i] = pv[i];
    }
    xform(pc.X_WA, C, P + 33);
    P[36] = pC;
    P[37] = nW[0], P[38] = nW[1], P[39] = nW[2];
    o.poly_keys[2 * o.npoly] = ka;
    o.poly_keys[2 * o.npoly + 1] = kb;
  }
  o.npoly++;
  const double third = 1.0 / 3.0;
  if (pc.repr == 0) {
    for (int k = 0; k < n; ++k) {
      const double* a = v[k];
      const double* b = v[(k + 1) % n];
      double e1[3], e2[3], cr[3], cS[3], cW[3];
      sub3(b, a, e1);
      sub3(C, a, e2);
      cross3(e1, e2, cr);
      double area = 0.5 * std::sqrt(dot3(cr, cr));
      for (int c = 0; c < 3; ++c) cS[c] = ((a[c] + b[c]) + C[c]) / 3.0;
      double p0 = (third * pv[k] + third * pv[(k + 1) % n]) + third * pC;
      xform(pc.X_WA, cS, cW);
      double* fr = (o.faces && o.nface < o.face_cap) ? o.faces + o.nface * ORC_FREC : nullptr;
      if (eval_face(pc, area, cW, p0, nW, gAW, gBW, b_soft, w, fr)) {
        if (fr) {
          o.face_keys[3 * o.nface] = ka;
          o.face_keys[3 * o.nface + 1] = kb;
          o.face_keys[3 * o.nface + 2] = k;
        }
        o.nface++;
      }
    }
  } else {
    double area = 0.5 * W2;
    double cW[3];
    xform(pc.X_WA, C, cW);
    double* fr = (o.faces && o.nface < o.face_cap) ? o.faces + o.nface * ORC_FREC : nullptr;
    if (eval_face(pc, area, cW, pC, nW, gAW, gBW, b_soft, w, fr)) {
      if (fr) {
        o.face_keys[3 * o.nface] = ka;
        o.face_keys[3 * o.nface + 1] = kb;
        o.face_keys[3 * o.nface + 2] = 0;
      }
      o.nface++;
    }
  }
}

/* ------------------------------------------------------ candidate search */

/* Median-split AABB tree over boxes (lo[3], hi[3]) used only to make the
 * oracle fast; the set it reports equals the brute-force set. */
struct CpuBvh {
  struct Node {
    double lo[3], hi[3];
    int left, right, first, count;
  };
  std::vector<Node> nodes;
  std::vector<int> idx;
  const double* box;  // stride 6: lo, hi

  void build(const double* boxes, int n) {
    box = boxes;
    idx.resize(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    nodes.clear();
    nodes.reserve(2 * n / 2 + 8);
    if (n > 0) rec(0, n);
  }
  int rec(int first, int count) {
    Node nd;
    for (int c = 0; c < 3; ++c) {
      nd.lo[c] = kInf;
      nd.hi[c] = -kInf;
    }
    for (int i = first; i < first + count; ++i) {
      const double* b = box + 6 * idx[i];
      for (int c = 0; c < 3; ++c) {
        nd.lo[c] = std::min(nd.lo[c], b[c]);
        nd.hi[c] = std::max(nd.hi[c], b[3 + c]);
      }
    }
    int me = (int)nodes.size();
    nodes.push_back(nd);
    if (count <= 4) {
      nodes[me].left = nodes[me].right = -1;
      nodes[me].first = first;
      nodes[me].count = count;
      return me;
    }
    int axis = 0;
    double ext = -1;
    for (int c = 0; c < 3; ++c)
      if (nd.hi[c] - nd.lo[c] > ext) ext = nd.hi[c] - nd.lo[c], axis = c;
    int mid = first + count / 2;
    std::nth_element(idx.begin() + first, idx.begin() + mid, idx.begin() + first + count, [&](int a, int b) {
      return box[6 * a + axis] + box[6 * a + 3 + axis] < box[6 * b + axis] + box[6 * b + 3 + axis];
    });
    int l = rec(first, mid - first);
    int r = rec(mid, first + count - mid);
    nodes[me].left = l;
    nodes[me].right = r;
    nodes[me].first = nodes[me].count = 0;
    return me;
  }
  template <class F>
  void query(const double* lo, const double* hi, F&& fn) const {
    if (nodes.empty()) return;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node& nd = nodes[stack[--sp]];
      bool ov = true;
      for (int c = 0; c < 3; ++c) ov = ov && nd.lo[c] <= hi[c] && lo[c] <= nd.hi[c];
      if (!ov) continue;
      if (nd.left < 0) {
        for (int i = nd.first; i < nd.first + nd.count; ++i) {
          const double* b = box + 6 * idx[i];
          bool o2 = true;
          for (int c = 0; c < 3; ++c) o2 = o2 && b[c] <= hi[c] && lo[c] <= b[3 + c];
          if (o2) fn(idx[i]);
        }
      } else {
        stack[sp++] = nd.left;
        stack[sp++] = nd.right;
      }
    }
  }
};

inline bool overlap(const double* alo, const double* ahi, const double* blo, const double* bhi) {
  return alo[0] <= bhi[0] && blo[0] <= ahi[0] && alo[1] <= bhi[1] && blo[1] <= ahi[1] && alo[2] <= bhi[2] &&
         blo[2] <= ahi[2];
}

/* Sorted canonical candidate keys a * nb + b for boxes A (soft frame) and
 * boxes B (transformed into the soft frame). */
std::vector<int64_t> candidates(const std::vector<double>& boxA, int na, const std::vector<double>& boxB, int nb,
                                int brute) {
  std::vector<int64_t> keys;
  if (brute) {
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < nb; ++b)
        if (overlap(&boxA[6 * a], &boxA[6 * a + 3], &boxB[6 * b], &boxB[6 * b + 3])) keys.push_back((int64_t)a * nb + b);
  } else {
    CpuBvh bvh;
    bvh.build(boxA.data(), na);
    for (int b = 0; b < nb; ++b)
      bvh.query(&boxB[6 * b], &boxB[6 * b + 3], [&](int a) { keys.push_back((int64_t)a * nb + b); });
    std::sort(keys.begin(), keys.end());
  }
  return keys;
}

void tet_boxes(const std::vector<double>& rec, int nt, std::vector<double>& box) {
  box.resize(6 * (size_t)nt);
  for (int t = 0; t < nt; ++t)
    for (int c = 0; c < 6; ++c) box[6 * t + c] = rec[(size_t)ORC_TREC * t + 23 + c];
}

void finish_wrench(const Wrench& w, double* out) {
  for (int c = 0; c < 3; ++c) {
    double F = w.F[c].get();
    out[c] = F;
    out[3 + c] = w.tA[c].get();
    out[6 + c] = -F;
    out[9 + c] = w.tB[c].get();
  }
}

/* Marching-tets slice of one tet by the plane s(x) = nh.x + off = 0
 * (vertices with s > 0 are above): the 3- or 4-gon on the crossed edges,
 * each point v_i + t (v_j - v_i) with t = s_i / (s_i - s_j) along the edge
 * (i < j), oriented counter-clockwise about nh.  Returns its vertex count,
 * 0 when the tet does not cross the plane. */
int slice_tet(const double* V, const int32_t* tet, const double* nh, double off, double (*poly)[3]) {
  const double* v[4];
  double s[4];
  int above = 0;
  for (int i = 0; i < 4; ++i) {
    v[i] = V + 3 * tet[i];
    s[i] = dot3(nh, v[i]) + off;
    above += s[i] > 0;
  }
  if (above == 0 || above == 4) return 0;
  int np = 0;
  int up[4], dn[4], nu = 0, nd = 0;
  for (int i = 0; i < 4; ++i) (s[i] > 0 ? up[nu++] : dn[nd++]) = i;
  int edges[4][2];
  int ne = 0;
  if (nu == 1 || nd == 1) {
    int lone = nu == 1 ? up[0] : dn[0];
    for (int i = 0; i < 4; ++i)
      if (i != lone) edges[ne][0] = std::min(lone, i), edges[ne][1] = std::max(lone, i), ++ne;
  } else {
    /* cyclic order (u0,d0) (u0,d1) (u1,d1) (u1,d0) */
    int pr[4][2] = {{up[0], dn[0]}, {up[0], dn[1]}, {up[1], dn[1]}, {up[1], dn[0]}};
    for (int e = 0; e < 4; ++e) edges[e][0] = std::min(pr[e][0], pr[e][1]), edges[e][1] = std::max(pr[e][0], pr[e][1]);
    ne = 4;
  }
  for (int e = 0; e < ne; ++e) {
    int i = edges[e][0], j = edges[e][1];
    double t = s[i] / (s[i] - s[j]);
    for (int c = 0; c < 3; ++c) poly[np][c] = v[i][c] + t * (v[j][c] - v[i][c]);
    ++np;
  }
  /* Orient counter-clockwise about nh. */
  double W = 0;
  for (int i = 1; i + 1 < np; ++i) {
    double e1[3], e2[3], cr[3];
    sub3(poly[i], poly[0], e1);
    sub3(poly[i + 1], poly[0], e2);
    cross3(e1, e2, cr);
    W += dot3(cr, nh);
  }
  if (W < 0)
    for (int i = 1, j = np - 1; i < j; ++i, --j)
      for (int c = 0; c < 3; ++c) std::swap(poly[i][c], poly[j][c]);
  return np;
}

}  // namespace

extern "C" {

void orc_tet_precompute(const double* V, int nv, const int32_t* T, int nt, const double* p, double* rec) {
  (void)nv;
  for (int t = 0; t < nt; ++t) tet_record(V, T + 4 * t, p, rec + (size_t)ORC_TREC * t);
}

/* Drake semantics (TriangleSurfaceMesh): face normal = normalised
 * (v1-v0)x(v2-v0); the zero vector for a degenerate triangle. */
void orc_tri_normals(const double* V, int nv, const int32_t* F, int ntri, double* nrm) {
  (void)nv;
  for (int f = 0; f < ntri; ++f) {
    const double *a = V + 3 * F[3 * f], *b = V + 3 * F[3 * f + 1], *c = V + 3 * F[3 * f + 2];
    double e1[3], e2[3], n[3];
    sub3(b, a, e1);
    sub3(c, a, e2);
    cross3(e1, e2, n);
    double z = dot3(n, n);
    if (z > 0) {
      double l = std::sqrt(z);
      n[0] /= l, n[1] /= l, n[2] /= l;
    }
    nrm[3 * f] = n[0], nrm[3 * f + 1] = n[1], nrm[3 * f + 2] = n[2];
  }
}

/* X_SR = X_WS^-1 X_WR (both poses as geom_xmat/geom_xpos, plugin.cpp:116-126). */
void orc_compose_SR(const double* X_WS, const double* X_WR, double* X_SR) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      X_SR[3 * i + j] = (X_WS[i] * X_WR[j] + X_WS[3 + i] * X_WR[3 + j]) + X_WS[6 + i] * X_WR[6 + j];
  double d[3] = {X_WR[9] - X_WS[9], X_WR[10] - X_WS[10], X_WR[11] - X_WS[11]};
  for (int i = 0; i < 3; ++i) X_SR[9 + i] = (X_WS[i] * d[0] + X_WS[3 + i] * d[1]) + X_WS[6 + i] * d[2];
}

int64_t orc_candidates(const double* VS, int nvs, const int32_t* T, int nt, const double* VR, int nvr,
                       const int32_t* F, int ntri, const double* X_WS, const double* X_WR, int brute, int64_t* keys,
                       int64_t cap) {
  (void)nvs;
  std::vector<double> boxA(6 * (size_t)nt), boxB(6 * (size_t)ntri);
  for (int t = 0; t < nt; ++t)
    for (int c = 0; c < 3; ++c) {
      double lo = VS[3 * T[4 * t] + c], hi = lo;
      for (int j = 1; j < 4; ++j) {
        lo = std::min(lo, VS[3 * T[4 * t + j] + c]);
        hi = std::max(hi, VS[3 * T[4 * t + j] + c]);
      }
      boxA[6 * t + c] = lo;
      boxA[6 * t + 3 + c] = hi;
    }
  double X_SR[12];
  orc_compose_SR(X_WS, X_WR, X_SR);
  std::vector<double> VRS(3 * (size_t)nvr);
  for (int v = 0; v < nvr; ++v) xform(X_SR, VR + 3 * v, &VRS[3 * v]);
  for (int f = 0; f < ntri; ++f)
    for (int c = 0; c < 3; ++c) {
      double lo = VRS[3 * F[3 * f] + c], hi = lo;
      for (int j = 1; j < 3; ++j) {
        lo = std::min(lo, VRS[3 * F[3 * f + j] + c]);
        hi = std::max(hi, VRS[3 * F[3 * f + j] + c]);
      }
      boxB[6 * f + c] = lo;
      boxB[6 * f + 3 + c] = hi;
    }
  std::vector<int64_t> k = candidates(boxA, nt, boxB, ntri, brute);
  int64_t n = (int64_t)k.size();
  if (keys)
    for (int64_t i = 0; i < n && i < cap; ++i) keys[i] = k[i];
  return n;
}

int orc_clip_tri_tet(const double* tri, const double* rec, double* out) {
  double poly[8][3];
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) poly[i][c] = tri[3 * i + c];
  int n = clip_by_planes(poly, 3, rec, rec + 12, 3);
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) out[3 * i + c] = poly[i][c];
  return n;
}

/* plugin.cpp:128-136 */
double orc_combined_modulus(double E_A, double E_B) {
  if (E_A == kInf) return E_B;
  if (E_B == kInf) return E_A;
  return E_A * E_B / (E_A + E_B);
}

/* plugin.cpp:138-159 */
double orc_combined_dissipation(double E_A, double d_A, double E_B, double d_B) {
  const double Estar = orc_combined_modulus(E_A, E_B);
  if (Estar == kInf) return 0.5 * (d_A + d_B);
  double d_star = 0;
  if (E_A != kInf) d_star += Estar / E_A * d_A;
  if (E_B != kInf) d_star += Estar / E_B * d_B;
  return d_star;
}

/* Drake semantics (CalcContactFrictionFromSurfaceProperties), the dynamic
 * coefficient used at plugin.cpp:427-433: 2 mu_A mu_B / (mu_A + mu_B), 0 when
 * the denominator is 0. */
double orc_combined_friction(double mu_A, double mu_B) {
  double den = mu_A + mu_B;
  return den == 0.0 ? 0.0 : 2 * mu_A * mu_B / den;
}

double orc_point_force(const double* pc, const double* oA, const double* VA, const double* oB, const double* VB,
                       double mu, double dt, int apply, double* f_out) {
  return point_force(pc, pc + 3, pc[6], pc[7], pc[8], oA, VA, oB, VB, mu, dt, apply, f_out);
}

/* collision_cb (plugin.cpp:255-318) for a SOFT-RIGID mesh pair, the Drake
 * tet-tri query it calls (plugin.cpp:300-304), evaluateContactSurface
 * (plugin.cpp:320-409) and the passiveCallback force loop (:411-483). */
}  // extern "C"

namespace {

/* Load-time state of a soft mesh: tet records, boxes and the CPU BVH (the
 * reference builds its Bvh<Obb, VolumeMesh> once at load,
 * plugin.cpp:656,680,704,730,790). */
struct SoftPrep {
  int nt = 0;
  std::vector<double> rec, box;
  CpuBvh bvh;
};

void soft_prepare(SoftPrep& sp, const double* VS, int nvs, const int32_t* T, int nt, const double* pS) {
  sp.nt = nt;
  sp.rec.resize((size_t)ORC_TREC * nt);
  orc_tet_precompute(VS, nvs, T, nt, pS, sp.rec.data());
  tet_boxes(sp.rec, nt, sp.box);
  sp.bvh.build(sp.box.data(), nt);
}

int soft_rigid_step(const SoftPrep& sp, const double* VR, int nvr, const int32_t* F, int ntri, const double* pose,
                    const double* params, int brute, double* wrench, int64_t* counts, double* faces,
                    int64_t* face_keys, int64_t face_cap, double* polys, int64_t* poly_keys, int64_t poly_cap) {
  const int nt = sp.nt;
  std::vector<double> nR(3 * (size_t)ntri);
  orc_tri_normals(VR, nvr, F, ntri, nR.data());
  PairCtx pc;
  pc.X_WA = pose;
  pc.X_WB = pose + 12;
  pc.V_WA = pose + 24;
  pc.V_WB = pose + 30;
  pc.dstar = orc_combined_dissipation(params[0], params[1], kInf, 1.0);
  pc.mu = orc_combined_friction(params[2], params[3]);
  pc.dt = params[4];
  pc.apply = params[5] != 0;
  pc.repr = (int)params[6];
  pc.a_is_M = params[7] != 0;

  double X_SR[12];
  orc_compose_SR(pc.X_WA, pc.X_WB, X_SR);
  std::vector<double> VRS(3 * (size_t)nvr);
  for (int v = 0; v < nvr; ++v) xform(X_SR, VR + 3 * v, &VRS[3 * v]);
  std::vector<double> boxB(6 * (size_t)ntri);
  for (int f = 0; f < ntri; ++f)
    for (int c = 0; c < 3; ++c) {
      double lo = VRS[3 * F[3 * f] + c], hi = lo;
      for (int j = 1; j < 3; ++j) {
        lo = std::min(lo, VRS[3 * F[3 * f + j] + c]);
        hi = std::max(hi, VRS[3 * F[3 * f + j] + c]);
      }
      boxB[6 * f + c] = lo;
      boxB[6 * f + 3 + c] = hi;
    }
  std::vector<int64_t> keys;
  if (brute) {
    keys = candidates(sp.box, nt, boxB, ntri, 1);
  } else {
    for (int b = 0; b < ntri; ++b)
      sp.bvh.query(&boxB[6 * b], &boxB[6 * b + 3], [&](int a) { keys.push_back((int64_t)a * ntri + b); });
    std::sort(keys.begin(), keys.end());
  }
  std::vector<double> nS(3 * (size_t)ntri);
  for (int f = 0; f < ntri; ++f) rot(X_SR, &nR[3 * f], &nS[3 * f]);

  Wrench w;
  Outputs o{faces, face_keys, face_cap, polys, poly_keys, poly_cap};
  const double zero3[3] = {0, 0, 0};
  for (int64_t key : keys) {
    int t = (int)(key / ntri), f = (int)(key % ntri);
    const double* r = &sp.rec[(size_t)ORC_TREC * t];
    const double* n = &nS[3 * f];
    /* Drake semantics (IsFaceNormalAlongPressureGradient). */
    if (!(dot3(r + 20, n) > kCosAlpha)) continue;
    double poly[8][3];
    for (int i = 0; i < 3; ++i)
      for (int c = 0; c < 3; ++c) poly[i][c] = VRS[3 * F[3 * f + i] + c];
    int nv = clip_by_planes(poly, 3, r, r + 12, 3);
    if (nv < 3) continue;
    emit_polygon(pc, poly, nv, n, r, zero3, false, t, f, w, o);
  }
  finish_wrench(w, wrench);
  counts[0] = (int64_t)keys.size();
  counts[1] = o.npoly;
  counts[2] = o.nface;
  return 0;
}

}  // namespace

extern "C" {

/* collision_cb (plugin.cpp:255-318) for a SOFT-RIGID mesh pair, the Drake
 * tet-tri query it calls (plugin.cpp:300-304), evaluateContactSurface
 * (plugin.cpp:320-409) and the passiveCallback force loop (:411-483). */
int orc_soft_rigid(const double* VS, int nvs, const int32_t* T, int nt, const double* pS, const double* VR, int nvr,
                   const int32_t* F, int ntri, const double* pose, const double* params, int brute, double* wrench,
                   int64_t* counts, double* faces, int64_t* face_keys, int64_t face_cap, double* polys,
                   int64_t* poly_keys, int64_t poly_cap) {
  SoftPrep sp;
  soft_prepare(sp, VS, nvs, T, nt, pS);
  return soft_rigid_step(sp, VR, nvr, F, ntri, pose, params, brute, wrench, counts, faces, face_keys, face_cap,
                         polys, poly_keys, poly_cap);
}

void* orc_soft_prepare(const double* VS, int nvs, const int32_t* T, int nt, const double* pS) {
  SoftPrep* sp = new SoftPrep();
  soft_prepare(*sp, VS, nvs, T, nt, pS);
  return sp;
}

void orc_soft_free(void* sp) { delete (SoftPrep*)sp; }

int orc_soft_rigid_step(void* sp, const double* VR, int nvr, const int32_t* F, int ntri, const double* pose,
                        const double* params, double* wrench, int64_t* counts) {
  return soft_rigid_step(*(SoftPrep*)sp, VR, nvr, F, ntri, pose, params, 0, wrench, counts, nullptr, nullptr, 0,
                         nullptr, nullptr, 0);
}

/* FlatTactileSensor::bvh_update (S/src/flat_tactile_sensor.cpp:262-402) in
 * FP64, with the Moeller-Trumbore test of S/src/bvh.cpp:49-74 and the
 * nearest hit over all contact-surface triangles (the TLAS of
 * S/src/bvh.cpp:438-476 returns the nearest).  Triangles are the kTriangle
 * fan (v_k, v_{k+1}, C) of every polygon in record order; equal-t ties keep
 * the first triangle in that order. */
int orc_taxels(const double* polys, int64_t npoly, const double* sensor, int S, int window, double sigma, int cx,
               int cy, double* out) {
  const double* X = sensor;
  const double xs = sensor[12], ys = sensor[13], zs = sensor[14], res = sensor[15];
  const double nrm[3] = {X[2], X[5], X[8]};
  const double rSR = res / S;
  const double rmean = 1. / (S * S);
  const double di_factor = res / S;
  const double sub_halfwidth = di_factor / 2.0 - res / 2.0;
  const double max_dist = 1.41421356237 * res / 2.0;
  double rsigma_squared = 0, rtukey = 0;
  if (window == 1) rsigma_squared = 0.5 / (sigma * sigma);
  if (window == 2) rtukey = 1. / sigma * S * S;
  const double PI = 3.14159265358979323846;
  /* Acceleration only: bin polygons by their footprint in the sensor plane
   * (1e-9 margin) so each taxel tests the polygons that can be hit; bins keep
   * record order, so the first-hit tie rule is unchanged. */
  std::vector<std::vector<int64_t>> bins((size_t)cx * cy);
  for (int64_t q = 0; q < npoly; ++q) {
    const double* P = polys + q * ORC_PREC;
    int n = (int)P[0];
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (int i = 0; i < n; ++i) {
      const double* v = P + 1 + 3 * i;
      double d[3] = {v[0] - X[9], v[1] - X[10], v[2] - X[11]};
      for (int c = 0; c < 2; ++c) {
        double l = (X[c] * d[0] + X[3 + c] * d[1]) + X[6 + c] * d[2];
        lo[c] = std::min(lo[c], l), hi[c] = std::max(hi[c], l);
      }
    }
    int x0 = (int)std::floor((lo[0] - 1e-9 + xs) / res), x1 = (int)std::floor((hi[0] + 1e-9 + xs) / res);
    int y0 = (int)std::floor((lo[1] - 1e-9 + ys) / res), y1 = (int)std::floor((hi[1] + 1e-9 + ys) / res);
    x0 = std::max(x0, 0), y0 = std::max(y0, 0), x1 = std::min(x1, cx - 1), y1 = std::min(y1, cy - 1);
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) bins[(size_t)x + (size_t)cx * y].push_back(q);
  }
  for (int x = 0; x < cx; ++x)
    for (int y = 0; y < cy; ++y) {
      double avg = 0;
      const std::vector<int64_t>& bin = bins[(size_t)x + (size_t)cx * y];
      for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) {
          double loc[3] = {-xs + x * res + i * rSR + 0.5 * rSR, -ys + y * res + j * rSR + 0.5 * rSR, 1.5 * zs};
          double pt[3];
          xform(X, loc, pt);
          double O[3], D[3];
          for (int c = 0; c < 3; ++c) {
            O[c] = pt[c] + nrm[c] * 1e-8;
            D[c] = -nrm[c];
          }
          double best_t = 1e30, bu = 0, bv = 0;
          const double* bp[3] = {nullptr, nullptr, nullptr};
          for (int64_t q : bin) {
            const double* P = polys + q * ORC_PREC;
            int n = (int)P[0];
            for (int k = 0; k < n; ++k) {
              int k1 = (k + 1) % n;
              const double* v0 = P + 1 + 3 * k;
              const double* v1 = P + 1 + 3 * k1;
              const double* v2 = P + 33;
              double e1[3], e2[3], h[3], s[3], qv[3];
              sub3(v1, v0, e1);
              sub3(v2, v0, e2);
              cross3(D, e2, h);
              double a = dot3(e1, h);
              if (std::fabs(a) < 1e-10) continue;
              double f = 1.0 / a;
              sub3(O, v0, s);
              double u = f * dot3(s, h);
              if (u < 0.0 || u > 1.0) continue;
              cross3(s, e1, qv);
              double v = f * dot3(D, qv);
              if (v < 0.0 || u + v > 1.0) continue;
              double t = f * dot3(e2, qv);
              if (t < best_t && t > 0.0) {
                best_t = t, bu = u, bv = v;
                bp[0] = P + 25 + k;
                bp[1] = P + 25 + k1;
                bp[2] = P + 36;
              }
            }
          }
          if (best_t < 1.5 * zs && best_t > 0.0) {
            double b0 = 1 - bu - bv;
            double raw = ((b0 * *bp[0] + bu * *bp[1]) + bv * *bp[2]) * rmean;
            double weight = 1.0;
            if (window == 1) {
              double dist = std::hypot(di_factor * i + sub_halfwidth, di_factor * j + sub_halfwidth) / max_dist;
              weight = std::exp(-(dist * dist) * rsigma_squared);
            } else if (window == 2) {
              int ni = S / 2 - std::abs(S / 2 - i), nj = S / 2 - std::abs(S / 2 - j);
              if (ni <= sigma * S / 2) weight *= 0.5 * (1. - std::cos(2. * PI * ni * rtukey));
              if (nj <= sigma * S / 2) weight *= 0.5 * (1. - std::cos(2. * PI * nj * rtukey));
            } else if (window == 3) {
              double inv = 1 - std::hypot(di_factor * i + sub_halfwidth, di_factor * j + sub_halfwidth) / max_dist;
              weight = inv * inv;
            }
            avg += weight * raw;
          }
        }
      out[x + cx * y] = avg;
    }
  return 0;
}

/* std::default_random_engine of libstdc++ (minstd_rand0: x <- 16807 x mod
 * 2^31 - 1, default seed 1) feeding std::uniform_real_distribution<double>
 * (0, 1), i.e. generate_canonical<double, 53>: two draws (k = max(1,
 * (53 + 30 - 1) / 30)), sum = (g1 - 1) + (g2 - 1) r with r = 2^31 - 2,
 * divided by r^2 (r^2 formed in long double, then rounded to double);
 * a result >= 1 becomes the largest double below 1. */
struct MinstdUniform {
  uint64_t x = 1;
  uint64_t next() {
    x = (x * 16807u) % 2147483647u;
    return x;
  }
  double operator()() {
    const long double r = 2147483646.0L;
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; ++k) {
      sum += (double)(next() - 1) * tmp;
      tmp = (double)(tmp * r);
    }
    double u = sum / tmp;
    if (u >= 1.0) u = std::nextafter(1.0, 0.0);
    return u;
  }
};

/* TaxelSensor::internal_update (S/src/taxel_sensor.cpp:158-478), FP64.
 * tris: kTriangle fan triangles of the contact surfaces touching the
 * sensor, in order, each [v0, v1, v2 (W), e0, e1, e2] (12 doubles), with
 * surf[t] the index of the surface the triangle belongs to (non-decreasing).
 * Samples: DEFAULT (:174-203, st1 and st2 both from |v2 - v0|) or
 * AREA_IMPORTANCE (:205-246, one generator for the whole update, area
 * accumulators per surface).  Taxel positions: X_W_sensor applied to the
 * taxels (sensor frame).  dist = (-2 t.p + |t|^2) + |p|^2.  Methods:
 * CLOSEST (:285-319: the first nearest sample; within the margin the value
 * is its pressure only when visualize is set and |p| > 1e-6, else 0);
 * WEIGHTED and MEAN fall through to SQUARED (:320-457, no break), so all
 * three give SQUARED: sum over samples within the margin of
 * max(0, margin - sqrt(dist))^2 |p|, times sample_resolution.  Taxels no
 * sample reaches keep their value; with no samples at all every taxel is
 * 0 (:458-477).  Returns the number of samples. */
int64_t orc_taxel_sensor(const double* tris, const int32_t* surf, int64_t ntri, const double* X, const double* taxels,
                         int ntaxels, double margin, double res, int method, int sample_method, int visualize,
                         double* values) {
  std::vector<double> P, E;
  auto add = [&](const double* T, double b0, double b1, double b2) {
    for (int c = 0; c < 3; ++c) P.push_back((b0 * T[c] + b1 * T[3 + c]) + b2 * T[6 + c]);
    E.push_back((b0 * T[9] + b1 * T[10]) + b2 * T[11]);
  };
  auto dist3 = [](const double* a, const double* b) {
    const double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return std::sqrt(dot3(d, d));
  };
  if (sample_method == 0) {
    for (int64_t t = 0; t < ntri; ++t) {
      const double* T = tris + 12 * t;
      const int st0 = (int)(dist3(T + 3, T) / res) + 1;
      const int st1 = (int)(dist3(T + 6, T) / res) + 1;
      const int st2 = (int)(dist3(T + 6, T) / res) + 1;
      const int st = std::max(st0, st1);
      for (double a = 0; a <= 1; a += 1. / st)
        for (double b = 0; b <= 1; b += 1. / st2) add(T, a, (1 - a) * (1 - b), (1 - a) * b);
    }
  } else {
    MinstdUniform gen;
    for (int64_t t0 = 0; t0 < ntri;) {
      int64_t t1 = t0;
      double total = 0.0;
      std::vector<double> area;
      while (t1 < ntri && surf[t1] == surf[t0]) {
        const double* T = tris + 12 * t1;
        double e1[3], e2[3], cr[3];
        sub3(T + 3, T, e1);
        sub3(T + 6, T, e2);
        cross3(e1, e2, cr);
        area.push_back(0.5 * std::sqrt(dot3(cr, cr)));
        total += area.back();
        ++t1;
      }
      const double area_resolution = res * total;
      double acc = 0.0, at = 0.0;
      for (int64_t t = t0; t < t1; ++t) {
        at += area[t - t0];
        while (acc < at) {
          acc += area_resolution;
          const double u0 = gen(), u1 = gen();
          const double a = 1.0 - std::sqrt(u0);
          const double b = (1.0 - a) * u1;
          add(tris + 12 * t, a, (1 - a) * (1 - b), (1 - a) * b);
        }
      }
      t0 = t1;
    }
  }
  const int64_t m = (int64_t)E.size();
  if (m == 0) {
    for (int i = 0; i < ntaxels; ++i) values[i] = 0.0;
    return 0;
  }
  const double msq = margin * margin;
  for (int i = 0; i < ntaxels; ++i) {
    double tw[3];
    xform(X, taxels + 3 * i, tw);
    const double tt = dot3(tw, tw);
    auto dist = [&](int64_t j) {
      const double* p = &P[3 * j];
      return ((-2.0 * dot3(tw, p)) + tt) + dot3(p, p);
    };
    if (method == 0) {
      int64_t jm = 0;
      double dm = dist(0);
      for (int64_t j = 1; j < m; ++j) {
        const double d = dist(j);
        if (d < dm) dm = d, jm = j;
      }
      if (dm < msq) values[i] = (visualize && std::fabs(E[jm]) > 1e-6) ? E[jm] : 0.0;
    } else {
      double sum = 0.0;
      int64_t ws = 0;
      for (int64_t j = 0; j < m; ++j) {
        const double d = dist(j);
        if (d < msq) {
          const double w0 = std::max(0.0, margin - std::sqrt(d));
          sum += (w0 * w0) * std::fabs(E[j]);
          ++ws;
        }
      }
      if (ws > 0) values[i] = sum * res;
    }
  }
  return m;
}

/* Moeller-Trumbore of S/src/bvh.cpp:49-74 in FP64 (|det| < 1e-10 parallel;
 * u, v in the triangle; t > 0); returns t or -1. */
static double mt_ray(const double* O, const double* D, const double* v0, const double* v1, const double* v2, double* uo,
                     double* vo) {
  double e1[3], e2[3], h[3], sv[3], qv[3];
  sub3(v1, v0, e1);
  sub3(v2, v0, e2);
  cross3(D, e2, h);
  const double a = dot3(e1, h);
  if (std::fabs(a) < 1e-10) return -1;
  const double f = 1.0 / a;
  sub3(O, v0, sv);
  const double u = f * dot3(sv, h);
  if (u < 0.0 || u > 1.0) return -1;
  cross3(sv, e1, qv);
  const double v = f * dot3(D, qv);
  if (v < 0.0 || u + v > 1.0) return -1;
  const double t = f * dot3(e2, qv);
  if (!(t > 0.0)) return -1;
  *uo = u, *vo = v;
  return t;
}

/* CurvedSensor::internal_update (S/src/curved_sensor.cpp:388-481) in FP64:
 * the sensor's surface samples (points and normals in the sensor frame)
 * are moved by X_W_sensor (normals rotated, not renormalised); sample j
 * casts a ray from p_j + n_j 1e-8 along -n_j, nearest hit over the fan
 * triangles (first in order on equal t), accepted when 0 < t < margin,
 * raw = Evaluate(tri, (1 - u - v, u, v)); taxel i = sum over its list of
 * weight * raw, in list order.  With no triangles every taxel is 0. */
int orc_curved_sensor(const double* tris, int64_t ntri, const double* X, const double* pts, const double* nrms, int m,
                      const int32_t* ptr, const int32_t* idx, const double* w, int ntaxels, double margin,
                      double* values) {
  if (ntri == 0) {
    for (int i = 0; i < ntaxels; ++i) values[i] = 0.0;
    return 0;
  }
  std::vector<double> raw(m, 0.0);
  std::vector<char> hit(m, 0);
  for (int j = 0; j < m; ++j) {
    double pw[3], nw[3], O[3], D[3];
    xform(X, pts + 3 * j, pw);
    rot(X, nrms + 3 * j, nw);
    for (int c = 0; c < 3; ++c) O[c] = pw[c] + nw[c] * 1e-8, D[c] = -nw[c];
    double best = 1e30, bu = 0, bv = 0;
    int64_t bt = -1;
    for (int64_t t = 0; t < ntri; ++t) {
      const double* T = tris + 12 * t;
      double u, v;
      const double tt = mt_ray(O, D, T, T + 3, T + 6, &u, &v);
      if (tt > 0 && tt < best) best = tt, bu = u, bv = v, bt = t;
    }
    if (bt >= 0 && best < margin && best > 0.0) {
      const double* T = tris + 12 * bt;
      const double b0 = 1 - bu - bv;
      raw[j] = (b0 * T[9] + bu * T[10]) + bv * T[11];
      hit[j] = 1;
    }
  }
  int hits = 0;
  for (int i = 0; i < ntaxels; ++i) {
    double pr = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k)
      if (hit[idx[k]]) pr += w[k] * raw[idx[k]];
    values[i] = pr;
  }
  for (int j = 0; j < m; ++j) hits += hit[j];
  return hits;
}

/* The first n values of MinstdUniform (for pinning it against libstdc++). */
void orc_minstd_uniform(int n, double* out) {
  MinstdUniform g;
  for (int i = 0; i < n; ++i) out[i] = g();
}

/* collision_cb SOFT-SOFT branch (plugin.cpp:280-286): Drake semantics of
 * ComputeContactSurfaceFromCompliantVolumes -- per candidate tet pair the
 * equal-pressure plane f_A = f_B (normal along grad f_A - grad f_B, into A),
 * both gradient filters, the slice of tet A by that plane (marching tets)
 * clipped by the 4 half-spaces of tet B; pressure f_A on the polygon. */
int orc_soft_soft(const double* VA, int nva, const int32_t* TA, int nta, const double* pA, const double* VB, int nvb,
                  const int32_t* TB, int ntb, const double* pB, const double* pose, const double* params, int brute,
                  double* wrench, int64_t* counts, double* faces, int64_t* face_keys, int64_t face_cap, double* polys,
                  int64_t* poly_keys, int64_t poly_cap) {
  std::vector<double> recA((size_t)ORC_TREC * nta), recB((size_t)ORC_TREC * ntb);
  orc_tet_precompute(VA, nva, TA, nta, pA, recA.data());
  orc_tet_precompute(VB, nvb, TB, ntb, pB, recB.data());
  PairCtx pc;
  pc.X_WA = pose;
  pc.X_WB = pose + 12;
  pc.V_WA = pose + 24;
  pc.V_WB = pose + 30;
  pc.dstar = orc_combined_dissipation(params[0], params[1], params[3], params[4]);
  pc.mu = orc_combined_friction(params[2], params[5]);
  pc.dt = params[6];
  pc.apply = params[7] != 0;
  pc.repr = (int)params[8];
  pc.a_is_M = params[9] != 0;

  double X_AB[12];
  orc_compose_SR(pc.X_WA, pc.X_WB, X_AB);
  std::vector<double> VBA(3 * (size_t)nvb);
  for (int v = 0; v < nvb; ++v) xform(X_AB, VB + 3 * v, &VBA[3 * v]);
  std::vector<double> boxA, boxB(6 * (size_t)ntb);
  tet_boxes(recA, nta, boxA);
  for (int t = 0; t < ntb; ++t)
    for (int c = 0; c < 3; ++c) {
      double lo = VBA[3 * TB[4 * t] + c], hi = lo;
      for (int j = 1; j < 4; ++j) {
        lo = std::min(lo, VBA[3 * TB[4 * t + j] + c]);
        hi = std::max(hi, VBA[3 * TB[4 * t + j] + c]);
      }
      boxB[6 * t + c] = lo;
      boxB[6 * t + 3 + c] = hi;
    }
  std::vector<int64_t> keys = candidates(boxA, nta, boxB, ntb, brute);

  Wrench w;
  Outputs o{faces, face_keys, face_cap, polys, poly_keys, poly_cap};
  for (int64_t key : keys) {
    int a = (int)(key / ntb), b = (int)(key % ntb);
    const double* ra = &recA[(size_t)ORC_TREC * a];
    const double* rb = &recB[(size_t)ORC_TREC * b];
    /* Field of tet b in A's frame: f_b(x) = cb + gb.x. */
    double gb[3], gbh[3];
    rot(X_AB, rb + 16, gb);
    rot(X_AB, rb + 20, gbh);
    double cb = rb[19] - dot3(gb, X_AB + 9);
    double nn[3] = {ra[16] - gb[0], ra[17] - gb[1], ra[18] - gb[2]};
    double mag = std::sqrt(dot3(nn, nn));
    if (!(mag > 0.0)) continue;
    double nh[3] = {nn[0] / mag, nn[1] / mag, nn[2] / mag};
    double off = -((cb - ra[19]) / mag);
    if (!(dot3(nh, ra + 20) > kCosAlpha)) continue;
    double mnh[3] = {-nh[0], -nh[1], -nh[2]};
    if (!(dot3(mnh, gbh) > kCosAlpha)) continue;
    double poly[8][3];
    const int np = slice_tet(VA, TA + 4 * a, nh, off, poly);
    if (np == 0) continue;
    /* Tet b's half-spaces in A's frame. */
    double pn[12], po[4];
    for (int k = 0; k < 4; ++k) {
      rot(X_AB, rb + 3 * k, pn + 3 * k);
      po[k] = rb[12 + k] - dot3(pn + 3 * k, X_AB + 9);
    }
    int nv = clip_by_planes(poly, np, pn, po, 4);
    if (nv < 3) continue;
    double gBW[3];
    rot(pc.X_WB, rb + 16, gBW);
    emit_polygon(pc, poly, nv, nh, ra, gBW, true, a, b, w, o);
  }
  finish_wrench(w, wrench);
  counts[0] = (int64_t)keys.size();
  counts[1] = o.npoly;
  counts[2] = o.nface;
  return 0;
}

/* collision_cb SOFT-PLANE branch (plugin.cpp:296-299): Drake semantics of
 * ComputeContactSurfaceFromSoftVolumeRigidHalfSpace.  The rigid geometry
 * is the half-space z <= 0 of its frame; its boundary plane in the soft
 * frame has unit normal n = R_SR e_z (out of the rigid body, into the soft
 * one) through p_SR.  Candidates: tets whose AABB straddles the plane
 * (|n.c + off| <= sum |n_k| h_k for centre c and half-extents h).  Each
 * candidate passes the gradient filter of rigid faces (n.unit(grad p) >
 * cos 5pi/8), is sliced by the plane (marching tets) and its polygon is
 * evaluated with the soft field; keys (tet, 0). */
int orc_soft_plane(const double* VS, int nvs, const int32_t* T, int nt, const double* pS, const double* pose,
                   const double* params, double* wrench, int64_t* counts, double* faces, int64_t* face_keys,
                   int64_t face_cap, double* polys, int64_t* poly_keys, int64_t poly_cap) {
  std::vector<double> rec((size_t)ORC_TREC * nt), box;
  orc_tet_precompute(VS, nvs, T, nt, pS, rec.data());
  tet_boxes(rec, nt, box);
  PairCtx pc;
  pc.X_WA = pose;
  pc.X_WB = pose + 12;
  pc.V_WA = pose + 24;
  pc.V_WB = pose + 30;
  pc.dstar = orc_combined_dissipation(params[0], params[1], kInf, 1.0);
  pc.mu = orc_combined_friction(params[2], params[3]);
  pc.dt = params[4];
  pc.apply = params[5] != 0;
  pc.repr = (int)params[6];
  pc.a_is_M = params[7] != 0;
  double X_SR[12];
  orc_compose_SR(pc.X_WA, pc.X_WB, X_SR);
  const double ez[3] = {0, 0, 1};
  double n[3];
  rot(X_SR, ez, n);
  const double off = -dot3(n, X_SR + 9);
  Wrench w;
  Outputs o{faces, face_keys, face_cap, polys, poly_keys, poly_cap};
  const double zero3[3] = {0, 0, 0};
  int64_t ncand = 0;
  for (int t = 0; t < nt; ++t) {
    const double* b = &box[6 * (size_t)t];
    double dist = off, rad = 0.0;
    for (int c = 0; c < 3; ++c) {
      dist += n[c] * (0.5 * (b[c] + b[3 + c]));
      rad += std::fabs(n[c]) * (0.5 * (b[3 + c] - b[c]));
    }
    if (std::fabs(dist) > rad) continue;
    ++ncand;
    const double* r = &rec[(size_t)ORC_TREC * t];
    /* Drake semantics (IsPlaneNormalAlongPressureGradient), as for faces */
    if (!(dot3(r + 20, n) > kCosAlpha)) continue;
    double poly[8][3];
    const int np = slice_tet(VS, T + 4 * t, n, off, poly);
    if (np == 0) continue;
    emit_polygon(pc, poly, np, n, r, zero3, false, t, 0, w, o);
  }
  finish_wrench(w, wrench);
  counts[0] = ncand;
  counts[1] = o.npoly;
  counts[2] = o.nface;
  return 0;
}

}  // extern "C"
