/*
 * hcs_device.h -- FP64 contact geometry evaluated per candidate in the HIP
 * kernels.  Templated on the scalar type so the tangent-linear backward
 * (Dual) runs the same arithmetic as the forward pass (double).
 *
 * Arithmetic contract: no contraction (the file is compiled with
 * -ffp-contract=off and the pragma below), 3-term sums as ((a+b)+c), rigid
 * transforms x' = ((R_i0 x0 + R_i1 x1) + R_i2 x2) + t_i with R row-major.
 * Polygons live in registers: every array index is a compile-time constant
 * (selects replace dynamic indexing), so nothing spills to scratch.
 *
 * Reference semantics restated here (citations relative to the reference
 * repository; "Drake semantics" marks Drake v1.8.0 behaviour the plugin
 * calls and which is not vendored):
 *   - tet-tri / tet-tet filters, Sutherland-Hodgman clipping, near-duplicate
 *     removal, polygon centroid, kTriangle fan (Drake semantics, called at
 *     mujoco_contact_surfaces/src/mujoco_contact_surfaces_plugin.cpp:284,301);
 *   - face evaluation, plugin.cpp:320-409;
 *   - force law, plugin.cpp:435-475.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace hcs {

/* Drake semantics: gradient filters accept cos(angle) > cos(5 pi / 8). */
#define HCS_COS_ALPHA (-0.38268343236508978)
/* Drake semantics: near-duplicate vertices closer than 1e-14. */
#define HCS_EPS2 (1e-14 * 1e-14)
#define HCS_MAXV 8

/* ------------------------------------------------------------ scalars */

struct Dual {
  double v, d;
  Dual() = default;
  __host__ __device__ Dual(double a) : v(a), d(0) {}
  __host__ __device__ Dual(double a, double b) : v(a), d(b) {}
};
__device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline Dual operator/(Dual a, Dual b) {
  double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ inline Dual dsqrt(Dual a) {
  double s = sqrt(a.v);
  return Dual(s, s > 0 ? a.d / (2.0 * s) : 0.0);
}
__device__ inline double dsqrt(double a) { return sqrt(a); }
__device__ inline double val(double a) { return a; }
__device__ inline double val(Dual a) { return a.v; }
__device__ inline Dual dmax0(Dual a) { return a.v > 0. ? a : Dual(0.); }
__device__ inline double tangent(double) { return 0.0; }
__device__ inline double tangent(Dual a) { return a.d; }
__device__ inline double dmax0(double a) { return a > 0. ? a : 0.; }

template <class T>
__device__ inline T sel(bool c, T a, T b) {
  return c ? a : b;
}

template <class T>
__device__ inline T dot3(const T* a, const T* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
template <class T>
__device__ inline void cross3(const T* a, const T* b, T* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
template <class T, class X>
__device__ inline void xform(const X* M, const T* x, T* y) {
#pragma unroll
  for (int i = 0; i < 3; ++i) y[i] = ((M[3 * i] * x[0] + M[3 * i + 1] * x[1]) + M[3 * i + 2] * x[2]) + M[9 + i];
}
template <class T, class X>
__device__ inline void rot(const X* M, const T* x, T* y) {
#pragma unroll
  for (int i = 0; i < 3; ++i) y[i] = (M[3 * i] * x[0] + M[3 * i + 1] * x[1]) + M[3 * i + 2] * x[2];
}

/* X_SR = X_WS^-1 X_WR (plugin.cpp:116-126 poses composed). */
template <class T>
__device__ inline void compose_SR(const T* X_WS, const T* X_WR, T* X_SR) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      X_SR[3 * i + j] = (X_WS[i] * X_WR[j] + X_WS[3 + i] * X_WR[3 + j]) + X_WS[6 + i] * X_WR[6 + j];
  T d[3] = {X_WR[9] - X_WS[9], X_WR[10] - X_WS[10], X_WR[11] - X_WS[11]};
#pragma unroll
  for (int i = 0; i < 3; ++i) X_SR[9 + i] = (X_WS[i] * d[0] + X_WS[3 + i] * d[1]) + X_WS[6 + i] * d[2];
}

/* ------------------------------------------------------------ polygons */

/* Register polygon: SoA coordinates, count n <= HCS_MAXV. */
template <class T>
struct Poly {
  T x[HCS_MAXV], y[HCS_MAXV], z[HCS_MAXV];
  int n;
};

template <class T, int N>
__device__ inline T pick(const T* a, int i) {
  T r = a[0];
#pragma unroll
  for (int k = 1; k < N; ++k) r = sel(i == k, a[k], r);
  return r;
}
template <class T, int LO, int HI>
__device__ inline void put(T* a, int i, T v) {
#pragma unroll
  for (int k = LO; k < HI; ++k) a[k] = sel(i == k, v, a[k]);
}

/* Drake semantics (ClipPolygonByHalfSpace): for i = 0..n-1 with previous
 * vertex (i-1) mod n, a vertex is inside when n.x + off <= 0; the crossing
 * point is wa*cur + (1-wa)*prev with wa = s_prev / (s_prev - s_cur).  Input
 * holds at most NIN vertices; output at most NIN+1. */
template <class T, int NIN>
__device__ inline void clip_halfspace(const Poly<T>& in, const T* nh, T off, Poly<T>& out) {
  const int n = in.n;
#pragma unroll
  for (int k = 0; k < HCS_MAXV; ++k) out.x[k] = out.y[k] = out.z[k] = T(0.0);
  T s[NIN];
#pragma unroll
  for (int i = 0; i < NIN; ++i) s[i] = ((nh[0] * in.x[i] + nh[1] * in.y[i]) + nh[2] * in.z[i]) + off;
  T sl = pick<T, NIN>(s, n - 1);
  T xl = pick<T, NIN>(in.x, n - 1), yl = pick<T, NIN>(in.y, n - 1), zl = pick<T, NIN>(in.z, n - 1);
  int m = 0;
#pragma unroll
  for (int i = 0; i < NIN; ++i) {
    const bool valid = i < n;
    const T sp = (i == 0) ? sl : s[i - 1];
    const T px = (i == 0) ? xl : in.x[i - 1];
    const T py = (i == 0) ? yl : in.y[i - 1];
    const T pz = (i == 0) ? zl : in.z[i - 1];
    const bool cin = val(s[i]) <= 0, pin = val(sp) <= 0;
    const bool em_int = valid && (cin != pin);
    /* At most NIN+1 vertices are kept; in exact arithmetic a convex
     * polygon never produces more, rounding noise on near-coplanar input
     * could, and the surplus is dropped (the oracle does the same). */
    if (em_int && m < NIN + 1) {
      T wa = sp / (sp - s[i]);
      T wb = T(1.0) - wa;
      put<T, 0, NIN + 1>(out.x, m, wa * in.x[i] + wb * px);
      put<T, 0, NIN + 1>(out.y, m, wa * in.y[i] + wb * py);
      put<T, 0, NIN + 1>(out.z, m, wa * in.z[i] + wb * pz);
      ++m;
    }
    if (valid && cin && m < NIN + 1) {
      put<T, 0, NIN + 1>(out.x, m, in.x[i]);
      put<T, 0, NIN + 1>(out.y, m, in.y[i]);
      put<T, 0, NIN + 1>(out.z, m, in.z[i]);
      ++m;
    }
  }
  out.n = m;
}

/* Drake semantics (RemoveNearlyDuplicateVertices): keep v_i unless it is
 * within 1e-14 of the last kept vertex, then drop the last if it is within
 * 1e-14 of the first; fewer than 3 vertices is no polygon (n = 0). */
template <class T, int NMAX>
__device__ inline void dedupe(Poly<T>& p) {
  const int n = p.n;
  if (n <= 0) {
    p.n = 0;
    return;
  }
  T lx = p.x[0], ly = p.y[0], lz = p.z[0];
  int m = 1;
#pragma unroll
  for (int i = 1; i < NMAX; ++i) {
    if (i < n) {
      T dx = lx - p.x[i], dy = ly - p.y[i], dz = lz - p.z[i];
      bool nearv = val((dx * dx + dy * dy) + dz * dz) < HCS_EPS2;
      if (!nearv) {
        put<T, 1, NMAX>(p.x, m, p.x[i]);
        put<T, 1, NMAX>(p.y, m, p.y[i]);
        put<T, 1, NMAX>(p.z, m, p.z[i]);
        lx = p.x[i], ly = p.y[i], lz = p.z[i];
        ++m;
      }
    }
  }
  if (m >= 3) {
    T dx = lx - p.x[0], dy = ly - p.y[0], dz = lz - p.z[0];
    if (val((dx * dx + dy * dy) + dz * dz) < HCS_EPS2) --m;
  }
  p.n = m >= 3 ? m : 0;
}

/* Clip a triangle (already in the soft frame) by the 4 tet half-spaces of
 * record rec (unit normals rec[0..11], offsets rec[12..15]). */
template <class T, class R>
__device__ inline void clip_tri_tet(Poly<T>& p, const R* rec) {
  Poly<T> q;
  T nh[3];
  nh[0] = rec[0], nh[1] = rec[1], nh[2] = rec[2];
  clip_halfspace<T, 3>(p, nh, T(rec[12]), q);
  if (q.n == 0) { p.n = 0; return; }
  nh[0] = rec[3], nh[1] = rec[4], nh[2] = rec[5];
  clip_halfspace<T, 4>(q, nh, T(rec[13]), p);
  if (p.n == 0) return;
  nh[0] = rec[6], nh[1] = rec[7], nh[2] = rec[8];
  clip_halfspace<T, 5>(p, nh, T(rec[14]), q);
  if (q.n == 0) { p.n = 0; return; }
  nh[0] = rec[9], nh[1] = rec[10], nh[2] = rec[11];
  clip_halfspace<T, 6>(q, nh, T(rec[15]), p);
  dedupe<T, 7>(p);
}

/* Clip a convex polygon of at most 4 vertices by 4 half-spaces given as
 * normals nrm[12] and offsets off[4] (tet-tet case). */
template <class T>
__device__ inline void clip_quad_tet(Poly<T>& p, const T* nrm, const T* off) {
  Poly<T> q;
  clip_halfspace<T, 4>(p, nrm, off[0], q);
  if (q.n == 0) { p.n = 0; return; }
  clip_halfspace<T, 5>(q, nrm + 3, off[1], p);
  if (p.n == 0) return;
  clip_halfspace<T, 6>(p, nrm + 6, off[2], q);
  if (q.n == 0) { p.n = 0; return; }
  clip_halfspace<T, 7>(q, nrm + 9, off[3], p);
  dedupe<T, 8>(p);
}

/* ------------------------------------------- polygons in a memory column */

/* A polygon of at most HCS_MAXV vertices kept in memory, one column per
 * thread (LDS in the clip kernel): coordinate c of vertex k at
 * col[(3 k + c) * st].  The clip below is clip_halfspace / dedupe with the
 * same arithmetic on the same values, but a vertex goes to its output
 * position with one store rather than a select over every position. */
struct ColPoly {
  double* col;
  int st, n;
  __device__ double& at(int k, int c) const { return col[(3 * k + c) * st]; }
};

template <int NIN>
__device__ inline void clip_halfspace_col(ColPoly& P, const double* nh, double off) {
  const int n = P.n;
  /* the whole input first: the outputs overwrite it */
  double x[NIN], y[NIN], z[NIN], s[NIN];
#pragma unroll
  for (int i = 0; i < NIN; ++i) {
    x[i] = P.at(i, 0), y[i] = P.at(i, 1), z[i] = P.at(i, 2);
    s[i] = ((nh[0] * x[i] + nh[1] * y[i]) + nh[2] * z[i]) + off;
  }
  const double sl = pick<double, NIN>(s, n - 1);
  const double xl = pick<double, NIN>(x, n - 1), yl = pick<double, NIN>(y, n - 1), zl = pick<double, NIN>(z, n - 1);
  int m = 0;
#pragma unroll
  for (int i = 0; i < NIN; ++i) {
    const bool valid = i < n;
    const double sp = (i == 0) ? sl : s[i - 1];
    const double px = (i == 0) ? xl : x[i - 1];
    const double py = (i == 0) ? yl : y[i - 1];
    const double pz = (i == 0) ? zl : z[i - 1];
    const bool cin = s[i] <= 0, pin = sp <= 0;
    if (valid && (cin != pin) && m < NIN + 1) {
      const double wa = sp / (sp - s[i]);
      const double wb = 1.0 - wa;
      P.at(m, 0) = wa * x[i] + wb * px;
      P.at(m, 1) = wa * y[i] + wb * py;
      P.at(m, 2) = wa * z[i] + wb * pz;
      ++m;
    }
    if (valid && cin && m < NIN + 1) {
      P.at(m, 0) = x[i], P.at(m, 1) = y[i], P.at(m, 2) = z[i];
      ++m;
    }
  }
  P.n = m;
}

template <int NMAX>
__device__ inline void dedupe_col(ColPoly& P) {
  const int n = P.n;
  if (n <= 0) {
    P.n = 0;
    return;
  }
  double x[NMAX], y[NMAX], z[NMAX];
#pragma unroll
  for (int i = 0; i < NMAX; ++i) x[i] = P.at(i, 0), y[i] = P.at(i, 1), z[i] = P.at(i, 2);
  double lx = x[0], ly = y[0], lz = z[0];
  int m = 1;
#pragma unroll
  for (int i = 1; i < NMAX; ++i) {
    if (i < n) {
      const double dx = lx - x[i], dy = ly - y[i], dz = lz - z[i];
      if (!(((dx * dx + dy * dy) + dz * dz) < HCS_EPS2)) {
        P.at(m, 0) = x[i], P.at(m, 1) = y[i], P.at(m, 2) = z[i];
        lx = x[i], ly = y[i], lz = z[i];
        ++m;
      }
    }
  }
  if (m >= 3) {
    const double dx = lx - x[0], dy = ly - y[0], dz = lz - z[0];
    if (((dx * dx + dy * dy) + dz * dz) < HCS_EPS2) --m;
  }
  P.n = m >= 3 ? m : 0;
}

/* clip_tri_tet on a column polygon holding the triangle */
template <class R>
__device__ inline void clip_tri_tet_col(ColPoly& P, const R* rec) {
  double nh[3] = {rec[0], rec[1], rec[2]};
  clip_halfspace_col<3>(P, nh, rec[12]);
  if (P.n == 0) return;
  nh[0] = rec[3], nh[1] = rec[4], nh[2] = rec[5];
  clip_halfspace_col<4>(P, nh, rec[13]);
  if (P.n == 0) return;
  nh[0] = rec[6], nh[1] = rec[7], nh[2] = rec[8];
  clip_halfspace_col<5>(P, nh, rec[14]);
  if (P.n == 0) return;
  nh[0] = rec[9], nh[1] = rec[10], nh[2] = rec[11];
  clip_halfspace_col<6>(P, nh, rec[15]);
  dedupe_col<7>(P);
}

/* clip_quad_tet on a column polygon */
__device__ inline void clip_quad_tet_col(ColPoly& P, const double* nrm, const double* off) {
  clip_halfspace_col<4>(P, nrm, off[0]);
  if (P.n == 0) return;
  clip_halfspace_col<5>(P, nrm + 3, off[1]);
  if (P.n == 0) return;
  clip_halfspace_col<6>(P, nrm + 6, off[2]);
  if (P.n == 0) return;
  clip_halfspace_col<7>(P, nrm + 9, off[3]);
  dedupe_col<8>(P);
}

/* The first NMAX vertices of a column polygon into registers (positions
 * at and past n hold whatever the column holds; no reader uses them). */
template <int NMAX>
__device__ inline void col_to_poly(const ColPoly& P, Poly<double>& p) {
#pragma unroll
  for (int k = 0; k < HCS_MAXV; ++k) {
    if (k < NMAX) p.x[k] = P.at(k, 0), p.y[k] = P.at(k, 1), p.z[k] = P.at(k, 2);
    else p.x[k] = p.y[k] = p.z[k] = 0.0;
  }
  p.n = P.n;
}

/* Drake semantics (CalcPolygonCentroid): area-weighted centroid of the fan
 * from v0 with weights (v_i - v0) x (v_{i+1} - v0) . n; the vertex mean when
 * the weights sum to zero.  Returns the weight sum (twice the area). */
template <class T, int NMAX>
__device__ inline T polygon_centroid(const Poly<T>& p, const T* nrm, T* C) {
  const int n = p.n;
  if (n == 3) {
    C[0] = ((p.x[0] + p.x[1]) + p.x[2]) / 3.0;
    C[1] = ((p.y[0] + p.y[1]) + p.y[2]) / 3.0;
    C[2] = ((p.z[0] + p.z[1]) + p.z[2]) / 3.0;
    T e1[3] = {p.x[1] - p.x[0], p.y[1] - p.y[0], p.z[1] - p.z[0]};
    T e2[3] = {p.x[2] - p.x[0], p.y[2] - p.y[0], p.z[2] - p.z[0]};
    T cr[3];
    cross3(e1, e2, cr);
    return dot3(cr, nrm);
  }
  T W = T(0.0), Sx = T(0.0), Sy = T(0.0), Sz = T(0.0);
#pragma unroll
  for (int i = 1; i + 1 < NMAX; ++i) {
    if (i + 1 < n) {
      T e1[3] = {p.x[i] - p.x[0], p.y[i] - p.y[0], p.z[i] - p.z[0]};
      T e2[3] = {p.x[i + 1] - p.x[0], p.y[i + 1] - p.y[0], p.z[i + 1] - p.z[0]};
      T cr[3];
      cross3(e1, e2, cr);
      T w = dot3(cr, nrm);
      W = W + w;
      Sx = Sx + w * ((p.x[0] + p.x[i]) + p.x[i + 1]);
      Sy = Sy + w * ((p.y[0] + p.y[i]) + p.y[i + 1]);
      Sz = Sz + w * ((p.z[0] + p.z[i]) + p.z[i + 1]);
    }
  }
  if (val(W) == 0) {
    T mx = T(0.0), my = T(0.0), mz = T(0.0);
#pragma unroll
    for (int i = 0; i < NMAX; ++i)
      if (i < n) mx = mx + p.x[i], my = my + p.y[i], mz = mz + p.z[i];
    C[0] = mx / (double)n, C[1] = my / (double)n, C[2] = mz / (double)n;
  } else {
    T W3 = 3.0 * W;
    C[0] = Sx / W3, C[1] = Sy / W3, C[2] = Sz / W3;
  }
  return W;
}

/* ------------------------------------------------------------ force law */

/* Everything of one pair a face evaluation needs (W frame). */
template <class T>
struct PairW {
  T X_WA[12], X_WB[12];
  T V_WA[6], V_WB[6];
  T dstar;
  double mu, dt;
  int apply, a_is_M, b_soft;
};

template <class T>
struct Acc {
  T F[3], tA[3], tB[3];
  __device__ void zero() {
#pragma unroll
    for (int c = 0; c < 3; ++c) F[c] = T(0.0), tA[c] = T(0.0), tB[c] = T(0.0);
  }
};

/* The quantities of plugin.cpp:345-404 that depend only on the polygon (its
 * normal and the two field gradients), shared by the polygon's faces:
 * n out of N into M, and the effective gradient g = 1/(1/gM + 1/gN) along
 * it.  ok is false when either side's gradient along n is below 1e-14 (no
 * face of the polygon is emitted).  nW: polygon normal out of B into A;
 * gAW/gBW: field gradients in W (gBW unused when B is rigid). */
template <class T>
struct FaceConst {
  T n[3], g;
  bool ok;
};
template <class T>
__device__ inline void face_const(const PairW<T>& pc, const T* nW, const T* gAW, const T* gBW, FaceConst<T>& fc) {
  /* M / N quantities picked by value (a pointer picked between two local
   * arrays would keep both in private memory) */
  const bool aM = pc.a_is_M;
  T gM[3], gN[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    fc.n[c] = aM ? nW[c] : -nW[c];
    gM[c] = aM ? gAW[c] : gBW[c];
    gN[c] = aM ? gBW[c] : gAW[c];
  }
  const bool M_soft = aM ? true : (bool)pc.b_soft, N_soft = aM ? (bool)pc.b_soft : true;
  const double kInf = __builtin_inf();
  const T gMv = M_soft ? dot3(gM, fc.n) : T(kInf);
  const T gNv = N_soft ? -dot3(gN, fc.n) : T(kInf);
  fc.ok = !(val(gMv) < 1.0e-14 || val(gNv) < 1.0e-14);
  /* 1/(1/gM + 1/gN); the rigid side contributes exactly 0 (value and tangent). */
  T inv = T(0.0);
  if (fc.ok && M_soft) inv = inv + T(1.0) / gMv;
  if (fc.ok && N_soft) inv = inv + T(1.0) / gNv;
  fc.g = fc.ok ? T(1.0) / inv : T(0.0);
}

/* plugin.cpp:345-404 (one face) followed by plugin.cpp:435-482 (its force),
 * with the polygon's constants fc (face_const); returns true when a point
 * collision is emitted.  rec (optional) receives p[3], n[3], fn0, k,
 * damping, area, p0, g. */
template <class T>
__device__ inline bool eval_face(const PairW<T>& pc, const FaceConst<T>& fc, T area, const T* cW, T p0, Acc<T>& acc,
                                 double* rec) {
  if (!(val(area) > 1.0e-14)) return false;
  if (!fc.ok) return false;
  const T* n = fc.n;
  const T g = fc.g;
  const T fn0 = area * p0;
  const T k = area * g;
  const bool aM = pc.a_is_M;
  T oM[3], oN[3], VM[6], VN[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    oM[c] = aM ? pc.X_WA[9 + c] : pc.X_WB[9 + c];
    oN[c] = aM ? pc.X_WB[9 + c] : pc.X_WA[9 + c];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    VM[c] = aM ? pc.V_WA[c] : pc.V_WB[c];
    VN[c] = aM ? pc.V_WB[c] : pc.V_WA[c];
  }
  /* Relative velocity at the quadrature point (plugin.cpp:440-456). */
  T rM[3] = {cW[0] - oM[0], cW[1] - oM[1], cW[2] - oM[2]};
  T rN[3] = {cW[0] - oN[0], cW[1] - oN[1], cW[2] - oN[2]};
  T wM[3], wN[3], v[3];
  cross3(VM, rM, wM);
  cross3(VN, rN, wN);
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = (VM[3 + c] + wM[c]) - (VN[3 + c] + wN[c]);
  const T vn = dot3(v, n);
  /* plugin.cpp:458 */
  const T fn = dmax0(T(1.) - pc.dstar * vn) * (fn0 - pc.dt * k * vn);
  T f[3] = {T(0.0), T(0.0), T(0.0)};
  if (pc.apply) {
    /* plugin.cpp:462-475 */
    const double stiction_tolerance = 1.0e-4;
    const double relative_tolerance = 1.0e-2;
    T vt[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) vt[c] = v[c] - n[c] * vn;
    double eps = stiction_tolerance * relative_tolerance;
    eps = eps * eps;
    const T v_slip = dsqrt(dot3(vt, vt) + eps);
    T mu_reg = T(pc.mu);
    const T s = v_slip / stiction_tolerance;
    if (val(s) < 1) mu_reg = pc.mu * s * (2.0 - s);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T that = vt[c] / v_slip;
      T f_slip = -mu_reg * that * fn;
      f[c] = f_slip + fn * n[c];
    }
  }
  /* +f on M, -f on N; reported on A (about o_A) and B (about o_B). */
  T fA[3], fB[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    fA[c] = pc.a_is_M ? f[c] : -f[c];
    fB[c] = -fA[c];
  }
  T rA[3] = {cW[0] - pc.X_WA[9], cW[1] - pc.X_WA[10], cW[2] - pc.X_WA[11]};
  T rB[3] = {cW[0] - pc.X_WB[9], cW[1] - pc.X_WB[10], cW[2] - pc.X_WB[11]};
  T tA[3], tB[3];
  cross3(rA, fA, tA);
  cross3(rB, fB, tB);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    acc.F[c] = acc.F[c] + fA[c];
    acc.tA[c] = acc.tA[c] + tA[c];
    acc.tB[c] = acc.tB[c] + tB[c];
  }
  if (rec) {
#pragma unroll
    for (int c = 0; c < 3; ++c) rec[c] = val(cW[c]), rec[3 + c] = val(n[c]);
    rec[6] = val(fn0), rec[7] = val(k), rec[8] = val(pc.dstar), rec[9] = val(area), rec[10] = val(p0), rec[11] = val(g);
  }
  return true;
}

}  // namespace hcs
