/*
 * hcs_query.h -- device bodies of the per-step contact pipeline, templated on
 * the scalar type where they do arithmetic on poses (double for the forward
 * step, Dual for the tangent-linear pass, one direction per grid.y slice).
 *
 *   1. run_cand: one wavefront owns one work item (a chunk of queries of one
 *      pair).  Each lane builds the AABB of its query element in the soft
 *      frame A (a rigid triangle moved by X_AB, or a tet of soft body B); in
 *      tet-query mode the wave walks A's Morton-ordered tets in tiles culled
 *      against the union box of the item's queries instead.  Lanes walk A's
 *      LBVH on private stacks; lanes that run dry take the bottom entry of a
 *      donor's stack (intra-wave work stealing).  Leaf hits go through the
 *      exact cheap rejections of survives(); survivors are compacted with a
 *      ballot + mbcnt prefix count and appended to the survivor list with a
 *      per-item sequence number.
 *   2. an exclusive scan over the per-item survivor counts gives each item
 *      its slot range, so contribution i of item k lands at off[k] + seq.
 *   3. clip_survivor: one survivor per thread -- clip, faces, force law,
 *      wrench contribution.
 *   4. per pair, contributions are summed lane-strided in slot order then by
 *      a fixed butterfly, so results do not depend on which wave ran which
 *      item or in what order survivors were appended.
 * Reference semantics per candidate: see hcs_device.h.
 */
#pragma once
#include "hcs_device.h"
#include "hcs_kernels.h"

#pragma clang fp contract(off)

namespace hcs {

/* Per-item uniform inputs. */
template <class T>
struct ItemCtx {
  PairW<T> pw;
  T X_AB[12];
  T escale; /* scales A's pressure field (E tangent); 1 in the forward pass */
  int repr;
};

template <class T>
__device__ inline T seed_val(double v) {
  return T(v);
}

/* Builds the per-item context.  dir < 0: forward; 0..2 rotation of B about
 * world axis dir, 3..5 translation of B along world axis dir-3, 6 modulus of
 * A (tangents 1/E_A on A's field). */
template <class T>
__device__ inline void make_item_ctx(const EnvDesc& E, const double* st, int dir, double dstar_dE, double invE,
                                     ItemCtx<T>& c);

template <>
__device__ inline void make_item_ctx<double>(const EnvDesc& E, const double* st, int, double, double,
                                             ItemCtx<double>& c) {
#pragma unroll
  for (int i = 0; i < 12; ++i) c.pw.X_WA[i] = st[i], c.pw.X_WB[i] = st[12 + i];
#pragma unroll
  for (int i = 0; i < 6; ++i) c.pw.V_WA[i] = st[24 + i], c.pw.V_WB[i] = st[30 + i];
  c.pw.mu = E.mu, c.pw.dstar = E.dstar, c.pw.dt = E.dt;
  c.pw.apply = E.apply, c.pw.a_is_M = E.a_is_M, c.pw.b_soft = E.b_soft;
  compose_SR(c.pw.X_WA, c.pw.X_WB, c.X_AB);
  c.escale = 1.0;
  c.repr = E.repr;
}

template <>
__device__ inline void make_item_ctx<Dual>(const EnvDesc& E, const double* st, int dir, double dstar_dE, double invE,
                                           ItemCtx<Dual>& c) {
#pragma unroll
  for (int i = 0; i < 12; ++i) c.pw.X_WA[i] = Dual(st[i]), c.pw.X_WB[i] = Dual(st[12 + i]);
#pragma unroll
  for (int i = 0; i < 6; ++i) c.pw.V_WA[i] = Dual(st[24 + i]), c.pw.V_WB[i] = Dual(st[30 + i]);
  if (dir >= 0 && dir < 3) {
    /* d/dtheta of exp([theta]) R at 0 is [e_dir]x R, row by row. */
    const double* R = st + 12;
    for (int j = 0; j < 3; ++j) {
      double r0 = R[j], r1 = R[3 + j], r2 = R[6 + j];
      double d0 = 0, d1 = 0, d2 = 0;
      if (dir == 0) d1 = -r2, d2 = r1;
      if (dir == 1) d0 = r2, d2 = -r0;
      if (dir == 2) d0 = -r1, d1 = r0;
      c.pw.X_WB[j].d = d0, c.pw.X_WB[3 + j].d = d1, c.pw.X_WB[6 + j].d = d2;
    }
  } else if (dir >= 3 && dir < 6) {
    c.pw.X_WB[9 + (dir - 3)].d = 1.0;
  }
  c.pw.mu = E.mu, c.pw.dt = E.dt;
  c.pw.dstar = Dual(E.dstar, dir == 6 ? dstar_dE : 0.0);
  c.pw.apply = E.apply, c.pw.a_is_M = E.a_is_M, c.pw.b_soft = E.b_soft;
  compose_SR(c.pw.X_WA, c.pw.X_WB, c.X_AB);
  c.escale = Dual(1.0, dir == 6 ? invE : 0.0);
  c.repr = E.repr;
}

/* Lanes of mask m below this lane. */
__device__ inline unsigned lane_prefix(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

struct OutPtrs {
  double* face_rec;
  int64_t* face_key;
  unsigned long long* face_cnt;
  int64_t face_cap;
  double* poly_rec;
  int64_t* poly_key;
  unsigned long long* poly_cnt;
  int64_t poly_cap;
  double* poly_tan;
  const unsigned char* poly_env; /* pairs whose polygons are kept (nullptr: every pair) */
};

/* Polygon (A frame) -> faces -> point collisions -> wrench; kTriangle fan
 * (v_k, v_{k+1}, C) with the field at C, or one kPolygon face. */
template <class T, int NMAX>
__device__ inline void emit_polygon(const ItemCtx<T>& ic, const Poly<T>& p, const T* nA, const T* gA, T cA,
                                    const T* gBW, Acc<T>& acc, int64_t& nf, const OutPtrs* o, int env, int ea,
                                    int eb) {
  T C[3];
  T W2 = polygon_centroid<T, NMAX>(p, nA, C);
  T pv[NMAX];
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    T x[3] = {p.x[i], p.y[i], p.z[i]};
    pv[i] = cA + dot3(gA, x);
  }
  const T pC = cA + dot3(gA, C);
  T nW[3], gAW[3];
  rot(ic.pw.X_WA, nA, nW);
  rot(ic.pw.X_WA, gA, gAW);
  FaceConst<T> fc;
  face_const(ic.pw, nW, gAW, gBW, fc);
  const int n = p.n;
  if (o && o->poly_rec && (!o->poly_env || o->poly_env[env])) {
    /* the lanes that keep a polygon here reserve their record slots with
     * one atomic (the lowest of them adds the count, the others read its
     * base): a per-polygon atomic on the one counter serialises when a
     * step keeps millions of polygons */
    const unsigned long long m = __ballot(1);
    const int lead = __builtin_ctzll(m);
    unsigned long long base = 0;
    if ((int)(threadIdx.x & 63) == lead) base = atomicAdd(o->poly_cnt, (unsigned long long)__popcll(m));
    base = __shfl(base, lead, 64);
    const unsigned long long slot = base + lane_prefix(m);
    if ((int64_t)slot < o->poly_cap) {
      double* P = o->poly_rec + slot * HCS_PREC;
      P[0] = n;
#pragma unroll
      for (int i = 0; i < NMAX; ++i) {
        T x[3] = {p.x[i], p.y[i], p.z[i]}, y[3];
        xform(ic.pw.X_WA, x, y);
        bool on = i < n;
        P[1 + 3 * i] = on ? val(y[0]) : 0.0;
        P[2 + 3 * i] = on ? val(y[1]) : 0.0;
        P[3 + 3 * i] = on ? val(y[2]) : 0.0;
        P[25 + i] = on ? val(pv[i]) : 0.0;
      }
#pragma unroll
      for (int i = NMAX; i < 8; ++i) P[1 + 3 * i] = P[2 + 3 * i] = P[3 + 3 * i] = P[25 + i] = 0.0;
      T y[3];
      xform(ic.pw.X_WA, C, y);
      P[33] = val(y[0]), P[34] = val(y[1]), P[35] = val(y[2]);
      P[36] = val(pC);
      P[37] = val(nW[0]), P[38] = val(nW[1]), P[39] = val(nW[2]);
      o->poly_key[3 * slot] = env, o->poly_key[3 * slot + 1] = ea, o->poly_key[3 * slot + 2] = eb;
      if (o->poly_tan) {
        double* Q = o->poly_tan + slot * HCS_PREC;
        Q[0] = 0.0;
#pragma unroll
        for (int i = 0; i < NMAX; ++i) {
          T x[3] = {p.x[i], p.y[i], p.z[i]}, y2[3];
          xform(ic.pw.X_WA, x, y2);
          bool on = i < n;
          Q[1 + 3 * i] = on ? tangent(y2[0]) : 0.0;
          Q[2 + 3 * i] = on ? tangent(y2[1]) : 0.0;
          Q[3 + 3 * i] = on ? tangent(y2[2]) : 0.0;
          Q[25 + i] = on ? tangent(pv[i]) : 0.0;
        }
#pragma unroll
        for (int i = NMAX; i < 8; ++i) Q[1 + 3 * i] = Q[2 + 3 * i] = Q[3 + 3 * i] = Q[25 + i] = 0.0;
        Q[33] = tangent(y[0]), Q[34] = tangent(y[1]), Q[35] = tangent(y[2]);
        Q[36] = tangent(pC);
        Q[37] = tangent(nW[0]), Q[38] = tangent(nW[1]), Q[39] = tangent(nW[2]);
      }
    }
  }
  const double third = 1.0 / 3.0;
  if (ic.repr == 0) {
#pragma unroll
    for (int k = 0; k < NMAX; ++k) {
      if (k < n) {
        const int k1 = (k + 1 == n) ? 0 : k + 1;
        T a[3] = {p.x[k], p.y[k], p.z[k]};
        T b[3] = {pick<T, NMAX>(p.x, k1), pick<T, NMAX>(p.y, k1), pick<T, NMAX>(p.z, k1)};
        T pb = pick<T, NMAX>(pv, k1);
        T e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        T e2[3] = {C[0] - a[0], C[1] - a[1], C[2] - a[2]};
        T cr[3];
        cross3(e1, e2, cr);
        T area = 0.5 * dsqrt(dot3(cr, cr));
        T cS[3] = {((a[0] + b[0]) + C[0]) / 3.0, ((a[1] + b[1]) + C[1]) / 3.0, ((a[2] + b[2]) + C[2]) / 3.0};
        T p0 = (third * pv[k] + third * pb) + third * pC;
        T cW[3];
        xform(ic.pw.X_WA, cS, cW);
        double rec[HCS_FREC];
        bool want = o && o->face_rec;
        if (eval_face(ic.pw, fc, area, cW, p0, acc, rec)) {
          ++nf;
          if (want) {
            unsigned long long slot = atomicAdd(o->face_cnt, 1ull);
            if ((int64_t)slot < o->face_cap) {
#pragma unroll
              for (int q = 0; q < HCS_FREC; ++q) o->face_rec[slot * HCS_FREC + q] = rec[q];
              o->face_key[4 * slot] = env, o->face_key[4 * slot + 1] = ea, o->face_key[4 * slot + 2] = eb,
                                o->face_key[4 * slot + 3] = k;
            }
          }
        }
      }
    }
  } else {
    T area = 0.5 * W2;
    T cW[3];
    xform(ic.pw.X_WA, C, cW);
    double rec[HCS_FREC];
    bool want = o && o->face_rec;
    if (eval_face(ic.pw, fc, area, cW, pC, acc, rec)) {
      ++nf;
      if (want) {
        unsigned long long slot = atomicAdd(o->face_cnt, 1ull);
        if ((int64_t)slot < o->face_cap) {
#pragma unroll
          for (int q = 0; q < HCS_FREC; ++q) o->face_rec[slot * HCS_FREC + q] = rec[q];
          o->face_key[4 * slot] = env, o->face_key[4 * slot + 1] = ea, o->face_key[4 * slot + 2] = eb,
                            o->face_key[4 * slot + 3] = 0;
        }
      }
    }
  }
}

/* One candidate (element a of A, element b of B). */
/* Marching-tets slice of tet a of A by the plane s(x) = nh.x + off = 0
 * (vertices with s > 0 are above): the 3- or 4-gon on the crossed edges,
 * each point v_i + t (v_j - v_i), t = s_i / (s_i - s_j) along edge i < j,
 * oriented counter-clockwise about nh.  False when the tet does not cross. */
template <class T>
__device__ inline bool slice_tet(const EnvDesc& E, int a, const T* nh, T off, Poly<T>& p) {
  const int32_t* ta = E.TA + 4 * (size_t)a;
  T vx[4], vy[4], vz[4], s[4];
  int above = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double* v = E.VA + 3 * (size_t)ta[i];
    vx[i] = T(v[0]), vy[i] = T(v[1]), vz[i] = T(v[2]);
    T x[3] = {vx[i], vy[i], vz[i]};
    s[i] = dot3(nh, x) + off;
    above += val(s[i]) > 0;
  }
  if (above == 0 || above == 4) return false;
  int up[4] = {0, 0, 0, 0}, dn[4] = {0, 0, 0, 0}, nu = 0, nd = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (val(s[i]) > 0) {
      up[0] = (nu == 0) ? i : up[0];
      up[1] = (nu == 1) ? i : up[1];
      up[2] = (nu == 2) ? i : up[2];
      ++nu;
    } else {
      dn[0] = (nd == 0) ? i : dn[0];
      dn[1] = (nd == 1) ? i : dn[1];
      dn[2] = (nd == 2) ? i : dn[2];
      ++nd;
    }
  }
  int ei[4], ej[4], ne;
  if (nu == 1 || nd == 1) {
    const int lone = nu == 1 ? up[0] : dn[0];
    int q = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i != lone) {
        const int lo = lone < i ? lone : i, hi = lone < i ? i : lone;
        ei[0] = q == 0 ? lo : ei[0], ej[0] = q == 0 ? hi : ej[0];
        ei[1] = q == 1 ? lo : ei[1], ej[1] = q == 1 ? hi : ej[1];
        ei[2] = q == 2 ? lo : ei[2], ej[2] = q == 2 ? hi : ej[2];
        ++q;
      }
    }
    ei[3] = ej[3] = 0;
    ne = 3;
  } else {
    const int pa[4] = {up[0], up[0], up[1], up[1]}, pb[4] = {dn[0], dn[1], dn[1], dn[0]};
#pragma unroll
    for (int e = 0; e < 4; ++e) ei[e] = pa[e] < pb[e] ? pa[e] : pb[e], ej[e] = pa[e] < pb[e] ? pb[e] : pa[e];
    ne = 4;
  }
#pragma unroll
  for (int i = 0; i < HCS_MAXV; ++i) p.x[i] = p.y[i] = p.z[i] = T(0.0);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (e < ne) {
      T si = pick<T, 4>(s, ei[e]), sj = pick<T, 4>(s, ej[e]);
      T t = si / (si - sj);
      T xi = pick<T, 4>(vx, ei[e]), yi = pick<T, 4>(vy, ei[e]), zi = pick<T, 4>(vz, ei[e]);
      T xj = pick<T, 4>(vx, ej[e]), yj = pick<T, 4>(vy, ej[e]), zj = pick<T, 4>(vz, ej[e]);
      p.x[e] = xi + t * (xj - xi);
      p.y[e] = yi + t * (yj - yi);
      p.z[e] = zi + t * (zj - zi);
    }
  }
  p.n = ne;
  /* orient counter-clockwise about nh */
  T W = T(0.0);
#pragma unroll
  for (int i = 1; i + 1 < 4; ++i) {
    if (i + 1 < ne) {
      T e1[3] = {p.x[i] - p.x[0], p.y[i] - p.y[0], p.z[i] - p.z[0]};
      T e2[3] = {p.x[i + 1] - p.x[0], p.y[i + 1] - p.y[0], p.z[i + 1] - p.z[0]};
      T cr[3];
      cross3(e1, e2, cr);
      W = W + dot3(cr, nh);
    }
  }
  if (val(W) < 0) {
    if (ne == 3) {
      T tx = p.x[1], ty = p.y[1], tz = p.z[1];
      p.x[1] = p.x[2], p.y[1] = p.y[2], p.z[1] = p.z[2];
      p.x[2] = tx, p.y[2] = ty, p.z[2] = tz;
    } else {
      T tx = p.x[1], ty = p.y[1], tz = p.z[1];
      p.x[1] = p.x[3], p.y[1] = p.y[3], p.z[1] = p.z[3];
      p.x[3] = tx, p.y[3] = ty, p.z[3] = tz;
    }
  }
  return true;
}

/* KS: the pair kind when every pair of the launch has it (the kind branches
 * fold away), or -1 for a launch that mixes kinds.  COL (double only): the
 * clip runs on a column polygon at colb with stride cols (ColPoly) instead
 * of in registers; the results are the same. */
template <class T, int KS = -1, bool COL = false>
__device__ inline void process_candidate(const EnvDesc& E, const ItemCtx<T>& ic, int env, int a, int b, Acc<T>& acc,
                                         int64_t& np, int64_t& nf, const OutPtrs* o, double* colb = nullptr,
                                         int cols = 1) {
  const int kind = KS >= 0 ? KS : E.kind;
  const double* __restrict__ rr = E.recA + (size_t)HCS_TREC * a;
  T gA[3] = {ic.escale * rr[16], ic.escale * rr[17], ic.escale * rr[18]};
  T cA = ic.escale * rr[19];
  if (kind == KIND_PLANE) {
    /* soft-half-space: B's plane z = 0 (normal +z, out of the rigid side)
     * in A's frame; the gradient filter of rigid faces, then the slice */
    const T ez[3] = {T(0.0), T(0.0), T(1.0)};
    T nS[3];
    rot(ic.X_AB, ez, nS);
    if (!(((rr[20] * val(nS[0]) + rr[21] * val(nS[1])) + rr[22] * val(nS[2])) > HCS_COS_ALPHA)) return;
    const T off = -dot3(nS, ic.X_AB + 9);
    Poly<T> p;
    if (!slice_tet<T>(E, a, nS, off, p)) return;
    ++np;
    T zero3[3] = {T(0.0), T(0.0), T(0.0)};
    emit_polygon<T, 4>(ic, p, nS, gA, cA, zero3, acc, nf, o, env, a, b);
  } else if (kind != KIND_SOFT) {
    /* soft-rigid: triangle b of the rigid mesh moved into A's frame */
    const int32_t* f = E.FB + 3 * (size_t)b;
    Poly<T> p;
#pragma unroll
    for (int i = 0; i < HCS_MAXV; ++i) p.x[i] = p.y[i] = p.z[i] = T(0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double* v = E.VB + 3 * (size_t)f[i];
      T x[3] = {T(v[0]), T(v[1]), T(v[2])}, y[3];
      xform(ic.X_AB, x, y);
      p.x[i] = y[0], p.y[i] = y[1], p.z[i] = y[2];
    }
    p.n = 3;
    const double* nr = E.nrmB + 3 * (size_t)b;
    T nR[3] = {T(nr[0]), T(nr[1]), T(nr[2])}, nS[3];
    rot(ic.X_AB, nR, nS);
    /* Drake semantics (IsFaceNormalAlongPressureGradient) */
    if (!(((rr[20] * val(nS[0]) + rr[21] * val(nS[1])) + rr[22] * val(nS[2])) > HCS_COS_ALPHA)) return;
    if constexpr (COL) {
      ColPoly P{colb, cols, 3};
#pragma unroll
      for (int i = 0; i < 3; ++i) P.at(i, 0) = p.x[i], P.at(i, 1) = p.y[i], P.at(i, 2) = p.z[i];
      clip_tri_tet_col(P, rr);
      if (P.n < 3) return;
      col_to_poly<7>(P, p);
    } else {
      clip_tri_tet(p, rr);
    }
    if (p.n < 3) return;
    ++np;
    T zero3[3] = {T(0.0), T(0.0), T(0.0)};
    emit_polygon<T, 7>(ic, p, nS, gA, cA, zero3, acc, nf, o, env, a, b);
  } else {
    /* soft-soft: equal-pressure plane of tets a (A) and b (B) */
    const double* __restrict__ rbb = E.recB + (size_t)HCS_TREC * b;
    T gbB[3] = {T(rbb[16]), T(rbb[17]), T(rbb[18])}, ghB[3] = {T(rbb[20]), T(rbb[21]), T(rbb[22])};
    T gb[3], gbh[3];
    rot(ic.X_AB, gbB, gb);
    rot(ic.X_AB, ghB, gbh);
    T cb = T(rbb[19]) - dot3(gb, ic.X_AB + 9);
    T nn[3] = {gA[0] - gb[0], gA[1] - gb[1], gA[2] - gb[2]};
    T mag = dsqrt(dot3(nn, nn));
    if (!(val(mag) > 0.0)) return;
    T nh[3] = {nn[0] / mag, nn[1] / mag, nn[2] / mag};
    T off = -((cb - cA) / mag);
    if (!(((val(nh[0]) * rr[20] + val(nh[1]) * rr[21]) + val(nh[2]) * rr[22]) > HCS_COS_ALPHA)) return;
    if (!(((-val(nh[0]) * val(gbh[0]) + -val(nh[1]) * val(gbh[1])) + -val(nh[2]) * val(gbh[2])) > HCS_COS_ALPHA))
      return;
    /* marching-tets slice of tet a; above the plane is s > 0 */
    Poly<T> p;
    if (!slice_tet<T>(E, a, nh, off, p)) return;
    /* tet b's half-spaces in A's frame */
    T pn[12], po[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      T nb[3] = {T(rbb[3 * k]), T(rbb[3 * k + 1]), T(rbb[3 * k + 2])};
      rot(ic.X_AB, nb, pn + 3 * k);
      po[k] = T(rbb[12 + k]) - dot3(pn + 3 * k, ic.X_AB + 9);
    }
    if constexpr (COL) {
      ColPoly P{colb, cols, p.n};
#pragma unroll
      for (int i = 0; i < 4; ++i) P.at(i, 0) = p.x[i], P.at(i, 1) = p.y[i], P.at(i, 2) = p.z[i];
      clip_quad_tet_col(P, pn, po);
      if (P.n < 3) return;
      col_to_poly<8>(P, p);
    } else {
      clip_quad_tet(p, pn, po);
    }
    if (p.n < 3) return;
    ++np;
    T gBW[3];
    rot(ic.pw.X_WB, gbB, gBW);
    emit_polygon<T, 8>(ic, p, nh, gA, cA, gBW, acc, nf, o, env, a, b);
  }
}

__device__ inline bool box_overlap(const double* a /*lo,hi*/, const double* lo, const double* hi) {
  return a[0] <= hi[0] && lo[0] <= a[3] && a[1] <= hi[1] && lo[1] <= a[4] && a[2] <= hi[2] && lo[2] <= a[5];
}

/* Whether a tet inside box bx (lo, hi) can meet triangle v (three vertices)
 * with unit normal n.  Separating axes: the triangle's normal (the same
 * test as the last of survives()) and, per edge, the in-plane direction
 * m = edge x n turned away from the opposite vertex, against the
 * triangle's largest projection on m.  A box that one axis separates from
 * the triangle by more than the margin holds no point within rounding of
 * the triangle, so clipping the triangle by the tet leaves nothing; the
 * axis itself may be rounded, since any axis separates.  The margin is
 * 1e-12 of the axis length plus 32 ulps of the coordinates the projections
 * add up (sum_k |m_k| (|c_k| + h_k + vmax)), so it stays above the rounding
 * of the projections and of the clip for meshes stored far from their
 * frame's origin or in large units. */
#define HCS_ULP32 (32.0 * 2.220446049250313e-16)
/* The triangle's half of the test, shared by the boxes tested against one
 * triangle: its largest coordinate magnitude, the plane offset -n.v0, and
 * per edge the in-plane axis m = edge x n turned away from the opposite
 * vertex (a negation, exact) and the triangle's largest projection on it. */
struct TriSat {
  double m[9], tmax[3], d0, vmax;
};
__device__ inline void tri_sat(const double* v, const double* n, TriSat& s) {
  double vmax = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) vmax = fmax(vmax, fabs(v[i]));
  s.vmax = vmax;
  s.d0 = -((n[0] * v[0] + n[1] * v[1]) + n[2] * v[2]);
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const double* a = v + 3 * e;
    const double* b = v + 3 * ((e + 1) % 3);
    const double* o = v + 3 * ((e + 2) % 3);
    const double ed[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double* m = s.m + 3 * e;
    cross3(ed, n, m);
    const double pa = dot3(m, a), pb = dot3(m, b), po = dot3(m, o);
    const bool flip = po > pa;
    s.tmax[e] = flip ? fmax(fmax(-pa, -pb), -po) : fmax(fmax(pa, pb), po);
    if (flip) m[0] = -m[0], m[1] = -m[1], m[2] = -m[2];
  }
}
/* The plane axis alone: whether box bx (lo, hi) can meet the plane
 * n.x + d0 = 0 of a triangle with unit normal n and coordinates of
 * magnitude at most vmax, within the margin above. */
__device__ inline bool box_may_meet_plane(const double* bx, const double* n, double d0, double vmax) {
  double dist = d0, rad = 0.0, sc = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double c = 0.5 * (bx[k] + bx[3 + k]), h = 0.5 * (bx[3 + k] - bx[k]);
    dist += n[k] * c, rad += fabs(n[k]) * h, sc += fabs(n[k]) * ((fabs(c) + h) + vmax);
  }
  return !(fabs(dist) > rad + (1e-12 + HCS_ULP32 * sc));
}
/* The box's half: bx (lo, hi) against the triangle of s with unit normal n. */
__device__ inline bool box_may_meet_tri(const double* bx, const double* n, const TriSat& s) {
  if (!box_may_meet_plane(bx, n, s.d0, s.vmax)) return false;
  double c[3], h[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) c[k] = 0.5 * (bx[k] + bx[3 + k]), h[k] = 0.5 * (bx[3 + k] - bx[k]);
  double mag[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) mag[k] = (fabs(c[k]) + h[k]) + s.vmax;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const double* m = s.m + 3 * e;
    double bmin = dot3(m, c), sc = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) bmin -= fabs(m[k]) * h[k], sc += fabs(m[k]) * mag[k];
    /* gap beyond the rounding margin, then beyond 1e-12 |m| (no square root) */
    const double gap = (bmin - s.tmax[e]) - HCS_ULP32 * sc;
    if (gap > 0.0 && gap * gap > 1e-24 * dot3(m, m)) return false;
  }
  return true;
}
__device__ inline bool box_may_meet_triangle(const double* bx, const double* v, const double* n) {
  TriSat s;
  tri_sat(v, n, s);
  return box_may_meet_tri(bx, n, s);
}


template <class T>
__device__ inline T wave_sum(T v);
template <>
__device__ inline double wave_sum<double>(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}
template <>
__device__ inline Dual wave_sum<Dual>(Dual v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    double a = __shfl_xor(v.v, o, 64), b = __shfl_xor(v.d, o, 64);
    v.v = v.v + a;
    v.d = v.d + b;
  }
  return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int64_t wave_sum_i(int64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

/* Early rejection, evaluated on primal values before the full clip: it
 * returns false only for candidates the clip leaves without a polygon, so
 * it saves work and never changes a result (the clip re-tests whatever is
 * kept).
 * soft-rigid:
 *  - all three triangle vertices farther than 1e-15 + 32 ulps of
 *    3 max|v| + |d| outside one tet face plane n.x + d: every vertex the
 *    clip creates is a convex combination of them, rounded by a few ulps
 *    of those magnitudes at each of the four stages, so that stage leaves
 *    nothing (tested on the float copy of the planes when the mesh has
 *    one);
 *  - the gradient filter of process_candidate (same arithmetic);
 *  - (before queueing, box_may_meet_triangle) the tet's AABB beyond the
 *    triangle's plane or an in-plane edge axis by more than its margin.
 * soft-soft: the equal-pressure plane, both gradient filters and the
 * marching-tets sign test of process_candidate (same arithmetic). */
/* The triangle case of survives() in two phases, so a caller can issue the
 * loads early and test later: the tet's float face planes (64 B) and unit
 * pressure gradient (three doubles of its record). */
struct TriSurvIn {
  float4 pl[4];
  double g[3];
};
__device__ inline void tri_surv_load(const EnvDesc& E, int a, TriSurvIn& s) {
  const HCS_GPTR float4* fp = E.fplA + 4 * (size_t)a;
#pragma unroll
  for (int k = 0; k < 4; ++k) s.pl[k] = fp[k];
  const HCS_GPTR double* r = E.recA + (size_t)HCS_TREC * a;
#pragma unroll
  for (int k = 0; k < 3; ++k) s.g[k] = r[20 + k];
}
/* tv, tn: the triangle's vertices and unit normal already moved into A's
 * frame (xform / rot with X, the same arithmetic as process_candidate).
 * The gradient filter is the one of process_candidate (same arithmetic).
 * The face-plane test runs on the float planes: each float differs from its
 * double by at most 2^-24 of it, so n.v + d moves by at most
 * 2^-24 (3 vmax + |d|) (unit n); a vertex more than twice that (plus the
 * double test's own margin) outside is outside in the double test too, so
 * every rejection here is one the double test makes.  The double test is
 * not repeated for the rest: it only saves clip work, and the clip decides
 * exactly whatever it keeps. */
__device__ inline bool tri_surv_test(const TriSurvIn& s, const double* tv, const double* tn) {
  double v[3][3];
  double vmax = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v[i][0] = tv[3 * i], v[i][1] = tv[3 * i + 1], v[i][2] = tv[3 * i + 2];
    vmax = fmax(vmax, fmax(fabs(v[i][0]), fmax(fabs(v[i][1]), fabs(v[i][2]))));
  }
  const bool filt = ((s.g[0] * tn[0] + s.g[1] * tn[1]) + s.g[2] * tn[2]) > HCS_COS_ALPHA;
  bool sep = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double n0 = s.pl[k].x, n1 = s.pl[k].y, n2 = s.pl[k].z, d = s.pl[k].w;
    const double tol = (1.2e-7 + HCS_ULP32) * (3.0 * vmax + fabs(d)) + 1e-15;
    bool out = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) out = out && (((n0 * v[i][0] + n1 * v[i][1]) + n2 * v[i][2]) + d > tol);
    sep = sep || out;
  }
  return filt && !sep;
}

template <int KS = -1>
__device__ inline bool survives(const EnvDesc& E, const double* X, int a, int b, const double* tv, const double* tn) {
  const int kind = KS >= 0 ? KS : E.kind;
  const double* r = E.recA + (size_t)HCS_TREC * a;
  if (kind == KIND_PLANE) {
    /* the gradient filter and the slice's sign test of process_candidate */
    const double ez[3] = {0.0, 0.0, 1.0};
    double nS[3];
    rot(X, ez, nS);
    if (!(((r[20] * nS[0] + r[21] * nS[1]) + r[22] * nS[2]) > HCS_COS_ALPHA)) return false;
    const double off = -dot3(nS, X + 9);
    const int32_t* ta = E.TA + 4 * (size_t)a;
    int above = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) above += (dot3(nS, hcs_plain(E.VA + 3 * (size_t)ta[i])) + off) > 0;
    return above != 0 && above != 4;
  }
  if (kind != KIND_SOFT && E.fplA) {
    TriSurvIn s;
    tri_surv_load(E, a, s);
    return tri_surv_test(s, tv, tn);
  }
  if (kind != KIND_SOFT) {
    /* tv, tn: triangle b's vertices and unit normal already moved into A's
     * frame (xform / rot with X, the same arithmetic as process_candidate) */
    double v[3][3];
    double vmax = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      v[i][0] = tv[3 * i], v[i][1] = tv[3 * i + 1], v[i][2] = tv[3 * i + 2];
      vmax = fmax(vmax, fmax(fabs(v[i][0]), fmax(fabs(v[i][1]), fabs(v[i][2]))));
    }
    const double nS[3] = {tn[0], tn[1], tn[2]};
    const bool filt = ((r[20] * nS[0] + r[21] * nS[1]) + r[22] * nS[2]) > HCS_COS_ALPHA;
    bool sep = false;
    { /* no float planes: the face-plane test on the record's double planes */
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double* n = r + 3 * k;
        const double d = r[12 + k];
        const double tol = HCS_ULP32 * (3.0 * vmax + fabs(d)) + 1e-15;
        bool out = true;
#pragma unroll
        for (int i = 0; i < 3; ++i) out = out && (((n[0] * v[i][0] + n[1] * v[i][1]) + n[2] * v[i][2]) + d > tol);
        sep = sep || out;
      }
    }
    return filt && !sep;
  }
  const double* rb = E.recB + (size_t)HCS_TREC * b;
  const double sc = 1.0;
  double gA[3] = {sc * r[16], sc * r[17], sc * r[18]};
  double cA = sc * r[19];
  double gbB[3] = {rb[16], rb[17], rb[18]}, ghB[3] = {rb[20], rb[21], rb[22]};
  double gb[3], gbh[3];
  rot(X, gbB, gb);
  rot(X, ghB, gbh);
  double cb = rb[19] - dot3(gb, X + 9);
  double nn[3] = {gA[0] - gb[0], gA[1] - gb[1], gA[2] - gb[2]};
  double mag = sqrt(dot3(nn, nn));
  if (!(mag > 0.0)) return false;
  double nh[3] = {nn[0] / mag, nn[1] / mag, nn[2] / mag};
  double off = -((cb - cA) / mag);
  if (!(((nh[0] * r[20] + nh[1] * r[21]) + nh[2] * r[22]) > HCS_COS_ALPHA)) return false;
  if (!(((-nh[0] * gbh[0] + -nh[1] * gbh[1]) + -nh[2] * gbh[2]) > HCS_COS_ALPHA)) return false;
  /* the slice of tet a by the plane (false: the plane misses it), then the
   * face-plane test of the triangle case on the slice against tet b's
   * half-spaces, placed in A's frame as process_candidate places them for
   * the clip */
  Poly<double> p;
  if (!slice_tet<double>(E, a, nh, off, p)) return false;
  double vmax = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) vmax = fmax(vmax, fmax(fabs(p.x[i]), fmax(fabs(p.y[i]), fabs(p.z[i]))));
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double nb[3] = {rb[3 * k], rb[3 * k + 1], rb[3 * k + 2]};
    double pn[3];
    rot(X, nb, pn);
    const double po = rb[12 + k] - dot3(pn, X + 9);
    const double tol = HCS_ULP32 * (3.0 * vmax + fabs(po)) + 1e-15;
    bool out = true;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      out = out && (i >= p.n || ((pn[0] * p.x[i] + pn[1] * p.y[i]) + pn[2] * p.z[i]) + po > tol);
    if (out) return false;
  }
  return true;
}

/* Whether an item can have a candidate: box P (B's frame) placed in A's
 * frame by X = [R | t] (y = R x + t) and bounded there by the axis-aligned
 * box of the rotated box, against box Q (A's frame).  The candidate test
 * compares axis-aligned boxes in A's frame, and every element box of P maps
 * into that bound, so this test alone is exact for it (a separating axis of
 * B's frame would not be: an element's A-frame box leaves P's B-frame box).
 * The tolerance (1e-9 of the coordinates involved) is many orders above
 * the rounding of the transformed element boxes. */
__device__ inline bool boxes_may_overlap(const double* X, const double* P, const double* Q) {
  double cp[3], hp[3], cq[3], hq[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    cp[c] = 0.5 * (P[c] + P[3 + c]), hp[c] = 0.5 * (P[3 + c] - P[c]);
    cq[c] = 0.5 * (Q[c] + Q[3 + c]), hq[c] = 0.5 * (Q[3 + c] - Q[c]);
  }
  double mag = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
    mag = fmax(mag, fmax(fabs(X[9 + c]), fmax(fabs(cp[c]) + hp[c], fabs(cq[c]) + hq[c])));
  const double tol = 1e-9 * mag + 1e-300;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double c1 = ((X[3 * i] * cp[0] + X[3 * i + 1] * cp[1]) + X[3 * i + 2] * cp[2]) + X[9 + i];
    const double h1 = (fabs(X[3 * i]) * hp[0] + fabs(X[3 * i + 1]) * hp[1]) + fabs(X[3 * i + 2]) * hp[2];
    if (fabs(c1 - cq[i]) > (h1 + hq[i]) + tol) return false;
  }
  return true;
}

/* Whether a box Q (A's frame) can reach B's plane z = 0 placed in A's frame
 * by X: |n.c + off| <= sum |n_k| h_k with a tolerance of 1e-9 of the
 * coordinates involved, so an item it rejects has no tet straddling the
 * plane under the exact per-tet test. */
__device__ inline bool box_may_meet_plane(const double* X, const double* Q) {
  const double ez[3] = {0.0, 0.0, 1.0};
  double n[3];
  rot(X, ez, n);
  const double off = -dot3(n, X + 9);
  double dist = off, rad = 0.0, mag = fabs(off);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double cq = 0.5 * (Q[c] + Q[3 + c]), hq = 0.5 * (Q[3 + c] - Q[c]);
    dist += n[c] * cq;
    rad += fabs(n[c]) * hq;
    mag = fmax(mag, fabs(cq) + hq);
  }
  return !(fabs(dist) > rad + (1e-9 * mag + 1e-300));
}

/* ===================================================================
 * k_cand: candidate generation, exact cheap tests, survivor compaction
 * =================================================================== */

/* Per-wave shared memory: candidate queue, survivor queue, rigid-triangle
 * tile (tet queries) or query boxes per slot (triangle / soft queries), the
 * work-donation mailbox and the pair transform. */
/* Traversal stack: the top HCS_LSTACK pending nodes of every lane live in
 * LDS (slot-major, so a wave's accesses to one slot are bank-conflict free);
 * deeper entries spill, oldest first, to a private overflow array of
 * HCS_STACK - HCS_LSTACK entries.  DFS over these LBVHs keeps well under 16
 * nodes pending, so the overflow is not touched in practice. */
#ifndef HCS_LSTACK
#define HCS_LSTACK 12
#endif

/* 13.1 KB, so 12 waves fit in a CU's 160 KB of LDS.  q1 takes up to 63
 * queued candidates plus 4 leaf hits from each of 64 lanes.  Query boxes are not
 * stored: a lane keeps its own in registers, and a box is rebuilt from the
 * vertices in tv / tn (min / max, exact) when a lane steals a subtree or a
 * tile triangle is tested. */
struct CandLds {
  int stk[HCS_LSTACK * 64];
  int2 q1[320];
  int2 q2[128]; /* while producing, q2 holds fewer than 64 entries and
                   q2[64..127] carries the work-stealing hand-over */
  double tv[64 * 9]; /* query / tile triangles (first three vertices of a
                        soft query tet) in A's frame */
  double tn[64 * 3]; /* triangles: unit normals in A's frame; soft queries:
                        the tet's fourth vertex */
  int qb[64];        /* triangle / soft queries: the element of each slot */
  double X[12];
};

/* The axis-aligned box of LDS slot k: its triangle, or with the fourth
 * vertex its soft query tet. */
__device__ inline void slot_box(const CandLds& L, int k, bool soft, double* lo, double* hi) {
  const double* v = &L.tv[9 * k];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    lo[c] = fmin(fmin(v[c], v[3 + c]), v[6 + c]);
    hi[c] = fmax(fmax(v[c], v[3 + c]), v[6 + c]);
    if (soft) lo[c] = fmin(lo[c], L.tn[3 * k + c]), hi[c] = fmax(hi[c], L.tn[3 * k + c]);
  }
}

/* Survivor slots are reserved per wave in chunks of this many (one atomic
 * per chunk); slots a wave leaves unused are filled with sentinels. */
#define HCS_SURV_CHUNK 256

__device__ inline void fill_sentinels(const QueryArgs& A, unsigned long long from, unsigned long long to) {
  const int lane = threadIdx.x & 63;
  for (unsigned long long k = from + lane; k < to; k += 64)
    if ((int64_t)k < A.surv_cap) A.surv[k] = make_int4(-1, -1, -1, -1);
}

/* One item: produce candidates (traversal with intra-wave work stealing, or
 * rigid-triangle tiles against 64 Morton-ordered tets), run the exact cheap
 * tests on every 64 queued candidates, and append survivors to HBM in
 * groups of up to 64 with a per-item sequence number, so the clip kernel can
 * place each contribution at item_off[item] + seq whatever order the groups
 * land in.  CAND_KEYS writes the canonical candidate keys instead and
 * CAND_COUNT only counts candidates per item (no tests, no survivors). */
template <int MODE, int KS = -1>
__device__ inline void run_cand(CandLds& L, const QueryArgs& A, int it, const Item item,
                                unsigned long long& s_next, unsigned long long& s_end) {
  const int lane = threadIdx.x & 63;
  const unsigned long long t_begin = A.item_stats ? clock64() : 0ull;
  const unsigned long long w_begin = A.item_stats ? wall_clock64() : 0ull;
  unsigned long long steps = 0;
  const int env = item.env;
  const EnvDesc E = A.envs[env];
  __syncthreads();
  if (lane < 12) L.X[lane] = A.xf[12 * (size_t)env + lane];
  __syncthreads();
  const double* X = L.X;
  int nc = 0;
  unsigned nsurv = 0;
  int q1n = 0, q2n = 0;
  const int kind = KS >= 0 ? KS : E.kind;
  const bool tet_mode = kind == KIND_TET;
  const bool plane_mode = kind == KIND_PLANE;
  bool producing = true;
  /* traversal state (triangle / soft queries) */
  int ovf[HCS_STACK - HCS_LSTACK];
  int bp = 0, sp = 0, on = 0, visits = 0, b = 0; /* LDS window [bp, sp) (circular), on spilled */
  const int visit_cap = E.ntA * 64;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  /* tile state (tet queries) */
  int t0 = 0, ntile = 0;
  unsigned long long tmask = 0;
  bool tile_loaded = false;
  bool valid = false;
  double bx[6] = {0, 0, 0, 0, 0, 0}, ulo[3] = {0, 0, 0}, uhi[3] = {0, 0, 0};
  if (tet_mode) {
    /* the lane's tet: 64 consecutive leaves in Morton order, so the item's
     * union box is compact and culls most rigid triangles */
    valid = lane < item.count;
    b = valid ? E.leafA[item.first + lane] : 0;
    const double* lb = E.boxLA + 6 * (size_t)(valid ? item.first + lane : 0); /* = boxA of tet b */
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bx[c] = valid ? lb[c] : __builtin_inf();
      bx[3 + c] = valid ? lb[3 + c] : -__builtin_inf();
      ulo[c] = wave_min(bx[c]);
      uhi[c] = wave_max(bx[3 + c]);
    }
  } else if (plane_mode) {
    /* 64 consecutive leaves; a tet is a candidate when its AABB straddles
     * B's plane: |n.c + off| <= sum |n_k| h_k (centre c, half-extents h) */
    valid = lane < item.count;
    b = valid ? E.leafA[item.first + lane] : 0;
    bool hit = false;
    if (valid) {
      const double* lb = E.boxLA + 6 * (size_t)(item.first + lane); /* = boxA of tet b */
      const double ez[3] = {0.0, 0.0, 1.0};
      double nS[3];
      rot(X, ez, nS);
      const double off = -dot3(nS, X + 9);
      double dist = off, rad = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        dist += nS[c] * (0.5 * (lb[c] + lb[3 + c]));
        rad += fabs(nS[c]) * (0.5 * (lb[3 + c] - lb[c]));
      }
      hit = !(fabs(dist) > rad);
    }
    const unsigned long long m = __ballot(hit);
    if (hit) L.q1[q1n + lane_prefix(m)] = make_int2(b, 0);
    q1n += __popcll(m);
    producing = false;
  } else {
    /* b is the lane's query slot; its element comes through qorder (a
     * spatially compact order of B's elements), and candidates carry the
     * slot until the exact rejections turn it into the element */
    valid = lane < item.count;
    b = lane;
    const int e = E.qorder ? E.qorder[item.first + lane < E.nB ? item.first + lane : 0] : item.first + lane;
    L.qb[lane] = e;
    if (valid) {
      const int nvq = kind == KIND_SOFT ? 4 : 3;
      const int32_t* f = E.FB + (size_t)nvq * e;
      for (int i = 0; i < nvq; ++i) {
        double y[3];
        xform(X, hcs_plain(E.VB + 3 * (size_t)f[i]), y);
        if (i < 3) {
#pragma unroll
          for (int c = 0; c < 3; ++c) L.tv[9 * lane + 3 * i + c] = y[c];
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c) L.tn[3 * lane + c] = y[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          lo[c] = i == 0 ? y[c] : fmin(lo[c], y[c]);
          hi[c] = i == 0 ? y[c] : fmax(hi[c], y[c]);
        }
      }
      if (kind != KIND_SOFT) {
        double nS[3];
        rot(X, hcs_plain(E.nrmB + 3 * (size_t)e), nS);
#pragma unroll
        for (int c = 0; c < 3; ++c) L.tn[3 * lane + c] = nS[c];
      }
    }
    if (E.ntA == 1 || item.root < 0) {
      /* a one-tet mesh, or a sub-item whose subtree is a single leaf */
      const int leaf = E.ntA == 1 ? 0 : ~item.root;
      const double* lb = E.boxA + 6 * (size_t)leaf;
      bool hit = valid && box_overlap(lb, lo, hi);
      if (hit && MODE == CAND_STEP && kind == KIND_TRI && !box_may_meet_triangle(lb, &L.tv[9 * lane], &L.tn[3 * lane]))
        hit = false, ++nc;
      const unsigned long long m = __ballot(hit);
      if (hit) L.q1[q1n + lane_prefix(m)] = make_int2(leaf, b);
      q1n += __popcll(m);
    } else if (valid) {
      L.stk[lane] = item.root;
      bp = 0, sp = 1, on = 0;
    }
  }
  for (;;) {
    /* ---- produce; in tet mode the queued candidates refer to the current
     * tile's triangles in LDS, so they are drained before the next tile */
    ++steps;
    const bool drain = tet_mode && producing && tmask == 0 && q1n > 0;
    if (producing && !drain) {
      if (tet_mode) {
        if (tmask == 0) {
          /* next tile of 64 rigid triangles: boxes into LDS, keep those that
           * touch the item's union box */
          if (tile_loaded) t0 += 64;
          tile_loaded = true;
          if (t0 >= E.nB) {
            producing = false;
          } else {
            ntile = (E.nB - t0) < 64 ? (E.nB - t0) : 64;
            bool touch = false;
            __syncthreads();
            if (lane < ntile) {
              double l3[3], h3[3];
              if (A.xtri && E.xtri_off >= 0) {
                /* moved once per step for the whole pair (k_xtri) */
                const double* o = A.xtri + HCS_XTRI * (size_t)(E.xtri_off + t0 + lane);
#pragma unroll
                for (int k = 0; k < 9; ++k) L.tv[9 * lane + k] = o[k];
#pragma unroll
                for (int c = 0; c < 3; ++c) L.tn[3 * lane + c] = o[9 + c], l3[c] = o[12 + c], h3[c] = o[15 + c];
              } else {
                const int32_t* f = E.FB + 3 * (size_t)(t0 + lane);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                  double y[3];
                  xform(X, hcs_plain(E.VB + 3 * (size_t)f[i]), y);
#pragma unroll
                  for (int c = 0; c < 3; ++c) {
                    L.tv[9 * lane + 3 * i + c] = y[c];
                    l3[c] = i == 0 ? y[c] : fmin(l3[c], y[c]);
                    h3[c] = i == 0 ? y[c] : fmax(h3[c], y[c]);
                  }
                }
                double nS[3];
                rot(X, hcs_plain(E.nrmB + 3 * (size_t)(t0 + lane)), nS);
#pragma unroll
                for (int c = 0; c < 3; ++c) L.tn[3 * lane + c] = nS[c];
              }
              const double ub[6] = {ulo[0], ulo[1], ulo[2], uhi[0], uhi[1], uhi[2]};
              touch = box_overlap(ub, l3, h3);
            }
            tmask = __ballot(touch);
            __syncthreads();
          }
        } else {
          const int jj = __builtin_ctzll(tmask);
          tmask &= tmask - 1;
          double tl[3], th[3];
          slot_box(L, jj, false, tl, th);
          bool hit = valid && box_overlap(bx, tl, th);
          if (hit && MODE == CAND_STEP && !box_may_meet_triangle(bx, &L.tv[9 * jj], &L.tn[3 * jj])) hit = false, ++nc;
          const unsigned long long m = __ballot(hit);
          if (hit) L.q1[q1n + lane_prefix(m)] = make_int2(b, t0 + jj);
          q1n += __popcll(m);
        }
      } else {
        int cnt = (sp - bp) + on;
        if (!__any(cnt > 0)) {
          producing = false;
        } else {
          /* work stealing: the k-th idle lane takes the bottom (shallowest,
           * largest) pending subtree in LDS of the k-th lane holding two or
           * more */
          const bool donor = cnt >= 2 && sp > bp;
          const unsigned long long idle = __ballot(cnt == 0), donors = __ballot(donor);
          const int npair = min(__popcll(idle), __popcll(donors));
          if (npair > 0) {
            int2* xfer = L.q2 + 64;
            if (donor) {
              const int r = (int)lane_prefix(donors);
              if (r < npair) {
                xfer[r] = make_int2(b, L.stk[(bp % HCS_LSTACK) * 64 + lane]);
                ++bp;
                --cnt;
              }
            }
            __syncthreads();
            if (cnt == 0) {
              const int r = (int)lane_prefix(idle);
              if (r < npair) {
                const int2 w = xfer[r];
                b = w.x;
                slot_box(L, w.x, kind == KIND_SOFT, lo, hi);
                L.stk[lane] = w.y;
                bp = 0, sp = 1, on = 0;
                cnt = 1;
              }
            }
            __syncthreads();
          }
          /* one step over a wide node (a tree of ntA leaves has ntA-1
           * internal nodes; the visit cap only guards against a malformed
           * tree) */
          int lh[4] = {-1, -1, -1, -1};
          if (cnt > 0 && ++visits > visit_cap) {
            atomicOr(A.error, 2);
            bp = sp = on = 0;
            cnt = 0;
          }
          if (cnt > 0) {
            const int nd = sp > bp ? L.stk[((--sp) % HCS_LSTACK) * 64 + lane] : ovf[--on];
            const WNode* N = E.wnodesA + nd;
            double bb[24];
#pragma unroll
            for (int i = 0; i < 24; ++i) bb[i] = N->b[i];
            const int4 ch = *(const int4*)N->ch;
            const int cc[4] = {ch.x, ch.y, ch.z, ch.w};
            /* push in reverse, so child 0 is walked first; a full LDS
             * window first spills its bottom entry */
#pragma unroll
            for (int k = 3; k >= 0; --k) {
              const int c = cc[k];
              if (c == HCS_WNONE || !box_overlap(bb + 6 * k, lo, hi)) continue;
              if (c < 0) {
                lh[k] = ~c;
                continue;
              }
              if (sp - bp == HCS_LSTACK) {
                if (on == HCS_STACK - HCS_LSTACK) {
                  atomicOr(A.error, 1);
                  continue;
                }
                ovf[on++] = L.stk[(bp % HCS_LSTACK) * 64 + lane];
                ++bp;
              }
              L.stk[(sp % HCS_LSTACK) * 64 + lane] = c;
              ++sp;
            }
            /* leaves whose tet box the triangle's plane separates from the
             * triangle: candidates the clip leaves empty, counted, not
             * queued.  Only the plane axis is tested here (the in-plane edge
             * axes of box_may_meet_triangle cost more at this point than
             * the few more rejections they make save later). */
            if (MODE == CAND_STEP && kind == KIND_TRI && (lh[0] >= 0 || lh[1] >= 0 || lh[2] >= 0 || lh[3] >= 0)) {
              const double* v = &L.tv[9 * b];
              const double* n = &L.tn[3 * b];
              double vmax = 0.0;
#pragma unroll
              for (int i = 0; i < 9; ++i) vmax = fmax(vmax, fabs(v[i]));
              const double d0 = -((n[0] * v[0] + n[1] * v[1]) + n[2] * v[2]);
#pragma unroll
              for (int k = 0; k < 4; ++k)
                if (lh[k] >= 0 && !box_may_meet_plane(bb + 6 * k, n, d0, vmax)) lh[k] = -1, ++nc;
            }
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const unsigned long long m = __ballot(lh[k] >= 0);
            if (lh[k] >= 0) L.q1[q1n + lane_prefix(m)] = make_int2(lh[k], b);
            q1n += __popcll(m);
          }
        }
      }
    }
    /* ---- consume: cheap tests on 64 candidates, flush 64 survivors;
     * q1 < 320 and q2 < 128 hold throughout */
    for (;;) {
      const bool do2 = q2n >= 64 || (!producing && q1n == 0 && q2n > 0);
      const bool do1 = q1n >= 64 || ((!producing || drain) && q1n > 0);
      if (!do1 && !do2) break;
      __syncthreads();
      if (do2) {
        const int take = q2n >= 64 ? 64 : q2n;
        const int base = q2n - take;
        if (s_next + take > s_end) {
          /* new chunk of survivor slots; the rest of the old one is padded */
          fill_sentinels(A, s_next, s_end);
          unsigned long long v = 0;
          if (lane == 0) v = atomicAdd(A.surv_cnt, (unsigned long long)HCS_SURV_CHUNK);
          s_next = __shfl(v, 0, 64);
          s_end = s_next + HCS_SURV_CHUNK;
        }
        const unsigned long long slot0 = s_next;
        s_next += take;
        if (lane < take) {
          const int2 c = L.q2[base + lane];
          const unsigned long long slot = slot0 + lane;
          if ((int64_t)slot < A.surv_cap) A.surv[slot] = make_int4(c.x, c.y, it, (int)(nsurv + lane));
        }
        nsurv += take;
        q2n = base;
      } else {
        const int take = q1n >= 64 ? 64 : q1n;
        const int base = q1n - take;
        bool keep = false;
        int2 c = make_int2(0, 0);
        if (lane < take) {
          c = L.q1[base + lane];
          ++nc;
          /* element of B and the LDS slot of its moved triangle */
          const bool by_slot = !tet_mode && !plane_mode;
          const int slot = tet_mode ? c.y - t0 : (plane_mode ? 0 : c.y);
          c.y = by_slot ? L.qb[c.y] : c.y;
          if (MODE == CAND_KEYS) {
            unsigned long long k = atomicAdd(A.cand_cnt, 1ull);
            if ((int64_t)k < A.cand_cap) A.cand_key[k] = (int64_t)c.x * E.nB + c.y;
          } else if (MODE == CAND_STEP) {
            keep = survives<KS>(E, X, c.x, c.y, &L.tv[9 * slot], &L.tn[3 * slot]);
          }
        }
        q1n = base;
        const unsigned long long m = __ballot(keep);
        __syncthreads();
        if (keep) L.q2[q2n + lane_prefix(m)] = c;
        q2n += __popcll(m);
      }
      __syncthreads();
    }
    if (!producing) break;
  }
  const int snc = (int)wave_sum_i(nc);
  if (lane == 0) {
    A.item_ncand[it] = snc;
    A.item_nsurv[it] = nsurv;
    if (A.item_stats) {
      unsigned long long* q = A.item_stats + 6 * (size_t)it;
      q[0] = clock64() - t_begin, q[1] = steps, q[2] = (unsigned long long)snc, q[3] = nsurv;
      q[4] = w_begin, q[5] = wall_clock64();
    }
  }
}

/* ===================================================================
 * k_clip: one survivor per thread -> contribution at item_off + seq
 * =================================================================== */

template <class T, int KS = -1, bool COL = false>
__device__ inline void clip_survivor(const QueryArgs& A, const ClipArgs& C, int64_t i, int dir, Contrib* out,
                                     const OutPtrs* o, double* colb = nullptr, int cols = 1) {
  const int4 s = A.surv[i];
  if (s.x < 0) return; /* unused slot of a wave's reserved chunk */
  const int64_t slot = (int64_t)C.item_off[s.z] + (unsigned)s.w;
  if (slot >= C.cap) return;
  const int env = A.items[s.z].env;
  const EnvDesc E = A.envs[env];
  ItemCtx<T> ic;
  make_item_ctx<T>(E, A.state + (size_t)HCS_STATE * env, dir, C.dstar_dE ? C.dstar_dE[env] : 0.0,
                   C.invE ? C.invE[env] : 0.0, ic);
  Acc<T> acc;
  acc.zero();
  int64_t np = 0, nf = 0;
  process_candidate<T, KS, COL>(E, ic, env, s.x, s.y, acc, np, nf, o, colb, cols);
  Contrib r;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    r.F[c] = dir < 0 ? val(acc.F[c]) : tangent(acc.F[c]);
    r.tA[c] = dir < 0 ? val(acc.tA[c]) : tangent(acc.tA[c]);
    r.tB[c] = dir < 0 ? val(acc.tB[c]) : tangent(acc.tB[c]);
  }
  r.np = (int)np, r.nf = (int)nf;
  out[slot] = r;
}

}  // namespace hcs
