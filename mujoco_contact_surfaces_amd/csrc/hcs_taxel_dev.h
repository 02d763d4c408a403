/*
 * hcs_taxel_dev.h -- device code of the flat tactile sensor sampling
 * (FlatTactileSensor::bvh_update,
 * mujoco_contact_surface_sensors/src/flat_tactile_sensor.cpp:262-402): the
 * per-taxel samplers shared by the plain and batched kernels of
 * hcs_taxels.hip, in a header so the CPU tests can build them too.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hcs_device.h"
#include "hcs_kernels.h"

#pragma clang fp contract(off)

namespace hcs {

__device__ inline void footprint(const double* P, const TaxelGridDev& g, int& x0, int& x1, int& y0, int& y1) {
  const int n = (int)P[0];
  double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
  for (int i = 0; i < n; ++i) {
    const double* v = P + 1 + 3 * i;
    double d[3] = {v[0] - g.X[9], v[1] - g.X[10], v[2] - g.X[11]};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      double l = (g.X[c] * d[0] + g.X[3 + c] * d[1]) + g.X[6 + c] * d[2];
      lo[c] = fmin(lo[c], l), hi[c] = fmax(hi[c], l);
    }
  }
  const double m = 1e-9;
  x0 = (int)floor((lo[0] - m + g.xs) / g.res);
  x1 = (int)floor((hi[0] + m + g.xs) / g.res);
  y0 = (int)floor((lo[1] - m + g.ys) / g.res);
  y1 = (int)floor((hi[1] + m + g.ys) / g.res);
  x0 = max(x0, 0), y0 = max(y0, 0), x1 = min(x1, g.cx - 1), y1 = min(y1, g.cy - 1);
}

__device__ inline int64_t pkey(const int64_t* key, int q) {
  return (key[3 * q] << 42) | (key[3 * q + 1] << 21) | key[3 * q + 2];
}

/* One workgroup per taxel; samples strided over the threads, each thread
 * sums its samples in index order, then a fixed tree reduction.  T = Dual
 * carries one tangent direction: polygon tangents in ptan, sensor pose
 * tangent in Xd (the hit triangle is chosen on primal values). */
template <class T>
__device__ inline T ld(const double* v, const double* d, size_t i);
template <>
__device__ inline double ld<double>(const double* v, const double*, size_t i) {
  return v[i];
}
template <>
__device__ inline Dual ld<Dual>(const double* v, const double* d, size_t i) {
  return Dual(v[i], d[i]);
}

/* Polygons of one chunk of a taxel's list, staged in LDS for its rays: the
 * vertices (v[8] = the centroid C) and, per fan triangle (v_k, v_k+1, C),
 * the sample-index ranges [i0, i1] x [j0, j1] (i0 / j0 in the low, i1 / j1
 * in the high 16 bits) outside of which no ray of the taxel can hit it, so a
 * ray runs the Moeller-Trumbore test only on the fans whose range holds it,
 * on vertices read from LDS. */
#define HCS_TX_CHUNK 64
struct TxStage {
  double v[HCS_TX_CHUNK][9][3];
  uint2 rng[HCS_TX_CHUNK][8];
  int q[HCS_TX_CHUNK], nv[HCS_TX_CHUNK];
  int64_t kq[HCS_TX_CHUNK];
};

/* The sample-index ranges of the fans of polygon P for taxel (x, y).
 *  - |a| < 1e-10 (a = e1 . (D x e2), the same expression the hit test
 *    rejects on): no ray can hit; empty range.
 *  - |a| < 1e-3 max(|e1|, |e2|)^2 (nearly edge-on to the rays): every ray.
 *  - otherwise the rays whose sensor-plane position lies within m of the
 *    fan's projected box, m = 1e-9 (1 + extent + height gap).  For such a
 *    triangle the computed barycentrics err by less than ~7e-13 (L + |s|) / L
 *    (L the longest edge, s = origin - v0), so an accepted ray lies within
 *    ~7e-13 (L + |s|) of the projected triangle: m is three orders wider, and
 *    the floor / ceil below add up to one more sample on each side. */
__device__ inline void stage_fans(const double (*vv)[3], int nv, const TaxelGridDev& g, int x, int y, uint2* rng) {
  const int S = g.S;
  const double rSR = g.res / S;
  const double Dv[3] = {-g.X[2], -g.X[5], -g.X[8]};
  const double ox = -g.xs + x * g.res + 0.5 * rSR, oy = -g.ys + y * g.res + 0.5 * rSR, oz = 1.5 * g.zs;
  for (int k = 0; k < nv; ++k) {
    const int k1 = (k + 1 == nv) ? 0 : k + 1;
    const double* v0 = vv[k];
    const double* v1 = vv[k1];
    const double* v2 = vv[8];
    double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    double h[3];
    cross3(Dv, e2, h);
    const double a = dot3(e1, h);
    const double L2 = fmax(dot3(e1, e1), dot3(e2, e2));
    uint2 r;
    if (fabs(a) < 1e-10) {
      r = make_uint2(1u, 1u); /* i0 = 1 > i1 = 0: empty */
    } else if (fabs(a) < 1e-3 * L2) {
      r = make_uint2((unsigned)(S - 1) << 16, (unsigned)(S - 1) << 16);
    } else {
      /* the fan's box in the sensor frame (x, y), and its height gap */
      double lx = 1e300, hx = -1e300, ly = 1e300, hy = -1e300, gap = 0.0;
      for (int t = 0; t < 3; ++t) {
        const double* w = t == 0 ? v0 : (t == 1 ? v1 : v2);
        const double d[3] = {w[0] - g.X[9], w[1] - g.X[10], w[2] - g.X[11]};
        const double px = (g.X[0] * d[0] + g.X[3] * d[1]) + g.X[6] * d[2];
        const double py = (g.X[1] * d[0] + g.X[4] * d[1]) + g.X[7] * d[2];
        const double pz = (g.X[2] * d[0] + g.X[5] * d[1]) + g.X[8] * d[2];
        lx = fmin(lx, px), hx = fmax(hx, px), ly = fmin(ly, py), hy = fmax(hy, py);
        gap = fmax(gap, fabs(oz - pz));
      }
      const double m = 1e-9 * (1.0 + fmax(hx - lx, hy - ly) + gap);
      /* clamp in double before the conversions */
      const double fi0 = fmax(floor((lx - m - ox) / rSR), 0.0), fi1 = fmin(ceil((hx + m - ox) / rSR), S - 1.0);
      const double fj0 = fmax(floor((ly - m - oy) / rSR), 0.0), fj1 = fmin(ceil((hy + m - oy) / rSR), S - 1.0);
      if (!(fi0 <= fi1) || !(fj0 <= fj1)) {
        r = make_uint2(1u, 1u);
      } else {
        r = make_uint2((unsigned)fi0 | ((unsigned)fi1 << 16), (unsigned)fj0 | ((unsigned)fj1 << 16));
      }
    }
    rng[k] = r;
  }
}

/* The value of taxel (x, y): one wave samples it, each lane the rays
 * lane, lane + 64, ... in order, then a fixed butterfly adds the lanes'
 * sums.  The taxel's candidate polygons are list[base .. base + n); when
 * they fit one chunk they are staged once for all rays. */
template <class T>
__device__ inline double sample_cell(const double* __restrict__ poly, const double* __restrict__ ptan,
                                     const int64_t* __restrict__ key, const TaxelGridDev& g,
                                     const double* __restrict__ Xd, unsigned base, unsigned n,
                                     const int* __restrict__ list, int x, int y) {
  __shared__ TxStage L;
  const int S = g.S;
  const double rSR = g.res / S;
  const double rmean = 1. / (S * S);
  const double di_factor = g.res / S;
  const double sub_halfwidth = di_factor / 2.0 - g.res / 2.0;
  const double max_dist = 1.41421356237 * g.res / 2.0;
  T X[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) X[i] = ld<T>(g.X, Xd, i);
  const T nrm[3] = {X[2], X[5], X[8]};
  double rsigma_squared = 0, rtukey = 0;
  if (g.window == 1) rsigma_squared = 0.5 / (g.sigma * g.sigma);
  if (g.window == 2) rtukey = 1. / g.sigma * S * S;
  const double PI = 3.14159265358979323846;
  T acc = T(0.0);
  const int lane = threadIdx.x & 63;
  const bool one_chunk = n <= HCS_TX_CHUNK;
  auto stage = [&](unsigned c0, unsigned cn) {
    for (unsigned l = lane; l < cn; l += 64) {
      const int q = list[base + c0 + l];
      const double* P = poly + (size_t)q * HCS_PREC;
      L.q[l] = q;
      L.nv[l] = min((int)P[0], 8);
      L.kq[l] = pkey(key, q);
      for (int k = 0; k < 27; ++k) L.v[l][k / 3][k % 3] = k < 24 ? P[1 + k] : P[33 + k - 24];
      stage_fans(L.v[l], L.nv[l], g, x, y, L.rng[l]);
    }
  };
  if (one_chunk) {
    stage(0, n);
    __syncthreads();
  }
  /* rays in rounds of 64 (every lane runs every round, so the staging
   * barriers are reached by all) */
  for (int s0 = 0; s0 < S * S; s0 += 64) {
    const int s = s0 + lane;
    const bool active = s < S * S;
    const int i = active ? s / S : 0, j = active ? s % S : 0;
    T loc[3] = {T(-g.xs + x * g.res + i * rSR + 0.5 * rSR), T(-g.ys + y * g.res + j * rSR + 0.5 * rSR),
                T(1.5 * g.zs)};
    T pt[3];
    xform(X, loc, pt);
    T O[3], D[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) O[c] = pt[c] + nrm[c] * 1e-8, D[c] = -nrm[c];
    double best_t = 1e30;
    int64_t bkey = INT64_MAX;
    int bk = 0, bq = -1;
    const double Ov[3] = {val(O[0]), val(O[1]), val(O[2])}, Dv[3] = {val(D[0]), val(D[1]), val(D[2])};
    for (unsigned c0 = 0; c0 < n; c0 += HCS_TX_CHUNK) {
      const unsigned cn = (n - c0) < HCS_TX_CHUNK ? (n - c0) : HCS_TX_CHUNK;
      if (!one_chunk) {
        __syncthreads();
        stage(c0, cn);
        __syncthreads();
      }
      if (!active) continue;
      for (unsigned l = 0; l < cn; ++l) {
        const int nv = L.nv[l];
        const int q = L.q[l];
        const int64_t kq = L.kq[l];
        for (int k = 0; k < nv; ++k) {
          const uint2 r = L.rng[l][k];
          if (i < (int)(r.x & 0xffffu) || i > (int)(r.x >> 16) || j < (int)(r.y & 0xffffu) || j > (int)(r.y >> 16))
            continue;
          const int k1 = (k + 1 == nv) ? 0 : k + 1;
          const double* v0 = L.v[l][k];
          const double* v1 = L.v[l][k1];
          const double* v2 = L.v[l][8];
          double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
          double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
          double h[3];
          cross3(Dv, e2, h);
          const double a = dot3(e1, h);
          if (fabs(a) < 1e-10) continue;
          const double f = 1.0 / a;
          double sv[3] = {Ov[0] - v0[0], Ov[1] - v0[1], Ov[2] - v0[2]};
          const double u = f * dot3(sv, h);
          if (u < 0.0 || u > 1.0) continue;
          double qv[3];
          cross3(sv, e1, qv);
          const double v = f * dot3(Dv, qv);
          if (v < 0.0 || u + v > 1.0) continue;
          const double t = f * dot3(e2, qv);
          if (!(t > 0.0)) continue;
          const bool better = t < best_t || (t == best_t && (kq < bkey || (kq == bkey && k < bk)));
          if (better) best_t = t, bkey = kq, bk = k, bq = q;
        }
      }
    }
    if (!active) continue;
    if (bq >= 0 && best_t < 1.5 * g.zs && best_t > 0.0) {
      /* Moeller-Trumbore on the chosen triangle, in T arithmetic */
      const size_t pb = (size_t)bq * HCS_PREC;
      const int nv = (int)poly[pb];
      const int k1 = (bk + 1 == nv) ? 0 : bk + 1;
      T v0[3], v1[3], v2[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        v0[c] = ld<T>(poly, ptan, pb + 1 + 3 * bk + c);
        v1[c] = ld<T>(poly, ptan, pb + 1 + 3 * k1 + c);
        v2[c] = ld<T>(poly, ptan, pb + 33 + c);
      }
      T e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
      T e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
      T h[3];
      cross3(D, e2, h);
      const T f = T(1.0) / dot3(e1, h);
      T sv[3] = {O[0] - v0[0], O[1] - v0[1], O[2] - v0[2]};
      const T u = f * dot3(sv, h);
      T qv[3];
      cross3(sv, e1, qv);
      const T v = f * dot3(D, qv);
      const T pa = ld<T>(poly, ptan, pb + 25 + bk), pbb = ld<T>(poly, ptan, pb + 25 + k1),
              pc = ld<T>(poly, ptan, pb + 36);
      const T b0 = T(1.0) - u - v;
      const T raw = ((b0 * pa + u * pbb) + v * pc) * rmean;
      double weight = 1.0;
      if (g.window == 1) {
        const double dist = hypot(di_factor * i + sub_halfwidth, di_factor * j + sub_halfwidth) / max_dist;
        weight = exp(-(dist * dist) * rsigma_squared);
      } else if (g.window == 2) {
        const int ni = S / 2 - abs(S / 2 - i), nj = S / 2 - abs(S / 2 - j);
        if (ni <= g.sigma * S / 2) weight *= 0.5 * (1. - cos(2. * PI * ni * rtukey));
        if (nj <= g.sigma * S / 2) weight *= 0.5 * (1. - cos(2. * PI * nj * rtukey));
      } else if (g.window == 3) {
        const double inv = 1 - hypot(di_factor * i + sub_halfwidth, di_factor * j + sub_halfwidth) / max_dist;
        weight = inv * inv;
      }
      acc = acc + weight * raw;
    }
  }
  double v = ptan ? tangent(acc) : val(acc);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return __shfl(v, 0, 64);
}

/* ---- fan-major sampling (double, S^2 <= HCS_TX_RAYS): a lane takes a fan
 * triangle and runs the hit test on the rays of its sample range only; the
 * nearest hit per ray is found with LDS atomics: the least t (its IEEE bits
 * order like t, t > 0), then among hits at that t the least (polygon key,
 * fan index), the tie rule of the ray-major loop above.  The packed key
 * kq * 8 + k needs kq < 2^61, i.e. pair indices below 2^19.
 *
 * The first pass also records, per ray, the fan whose atomic lowered t
 * (the owner) and flags a ray when a hit matched the t already there (a
 * tie).  A ray whose owner's recomputed t is its least t and that has no
 * tie is settled; only the other rays (and every ray of a taxel whose
 * polygon list takes more than one chunk) run the second pass, which finds
 * the least key among the hits at the least t.  Every hit test is the
 * ray-major loop's arithmetic on the same values, so the chosen triangle,
 * and the taxel, are the same. */
#define HCS_TX_RAYS 576
/* a lane takes HCS_TX_ROWS rows of a fan's range at a time; the rows past
 * a fan's first piece are queued for the whole wave, at most 11 pieces per
 * fan of a round of 64 (S <= 24 rows) */
#define HCS_TX_ROWS 2
#define HCS_TX_TASKS (64 * ((24 + HCS_TX_ROWS - 1) / HCS_TX_ROWS - 1))
static_assert(HCS_TX_RAYS == 24 * 24, "rays of the fan-major sampler: at most 24 rows of 24");
struct TxScatter {
  unsigned long long tmin[HCS_TX_RAYS], smin[HCS_TX_RAYS];
  /* first pass: list position * 8 + fan of the hit that lowered t last;
   * after the check, that of a settled ray, -1 for the others */
  int own[HCS_TX_RAYS];
  unsigned char flag[HCS_TX_RAYS]; /* the ray needs the second pass */
  int q[HCS_TX_CHUNK], fan0[HCS_TX_CHUNK + 1];
  long long kq[HCS_TX_CHUNK];
  /* pieces of HCS_TX_ROWS rows of the round's fans past their first
   * piece: fan * 32 + first row */
  int tq[HCS_TX_TASKS];
};

/* Hit-test constants of fan k of polygon P (the ray-major loop's e1, e2,
 * h = D x e2, f = 1 / a) and its sample range; false when no ray can hit it
 * (|a| < 1e-10, which the hit test rejects, or an empty range). */
struct TxFan {
  double v0[3], e1[3], e2[3], h[3], f;
  double px[3], py[3], m; /* the fan in the sensor's (x, y), the range margin */
  int i0, i1, j0, j1;
  bool rows; /* per-row ranges apply (false: the fan gets every ray) */
};
__device__ inline bool fan_setup(const double* P, int nv, int k, const TaxelGridDev& g, int x, int y, TxFan& c) {
  const int S = g.S;
  const double rSR = g.res / S;
  const double Dv[3] = {-g.X[2], -g.X[5], -g.X[8]};
  const int k1 = (k + 1 == nv) ? 0 : k + 1;
  const double* v0 = P + 1 + 3 * k;
  const double* v1 = P + 1 + 3 * k1;
  const double* v2 = P + 33;
#pragma unroll
  for (int d = 0; d < 3; ++d) c.v0[d] = v0[d], c.e1[d] = v1[d] - v0[d], c.e2[d] = v2[d] - v0[d];
  cross3(Dv, c.e2, c.h);
  const double a = dot3(c.e1, c.h);
  if (fabs(a) < 1e-10) return false;
  c.f = 1.0 / a;
  const double L2 = fmax(dot3(c.e1, c.e1), dot3(c.e2, c.e2));
  c.rows = false;
  if (fabs(a) < 1e-3 * L2) {
    c.i0 = c.j0 = 0, c.i1 = c.j1 = S - 1;
    return true;
  }
  /* the range of stage_fans (same bound, same margin) */
  const double ox = -g.xs + x * g.res + 0.5 * rSR, oy = -g.ys + y * g.res + 0.5 * rSR, oz = 1.5 * g.zs;
  double lx = 1e300, hx = -1e300, ly = 1e300, hy = -1e300, gap = 0.0;
  for (int t = 0; t < 3; ++t) {
    const double* w = t == 0 ? v0 : (t == 1 ? v1 : v2);
    const double d[3] = {w[0] - g.X[9], w[1] - g.X[10], w[2] - g.X[11]};
    const double px = (g.X[0] * d[0] + g.X[3] * d[1]) + g.X[6] * d[2];
    const double py = (g.X[1] * d[0] + g.X[4] * d[1]) + g.X[7] * d[2];
    const double pz = (g.X[2] * d[0] + g.X[5] * d[1]) + g.X[8] * d[2];
    c.px[t] = px, c.py[t] = py;
    lx = fmin(lx, px), hx = fmax(hx, px), ly = fmin(ly, py), hy = fmax(hy, py);
    gap = fmax(gap, fabs(oz - pz));
  }
  const double m = 1e-9 * (1.0 + fmax(hx - lx, hy - ly) + gap);
  c.m = m;
  c.rows = true;
  const double fi0 = fmax(floor((lx - m - ox) / rSR), 0.0), fi1 = fmin(ceil((hx + m - ox) / rSR), S - 1.0);
  const double fj0 = fmax(floor((ly - m - oy) / rSR), 0.0), fj1 = fmin(ceil((hy + m - oy) / rSR), S - 1.0);
  if (!(fi0 <= fi1) || !(fj0 <= fj1)) return false;
  c.i0 = (int)fi0, c.i1 = (int)fi1, c.j0 = (int)fj0, c.j1 = (int)fj1;
  return true;
}

/* The rays of row i (ray x index) that fan c can hold: the fan's y extent
 * over the slab of x within 2m of the row's x (each edge cut to the slab),
 * widened by 2m, within the fan's range [j0, j1]; empty when ja > jb.  The
 * margin m is the fan's range margin, so a row range holds every ray the
 * range holds that the hit test can accept. */
__device__ inline void fan_row(const TxFan& c, const TaxelGridDev& g, int x, int y, int i, int& ja, int& jb) {
  ja = c.j0, jb = c.j1;
  if (!c.rows) return;
  const double rSR = g.res / g.S;
  const double ox = -g.xs + x * g.res + 0.5 * rSR, oy = -g.ys + y * g.res + 0.5 * rSR;
  const double xa = ox + i * rSR - 2.0 * c.m, xb = ox + i * rSR + 2.0 * c.m;
  double lo = 1e300, hi = -1e300;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int e1 = e == 2 ? 0 : e + 1;
    const double x0 = c.px[e], y0 = c.py[e], x1 = c.px[e1], y1 = c.py[e1];
    if (fmax(x0, x1) < xa || fmin(x0, x1) > xb) continue;
    const double dx = x1 - x0;
    double ta = 0.0, tb = 1.0;
    if (dx != 0.0) {
      ta = (xa - x0) / dx, tb = (xb - x0) / dx;
      if (ta > tb) {
        const double t = ta;
        ta = tb, tb = t;
      }
      ta = fmin(fmax(ta, 0.0), 1.0), tb = fmin(fmax(tb, 0.0), 1.0);
    }
    const double ya = y0 + ta * (y1 - y0), yb = y0 + tb * (y1 - y0);
    lo = fmin(lo, fmin(ya, yb)), hi = fmax(hi, fmax(ya, yb));
  }
  if (!(lo <= hi)) {
    ja = 1, jb = 0;
    return;
  }
  const double fa = floor((lo - 2.0 * c.m - oy) / rSR), fb = ceil((hi + 2.0 * c.m - oy) / rSR);
  ja = fa > (double)ja ? (int)fmin(fa, (double)c.j1 + 1.0) : ja;
  jb = fb < (double)jb ? (int)fmax(fb, (double)c.j0 - 1.0) : jb;
}

/* The hit of ray (i, j) with fan k of polygon P without the fan's ray
 * range: fan_setup's hit constants, then fan_hit (the same arithmetic). */
__device__ inline double fan_hit_at(const double* P, int nv, int k, const TaxelGridDev& g, int x, int y, int i,
                                    int j);

/* The hit of ray (i, j) with fan c, as the ray-major loop computes it
 * (origin R loc + p + 1e-8 n, direction -n): t > 0, or a negative value. */
__device__ inline double fan_hit(const TxFan& c, const TaxelGridDev& g, int x, int y, int i, int j) {
  const double rSR = g.res / g.S;
  const double loc[3] = {-g.xs + x * g.res + i * rSR + 0.5 * rSR, -g.ys + y * g.res + j * rSR + 0.5 * rSR,
                         1.5 * g.zs};
  double pt[3];
  xform(g.X, loc, pt);
  const double nrm[3] = {g.X[2], g.X[5], g.X[8]};
  const double Dv[3] = {-nrm[0], -nrm[1], -nrm[2]};
  double O[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) O[d] = pt[d] + nrm[d] * 1e-8;
  double sv[3] = {O[0] - c.v0[0], O[1] - c.v0[1], O[2] - c.v0[2]};
  const double u = c.f * dot3(sv, c.h);
  if (u < 0.0 || u > 1.0) return -1.0;
  double qv[3];
  cross3(sv, c.e1, qv);
  const double v = c.f * dot3(Dv, qv);
  if (v < 0.0 || u + v > 1.0) return -1.0;
  const double t = c.f * dot3(c.e2, qv);
  return t > 0.0 ? t : -1.0;
}

__device__ inline double fan_hit_at(const double* P, int nv, int k, const TaxelGridDev& g, int x, int y, int i,
                                    int j) {
  const double Dv[3] = {-g.X[2], -g.X[5], -g.X[8]};
  const int k1 = (k + 1 == nv) ? 0 : k + 1;
  const double* v0 = P + 1 + 3 * k;
  const double* v1 = P + 1 + 3 * k1;
  const double* v2 = P + 33;
  TxFan c;
#pragma unroll
  for (int d = 0; d < 3; ++d) c.v0[d] = v0[d], c.e1[d] = v1[d] - v0[d], c.e2[d] = v2[d] - v0[d];
  cross3(Dv, c.e2, c.h);
  const double a = dot3(c.e1, c.h);
  if (fabs(a) < 1e-10) return -1.0;
  c.f = 1.0 / a;
  return fan_hit(c, g, x, y, i, j);
}

/* The entry of fan f of the staged chunk: the last entry whose first fan
 * is at most f (binary search; every entry has at least one fan). */
__device__ inline int fan_entry(const TxScatter& L, unsigned cn, int f) {
  int l = 0;
  for (int hi = (int)cn - 1; l < hi;) {
    const int mid = (l + hi + 1) >> 1;
    if (L.fan0[mid] <= f) l = mid;
    else hi = mid - 1;
  }
  return l;
}

/* Rows ia..ib of fan c (entry l, fan k of the chunk starting at list
 * position c0): PASS 0 keeps the least t per ray, PASS 1 the least (key,
 * fan) among hits at that t. */
template <int PASS>
__device__ inline void fan_rows(TxScatter& L, const TxFan& c, const TaxelGridDev& g, int x, int y, int ia, int ib,
                                unsigned c0, int l, int k) {
  const unsigned long long sk = ((unsigned long long)L.kq[l] << 3) | (unsigned)k;
  for (int i = ia; i <= ib; ++i) {
    int ja, jb;
    fan_row(c, g, x, y, i, ja, jb);
    for (int j = ja; j <= jb; ++j) {
      const int s = i * g.S + j;
      if (PASS == 1 && !L.flag[s]) continue;
      const double t = fan_hit(c, g, x, y, i, j);
      if (!(t > 0.0)) continue;
      const unsigned long long tb = (unsigned long long)__double_as_longlong(t);
      if (PASS == 0) {
        const unsigned long long old = atomicMin(&L.tmin[s], tb);
        if (tb < old) L.own[s] = (int)(((c0 + (unsigned)l) << 3) | (unsigned)k);
        else if (tb == old) L.flag[s] = 1;
      } else if (tb == L.tmin[s]) {
        atomicMin(&L.smin[s], sk);
      }
    }
  }
}

/* One pass over the fans of list[base + c0 ..] for the wave, 64 fans a
 * round: each lane takes the first HCS_TX_ROWS rows of its fan's range and
 * queues the rest in pieces of HCS_TX_ROWS rows, which the wave then
 * shares, so a large fan does not hold the round while the other lanes
 * idle.  Pieces cover disjoint rows of one fan, and the per-ray minima do
 * not depend on the order of the hits, so the result is that of one lane
 * per fan. */
template <int PASS>
__device__ inline void scatter_chunk(TxScatter& L, const double* __restrict__ poly, const TaxelGridDev& g, int x,
                                     int y, unsigned cn, unsigned c0) {
  const int lane = threadIdx.x & 63;
  const int nf = L.fan0[cn];
  for (int f0 = 0; f0 < nf; f0 += 64) {
    const int f = f0 + lane;
    unsigned extra = 0;
    int i0 = 0;
    if (f < nf) {
      const int l = fan_entry(L, cn, f);
      const int k = f - L.fan0[l];
      const double* P = poly + (size_t)L.q[l] * HCS_PREC;
      TxFan c;
      if (fan_setup(P, min((int)P[0], 8), k, g, x, y, c)) {
        i0 = c.i0;
        fan_rows<PASS>(L, c, g, x, y, c.i0, min(c.i1, c.i0 + HCS_TX_ROWS - 1), c0, l, k);
        extra = (unsigned)((c.i1 - c.i0) / HCS_TX_ROWS);
      }
    }
    /* queue the further pieces by a prefix sum over the wave */
    unsigned incl = extra;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned t = __shfl(incl, lane >= o ? lane - o : lane, 64);
      if (lane >= o) incl += t;
    }
    const int tot = (int)__shfl(incl, 63, 64);
    if (tot == 0) continue;
    for (unsigned m = 0; m < extra; ++m) L.tq[incl - extra + m] = (f << 5) | (i0 + (int)(m + 1) * HCS_TX_ROWS);
    __syncthreads();
    for (int t = lane; t < tot; t += 64) {
      const int ff = L.tq[t] >> 5, r = L.tq[t] & 31;
      const int l = fan_entry(L, cn, ff);
      const int k = ff - L.fan0[l];
      const double* P = poly + (size_t)L.q[l] * HCS_PREC;
      TxFan c;
      if (fan_setup(P, min((int)P[0], 8), k, g, x, y, c))
        fan_rows<PASS>(L, c, g, x, y, r, min(c.i1, r + HCS_TX_ROWS - 1), c0, l, k);
    }
    __syncthreads();
  }
}

__device__ inline double sample_cell_scatter(const double* __restrict__ poly, const int64_t* __restrict__ key,
                                             const TaxelGridDev& g, unsigned base, unsigned n,
                                             const int* __restrict__ list, int x, int y) {
  __shared__ TxScatter L;
  const int lane = threadIdx.x & 63;
  const int S = g.S, nr = S * S;
  for (int s = lane; s < nr; s += 64) L.tmin[s] = 0x7ff0000000000000ull, L.smin[s] = ~0ull, L.flag[s] = 0, L.own[s] = -1;
  auto stage = [&](unsigned c0, unsigned cn) {
    __syncthreads();
    /* one entry per lane (cn <= 64): its polygon, key and fan count, and
     * the first fan of each entry by a prefix sum over the wave */
    unsigned nf = 0;
    if ((unsigned)lane < cn) {
      const int q = list[base + c0 + lane];
      L.q[lane] = q;
      L.kq[lane] = pkey(key, q);
      nf = (unsigned)min((int)poly[(size_t)q * HCS_PREC], 8);
    }
    unsigned incl = nf;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned t = __shfl(incl, lane >= o ? lane - o : lane, 64);
      if (lane >= o) incl += t;
    }
    if ((unsigned)lane < cn) L.fan0[lane] = (int)(incl - nf);
    if ((unsigned)lane == cn - 1) L.fan0[cn] = (int)incl;
    __syncthreads();
  };
  const bool one_chunk = n <= HCS_TX_CHUNK;
  for (unsigned c0 = 0; c0 < n; c0 += HCS_TX_CHUNK) {
    const unsigned cn = (n - c0) < HCS_TX_CHUNK ? (n - c0) : HCS_TX_CHUNK;
    stage(c0, cn);
    scatter_chunk<0>(L, poly, g, x, y, cn, c0);
    __syncthreads();
  }
  /* settle the rays whose owner holds the least t; flag the rest for the
   * second pass */
  bool again = false;
  for (int s = lane; s < nr; s += 64) {
    if (L.tmin[s] == 0x7ff0000000000000ull) continue;
    bool settled = false;
    if (!L.flag[s]) {
      const int pos = L.own[s] >> 3, bk = L.own[s] & 7;
      const double* P = poly + (size_t)(one_chunk ? L.q[pos] : list[base + pos]) * HCS_PREC;
      const double t = fan_hit_at(P, min((int)P[0], 8), bk, g, x, y, s / g.S, s % g.S);
      settled = t > 0.0 && (unsigned long long)__double_as_longlong(t) == L.tmin[s];
    }
    if (!settled) L.flag[s] = 1, L.own[s] = -1, again = true;
  }
  if (__any(again)) {
    __syncthreads();
    for (unsigned c0 = 0; c0 < n; c0 += HCS_TX_CHUNK) {
      const unsigned cn = (n - c0) < HCS_TX_CHUNK ? (n - c0) : HCS_TX_CHUNK;
      if (!one_chunk) stage(c0, cn);
      scatter_chunk<1>(L, poly, g, x, y, cn, c0);
      __syncthreads();
    }
  }
  __syncthreads();
  /* the chosen triangle of every ray, evaluated as in the ray-major loop */
  const double rSR = g.res / S;
  const double rmean = 1. / (S * S);
  const double di_factor = g.res / S;
  const double sub_halfwidth = di_factor / 2.0 - g.res / 2.0;
  const double max_dist = 1.41421356237 * g.res / 2.0;
  double rsigma_squared = 0, rtukey = 0;
  if (g.window == 1) rsigma_squared = 0.5 / (g.sigma * g.sigma);
  if (g.window == 2) rtukey = 1. / g.sigma * S * S;
  const double PI = 3.14159265358979323846;
  double acc = 0.0;
  for (int s0 = 0; s0 < nr; s0 += 64) {
    const int s = s0 + lane;
    const bool active = s < nr;
    const unsigned long long tb = active ? L.tmin[s] : 0x7ff0000000000000ull;
    const double best_t = __longlong_as_double((long long)tb);
    /* a settled ray's owner, else the least key of the second pass */
    const int ow = active ? L.own[s] : -1;
    const unsigned long long sk = active ? L.smin[s] : ~0ull;
    const long long kq = (long long)(sk >> 3);
    const int bk = ow >= 0 ? (ow & 7) : (int)(sk & 7);
    /* the polygon: the settled ray's list position, else a search of the
     * list for the key (restaged per chunk when the list has more than one,
     * and only when some lane needs it) */
    const int pos = ow >= 0 ? ow >> 3 : -1;
    int bq = pos < 0 ? -1 : (one_chunk ? L.q[pos] : list[base + pos]);
    const bool search = tb != 0x7ff0000000000000ull && pos < 0;
    if (__any(search))
      for (unsigned c0 = 0; c0 < n; c0 += HCS_TX_CHUNK) {
        const unsigned cn = (n - c0) < HCS_TX_CHUNK ? (n - c0) : HCS_TX_CHUNK;
        if (!one_chunk) stage(c0, cn);
        if (search)
          for (unsigned l = 0; l < cn; ++l)
            if (L.kq[l] == kq) bq = L.q[l];
      }
    if (!active) continue;
    const int i = s / S, j = s % S;
    if (bq >= 0 && best_t < 1.5 * g.zs && best_t > 0.0) {
      const double loc[3] = {-g.xs + x * g.res + i * rSR + 0.5 * rSR, -g.ys + y * g.res + j * rSR + 0.5 * rSR,
                             1.5 * g.zs};
      double pt[3];
      xform(g.X, loc, pt);
      const double nrm[3] = {g.X[2], g.X[5], g.X[8]};
      double O[3], D[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) O[c] = pt[c] + nrm[c] * 1e-8, D[c] = -nrm[c];
      const size_t pb = (size_t)bq * HCS_PREC;
      const int nv = (int)poly[pb];
      const int k1 = (bk + 1 == nv) ? 0 : bk + 1;
      double v0[3], v1[3], v2[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        v0[c] = poly[pb + 1 + 3 * bk + c];
        v1[c] = poly[pb + 1 + 3 * k1 + c];
        v2[c] = poly[pb + 33 + c];
      }
      double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
      double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
      double h[3];
      cross3(D, e2, h);
      const double f = 1.0 / dot3(e1, h);
      double sv[3] = {O[0] - v0[0], O[1] - v0[1], O[2] - v0[2]};
      const double u = f * dot3(sv, h);
      double qv[3];
      cross3(sv, e1, qv);
      const double v = f * dot3(D, qv);
      const double pa = poly[pb + 25 + bk], pbb = poly[pb + 25 + k1], pc = poly[pb + 36];
      const double b0 = 1.0 - u - v;
      const double raw = ((b0 * pa + u * pbb) + v * pc) * rmean;
      double weight = 1.0;
      if (g.window == 1) {
        const double dist = hypot(di_factor * i + sub_halfwidth, di_factor * j + sub_halfwidth) / max_dist;
        weight = exp(-(dist * dist) * rsigma_squared);
      } else if (g.window == 2) {
        const int ni = S / 2 - abs(S / 2 - i), nj = S / 2 - abs(S / 2 - j);
        if (ni <= g.sigma * S / 2) weight *= 0.5 * (1. - cos(2. * PI * ni * rtukey));
        if (nj <= g.sigma * S / 2) weight *= 0.5 * (1. - cos(2. * PI * nj * rtukey));
      } else if (g.window == 3) {
        const double inv = 1 - hypot(di_factor * i + sub_halfwidth, di_factor * j + sub_halfwidth) / max_dist;
        weight = inv * inv;
      }
      acc = acc + weight * raw;
    }
  }
  double v = acc;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return __shfl(v, 0, 64);
}

}  // namespace hcs
