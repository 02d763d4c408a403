/*
 * hcs_taxels.hip -- flat tactile sensor sampling on the GPU, FP64.
 *
 * Restates FlatTactileSensor::bvh_update
 * (mujoco_contact_surface_sensors/src/flat_tactile_sensor.cpp:262-402): per
 * taxel (x, y) and sub-sample (i, j) a ray from
 *   R (-xs + x res + (i+1/2) res/S, -ys + y res + (j+1/2) res/S, 1.5 zs) + t
 * offset by 1e-8 along the sensor normal n, direction -n, nearest hit over
 * the kTriangle fans of every contact polygon of the pairs touching the
 * sensor geometry (Moeller-Trumbore as in
 * mujoco_contact_surface_sensors/src/bvh.cpp:49-74, |det| < 1e-10 rejected),
 * accepted when 0 < t < 1.5 zs; value Evaluate(tri, (1-u-v, u, v)) / S^2
 * times the window weight, summed per taxel.
 *
 * Instead of the reference's per-update float32 BVH + TLAS over the contact
 * triangles, all rays are parallel, so the polygons are binned once into the
 * taxel grid by their footprint in the sensor plane; one workgroup per taxel
 * then tests only its bin.  Ties in t are broken by (polygon key, fan index),
 * the canonical polygon order.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hcs_device.h"
#include "hcs_internal.h"
#include "hcs_prim.h"

#pragma clang fp contract(off)

using namespace hcs;

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

__global__ void k_tx_bin(const double* __restrict__ poly, const int64_t* __restrict__ key, int64_t np,
                         const unsigned char* __restrict__ env_mask, TaxelGridDev g, unsigned* __restrict__ cnt,
                         const unsigned* __restrict__ off, unsigned* __restrict__ cursor, int* __restrict__ list) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= np) return;
  if (!env_mask[key[3 * q]]) return;
  const double* P = poly + q * HCS_PREC;
  int x0, x1, y0, y1;
  footprint(P, g, x0, x1, y0, y1);
  for (int y = y0; y <= y1; ++y)
    for (int x = x0; x <= x1; ++x) {
      const int cell = x + g.cx * y;
      if (!off) {
        atomicAdd(cnt + cell, 1u);
      } else {
        const unsigned slot = atomicAdd(cursor + cell, 1u);
        list[off[cell] + slot] = (int)q;
      }
    }
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

/* Polygons of one chunk of a taxel's list, staged in LDS for its rays: per
 * fan triangle (v_k, v_k+1, C) the sample-index ranges [i0, i1] x [j0, j1]
 * (i in the low, i1 / j1 in the high 16 bits) outside of which no ray of the
 * taxel can hit it, so a ray runs the Moeller-Trumbore test only on the fans
 * whose range holds it. */
#define HCS_TX_CHUNK 128
struct TxStage {
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
__device__ inline void stage_fans(const double* P, const TaxelGridDev& g, int x, int y, uint2* rng) {
  const int S = g.S;
  const double rSR = g.res / S;
  const int nv = min((int)P[0], 8);
  const double Dv[3] = {-g.X[2], -g.X[5], -g.X[8]};
  double px[9], py[9], pz[9];
  for (int k = 0; k <= nv && k < 9; ++k) {
    const double* v = k < nv ? P + 1 + 3 * k : P + 33;
    const double d[3] = {v[0] - g.X[9], v[1] - g.X[10], v[2] - g.X[11]};
    const int o = k < nv ? k : 8;
    px[o] = (g.X[0] * d[0] + g.X[3] * d[1]) + g.X[6] * d[2];
    py[o] = (g.X[1] * d[0] + g.X[4] * d[1]) + g.X[7] * d[2];
    pz[o] = (g.X[2] * d[0] + g.X[5] * d[1]) + g.X[8] * d[2];
  }
  const double ox = -g.xs + x * g.res + 0.5 * rSR, oy = -g.ys + y * g.res + 0.5 * rSR, oz = 1.5 * g.zs;
  for (int k = 0; k < nv; ++k) {
    const int k1 = (k + 1 == nv) ? 0 : k + 1;
    const double* v0 = P + 1 + 3 * k;
    const double* v1 = P + 1 + 3 * k1;
    const double* v2 = P + 33;
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
      const int id[3] = {k, k1, 8};
      double lx = 1e300, hx = -1e300, ly = 1e300, hy = -1e300, gap = 0.0;
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        lx = fmin(lx, px[id[t]]), hx = fmax(hx, px[id[t]]);
        ly = fmin(ly, py[id[t]]), hy = fmax(hy, py[id[t]]);
        gap = fmax(gap, fabs(oz - pz[id[t]]));
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

/* The value of taxel (x, y): the block's 256 threads sample it; the block's
 * candidate polygons are list[base .. base + n). */
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
  /* rays in rounds of blockDim.x (every thread runs every round, so the
   * staging barriers are reached by all) */
  for (int s0 = 0; s0 < S * S; s0 += blockDim.x) {
    const int s = s0 + (int)threadIdx.x;
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
      __syncthreads();
      for (unsigned l = threadIdx.x; l < cn; l += blockDim.x) {
        const int q = list[base + c0 + l];
        const double* P = poly + (size_t)q * HCS_PREC;
        L.q[l] = q;
        L.nv[l] = min((int)P[0], 8);
        L.kq[l] = pkey(key, q);
        stage_fans(P, g, x, y, L.rng[l]);
      }
      __syncthreads();
      if (!active) continue;
      for (unsigned l = 0; l < cn; ++l) {
        const int nv = L.nv[l];
        const int q = L.q[l];
        const int64_t kq = L.kq[l];
        const double* P = poly + (size_t)q * HCS_PREC;
        for (int k = 0; k < nv; ++k) {
          const uint2 r = L.rng[l][k];
          if (i < (int)(r.x & 0xffffu) || i > (int)(r.x >> 16) || j < (int)(r.y & 0xffffu) || j > (int)(r.y >> 16))
            continue;
          const int k1 = (k + 1 == nv) ? 0 : k + 1;
          const double* v0 = P + 1 + 3 * k;
          const double* v1 = P + 1 + 3 * k1;
          const double* v2 = P + 33;
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
  __shared__ double red[256];
  red[threadIdx.x] = ptan ? tangent(acc) : val(acc);
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

template <class T>
__global__ __launch_bounds__(256) void k_tx_sample(const double* __restrict__ poly, const double* __restrict__ ptan,
                                                   const int64_t* __restrict__ key, TaxelGridDev g,
                                                   const double* __restrict__ Xd, const unsigned* __restrict__ off,
                                                   const unsigned* __restrict__ cnt, const int* __restrict__ list,
                                                   double* __restrict__ out) {
  const int cell = blockIdx.x;
  const int x = cell % g.cx, y = cell / g.cx;
  const double v = sample_cell<T>(poly, ptan, key, g, Xd, off[cell], cnt[cell], list, x, y);
  if (threadIdx.x == 0) out[x + g.cx * y] = v;
}

/* Bins polygons of the masked pairs into the taxel grid and samples it.
 * ptan/Xd select the tangent pass (nullptr: values). */
int run_taxels(Ctx* c, const TaxelGridDev& g, const unsigned char* d_mask, const double* poly, const double* ptan,
               const int64_t* key, int64_t np, const double* Xd, double* d_out) {
  const int ncell = g.cx * g.cy;
  int rc;
  if ((rc = ensure(c, c->d_tx_cnt, sizeof(unsigned) * 2 * ncell)) ||
      (rc = ensure(c, c->d_tx_off, sizeof(unsigned) * (ncell + 1))))
    return rc;
  unsigned* cnt = (unsigned*)c->d_tx_cnt.ptr;
  unsigned* cursor = cnt + ncell;
  unsigned* off = (unsigned*)c->d_tx_off.ptr;
  rc = hip_check(c, hipMemsetAsync(cnt, 0, sizeof(unsigned) * 2 * ncell, c->stream), "memset");
  if (rc) return rc;
  const int64_t nb = (np + 255) / 256;
  if (np > 0)
    hipLaunchKernelGGL(k_tx_bin, dim3((unsigned)nb), dim3(256), 0, c->stream, poly, key, np, d_mask, g, cnt,
                       (const unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr);
  std::vector<unsigned> hc(ncell), ho(ncell + 1);
  rc = hip_check(c, hipMemcpyAsync(hc.data(), cnt, sizeof(unsigned) * ncell, hipMemcpyDeviceToHost, c->stream),
                 "download counts");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  ho[0] = 0;
  for (int i = 0; i < ncell; ++i) ho[i + 1] = ho[i] + hc[i];
  if ((rc = ensure(c, c->d_tx_list, sizeof(int) * std::max<unsigned>(1, ho[ncell])))) return rc;
  rc = hip_check(c, hipMemcpyAsync(off, ho.data(), sizeof(unsigned) * (ncell + 1), hipMemcpyHostToDevice, c->stream),
                 "upload offsets");
  if (rc) return rc;
  if (np > 0)
    hipLaunchKernelGGL(k_tx_bin, dim3((unsigned)nb), dim3(256), 0, c->stream, poly, key, np, d_mask, g, cnt,
                       (const unsigned*)off, cursor, (int*)c->d_tx_list.ptr);
  if (ptan)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tx_sample<Dual>), dim3(ncell), dim3(256), 0, c->stream, poly, ptan, key, g,
                       Xd, (const unsigned*)off, (const unsigned*)cnt, (const int*)c->d_tx_list.ptr, d_out);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tx_sample<double>), dim3(ncell), dim3(256), 0, c->stream, poly,
                       (const double*)nullptr, key, g, (const double*)nullptr, (const unsigned*)off,
                       (const unsigned*)cnt, (const int*)c->d_tx_list.ptr, d_out);
  return hip_check(c, hipGetLastError(), "taxel kernels");
}

int taxel_grid_dev(const hcs_taxel_grid& G, TaxelGridDev& g) {
  int32_t cx, cy;
  if (hcs_taxel_dims(&G, &cx, &cy) != HCS_OK || G.sampling_resolution <= 0 || G.window < 0 || G.window > 3)
    return HCS_ERR_ARG;
  std::memcpy(g.X, G.X_W_sensor, sizeof g.X);
  g.xs = G.half_size[0], g.ys = G.half_size[1], g.zs = G.half_size[2], g.res = G.resolution, g.sigma = G.sigma;
  g.S = G.sampling_resolution, g.window = G.window, g.cx = cx, g.cy = cy;
  return HCS_OK;
}

/* ===================================================================
 * Batched sensors: every sensor of a batch in one pass per stage, inside
 * the step's graph (no host round trip; list sizes are checked after the
 * step, which is re-run when a buffer overflowed).
 * =================================================================== */

/* Sensor poses: the 12 doubles (rotation, position) at src[s] in the state. */
__global__ void k_txb_pose(TaxelGridDev* __restrict__ g, const int* __restrict__ src, const double* __restrict__ state,
                           int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 12 * n) return;
  const int s = t / 12, k = t % 12;
  g[s].X[k] = state[src[s] + k];
}

/* Polygons strided over the grid; each polygon goes to every sensor of its
 * pair, into the taxels its footprint covers.  FILL = 0 counts per taxel;
 * FILL = 1 writes the polygon into the taxel's list slice. */
template <int FILL>
__global__ __launch_bounds__(256) void k_txb_bin(const double* __restrict__ poly, const int64_t* __restrict__ key,
                                                 const unsigned long long* __restrict__ poly_cnt, int64_t poly_cap,
                                                 const int* __restrict__ psoff, const int* __restrict__ psens,
                                                 const TaxelGridDev* __restrict__ grids,
                                                 const int* __restrict__ cell_base, unsigned* __restrict__ cnt,
                                                 const unsigned* __restrict__ off, unsigned* __restrict__ cursor,
                                                 int* __restrict__ list, int64_t list_cap) {
  const int64_t np = min((int64_t)*poly_cnt, poly_cap);
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    const int e = (int)key[3 * q];
    const double* P = poly + q * HCS_PREC;
    for (int k = psoff[e]; k < psoff[e + 1]; ++k) {
      const int s = psens[k];
      const TaxelGridDev& g = grids[s];
      int x0, x1, y0, y1;
      footprint(P, g, x0, x1, y0, y1);
      for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
          const int cell = cell_base[s] + x + g.cx * y;
          if (!FILL) {
            atomicAdd(cnt + cell, 1u);
          } else {
            const int64_t slot = (int64_t)off[cell] + atomicAdd(cursor + cell, 1u);
            if (slot < list_cap) list[slot] = (int)q;
          }
        }
    }
  }
}

/* One workgroup per taxel of any sensor. */
__global__ __launch_bounds__(256) void k_txb_sample(const double* __restrict__ poly, const int64_t* __restrict__ key,
                                                    const TaxelGridDev* __restrict__ grids,
                                                    const int* __restrict__ cell_base,
                                                    const int* __restrict__ cell_sensor,
                                                    const unsigned* __restrict__ off,
                                                    const unsigned* __restrict__ cnt, const int* __restrict__ list,
                                                    int64_t list_cap, double* __restrict__ out) {
  const int cell = blockIdx.x;
  const int s = cell_sensor[cell];
  const unsigned n = cnt[cell], base = off[cell];
  /* an overflowed list is re-run after the step with a larger buffer */
  if (n == 0 || (int64_t)base + n > list_cap) {
    if (threadIdx.x == 0) out[cell] = 0.0;
    return;
  }
  const TaxelGridDev g = grids[s];
  const int local = cell - cell_base[s];
  const double v = sample_cell<double>(poly, nullptr, key, g, nullptr, base, n, list, local % g.cx, local / g.cx);
  if (threadIdx.x == 0) out[cell] = v;
}

int batch_taxels_setup(Ctx* c, BatchTaxels& t, int npairs, const int32_t* geoms, int n, const int32_t* pair,
                       const hcs_taxel_grid* grids) {
  std::vector<TaxelGridDev> g(std::max(1, n));
  std::vector<int> src(std::max(1, n)), base(n + 1, 0), psoff(npairs + 1, 0), psens(std::max(1, n));
  std::vector<unsigned char> mask(std::max(1, npairs), 0);
  for (int s = 0; s < n; ++s) {
    const int e = pair[s];
    if (e < 0 || e >= npairs) return set_err(c, HCS_ERR_ARG, "sensor " + std::to_string(s) + ": pair out of range");
    if (taxel_grid_dev(grids[s], g[s]) != HCS_OK)
      return set_err(c, HCS_ERR_ARG, "sensor " + std::to_string(s) + ": invalid taxel grid");
    /* the pose of the pair's geometry with the sensor's mesh handle (B when
     * both are) */
    int side;
    if (geoms[2 * e + 1] == grids[s].geom) side = 1;
    else if (geoms[2 * e] == grids[s].geom) side = 0;
    else return set_err(c, HCS_ERR_ARG, "sensor " + std::to_string(s) + ": its geom is not in its pair");
    src[s] = HCS_STATE * e + 12 * side;
    if ((int64_t)base[s] + (int64_t)g[s].cx * g[s].cy > (1 << 30))
      return set_err(c, HCS_ERR_ARG, "too many taxels in the batch's sensors");
    base[s + 1] = base[s] + g[s].cx * g[s].cy;
    ++psoff[e + 1];
    mask[e] = 1;
  }
  for (int e = 0; e < npairs; ++e) psoff[e + 1] += psoff[e];
  {
    std::vector<int> fill(psoff.begin(), psoff.end() - 1);
    for (int s = 0; s < n; ++s) psens[fill[pair[s]]++] = s;
  }
  const int ncell = base[n];
  std::vector<int> cs(std::max(1, ncell));
  for (int s = 0; s < n; ++s)
    for (int k = base[s]; k < base[s + 1]; ++k) cs[k] = s;
  int rc;
  if ((rc = ensure(c, t.grid, sizeof(TaxelGridDev) * g.size())) || (rc = ensure(c, t.src, sizeof(int) * src.size())) ||
      (rc = ensure(c, t.cell_base, sizeof(int) * base.size())) ||
      (rc = ensure(c, t.cell_sensor, sizeof(int) * cs.size())) || (rc = ensure(c, t.psoff, sizeof(int) * psoff.size())) ||
      (rc = ensure(c, t.psens, sizeof(int) * psens.size())) || (rc = ensure(c, t.mask, mask.size())) ||
      (rc = ensure(c, t.poly_cnt, 64)) || (rc = ensure(c, t.cnt, sizeof(unsigned) * (ncell + 1))) ||
      (rc = ensure(c, t.cursor, sizeof(unsigned) * std::max(1, ncell))) ||
      (rc = ensure(c, t.off, sizeof(unsigned) * (ncell + 1))) || (rc = ensure(c, t.out, sizeof(double) * std::max(1, ncell))))
    return rc;
  const size_t sb = scan_temp_bytes((long long)ncell + 1);
  if ((rc = ensure(c, t.scan_tmp, std::max<size_t>(sb, 8)))) return rc;
  t.scan_bytes = sb;
  struct Up {
    DBuf& d;
    const void* h;
    size_t bytes;
  } ups[] = {{t.grid, g.data(), sizeof(TaxelGridDev) * g.size()}, {t.src, src.data(), sizeof(int) * src.size()},
             {t.cell_base, base.data(), sizeof(int) * base.size()}, {t.cell_sensor, cs.data(), sizeof(int) * cs.size()},
             {t.psoff, psoff.data(), sizeof(int) * psoff.size()}, {t.psens, psens.data(), sizeof(int) * psens.size()},
             {t.mask, mask.data(), mask.size()}};
  for (const Up& u : ups)
    if ((rc = hip_check(c, hipMemcpyAsync(u.d.ptr, u.h, u.bytes, hipMemcpyHostToDevice, c->stream), "upload sensors")))
      return rc;
  t.n = n, t.ncell = ncell;
  t.bin_grid = std::max(1, c->num_cu * 4);
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

int batch_taxels_reserve(Ctx* c, BatchTaxels& t, int64_t poly_cap, int64_t list_cap) {
  poly_cap = std::max<int64_t>(poly_cap, 4096);
  list_cap = std::max<int64_t>(list_cap, 4096);
  /* list offsets are 32-bit */
  if (list_cap > 0xffffffffLL) return set_err(c, HCS_ERR_CAPACITY, "taxel polygon lists exceed 2^32 entries");
  int rc;
  if ((rc = ensure(c, t.poly_rec, sizeof(double) * HCS_PREC * poly_cap)) ||
      (rc = ensure(c, t.poly_key, sizeof(int64_t) * 3 * poly_cap)) || (rc = ensure(c, t.list, sizeof(int) * list_cap)))
    return rc;
  t.poly_cap = poly_cap, t.list_cap = list_cap;
  return HCS_OK;
}

void batch_taxels_enqueue(Ctx* c, BatchTaxels& t, const double* state) {
  if (t.n == 0) return;
  hipStream_t s = c->stream;
  TaxelGridDev* g = (TaxelGridDev*)t.grid.ptr;
  hipLaunchKernelGGL(k_txb_pose, dim3((12 * t.n + 255) / 256), dim3(256), 0, s, g, (const int*)t.src.ptr, state, t.n);
  launch_fill_u32(s, (unsigned*)t.cnt.ptr, 0u, t.ncell + 1);
  launch_fill_u32(s, (unsigned*)t.cursor.ptr, 0u, t.ncell);
  const unsigned long long* pc = (const unsigned long long*)t.poly_cnt.ptr;
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_txb_bin<0>), dim3(t.bin_grid), dim3(256), 0, s, (const double*)t.poly_rec.ptr,
                     (const int64_t*)t.poly_key.ptr, pc, t.poly_cap, (const int*)t.psoff.ptr, (const int*)t.psens.ptr,
                     (const TaxelGridDev*)g, (const int*)t.cell_base.ptr, (unsigned*)t.cnt.ptr,
                     (const unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr, (int64_t)0);
  (void)exclusive_scan_u32(s, (const unsigned*)t.cnt.ptr, (unsigned*)t.off.ptr, (long long)t.ncell + 1,
                           t.scan_tmp.ptr, t.scan_bytes);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_txb_bin<1>), dim3(t.bin_grid), dim3(256), 0, s, (const double*)t.poly_rec.ptr,
                     (const int64_t*)t.poly_key.ptr, pc, t.poly_cap, (const int*)t.psoff.ptr, (const int*)t.psens.ptr,
                     (const TaxelGridDev*)g, (const int*)t.cell_base.ptr, (unsigned*)t.cnt.ptr,
                     (const unsigned*)t.off.ptr, (unsigned*)t.cursor.ptr, (int*)t.list.ptr, t.list_cap);
  if (t.ncell > 0)
    hipLaunchKernelGGL(k_txb_sample, dim3(t.ncell), dim3(256), 0, s, (const double*)t.poly_rec.ptr,
                       (const int64_t*)t.poly_key.ptr, (const TaxelGridDev*)g, (const int*)t.cell_base.ptr,
                       (const int*)t.cell_sensor.ptr, (const unsigned*)t.off.ptr, (const unsigned*)t.cnt.ptr,
                       (const int*)t.list.ptr, t.list_cap, (double*)t.out.ptr);
}

int batch_taxels_demand(Ctx* c, BatchTaxels& t, int64_t& npoly, int64_t& nlist) {
  npoly = nlist = 0;
  if (t.n == 0) return HCS_OK;
  unsigned long long np = 0;
  unsigned nl = 0;
  int rc = hip_check(c, hipMemcpyAsync(&np, t.poly_cnt.ptr, 8, hipMemcpyDeviceToHost, c->stream), "dl");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(&nl, (const unsigned*)t.off.ptr + t.ncell, 4, hipMemcpyDeviceToHost, c->stream),
                   "dl");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  npoly = (int64_t)np, nlist = (int64_t)nl;
  return rc;
}

void batch_taxels_free(BatchTaxels& t) {
  for (DBuf* b : {&t.grid, &t.src, &t.cell_base, &t.cell_sensor, &t.psoff, &t.psens, &t.mask, &t.poly_rec, &t.poly_key,
                  &t.poly_cnt, &t.cnt, &t.cursor, &t.off, &t.scan_tmp, &t.list, &t.out})
    if (b->ptr) (void)hipFree(b->ptr), b->ptr = nullptr, b->bytes = 0;
  t = BatchTaxels();
}

}  // namespace hcs

extern "C" {

int hcs_taxel_dims(const hcs_taxel_grid* g, int32_t* cx, int32_t* cy) {
  if (!g || !cx || !cy || !(g->resolution > 0)) return HCS_ERR_ARG;
  /* ::floorl(2 * xs / resolution + 0.1), flat_tactile_sensor.cpp:196-197 */
  *cx = (int32_t)std::floor(2 * (long double)g->half_size[0] / g->resolution + 0.1L);
  *cy = (int32_t)std::floor(2 * (long double)g->half_size[1] / g->resolution + 0.1L);
  if (*cx <= 0 || *cy <= 0) return HCS_ERR_ARG;
  return HCS_OK;
}

int hcs_taxels(hcs_ctx* h, int32_t nsensors, const hcs_taxel_grid* grids, double* out) {
  if (!h || nsensors < 0 || (nsensors > 0 && (!grids || !out))) return HCS_ERR_ARG;
  Ctx* c = &h->c;
  if (!c->last_valid) return set_err(c, HCS_ERR_STATE, "hcs_taxels needs a contact step run with keep_polygons = 1");
  int rc = hip_check(c, hipSetDevice(c->device), "hipSetDevice");
  if (rc) return rc;
  const int npairs = (int)c->last_pairs.size();
  size_t off_out = 0;
  for (int s = 0; s < nsensors; ++s) {
    const hcs_taxel_grid& G = grids[s];
    TaxelGridDev g;
    if (taxel_grid_dev(G, g) != HCS_OK) return set_err(c, HCS_ERR_ARG, "invalid taxel grid");
    const int ncell = g.cx * g.cy;
    std::vector<unsigned char> mask(std::max(1, npairs), 0);
    for (int e = 0; e < npairs; ++e)
      mask[e] = (c->last_geoms[2 * e] == G.geom || c->last_geoms[2 * e + 1] == G.geom) ? 1 : 0;
    if ((rc = ensure(c, c->d_tx_grid, mask.size())) || (rc = ensure(c, c->d_tx_out, sizeof(double) * ncell)))
      return rc;
    rc = hip_check(c, hipMemcpyAsync(c->d_tx_grid.ptr, mask.data(), mask.size(), hipMemcpyHostToDevice, c->stream),
                   "upload mask");
    if (!rc)
      rc = run_taxels(c, g, (const unsigned char*)c->d_tx_grid.ptr, (const double*)c->d_poly_rec.ptr, nullptr,
                      (const int64_t*)c->d_poly_key.ptr, c->last_npoly, nullptr, (double*)c->d_tx_out.ptr);
    if (!rc)
      rc = hip_check(c, hipMemcpyAsync(out + off_out, c->d_tx_out.ptr, sizeof(double) * ncell, hipMemcpyDeviceToHost,
                                       c->stream),
                     "download taxels");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
    if (rc) return rc;
    off_out += ncell;
  }
  return HCS_OK;
}

}  // extern "C"
