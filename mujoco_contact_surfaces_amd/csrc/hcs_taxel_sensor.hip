/*
 * hcs_taxel_sensor.hip -- the taxel sensor on the GPU, FP64.
 *
 * Restates TaxelSensor::internal_update
 * (mujoco_contact_surface_sensors/src/taxel_sensor.cpp:158-478) over the
 * contact polygons kept by the last contact step:
 *   1. the polygons of the pairs touching the sensor geometry, in key order
 *      (pair, element of A, element of B) -- two stable radix passes -- and
 *      their kTriangle fans (v_k, v_{k+1}, C), the surface's triangles;
 *   2. samples on every triangle: DEFAULT (:174-203), the reference's loops
 *      a, b += 1/st from 0 while <= 1 with st from |v1 - v0| and |v2 - v0|
 *      (both st1 and st2 from v2 - v0), one thread per triangle after a
 *      count and a scan; or AREA_IMPORTANCE (:205-246), a single sequential
 *      stream of libstdc++'s default_random_engine through
 *      uniform_real_distribution<double>(0, 1), one thread;
 *   3. one wavefront per taxel: dist = (-2 t.p + |t|^2) + |p|^2 to every
 *      sample; CLOSEST (:285-319) takes the first nearest sample, and within
 *      the margin writes its pressure when visualize is set and |p| > 1e-6,
 *      else 0; WEIGHTED and MEAN fall through to SQUARED (:320-457, no
 *      break): sum of max(0, margin - sqrt(dist))^2 |p| over the samples in
 *      the margin, times sample_resolution.  A taxel no sample reaches keeps
 *      its value; with no sample at all every taxel is 0 (:458-477).
 * Sums over samples run lane-strided in sample order and then in a fixed
 * butterfly, so results are deterministic.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hcs_device.h"
#include "hcs_internal.h"
#include "hcs_prim.h"
#include "hcs_query.h"

#pragma clang fp contract(off)

using namespace hcs;

namespace hcs {

#define HCS_TS_TRI 12 /* fan triangle: v0, v1, v2 (W), e0, e1, e2 */

/* selected polygons get their key, others sort last */
__global__ void k_ts_keys(const int64_t* __restrict__ key, int64_t np, const unsigned char* __restrict__ mask,
                          unsigned long long* __restrict__ k64, unsigned* __restrict__ kp, int* __restrict__ idx,
                          unsigned* __restrict__ cnt) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int64_t pair = key[3 * i], a = key[3 * i + 1], b = key[3 * i + 2];
  const bool on = mask[pair] != 0;
  k64[i] = ((unsigned long long)(unsigned)a << 32) | (unsigned long long)(unsigned)b;
  kp[i] = on ? (unsigned)pair : 0xffffffffu;
  idx[i] = (int)i;
  if (on) atomicAdd(cnt, 1u);
}

__global__ void k_ts_gather_u32(const unsigned* __restrict__ src, const int* __restrict__ idx, int64_t n,
                                unsigned* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_ts_nfan(const double* __restrict__ poly, const int* __restrict__ order, int n,
                          unsigned* __restrict__ nt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) nt[i] = (unsigned)poly[HCS_PREC * (size_t)order[i]];
  if (i == n) nt[i] = 0;
}

__global__ void k_ts_fans(const double* __restrict__ poly, const int64_t* __restrict__ key,
                          const int* __restrict__ order, int n, const unsigned* __restrict__ off,
                          double* __restrict__ tris, int* __restrict__ surf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* P = poly + HCS_PREC * (size_t)order[i];
  const int nv = (int)P[0];
  for (int k = 0; k < nv; ++k) {
    const int k2 = k + 1 < nv ? k + 1 : 0;
    double* T = tris + HCS_TS_TRI * (size_t)(off[i] + k);
#pragma unroll
    for (int c = 0; c < 3; ++c) T[c] = P[1 + 3 * k + c], T[3 + c] = P[1 + 3 * k2 + c], T[6 + c] = P[33 + c];
    T[9] = P[25 + k], T[10] = P[25 + k2], T[11] = P[36];
    surf[off[i] + k] = (int)key[3 * (size_t)order[i]];
  }
}

__device__ inline double ts_len(const double* a, const double* b) {
  const double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  return sqrt(dot3(d, d));
}

/* the reference's loop counts of one triangle */
__device__ inline void ts_steps(const double* T, double res, int& st, int& st2) {
  const int st0 = (int)(ts_len(T + 3, T) / res) + 1;
  const int st1 = (int)(ts_len(T + 6, T) / res) + 1;
  st2 = (int)(ts_len(T + 6, T) / res) + 1;
  st = st0 > st1 ? st0 : st1;
}

__device__ inline void ts_sample(const double* T, double b0, double b1, double b2, double* p, double* e) {
#pragma unroll
  for (int c = 0; c < 3; ++c) p[c] = (b0 * T[c] + b1 * T[3 + c]) + b2 * T[6 + c];
  *e = (b0 * T[9] + b1 * T[10]) + b2 * T[11];
}

__global__ void k_ts_count(const double* __restrict__ tris, int64_t nt, double res, unsigned* __restrict__ cnt) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t > nt) return;
  if (t == nt) {
    cnt[t] = 0;
    return;
  }
  int st, st2;
  ts_steps(tris + HCS_TS_TRI * t, res, st, st2);
  unsigned na = 0, nb = 0;
  for (double a = 0; a <= 1; a += 1. / st) ++na;
  for (double b = 0; b <= 1; b += 1. / st2) ++nb;
  cnt[t] = na * nb;
}

__global__ void k_ts_gen(const double* __restrict__ tris, int64_t nt, double res, const unsigned* __restrict__ off,
                         double* __restrict__ P, double* __restrict__ E) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const double* T = tris + HCS_TS_TRI * t;
  int st, st2;
  ts_steps(T, res, st, st2);
  size_t j = off[t];
  for (double a = 0; a <= 1; a += 1. / st)
    for (double b = 0; b <= 1; b += 1. / st2) {
      ts_sample(T, a, (1 - a) * (1 - b), (1 - a) * b, P + 3 * j, E + j);
      ++j;
    }
}

/* libstdc++ default_random_engine (minstd_rand0, seed 1) through
 * generate_canonical<double, 53>: two draws, sum = (g1 - 1) + (g2 - 1) r,
 * r = 2^31 - 2, over r^2; a result >= 1 becomes the largest double below 1.
 * The long double products of the library are exact here, so double
 * arithmetic gives the same bits. */
struct TsRng {
  unsigned long long x = 1;
  __device__ double next() {
    const double r = 2147483646.0;
    double sum = 0.0, tmp = 1.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      x = (x * 16807ull) % 2147483647ull;
      sum += (double)(x - 1) * tmp;
      tmp = tmp * r;
    }
    const double u = sum / tmp;
    return u >= 1.0 ? 0x1.fffffffffffffp-1 : u;
  }
};

/* AREA_IMPORTANCE, one thread: write = false counts the samples into *m */
__global__ void k_ts_area(const double* __restrict__ tris, const int* __restrict__ surf, int64_t nt, double res,
                          bool write, double* __restrict__ P, double* __restrict__ E,
                          unsigned long long* __restrict__ m) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  TsRng gen;
  unsigned long long j = 0;
  for (int64_t t0 = 0; t0 < nt;) {
    int64_t t1 = t0;
    double total = 0.0;
    while (t1 < nt && surf[t1] == surf[t0]) {
      const double* T = tris + HCS_TS_TRI * t1;
      double e1[3] = {T[3] - T[0], T[4] - T[1], T[5] - T[2]}, e2[3] = {T[6] - T[0], T[7] - T[1], T[8] - T[2]}, cr[3];
      cross3(e1, e2, cr);
      total += 0.5 * sqrt(dot3(cr, cr));
      ++t1;
    }
    const double area_resolution = res * total;
    double acc = 0.0, at = 0.0;
    for (int64_t t = t0; t < t1; ++t) {
      const double* T = tris + HCS_TS_TRI * t;
      double e1[3] = {T[3] - T[0], T[4] - T[1], T[5] - T[2]}, e2[3] = {T[6] - T[0], T[7] - T[1], T[8] - T[2]}, cr[3];
      cross3(e1, e2, cr);
      at += 0.5 * sqrt(dot3(cr, cr));
      while (acc < at) {
        acc += area_resolution;
        const double u0 = gen.next(), u1 = gen.next();
        const double a = 1.0 - sqrt(u0);
        const double b = (1.0 - a) * u1;
        if (write) ts_sample(T, a, (1 - a) * (1 - b), (1 - a) * b, P + 3 * j, E + j);
        ++j;
      }
    }
    t0 = t1;
  }
  if (!write) *m = j;
}

struct CsArgs {
  double X[12];
  double margin;
};

struct TsTaxelArgs {
  double X[12];
  double margin, res;
  int method, visualize;
};

__global__ __launch_bounds__(64) void k_ts_taxel(const double* __restrict__ P, const double* __restrict__ E,
                                                 int64_t m, TsTaxelArgs a, const double* __restrict__ taxels,
                                                 double* __restrict__ values) {
  const int i = blockIdx.x, lane = threadIdx.x;
  double tw[3];
  xform(a.X, taxels + 3 * (size_t)i, tw);
  const double tt = dot3(tw, tw), msq = a.margin * a.margin;
  if (a.method == 0) {
    double bd = __builtin_inf();
    long long bj = 0x7fffffffffffffffll;
    for (int64_t j = lane; j < m; j += 64) {
      const double* p = P + 3 * j;
      const double d = ((-2.0 * dot3(tw, p)) + tt) + dot3(p, p);
      if (d < bd) bd = d, bj = j;
    }
    /* first nearest sample: (distance, index) lexicographic minimum */
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const double od = __shfl_xor(bd, o, 64);
      const long long oj = __shfl_xor(bj, o, 64);
      if (od < bd || (od == bd && oj < bj)) bd = od, bj = oj;
    }
    if (lane == 0 && bd < msq) values[i] = (a.visualize && fabs(E[bj]) > 1e-6) ? E[bj] : 0.0;
  } else {
    double sum = 0.0;
    long long ws = 0;
    for (int64_t j = lane; j < m; j += 64) {
      const double* p = P + 3 * j;
      const double d = ((-2.0 * dot3(tw, p)) + tt) + dot3(p, p);
      if (d < msq) {
        const double w0 = fmax(0.0, a.margin - sqrt(d));
        sum += (w0 * w0) * fabs(E[j]);
        ++ws;
      }
    }
    sum = wave_sum<double>(sum);
    ws = wave_sum_i(ws);
    if (lane == 0 && ws > 0) values[i] = sum * a.res;
  }
}

/* CurvedSensor ray of one surface sample (one thread): the sample moved by
 * X (normal rotated, not renormalised), ray from p + n 1e-8 along -n,
 * Moeller-Trumbore (S/src/bvh.cpp:49-74) in FP64 over the fan triangles,
 * nearest t (first triangle in order on ties); raw pressure when 0 < t <
 * margin.  Triangles whose box misses the box of the ray segment
 * [0, margin] (widened by 1e-9 of its largest coordinate) are skipped: they
 * cannot give an accepted hit, and one with t >= margin never changes
 * whether the nearest hit is accepted. */
__global__ __launch_bounds__(256) void k_cs_rays(const double* __restrict__ tris, int64_t nt, CsArgs a,
                                                 const double* __restrict__ pts, const double* __restrict__ nrms,
                                                 int m, double* __restrict__ raw, unsigned char* __restrict__ hit) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  double pw[3], nw[3], O[3], D[3];
  xform(a.X, pts + 3 * (size_t)j, pw);
  rot(a.X, nrms + 3 * (size_t)j, nw);
#pragma unroll
  for (int c = 0; c < 3; ++c) O[c] = pw[c] + nw[c] * 1e-8, D[c] = -nw[c];
  double slo[3], shi[3], E[3], sc = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) E[c] = O[c] + D[c] * a.margin, sc = fmax(sc, fmax(fabs(O[c]), fabs(E[c])));
  /* one slack for all axes, on the scale of the largest coordinate: the
   * rounding of t moves the computed hit along every axis by that much */
  const double sl = 1e-9 * sc + 1e-300;
#pragma unroll
  for (int c = 0; c < 3; ++c) slo[c] = fmin(O[c], E[c]) - sl, shi[c] = fmax(O[c], E[c]) + sl;
  double best = 1e30, bu = 0.0, bv = 0.0;
  int64_t bt = -1;
  for (int64_t t = 0; t < nt; ++t) {
    const double* T = tris + HCS_TS_TRI * t;
    bool out = false;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double lo = fmin(fmin(T[c], T[3 + c]), T[6 + c]), hi = fmax(fmax(T[c], T[3 + c]), T[6 + c]);
      out = out || hi < slo[c] || lo > shi[c];
    }
    if (out) continue;
    const double e1[3] = {T[3] - T[0], T[4] - T[1], T[5] - T[2]}, e2[3] = {T[6] - T[0], T[7] - T[1], T[8] - T[2]};
    double h[3];
    cross3(D, e2, h);
    const double det = dot3(e1, h);
    if (fabs(det) < 1e-10) continue;
    const double f = 1.0 / det;
    const double sv[3] = {O[0] - T[0], O[1] - T[1], O[2] - T[2]};
    const double u = f * dot3(sv, h);
    if (u < 0.0 || u > 1.0) continue;
    double qv[3];
    cross3(sv, e1, qv);
    const double v = f * dot3(D, qv);
    if (v < 0.0 || u + v > 1.0) continue;
    const double tt = f * dot3(e2, qv);
    if (tt > 0.0 && tt < best) best = tt, bu = u, bv = v, bt = t;
  }
  const bool ok = bt >= 0 && best < a.margin && best > 0.0;
  double r = 0.0;
  if (ok) {
    const double* T = tris + HCS_TS_TRI * bt;
    const double b0 = 1 - bu - bv;
    r = (b0 * T[9] + bu * T[10]) + bv * T[11];
  }
  raw[j] = r;
  hit[j] = ok ? 1 : 0;
}

/* taxel i: sum over its samples of weight * raw, in list order */
__global__ void k_cs_taxels(const int* __restrict__ ptr, const int* __restrict__ idx, const double* __restrict__ w,
                            int n, const double* __restrict__ raw, const unsigned char* __restrict__ hit,
                            double* __restrict__ values) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double pr = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k)
    if (hit[idx[k]]) pr += w[k] * raw[idx[k]];
  values[i] = pr;
}

}  // namespace hcs

extern "C" {

/* The kTriangle fan triangles of the kept polygons of the pairs touching
 * geometry geom, in key order, into c->ts.tris / c->ts.surf. */
static int sensor_fans(Ctx* c, int32_t geom, unsigned& ntri) {
  int rc;
  hipStream_t s = c->stream;
  const int npairs = (int)c->last_pairs.size();
  const int64_t np = c->last_npoly;
  std::vector<unsigned char> mask(std::max(1, npairs), 0);
  for (int e = 0; e < npairs; ++e)
    mask[e] = (c->last_geoms[2 * e] == geom || c->last_geoms[2 * e + 1] == geom) ? 1 : 0;
  TaxelSensorBufs& B = c->ts;
  const size_t n1 = (size_t)std::max<int64_t>(1, np);
  if ((rc = ensure(c, B.mask, mask.size())) || (rc = ensure(c, B.k64, 2 * sizeof(unsigned long long) * n1)) ||
      (rc = ensure(c, B.kp, 2 * sizeof(unsigned) * n1)) || (rc = ensure(c, B.idx, 2 * sizeof(int) * n1)) ||
      (rc = ensure(c, B.cnt, 64)) || (rc = ensure(c, B.nfan, sizeof(unsigned) * (n1 + 1))) ||
      (rc = ensure(c, B.fan_off, sizeof(unsigned) * (n1 + 1))))
    return rc;
  rc = hip_check(c, hipMemcpyAsync(B.mask.ptr, mask.data(), mask.size(), hipMemcpyHostToDevice, s), "upload mask");
  if (!rc) rc = hip_check(c, hipMemsetAsync(B.cnt.ptr, 0, 64, s), "memset");
  if (rc) return rc;
  unsigned long long* k64 = (unsigned long long*)B.k64.ptr;
  unsigned* kp = (unsigned*)B.kp.ptr;
  int* idx = (int*)B.idx.ptr;
  unsigned nsel = 0;
  if (np > 0) {
    hipLaunchKernelGGL(k_ts_keys, dim3((np + 255) / 256), dim3(256), 0, s, (const int64_t*)c->d_poly_key.ptr, np,
                       (const unsigned char*)B.mask.ptr, k64, kp, idx, (unsigned*)B.cnt.ptr);
    /* stable LSD passes: (element of A, element of B), then pair */
    const size_t b1 = sort_temp_bytes(np, 8), b2 = sort_temp_bytes(np, 4);
    if ((rc = ensure(c, B.tmp, std::max<size_t>(8, std::max(b1, b2))))) return rc;
    if (radix_sort_pairs_u64(s, k64, k64 + n1, idx, idx + n1, np, 0, 64, B.tmp.ptr, B.tmp.bytes) != 0)
      return set_err(c, HCS_ERR_LIMIT, "sensor: too many polygons to sort");
    hipLaunchKernelGGL(k_ts_gather_u32, dim3((np + 255) / 256), dim3(256), 0, s, kp, idx + n1, np, kp + n1);
    if (radix_sort_pairs_u32(s, kp + n1, kp, idx + n1, idx, np, 0, 32, B.tmp.ptr, B.tmp.bytes) != 0)
      return set_err(c, HCS_ERR_LIMIT, "sensor: too many polygons to sort");
    rc = hip_check(c, hipGetLastError(), "select polygons");
    if (!rc) rc = hip_check(c, hipMemcpyAsync(&nsel, B.cnt.ptr, sizeof(unsigned), hipMemcpyDeviceToHost, s), "dl");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    if (rc) return rc;
  }
  /* fan triangles of the selected polygons, in order */
  ntri = 0;
  if (nsel > 0) {
    hipLaunchKernelGGL(k_ts_nfan, dim3((nsel + 1 + 255) / 256), dim3(256), 0, s, (const double*)c->d_poly_rec.ptr,
                       idx, (int)nsel, (unsigned*)B.nfan.ptr);
    if ((rc = ensure(c, B.tmp, std::max<size_t>(scan_temp_bytes((long long)nsel + 1), 8)))) return rc;
    if (exclusive_scan_u32(s, (const unsigned*)B.nfan.ptr, (unsigned*)B.fan_off.ptr, (long long)nsel + 1, B.tmp.ptr,
                           B.tmp.bytes) != 0)
      return set_err(c, HCS_ERR_LIMIT, "sensor: too many polygons to scan");
    rc = hip_check(c, hipMemcpyAsync(&ntri, (unsigned*)B.fan_off.ptr + nsel, sizeof(unsigned), hipMemcpyDeviceToHost,
                                     s), "dl");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    if (rc) return rc;
  }
  const size_t nt1 = std::max<size_t>(1, ntri);
  if ((rc = ensure(c, B.tris, sizeof(double) * HCS_TS_TRI * nt1)) || (rc = ensure(c, B.surf, sizeof(int) * nt1)) ||
      (rc = ensure(c, B.scnt, sizeof(unsigned) * (nt1 + 1))) || (rc = ensure(c, B.soff, sizeof(unsigned) * (nt1 + 1))))
    return rc;
  if (ntri > 0)
    hipLaunchKernelGGL(k_ts_fans, dim3((nsel + 255) / 256), dim3(256), 0, s, (const double*)c->d_poly_rec.ptr,
                       (const int64_t*)c->d_poly_key.ptr, idx, (int)nsel, (const unsigned*)B.fan_off.ptr,
                       (double*)B.tris.ptr, (int*)B.surf.ptr);
  return hip_check(c, hipGetLastError(), "sensor triangles");
}

int hcs_taxel_sensor_update(hcs_ctx* h, const hcs_taxel_sensor* S, double* values, int64_t* nsamples) {
  if (!h || !S || !values || S->ntaxels < 0 || (S->ntaxels > 0 && !S->taxels) || S->method < 0 || S->method > 3 ||
      S->sample_method < 0 || S->sample_method > 1 || !(S->sample_resolution > 0) || !(S->include_margin >= 0))
    return HCS_ERR_ARG;
  Ctx* c = &h->c;
  if (!c->last_valid)
    return set_err(c, HCS_ERR_STATE, "hcs_taxel_sensor_update needs a contact step run with keep_polygons = 1");
  int rc = hip_check(c, hipSetDevice(c->device), "hipSetDevice");
  if (rc) return rc;
  hipStream_t s = c->stream;
  TaxelSensorBufs& B = c->ts;
  unsigned ntri = 0;
  if ((rc = sensor_fans(c, S->geom, ntri))) return rc;
  /* samples */
  unsigned long long m = 0;
  if (ntri > 0 && S->sample_method == 0) {
    hipLaunchKernelGGL(k_ts_count, dim3((ntri + 1 + 255) / 256), dim3(256), 0, s, (const double*)B.tris.ptr,
                       (int64_t)ntri, S->sample_resolution, (unsigned*)B.scnt.ptr);
    if ((rc = ensure(c, B.tmp, std::max<size_t>(scan_temp_bytes((long long)ntri + 1), 8)))) return rc;
    if (exclusive_scan_u32(s, (const unsigned*)B.scnt.ptr, (unsigned*)B.soff.ptr, (long long)ntri + 1, B.tmp.ptr,
                           B.tmp.bytes) != 0)
      return set_err(c, HCS_ERR_LIMIT, "sensor: too many triangles to scan");
    unsigned m32 = 0;
    rc = hip_check(c, hipMemcpyAsync(&m32, (unsigned*)B.soff.ptr + ntri, sizeof(unsigned), hipMemcpyDeviceToHost, s),
                   "dl");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    m = m32;
  } else if (ntri > 0) {
    hipLaunchKernelGGL(k_ts_area, dim3(1), dim3(64), 0, s, (const double*)B.tris.ptr, (const int*)B.surf.ptr,
                       (int64_t)ntri, S->sample_resolution, false, nullptr, nullptr,
                       (unsigned long long*)B.cnt.ptr);
    rc = hip_check(c, hipMemcpyAsync(&m, B.cnt.ptr, sizeof(m), hipMemcpyDeviceToHost, s), "dl");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    if (rc) return rc;
  }
  if (nsamples) *nsamples = (int64_t)m;
  if (m == 0) {
    for (int i = 0; i < S->ntaxels; ++i) values[i] = 0.0;
    return HCS_OK;
  }
  if ((rc = ensure(c, B.P, sizeof(double) * 3 * m)) || (rc = ensure(c, B.E, sizeof(double) * m)) ||
      (rc = ensure(c, B.taxel, sizeof(double) * 3 * std::max(1, S->ntaxels))) ||
      (rc = ensure(c, B.val, sizeof(double) * std::max(1, S->ntaxels))))
    return rc;
  if (S->sample_method == 0)
    hipLaunchKernelGGL(k_ts_gen, dim3((ntri + 255) / 256), dim3(256), 0, s, (const double*)B.tris.ptr, (int64_t)ntri,
                       S->sample_resolution, (const unsigned*)B.soff.ptr, (double*)B.P.ptr, (double*)B.E.ptr);
  else
    hipLaunchKernelGGL(k_ts_area, dim3(1), dim3(64), 0, s, (const double*)B.tris.ptr, (const int*)B.surf.ptr,
                       (int64_t)ntri, S->sample_resolution, true, (double*)B.P.ptr, (double*)B.E.ptr,
                       (unsigned long long*)B.cnt.ptr);
  if (S->ntaxels == 0) return hip_check(c, hipGetLastError(), "samples");
  TsTaxelArgs a;
  std::memcpy(a.X, S->X_W_sensor, sizeof a.X);
  a.margin = S->include_margin, a.res = S->sample_resolution, a.method = S->method, a.visualize = S->visualize != 0;
  rc = hip_check(c, hipMemcpyAsync(B.taxel.ptr, S->taxels, sizeof(double) * 3 * S->ntaxels, hipMemcpyHostToDevice, s),
                 "upload taxels");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(B.val.ptr, values, sizeof(double) * S->ntaxels, hipMemcpyHostToDevice, s),
                   "upload values");
  if (rc) return rc;
  hipLaunchKernelGGL(k_ts_taxel, dim3(S->ntaxels), dim3(64), 0, s, (const double*)B.P.ptr, (const double*)B.E.ptr,
                     (int64_t)m, a, (const double*)B.taxel.ptr, (double*)B.val.ptr);
  rc = hip_check(c, hipGetLastError(), "taxel sensor");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(values, B.val.ptr, sizeof(double) * S->ntaxels, hipMemcpyDeviceToHost, s), "dl");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
  return rc;
}

int hcs_curved_sensor_update(hcs_ctx* h, const hcs_curved_sensor* S, double* values) {
  if (!h || !S || !values || S->ntaxels < 0 || S->nsamples < 0 || !(S->include_margin >= 0) ||
      (S->ntaxels > 0 && !S->taxel_ptr) || (S->nsamples > 0 && (!S->points || !S->normals)))
    return HCS_ERR_ARG;
  Ctx* c = &h->c;
  if (S->ntaxels == 0) return HCS_OK;
  const int nnz = S->taxel_ptr[S->ntaxels];
  if (S->taxel_ptr[0] != 0 || nnz < 0 || (nnz > 0 && (!S->sample_idx || !S->weights)))
    return set_err(c, HCS_ERR_ARG, "invalid taxel sample lists");
  for (int i = 0; i < S->ntaxels; ++i)
    if (S->taxel_ptr[i + 1] < S->taxel_ptr[i]) return set_err(c, HCS_ERR_ARG, "taxel_ptr must be non-decreasing");
  for (int k = 0; k < nnz; ++k)
    if (S->sample_idx[k] < 0 || S->sample_idx[k] >= S->nsamples)
      return set_err(c, HCS_ERR_ARG, "sample index out of range");
  if (!c->last_valid)
    return set_err(c, HCS_ERR_STATE, "hcs_curved_sensor_update needs a contact step run with keep_polygons = 1");
  int rc = hip_check(c, hipSetDevice(c->device), "hipSetDevice");
  if (rc) return rc;
  hipStream_t s = c->stream;
  TaxelSensorBufs& B = c->ts;
  unsigned ntri = 0;
  if ((rc = sensor_fans(c, S->geom, ntri))) return rc;
  if (ntri == 0) {
    /* no contact surface touches the sensor (curved_sensor.cpp:434-438) */
    for (int i = 0; i < S->ntaxels; ++i) values[i] = 0.0;
    return HCS_OK;
  }
  const int m = S->nsamples;
  const size_t m1 = (size_t)std::max(1, m), n1 = (size_t)S->ntaxels, z1 = (size_t)std::max(1, nnz);
  if ((rc = ensure(c, B.P, sizeof(double) * 6 * m1)) || (rc = ensure(c, B.E, sizeof(double) * m1)) ||
      (rc = ensure(c, B.mask, m1)) || (rc = ensure(c, B.taxel, sizeof(double) * z1)) ||
      (rc = ensure(c, B.val, sizeof(double) * n1)) || (rc = ensure(c, B.scnt, sizeof(int) * (n1 + 1))) ||
      (rc = ensure(c, B.soff, sizeof(int) * z1)))
    return rc;
  double* pts = (double*)B.P.ptr;
  rc = hip_check(c, hipMemcpyAsync(pts, S->points, sizeof(double) * 3 * m, hipMemcpyHostToDevice, s), "ul");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(pts + 3 * m1, S->normals, sizeof(double) * 3 * m, hipMemcpyHostToDevice, s), "ul");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(B.scnt.ptr, S->taxel_ptr, sizeof(int) * (n1 + 1), hipMemcpyHostToDevice, s), "ul");
  if (!rc && nnz > 0)
    rc = hip_check(c, hipMemcpyAsync(B.soff.ptr, S->sample_idx, sizeof(int) * nnz, hipMemcpyHostToDevice, s), "ul");
  if (!rc && nnz > 0)
    rc = hip_check(c, hipMemcpyAsync(B.taxel.ptr, S->weights, sizeof(double) * nnz, hipMemcpyHostToDevice, s), "ul");
  if (rc) return rc;
  CsArgs a;
  std::memcpy(a.X, S->X_W_sensor, sizeof a.X);
  a.margin = S->include_margin;
  if (m > 0)
    hipLaunchKernelGGL(k_cs_rays, dim3((m + 255) / 256), dim3(256), 0, s, (const double*)B.tris.ptr, (int64_t)ntri, a,
                       pts, pts + 3 * m1, m, (double*)B.E.ptr, (unsigned char*)B.mask.ptr);
  hipLaunchKernelGGL(k_cs_taxels, dim3((S->ntaxels + 255) / 256), dim3(256), 0, s, (const int*)B.scnt.ptr,
                     (const int*)B.soff.ptr, (const double*)B.taxel.ptr, S->ntaxels, (const double*)B.E.ptr,
                     (const unsigned char*)B.mask.ptr, (double*)B.val.ptr);
  rc = hip_check(c, hipGetLastError(), "curved sensor");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(values, B.val.ptr, sizeof(double) * n1, hipMemcpyDeviceToHost, s), "dl");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
  return rc;
}

}  // extern "C"
