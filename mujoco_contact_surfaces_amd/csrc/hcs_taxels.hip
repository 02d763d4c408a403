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
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "hcs_device.h"
#include "hcs_internal.h"
#include "hcs_prim.h"
#include "hcs_taxel_dev.h"

#pragma clang fp contract(off)

using namespace hcs;

namespace hcs {

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

template <class T>
__global__ __launch_bounds__(64) void k_tx_sample(const double* __restrict__ poly, const double* __restrict__ ptan,
                                                   const int64_t* __restrict__ key, TaxelGridDev g,
                                                   const double* __restrict__ Xd, const unsigned* __restrict__ off,
                                                   const unsigned* __restrict__ cnt, const int* __restrict__ list,
                                                   double* __restrict__ out) {
  const int cell = blockIdx.x;
  const int x = cell % g.cx, y = cell / g.cx;
  const double v = sample_cell<T>(poly, ptan, key, g, Xd, off[cell], cnt[cell], list, x, y);
  if (threadIdx.x == 0) out[x + g.cx * y] = v;
}

/* The same, fan-major (double, S^2 <= HCS_TX_RAYS). */
__global__ __launch_bounds__(64) void k_tx_sample_fan(const double* __restrict__ poly, const int64_t* __restrict__ key,
                                                      TaxelGridDev g, const unsigned* __restrict__ off,
                                                      const unsigned* __restrict__ cnt, const int* __restrict__ list,
                                                      double* __restrict__ out) {
  const int cell = blockIdx.x;
  const int x = cell % g.cx, y = cell / g.cx;
  const double v = sample_cell_scatter(poly, key, g, off[cell], cnt[cell], list, x, y);
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
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tx_sample<Dual>), dim3(ncell), dim3(64), 0, c->stream, poly, ptan, key, g,
                       Xd, (const unsigned*)off, (const unsigned*)cnt, (const int*)c->d_tx_list.ptr, d_out);
  else if (g.S * g.S <= HCS_TX_RAYS && c->last_pairs.size() < (1u << 19))
    hipLaunchKernelGGL(k_tx_sample_fan, dim3(ncell), dim3(64), 0, c->stream, poly, key, g, (const unsigned*)off,
                       (const unsigned*)cnt, (const int*)c->d_tx_list.ptr, d_out);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tx_sample<double>), dim3(ncell), dim3(64), 0, c->stream, poly,
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

/* The taxels that have candidate polygons (and whose list fits the list
 * buffer; an overflowed list is re-run after the step with a larger one),
 * appended to live[] with one atomic per block of HCS_TXB_LIVE threads
 * (atomics on the one counter complete one after another, so one per wave
 * would queue thousands); every other taxel is 0 (the output is cleared
 * before). */
#define HCS_TXB_LIVE 1024
__global__ __launch_bounds__(HCS_TXB_LIVE) void k_txb_live(const unsigned* __restrict__ cnt,
                                                           const unsigned* __restrict__ off, int ncell,
                                                           int64_t list_cap, int* __restrict__ live,
                                                           unsigned* __restrict__ nlive) {
  __shared__ unsigned s_off[HCS_TXB_LIVE / 64 + 1];
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = cell < ncell && cnt[cell] > 0 && (int64_t)off[cell] + cnt[cell] <= list_cap;
  const unsigned long long m = __ballot(on);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s_off[w] = (unsigned)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tot = 0;
    for (int x = 0; x < (int)(blockDim.x >> 6); ++x) {
      const unsigned n = s_off[x];
      s_off[x] = tot;
      tot += n;
    }
    s_off[HCS_TXB_LIVE / 64] = tot ? atomicAdd(nlive, tot) : 0u;
  }
  __syncthreads();
  if (on)
    live[s_off[HCS_TXB_LIVE / 64] + s_off[w] +
         __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = cell;
}

/* One wave per live taxel of any sensor, strided over the grid; FAN: the
 * fan-major sampler (every sensor has S^2 <= HCS_TX_RAYS). */
template <bool FAN>
__global__ __launch_bounds__(64) void k_txb_sample(const double* __restrict__ poly, const int64_t* __restrict__ key,
                                                   const TaxelGridDev* __restrict__ grids,
                                                   const int* __restrict__ cell_base, const int* __restrict__ cell_sensor,
                                                   const unsigned* __restrict__ off, const unsigned* __restrict__ cnt,
                                                   const int* __restrict__ list, const int* __restrict__ live,
                                                   const unsigned* __restrict__ nlive, double* __restrict__ out) {
  const unsigned nl = *nlive;
  for (unsigned k = blockIdx.x; k < nl; k += gridDim.x) {
    const int cell = live[k];
    const int s = cell_sensor[cell];
    const TaxelGridDev g = grids[s];
    const int local = cell - cell_base[s];
    const double v =
        FAN ? sample_cell_scatter(poly, key, g, off[cell], cnt[cell], list, local % g.cx, local / g.cx)
            : sample_cell<double>(poly, nullptr, key, g, nullptr, off[cell], cnt[cell], list, local % g.cx,
                                  local / g.cx);
    if (threadIdx.x == 0) out[cell] = v;
  }
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
      (rc = ensure(c, t.off, sizeof(unsigned) * (ncell + 1))) || (rc = ensure(c, t.out, sizeof(double) * std::max(1, ncell))) ||
      (rc = ensure(c, t.live, sizeof(int) * std::max(1, ncell))))
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
  t.max_rays = 0;
  for (int k = 0; k < n; ++k) t.max_rays = std::max(t.max_rays, g[k].S * g[k].S);
  /* the fan-major sampler packs pair indices into 19 bits */
  if (npairs >= (1 << 19)) t.max_rays = INT_MAX;
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
  unsigned* nlive = (unsigned*)t.poly_cnt.ptr + 2;
  launch_fill_u32(s, nlive, 0u, 1);
  launch_fill_u32(s, (unsigned*)t.out.ptr, 0u, 2 * t.ncell);
  hipLaunchKernelGGL(k_txb_live, dim3((t.ncell + HCS_TXB_LIVE - 1) / HCS_TXB_LIVE), dim3(HCS_TXB_LIVE), 0, s,
                     (const unsigned*)t.cnt.ptr,
                     (const unsigned*)t.off.ptr, t.ncell, t.list_cap, (int*)t.live.ptr, nlive);
  const dim3 grid(std::max(1, std::min(t.ncell, c->num_cu * 16)));
  if (t.max_rays <= HCS_TX_RAYS)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_txb_sample<true>), grid, dim3(64), 0, s, (const double*)t.poly_rec.ptr,
                       (const int64_t*)t.poly_key.ptr, (const TaxelGridDev*)g, (const int*)t.cell_base.ptr,
                       (const int*)t.cell_sensor.ptr, (const unsigned*)t.off.ptr, (const unsigned*)t.cnt.ptr,
                       (const int*)t.list.ptr, (const int*)t.live.ptr, (const unsigned*)nlive, (double*)t.out.ptr);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_txb_sample<false>), grid, dim3(64), 0, s, (const double*)t.poly_rec.ptr,
                       (const int64_t*)t.poly_key.ptr, (const TaxelGridDev*)g, (const int*)t.cell_base.ptr,
                       (const int*)t.cell_sensor.ptr, (const unsigned*)t.off.ptr, (const unsigned*)t.cnt.ptr,
                       (const int*)t.list.ptr, (const int*)t.live.ptr, (const unsigned*)nlive, (double*)t.out.ptr);
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
                  &t.poly_cnt, &t.cnt, &t.cursor, &t.off, &t.scan_tmp, &t.list, &t.out, &t.live})
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
