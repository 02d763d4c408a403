/*
 * hcs_engine.hip -- device kernels and host orchestration of the contact
 * step: mesh upload + static LBVH build, the fused per-step query kernel,
 * per-pair reduction, the plain per-call path and the batched persistent
 * path captured in a hipGraph.
 *
 * Replaces, per step, what the reference does on the MuJoCo physics thread
 * (citations relative to the reference repository):
 *   collision_cb            mujoco_contact_surfaces/src/mujoco_contact_surfaces_plugin.cpp:255-318
 *   Drake tet-tri / tet-tet queries                                  :284-304
 *   evaluateContactSurface                                           :320-409
 *   passiveCallback force loop                                       :411-483
 * and at load time the Bvh<Obb, VolumeMesh> builds of
 * parseMujocoCustomFields (:656,680,704,730,790).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "hcs_internal.h"
#include "hcs_prim.h"
#include "hcs_query.h"

#pragma clang fp contract(off)

namespace hcs {

static thread_local std::string g_create_err;

int set_err(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_check(Ctx* c, hipError_t e, const char* what) {
  if (e == hipSuccess) return HCS_OK;
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return set_err(c, e == hipErrorOutOfMemory ? HCS_ERR_OOM : HCS_ERR_HIP, m);
}

#define HCHK(c, expr)                                  \
  do {                                                 \
    int _rc = hip_check((c), (expr), #expr);           \
    if (_rc != HCS_OK) return _rc;                     \
  } while (0)

int ensure(Ctx* c, DBuf& b, size_t bytes) {
  if (bytes <= b.bytes && b.ptr) return HCS_OK;
  if (b.fixed)
    return set_err(c, HCS_ERR_CAPACITY, "request of " + std::to_string(bytes) + " bytes exceeds a fixed " +
                                            std::to_string(b.bytes) + "-byte slice of the scratch array");
  if (b.ptr) {
    HCHK(c, hipStreamSynchronize(c->stream));
    HCHK(c, hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t n = std::max<size_t>(bytes, 256);
  HCHK(c, hipMalloc(&b.ptr, n));
  b.bytes = n;
  return HCS_OK;
}

/* ===================================================================
 * Mesh preparation kernels (load time)
 * =================================================================== */

__constant__ int kFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

/* Tet record: face half-spaces, linear pressure field, AABB.  Same
 * arithmetic as the oracle's tet record. */
__global__ void k_tet_precompute(const double* __restrict__ V, const int32_t* __restrict__ T,
                                 const double* __restrict__ p, int nt, double* __restrict__ rec,
                                 double* __restrict__ box, float4* __restrict__ fpl) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t* tet = T + 4 * (size_t)t;
  double v[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) v[i][c] = V[3 * (size_t)tet[i] + c];
  double* r = rec + (size_t)HCS_TREC * t;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double* A = v[kFace[k][0]];
    const double* B = v[kFace[k][1]];
    const double* C = v[kFace[k][2]];
    double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    double n[3];
    cross3(e1, e2, n);
    double len = sqrt(dot3(n, n));
    double nh[3] = {n[0] / len, n[1] / len, n[2] / len};
    r[3 * k] = nh[0], r[3 * k + 1] = nh[1], r[3 * k + 2] = nh[2];
    r[12 + k] = -dot3(nh, A);
    /* the same plane rounded to float (nearest) */
    fpl[4 * (size_t)t + k] = make_float4((float)nh[0], (float)nh[1], (float)nh[2], (float)r[12 + k]);
  }
  double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
  double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
  double e3[3] = {v[3][0] - v[0][0], v[3][1] - v[0][1], v[3][2] - v[0][2]};
  double c23[3], c31[3], c12[3];
  cross3(e2, e3, c23);
  cross3(e3, e1, c31);
  cross3(e1, e2, c12);
  const double det = dot3(e1, c23);
  const double p0 = p[tet[0]];
  const double dp1 = p[tet[1]] - p0, dp2 = p[tet[2]] - p0, dp3 = p[tet[3]] - p0;
  double g[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) g[i] = ((dp1 * c23[i] + dp2 * c31[i]) + dp3 * c12[i]) / det;
  r[16] = g[0], r[17] = g[1], r[18] = g[2];
  r[19] = p0 - dot3(g, v[0]);
  const double gg = dot3(g, g);
  if (gg > 0) {
    const double gl = sqrt(gg);
    r[20] = g[0] / gl, r[21] = g[1] / gl, r[22] = g[2] / gl;
  } else {
    r[20] = g[0], r[21] = g[1], r[22] = g[2];
  }
  r[23] = 0.0;
  double* b = box + 6 * (size_t)t;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double lo = v[0][c], hi = v[0][c];
#pragma unroll
    for (int j = 1; j < 4; ++j) lo = fmin(lo, v[j][c]), hi = fmax(hi, v[j][c]);
    b[c] = lo, b[3 + c] = hi;
  }
}

/* 4-wide nodes of the LBVH's n_int internal nodes. */
__global__ void k_wide_nodes(const Node* __restrict__ nodes, int n_int, WNode* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_int) make_wide_node(nodes, i, w[i]);
}

/* Unit triangle normals (Drake's TriangleSurfaceMesh face normals). */
__global__ void k_tri_normals(const double* __restrict__ V, const int32_t* __restrict__ F, int ntri,
                              double* __restrict__ nrm) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= ntri) return;
  const double* a = V + 3 * (size_t)F[3 * f];
  const double* b = V + 3 * (size_t)F[3 * f + 1];
  const double* c = V + 3 * (size_t)F[3 * f + 2];
  double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double n[3];
  cross3(e1, e2, n);
  const double z = dot3(n, n);
  if (z > 0) {
    const double l = sqrt(z);
    n[0] /= l, n[1] /= l, n[2] /= l;
  }
  nrm[3 * f] = n[0], nrm[3 * f + 1] = n[1], nrm[3 * f + 2] = n[2];
}

__device__ inline unsigned expand_bits(unsigned v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void k_morton(const double* __restrict__ box, int n, double lx, double ly, double lz, double sx,
                         double sy, double sz, unsigned* __restrict__ codes, int* __restrict__ idx) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double* b = box + 6 * (size_t)t;
  double u[3] = {((b[0] + b[3]) * 0.5 - lx) * sx, ((b[1] + b[4]) * 0.5 - ly) * sy, ((b[2] + b[5]) * 0.5 - lz) * sz};
  unsigned q[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) q[c] = (unsigned)fmin(fmax(u[c], 0.0), 1023.0);
  codes[t] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
  idx[t] = t;
}

__device__ inline int delta(const unsigned* codes, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  unsigned a = codes[i], b = codes[j];
  if (a == b) return 32 + __clz((unsigned)(i ^ j));
  return __clz(a ^ b);
}

/* Karras radix-tree construction over sorted Morton codes: internal node i
 * covers a key range whose split is the highest differing bit. */
__global__ void k_karras(const unsigned* __restrict__ codes, const int* __restrict__ idx, int n,
                         Node* __restrict__ nodes, int* __restrict__ node_parent, int* __restrict__ leaf_parent) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(codes, n, i, i - d);
  int lmax = 2;
  while (delta(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(codes, n, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? -1 : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  Node* N = nodes + i;
  if (lo == gamma) {
    N->c[0] = ~idx[gamma];
    leaf_parent[gamma] = 2 * i;
  } else {
    N->c[0] = gamma;
    node_parent[gamma] = 2 * i;
  }
  if (hi == gamma + 1) {
    N->c[1] = ~idx[gamma + 1];
    leaf_parent[gamma + 1] = 2 * i + 1;
  } else {
    N->c[1] = gamma + 1;
    node_parent[gamma + 1] = 2 * i + 1;
  }
}

/* Bottom-up refit: the second thread to reach a node writes its union box
 * and height into the parent's child slot.  Counter RMWs are acq_rel at
 * agent scope, which orders the slot stores/loads between the two threads. */
__global__ void k_refit(const double* __restrict__ box, const int* __restrict__ idx, int n, Node* nodes,
                        const int* __restrict__ node_parent, const int* __restrict__ leaf_parent, int* flags,
                        int* root_height) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double* b = box + 6 * (size_t)idx[j];
  int pc = leaf_parent[j];
  int p = pc >> 1, side = pc & 1;
  Node* N = nodes + p;
#pragma unroll
  for (int c = 0; c < 6; ++c) N->b[6 * side + c] = b[c];
  N->h[side] = 0;
  for (int guard = 0; guard < n; ++guard) {
    const int old = __hip_atomic_fetch_add(flags + p, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) return;
    N = nodes + p;
    double u[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      u[c] = fmin(N->b[c], N->b[6 + c]);
      u[3 + c] = fmax(N->b[3 + c], N->b[9 + c]);
    }
    const int h = 1 + (N->h[0] > N->h[1] ? N->h[0] : N->h[1]);
    if (p == 0) {
      *root_height = h;
      return;
    }
    pc = node_parent[p];
    p = pc >> 1, side = pc & 1;
    Node* P = nodes + p;
#pragma unroll
    for (int c = 0; c < 6; ++c) P->b[6 * side + c] = u[c];
    P->h[side] = h;
  }
}

/* ===================================================================
 * Per-step kernels
 * =================================================================== */

__global__ void k_fill_u32(unsigned int* p, unsigned int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_prep(const EnvDesc* __restrict__ envs, const double* __restrict__ state, double* __restrict__ xf,
                       int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double* st = state + (size_t)HCS_STATE * e;
  double X[12];
  compose_SR(st, st + 12, X);
#pragma unroll
  for (int i = 0; i < 12; ++i) xf[12 * (size_t)e + i] = X[i];
}

/* AABBs of a soft mesh's tets in Morton (leaf) order. */
__global__ void k_gather_boxes(const double* __restrict__ box, const int* __restrict__ idx, int n,
                               double* __restrict__ boxL) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) boxL[6 * (size_t)i + c] = box[6 * (size_t)idx[i] + c];
}

/* Per step, one block per tet-query pair: its rigid triangles moved into
 * A's frame (vertices by xform, unit normal by rot, and their AABB), the
 * same arithmetic the candidate and clip stages use. */
__global__ __launch_bounds__(64) void k_xtri(const EnvDesc* __restrict__ envs, const int* __restrict__ xenv,
                                             const double* __restrict__ xf, double* __restrict__ xtri) {
  const int e = xenv[blockIdx.x];
  const EnvDesc& E = envs[e];
  const double* X = xf + 12 * (size_t)e;
  for (int t = threadIdx.x; t < E.nB; t += blockDim.x) {
    double* o = xtri + HCS_XTRI * (size_t)(E.xtri_off + t);
    const int32_t* f = E.FB + 3 * (size_t)t;
    double lo[3], hi[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double y[3];
      xform(X, hcs_plain(E.VB + 3 * (size_t)f[i]), y);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        o[3 * i + c] = y[c];
        lo[c] = i == 0 ? y[c] : fmin(lo[c], y[c]);
        hi[c] = i == 0 ? y[c] : fmax(hi[c], y[c]);
      }
    }
    double nS[3];
    rot(X, hcs_plain(E.nrmB + 3 * (size_t)t), nS);
#pragma unroll
    for (int c = 0; c < 3; ++c) o[9 + c] = nS[c], o[12 + c] = lo[c], o[15 + c] = hi[c];
  }
}

/* Static boxes of one item per wave: [0..5] its query elements in their own
 * frame (tets of A in A's frame; triangles or tets of B in B's frame), then
 * [6..11] the target (B's vertex AABB in B's frame for tet queries, the
 * item's subtree of A's LBVH otherwise). */
__global__ __launch_bounds__(64) void k_item_boxes(const EnvDesc* __restrict__ envs, const Item* __restrict__ items,
                                                   const int* __restrict__ order, int nitems,
                                                   double* __restrict__ ibox) {
  /* block k: the k-th item of the work order, its boxes at position k */
  const int k = blockIdx.x, lane = threadIdx.x;
  if (k >= nitems) return;
  const int i = order ? order[k] : k;
  const Item it = items[i];
  const EnvDesc& E = envs[it.env];
  const double inf = __builtin_inf();
  double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
  /* every query of the item (64 per lane-batch, more with queries_per_lane) */
  for (int l = lane; l < it.count; l += 64) {
    if (E.kind == KIND_TET || E.kind == KIND_PLANE) {
      const double* b = E.boxLA + 6 * (size_t)(it.first + l);
#pragma unroll
      for (int c = 0; c < 3; ++c) lo[c] = fmin(lo[c], b[c]), hi[c] = fmax(hi[c], b[3 + c]);
    } else {
      const int nvq = E.kind == KIND_SOFT ? 4 : 3;
      const int q = it.first + l;
      const int32_t* f = E.FB + (size_t)nvq * (E.qorder ? E.qorder[q] : q);
      for (int k = 0; k < nvq; ++k) {
        const double* v = E.VB + 3 * (size_t)f[k];
#pragma unroll
        for (int c = 0; c < 3; ++c) lo[c] = fmin(lo[c], v[c]), hi[c] = fmax(hi[c], v[c]);
      }
    }
  }
  double t[6] = {inf, inf, inf, -inf, -inf, -inf};
  if (E.kind == KIND_PLANE) {
#pragma unroll
    for (int c = 0; c < 6; ++c) t[c] = 0.0; /* the plane needs no target box */
  } else if (E.kind == KIND_TET) {
    for (int v = lane; v < E.nvB; v += 64) {
#pragma unroll
      for (int c = 0; c < 3; ++c) t[c] = fmin(t[c], E.VB[3 * (size_t)v + c]), t[3 + c] = fmax(t[3 + c], E.VB[3 * (size_t)v + c]);
    }
  } else if (E.ntA == 1 || it.root < 0) {
    const double* b = E.boxA + 6 * (size_t)(E.ntA == 1 ? 0 : ~it.root);
#pragma unroll
    for (int c = 0; c < 6; ++c) t[c] = b[c];
  } else {
    const Node& N = E.nodesA[it.root];
#pragma unroll
    for (int c = 0; c < 3; ++c) t[c] = fmin(N.b[c], N.b[6 + c]), t[3 + c] = fmax(N.b[3 + c], N.b[9 + c]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    lo[c] = wave_min(lo[c]), hi[c] = wave_max(hi[c]);
    t[c] = wave_min(t[c]), t[3 + c] = wave_max(t[3 + c]);
  }
  if (lane == 0) {
    double* o = ibox + 12 * (size_t)k;
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = lo[c], o[3 + c] = hi[c], o[6 + c] = t[c], o[9 + c] = t[3 + c];
  }
}

/* Per step, one thread per item: keep the items whose static boxes may
 * overlap under this step's pose, appended to their queue's segment of the
 * live list (one atomic per queue present in the wave); a rejected item gets
 * zero counts here. */
#define HCS_CULL_THREADS 1024
__global__ __launch_bounds__(HCS_CULL_THREADS) void k_cull(const EnvDesc* __restrict__ envs, const Item* __restrict__ items,
                                              const int* __restrict__ order, const Item* __restrict__ items_o,
                                              int nitems,
                                              const double* __restrict__ ibox,
                                              const double* __restrict__ xf, const int* __restrict__ qseg,
                                              int nq, int* __restrict__ live, int4* __restrict__ live_item,
                                              unsigned* __restrict__ live_cnt, int* __restrict__ ncand,
                                              unsigned* __restrict__ nsurv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  /* threads visit the items in the plan's work order (heaviest first), so
   * the live queues hand out long items before short ones; the item records
   * and boxes are stored in that order, so the reads are contiguous */
  const int i = (k < nitems && order) ? order[k] : k;
  bool keep = false;
  int q = -1;
  Item it{0, 0, 0, 0};
  if (k < nitems) {
    it = order ? items_o[k] : items[k];
    const double* bx = ibox + 12 * (size_t)k;
    const int kind = envs[it.env].kind;
    const double* X = xf + 12 * (size_t)it.env;
    if (kind == KIND_PLANE) {
      keep = box_may_meet_plane(X, bx);
    } else {
      const bool tet = kind == KIND_TET;
      keep = boxes_may_overlap(X, tet ? bx + 6 : bx, tet ? bx : bx + 6);
    }
    if (!keep) ncand[i] = 0, nsurv[i] = 0;
    q = it.env % nq;
  }
  if (nq == 1) {
    /* one queue: one atomic per block on its counter (one per wave
     * contends on that single address), the block's waves in order */
    __shared__ unsigned s_off[HCS_CULL_THREADS / 64 + 1];
    const int w = threadIdx.x >> 6;
    const unsigned long long m = __ballot(keep);
    if (lane == 0) s_off[w] = (unsigned)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned tot = 0;
      for (int x = 0; x < (int)(blockDim.x >> 6); ++x) {
        const unsigned n = s_off[x];
        s_off[x] = tot;
        tot += n;
      }
      s_off[HCS_CULL_THREADS / 64] = atomicAdd(live_cnt, tot);
    }
    __syncthreads();
    if (keep) {
      const int pos = qseg[0] + s_off[HCS_CULL_THREADS / 64] + s_off[w] + lane_prefix(m);
      live[pos] = i;
      live_item[pos] = make_int4(it.env, it.first, it.count, it.root);
    }
    return;
  }
  unsigned long long rem = __ballot(keep);
  while (rem) {
    const int lead = __builtin_ctzll(rem);
    const int qq = __shfl(q, lead, 64);
    const unsigned long long m = __ballot(keep && q == qq);
    unsigned base = 0;
    if (lane == lead) base = atomicAdd(live_cnt + qq, (unsigned)__popcll(m));
    base = __shfl(base, lead, 64);
    if (keep && q == qq) {
      const int pos = qseg[qq] + base + lane_prefix(m);
      live[pos] = i;
      live_item[pos] = make_int4(it.env, it.first, it.count, it.root);
    }
    rem &= ~m;
  }
}

/* Sub-roots for splitting (one thread per request): the frontier of the LBVH
 * of the request's item below the item's root at depth[r] <= 6; out[65 r] =
 * count, then node codes (a leaf reached early stays a leaf).  The
 * frontier's subtrees hold each leaf of the root's subtree exactly once. */
__global__ void k_split_roots(const EnvDesc* __restrict__ envs, const Item* __restrict__ items,
                              const int* __restrict__ req, const int* __restrict__ depth, int nreq,
                              int32_t* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nreq) return;
  const Item it = items[req[r]];
  const Node* nodes = envs[it.env].nodesA;
  int32_t cur[64], nxt[64];
  int nc = 1;
  cur[0] = it.root;
  const int dmax = depth[r] < 6 ? depth[r] : 6;
  for (int d = 0; d < dmax; ++d) {
    int nn = 0;
    for (int k = 0; k < nc; ++k) {
      if (cur[k] < 0) {
        nxt[nn++] = cur[k];
      } else {
        nxt[nn++] = nodes[cur[k]].c[0];
        nxt[nn++] = nodes[cur[k]].c[1];
      }
    }
    for (int k = 0; k < nn; ++k) cur[k] = nxt[k];
    nc = nn;
  }
  out[65 * (size_t)r] = nc;
  for (int k = 0; k < nc; ++k) out[65 * (size_t)r + 1 + k] = cur[k];
}

/* Candidate generation: one wave per item; persistent when a work-queue
 * counter is given (the counter is reset by a fill kernel in the same
 * graph).  The triangle build fits three waves per SIMD; the general build
 * is held to two. */
template <int MODE, int KS>
__global__ __launch_bounds__(64, KS == KIND_TRI ? 3 : 2) void k_cand(QueryArgs A) {
  __shared__ CandLds L;
  unsigned long long s_next = 0, s_end = 0; /* this wave's survivor slots */
  const int nq = A.live ? A.nqueues : 1;
  const int lane = threadIdx.x & 63;
  if (A.counter == nullptr) {
    /* one item per block, queues back to back */
    unsigned v = blockIdx.x;
    for (int x = 0; x < nq; ++x) {
      const unsigned n = A.live ? A.live_cnt[x] : (unsigned)A.nitems;
      if (v < n) {
        if (A.live) {
          const int4 r = A.live_item[A.qseg[x] + v];
          run_cand<MODE, KS>(L, A, A.live[A.qseg[x] + v], Item{r.x, r.y, r.z, r.w}, s_next, s_end);
        } else {
          run_cand<MODE, KS>(L, A, (int)v, A.items[v], s_next, s_end);
        }
        break;
      }
      v -= n;
    }
  } else if (A.live && nq == 1) {
    /* persistent, one queue.  Block b's first three items are items b,
     * b + G and b + 2G (G = gridDim.x), and the counter hands out the
     * items from 3G on: every wave starts at the launch, and one atomic per
     * wave on a single address there would hold the first items back
     * (atomics on one address complete one after another).  While an item
     * runs, the next item's record is loaded, and from the second item on
     * the counter is advanced for the one after; a wave's first atomic is
     * issued after its first item, when the waves no longer run in step,
     * and its result is needed only after the second. */
    const unsigned n = A.live_cnt[0], G = gridDim.x;
    const int base = A.qseg[0];
    const unsigned v = blockIdx.x;
    if (v < n) {
      int idx = A.live[base + v];
      int4 rec = A.live_item[base + v];
      unsigned vn = v + G;
      for (int k = 0;; ++k) {
        int nidx = 0;
        int4 nrec = make_int4(0, 0, 0, 0);
        unsigned raw = 0;
        if (vn < n) {
          nidx = A.live[base + vn];
          nrec = A.live_item[base + vn];
          if (k > 0 && lane == 0) raw = atomicAdd(A.counter, 1u);
        }
        run_cand<MODE, KS>(L, A, idx, Item{rec.x, rec.y, rec.z, rec.w}, s_next, s_end);
        if (vn >= n) break;
        idx = nidx;
        rec = nrec;
        vn = k == 0 ? v + 2 * G : __shfl(raw, 0, 64) + 3 * G;
      }
    }
  } else if (A.live) {
    /* persistent over several queues: the block's own queue first, then
     * the others.  The next item's record is loaded, and the work counter
     * advanced for the one after, while the current item runs, so neither
     * round trip is on an item's critical path. */
    const int x0 = (int)(blockIdx.x % nq);
    for (int k = 0; k < nq; ++k) {
      const int x = (x0 + k) % nq;
      const unsigned n = A.live_cnt[x];
      const int base = A.qseg[x];
      unsigned raw = 0;
      if (lane == 0) raw = atomicAdd(A.counter + x, 1u);
      const unsigned v = __shfl(raw, 0, 64);
      if (v >= n) continue;
      int idx = A.live[base + v];
      int4 rec = A.live_item[base + v];
      if (lane == 0) raw = atomicAdd(A.counter + x, 1u);
      for (;;) {
        const unsigned vn = __shfl(raw, 0, 64);
        int nidx = 0;
        int4 nrec = make_int4(0, 0, 0, 0);
        if (vn < n) {
          nidx = A.live[base + vn];
          nrec = A.live_item[base + vn];
          if (lane == 0) raw = atomicAdd(A.counter + x, 1u);
        }
        run_cand<MODE, KS>(L, A, idx, Item{rec.x, rec.y, rec.z, rec.w}, s_next, s_end);
        if (vn >= n) break;
        idx = nidx;
        rec = nrec;
      }
    }
  } else {
    for (;;) {
      unsigned v = 0;
      if (lane == 0) v = atomicAdd(A.counter, 1u);
      v = __shfl(v, 0, 64);
      if (v >= (unsigned)A.nitems) break;
      run_cand<MODE, KS>(L, A, (int)v, A.items[v], s_next, s_end);
    }
  }
  fill_sentinels(A, s_next, s_end);
}

/* Clip, faces and force of every survivor (grid-stride); grid.y selects the
 * tangent direction in the Dual instantiation. */
template <class T, int KS>
__global__ __launch_bounds__(256, 2) void k_clip(QueryArgs A, ClipArgs C, int dir_base, Contrib* out,
                                              int64_t out_stride) {
  /* always passed by address (a null record pointer switches its output
   * off), so the struct stays in registers */
  const OutPtrs o{A.face_rec, A.face_key, A.face_cnt, A.face_cap, A.poly_rec,
                  A.poly_key, A.poly_cnt, A.poly_cap, A.poly_tan, A.poly_env};
  const OutPtrs* op = &o;
  const int dir = dir_base < 0 ? -1 : dir_base + (int)blockIdx.y;
  Contrib* dst = out + (dir < 0 ? 0 : (int64_t)(dir - dir_base) * out_stride);
  const unsigned long long cnt = *A.surv_cnt;
  const int64_t n = (int64_t)cnt < A.surv_cap ? (int64_t)cnt : A.surv_cap;
  /* the double build clips in LDS, one column of HCS_MAXV vertices per
   * thread (48 KB per block) */
  constexpr bool kCol = std::is_same<T, double>::value;
  __shared__ double col[kCol ? 3 * HCS_MAXV * 256 : 1];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    clip_survivor<T, KS, kCol>(A, C, i, dir, dst, op, col + threadIdx.x, 256);
}

/* Per pair, in K_e waves: wave w = (pair e, slice j, K_e, first partial)
 * from the plan's table sums slice j of the pair's contributions (a fixed
 * split of its slot range) lane-strided in order, then a fixed butterfly,
 * and the slice's items' candidate counts likewise; grid.y = direction.
 * k_env_combine adds a pair's slices in order, so results do not depend on
 * timing, and a heavy pair (K_e grows with its item count) no longer sets
 * the time of the whole reduction. */
__global__ __launch_bounds__(64) void k_env_reduce(const EnvDesc* __restrict__ envs, const int4* __restrict__ rtab,
                                                   int nred, const unsigned* __restrict__ off,
                                                   const int* __restrict__ item_ncand,
                                                   const unsigned long long* __restrict__ surv_cnt, int64_t surv_cap,
                                                   const Contrib* __restrict__ contrib, int64_t stride,
                                                   RedPart* __restrict__ part) {
  const int4 w = rtab[blockIdx.x];
  const int e = w.x, j = w.y, K = w.z, d = blockIdx.y, lane = threadIdx.x;
  const int first = envs[e].item_first, cnt = envs[e].item_count;
  /* on survivor-list overflow the offsets run past the contribution slots:
   * sum nothing, the status reports the overflow */
  const bool over = (int64_t)*surv_cnt > surv_cap;
  /* (slots never reach past the buffer; the clamp keeps every read inside
   * it whatever the offsets hold) */
  const unsigned capu = (unsigned)(surv_cap < 0xffffffffLL ? surv_cap : 0xffffffffLL);
  const unsigned lo = min(off[first], capu), hi = over ? lo : max(lo, min(off[first + cnt], capu));
  const uint64_t len = hi - lo;
  const unsigned s0 = lo + (unsigned)(len * j / K), s1 = lo + (unsigned)(len * (j + 1) / K);
  const int i0 = first + (int)((int64_t)cnt * j / K), i1 = first + (int)((int64_t)cnt * (j + 1) / K);
  const Contrib* cb = contrib + (int64_t)d * stride;
  double F[3] = {0, 0, 0}, tA[3] = {0, 0, 0}, tB[3] = {0, 0, 0};
  int64_t np = 0, nf = 0, nc = 0;
  auto add = [&](const Contrib& q) {
#pragma unroll
    for (int c = 0; c < 3; ++c) F[c] = F[c] + q.F[c], tA[c] = tA[c] + q.tA[c], tB[c] = tB[c] + q.tB[c];
    np += q.np, nf += q.nf;
  };
  /* lane-strided in slot order; four slots loaded before they are added,
   * so four loads are in flight per lane */
  unsigned i = s0 + lane;
  for (; i < s1 && s1 - i > 192; i += 256) {
    const Contrib q0 = cb[i], q1 = cb[i + 64], q2 = cb[i + 128], q3 = cb[i + 192];
    add(q0), add(q1), add(q2), add(q3);
  }
  for (; i < s1; i += 64) add(cb[i]);
  for (int i = i0 + lane; i < i1; i += 64) nc += item_ncand[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) F[c] = wave_sum(F[c]), tA[c] = wave_sum(tA[c]), tB[c] = wave_sum(tB[c]);
  np = wave_sum_i(np), nf = wave_sum_i(nf), nc = wave_sum_i(nc);
  if (lane == 0) {
    RedPart r;
#pragma unroll
    for (int c = 0; c < 3; ++c) r.F[c] = F[c], r.tA[c] = tA[c], r.tB[c] = tB[c];
    r.np = np, r.nf = nf, r.nc = nc;
    part[(int64_t)d * nred + w.w + j] = r;
  }
}

/* One thread per (pair, direction): the pair's slices in order. */
__global__ void k_env_combine(const int2* __restrict__ ptab, int nred, const RedPart* __restrict__ part,
                              const unsigned long long* __restrict__ surv_cnt, int64_t surv_cap,
                              const int* __restrict__ err, PairOut* __restrict__ out, int n, int ndir) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * ndir) return;
  const int e = (int)(t % n), d = (int)(t / n);
  const int2 pe = ptab[e];
  const RedPart* p = part + (int64_t)d * nred + pe.x;
  double F[3] = {0, 0, 0}, tA[3] = {0, 0, 0}, tB[3] = {0, 0, 0};
  int64_t np = 0, nf = 0, nc = 0;
  auto add = [&](const RedPart& q) {
#pragma unroll
    for (int c = 0; c < 3; ++c) F[c] = F[c] + q.F[c], tA[c] = tA[c] + q.tA[c], tB[c] = tB[c] + q.tB[c];
    np += q.np, nf += q.nf, nc += q.nc;
  };
  /* in slice order; four slices loaded before they are added */
  int j = 0;
  for (; j + 3 < pe.y; j += 4) {
    const RedPart q0 = p[j], q1 = p[j + 1], q2 = p[j + 2], q3 = p[j + 3];
    add(q0), add(q1), add(q2), add(q3);
  }
  for (; j < pe.y; ++j) add(p[j]);
  const bool over = (int64_t)*surv_cnt > surv_cap;
  PairOut r;
#pragma unroll
  for (int c = 0; c < 3; ++c) r.F_A[c] = F[c], r.tau_A[c] = tA[c], r.F_B[c] = -F[c], r.tau_B[c] = tB[c];
  r.n_candidates = nc, r.n_polygons = np, r.n_faces = nf;
  r.status = (err && *err) ? HCS_ERR_LIMIT : (over ? HCS_ERR_CAPACITY : 0);
  out[t] = r;
}

void launch_prep(hipStream_t s, const EnvDesc* envs, const double* state, double* xf, int n) {
  if (n > 0) hipLaunchKernelGGL(k_prep, dim3((n + 127) / 128), dim3(128), 0, s, envs, state, xf, n);
}
void launch_fill_u32(hipStream_t s, unsigned int* p, unsigned int v, int n) {
  if (n > 0) hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
}

/* Pipeline buffers (grow-only) and the enqueue of one pass. */
int pipe_reserve(Ctx* c, Pipe& p, int nitems, int npairs, int64_t cap, int ndir) {
  int rc;
  const size_t ni = (size_t)std::max(1, nitems);
  if ((rc = ensure(c, p.counter, 64)) || (rc = ensure(c, p.surv_cnt, 64)) || (rc = ensure(c, p.err, 64)) ||
      (rc = ensure(c, p.xf, sizeof(double) * 12 * std::max(1, npairs))) ||
      (rc = ensure(c, p.ncand, sizeof(int) * ni)) || (rc = ensure(c, p.nsurv, sizeof(unsigned) * (ni + 1))) ||
      (rc = ensure(c, p.off, sizeof(unsigned) * (ni + 1))) ||
      (rc = ensure(c, p.out, sizeof(PairOut) * std::max(1, npairs) * std::max(1, ndir))))
    return rc;
  cap = std::max<int64_t>(cap, 4096);
  if ((rc = ensure(c, p.surv, sizeof(int4) * cap)) ||
      (rc = ensure(c, p.contrib, sizeof(Contrib) * cap * std::max(1, ndir))))
    return rc;
  p.cap = cap;
  const size_t bytes = scan_temp_bytes((long long)ni + 1);
  if ((rc = ensure(c, p.scan_tmp, std::max<size_t>(bytes, 8)))) return rc;
  p.scan_bytes = bytes;
  return HCS_OK;
}

QueryArgs pipe_args(const Pipe& p, const EnvDesc* envs, const double* state, const Item* items, int nitems,
                    bool persistent) {
  QueryArgs a{};
  a.envs = envs;
  a.state = state;
  a.xf = (const double*)p.xf.ptr;
  a.items = items;
  a.nitems = nitems;
  a.queries_per_lane = 1;
  a.counter = persistent ? (unsigned*)p.counter.ptr : nullptr;
  a.error = (int*)p.err.ptr;
  a.surv = (int4*)p.surv.ptr;
  a.surv_cnt = (unsigned long long*)p.surv_cnt.ptr;
  a.surv_cap = p.cap;
  a.item_ncand = (int*)p.ncand.ptr;
  a.item_nsurv = (unsigned*)p.nsurv.ptr;
  a.live = p.live.ptr ? (const int*)p.live.ptr : nullptr;
  a.live_cnt = p.live.ptr ? (const unsigned*)p.counter.ptr + HCS_QUEUES : nullptr;
  a.qseg = p.live.ptr ? (const int*)p.qseg.ptr : nullptr;
  a.nqueues = p.nqueues;
  a.live_item = p.live.ptr ? (const int4*)p.live_item.ptr : nullptr;
  a.xtri = (const double*)p.xtri.ptr;
  a.kinds = p.kinds;
  return a;
}

void pipe_front(Ctx* c, Pipe& p, const QueryArgs& a, int npairs) {
  hipStream_t s = c->stream;
  launch_fill_u32(s, (unsigned*)p.counter.ptr, 0u, 2 * HCS_QUEUES);
  launch_fill_u32(s, (unsigned*)p.surv_cnt.ptr, 0u, 2);
  launch_fill_u32(s, (unsigned*)p.err.ptr, 0u, 1);
  launch_fill_u32(s, (unsigned*)p.nsurv.ptr + a.nitems, 0u, 1);
  launch_prep(s, a.envs, a.state, (double*)p.xf.ptr, npairs);
  if (p.nxenv > 0)
    hipLaunchKernelGGL(k_xtri, dim3(p.nxenv), dim3(64), 0, s, a.envs, (const int*)p.xenv.ptr,
                       (const double*)p.xf.ptr, (double*)p.xtri.ptr);
  if (a.nitems > 0 && a.live)
    hipLaunchKernelGGL(k_cull, dim3((a.nitems + HCS_CULL_THREADS - 1) / HCS_CULL_THREADS), dim3(HCS_CULL_THREADS), 0,
                       s, a.envs, a.items,
                       p.ordered ? (const int*)p.order.ptr : nullptr, (const Item*)p.item_o.ptr, a.nitems,
                       (const double*)p.ibox.ptr, (const double*)p.xf.ptr, (const int*)p.qseg.ptr, p.nqueues,
                       (int*)p.live.ptr, (int4*)p.live_item.ptr, (unsigned*)p.counter.ptr + HCS_QUEUES,
                       a.item_ncand, a.item_nsurv);
}

/* The kind a launch is specialised for: the one kind all its pairs share,
 * triangle or soft queries (the common single-kind batches), else -1. */
static int launch_kind(const QueryArgs& a) {
  if (a.kinds == (1 << KIND_TRI)) return KIND_TRI;
  if (a.kinds == (1 << KIND_SOFT)) return KIND_SOFT;
  return -1;
}

template <int MODE>
static void launch_cand_mode(hipStream_t s, const QueryArgs& a, int grid) {
  switch (launch_kind(a)) {
    case KIND_TRI: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cand<MODE, KIND_TRI>), dim3(grid), dim3(64), 0, s, a); break;
    case KIND_SOFT: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cand<MODE, KIND_SOFT>), dim3(grid), dim3(64), 0, s, a); break;
    default: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cand<MODE, -1>), dim3(grid), dim3(64), 0, s, a); break;
  }
}

void launch_cand(hipStream_t s, const QueryArgs& a, int grid, int mode) {
  if (a.nitems <= 0 || grid <= 0) return;
  if (mode == CAND_KEYS) launch_cand_mode<CAND_KEYS>(s, a, grid);
  else if (mode == CAND_COUNT) launch_cand_mode<CAND_COUNT>(s, a, grid);
  else launch_cand_mode<CAND_STEP>(s, a, grid);
}

/* Scan of the per-item survivor counts, [event e_clip], clip (double, or
 * Dual over ndir directions), [event e_reduce], per-pair reduction. */
static void pipe_back(Ctx* c, Pipe& p, const QueryArgs& a, int npairs, int clip_grid, int ndir,
                      const double* dstar_dE, const double* invE, int dir0, hipEvent_t e_clip, hipEvent_t e_reduce) {
  hipStream_t s = c->stream;
  /* sizes were checked when the buffers were reserved */
  (void)exclusive_scan_u32(s, (const unsigned*)p.nsurv.ptr, (unsigned*)p.off.ptr, (long long)a.nitems + 1,
                           p.scan_tmp.ptr, p.scan_bytes);
  if (e_clip) (void)hipEventRecord(e_clip, s);
  ClipArgs C{(const unsigned*)p.off.ptr, dstar_dE, invE, p.cap};
  const int g = std::max(1, clip_grid);
  const int kind = launch_kind(a);
  if (ndir > 0)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_clip<Dual, -1>), dim3(g, ndir), dim3(256), 0, s, a, C, dir0,
                       (Contrib*)p.contrib.ptr, (int64_t)p.cap);
  else if (kind == KIND_TRI)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_clip<double, KIND_TRI>), dim3(g), dim3(256), 0, s, a, C, -1,
                       (Contrib*)p.contrib.ptr, (int64_t)0);
  else if (kind == KIND_SOFT)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_clip<double, KIND_SOFT>), dim3(g), dim3(256), 0, s, a, C, -1,
                       (Contrib*)p.contrib.ptr, (int64_t)0);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_clip<double, -1>), dim3(g), dim3(256), 0, s, a, C, -1,
                       (Contrib*)p.contrib.ptr, (int64_t)0);
  if (e_reduce) (void)hipEventRecord(e_reduce, s);
  if (npairs > 0 && p.nred > 0) {
    const int nd = std::max(1, ndir);
    hipLaunchKernelGGL(k_env_reduce, dim3(p.nred, nd), dim3(64), 0, s, a.envs, (const int4*)p.rtab.ptr, p.nred,
                       (const unsigned*)p.off.ptr, (const int*)p.ncand.ptr, (const unsigned long long*)p.surv_cnt.ptr,
                       p.cap, (const Contrib*)p.contrib.ptr, (int64_t)p.cap, (RedPart*)p.part.ptr);
    hipLaunchKernelGGL(k_env_combine, dim3((npairs * nd + 127) / 128), dim3(128), 0, s, (const int2*)p.ptab.ptr,
                       p.nred, (const RedPart*)p.part.ptr, (const unsigned long long*)p.surv_cnt.ptr, p.cap,
                       (const int*)p.err.ptr, (PairOut*)p.out.ptr, npairs, nd);
  }
}

void pipe_enqueue(Ctx* c, Pipe& p, const QueryArgs& a, int npairs, int cand_grid, int clip_grid, int ndir,
                  const double* dstar_dE, const double* invE, bool cand_only, int dir0) {
  pipe_front(c, p, a, npairs);
  launch_cand(c->stream, a, cand_grid, a.cand_key ? CAND_KEYS : CAND_STEP);
  if (cand_only) return;
  pipe_back(c, p, a, npairs, clip_grid, ndir, dstar_dE, invE, dir0, nullptr, nullptr);
}

void launch_clip_reduce(Ctx* c, Pipe& p, const QueryArgs& a, int npairs, int clip_grid, hipEvent_t e_clip,
                        hipEvent_t e_reduce) {
  pipe_back(c, p, a, npairs, clip_grid, 0, nullptr, nullptr, 0, e_clip, e_reduce);
}

/* Reduction slices: a pair of n items gets ceil(n / 8) waves (at most 64;
 * a split heavy pair has many items).  rt: per wave (pair, slice, slices,
 * first partial of the pair); pt: per pair (first partial, slices). */
void reduction_slices(const Plan& plan, std::vector<int4>& rt, std::vector<int2>& pt) {
  rt.clear();
  pt.assign(std::max<size_t>(1, plan.envs.size()), make_int2(0, 0));
  for (size_t e = 0; e < plan.envs.size(); ++e) {
    const int K = std::max(1, std::min(64, (plan.envs[e].item_count + 7) / 8));
    pt[e] = make_int2((int)rt.size(), K);
    for (int j = 0; j < K; ++j) rt.push_back(make_int4((int)e, j, K, (int)rt.size() - j));
  }
}

int pipe_plan(Ctx* c, Pipe& p, const EnvDesc* envs, const Item* items, const Plan& plan) {
  int rc;
  const std::vector<Item>& host_items = plan.items;
  const int nitems = (int)host_items.size();
  const size_t ni = (size_t)std::max(1, nitems);
  if ((rc = ensure(c, p.ibox, sizeof(double) * 12 * ni)) || (rc = ensure(c, p.live, sizeof(int) * ni)) ||
      (rc = ensure(c, p.live_item, sizeof(Item) * ni)) || (rc = ensure(c, p.qseg, sizeof(int) * (HCS_QUEUES + 1))) ||
      (rc = ensure(c, p.xtri, sizeof(double) * HCS_XTRI * std::max<int64_t>(1, plan.ntx))) ||
      (rc = ensure(c, p.xenv, sizeof(int) * std::max<size_t>(1, plan.xenv.size()))))
    return rc;
  /* queue segments of the live list: queue x holds the items of pairs
   * e % nqueues == x */
  p.nqueues = std::max(1, std::min(p.nqueues, HCS_QUEUES));
  std::vector<int> seg(HCS_QUEUES + 1, 0);
  for (const Item& it : host_items) ++seg[it.env % p.nqueues + 1];
  for (int x = 0; x < HCS_QUEUES; ++x) seg[x + 1] += seg[x];
  p.qseg_host = seg;
  p.nxenv = (int)plan.xenv.size();
  p.kinds = 0;
  for (const EnvDesc& E : plan.envs) p.kinds |= 1 << E.kind;
  {
    std::vector<int4> rt;
    std::vector<int2> pt;
    reduction_slices(plan, rt, pt);
    p.nred = (int)rt.size();
    if ((rc = ensure(c, p.rtab, sizeof(int4) * std::max<size_t>(1, rt.size()))) ||
        (rc = ensure(c, p.ptab, sizeof(int2) * pt.size())) ||
        (rc = ensure(c, p.part, sizeof(RedPart) * HCS_MAXDIR * std::max<size_t>(1, rt.size()))))
      return rc;
    p.rtab_host = rt;
    p.ptab_host = pt;
    if (!rt.empty())
      rc = hip_check(c, hipMemcpyAsync(p.rtab.ptr, p.rtab_host.data(), sizeof(int4) * rt.size(), hipMemcpyHostToDevice,
                                       c->stream), "upload reduction slices");
    if (!rc)
      rc = hip_check(c, hipMemcpyAsync(p.ptab.ptr, p.ptab_host.data(), sizeof(int2) * pt.size(),
                                       hipMemcpyHostToDevice, c->stream), "upload reduction slices");
    if (rc) return rc;
  }
  p.ordered = (int)plan.order.size() == nitems && nitems > 0;
  if (p.ordered) {
    if ((rc = ensure(c, p.order, sizeof(int) * ni)) || (rc = ensure(c, p.item_o, sizeof(Item) * ni))) return rc;
    p.item_o_host.resize(nitems);
    for (int k = 0; k < nitems; ++k) p.item_o_host[k] = host_items[plan.order[k]];
    rc = hip_check(c, hipMemcpyAsync(p.order.ptr, plan.order.data(), sizeof(int) * nitems, hipMemcpyHostToDevice,
                                     c->stream), "upload order");
    if (!rc)
      rc = hip_check(c, hipMemcpyAsync(p.item_o.ptr, p.item_o_host.data(), sizeof(Item) * nitems,
                                       hipMemcpyHostToDevice, c->stream), "upload ordered items");
    if (rc) return rc;
  }
  rc = hip_check(c, hipMemcpyAsync(p.qseg.ptr, p.qseg_host.data(), sizeof(int) * (HCS_QUEUES + 1),
                                   hipMemcpyHostToDevice, c->stream), "upload queues");
  if (!rc && p.nxenv > 0)
    rc = hip_check(c, hipMemcpyAsync(p.xenv.ptr, plan.xenv.data(), sizeof(int) * p.nxenv, hipMemcpyHostToDevice,
                                     c->stream), "upload tet-query pairs");
  if (rc) return rc;
  if (nitems > 0)
    hipLaunchKernelGGL(k_item_boxes, dim3(nitems), dim3(64), 0, c->stream, envs, items,
                       p.ordered ? (const int*)p.order.ptr : nullptr, nitems, (double*)p.ibox.ptr);
  rc = hip_check(c, hipGetLastError(), "item boxes");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  return rc;
}

void pipe_free(Pipe& p) {
  for (DBuf* b : {&p.counter, &p.surv, &p.surv_cnt, &p.ncand, &p.nsurv, &p.off, &p.scan_tmp, &p.contrib, &p.out, &p.err,
                  &p.xf, &p.ibox, &p.live, &p.live_item, &p.qseg, &p.xtri, &p.xenv, &p.order, &p.item_o, &p.part, &p.rtab,
                  &p.ptab})
    if (b->ptr) (void)hipFree(b->ptr), b->ptr = nullptr, b->bytes = 0;
}

int split_roots(Ctx* c, const EnvDesc* envs, const Item* items, const std::vector<int>& req,
                const std::vector<int>& depth, std::vector<int32_t>& out) {
  const int n = (int)req.size();
  out.assign(65 * (size_t)n, 0);
  if (n == 0) return HCS_OK;
  DBuf dreq, ddep, dout;
  auto done = [&](int rc) {
    for (DBuf* b : {&dreq, &ddep, &dout})
      if (b->ptr) (void)hipFree(b->ptr);
    return rc;
  };
  int rc;
  if ((rc = ensure(c, dreq, sizeof(int) * n)) || (rc = ensure(c, ddep, sizeof(int) * n)) ||
      (rc = ensure(c, dout, sizeof(int32_t) * out.size())))
    return done(rc);
  rc = hip_check(c, hipMemcpyAsync(dreq.ptr, req.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream), "ul");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(ddep.ptr, depth.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream), "ul");
  if (rc) return done(rc);
  hipLaunchKernelGGL(k_split_roots, dim3((n + 63) / 64), dim3(64), 0, c->stream, envs, items, (const int*)dreq.ptr,
                     (const int*)ddep.ptr, n, (int32_t*)dout.ptr);
  rc = hip_check(c, hipGetLastError(), "split roots");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(out.data(), dout.ptr, sizeof(int32_t) * out.size(), hipMemcpyDeviceToHost,
                                     c->stream), "dl");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  return done(rc);
}

/* ===================================================================
 * Host side
 * =================================================================== */

static bool is_rigid(const hcs_props& p) {
  return !(p.hydroelastic_modulus > 0) || std::isinf(p.hydroelastic_modulus);
}

static int upload(Ctx* c, void** dst, const void* src, size_t bytes) {
  HCHK(c, hipMalloc(dst, std::max<size_t>(bytes, 8)));
  if (bytes) HCHK(c, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return HCS_OK;
}

static int build_bvh(Ctx* c, Mesh& m) {
  const int n = m.nelem;
  m.height = 0;
  if (n == 1) {
    HCHK(c, hipMalloc((void**)&m.nodes, sizeof(Node)));
    const int32_t zero = 0;
    int rc = upload(c, (void**)&m.leaf, &zero, sizeof(int32_t));
    if (!rc) rc = hip_check(c, hipMalloc((void**)&m.boxL, 6 * sizeof(double)), "hipMalloc");
    if (!rc)
      rc = hip_check(c, hipMemcpyAsync(m.boxL, m.box, 6 * sizeof(double), hipMemcpyDeviceToDevice, c->stream), "copy");
    return rc ? rc : hip_check(c, hipStreamSynchronize(c->stream), "sync");
  }
  /* Scene bounds from the vertices (they bound every tet AABB). */
  std::vector<double> Vh(3 * (size_t)m.nv);
  HCHK(c, hipMemcpyAsync(Vh.data(), m.V, Vh.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCHK(c, hipStreamSynchronize(c->stream));
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int v = 0; v < m.nv; ++v)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], Vh[3 * v + k]), hi[k] = std::max(hi[k], Vh[3 * v + k]);
  double sc[3];
  for (int k = 0; k < 3; ++k) sc[k] = hi[k] > lo[k] ? 1024.0 / (hi[k] - lo[k]) : 0.0;
  unsigned *codes = nullptr, *codes2 = nullptr;
  int *idx = nullptr, *idx2 = nullptr, *np = nullptr, *lp = nullptr, *flags = nullptr, *rh = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = HCS_OK;
  auto fail = [&](hipError_t e, const char* w) {
    if (rc == HCS_OK) rc = hip_check(c, e, w);
  };
  fail(hipMalloc((void**)&codes, n * sizeof(unsigned)), "hipMalloc");
  fail(hipMalloc((void**)&codes2, n * sizeof(unsigned)), "hipMalloc");
  fail(hipMalloc((void**)&idx, n * sizeof(int)), "hipMalloc");
  fail(hipMalloc((void**)&idx2, n * sizeof(int)), "hipMalloc");
  fail(hipMalloc((void**)&np, n * sizeof(int)), "hipMalloc");
  fail(hipMalloc((void**)&lp, n * sizeof(int)), "hipMalloc");
  fail(hipMalloc((void**)&flags, n * sizeof(int)), "hipMalloc");
  fail(hipMalloc((void**)&rh, sizeof(int)), "hipMalloc");
  fail(hipMalloc((void**)&m.nodes, (size_t)(n - 1) * sizeof(Node)), "hipMalloc");
  if (rc == HCS_OK) {
    hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, c->stream, m.box, n, lo[0], lo[1], lo[2],
                       sc[0], sc[1], sc[2], codes, idx);
    fail(hipGetLastError(), "k_morton");
    tmp_bytes = sort_temp_bytes(n, sizeof(unsigned));
    fail(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 8)), "hipMalloc");
  }
  if (rc == HCS_OK) {
    if (radix_sort_pairs_u32(c->stream, codes, codes2, idx, idx2, n, 0, 30, tmp, tmp_bytes) != 0)
      fail(hipErrorInvalidValue, "radix sort size");
  }
  if (rc == HCS_OK) {
    /* the leaf order must be a permutation in code order before the tree
     * is built on it */
    std::vector<unsigned> hc(n);
    std::vector<int> hi(n);
    fail(hipMemcpyAsync(hc.data(), codes2, n * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream), "memcpy");
    fail(hipMemcpyAsync(hi.data(), idx2, n * sizeof(int), hipMemcpyDeviceToHost, c->stream), "memcpy");
    fail(hipStreamSynchronize(c->stream), "sync");
    if (rc == HCS_OK) {
      std::vector<char> seen(n, 0);
      bool ok = true;
      for (int i = 0; i < n && ok; ++i) {
        ok = hi[i] >= 0 && hi[i] < n && !seen[hi[i]] && (i == 0 || hc[i - 1] <= hc[i]);
        if (ok) seen[hi[i]] = 1;
      }
      if (!ok) rc = set_err(c, HCS_ERR_HIP, "LBVH build: the Morton sort returned an invalid order");
    }
  }
  if (rc == HCS_OK) {
    fail(hipMemsetAsync(flags, 0, n * sizeof(int), c->stream), "memset");
    hipLaunchKernelGGL(k_karras, dim3((n - 1 + 255) / 256), dim3(256), 0, c->stream, codes2, idx2, n, m.nodes, np,
                       lp);
    fail(hipGetLastError(), "k_karras");
    hipLaunchKernelGGL(k_refit, dim3((n + 255) / 256), dim3(256), 0, c->stream, m.box, idx2, n, m.nodes, np, lp,
                       flags, rh);
    fail(hipGetLastError(), "k_refit");
    int h = 0;
    fail(hipMemcpyAsync(&h, rh, sizeof(int), hipMemcpyDeviceToHost, c->stream), "memcpy");
    fail(hipStreamSynchronize(c->stream), "sync");
    m.height = h;
  }
  if (rc == HCS_OK) {
    fail(hipMalloc((void**)&m.wnodes, (size_t)(n - 1) * sizeof(WNode)), "hipMalloc");
    if (rc == HCS_OK) {
      hipLaunchKernelGGL(k_wide_nodes, dim3((n - 1 + 255) / 256), dim3(256), 0, c->stream, m.nodes, n - 1, m.wnodes);
      fail(hipGetLastError(), "k_wide_nodes");
    }
  }
  if (rc == HCS_OK) {
    fail(hipMalloc((void**)&m.boxL, 6 * sizeof(double) * (size_t)n), "hipMalloc");
    if (rc == HCS_OK) {
      hipLaunchKernelGGL(k_gather_boxes, dim3((n + 255) / 256), dim3(256), 0, c->stream, m.box, idx2, n, m.boxL);
      fail(hipGetLastError(), "k_gather_boxes");
      fail(hipStreamSynchronize(c->stream), "sync");
    }
  }
  if (rc == HCS_OK) {
    m.leaf = idx2; /* kept: Morton order of the tets */
    idx2 = nullptr;
  }
  for (void* ptr : {(void*)codes, (void*)codes2, (void*)idx, (void*)idx2, (void*)np, (void*)lp, (void*)flags,
                    (void*)rh, tmp})
    if (ptr) (void)hipFree(ptr);
  if (rc == HCS_OK && 3 * ((m.height + 1) / 2) + 1 > HCS_STACK)
    rc = set_err(c, HCS_ERR_LIMIT, "LBVH height " + std::to_string(m.height) + " exceeds the traversal stack");
  return rc;
}

static void free_mesh(Mesh& m) {
  for (void* p : {(void*)m.V, (void*)m.E, (void*)m.p, (void*)m.rec, (void*)m.box, (void*)m.nodes, (void*)m.nrm,
                  (void*)m.leaf, (void*)m.boxL, (void*)m.qorder, (void*)m.fpl, (void*)m.wnodes})
    if (p) (void)hipFree(p);
  m = Mesh();
}

int build_plan(Ctx* c, int npairs, const int32_t* geoms, const hcs_step_opts* o, int tet_max, Plan& plan) {
  plan.envs.assign(npairs, EnvDesc{});
  plan.items.clear();
  plan.xenv.clear();
  plan.ntx = 0;
  plan.pair_kind.assign(npairs, 0);
  const int per_item = 64;
  for (int e = 0; e < npairs; ++e) {
    const int ha = geoms[2 * e], hb = geoms[2 * e + 1];
    if (ha < 0 || hb < 0 || ha >= (int)c->meshes.size() || hb >= (int)c->meshes.size() ||
        c->meshes[ha].kind == 0 || c->meshes[hb].kind == 0)
      return set_err(c, HCS_ERR_ARG, "pair " + std::to_string(e) + ": invalid mesh handle");
    const Mesh& A = c->meshes[ha];
    const Mesh& B = c->meshes[hb];
    if (A.kind != 1)
      return set_err(c, HCS_ERR_ARG,
                     "pair " + std::to_string(e) + ": geom_a must be soft (pass soft-rigid pairs soft first)");
    EnvDesc& E = plan.envs[e];
    E.recA = hcs_gptr(A.rec), E.fplA = hcs_gptr((const float4*)A.fpl), E.boxA = hcs_gptr(A.box),
    E.nodesA = hcs_gptr(A.nodes), E.wnodesA = hcs_gptr(A.wnodes), E.leafA = hcs_gptr(A.leaf),
    E.boxLA = hcs_gptr(A.boxL), E.TA = hcs_gptr(A.E), E.VA = hcs_gptr(A.V),
    E.ntA = A.nelem;
    E.xtri_off = -1;
    E.VB = hcs_gptr(B.V), E.FB = hcs_gptr(B.E), E.nrmB = hcs_gptr(B.nrm), E.recB = hcs_gptr(B.rec), E.nB = B.nelem,
    E.nvB = B.nv;
    E.qorder = hcs_gptr(B.kind == 2 ? B.qorder : (B.kind == 1 ? B.leaf : nullptr));
    E.b_soft = B.kind == 1;
    E.a_is_M = A.geom_id < B.geom_id;
    E.repr = o ? o->representation : 0;
    E.apply = o ? o->apply_forces : 1;
    E.dt = o ? o->dt : 0.001;
    const double inf = std::numeric_limits<double>::infinity();
    const double EA = A.props.hydroelastic_modulus, dA = A.props.dissipation;
    const double EB = E.b_soft ? B.props.hydroelastic_modulus : inf, dB = E.b_soft ? B.props.dissipation : 1.0;
    /* calcCombinedDissipation (plugin.cpp:138-159) */
    {
      double Estar = (EA == inf) ? EB : (EB == inf ? EA : EA * EB / (EA + EB));
      double ds;
      if (Estar == inf) {
        ds = 0.5 * (dA + dB);
      } else {
        ds = 0;
        if (EA != inf) ds += Estar / EA * dA;
        if (EB != inf) ds += Estar / EB * dB;
      }
      E.dstar = ds;
    }
    /* CalcContactFrictionFromSurfaceProperties, dynamic coefficient (plugin.cpp:427-433) */
    {
      const double m1 = A.props.dynamic_friction, m2 = B.props.dynamic_friction;
      const double den = m1 + m2;
      E.mu = den == 0.0 ? 0.0 : 2 * m1 * m2 / den;
    }
    int nq, chunk;
    if (B.kind == 3) {
      /* rigid half-space (plane geom): A's tets in Morton chunks */
      E.kind = KIND_PLANE;
      nq = A.nelem;
      chunk = 64;
    } else if (E.b_soft) {
      E.kind = KIND_SOFT;
      nq = B.nelem;
      chunk = per_item;
    } else if (B.nelem <= tet_max) {
      E.kind = KIND_TET;
      nq = A.nelem;
      chunk = 64;
      E.xtri_off = plan.ntx;
      plan.ntx += B.nelem;
      plan.xenv.push_back(e);
    } else {
      E.kind = KIND_TRI;
      nq = B.nelem;
      chunk = per_item;
    }
    plan.pair_kind[e] = E.kind;
    E.item_first = (int)plan.items.size();
    for (int q = 0; q < nq; q += chunk) plan.items.push_back(Item{e, q, std::min(chunk, nq - q), 0});
    /* (root 0: the whole LBVH of A) */
    E.item_count = (int)plan.items.size() - E.item_first;
  }
  return HCS_OK;
}

}  // namespace hcs

using namespace hcs;

/* ===================================================================
 * C-ABI
 * =================================================================== */

extern "C" {

int hcs_create(hcs_ctx** out, int device) {
  if (!out) return HCS_ERR_ARG;
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    g_create_err = "no HIP device available";
    (void)hipGetLastError();
    return HCS_ERR_HIP;
  }
  if (device < 0 || device >= count) {
    g_create_err = "device ordinal out of range";
    return HCS_ERR_ARG;
  }
  hcs_ctx* h = new (std::nothrow) hcs_ctx();
  if (!h) return HCS_ERR_OOM;
  Ctx* c = &h->c;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    g_create_err = "hipSetDevice/hipStreamCreate failed";
    delete h;
    return HCS_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
  if (c->num_cu <= 0) c->num_cu = 256;
  if (scratch_setup(c) != HCS_OK) {
    g_create_err = c->err;
    (void)hipStreamDestroy(c->stream);
    delete h;
    return HCS_ERR_HIP;
  }
  *out = h;
  return HCS_OK;
}

int hcs_destroy(hcs_ctx* h) {
  if (!h) return HCS_ERR_ARG;
  Ctx* c = &h->c;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  while (!c->batches.empty()) hcs_batch_destroy(*c->batches.begin());
  for (auto& m : c->meshes) free_mesh(m);
  pipe_free(c->pipe);
  for (DBuf* b : {&c->d_envs, &c->d_state, &c->d_items, &c->d_face_rec, &c->d_face_key, &c->d_face_cnt, &c->d_poly_rec, &c->d_poly_key, &c->d_poly_cnt,
                  &c->d_cand, &c->d_cand_cnt, &c->d_tx_cnt, &c->d_tx_off, &c->d_tx_list, &c->d_tx_out,
                  &c->d_tx_grid})
    if (b->ptr) (void)hipFree(b->ptr);
  TaxelSensorBufs& T = c->ts;
  for (DBuf* b : {&T.mask, &T.k64, &T.kp, &T.idx, &T.cnt, &T.nfan, &T.fan_off, &T.tmp, &T.tris, &T.surf, &T.scnt, &T.soff,
                  &T.P, &T.E, &T.taxel, &T.val})
    if (b->ptr) (void)hipFree(b->ptr);
  (void)hipStreamDestroy(c->stream);
  delete h;
  return HCS_OK;
}

const char* hcs_last_error(const hcs_ctx* h) { return h ? h->c.err.c_str() : g_create_err.c_str(); }

int hcs_synchronize(hcs_ctx* h) {
  if (!h) return HCS_ERR_ARG;
  return hip_check(&h->c, hipStreamSynchronize(h->c.stream), "hipStreamSynchronize");
}

void* hcs_stream(hcs_ctx* h) { return h ? (void*)h->c.stream : nullptr; }

}  // extern "C"

namespace hcs {

/* The device data of a soft mesh (tet records, float face planes, boxes,
 * LBVH); m is left empty on failure. */
static int make_soft(Ctx* c, const double* V, int32_t nv, const int32_t* T, int32_t nt, const double* pr,
                     const hcs_props* props, int32_t geom_id, Mesh& m) {
  if (!V || !T || !pr || !props || nv < 4 || nt < 1) return HCS_ERR_ARG;
  if (!(props->hydroelastic_modulus > 0) || std::isinf(props->hydroelastic_modulus))
    return set_err(c, HCS_ERR_ARG, "soft mesh needs a finite hydroelastic_modulus > 0");
  for (int64_t i = 0; i < 4 * (int64_t)nt; ++i)
    if (T[i] < 0 || T[i] >= nv) return set_err(c, HCS_ERR_ARG, "tet vertex index out of range");
  HCHK(c, hipSetDevice(c->device));
  m = Mesh();
  m.kind = 1;
  m.geom_id = geom_id;
  m.props = *props;
  m.nv = nv;
  m.nelem = nt;
  int rc = upload(c, (void**)&m.V, V, 3 * sizeof(double) * nv);
  if (rc == HCS_OK) rc = upload(c, (void**)&m.E, T, 4 * sizeof(int32_t) * nt);
  if (rc == HCS_OK) rc = upload(c, (void**)&m.p, pr, sizeof(double) * nv);
  if (rc == HCS_OK) rc = hip_check(c, hipMalloc((void**)&m.rec, (size_t)HCS_TREC * sizeof(double) * nt), "hipMalloc");
  if (rc == HCS_OK) rc = hip_check(c, hipMalloc((void**)&m.box, 6 * sizeof(double) * (size_t)nt), "hipMalloc");
  if (rc == HCS_OK) rc = hip_check(c, hipMalloc((void**)&m.fpl, 16 * sizeof(float) * (size_t)nt), "hipMalloc");
  if (rc == HCS_OK) {
    hipLaunchKernelGGL(k_tet_precompute, dim3((nt + 127) / 128), dim3(128), 0, c->stream, m.V, m.E, m.p, nt, m.rec,
                       m.box, (float4*)m.fpl);
    rc = hip_check(c, hipGetLastError(), "k_tet_precompute");
  }
  if (rc == HCS_OK) rc = build_bvh(c, m);
  if (rc != HCS_OK) {
    (void)hipStreamSynchronize(c->stream);
    free_mesh(m);
  }
  return rc;
}

/* The device data of a rigid triangle mesh (triangles in a Morton query
 * order, unit normals); m is left empty on failure. */
static int make_rigid(Ctx* c, const double* V, int32_t nv, const int32_t* F, int32_t ntri, const hcs_props* props,
                      int32_t geom_id, Mesh& m) {
  if (!V || !F || !props || nv < 3 || ntri < 1) return HCS_ERR_ARG;
  if (!is_rigid(*props)) return set_err(c, HCS_ERR_ARG, "rigid mesh needs hydroelastic_modulus 0 or +inf");
  for (int64_t i = 0; i < 3 * (int64_t)ntri; ++i)
    if (F[i] < 0 || F[i] >= nv) return set_err(c, HCS_ERR_ARG, "triangle vertex index out of range");
  HCHK(c, hipSetDevice(c->device));
  m = Mesh();
  m.kind = 2;
  m.geom_id = geom_id;
  m.props = *props;
  m.nv = nv;
  m.nelem = ntri;
  int rc = upload(c, (void**)&m.V, V, 3 * sizeof(double) * nv);
  if (rc == HCS_OK) rc = upload(c, (void**)&m.E, F, 3 * sizeof(int32_t) * ntri);
  if (rc == HCS_OK) {
    /* query order: triangles sorted by the Morton code of their centroids,
     * so a work item's 64 triangles are spatially compact */
    std::vector<double> cen(3 * (size_t)ntri);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t t = 0; t < ntri; ++t)
      for (int k = 0; k < 3; ++k) {
        const double x = (V[3 * F[3 * t] + k] + V[3 * F[3 * t + 1] + k] + V[3 * F[3 * t + 2] + k]) / 3.0;
        cen[3 * (size_t)t + k] = x;
        lo[k] = std::min(lo[k], x), hi[k] = std::max(hi[k], x);
      }
    auto spread = [](uint32_t v) {
      v = (v * 0x00010001u) & 0xFF0000FFu;
      v = (v * 0x00000101u) & 0x0F00F00Fu;
      v = (v * 0x00000011u) & 0xC30C30C3u;
      v = (v * 0x00000005u) & 0x49249249u;
      return v;
    };
    std::vector<uint32_t> code(ntri);
    for (int32_t t = 0; t < ntri; ++t) {
      uint32_t q[3];
      for (int k = 0; k < 3; ++k) {
        const double u = hi[k] > lo[k] ? (cen[3 * (size_t)t + k] - lo[k]) / (hi[k] - lo[k]) * 1023.0 : 0.0;
        q[k] = (uint32_t)std::min(1023.0, std::max(0.0, u));
      }
      code[t] = (spread(q[0]) << 2) | (spread(q[1]) << 1) | spread(q[2]);
    }
    std::vector<int32_t> order(ntri);
    for (int32_t t = 0; t < ntri; ++t) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return code[a] < code[b]; });
    rc = upload(c, (void**)&m.qorder, order.data(), sizeof(int32_t) * (size_t)ntri);
  }
  if (rc == HCS_OK) rc = hip_check(c, hipMalloc((void**)&m.nrm, 3 * sizeof(double) * (size_t)ntri), "hipMalloc");
  if (rc == HCS_OK) {
    hipLaunchKernelGGL(k_tri_normals, dim3((ntri + 127) / 128), dim3(128), 0, c->stream, m.V, m.E, ntri, m.nrm);
    rc = hip_check(c, hipGetLastError(), "k_tri_normals");
  }
  if (rc == HCS_OK) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (rc != HCS_OK) free_mesh(m);
  return rc;
}

/* Installs m as the geometry of `handle` (onGeomChanged): the old device
 * data is freed once nothing in flight uses it, and every batch that steps
 * the handle is planned and captured again. */
static int install_mesh(Ctx* c, int32_t handle, Mesh& m) {
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (rc) {
    free_mesh(m);
    return rc;
  }
  free_mesh(c->meshes[handle]);
  c->meshes[handle] = m;
  c->last_valid = 0;
  return batches_refresh(c, handle);
}

}  // namespace hcs

extern "C" {

int hcs_add_soft_mesh(hcs_ctx* h, const double* V, int32_t nv, const int32_t* T, int32_t nt, const double* pr,
                      const hcs_props* props, int32_t geom_id, int32_t* handle) {
  if (!h || !handle) return HCS_ERR_ARG;
  Mesh m;
  const int rc = make_soft(&h->c, V, nv, T, nt, pr, props, geom_id, m);
  if (rc) return rc;
  *handle = (int32_t)h->c.meshes.size();
  h->c.meshes.push_back(m);
  return HCS_OK;
}

int hcs_add_rigid_mesh(hcs_ctx* h, const double* V, int32_t nv, const int32_t* F, int32_t ntri,
                       const hcs_props* props, int32_t geom_id, int32_t* handle) {
  if (!h || !handle) return HCS_ERR_ARG;
  Mesh m;
  const int rc = make_rigid(&h->c, V, nv, F, ntri, props, geom_id, m);
  if (rc) return rc;
  *handle = (int32_t)h->c.meshes.size();
  h->c.meshes.push_back(m);
  return HCS_OK;
}

int hcs_replace_soft_mesh(hcs_ctx* h, int32_t handle, const double* V, int32_t nv, const int32_t* T, int32_t nt,
                          const double* pressure) {
  if (!h || handle < 0 || handle >= (int)h->c.meshes.size()) return HCS_ERR_ARG;
  Ctx* c = &h->c;
  if (c->meshes[handle].kind != 1) return set_err(c, HCS_ERR_ARG, "handle is not a soft mesh");
  Mesh m;
  const hcs_props props = c->meshes[handle].props;
  int rc = make_soft(c, V, nv, T, nt, pressure, &props, c->meshes[handle].geom_id, m);
  return rc ? rc : install_mesh(c, handle, m);
}

int hcs_replace_rigid_mesh(hcs_ctx* h, int32_t handle, const double* V, int32_t nv, const int32_t* F,
                           int32_t ntri) {
  if (!h || handle < 0 || handle >= (int)h->c.meshes.size()) return HCS_ERR_ARG;
  Ctx* c = &h->c;
  if (c->meshes[handle].kind != 2) return set_err(c, HCS_ERR_ARG, "handle is not a rigid triangle mesh");
  Mesh m;
  const hcs_props props = c->meshes[handle].props;
  int rc = make_rigid(c, V, nv, F, ntri, &props, c->meshes[handle].geom_id, m);
  return rc ? rc : install_mesh(c, handle, m);
}

int hcs_add_rigid_plane(hcs_ctx* h, const hcs_props* props, int32_t geom_id, int32_t* handle) {
  if (!h || !props || !handle) return HCS_ERR_ARG;
  Ctx* c = &h->c;
  if (!is_rigid(*props)) return set_err(c, HCS_ERR_ARG, "rigid plane needs hydroelastic_modulus 0 or +inf");
  Mesh m;
  m.kind = 3;
  m.geom_id = geom_id;
  m.props = *props;
  m.nv = 0;
  m.nelem = 1; /* one element: the plane (candidate keys tet * 1 + 0) */
  *handle = (int32_t)c->meshes.size();
  c->meshes.push_back(m);
  return HCS_OK;
}

int hcs_mesh_info(hcs_ctx* h, int32_t handle, int64_t* info) {
  if (!h || !info || handle < 0 || handle >= (int)h->c.meshes.size()) return HCS_ERR_ARG;
  const Mesh& m = h->c.meshes[handle];
  info[0] = m.kind, info[1] = m.nv, info[2] = m.nelem, info[3] = m.height;
  return HCS_OK;
}

int hcs_mesh_tet_records(hcs_ctx* h, int32_t handle, double* out) {
  if (!h || !out || handle < 0 || handle >= (int)h->c.meshes.size() || h->c.meshes[handle].kind != 1)
    return HCS_ERR_ARG;
  Ctx* c = &h->c;
  const Mesh& m = c->meshes[handle];
  HCHK(c, hipMemcpyAsync(out, m.rec, (size_t)HCS_TREC * sizeof(double) * m.nelem, hipMemcpyDeviceToHost, c->stream));
  HCHK(c, hipStreamSynchronize(c->stream));
  return HCS_OK;
}

}  // extern "C"

namespace hcs {

/* Uploads plan and state of npairs pairs into the context scratch. */
int stage_pairs(Ctx* c, int npairs, const hcs_pair_in* in, const hcs_step_opts* opts, Plan& plan, DBuf* dst) {
  HCHK(c, hipSetDevice(c->device));
  std::vector<int32_t> geoms(2 * (size_t)npairs);
  std::vector<double> state((size_t)HCS_STATE * npairs);
  for (int e = 0; e < npairs; ++e) {
    geoms[2 * e] = in[e].geom_a, geoms[2 * e + 1] = in[e].geom_b;
    double* st = &state[(size_t)HCS_STATE * e];
    std::memcpy(st, in[e].X_WA, 12 * sizeof(double));
    std::memcpy(st + 12, in[e].X_WB, 12 * sizeof(double));
    std::memcpy(st + 24, in[e].V_WA, 6 * sizeof(double));
    std::memcpy(st + 30, in[e].V_WB, 6 * sizeof(double));
  }
  int rc = build_plan(c, npairs, geoms.data(), opts, c->tet_mode_max_tris, plan);
  if (rc) return rc;
  const int nitems = (int)plan.items.size();
  DBuf& envs = dst ? dst[0] : c->d_envs;
  DBuf& st = dst ? dst[1] : c->d_state;
  DBuf& items = dst ? dst[2] : c->d_items;
  if ((rc = ensure(c, envs, sizeof(EnvDesc) * npairs)) || (rc = ensure(c, st, state.size() * 8)) ||
      (rc = ensure(c, items, sizeof(Item) * std::max(1, nitems))))
    return rc;
  HCHK(c, hipMemcpyAsync(envs.ptr, plan.envs.data(), sizeof(EnvDesc) * npairs, hipMemcpyHostToDevice, c->stream));
  HCHK(c, hipMemcpyAsync(st.ptr, state.data(), state.size() * 8, hipMemcpyHostToDevice, c->stream));
  if (nitems)
    HCHK(c, hipMemcpyAsync(items.ptr, plan.items.data(), sizeof(Item) * nitems, hipMemcpyHostToDevice, c->stream));
  c->last_geoms = geoms;
  return HCS_OK;
}

int clip_grid_for(const Ctx* c, int64_t cap) {
  const int64_t blocks = (cap + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, (int64_t)c->num_cu * 16));
}

/* Shared body of hcs_contact_step / hcs_candidates: one pass of the
 * pipeline over npairs pairs (keys: candidate-key mode). */
static int run_pairs(Ctx* c, int npairs, const hcs_pair_in* in, const hcs_step_opts* opts, hcs_pair_out* out,
                     hcs_faces* faces, int64_t* keys, int64_t key_cap, int64_t* key_count) {
  Plan plan;
  int rc = stage_pairs(c, npairs, in, opts, plan);
  if (rc) return rc;
  const int nitems = (int)plan.items.size();
  Pipe& P = c->pipe;
  if ((rc = pipe_reserve(c, P, nitems, npairs, std::max<int64_t>(P.cap, 4096), 0))) return rc;
  if ((rc = pipe_plan(c, P, (const EnvDesc*)c->d_envs.ptr, (const Item*)c->d_items.ptr, plan))) return rc;
  if ((rc = ensure(c, c->d_face_cnt, 64)) || (rc = ensure(c, c->d_poly_cnt, 64)) || (rc = ensure(c, c->d_cand_cnt, 64)))
    return rc;
  const bool keep = opts && opts->keep_polygons && !keys;
  const bool want_faces = faces && faces->capacity > 0 && faces->rec && faces->key && !keys;
  if (want_faces) {
    if ((rc = ensure(c, c->d_face_rec, sizeof(double) * HCS_FREC * faces->capacity)) ||
        (rc = ensure(c, c->d_face_key, sizeof(int64_t) * 4 * faces->capacity)))
      return rc;
  }
  if (keys && (rc = ensure(c, c->d_cand, sizeof(int64_t) * std::max<int64_t>(key_cap, 1)))) return rc;
  int64_t poly_cap = 0;
  if (keep) {
    poly_cap = std::max<int64_t>(c->d_poly_rec.bytes / (sizeof(double) * HCS_PREC), 4096);
    if ((rc = ensure(c, c->d_poly_rec, sizeof(double) * HCS_PREC * poly_cap)) ||
        (rc = ensure(c, c->d_poly_key, sizeof(int64_t) * 3 * poly_cap)))
      return rc;
  }
  std::vector<PairOut> res(npairs);
  unsigned long long nface = 0, ncand = 0, npoly = 0, nsurv = 0;
  for (int attempt = 0; attempt < 3; ++attempt) {
    HCHK(c, hipMemsetAsync(c->d_face_cnt.ptr, 0, 64, c->stream));
    HCHK(c, hipMemsetAsync(c->d_poly_cnt.ptr, 0, 64, c->stream));
    HCHK(c, hipMemsetAsync(c->d_cand_cnt.ptr, 0, 64, c->stream));
    QueryArgs a = pipe_args(P, (const EnvDesc*)c->d_envs.ptr, (const double*)c->d_state.ptr,
                            (const Item*)c->d_items.ptr, nitems, false);
    if (want_faces) {
      a.face_rec = (double*)c->d_face_rec.ptr;
      a.face_key = (int64_t*)c->d_face_key.ptr;
      a.face_cnt = (unsigned long long*)c->d_face_cnt.ptr;
      a.face_cap = faces->capacity;
    }
    if (keep) {
      a.poly_rec = (double*)c->d_poly_rec.ptr;
      a.poly_key = (int64_t*)c->d_poly_key.ptr;
      a.poly_cnt = (unsigned long long*)c->d_poly_cnt.ptr;
      a.poly_cap = poly_cap;
    }
    if (keys) {
      a.cand_key = (int64_t*)c->d_cand.ptr;
      a.cand_cnt = (unsigned long long*)c->d_cand_cnt.ptr;
      a.cand_cap = key_cap;
    }
    pipe_enqueue(c, P, a, npairs, nitems, clip_grid_for(c, P.cap), 0, nullptr, nullptr, keys != nullptr, 0);
    HCHK(c, hipGetLastError());
    HCHK(c, hipMemcpyAsync(&nsurv, P.surv_cnt.ptr, 8, hipMemcpyDeviceToHost, c->stream));
    HCHK(c, hipMemcpyAsync(&npoly, c->d_poly_cnt.ptr, 8, hipMemcpyDeviceToHost, c->stream));
    HCHK(c, hipMemcpyAsync(&nface, c->d_face_cnt.ptr, 8, hipMemcpyDeviceToHost, c->stream));
    HCHK(c, hipMemcpyAsync(&ncand, c->d_cand_cnt.ptr, 8, hipMemcpyDeviceToHost, c->stream));
    if (!keys)
      HCHK(c, hipMemcpyAsync(res.data(), P.out.ptr, sizeof(PairOut) * npairs, hipMemcpyDeviceToHost, c->stream));
    HCHK(c, hipStreamSynchronize(c->stream));
    bool again = false;
    if (!keys && (int64_t)nsurv > P.cap) {
      if ((rc = pipe_reserve(c, P, nitems, npairs, (int64_t)nsurv + (int64_t)nsurv / 4 + 1024, 0))) return rc;
      again = true;
    }
    if (keep && (int64_t)npoly > poly_cap) {
      poly_cap = (int64_t)npoly + (int64_t)npoly / 4 + 1024;
      if ((rc = ensure(c, c->d_poly_rec, sizeof(double) * HCS_PREC * poly_cap)) ||
          (rc = ensure(c, c->d_poly_key, sizeof(int64_t) * 3 * poly_cap)))
        return rc;
      again = true;
    }
    if (!again) break;
  }
  if (keep) c->last_npoly = (int64_t)npoly;
  if (want_faces) {
    const int64_t nw = std::min<int64_t>((int64_t)nface, faces->capacity);
    if (nw > 0) {
      HCHK(c, hipMemcpyAsync(faces->rec, c->d_face_rec.ptr, sizeof(double) * HCS_FREC * nw, hipMemcpyDeviceToHost,
                             c->stream));
      HCHK(c, hipMemcpyAsync(faces->key, c->d_face_key.ptr, sizeof(int64_t) * 4 * nw, hipMemcpyDeviceToHost,
                             c->stream));
      HCHK(c, hipStreamSynchronize(c->stream));
    }
    faces->count = (int64_t)nface;
  } else if (faces) {
    faces->count = 0;
  }
  int err = 0;
  HCHK(c, hipMemcpyAsync(&err, P.err.ptr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HCHK(c, hipStreamSynchronize(c->stream));
  if (keys) {
    const int64_t nw = std::min<int64_t>((int64_t)ncand, key_cap);
    if (nw > 0) {
      HCHK(c, hipMemcpyAsync(keys, c->d_cand.ptr, sizeof(int64_t) * nw, hipMemcpyDeviceToHost, c->stream));
      HCHK(c, hipStreamSynchronize(c->stream));
      std::sort(keys, keys + nw);
    }
    if (key_count) *key_count = (int64_t)ncand;
    if (err) return set_err(c, HCS_ERR_LIMIT, "traversal stack overflow");
    if ((int64_t)ncand > key_cap)
      return set_err(c, HCS_ERR_CAPACITY, "key buffer too small: " + std::to_string(ncand) + " keys");
    return HCS_OK;
  }
  if (out) std::memcpy(out, res.data(), sizeof(PairOut) * npairs);
  if (keep) {
    c->last_valid = 1;
    c->last_pairs.assign(in, in + npairs);
  }
  if (err) return set_err(c, HCS_ERR_LIMIT, "traversal stack overflow");
  if (want_faces && (int64_t)nface > faces->capacity)
    return set_err(c, HCS_ERR_CAPACITY, "face buffer too small: " + std::to_string(nface) + " faces");
  return HCS_OK;
}

}  // namespace hcs

extern "C" {

int hcs_contact_step(hcs_ctx* h, int32_t npairs, const hcs_pair_in* in, const hcs_step_opts* opts,
                     hcs_pair_out* out, hcs_faces* faces) {
  if (!h || npairs < 0 || (npairs > 0 && (!in || !out))) return HCS_ERR_ARG;
  if (npairs == 0) return HCS_OK;
  static_assert(sizeof(hcs_pair_out) == sizeof(PairOut), "pair_out layout");
  return run_pairs(&h->c, npairs, in, opts, out, faces, nullptr, 0, nullptr);
}

int hcs_candidates(hcs_ctx* h, const hcs_pair_in* pair, int64_t* keys, int64_t capacity, int64_t* count) {
  if (!h || !pair || !count || capacity < 0 || (capacity > 0 && !keys)) return HCS_ERR_ARG;
  int64_t dummy = 0;
  /* keys == NULL: count-only query */
  const int rc = run_pairs(&h->c, 1, pair, nullptr, nullptr, nullptr, keys ? keys : &dummy, keys ? capacity : 0, count);
  return (!keys && rc == HCS_ERR_CAPACITY) ? HCS_OK : rc;
}

}  // extern "C"
