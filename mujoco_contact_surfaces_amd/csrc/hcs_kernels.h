/*
 * hcs_kernels.h -- device data layout shared by the engine's kernels and its
 * host orchestration (hcs_engine.hip, hcs_taxels.hip, hcs_adjoint.hip).
 *
 * HBM layout (all FP64 unless noted; one set per mesh, allocated at load):
 *   soft mesh:  rec   nt x 24  tet records (4 unit outward face normals,
 *                               4 plane offsets, pressure gradient g,
 *                               constant c with p(x) = c + g.x, unit g, pad)
 *               box   nt x 6   AABB lo, hi (soft frame)
 *               nodes (nt-1) x Node   binary LBVH, child boxes stored in the
 *                               parent so one 128-byte fetch (one cache
 *                               line) tests both
 *               leaf  nt int32 tets in Morton order
 *               T     nt x 4 int32, V nv x 3, p nv
 *   rigid mesh: V nv x 3, F ntri x 3 int32, nrm ntri x 3 (unit normals)
 * Per plan:    items (16 B), static item boxes (12 doubles: the query
 *              elements' box in their own frame, then the target's box)
 * Per step:    state  npairs x 36 (X_WA, X_WB, V_WA, V_WB), xf npairs x 12
 *              (X_AB = X_WA^-1 X_WB), live item list, survivors (16 B),
 *              per-item counts and scan, contributions (80 B), per-pair
 *              outputs.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hcs {

#define HCS_TREC 24
/* traversal stack entries per lane: a 4-wide walk of a binary LBVH of
 * height h keeps at most 3 ceil(h / 2) + 1 */
#define HCS_STACK 96
#define HCS_FREC 12
#define HCS_PREC 40
#define HCS_STATE 36
/* tangent directions of the differentiation pass: 6 of the pose of B, E_A */
#define HCS_MAXDIR 7
/* a rigid triangle moved into A's frame: 3 vertices, unit normal, AABB */
#define HCS_XTRI 18
/* at most this many work queues for the candidate kernel: with one per XCD
 * the waves of one XCD (workgroups are dispatched to the 8 XCDs round-robin
 * by block index) work on the same pairs and share that XCD's L2 */
#define HCS_QUEUES 8

/* 128 bytes, so a node on a 128-byte boundary is one cache line. */
struct Node {
  double b[12];   /* child 0 lo[3] hi[3], child 1 lo[3] hi[3] */
  int32_t c[2];   /* >= 0 internal node, < 0 leaf: ~element */
  int32_t h[2];   /* subtree heights (leaf 0) */
  int32_t pad[4];
};
static_assert(sizeof(Node) == 128, "one LBVH node per cache line");

/* The 4-wide node of binary internal node i: its grandchildren (a leaf
 * child stands for itself), so a walk over wide nodes visits every other
 * level of the binary tree.  b[6k..6k+5]: child k's box (lo, hi, the
 * binary node's double box); ch[k]: >= 0 the wide node of binary node
 * ch[k], < 0 the leaf ~ch[k], HCS_WNONE an empty slot (whose box is
 * empty). */
struct WNode {
  double b[24];
  int32_t ch[4];
  int32_t pad[12];
};
static_assert(sizeof(WNode) == 256, "one wide node per two cache lines");
#define HCS_WNONE INT32_MIN

/* Builds the wide node of binary internal node i. */
__host__ __device__ inline void make_wide_node(const Node* nodes, int i, WNode& w) {
  const Node& n = nodes[i];
  int m = 0;
  auto put = [&](const double* b, int ch) {
    for (int c = 0; c < 6; ++c) w.b[6 * m + c] = b[c];
    w.ch[m++] = ch;
  };
  for (int s = 0; s < 2; ++s) {
    const int c = n.c[s];
    if (c < 0) {
      put(n.b + 6 * s, c);
    } else {
      const Node& g = nodes[c];
      for (int t = 0; t < 2; ++t) put(g.b + 6 * t, g.c[t]);
    }
  }
  for (; m < 4; ++m) {
    for (int c = 0; c < 3; ++c) w.b[6 * m + c] = INFINITY, w.b[6 * m + 3 + c] = -INFINITY;
    w.ch[m] = HCS_WNONE;
  }
  for (int k = 0; k < 12; ++k) w.pad[k] = 0;
}

/* query kinds: rigid triangles against A's LBVH, A's tets against a small
 * rigid mesh, tets of a soft B against A's LBVH, A's tets against the rigid
 * half-space z <= 0 of B's frame (a plane geom) */
enum EnvKind { KIND_TRI = 0, KIND_TET = 1, KIND_SOFT = 2, KIND_PLANE = 3 };

/* candidate kernel modes: a contact step (exact rejections, survivors), the
 * canonical candidate keys, or per-item candidate counts only (planning) */
enum CandMode { CAND_STEP = 0, CAND_KEYS = 1, CAND_COUNT = 2 };

/* Device code sees the mesh pointers of an env descriptor in the global
 * address space: the descriptor is read from memory, so the compiler could
 * not otherwise tell them from LDS or private pointers and would load
 * through them with flat instructions, which count against the LDS wait
 * counter and so make every LDS wait also wait for those loads.  Host code
 * (and the CPU build of the device code) sees plain pointers; hcs_gptr()
 * converts a plain pointer for assignment. */
#if defined(__HIP_DEVICE_COMPILE__)
#define HCS_GPTR __attribute__((address_space(1)))
#else
#define HCS_GPTR
#endif
template <class T>
__host__ __device__ inline const HCS_GPTR T* hcs_gptr(const T* p) {
  return (const HCS_GPTR T*)p;
}
/* The same pointer as a plain one (for templates that deduce the element
 * type); loads through it stay global after inlining. */
template <class T>
__host__ __device__ inline const T* hcs_plain(const HCS_GPTR T* p) {
  return (const T*)p;
}

struct EnvDesc {
  const HCS_GPTR double* recA;
  const HCS_GPTR float4* fplA; /* A's face planes rounded to float (n, d), 4 per tet */
  const HCS_GPTR double* boxA;
  const HCS_GPTR Node* nodesA;
  const HCS_GPTR WNode* wnodesA; /* the 4-wide nodes of A's LBVH, one per internal node */
  const HCS_GPTR int32_t* leafA; /* A's tets in Morton order (tet-query items) */
  const HCS_GPTR double* boxLA;  /* their AABBs in that order */
  const HCS_GPTR int32_t* TA;
  const HCS_GPTR double* VA;
  const HCS_GPTR double* VB;
  const HCS_GPTR int32_t* FB;  /* rigid B triangles, or soft B tets */
  const HCS_GPTR double* nrmB; /* rigid B normals */
  const HCS_GPTR double* recB; /* soft B records */
  const HCS_GPTR int32_t* qorder; /* triangle / soft queries: B's elements in a
                                     spatially compact order (nullptr: identity) */
  int32_t ntA, nB, kind, nvB;
  int32_t a_is_M, repr, apply, b_soft;
  double mu, dstar, dt;
  int32_t item_first, item_count;
  int64_t xtri_off; /* tet-query mode: first of this pair's triangles in the
                       per-step moved-triangle buffer (HCS_XTRI doubles each) */
};

/* A work item: queries first .. first+count-1 of pair env.  Triangle and
 * soft queries walk A's LBVH from node root (0: the whole tree; < 0: the
 * single leaf ~root), so a heavy item can be split into sub-items over
 * disjoint subtrees. */
struct Item {
  int32_t env, first, count, root;
};

struct PairOut {
  double F_A[3], tau_A[3], F_B[3], tau_B[3];
  int64_t n_candidates, n_polygons, n_faces, status;
};

struct QueryArgs {
  const EnvDesc* envs;
  const double* state;
  const double* xf;
  const Item* items;
  int32_t nitems;
  int32_t queries_per_lane;
  unsigned int* counter; /* per-queue work counters (persistent kernel); nullptr: one item per block */
  double* face_rec;
  int64_t* face_key;
  unsigned long long* face_cnt;
  int64_t face_cap;
  double* poly_rec;
  int64_t* poly_key;
  unsigned long long* poly_cnt;
  int64_t poly_cap;
  double* poly_tan; /* tangent copy of poly_rec (tangent pass only) */
  const unsigned char* poly_env; /* pairs whose polygons are kept (nullptr: every pair) */
  int64_t* cand_key; /* keys-only mode when non-null */
  unsigned long long* cand_cnt;
  int64_t cand_cap;
  int* error;
  /* optional per-item profile: cycles, production steps, candidates,
   * survivors of the cheap tests, wall-clock start and end (6 x uint64 per
   * item) */
  unsigned long long* item_stats;
  /* survivors of the cheap tests: (elem A, elem B, item, seq in item) */
  int4* surv;
  unsigned long long* surv_cnt;
  int64_t surv_cap;
  int* item_ncand;
  unsigned* item_nsurv;
  /* items whose static boxes overlap this step, in nqueues queues (the
   * items of pair e go to queue e % nqueues; queue x holds live_cnt[x]
   * items from live + qseg[x]); nullptr: all items, one queue */
  const int* live;
  const unsigned* live_cnt;
  const int* qseg;
  int32_t nqueues;
  const int4* live_item; /* the live items' Item records, same order as live */
  const double* xtri;    /* per-step moved triangles of tet-query pairs */
  int32_t kinds;         /* bit k set: some pair is of EnvKind k; a launch
                            whose pairs share one kind runs a kernel built
                            for it (0: not known, the general kernel) */
};

/* One survivor's wrench contribution (clip kernel output, 80 B). */
struct Contrib {
  double F[3], tA[3], tB[3];
  int np, nf;
};

/* One reduction wave's sums over its slice of a pair's contributions. */
struct RedPart {
  double F[3], tA[3], tB[3];
  int64_t np, nf, nc;
};

/* Inputs of the clip kernel beyond QueryArgs. */
struct ClipArgs {
  const unsigned* item_off; /* exclusive scan of item_nsurv, nitems + 1 */
  const double* dstar_dE;   /* tangent seeds (tangent pass only) */
  const double* invE;
  int64_t cap; /* contribution slots per direction; a survivor whose slot
                  is beyond it (survivor list overflow) writes nothing */
};

/* A flat sensor's taxel grid on the device: pose (rotation row-major,
 * position), half sizes, taxel pitch, window sigma, S rays per taxel side,
 * window kind, taxels along x and y. */
struct TaxelGridDev {
  double X[12];
  double xs, ys, zs, res, sigma;
  int S, window, cx, cy;
};

}  // namespace hcs
