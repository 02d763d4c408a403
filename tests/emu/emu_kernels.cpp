// CPU build of the engine's candidate and clip device code (hcs_query.h),
// test infrastructure only: one pair, work items of 64 queries (optionally
// split over the LBVH frontier), each item run by one emulated 64-lane wave
// (run_cand), then every survivor clipped in double and, per tangent
// direction, in the Dual scalar type (clip_survivor), and contributions
// summed in slot order as the per-pair reduction does.  The LBVH is built
// sequentially with the engine's algorithm (Morton codes, stable sort,
// Karras tree, bottom-up boxes); tet records and combined material
// parameters come from the oracle, whose arithmetic matches the engine's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <thread>
#include <vector>

#include "hcs_oracle.h"
#include "hcs_query.h"

thread_local Dim3 threadIdx;
thread_local int g_lane;
WaveEmu* g_wave;
/* emu_set_options: run the general candidate kernel (the one mixed-kind
 * launches use) instead of the kind's own */
int g_generic = 0;

using namespace hcs;

namespace {

unsigned expand_bits(unsigned v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

int delta(const std::vector<unsigned>& c, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const unsigned a = c[i], b = c[j];
  if (a == b) return 32 + __builtin_clz((unsigned)(i ^ j));
  return __builtin_clz(a ^ b);
}

struct Soft {
  std::vector<double> rec, box, boxL;
  std::vector<float4> fpl;
  std::vector<Node> nodes;
  std::vector<WNode> wnodes;
  std::vector<int> leaf;
};

/* tet records (24 doubles, the device layout), boxes and the LBVH */
void prepare(const double* V, int nv, const int32_t* T, int nt, const double* p, Soft& s) {
  std::vector<double> r32(32 * (size_t)nt);
  orc_tet_precompute(V, nv, T, nt, p, r32.data());
  s.rec.assign(24 * (size_t)nt, 0.0);
  s.box.resize(6 * (size_t)nt);
  s.fpl.resize(4 * (size_t)nt);
  for (int t = 0; t < nt; ++t) {
    for (int k = 0; k < 23; ++k) s.rec[24 * (size_t)t + k] = r32[32 * (size_t)t + k];
    const double* r = &s.rec[24 * (size_t)t];
    for (int k = 0; k < 4; ++k)
      s.fpl[4 * (size_t)t + k] = make_float4((float)r[3 * k], (float)r[3 * k + 1], (float)r[3 * k + 2], (float)r[12 + k]);
    for (int k = 0; k < 6; ++k) s.box[6 * (size_t)t + k] = r32[32 * (size_t)t + 23 + k];
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int v = 0; v < nv; ++v)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], V[3 * v + k]), hi[k] = std::max(hi[k], V[3 * v + k]);
  std::vector<unsigned> code(nt);
  s.leaf.resize(nt);
  for (int t = 0; t < nt; ++t) {
    unsigned q[3];
    for (int c = 0; c < 3; ++c) {
      const double sc = hi[c] > lo[c] ? 1024.0 / (hi[c] - lo[c]) : 0.0;
      const double u = ((s.box[6 * t + c] + s.box[6 * t + 3 + c]) * 0.5 - lo[c]) * sc;
      q[c] = (unsigned)std::fmin(std::fmax(u, 0.0), 1023.0);
    }
    code[t] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    s.leaf[t] = t;
  }
  std::stable_sort(s.leaf.begin(), s.leaf.end(), [&](int a, int b) { return code[a] < code[b]; });
  s.boxL.resize(6 * (size_t)nt);
  for (int i = 0; i < nt; ++i)
    for (int c = 0; c < 6; ++c) s.boxL[6 * (size_t)i + c] = s.box[6 * (size_t)s.leaf[i] + c];
  const int n = nt;
  s.nodes.assign(std::max(1, n - 1), Node{});
  if (n == 1) return;
  std::vector<unsigned> sc(n);
  for (int i = 0; i < n; ++i) sc[i] = code[s.leaf[i]];
  std::vector<int> np(n, -1), lp(n, -1);
  for (int i = 0; i < n - 1; ++i) {
    const int d = (delta(sc, n, i, i + 1) - delta(sc, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(sc, n, i, i - d);
    int lmax = 2;
    while (delta(sc, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (delta(sc, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d, dn = delta(sc, n, i, j);
    int sp = 0, t = l;
    do {
      t = (t + 1) >> 1;
      if (delta(sc, n, i, i + (sp + t) * d) > dn) sp += t;
    } while (t > 1);
    const int g = i + sp * d + (d < 0 ? -1 : 0), lo_ = std::min(i, j), hi_ = std::max(i, j);
    if (lo_ == g) s.nodes[i].c[0] = ~s.leaf[g], lp[g] = 2 * i;
    else s.nodes[i].c[0] = g, np[g] = 2 * i;
    if (hi_ == g + 1) s.nodes[i].c[1] = ~s.leaf[g + 1], lp[g + 1] = 2 * i + 1;
    else s.nodes[i].c[1] = g + 1, np[g + 1] = 2 * i + 1;
  }
  std::vector<int> flags(n, 0);
  for (int j = 0; j < n; ++j) {
    const double* b = &s.box[6 * (size_t)s.leaf[j]];
    int pc = lp[j], pp = pc >> 1, side = pc & 1;
    for (int c = 0; c < 6; ++c) s.nodes[pp].b[6 * side + c] = b[c];
    s.nodes[pp].h[side] = 0;
    for (int guard = 0; guard < n; ++guard) {
      if (flags[pp]++ == 0) break;
      Node& N = s.nodes[pp];
      double u[6];
      for (int c = 0; c < 3; ++c) u[c] = std::fmin(N.b[c], N.b[6 + c]), u[3 + c] = std::fmax(N.b[3 + c], N.b[9 + c]);
      const int h = 1 + std::max(N.h[0], N.h[1]);
      if (pp == 0) break;
      pc = np[pp], pp = pc >> 1, side = pc & 1;
      for (int c = 0; c < 6; ++c) s.nodes[pp].b[6 * side + c] = u[c];
      s.nodes[pp].h[side] = h;
    }
  }
  s.wnodes.resize(std::max(1, n - 1));
  for (int i = 0; i + 1 < n; ++i) make_wide_node(s.nodes.data(), i, s.wnodes[i]);
}

}  // namespace

extern "C" {

/* One pair through the candidate kernel body and the clip body.
 *   kind: 0 rigid-triangle queries, 1 tet queries (small rigid mesh),
 *         2 soft-soft (B = tets VB/FB with pressures pB), 3 rigid half-space
 *   params: E_A, d_A, E_B, d_B, mu_A, mu_B, dt, apply, representation, a_is_M
 *           (E_B = +inf, d_B = 1 for a rigid B)
 *   split: LBVH frontier depth of the items (0: whole tree; kinds 0 and 2)
 *   ndir, seeds: tangent directions 0..ndir-1 (0-2 rotation of B about the
 *           world axes, 3-5 translation of B, 6 modulus of A), seeds =
 *           [d(d*)/dE_A, 1/E_A]
 *   wrench: (1 + ndir) x 12 = F_A, tau_A, F_B, tau_B (value, then tangents)
 *   counts: candidates, polygons, faces
 *   stats: items, items whose static boxes reject them under this pose,
 *          such items that nevertheless have candidates (must be 0),
 *          survivors of the exact rejections, traversal error flags */
int emu_pair(const double* VA, int nva, const int32_t* TA, int nta, const double* pA, const double* VB, int nvb,
             const int32_t* FB, int nb, const double* pB, int kind, const double* state, const double* params,
             int split, int ndir, const double* seeds, double* wrench, int64_t* counts, int64_t* stats) {
  Soft A;
  prepare(VA, nva, TA, nta, pA, A);
  Soft Bs;
  std::vector<double> nrm;
  if (kind == KIND_SOFT) {
    prepare(VB, nvb, FB, nb, pB, Bs);
  } else if (kind != KIND_PLANE) {
    nrm.resize(3 * (size_t)nb);
    orc_tri_normals(VB, nvb, FB, nb, nrm.data());
  }
  EnvDesc E{};
  E.recA = A.rec.data(), E.fplA = A.fpl.data(), E.boxA = A.box.data(), E.nodesA = A.nodes.data(), E.wnodesA = A.wnodes.data(), E.leafA = A.leaf.data();
  E.boxLA = A.boxL.data(), E.TA = TA, E.VA = VA, E.ntA = nta;
  E.VB = VB, E.FB = FB, E.nrmB = nrm.empty() ? nullptr : nrm.data();
  E.recB = kind == KIND_SOFT ? Bs.rec.data() : nullptr;
  E.nB = kind == KIND_PLANE ? 1 : nb, E.nvB = nvb, E.kind = kind;
  E.b_soft = kind == KIND_SOFT, E.a_is_M = params[9] != 0, E.repr = (int)params[8], E.apply = params[7] != 0;
  E.dt = params[6];
  E.dstar = orc_combined_dissipation(params[0], params[1], params[2], params[3]);
  E.mu = orc_combined_friction(params[4], params[5]);
  E.xtri_off = -1;
  /* query order of B's elements: reversed, so slot -> element mapping is
   * exercised (the engine uses a Morton order) */
  std::vector<int32_t> qorder(std::max(1, nb));
  for (int i = 0; i < nb; ++i) qorder[i] = nb - 1 - i;
  E.qorder = (kind == KIND_TRI || kind == KIND_SOFT) ? qorder.data() : nullptr;
  /* items and their LBVH roots */
  const bool chunks_of_A = kind == KIND_TET || kind == KIND_PLANE;
  const int nq = chunks_of_A ? nta : nb;
  std::vector<int> roots{0};
  if (!chunks_of_A && nta > 1)
    for (int d = 0; d < std::max(split, 0); ++d) {
      std::vector<int> nx;
      for (int r : roots) {
        if (r < 0) nx.push_back(r);
        else nx.push_back(A.nodes[r].c[0]), nx.push_back(A.nodes[r].c[1]);
      }
      roots.swap(nx);
    }
  std::vector<Item> items;
  const int per = 64;
  for (int q = 0; q < nq; q += per)
    for (int r : roots) items.push_back(Item{0, q, std::min(per, nq - q), chunks_of_A ? 0 : r});
  E.item_first = 0, E.item_count = (int)items.size();
  const int nitems = (int)items.size();
  double X[12];
  orc_compose_SR(state, state + 12, X);
  /* the cull decision of each item (k_item_boxes + k_cull) */
  std::vector<char> culled(nitems, 0);
  for (int i = 0; i < nitems; ++i) {
    const Item& it = items[i];
    const double inf = std::numeric_limits<double>::infinity();
    double q[6] = {inf, inf, inf, -inf, -inf, -inf}, t[6] = {inf, inf, inf, -inf, -inf, -inf};
    for (int l = 0; l < it.count; ++l) {
      if (chunks_of_A) {
        const double* b = &A.boxL[6 * (size_t)(it.first + l)];
        for (int c = 0; c < 3; ++c) q[c] = std::fmin(q[c], b[c]), q[3 + c] = std::fmax(q[3 + c], b[3 + c]);
      } else {
        const int nvq = kind == KIND_SOFT ? 4 : 3;
        const int32_t* f = FB + nvq * (size_t)qorder[it.first + l];
        for (int k = 0; k < nvq; ++k)
          for (int c = 0; c < 3; ++c)
            q[c] = std::fmin(q[c], VB[3 * f[k] + c]), q[3 + c] = std::fmax(q[3 + c], VB[3 * f[k] + c]);
      }
    }
    if (kind == KIND_PLANE) {
      culled[i] = !box_may_meet_plane(X, q);
      continue;
    }
    if (kind == KIND_TET) {
      for (int v = 0; v < nvb; ++v)
        for (int c = 0; c < 3; ++c) t[c] = std::fmin(t[c], VB[3 * v + c]), t[3 + c] = std::fmax(t[3 + c], VB[3 * v + c]);
      culled[i] = !boxes_may_overlap(X, t, q);
      continue;
    }
    if (nta == 1 || it.root < 0) {
      for (int c = 0; c < 6; ++c) t[c] = A.box[6 * (size_t)(nta == 1 ? 0 : ~it.root) + c];
    } else {
      const Node& N = A.nodes[it.root];
      for (int c = 0; c < 3; ++c) t[c] = std::fmin(N.b[c], N.b[6 + c]), t[3 + c] = std::fmax(N.b[3 + c], N.b[9 + c]);
    }
    culled[i] = !boxes_may_overlap(X, q, t);
  }
  /* candidates: every item (culled ones too, to check the cull) */
  const int64_t cap = 1 << 24;
  std::vector<int4> surv(cap);
  unsigned long long scnt = 0;
  std::vector<int> ncand(nitems, 0);
  std::vector<unsigned> nsurv(nitems + 1, 0), off(nitems + 1, 0);
  int err = 0;
  QueryArgs Q{};
  Q.envs = &E, Q.state = state, Q.xf = X, Q.items = items.data(), Q.nitems = nitems, Q.queries_per_lane = 1;
  Q.error = &err, Q.surv = surv.data(), Q.surv_cnt = &scnt, Q.surv_cap = cap;
  Q.item_ncand = ncand.data(), Q.item_nsurv = nsurv.data();
  WaveEmu wave;
  g_wave = &wave;
  /* items in reverse order, so the append order of survivors differs from
   * item order as it does on the GPU */
  for (int it = nitems - 1; it >= 0; --it) {
    CandLds* L = new CandLds();
    std::vector<std::thread> th;
    for (int lane = 0; lane < 64; ++lane)
      th.emplace_back([&, lane] {
        g_lane = lane;
        threadIdx.x = lane;
        unsigned long long sn = 0, se = 0;
        /* the kernels the engine launches: built for the pair's kind for
         * triangle and soft queries, the general one otherwise */
        if (E.kind == KIND_TRI && !g_generic) run_cand<CAND_STEP, KIND_TRI>(*L, Q, it, items[it], sn, se);
        else if (E.kind == KIND_SOFT && !g_generic) run_cand<CAND_STEP, KIND_SOFT>(*L, Q, it, items[it], sn, se);
        else run_cand<CAND_STEP>(*L, Q, it, items[it], sn, se);
        fill_sentinels(Q, sn, se);
      });
    for (auto& t : th) t.join();
    delete L;
  }
  for (int it = 0; it < nitems; ++it) off[it + 1] = off[it] + nsurv[it];
  int64_t nc = 0, bad = 0, ncull = 0;
  for (int it = 0; it < nitems; ++it) nc += ncand[it], ncull += culled[it], bad += culled[it] && ncand[it] > 0;
  const double dsE = seeds ? seeds[0] : 0.0, invE = seeds ? seeds[1] : 0.0;
  ClipArgs C{off.data(), &dsE, &invE, (int64_t)off[nitems]};
  std::vector<Contrib> con(std::max<unsigned>(1, off[nitems]));
  for (int dir = -1; dir < ndir; ++dir) {
    for (auto& q : con) q = Contrib{};
    /* the double clips run on a column polygon (as in the clip kernel), the
     * general build and the tangent pass on register polygons */
    double colbuf[3 * HCS_MAXV] = {};
    for (unsigned long long i = 0; i < scnt; ++i) {
      if (dir < 0 && E.kind == KIND_TRI && !g_generic)
        clip_survivor<double, KIND_TRI, true>(Q, C, (int64_t)i, dir, con.data(), nullptr, colbuf, 1);
      else if (dir < 0 && E.kind == KIND_SOFT && !g_generic)
        clip_survivor<double, KIND_SOFT, true>(Q, C, (int64_t)i, dir, con.data(), nullptr, colbuf, 1);
      else if (dir < 0) clip_survivor<double>(Q, C, (int64_t)i, dir, con.data(), nullptr);
      else clip_survivor<Dual>(Q, C, (int64_t)i, dir, con.data(), nullptr);
    }
    double F[3] = {0, 0, 0}, tA[3] = {0, 0, 0}, tB[3] = {0, 0, 0};
    int64_t np = 0, nf = 0;
    for (unsigned k = 0; k < off[nitems]; ++k) {
      for (int c = 0; c < 3; ++c) F[c] += con[k].F[c], tA[c] += con[k].tA[c], tB[c] += con[k].tB[c];
      np += con[k].np, nf += con[k].nf;
    }
    double* w = wrench + 12 * (size_t)(dir + 1);
    for (int c = 0; c < 3; ++c) w[c] = F[c], w[3 + c] = tA[c], w[6 + c] = -F[c], w[9 + c] = tB[c];
    if (dir < 0) counts[0] = nc, counts[1] = np, counts[2] = nf;
  }
  stats[0] = nitems, stats[1] = ncull, stats[2] = bad, stats[3] = off[nitems], stats[4] = err;
  return 0;
}

}  // extern "C"

extern "C" {

void emu_set_options(int generic) { g_generic = generic; }

/* A wave whose even and odd lanes reach two different ballots: the
 * emulation must abort on it (a collective under divergent control flow). */
int emu_divergent_ballot_probe() {
  WaveEmu wave;
  g_wave = &wave;
  std::vector<std::thread> th;
  unsigned long long r[64];
  for (int lane = 0; lane < 64; ++lane)
    th.emplace_back([&, lane] {
      g_lane = lane;
      threadIdx.x = lane;
      if (lane & 1) r[lane] = __ballot(1);
      else r[lane] = __ballot(0);
    });
  for (auto& t : th) t.join();
  return (int)(r[0] & 1);
}

/* The separating-axis prefilter of the candidate kernel on one box
 * (lo[3], hi[3]) and one triangle (3 x 3 vertices, unit normal n). */
int emu_box_may_meet_triangle(const double* box, const double* tri, const double* n) {
  return box_may_meet_triangle(box, tri, n) ? 1 : 0;
}

}  // extern "C"
