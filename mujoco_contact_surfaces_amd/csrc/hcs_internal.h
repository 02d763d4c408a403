/*
 * hcs_internal.h -- host-side context shared by the engine translation units.
 */
#pragma once
#include <hip/hip_runtime.h>

#include <set>
#include <string>
#include <vector>

#include "../../include/hcs.h"
#include "hcs_kernels.h"

namespace hcs {

struct Mesh {
  int kind = 0; /* 1 soft, 2 rigid */
  int32_t geom_id = 0;
  hcs_props props{};
  int32_t nv = 0, nelem = 0, height = 0;
  double* V = nullptr;
  int32_t* E = nullptr; /* T (soft) or F (rigid) */
  double* p = nullptr;
  double* rec = nullptr;
  float* fpl = nullptr; /* soft: the record's four face planes in float, 64 B per tet */
  double* box = nullptr;
  Node* nodes = nullptr;
  WNode* wnodes = nullptr; /* soft: 4-wide nodes, one per internal LBVH node */
  int32_t* leaf = nullptr; /* tets in Morton (BVH leaf) order */
  int32_t* qorder = nullptr; /* rigid: triangles in Morton order of their centroids */
  double* boxL = nullptr;   /* their AABBs in that order */
  double* nrm = nullptr;
};

/* Grow-only device buffer.  A fixed buffer is a view of memory the library
 * does not own (a slice of the static scratch array of the no-allocation
 * entry point): ensure() never reallocates or frees it, and reports a
 * request beyond its size as HCS_ERR_CAPACITY. */
struct DBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  bool fixed = false;
};

/* The per-step pipeline (hcs_engine.hip): k_prep, k_cand, scan of
 * per-item survivor counts, k_clip, k_env_reduce. */
struct Pipe {
  DBuf counter /* [0..7] work counters, [8..15] live items per queue */, surv, surv_cnt, ncand, nsurv, off, scan_tmp,
      contrib, out, err, xf, ibox /* item boxes, in the work order when ordered */, live, live_item, qseg, xtri, xenv,
      order, item_o /* the items in the work order (ordered plans) */, part /* per-slice sums */,
      rtab /* reduction waves (pair, slice, slices, first partial) */, ptab /* per pair (first partial, slices) */;
  int64_t cap = 0;
  size_t scan_bytes = 0;
  std::vector<int> qseg_host; /* host copy of the queue segments (the upload source) */
  std::vector<Item> item_o_host; /* upload source of item_o */
  int nqueues = 1;            /* candidate-kernel work queues (1..HCS_QUEUES) */
  int nxenv = 0;              /* pairs in tet-query mode (their triangles are moved per step) */
  bool ordered = false;       /* order holds a work order of the items */
  int kinds = 0;
  int nred = 0;                 /* reduction waves */
  std::vector<int4> rtab_host;  /* upload sources of rtab / ptab */
  std::vector<int2> ptab_host;              /* bit k set: some pair of the plan is of EnvKind k (0: not known) */
};

struct Ctx;
int set_err(Ctx* c, int code, const std::string& msg);
int hip_check(Ctx* c, hipError_t e, const char* what);
int ensure(Ctx* c, DBuf& b, size_t bytes);

/* Per-call plan: env descriptors and items for a set of pairs. */
struct Plan {
  std::vector<EnvDesc> envs;
  std::vector<Item> items;
  std::vector<int> pair_kind;
  std::vector<int> xenv; /* pairs in tet-query mode */
  int64_t ntx = 0;       /* their triangles in total */
  std::vector<int> order; /* items by decreasing measured work (empty: item order) */
};

/* Scratch of hcs_taxel_sensor_update. */
struct TaxelSensorBufs {
  DBuf mask, k64, kp, idx, cnt, nfan, fan_off, tmp, tris, surf, scnt, soff, P, E, taxel, val;
};

struct Ctx {
  int device = 0;
  int num_cu = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<Mesh> meshes;
  /* scratch for hcs_contact_step / hcs_candidates / hcs_taxels */
  DBuf d_envs, d_state, d_items;
  Pipe pipe;
  DBuf d_face_rec, d_face_key, d_face_cnt;
  DBuf d_poly_rec, d_poly_key, d_poly_cnt;
  DBuf d_cand, d_cand_cnt;
  DBuf d_tx_cnt, d_tx_off, d_tx_list, d_tx_out, d_tx_grid;
  TaxelSensorBufs ts;
  /* last contact step's polygons (for taxels) */
  int64_t last_npoly = 0;
  int last_valid = 0;
  std::vector<hcs_pair_in> last_pairs;
  std::vector<int32_t> last_geoms; /* per pair: geom_a, geom_b */
  int tet_mode_max_tris = 64;
  /* device address and size of the static scratch array on this device */
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  /* live batches of this context: hcs_destroy releases them first */
  std::set<hcs_batch*> batches;
};

/* Builds env descriptors and work items (chunks of 64 queries, one per
 * lane of the candidate kernel's wave).  Returns HCS_OK or an error. */
int build_plan(Ctx* c, int npairs, const int32_t* geoms, const hcs_step_opts* o, int tet_mode_max_tris,
               Plan& plan);

int pipe_reserve(Ctx* c, Pipe& p, int nitems, int npairs, int64_t cap, int ndir);
QueryArgs pipe_args(const Pipe& p, const EnvDesc* envs, const double* state, const Item* items, int nitems,
                    bool persistent);
/* ndir = 0: forward (double); ndir > 0: tangent directions dir0 ..
 * dir0+ndir-1 (contributions and outputs indexed by direction - dir0). */
void pipe_enqueue(Ctx* c, Pipe& p, const QueryArgs& a, int npairs, int cand_grid, int clip_grid, int ndir,
                  const double* dstar_dE, const double* invE, bool cand_only, int dir0);
/* Static boxes of the current items (once per plan, after the items and
 * env descriptors are on the device). */
struct Plan;
int pipe_plan(Ctx* c, Pipe& p, const EnvDesc* envs, const Item* items, const Plan& plan);
/* Counter resets, pose composition and the item cull of one step. */
void pipe_front(Ctx* c, Pipe& p, const QueryArgs& a, int npairs);
/* For each requested item (index into the device items), the LBVH
 * frontier below its root at the requested depth (<= 6): out[65 r] = count,
 * then node codes. */
int split_roots(Ctx* c, const EnvDesc* envs, const Item* items, const std::vector<int>& req,
                const std::vector<int>& depth, std::vector<int32_t>& out);
void pipe_free(Pipe& p);
/* mode: CandMode */
void launch_cand(hipStream_t s, const QueryArgs& a, int grid, int mode);
void launch_clip_reduce(Ctx* c, Pipe& p, const QueryArgs& a, int npairs, int clip_grid, hipEvent_t e_clip,
                        hipEvent_t e_reduce);
/* Builds the plan of npairs pairs and uploads env descriptors, state and
 * items into the context's scratch buffers, or into dst[0..2] (envs, state,
 * items) when given. */
int stage_pairs(Ctx* c, int npairs, const hcs_pair_in* in, const hcs_step_opts* opts, Plan& plan,
                DBuf* dst = nullptr);
/* The per-pair reduction waves of a plan (see pipe_plan). */
void reduction_slices(const Plan& plan, std::vector<int4>& rt, std::vector<int2>& pt);
/* The static scratch array of the no-allocation entry point on the
 * context's device (set up by hcs_create). */
int scratch_setup(Ctx* c);
/* Plans and captures again every batch of the context that steps mesh
 * `handle` (its geometry was replaced). */
int batches_refresh(Ctx* c, int handle);
int clip_grid_for(const Ctx* c, int64_t cap);
void launch_prep(hipStream_t s, const EnvDesc* envs, const double* state, double* xf, int n);
void launch_fill_u32(hipStream_t s, unsigned int* p, unsigned int v, int n);

int taxel_grid_dev(const hcs_taxel_grid& G, TaxelGridDev& g);
int run_taxels(Ctx* c, const TaxelGridDev& g, const unsigned char* d_mask, const double* poly, const double* ptan,
               const int64_t* key, int64_t np, const double* Xd, double* d_out);

/* Flat tactile sensors of a batch (hcs_batch_set_sensors): sensor s samples
 * the polygons of its pair only, at the pose of one of the pair's two
 * geometries in the batch state.  Every buffer is sized when the sensors are
 * set, except the polygon records and the per-taxel polygon lists, which
 * grow when a step overflows them (checked after the step). */
struct BatchTaxels {
  int n = 0;     /* sensors */
  int ncell = 0; /* taxels of all sensors */
  DBuf grid /* TaxelGridDev per sensor, pose written per step */, src /* int per sensor: 36 pair + 12 side */,
      cell_base /* int[n + 1] */, cell_sensor /* int per taxel */, psoff /* int[npairs + 1] */, psens /* int[n] */,
      mask /* uchar per pair: has a sensor */, poly_rec, poly_key, poly_cnt /* ull */, cnt /* unsigned[ncell + 1] */,
      cursor /* unsigned[ncell] */, off /* unsigned[ncell + 1] */, scan_tmp, list /* int[list_cap] */,
      out /* double per taxel */, live /* int[ncell]: taxels with candidate polygons */;
  int64_t poly_cap = 0, list_cap = 0;
  size_t scan_bytes = 0;
  int bin_grid = 1;
  int max_rays = 0; /* the largest S^2 of the sensors */
};
int batch_taxels_setup(Ctx* c, BatchTaxels& t, int npairs, const int32_t* geoms, int n, const int32_t* pair,
                       const hcs_taxel_grid* grids);
int batch_taxels_reserve(Ctx* c, BatchTaxels& t, int64_t poly_cap, int64_t list_cap);
/* Sensor poses from the state, binning and sampling of the kept polygons
 * (after the clip kernel of the same step). */
void batch_taxels_enqueue(Ctx* c, BatchTaxels& t, const double* state);
/* Polygons and list entries the last step asked for (to detect overflow). */
int batch_taxels_demand(Ctx* c, BatchTaxels& t, int64_t& npoly, int64_t& nlist);
void batch_taxels_free(BatchTaxels& t);

}  // namespace hcs

struct hcs_ctx {
  hcs::Ctx c;
};
