/*
 * hcs_noalloc.hip -- a contact step that allocates nothing
 * (hcs_contact_step_noalloc).  No reference counterpart: it is
 * hcs_contact_step for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc).
 *
 * Every device buffer of the step (env descriptors, state, work items, the
 * pipeline's counters, item boxes, live lists, survivors, contributions,
 * scan storage, reduction partials and per-pair outputs) is a slice of one
 * file-scope array, `scratch`, of HCS_SCRATCH_BYTES on each device.  Its
 * device address comes from hipGetSymbolAddress, taken once per device
 * with that device current (hcs_create calls scratch_setup); host code only
 * ever uses that address.  The slices are laid out per call from the call's
 * sizes; a call that does not fit returns HCS_ERR_CAPACITY before anything
 * runs, or (survivors beyond the space left) after the step with every
 * pair's counts valid and its status HCS_ERR_CAPACITY.
 *
 * The array is shared by every call on one device and keeps the last call's
 * contents; every counter is cleared by a fill kernel at the start of each
 * step.  Calls that overlap on one device (two contexts, two host threads)
 * race on it.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "hcs_internal.h"
#include "hcs_prim.h"

/* 64 MiB per device: room for about 690k survivors of the exact
 * rejections after the per-call tables of a few hundred pairs. */
#define HCS_SCRATCH_BYTES (64ull << 20)

static __device__ __align__(256) unsigned char scratch[HCS_SCRATCH_BYTES];

namespace hcs {

namespace {
std::mutex g_scratch_mu;
/* per device ordinal: the device address of `scratch` (nullptr: not yet) */
std::vector<void*> g_scratch_ptr;
}  // namespace

int scratch_setup(Ctx* c) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= c->device) g_scratch_ptr.resize(c->device + 1, nullptr);
  if (!g_scratch_ptr[c->device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != c->device) e = hipSetDevice(c->device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_check(c, e == hipSuccess ? hipErrorInvalidSymbol : e, "hipGetSymbolAddress");
    g_scratch_ptr[c->device] = p;
  }
  c->scratch = g_scratch_ptr[c->device];
  c->scratch_bytes = HCS_SCRATCH_BYTES;
  return HCS_OK;
}

namespace {

/* Carves 256-byte-aligned fixed slices off the scratch array. */
struct Carver {
  unsigned char* base;
  size_t size, used = 0;
  bool ok = true;
  void take(DBuf& b, size_t bytes) {
    bytes = std::max<size_t>(bytes, 8);
    const size_t at = (used + 255) & ~(size_t)255;
    b.fixed = true;
    if (at + bytes > size) {
      ok = false;
      b.ptr = nullptr, b.bytes = 0;
      return;
    }
    b.ptr = base + at, b.bytes = bytes;
    used = at + bytes;
  }
  size_t left() const { return size > ((used + 255) & ~(size_t)255) ? size - ((used + 255) & ~(size_t)255) : 0; }
};

}  // namespace

}  // namespace hcs

using namespace hcs;

extern "C" int hcs_contact_step_noalloc(hcs_ctx* h, int32_t npairs, const hcs_pair_in* in, const hcs_step_opts* opts,
                                        hcs_pair_out* out) {
  if (!h || npairs < 0 || (npairs > 0 && (!in || !out))) return HCS_ERR_ARG;
  Ctx* c = &h->c;
  if (opts && opts->keep_polygons)
    return set_err(c, HCS_ERR_ARG, "hcs_contact_step_noalloc keeps no polygons (keep_polygons must be 0)");
  if (!c->scratch) return set_err(c, HCS_ERR_STATE, "scratch array not set up on this device");
  if (npairs == 0) return HCS_OK;
  int rc = hip_check(c, hipSetDevice(c->device), "hipSetDevice");
  if (rc) return rc;
  /* the plan on the host, then the slices for its sizes */
  std::vector<int32_t> geoms(2 * (size_t)npairs);
  for (int e = 0; e < npairs; ++e) geoms[2 * e] = in[e].geom_a, geoms[2 * e + 1] = in[e].geom_b;
  Plan plan;
  if ((rc = build_plan(c, npairs, geoms.data(), opts, c->tet_mode_max_tris, plan))) return rc;
  const size_t ni = std::max<size_t>(1, plan.items.size());
  std::vector<int4> rt;
  std::vector<int2> pt;
  reduction_slices(plan, rt, pt);
  const size_t scan_bytes = scan_temp_bytes((long long)ni + 1);
  Carver cv{(unsigned char*)c->scratch, c->scratch_bytes};
  DBuf staged[3];
  Pipe P;
  cv.take(staged[0], sizeof(EnvDesc) * npairs);
  cv.take(staged[1], sizeof(double) * HCS_STATE * npairs);
  cv.take(staged[2], sizeof(Item) * ni);
  cv.take(P.counter, 64);
  cv.take(P.surv_cnt, 64);
  cv.take(P.err, 64);
  cv.take(P.xf, sizeof(double) * 12 * npairs);
  cv.take(P.ncand, sizeof(int) * ni);
  cv.take(P.nsurv, sizeof(unsigned) * (ni + 1));
  cv.take(P.off, sizeof(unsigned) * (ni + 1));
  cv.take(P.out, sizeof(PairOut) * npairs);
  cv.take(P.scan_tmp, std::max<size_t>(scan_bytes, 8));
  cv.take(P.ibox, sizeof(double) * 12 * ni);
  cv.take(P.live, sizeof(int) * ni);
  cv.take(P.live_item, sizeof(Item) * ni);
  cv.take(P.qseg, sizeof(int) * (HCS_QUEUES + 1));
  cv.take(P.xtri, sizeof(double) * HCS_XTRI * std::max<int64_t>(1, plan.ntx));
  cv.take(P.xenv, sizeof(int) * std::max<size_t>(1, plan.xenv.size()));
  cv.take(P.order, 8);
  cv.take(P.rtab, sizeof(int4) * std::max<size_t>(1, rt.size()));
  cv.take(P.ptab, sizeof(int2) * pt.size());
  cv.take(P.part, sizeof(RedPart) * HCS_MAXDIR * std::max<size_t>(1, rt.size()));
  /* the rest: survivors (16 B) and their contributions (80 B), less the
   * alignment of the two slices */
  const int64_t cap = (int64_t)(cv.left() > 512 ? (cv.left() - 512) / (sizeof(int4) + sizeof(Contrib)) : 0);
  if (!cv.ok || cap < 4096)
    return set_err(c, HCS_ERR_CAPACITY,
                   "hcs_contact_step_noalloc: " + std::to_string(npairs) + " pairs and " + std::to_string(ni) +
                       " work items do not fit the " + std::to_string(c->scratch_bytes >> 20) + " MiB scratch array");
  cv.take(P.surv, sizeof(int4) * cap);
  cv.take(P.contrib, sizeof(Contrib) * cap);
  if (!cv.ok) return set_err(c, HCS_ERR_CAPACITY, "hcs_contact_step_noalloc: scratch layout");
  if ((rc = stage_pairs(c, npairs, in, opts, plan, staged))) return rc;
  const int nitems = (int)plan.items.size();
  /* every request below fits its slice (ensure() on a fixed slice never
   * allocates and reports anything larger as HCS_ERR_CAPACITY) */
  if ((rc = pipe_reserve(c, P, nitems, npairs, cap, 0))) return rc;
  if ((rc = pipe_plan(c, P, (const EnvDesc*)staged[0].ptr, (const Item*)staged[2].ptr, plan))) return rc;
  QueryArgs a = pipe_args(P, (const EnvDesc*)staged[0].ptr, (const double*)staged[1].ptr, (const Item*)staged[2].ptr,
                          nitems, false);
  pipe_enqueue(c, P, a, npairs, nitems, clip_grid_for(c, P.cap), 0, nullptr, nullptr, false, 0);
  rc = hip_check(c, hipGetLastError(), "noalloc step");
  unsigned long long nsurv = 0;
  int err = 0;
  if (!rc) rc = hip_check(c, hipMemcpyAsync(&nsurv, P.surv_cnt.ptr, 8, hipMemcpyDeviceToHost, c->stream), "dl");
  if (!rc) rc = hip_check(c, hipMemcpyAsync(&err, P.err.ptr, sizeof(int), hipMemcpyDeviceToHost, c->stream), "dl");
  static_assert(sizeof(hcs_pair_out) == sizeof(PairOut), "pair_out layout");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(out, P.out.ptr, sizeof(PairOut) * npairs, hipMemcpyDeviceToHost, c->stream),
                   "dl");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  if (err) return set_err(c, HCS_ERR_LIMIT, "traversal stack overflow");
  if ((int64_t)nsurv > P.cap)
    return set_err(c, HCS_ERR_CAPACITY, "hcs_contact_step_noalloc: " + std::to_string(nsurv) +
                                            " survivors exceed the scratch array's room for " +
                                            std::to_string(P.cap));
  return HCS_OK;
}
