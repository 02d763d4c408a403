/*
 * hcs_adjoint.hip -- differentiation of the contact step at fixed clipping
 * topology (no reference counterpart).
 *
 * Tangent-linear recomputation with no stored polygons: the candidate kernel
 * runs once on primal values (same survivors as the forward step), then the
 * clip kernel is instantiated with the Dual scalar type and run once per
 * tangent direction (grid.y): rotation and translation of geometry B's pose,
 * and the modulus of geometry A.  Inside/outside tests, filters and branches
 * follow the primal values, so the result is the exact derivative of the
 * piecewise-smooth map on the current piece.  The vector-Jacobian product
 * (the adjoint) is J^T applied to the output cotangent; with 7 inputs and 12
 * (+ taxel) outputs the tangent form is the cheaper way to get it.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "hcs_device.h"
#include "hcs_internal.h"
#include "hcs_query.h"

#pragma clang fp contract(off)

using namespace hcs;

namespace hcs {

static const int kDirs = 7;

/* The tangent pass stays switched off: an earlier fused tangent kernel went
 * 180 s without finishing on gfx950 and its cause has not been shown, so
 * every differentiation entry point returns HCS_ERR_STATE until it is.  The
 * code below (forward candidate kernel + Dual clip kernel) is exercised on
 * the CPU only. */
static const bool kTangentEnabled = false;

/* Per-pair tangent seeds: d(d*)/dE_A then 1/E_A. */
static void seeds(Ctx* c, const hcs_pair_in* in, int npairs, std::vector<double>& sd) {
  sd.assign(2 * (size_t)npairs, 0.0);
  for (int e = 0; e < npairs; ++e) {
    const Mesh& A = c->meshes[in[e].geom_a];
    const Mesh& B = c->meshes[in[e].geom_b];
    const double EA = A.props.hydroelastic_modulus;
    sd[npairs + e] = 1.0 / EA;
    if (B.kind == 1) {
      /* d* = E_B/(E_A+E_B) d_A + E_A/(E_A+E_B) d_B (plugin.cpp:138-159) */
      const double EB = B.props.hydroelastic_modulus, s = EA + EB;
      sd[e] = EB * (B.props.dissipation - A.props.dissipation) / (s * s);
    }
  }
}

/* One tangent pass over npairs pairs for directions dir0..dir0+ndir-1;
 * per-direction outputs land in c->pipe.out.  Optional polygon output with
 * tangents (taxel Jacobian). */
static int tangent_pass(Ctx* c, int npairs, const hcs_pair_in* in, const hcs_step_opts* opts, int dir0, int ndir,
                        double* poly_rec, double* poly_tan, int64_t* poly_key, int64_t poly_cap, Plan& plan,
                        unsigned long long* npoly) {
  if (ndir < 1 || ndir > HCS_MAXDIR) return set_err(c, HCS_ERR_ARG, "tangent directions out of range");
  int rc = stage_pairs(c, npairs, in, opts, plan);
  if (rc) return rc;
  const int nitems = (int)plan.items.size();
  std::vector<double> sd;
  seeds(c, in, npairs, sd);
  DBuf dseed;
  if ((rc = ensure(c, dseed, sizeof(double) * sd.size())) || (rc = ensure(c, c->d_poly_cnt, 64))) return rc;
  rc = hip_check(c, hipMemcpyAsync(dseed.ptr, sd.data(), sd.size() * 8, hipMemcpyHostToDevice, c->stream), "upload");
  Pipe& P = c->pipe;
  for (int attempt = 0; attempt < 3 && !rc; ++attempt) {
    rc = pipe_reserve(c, P, nitems, npairs, std::max<int64_t>(P.cap, 4096), ndir);
    if (!rc) rc = pipe_plan(c, P, (const EnvDesc*)c->d_envs.ptr, (const Item*)c->d_items.ptr, plan);
    if (rc) break;
    rc = hip_check(c, hipMemsetAsync(c->d_poly_cnt.ptr, 0, 64, c->stream), "memset");
    if (rc) break;
    QueryArgs a = pipe_args(P, (const EnvDesc*)c->d_envs.ptr, (const double*)c->d_state.ptr,
                            (const Item*)c->d_items.ptr, nitems, false);
    if (poly_rec) {
      a.poly_rec = poly_rec, a.poly_tan = poly_tan, a.poly_key = poly_key, a.poly_cap = poly_cap;
      a.poly_cnt = (unsigned long long*)c->d_poly_cnt.ptr;
    }
    pipe_enqueue(c, P, a, npairs, nitems, clip_grid_for(c, P.cap), ndir, (const double*)dseed.ptr,
                 (const double*)dseed.ptr + npairs, false, dir0);
    rc = hip_check(c, hipGetLastError(), "tangent pass");
    unsigned long long ns = 0;
    if (!rc) rc = hip_check(c, hipMemcpyAsync(&ns, P.surv_cnt.ptr, 8, hipMemcpyDeviceToHost, c->stream), "dl");
    if (!rc && npoly) rc = hip_check(c, hipMemcpyAsync(npoly, c->d_poly_cnt.ptr, 8, hipMemcpyDeviceToHost, c->stream), "dl");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
    if (rc) break;
    if ((int64_t)ns <= P.cap) break;
    rc = pipe_reserve(c, P, nitems, npairs, (int64_t)ns + (int64_t)ns / 4 + 1024, ndir);
  }
  (void)hipFree(dseed.ptr);
  if (rc) return rc;
  int err = 0;
  rc = hip_check(c, hipMemcpyAsync(&err, P.err.ptr, sizeof(int), hipMemcpyDeviceToHost, c->stream), "dl");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (!rc && err) rc = set_err(c, HCS_ERR_LIMIT, "traversal stack overflow");
  return rc;
}

}  // namespace hcs

extern "C" {

int hcs_contact_jacobian(hcs_ctx* h, int32_t npairs, const hcs_pair_in* in, const hcs_step_opts* opts, double* J) {
  if (!h || npairs < 0 || (npairs > 0 && (!in || !J))) return HCS_ERR_ARG;
  Ctx* c = &h->c;
  if (!kTangentEnabled) return set_err(c, HCS_ERR_STATE, "differentiation is switched off in this build");
  if (npairs == 0) return HCS_OK;
  Plan plan;
  int rc = tangent_pass(c, npairs, in, opts, 0, kDirs, nullptr, nullptr, nullptr, 0, plan, nullptr);
  if (rc) return rc;
  std::vector<PairOut> o((size_t)npairs * kDirs);
  rc = hip_check(c, hipMemcpyAsync(o.data(), c->pipe.out.ptr, sizeof(PairOut) * o.size(), hipMemcpyDeviceToHost,
                                   c->stream),
                 "download");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  for (int e = 0; e < npairs; ++e) {
    double* Je = J + (size_t)e * 12 * kDirs;
    for (int d = 0; d < kDirs; ++d) {
      const PairOut& r = o[(size_t)d * npairs + e];
      for (int k = 0; k < 3; ++k) {
        Je[(0 + k) * kDirs + d] = r.F_A[k];
        Je[(3 + k) * kDirs + d] = r.tau_A[k];
        Je[(6 + k) * kDirs + d] = r.F_B[k];
        Je[(9 + k) * kDirs + d] = r.tau_B[k];
      }
    }
  }
  return HCS_OK;
}

int hcs_contact_backward(hcs_ctx* h, int32_t npairs, const hcs_pair_in* in, const hcs_step_opts* opts,
                         const double* dL_dout, double* dL_dxi) {
  if (!h || npairs < 0 || (npairs > 0 && (!in || !dL_dout || !dL_dxi))) return HCS_ERR_ARG;
  std::vector<double> J((size_t)npairs * 12 * kDirs);
  int rc = hcs_contact_jacobian(h, npairs, in, opts, J.data());
  if (rc) return rc;
  for (int e = 0; e < npairs; ++e)
    for (int d = 0; d < kDirs; ++d) {
      double s = 0;
      for (int r = 0; r < 12; ++r) s += J[((size_t)e * 12 + r) * kDirs + d] * dL_dout[12 * (size_t)e + r];
      dL_dxi[kDirs * (size_t)e + d] = s;
    }
  return HCS_OK;
}

int hcs_taxel_jacobian(hcs_ctx* h, const hcs_pair_in* pair, const hcs_step_opts* opts, const hcs_taxel_grid* G,
                       double* Jt) {
  if (!h || !pair || !G || !Jt) return HCS_ERR_ARG;
  Ctx* c = &h->c;
  if (!kTangentEnabled) return set_err(c, HCS_ERR_STATE, "differentiation is switched off in this build");
  TaxelGridDev g;
  if (taxel_grid_dev(*G, g) != HCS_OK) return set_err(c, HCS_ERR_ARG, "invalid taxel grid");
  const int ncell = g.cx * g.cy;
  hcs_step_opts o = opts ? *opts : hcs_step_opts{0, 1, 0.001, 0, 0};
  /* polygon capacity from a primal run */
  hcs_pair_out po;
  hcs_step_opts oc = o;
  oc.keep_polygons = 0;
  int rc = hcs_contact_step(h, 1, pair, &oc, &po, nullptr);
  if (rc) return rc;
  const int64_t cap = std::max<int64_t>(1, po.n_polygons);
  DBuf prec, ptan, pkey, dXd, dmask, dout;
  auto cleanup = [&]() {
    for (DBuf* b : {&prec, &ptan, &pkey, &dXd, &dmask, &dout})
      if (b->ptr) (void)hipFree(b->ptr);
  };
  if ((rc = ensure(c, prec, sizeof(double) * HCS_PREC * cap)) || (rc = ensure(c, ptan, sizeof(double) * HCS_PREC * cap)) ||
      (rc = ensure(c, pkey, sizeof(int64_t) * 3 * cap)) || (rc = ensure(c, dXd, sizeof(double) * 12)) ||
      (rc = ensure(c, dmask, 8)) || (rc = ensure(c, dout, sizeof(double) * ncell))) {
    cleanup();
    return rc;
  }
  const unsigned char one = 1;
  rc = hip_check(c, hipMemcpyAsync(dmask.ptr, &one, 1, hipMemcpyHostToDevice, c->stream), "upload");
  const double* R = pair->X_WB;
  for (int d = 0; d < kDirs && !rc; ++d) {
    /* tangent of the sensor pose (the sensor rides on geometry B) */
    double Xd[12] = {0};
    if (d < 3) {
      for (int j = 0; j < 3; ++j) {
        const double r0 = R[j], r1 = R[3 + j], r2 = R[6 + j];
        double d0 = 0, d1 = 0, d2 = 0;
        if (d == 0) d1 = -r2, d2 = r1;
        if (d == 1) d0 = r2, d2 = -r0;
        if (d == 2) d0 = -r1, d1 = r0;
        Xd[j] = d0, Xd[3 + j] = d1, Xd[6 + j] = d2;
      }
    } else if (d < 6) {
      Xd[9 + (d - 3)] = 1.0;
    }
    rc = hip_check(c, hipMemcpyAsync(dXd.ptr, Xd, sizeof Xd, hipMemcpyHostToDevice, c->stream), "upload");
    if (rc) break;
    Plan plan;
    unsigned long long np = 0;
    rc = tangent_pass(c, 1, pair, &o, d, 1, (double*)prec.ptr, (double*)ptan.ptr, (int64_t*)pkey.ptr, cap, plan, &np);
    if (rc) break;
    const int64_t n = std::min<int64_t>((int64_t)np, cap);
    rc = run_taxels(c, g, (const unsigned char*)dmask.ptr, (const double*)prec.ptr, (const double*)ptan.ptr,
                    (const int64_t*)pkey.ptr, n, (const double*)dXd.ptr, (double*)dout.ptr);
    std::vector<double> col(ncell);
    if (!rc)
      rc = hip_check(c, hipMemcpyAsync(col.data(), dout.ptr, sizeof(double) * ncell, hipMemcpyDeviceToHost, c->stream),
                     "dl");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
    for (int k = 0; k < ncell && !rc; ++k) Jt[(size_t)k * kDirs + d] = col[k];
  }
  cleanup();
  return rc;
}

}  // extern "C"
