"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker or the timed CPU baseline.  The product
(mujoco_contact_surfaces_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
TREC, FREC, PREC = 32, 12, 40

_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_tet_precompute.argtypes = [_D, C.c_int, _I32, C.c_int, _D, _D]
        L.orc_tri_normals.argtypes = [_D, C.c_int, _I32, C.c_int, _D]
        L.orc_compose_SR.argtypes = [_D, _D, _D]
        L.orc_candidates.restype = C.c_int64
        L.orc_candidates.argtypes = [_D, C.c_int, _I32, C.c_int, _D, C.c_int, _I32, C.c_int, _D, _D, C.c_int, _I64,
                                     C.c_int64]
        L.orc_clip_tri_tet.restype = C.c_int
        L.orc_clip_tri_tet.argtypes = [_D, _D, _D]
        for f in ("orc_combined_modulus", "orc_combined_friction"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double, C.c_double]
        L.orc_combined_dissipation.restype = C.c_double
        L.orc_combined_dissipation.argtypes = [C.c_double] * 4
        L.orc_point_force.restype = C.c_double
        L.orc_point_force.argtypes = [_D, _D, _D, _D, _D, C.c_double, C.c_double, C.c_int, _D]
        L.orc_soft_rigid.argtypes = [_D, C.c_int, _I32, C.c_int, _D, _D, C.c_int, _I32, C.c_int, _D, _D, C.c_int, _D,
                                     _I64, _D, _I64, C.c_int64, _D, _I64, C.c_int64]
        L.orc_soft_soft.argtypes = [_D, C.c_int, _I32, C.c_int, _D, _D, C.c_int, _I32, C.c_int, _D, _D, _D, C.c_int,
                                    _D, _I64, _D, _I64, C.c_int64, _D, _I64, C.c_int64]
        L.orc_soft_plane.argtypes = [_D, C.c_int, _I32, C.c_int, _D, _D, _D, _D, _I64, _D, _I64, C.c_int64, _D,
                                     _I64, C.c_int64]
        L.orc_taxel_sensor.restype = C.c_int64
        L.orc_taxel_sensor.argtypes = [_D, _I32, C.c_int64, _D, _D, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.c_int, _D]
        L.orc_minstd_uniform.argtypes = [C.c_int, _D]
        L.orc_curved_sensor.restype = C.c_int
        L.orc_curved_sensor.argtypes = [_D, C.c_int64, _D, _D, _D, C.c_int, _I32, _I32, _D, C.c_int, C.c_double, _D]
        L.orc_soft_prepare.restype = C.c_void_p
        L.orc_soft_prepare.argtypes = [_D, C.c_int, _I32, C.c_int, _D]
        L.orc_soft_free.argtypes = [C.c_void_p]
        L.orc_soft_rigid_step.argtypes = [C.c_void_p, _D, C.c_int, _I32, C.c_int, _D, _D, _D, _I64]
        L.orc_taxels.argtypes = [_D, C.c_int64, _D, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _D]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I32)


def lp(a):
    return a.ctypes.data_as(_I64)


def tet_records(V, T, p):
    V, T, p = _d(V), _i(T), _d(p)
    out = np.zeros((T.shape[0], TREC))
    lib().orc_tet_precompute(dp(V), V.shape[0], ip(T), T.shape[0], dp(p), dp(out))
    return out


def tri_normals(V, F):
    V, F = _d(V), _i(F)
    out = np.zeros((F.shape[0], 3))
    lib().orc_tri_normals(dp(V), V.shape[0], ip(F), F.shape[0], dp(out))
    return out


def compose_SR(X_WS, X_WR):
    out = np.zeros(12)
    lib().orc_compose_SR(dp(_d(X_WS)), dp(_d(X_WR)), dp(out))
    return out


def candidates(VS, T, VR, F, X_WS, X_WR, brute=False):
    VS, T, VR, F = _d(VS), _i(T), _d(VR), _i(F)
    X_WS, X_WR = _d(X_WS), _d(X_WR)
    n = lib().orc_candidates(dp(VS), VS.shape[0], ip(T), T.shape[0], dp(VR), VR.shape[0], ip(F), F.shape[0],
                             dp(X_WS), dp(X_WR), int(brute), None, 0)
    keys = np.zeros(max(1, n), dtype=np.int64)
    lib().orc_candidates(dp(VS), VS.shape[0], ip(T), T.shape[0], dp(VR), VR.shape[0], ip(F), F.shape[0], dp(X_WS),
                         dp(X_WR), int(brute), lp(keys), keys.size)
    return keys[:n]


def clip_tri_tet(tri, rec):
    out = np.zeros((8, 3))
    n = lib().orc_clip_tri_tet(dp(_d(tri)), dp(_d(rec)), dp(out))
    return out[:n].copy()


def point_force(p, n, fn0, k, damping, oA, VA, oB, VB, mu, dt=0.001, apply=True):
    pc = _d(list(p) + list(n) + [fn0, k, damping])
    f = np.zeros(3)
    fn = lib().orc_point_force(dp(pc), dp(_d(oA)), dp(_d(VA)), dp(_d(oB)), dp(_d(VB)), mu, dt, int(apply), dp(f))
    return f, fn


def soft_rigid(VS, T, pS, VR, F, X_WS, X_WR, V_WS, V_WR, E, d, mu_s, mu_r, representation=0, apply=True,
               soft_is_M=True, dt=0.001, brute=False, want=True, cap=None):
    """Returns dict(wrench (12), counts (3), faces, face_keys, polys, poly_keys)."""
    VS, T, pS, VR, F = _d(VS), _i(T), _d(pS), _d(VR), _i(F)
    pose = _d(np.concatenate([X_WS, X_WR, V_WS, V_WR]))
    params = _d([E, d, mu_s, mu_r, dt, int(apply), representation, int(soft_is_M)])
    wrench = np.zeros(12)
    counts = np.zeros(3, dtype=np.int64)
    if cap is None:
        lib().orc_soft_rigid(dp(VS), VS.shape[0], ip(T), T.shape[0], dp(pS), dp(VR), VR.shape[0], ip(F), F.shape[0],
                             dp(pose), dp(params), int(brute), dp(wrench), lp(counts), None, None, 0, None, None, 0)
        if not want:
            return dict(wrench=wrench, counts=counts)
        cap = (int(counts[2]) + 1, int(counts[1]) + 1)
    faces = np.zeros((cap[0], FREC))
    fk = np.zeros((cap[0], 3), dtype=np.int64)
    polys = np.zeros((cap[1], PREC))
    pk = np.zeros((cap[1], 2), dtype=np.int64)
    lib().orc_soft_rigid(dp(VS), VS.shape[0], ip(T), T.shape[0], dp(pS), dp(VR), VR.shape[0], ip(F), F.shape[0],
                         dp(pose), dp(params), int(brute), dp(wrench), lp(counts), dp(faces), lp(fk), cap[0],
                         dp(polys), lp(pk), cap[1])
    nf, npl = int(counts[2]), int(counts[1])
    return dict(wrench=wrench, counts=counts, faces=faces[:nf], face_keys=fk[:nf], polys=polys[:npl],
                poly_keys=pk[:npl])


def soft_plane(VS, T, pS, X_WS, X_WR, V_WS, V_WR, E, d, mu_s, mu_r, representation=0, apply=True, soft_is_M=True,
               dt=0.001, want=True):
    """Soft body vs the rigid half-space z <= 0 of frame R (a MuJoCo plane
    geom; plugin.cpp:296-299).  Returns dict(wrench, counts, faces,
    face_keys, polys, poly_keys); keys are (tet, 0, ...)."""
    VS, T, pS = _d(VS), _i(T), _d(pS)
    pose = _d(np.concatenate([X_WS, X_WR, V_WS, V_WR]))
    params = _d([E, d, mu_s, mu_r, dt, int(apply), representation, int(soft_is_M)])
    wrench = np.zeros(12)
    counts = np.zeros(3, dtype=np.int64)
    args = (dp(VS), VS.shape[0], ip(T), T.shape[0], dp(pS), dp(pose), dp(params), dp(wrench), lp(counts))
    lib().orc_soft_plane(*args, None, None, 0, None, None, 0)
    if not want:
        return dict(wrench=wrench, counts=counts)
    nf, npl = int(counts[2]), int(counts[1])
    faces = np.zeros((nf + 1, FREC))
    fk = np.zeros((nf + 1, 3), dtype=np.int64)
    polys = np.zeros((npl + 1, PREC))
    pk = np.zeros((npl + 1, 2), dtype=np.int64)
    lib().orc_soft_plane(*args, dp(faces), lp(fk), nf + 1, dp(polys), lp(pk), npl + 1)
    return dict(wrench=wrench, counts=counts, faces=faces[:nf], face_keys=fk[:nf], polys=polys[:npl],
                poly_keys=pk[:npl])


def soft_soft(VA, TA, pA, VB, TB, pB, X_WA, X_WB, V_WA, V_WB, EA, dA, muA, EB, dB, muB, representation=0,
              apply=True, a_is_M=True, dt=0.001, brute=False, want=True):
    VA, TA, pA, VB, TB, pB = _d(VA), _i(TA), _d(pA), _d(VB), _i(TB), _d(pB)
    pose = _d(np.concatenate([X_WA, X_WB, V_WA, V_WB]))
    params = _d([EA, dA, muA, EB, dB, muB, dt, int(apply), representation, int(a_is_M)])
    wrench = np.zeros(12)
    counts = np.zeros(3, dtype=np.int64)
    args = (dp(VA), VA.shape[0], ip(TA), TA.shape[0], dp(pA), dp(VB), VB.shape[0], ip(TB), TB.shape[0], dp(pB),
            dp(pose), dp(params), int(brute), dp(wrench), lp(counts))
    lib().orc_soft_soft(*args, None, None, 0, None, None, 0)
    if not want:
        return dict(wrench=wrench, counts=counts)
    nf, npl = int(counts[2]), int(counts[1])
    faces = np.zeros((nf + 1, FREC))
    fk = np.zeros((nf + 1, 3), dtype=np.int64)
    polys = np.zeros((npl + 1, PREC))
    pk = np.zeros((npl + 1, 2), dtype=np.int64)
    lib().orc_soft_soft(*args, dp(faces), lp(fk), nf + 1, dp(polys), lp(pk), npl + 1)
    return dict(wrench=wrench, counts=counts, faces=faces[:nf], face_keys=fk[:nf], polys=polys[:npl],
                poly_keys=pk[:npl])


def taxel_dims(half_size, resolution):
    import math
    return (int(math.floor(2 * half_size[0] / resolution + 0.1)), int(math.floor(2 * half_size[1] / resolution + 0.1)))


def taxels(polys, X_W_sensor, half_size, resolution, sampling_resolution=20, window=0, sigma=0.0):
    polys = _d(polys).reshape(-1, PREC)
    cx, cy = taxel_dims(half_size, resolution)
    sensor = _d(list(X_W_sensor) + list(half_size) + [resolution])
    out = np.zeros(cx * cy)
    lib().orc_taxels(dp(polys), polys.shape[0], dp(sensor), int(sampling_resolution), int(window), float(sigma), cx,
                     cy, dp(out))
    return out.reshape(cy, cx)


class PreparedSoft:
    """Load-time oracle state of one soft mesh; step() is the per-step work."""

    def __init__(self, V, T, p):
        self.V, self.T, self.p = _d(V), _i(T), _d(p)
        self.h = lib().orc_soft_prepare(dp(self.V), self.V.shape[0], ip(self.T), self.T.shape[0], dp(self.p))

    def step(self, VR, F, pose12x2_twists, params):
        VR, F = _d(VR), _i(F)
        pose = _d(pose12x2_twists)
        params = _d(params)
        wrench = np.zeros(12)
        counts = np.zeros(3, dtype=np.int64)
        lib().orc_soft_rigid_step(self.h, dp(VR), VR.shape[0], ip(F), F.shape[0], dp(pose), dp(params), dp(wrench),
                                  lp(counts))
        return wrench, counts

    def close(self):
        if self.h:
            lib().orc_soft_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


METHODS = {"closest": 0, "weighted": 1, "mean": 2, "squared": 3}
SAMPLE_METHODS = {"default": 0, "area_importance": 1}


def fan_triangles(polys, surface_ids):
    """kTriangle fan (v_k, v_k+1, C) of polygon records in order: (ntri, 12)
    rows [v0, v1, v2, e0, e1, e2] and the surface index of each."""
    rows, surf = [], []
    for P, sid in zip(np.asarray(polys), surface_ids):
        n = int(P[0])
        V = P[1:25].reshape(8, 3)
        e = P[25:33]
        C_, eC = P[33:36], P[36]
        for k in range(n):
            k2 = (k + 1) % n
            rows.append(np.concatenate([V[k], V[k2], C_, [e[k], e[k2], eC]]))
            surf.append(sid)
    return np.array(rows, float).reshape(-1, 12), np.array(surf, np.int32)


def taxel_sensor(tris, surf, X_W_sensor, taxels, include_margin, sample_resolution, method="squared",
                 sample_method="default", visualize=False, values=None):
    """TaxelSensor::internal_update (S/src/taxel_sensor.cpp:158-478) over
    fan triangles; returns (values, number of samples)."""
    tris, X, tx = _d(tris).reshape(-1, 12), _d(X_W_sensor), _d(taxels).reshape(-1, 3)
    surf = _i(surf)
    vals = np.zeros(tx.shape[0]) if values is None else _d(values).copy()
    m = lib().orc_taxel_sensor(dp(tris), ip(surf), tris.shape[0], dp(X), dp(tx), tx.shape[0], float(include_margin),
                               float(sample_resolution), METHODS[method], SAMPLE_METHODS[sample_method],
                               int(visualize), dp(vals))
    return vals, int(m)


def minstd_uniform(n):
    out = np.zeros(n)
    lib().orc_minstd_uniform(int(n), dp(out))
    return out


def curved_sensor(tris, X_W_sensor, points, normals, taxel_ptr, sample_idx, weights, include_margin):
    """CurvedSensor::internal_update (S/src/curved_sensor.cpp:388-481) over
    fan triangles; returns (values, rays that hit)."""
    tris, X = _d(tris).reshape(-1, 12), _d(X_W_sensor)
    P, N = _d(points).reshape(-1, 3), _d(normals).reshape(-1, 3)
    ptr, idx, w = _i(taxel_ptr), _i(sample_idx), _d(weights)
    n = ptr.shape[0] - 1
    vals = np.zeros(n)
    hits = lib().orc_curved_sensor(dp(tris), tris.shape[0], dp(X), dp(P), dp(N), P.shape[0], ip(ptr), ip(idx), dp(w),
                                   n, float(include_margin), dp(vals))
    return vals, int(hits)
