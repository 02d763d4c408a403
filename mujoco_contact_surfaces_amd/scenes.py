"""Seeded synthetic workloads for the engine's benchmark configurations.

Every random number comes from SplitMix64 keyed by (seed, config, env), so
the same bytes are produced on any host.  Geometry comes from the engine's
own generators (hcs_gen_*), which stand in for the Drake mesh and pressure
field generators the plugin calls (mujoco_contact_surfaces_plugin.cpp:648-796).

Configurations (sizes as generated; tet counts are recorded, not assumed):
  sphere_box      soft sphere r = 0.05 m (icosphere L3, K = 2: 5,120 tets,
                  E = 1e5 Pa) pressed 5 mm into a 0.2 m rigid box (12
                  triangles), pose jitter +-1 mm and +-2 deg.
  tactile_pad     soft ellipsoid 40/30/20 mm (L4, K = 4: 51,200 tets,
                  E = 5e4 Pa) pressed 3 mm into an 80 mm pad meshed 64 x 64
                  (8,192 triangles) under a random tilt <= 10 deg; 16 x 16
                  taxels, 20 x 20 rays each.
  batched         N environments, soft tet count log-uniform in [2k, 100k]
                  (sphere, ellipsoid or capsule), rigid triangle count
                  log-uniform in [12, 20k] (box, plane grid or icosphere),
                  penetration uniform in [0.5, 8] mm.
  differentiation N environments of the tactile_pad generator with <= 20k
                  tets and <= 2k triangles.
  soft_soft       N pairs of soft spheres/ellipsoids, 20k-200k tets each,
                  E log-uniform in [1e4, 1e6], overlap uniform in [2, 10] mm.
"""
import math

import numpy as np

from . import _abi as A

BASE_SEED = 20261202
MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, *keys):
        s = 0
        for k in keys:
            s = self._mix((s ^ (int(k) & MASK)) + 0x9E3779B97F4A7C15)
        self.state = s

    @staticmethod
    def _mix(z):
        z &= MASK
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        return self._mix(self.state)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def loguniform(self, lo, hi):
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def choice(self, n):
        return int(self.next() % n)


def axis_angle(axis, angle):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * (K @ K)


def random_axis(rng):
    while True:
        v = np.array([rng.uniform(-1, 1) for _ in range(3)])
        n = np.linalg.norm(v)
        if 1e-3 < n <= 1:
            return v / n


def random_rotation(rng):
    """Uniform rotation from a random unit quaternion."""
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def pose(R, t):
    X = np.zeros(12)
    X[:9] = np.asarray(R, dtype=np.float64).reshape(9)
    X[9:] = t
    return X


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with %d" % (what, rc))


def soft_mesh(L, K, shape, a, b, c, E):
    """Icosphere level L with K radial layers; shape 0 sphere, 1 ellipsoid, 2 capsule."""
    lib = A.lib()
    nv, nt = A.C.c_int32(), A.C.c_int32()
    _check(lib.hcs_gen_soft_size(L, K, A.C.byref(nv), A.C.byref(nt)), "hcs_gen_soft_size")
    V = np.zeros((nv.value, 3))
    T = np.zeros((nt.value, 4), dtype=np.int32)
    p = np.zeros(nv.value)
    _check(lib.hcs_gen_soft(L, K, shape, a, b, c, E, A.dptr(V), A.i32ptr(T), A.dptr(p)), "hcs_gen_soft")
    return V, T, p


def rigid_mesh(kind, m, a, b=0.0, c=0.0):
    """kind 0 box (half sizes a, b, c), 1 plane grid of side a, 2 icosphere radius a."""
    lib = A.lib()
    nv, nt = A.C.c_int32(), A.C.c_int32()
    _check(lib.hcs_gen_rigid_size(kind, m, A.C.byref(nv), A.C.byref(nt)), "hcs_gen_rigid_size")
    V = np.zeros((nv.value, 3))
    F = np.zeros((nt.value, 3), dtype=np.int32)
    _check(lib.hcs_gen_rigid(kind, m, a, b, c, A.dptr(V), A.i32ptr(F)), "hcs_gen_rigid")
    return V, F


def soft_from_surface(V, F, E):
    lib = A.lib()
    V = A.as_f64(V)
    F = A.as_i32(F)
    nv, nt = V.shape[0], F.shape[0]
    Vo = np.zeros((nv + 1, 3))
    To = np.zeros((nt, 4), dtype=np.int32)
    po = np.zeros(nv + 1)
    _check(lib.hcs_gen_soft_from_surface(A.dptr(V), nv, A.i32ptr(F), nt, E, A.dptr(Vo), A.i32ptr(To), A.dptr(po)),
           "hcs_gen_soft_from_surface")
    return Vo, To, po


def soft_size(L, K):
    lib = A.lib()
    nv, nt = A.C.c_int32(), A.C.c_int32()
    _check(lib.hcs_gen_soft_size(L, K, A.C.byref(nv), A.C.byref(nt)), "hcs_gen_soft_size")
    return nv.value, nt.value


def choose_LK(nt_target, Ls=(2, 3, 4, 5)):
    """Icosphere level and layer count whose tet count is closest to nt_target
    (log error), preferring layer spacing comparable to the surface edge."""
    best = None
    for L in Ls:
        F = 20 * 4 ** L
        for K in range(1, 129):
            nt = F * (1 + 3 * (K - 1))
            err = abs(math.log(nt / nt_target)) + 0.02 * abs(math.log2(K / (0.9 * 2 ** L)))
            if best is None or err < best[0]:
                best = (err, L, K)
    return best[1], best[2]


def twist(rng, w=0.5, v=0.02):
    return np.array([rng.uniform(-w, w) for _ in range(3)] + [rng.uniform(-v, v) for _ in range(3)])


SOFT_PROPS = dict(dissipation=3.0, resolution_hint=0.0, static_friction=0.3, dynamic_friction=0.3)
RIGID_PROPS = dict(hydroelastic_modulus=0.0, dissipation=1.0, resolution_hint=0.0, static_friction=0.3,
                   dynamic_friction=0.3)


def sphere_box(seed=0):
    """Soft sphere on a rigid box, one pair.  Returns a scene dict."""
    rng = SplitMix64(BASE_SEED, 1, seed)
    V, T, p = soft_mesh(3, 2, 0, 0.05, 0.05, 0.05, 1e5)
    VR, FR = rigid_mesh(0, 1, 0.1, 0.1, 0.1)
    R = axis_angle(random_axis(rng), math.radians(rng.uniform(-2, 2)))
    c = np.array([rng.uniform(-1e-3, 1e-3), rng.uniform(-1e-3, 1e-3), 0.045 + rng.uniform(-1e-3, 1e-3)])
    return dict(
        soft=dict(V=V, T=T, p=p, props=dict(hydroelastic_modulus=1e5, **SOFT_PROPS), geom_id=0),
        rigid=dict(V=VR, F=FR, props=dict(RIGID_PROPS), geom_id=1),
        X_WA=pose(R, c), X_WB=pose(np.eye(3), [0, 0, -0.1]), V_WA=twist(rng), V_WB=twist(rng, 0.1, 0.005))


def tactile_pad(seed=0, L=4, K=4, semi=(0.04, 0.03, 0.02), m=64, E=5e4, depth=0.003):
    """Soft ellipsoid on a rigid 80 mm tactile pad (16 x 16 taxels)."""
    rng = SplitMix64(BASE_SEED, 2, seed)
    V, T, p = soft_mesh(L, K, 1, semi[0], semi[1], semi[2], E)
    zs = 0.002
    VR, FR = rigid_mesh(1, m, 0.08)
    VR[:, 2] = zs
    R = axis_angle(random_axis(rng), math.radians(rng.uniform(0, 10)))
    zmin = float((V @ R.T)[:, 2].min())
    c = np.array([rng.uniform(-2e-3, 2e-3), rng.uniform(-2e-3, 2e-3), zs - zmin - depth])
    grid = dict(sampling_resolution=20, window=0, sigma=0.0, half_size=(0.04, 0.04, zs), resolution=0.005)
    return dict(
        soft=dict(V=V, T=T, p=p, props=dict(hydroelastic_modulus=E, **SOFT_PROPS), geom_id=0),
        rigid=dict(V=VR, F=FR, props=dict(RIGID_PROPS), geom_id=1),
        X_WA=pose(R, c), X_WB=pose(np.eye(3), [0, 0, 0]), V_WA=twist(rng), V_WB=np.zeros(6), grid=grid)


def _rigid_for(rng, ntri_target):
    kind = rng.choice(3)
    if kind == 0:
        m = max(1, int(round(math.sqrt(ntri_target / 12.0))))
        VR, FR = rigid_mesh(0, m, 0.1, 0.1, 0.1)
        XR = pose(np.eye(3), [0, 0, -0.1])
        top = 0.0
    elif kind == 1:
        m = max(1, int(round(math.sqrt(ntri_target / 2.0))))
        VR, FR = rigid_mesh(1, m, 0.2)
        XR = pose(np.eye(3), [0, 0, 0])
        top = 0.0
    else:
        L = int(min(5, max(0, round(math.log(ntri_target / 20.0, 4)))))
        r = rng.uniform(0.05, 0.1)
        VR, FR = rigid_mesh(2, L, r)
        XR = pose(np.eye(3), [0, 0, -r])
        top = 0.0
    return kind, VR, FR, XR, top


def batched_env(seed, env, nt_range=(2e3, 1e5), ntri_range=(12, 2e4)):
    """One environment of the batched workload."""
    rng = SplitMix64(BASE_SEED, 3, seed, env)
    nt_target = rng.loguniform(*nt_range)
    ntri_target = rng.loguniform(*ntri_range)
    shape = rng.choice(3)
    L, K = choose_LK(nt_target)
    E = 1e5
    if shape == 0:
        r = rng.uniform(0.03, 0.06)
        V, T, p = soft_mesh(L, K, 0, r, r, r, E)
    elif shape == 1:
        V, T, p = soft_mesh(L, K, 1, rng.uniform(0.02, 0.06), rng.uniform(0.02, 0.06), rng.uniform(0.02, 0.06), E)
    else:
        V, T, p = soft_mesh(L, K, 2, rng.uniform(0.02, 0.04), rng.uniform(0.01, 0.04), 0.0, E)
    kind, VR, FR, XR, top = _rigid_for(rng, ntri_target)
    R = random_rotation(rng)
    pen = rng.uniform(0.5e-3, 8e-3)
    zmin = float((V @ R.T)[:, 2].min())
    off = 0.0 if kind == 2 else 0.02
    c = np.array([rng.uniform(-off, off), rng.uniform(-off, off), top - zmin - pen])
    return dict(
        soft=dict(V=V, T=T, p=p, props=dict(hydroelastic_modulus=E, **SOFT_PROPS), geom_id=2 * env),
        rigid=dict(V=VR, F=FR, props=dict(RIGID_PROPS), geom_id=2 * env + 1),
        X_WA=pose(R, c), X_WB=XR, V_WA=twist(rng), V_WB=twist(rng, 0.1, 0.005), shape=shape, rigid_kind=kind)


def differentiation_env(seed, env):
    rng = SplitMix64(BASE_SEED, 4, seed, env)
    K = 2 + rng.choice(3)
    m = 16 + rng.choice(16)
    return tactile_pad(seed=(seed << 20) + env, L=3, K=K, m=m)


def soft_soft_env(seed, env, nt_range=(2e4, 2e5)):
    rng = SplitMix64(BASE_SEED, 5, seed, env)
    bodies = []
    for k in range(2):
        nt_target = rng.loguniform(*nt_range)
        L, K = choose_LK(nt_target)
        E = rng.loguniform(1e4, 1e6)
        if rng.choice(2) == 0:
            r = rng.uniform(0.03, 0.06)
            V, T, p = soft_mesh(L, K, 0, r, r, r, E)
        else:
            V, T, p = soft_mesh(L, K, 1, rng.uniform(0.03, 0.06), rng.uniform(0.03, 0.06), rng.uniform(0.03, 0.06), E)
        bodies.append((V, T, p, E))
    RA, RB = random_rotation(rng), random_rotation(rng)
    ov = rng.uniform(2e-3, 10e-3)
    VA = bodies[0][0] @ RA.T
    VB = bodies[1][0] @ RB.T
    # A below B along z: B's lowest point ov below A's highest point
    zA = float(VA[:, 2].max())
    zB = float(VB[:, 2].min())
    cB = np.array([rng.uniform(-2e-3, 2e-3), rng.uniform(-2e-3, 2e-3), zA - zB - ov])
    mk = lambda b, gid: dict(V=b[0], T=b[1], p=b[2], props=dict(hydroelastic_modulus=b[3], **SOFT_PROPS),
                             geom_id=gid)
    return dict(soft=mk(bodies[0], 2 * env), soft_b=mk(bodies[1], 2 * env + 1),
                X_WA=pose(RA, [0, 0, 0]), X_WB=pose(RB, cB), V_WA=twist(rng), V_WB=twist(rng))


def soft_plane_env(seed, env, nt_range=(2e3, 3e4), geom_order=0):
    """A soft body resting on a MuJoCo ground plane (rigid half-space z <= 0
    of a tilted frame): the soft-plane branch of collision_cb."""
    rng = SplitMix64(BASE_SEED, 6, seed, env)
    nt_target = rng.loguniform(*nt_range)
    L, K = choose_LK(nt_target)
    E = rng.loguniform(3e4, 3e5)
    if rng.choice(2) == 0:
        r = rng.uniform(0.03, 0.06)
        V, T, p = soft_mesh(L, K, 0, r, r, r, E)
    else:
        V, T, p = soft_mesh(L, K, 1, rng.uniform(0.02, 0.06), rng.uniform(0.02, 0.06), rng.uniform(0.02, 0.06), E)
    R = random_rotation(rng)
    RP = axis_angle(random_axis(rng), math.radians(rng.uniform(0, 20)))  # the plane's frame
    n = RP[:, 2]
    pen = rng.uniform(0.5e-3, 6e-3)
    cP = np.array([rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1), rng.uniform(-0.05, 0.05)])
    VW = V @ R.T
    lowest = float((VW @ n).min())
    c = cP - (lowest + pen) * n
    gid_s, gid_p = (2 * env, 2 * env + 1) if geom_order == 0 else (2 * env + 1, 2 * env)
    return dict(soft=dict(V=V, T=T, p=p, props=dict(hydroelastic_modulus=E, **SOFT_PROPS), geom_id=gid_s),
                plane=dict(props=dict(RIGID_PROPS), geom_id=gid_p),
                X_WA=pose(R, c), X_WB=pose(RP, cP), V_WA=twist(rng), V_WB=twist(rng, 0.1, 0.005))


def upload_batched(engine, n_env, seed=0, log=None, every=512):
    """Builds the batched workload (batched_env for env 0..n_env-1) and
    uploads every mesh.  Returns (scenes, handles (n, 2) int32, state (n, 36):
    X_WA, X_WB, V_WA, V_WB per env)."""
    scs, handles = [], []
    for e in range(n_env):
        sc = batched_env(seed, e)
        s, r = sc["soft"], sc["rigid"]
        ha = engine.add_soft_mesh(s["V"], s["T"], s["p"], s["props"], s["geom_id"])
        hb = engine.add_rigid_mesh(r["V"], r["F"], r["props"], r["geom_id"])
        scs.append(sc)
        handles.append((ha, hb))
        if log and (e + 1) % every == 0:
            log("uploaded %d/%d envs" % (e + 1, n_env))
    state = np.array([np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]]) for sc in scs])
    return scs, np.array(handles, dtype=np.int32).reshape(-1, 2), state


def upload_soft_soft(engine, n_pairs, seed=0, nt_range=(2e4, 2e5), log=None, every=32):
    """Builds soft_soft_env for pairs 0..n_pairs-1 and uploads both bodies.
    Returns (scenes, handles, state) as upload_batched."""
    scs, handles = [], []
    for e in range(n_pairs):
        sc = soft_soft_env(seed, e, nt_range=nt_range)
        a, b = sc["soft"], sc["soft_b"]
        ha = engine.add_soft_mesh(a["V"], a["T"], a["p"], a["props"], a["geom_id"])
        hb = engine.add_soft_mesh(b["V"], b["T"], b["p"], b["props"], b["geom_id"])
        scs.append(sc)
        handles.append((ha, hb))
        if log and (e + 1) % every == 0:
            log("uploaded %d/%d pairs" % (e + 1, n_pairs))
    state = np.array([np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]]) for sc in scs])
    return scs, np.array(handles, dtype=np.int32).reshape(-1, 2), state


GEOM = dict(sphere=2, ellipsoid=4, cylinder=5, box=6)


def geom_mesh(kind, size, resolution_hint, E=None):
    """Mesh of a MuJoCo geom as the plugin builds it for a cs:: geom
    (hcs_gen_geom): kind 'sphere' | 'ellipsoid' | 'cylinder' | 'box', size =
    geom_size; with E a soft tet mesh (V, T, p), else the rigid surface
    (V, F)."""
    lib = A.lib()
    t = GEOM[kind]
    s = np.zeros(3)
    s[:len(size)] = size
    soft = E is not None
    nv, ne = A.C.c_int32(), A.C.c_int32()
    _check(lib.hcs_gen_geom_size(t, int(soft), A.dptr(s), float(resolution_hint), A.C.byref(nv), A.C.byref(ne)),
           "hcs_gen_geom_size")
    V = np.zeros((nv.value, 3))
    El = np.zeros((ne.value, 4 if soft else 3), dtype=np.int32)
    p = np.zeros(nv.value)
    _check(lib.hcs_gen_geom(t, int(soft), A.dptr(s), float(resolution_hint), float(E if soft else 0.0), A.dptr(V),
                            A.i32ptr(El), A.dptr(p)), "hcs_gen_geom")
    return (V, El, p) if soft else (V, El)
