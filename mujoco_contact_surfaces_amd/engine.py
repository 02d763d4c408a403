"""Host-side mirror of the reference plugin's per-step interface over libhcs.

The reference (ubi-agni/mujoco_contact_surfaces) runs, per MuJoCo step:
  collision_cb(m, d, con, g1, g2, margin)       mujoco_contact_surfaces_plugin.cpp:255-318
    -> Drake contact-surface query + evaluateContactSurface (:320-409)
  passiveCallback(m, d)                          :411-523
    -> Hunt-Crossley + regularised friction per PointCollision, mj_applyFT
  SurfacePlugin::update(m, d, geomCollisions)     plugin_utils.h:94
    -> FlatTactileSensor::bvh_update              flat_tactile_sensor.cpp:262-402
ContactSurfacesPlugin below keeps that shape (collision_cb queues a pair,
passive_callback runs the whole step on the GPU and returns one wrench per
geometry, sensors sample the kept contact polygons); Engine is the direct
C-ABI wrapper.  Everything runs through libhcs.so; there is no CPU path.
"""
import atexit
import ctypes as C
import weakref

import numpy as np

from . import _abi as A

KTRIANGLE = 0
KPOLYGON = 1
WINDOWS = {"none": 0, "gauss": 1, "tukey": 2, "square": 3}


# Engines still open when the interpreter exits are closed (batches first,
# then the context) by an exit handler, while every module is intact, rather
# than by garbage collection in whatever order interpreter teardown reaches
# them.
_live_engines = weakref.WeakSet()


@atexit.register
def _close_live_engines():
    for e in list(_live_engines):
        try:
            e.close()
        except Exception:
            pass


class HcsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hcs error %d: %s" % (code, msg))
        self.code = code


def _props(d):
    p = A.Props()
    p.hydroelastic_modulus = d.get("hydroelastic_modulus", 0.0)
    p.dissipation = d.get("dissipation", 1.0)
    p.resolution_hint = d.get("resolution_hint", 0.0)
    p.static_friction = d.get("static_friction", 0.0)
    p.dynamic_friction = d.get("dynamic_friction", 0.0)
    return p


def props_from_cs(values):
    """The cs::<geom> 5-tuple (E, dissipation, resolution hint, mu_static,
    mu_dynamic) as parsed at mujoco_contact_surfaces_plugin.cpp:613-626."""
    E, d, hint, ms, md = [float(v) for v in values]
    return dict(hydroelastic_modulus=E, dissipation=d, resolution_hint=hint, static_friction=ms, dynamic_friction=md)


def _opts(representation=KTRIANGLE, apply_forces=True, dt=0.001, keep_polygons=False):
    o = A.StepOpts()
    o.representation = int(representation)
    o.apply_forces = int(bool(apply_forces))
    o.dt = float(dt)
    o.keep_polygons = int(bool(keep_polygons))
    return o


def _pair(ha, hb, X_WA, X_WB, V_WA=None, V_WB=None):
    p = A.PairIn()
    p.geom_a, p.geom_b = int(ha), int(hb)
    p.X_WA[:] = [float(v) for v in np.asarray(X_WA, dtype=np.float64).reshape(12)]
    p.X_WB[:] = [float(v) for v in np.asarray(X_WB, dtype=np.float64).reshape(12)]
    p.V_WA[:] = [float(v) for v in (np.zeros(6) if V_WA is None else np.asarray(V_WA, dtype=np.float64).reshape(6))]
    p.V_WB[:] = [float(v) for v in (np.zeros(6) if V_WB is None else np.asarray(V_WB, dtype=np.float64).reshape(6))]
    return p


def _grid(geom, X_W_sensor, half_size, resolution, sampling_resolution=20, window="none", sigma=0.0):
    g = A.TaxelGrid()
    g.geom = int(geom)
    g.sampling_resolution = int(sampling_resolution)
    g.window = WINDOWS[window] if isinstance(window, str) else int(window)
    g.sigma = float(sigma)
    g.X_W_sensor[:] = [float(v) for v in np.asarray(X_W_sensor, dtype=np.float64).reshape(12)]
    g.half_size[:] = [float(v) for v in half_size]
    g.resolution = float(resolution)
    return g


class Engine:
    """One libhcs context on one GPU."""

    def __init__(self, device=0):
        self.lib = A.lib()
        h = C.c_void_p()
        rc = self.lib.hcs_create(C.byref(h), int(device))
        if rc != 0:
            raise HcsError(rc, self.lib.hcs_last_error(None).decode())
        self.h = h
        self._kind = {}
        self._batches = weakref.WeakSet()
        _live_engines.add(self)

    def _chk(self, rc):
        if rc != 0:
            raise HcsError(rc, self.lib.hcs_last_error(self.h).decode())

    def close(self):
        if self.h:
            # a batch must not outlive its context: release them first
            for b in list(self._batches):
                b.close()
            self.lib.hcs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self._chk(self.lib.hcs_synchronize(self.h))

    def stream(self):
        return self.lib.hcs_stream(self.h)

    # ---- geometry
    def add_soft_mesh(self, V, T, p, props, geom_id):
        V, T, p = A.as_f64(V, (-1, 3)), A.as_i32(T, (-1, 4)), A.as_f64(p)
        h = C.c_int32()
        pr = _props(props)
        self._chk(self.lib.hcs_add_soft_mesh(self.h, A.dptr(V), V.shape[0], A.i32ptr(T), T.shape[0], A.dptr(p),
                                             C.byref(pr), int(geom_id), C.byref(h)))
        self._kind[h.value] = "soft"
        return h.value

    def add_rigid_mesh(self, V, F, props, geom_id):
        V, F = A.as_f64(V, (-1, 3)), A.as_i32(F, (-1, 3))
        h = C.c_int32()
        pr = _props(props)
        self._chk(self.lib.hcs_add_rigid_mesh(self.h, A.dptr(V), V.shape[0], A.i32ptr(F), F.shape[0], C.byref(pr),
                                              int(geom_id), C.byref(h)))
        self._kind[h.value] = "rigid"
        return h.value

    def replace_soft_mesh(self, handle, V, T, p):
        """onGeomChanged for a soft geom: new tets and pressures, same
        properties and id (batches stepping the handle are re-planned)."""
        V, T, p = A.as_f64(V, (-1, 3)), A.as_i32(T, (-1, 4)), A.as_f64(p)
        self._chk(self.lib.hcs_replace_soft_mesh(self.h, int(handle), A.dptr(V), V.shape[0], A.i32ptr(T), T.shape[0],
                                                 A.dptr(p)))

    def replace_rigid_mesh(self, handle, V, F):
        """onGeomChanged for a rigid mesh geom."""
        V, F = A.as_f64(V, (-1, 3)), A.as_i32(F, (-1, 3))
        self._chk(self.lib.hcs_replace_rigid_mesh(self.h, int(handle), A.dptr(V), V.shape[0], A.i32ptr(F),
                                                  F.shape[0]))

    def add_rigid_plane(self, props, geom_id):
        """A MuJoCo plane geom: the rigid half-space z <= 0 of its frame."""
        h = C.c_int32()
        pr = _props(props)
        self._chk(self.lib.hcs_add_rigid_plane(self.h, C.byref(pr), int(geom_id), C.byref(h)))
        self._kind[h.value] = "rigid"
        return h.value

    def mesh_info(self, handle):
        info = np.zeros(4, dtype=np.int64)
        self._chk(self.lib.hcs_mesh_info(self.h, int(handle), A.i64ptr(info)))
        return dict(kind=int(info[0]), nv=int(info[1]), nelem=int(info[2]), bvh_height=int(info[3]))

    def tet_records(self, handle):
        n = self.mesh_info(handle)["nelem"]
        out = np.zeros((n, 24))
        self._chk(self.lib.hcs_mesh_tet_records(self.h, int(handle), A.dptr(out)))
        return out

    # ---- per-step contact
    def contact_step(self, pairs, representation=KTRIANGLE, apply_forces=True, dt=0.001, keep_polygons=False,
                     face_capacity=0):
        """pairs: list of (handle_a, handle_b, X_WA, X_WB, V_WA, V_WB).
        Returns wrench (n, 12) = F_A, tau_A, F_B, tau_B; counts (n, 3) =
        candidates, polygons, faces; and (rec, key) of the faces when
        face_capacity > 0."""
        n = len(pairs)
        arr = (A.PairIn * n)(*[_pair(*p) for p in pairs])
        out = (A.PairOut * n)()
        o = _opts(representation, apply_forces, dt, keep_polygons)
        faces = None
        fptr = None
        if face_capacity > 0:
            rec = np.zeros((face_capacity, 12))
            key = np.zeros((face_capacity, 4), dtype=np.int64)
            faces = A.Faces(face_capacity, 0, A.dptr(rec), A.i64ptr(key))
            fptr = C.byref(faces)
        rc = self.lib.hcs_contact_step(self.h, n, arr, C.byref(o), out, fptr)
        if rc == A.HCS_ERR_CAPACITY and faces is not None and faces.count > face_capacity:
            return self.contact_step(pairs, representation, apply_forces, dt, keep_polygons, int(faces.count))
        self._chk(rc)
        W = np.array([list(r.F_A) + list(r.tau_A) + list(r.F_B) + list(r.tau_B) for r in out]).reshape(n, 12)
        cnt = np.array([[r.n_candidates, r.n_polygons, r.n_faces] for r in out], dtype=np.int64).reshape(n, 3)
        if faces is not None:
            m = int(faces.count)
            return W, cnt, (rec[:m].copy(), key[:m].copy())
        return W, cnt

    def candidates(self, pair):
        p = _pair(*pair)
        cnt = C.c_int64()
        self._chk(self.lib.hcs_candidates(self.h, C.byref(p), None, 0, C.byref(cnt)))
        keys = np.zeros(max(1, cnt.value), dtype=np.int64)
        self._chk(self.lib.hcs_candidates(self.h, C.byref(p), A.i64ptr(keys), keys.size, C.byref(cnt)))
        return keys[:cnt.value]

    def taxels(self, grids):
        """grids: list of dicts (geom, X_W_sensor, half_size, resolution,
        sampling_resolution, window, sigma).  Needs a preceding contact_step
        with keep_polygons=True."""
        gs = [_grid(**g) for g in grids]
        dims = []
        for g in gs:
            cx, cy = C.c_int32(), C.c_int32()
            self._chk(self.lib.hcs_taxel_dims(C.byref(g), C.byref(cx), C.byref(cy)))
            dims.append((cx.value, cy.value))
        out = np.zeros(sum(a * b for a, b in dims))
        arr = (A.TaxelGrid * len(gs))(*gs)
        self._chk(self.lib.hcs_taxels(self.h, len(gs), arr, A.dptr(out)))
        res, o = [], 0
        for cx, cy in dims:
            res.append(out[o:o + cx * cy].reshape(cy, cx).copy())
            o += cx * cy
        return res

    def taxel_sensor(self, geom, X_W_sensor, taxels, include_margin, sample_resolution, method="squared",
                     sample_method="default", visualize=False, values=None):
        """TaxelSensor::internal_update (S/src/taxel_sensor.cpp:158-478) on
        the polygons of the preceding contact_step(keep_polygons=True).
        Returns (values, number of samples); values no sample reaches keep
        their previous value (default 0)."""
        tx = A.as_f64(taxels, (-1, 3))
        vals = np.zeros(tx.shape[0]) if values is None else A.as_f64(values).copy()
        s = A.TaxelSensor()
        s.geom, s.ntaxels = int(geom), tx.shape[0]
        s.method = {"closest": 0, "weighted": 1, "mean": 2, "squared": 3}[method]
        s.sample_method = {"default": 0, "area_importance": 1}[sample_method]
        s.visualize = int(bool(visualize))
        s.include_margin, s.sample_resolution = float(include_margin), float(sample_resolution)
        for i, v in enumerate(A.as_f64(X_W_sensor)):
            s.X_W_sensor[i] = v
        s.taxels = A.dptr(tx)
        n = C.c_int64()
        self._chk(self.lib.hcs_taxel_sensor_update(self.h, C.byref(s), A.dptr(vals), C.byref(n)))
        return vals, n.value

    def curved_sensor(self, geom, X_W_sensor, points, normals, taxel_ptr, sample_idx, weights, include_margin):
        """CurvedSensor::internal_update (S/src/curved_sensor.cpp:388-481) on
        the polygons of the preceding contact_step(keep_polygons=True)."""
        P, N = A.as_f64(points, (-1, 3)), A.as_f64(normals, (-1, 3))
        ptr, idx, w = A.as_i32(taxel_ptr), A.as_i32(sample_idx), A.as_f64(weights)
        s = A.CurvedSensor()
        s.geom, s.ntaxels, s.nsamples = int(geom), ptr.shape[0] - 1, P.shape[0]
        s.include_margin = float(include_margin)
        for i, v in enumerate(A.as_f64(X_W_sensor)):
            s.X_W_sensor[i] = v
        s.points, s.normals = A.dptr(P), A.dptr(N)
        s.taxel_ptr, s.sample_idx, s.weights = A.i32ptr(ptr), A.i32ptr(idx), A.dptr(w)
        vals = np.zeros(max(0, ptr.shape[0] - 1))
        self._chk(self.lib.hcs_curved_sensor_update(self.h, C.byref(s), A.dptr(vals)))
        return vals

    # ---- differentiation
    def jacobian(self, pairs, representation=KTRIANGLE, apply_forces=True, dt=0.001):
        n = len(pairs)
        arr = (A.PairIn * n)(*[_pair(*p) for p in pairs])
        o = _opts(representation, apply_forces, dt)
        J = np.zeros((n, 12, 7))
        self._chk(self.lib.hcs_contact_jacobian(self.h, n, arr, C.byref(o), A.dptr(J)))
        return J

    def backward(self, pairs, dL_dout, representation=KTRIANGLE, apply_forces=True, dt=0.001):
        n = len(pairs)
        arr = (A.PairIn * n)(*[_pair(*p) for p in pairs])
        o = _opts(representation, apply_forces, dt)
        g = A.as_f64(dL_dout, (n, 12))
        out = np.zeros((n, 7))
        self._chk(self.lib.hcs_contact_backward(self.h, n, arr, C.byref(o), A.dptr(g), A.dptr(out)))
        return out

    def taxel_jacobian(self, pair, grid, representation=KTRIANGLE):
        p = _pair(*pair)
        g = _grid(**grid)
        cx, cy = C.c_int32(), C.c_int32()
        self._chk(self.lib.hcs_taxel_dims(C.byref(g), C.byref(cx), C.byref(cy)))
        o = _opts(representation)
        J = np.zeros((cx.value * cy.value, 7))
        self._chk(self.lib.hcs_taxel_jacobian(self.h, C.byref(p), C.byref(o), C.byref(g), A.dptr(J)))
        return J

    # ---- batched
    def batch(self, geoms, representation=KTRIANGLE, apply_forces=True, dt=0.001):
        return Batch(self, geoms, _opts(representation, apply_forces, dt))


class Batch:
    """Many independent pairs stepped by one graph-replayed persistent launch."""

    def __init__(self, engine, geoms, opts):
        self.e = engine
        self.b = None
        engine._batches.add(self)
        g = A.as_i32(geoms, (-1, 2))
        self.n = g.shape[0]
        b = C.c_void_p()
        engine._chk(engine.lib.hcs_batch_create(engine.h, self.n, A.i32ptr(g), C.byref(opts), C.byref(b)))
        self.b = b

    def set_state(self, state):
        s = A.as_f64(state, (self.n, 36))
        self.e._chk(self.e.lib.hcs_batch_set_state(self.b, A.dptr(s)))

    def step(self):
        self.e._chk(self.e.lib.hcs_batch_step(self.b))

    def get(self):
        out = (A.PairOut * self.n)()
        self.e._chk(self.e.lib.hcs_batch_get(self.b, out))
        W = np.array([list(r.F_A) + list(r.tau_A) + list(r.F_B) + list(r.tau_B) for r in out]).reshape(self.n, 12)
        cnt = np.array([[r.n_candidates, r.n_polygons, r.n_faces] for r in out], dtype=np.int64).reshape(self.n, 3)
        return W, cnt

    def tune(self, reps=3):
        ms = C.c_double()
        self.e._chk(self.e.lib.hcs_batch_tune(self.b, int(reps), C.byref(ms)))
        return ms.value

    def config(self):
        cfg = np.zeros(6, dtype=np.int32)
        self.e._chk(self.e.lib.hcs_batch_config(self.b, A.i32ptr(cfg)))
        return dict(blocks_per_cu=int(cfg[0]), queries_per_lane=int(cfg[1]), tet_query_max_tris=int(cfg[2]),
                    split_target=int(cfg[3]), queues=int(cfg[4]), order=int(cfg[5]))

    def set_config(self, blocks_per_cu, queries_per_lane, tet_query_max_tris, split_target=2048, queues=1, order=0):
        cfg = np.array([blocks_per_cu, queries_per_lane, tet_query_max_tris, split_target, queues, order],
                       dtype=np.int32)
        self.e._chk(self.e.lib.hcs_batch_set_config(self.b, A.i32ptr(cfg)))

    def step_times(self, reps=50):
        """Milliseconds of each of reps back-to-back graph replays (HIP
        events between consecutive replays)."""
        ms = np.zeros(int(reps))
        self.e._chk(self.e.lib.hcs_batch_time_steps(self.b, int(reps), A.dptr(ms)))
        return ms

    def kernel_times(self, reps=5):
        ms = np.zeros(3)
        self.e._chk(self.e.lib.hcs_batch_kernel_times(self.b, int(reps), A.dptr(ms)))
        return dict(cand=ms[0], clip=ms[1], rest=ms[2])

    def item_stats(self):
        """Per-item profile of one step: columns pair, kind, queries, cycles,
        production steps, candidates, survivors, elements of B, wall-clock
        start and end."""
        n = C.c_int64()
        self.e._chk(self.e.lib.hcs_batch_item_stats(self.b, None, 0, C.byref(n)))
        out = np.zeros((max(1, n.value), 10), dtype=np.int64)
        self.e._chk(self.e.lib.hcs_batch_item_stats(self.b, A.i64ptr(out), out.shape[0], C.byref(n)))
        return out[:n.value]

    def set_sensors(self, pairs, grids):
        """Flat tactile sensors inside the batch step: sensor s belongs to
        pair pairs[s] and takes the pose of that pair's geometry grids[s]
        ['geom'] from the batch state (X_W_sensor is not needed)."""
        pairs = A.as_i32(pairs, (-1,))
        gs = []
        self._dims = []
        for g in grids:
            g = dict(g)
            g.setdefault("X_W_sensor", np.r_[np.eye(3).reshape(9), 0.0, 0.0, 0.0])
            gg = _grid(**g)
            cx, cy = C.c_int32(), C.c_int32()
            self.e._chk(self.e.lib.hcs_taxel_dims(C.byref(gg), C.byref(cx), C.byref(cy)))
            self._dims.append((cx.value, cy.value))
            gs.append(gg)
        if len(gs) != pairs.shape[0]:
            raise ValueError("one pair per sensor")
        arr = (A.TaxelGrid * max(1, len(gs)))(*gs)
        self.e._chk(self.e.lib.hcs_batch_set_sensors(self.b, len(gs), A.i32ptr(pairs), arr))

    def taxels(self):
        """Taxels of the last step, one (cy, cx) array per sensor."""
        n = C.c_int64()
        self.e._chk(self.e.lib.hcs_batch_taxel_count(self.b, C.byref(n)))
        out = np.zeros(max(1, n.value))
        self.e._chk(self.e.lib.hcs_batch_get_taxels(self.b, A.dptr(out)))
        res, o = [], 0
        for cx, cy in self._dims:
            res.append(out[o:o + cx * cy].reshape(cy, cx).copy())
            o += cx * cy
        return res

    def counts(self):
        c = np.zeros(3, dtype=np.int64)
        self.e._chk(self.e.lib.hcs_batch_counts(self.b, A.i64ptr(c)))
        return c

    def close(self):
        if self.b:
            self.e.lib.hcs_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ContactSurfacesPlugin:
    """Per-step flow of MujocoContactSurfacesPlugin (plugin.cpp:255-523).

    add_geom registers a geometry with its cs:: 5-tuple; registration order
    plays the role of Drake's GeometryId (plugin.h:175,191,205), so the first
    registered of a pair is M.  collision_cb queues a pair (soft first, as
    plugin.cpp:288-295 swaps); passive_callback evaluates every queued pair
    in one GPU step and returns, per pair, the wrench on each geometry about
    its origin, which a MuJoCo host applies with one mj_applyFT per body.
    """

    def __init__(self, engine=None, representation=KTRIANGLE, apply_forces=True):
        self.engine = engine or Engine()
        self.representation = representation
        self.apply_forces = apply_forces
        self.geoms = {}
        self.queue = []
        self.sensors = []

    def add_geom(self, name, cs, V=None, elements=None, pressure=None, plane=False):
        """plane=True registers a MuJoCo plane geom (rigid half-space z <= 0
        of its frame; plugin.cpp:296-299); V and elements are then unused."""
        props = props_from_cs(cs)
        gid = len(self.geoms)
        if plane:
            if props["hydroelastic_modulus"] > 0:
                raise ValueError("a plane geom is rigid only (plugin.cpp:634-647)")
            h = self.engine.add_rigid_plane(props, gid)
            kind = "rigid"
        elif props["hydroelastic_modulus"] > 0:
            h = self.engine.add_soft_mesh(V, elements, pressure, props, gid)
            kind = "soft"
        else:
            h = self.engine.add_rigid_mesh(V, elements, props, gid)
            kind = "rigid"
        self.geoms[name] = dict(handle=h, kind=kind, id=gid)
        return h

    def add_flat_sensor(self, geom_name, half_size, resolution, sampling_resolution=20, window="none", sigma=0.0):
        self.sensors.append(dict(geom=geom_name, half_size=half_size, resolution=resolution,
                                 sampling_resolution=sampling_resolution, window=window, sigma=sigma))

    def collision_cb(self, g1, g2, X1, X2, V1=None, V2=None):
        a, b = self.geoms[g1], self.geoms[g2]
        if a["kind"] == "rigid" and b["kind"] == "rigid":
            return False
        if a["kind"] == "rigid":
            g1, g2, X1, X2, V1, V2 = g2, g1, X2, X1, V2, V1
        self.queue.append((g1, g2, X1, X2, V1, V2))
        return True

    def passive_callback(self, sensor_poses=None):
        if not self.queue:
            return [], []
        pairs = [(self.geoms[g1]["handle"], self.geoms[g2]["handle"], X1, X2, V1, V2)
                 for g1, g2, X1, X2, V1, V2 in self.queue]
        W, cnt = self.engine.contact_step(pairs, self.representation, self.apply_forces,
                                          keep_polygons=bool(self.sensors))
        res = [dict(g1=q[0], g2=q[1], wrench1=W[i, :6], wrench2=W[i, 6:], counts=cnt[i])
               for i, q in enumerate(self.queue)]
        tax = []
        if self.sensors:
            grids = []
            for s, X in zip(self.sensors, sensor_poses or []):
                g = dict(s)
                g["geom"] = self.geoms[s["geom"]]["handle"]
                g["X_W_sensor"] = X
                grids.append(g)
            tax = self.engine.taxels(grids)
        self.queue = []
        return res, tax
