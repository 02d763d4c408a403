"""Meshes of the cs:: geom types (hcs_gen_geom; the plugin's Drake
generator calls at mujoco_contact_surfaces/src/mujoco_contact_surfaces_plugin.cpp:
634-742 and 828-915), checked against what the geometry fixes.  CPU only.

Soft meshes: positive tets, conforming (every interior face in exactly two
tets, oppositely oriented), boundary faces on the shape's surface carrying
pressure 0, total volume equal to the polyhedron's (exact for boxes and
prisms), pressure E times the distance to the boundary over its largest
value.  The box around its medial axis carries that field exactly in every
tet, which gives closed forms for the pressure at any point and for the
force of a flat rigid indenter.  Rigid surfaces: closed, consistently
oriented outward, enclosing the same volume.
"""
import math
from collections import Counter

import numpy as np
import pytest

import oracle as O
from meshes import pose
from mujoco_contact_surfaces_amd import scenes as S

E = 1e5


def _vol6(V, T):
    a, b, c, d = (V[T[:, i]] for i in range(4))
    return np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))


def _faces(T):
    # faces of each tet, oriented outward for a positive tet (ccw seen from outside)
    return np.concatenate([T[:, [1, 2, 3]], T[:, [0, 3, 2]], T[:, [0, 1, 3]], T[:, [0, 2, 1]]])


def _boundary(T):
    F = _faces(T)
    keys = Counter(tuple(sorted(f)) for f in F)
    assert max(keys.values()) <= 2, "a face in more than two tets"
    # interior faces appear once in each orientation
    directed = Counter(tuple(f[np.argsort(f)]) + (_parity(f),) for f in F)
    for k, n in keys.items():
        if n == 2:
            assert directed[k + (0,)] == 1 and directed[k + (1,)] == 1, "interior face not oppositely oriented"
    return np.array([f for f in F if keys[tuple(sorted(f))] == 1])


def _parity(f):
    # parity of the permutation sorting f
    p = list(np.argsort(f))
    s = 0
    for i in range(len(p)):
        for j in range(i + 1, len(p)):
            s ^= p[i] > p[j]
    return int(s)


def _surface_volume(V, F):
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0


def _check_soft(V, T, p, analytic_volume=None, rtol=1e-12):
    v6 = _vol6(V, T)
    assert v6.min() > 0
    B = _boundary(T)
    assert B.shape[0] > 0
    # the boundary is closed and encloses the mesh's volume
    assert _surface_volume(V, B) == pytest.approx(v6.sum() / 6, rel=1e-12)
    if analytic_volume is not None:
        assert v6.sum() / 6 == pytest.approx(analytic_volume, rel=rtol)
    bv = np.unique(B)
    np.testing.assert_array_equal(p[bv], 0.0)
    assert p.min() >= 0 and p.max() == pytest.approx(E, rel=1e-15)
    interior = np.setdiff1d(np.arange(V.shape[0]), bv)
    assert (p[interior] > 0).all()
    return B


def _check_surface(V, F, analytic_volume=None, rtol=1e-12):
    # vertices merged by position (a box surface repeats its edge vertices per face)
    V, canon = np.unique(V, axis=0, return_inverse=True)
    F = canon.reshape(-1)[F]
    directed = Counter((int(F[i, k]), int(F[i, (k + 1) % 3])) for i in range(F.shape[0]) for k in range(3))
    for (a, b), n in directed.items():
        assert n == 1 and directed[(b, a)] == 1, "surface not closed and consistently oriented"
    vol = _surface_volume(V, F)
    assert vol > 0
    if analytic_volume is not None:
        assert vol == pytest.approx(analytic_volume, rel=rtol)


@pytest.mark.parametrize("r,hint", [(0.05, 0.05), (0.05, 0.02), (0.05, 0.006)])
def test_sphere_single_interior_vertex(r, hint):
    V, T, p = S.geom_mesh("sphere", [r], hint, E)
    B = _check_soft(V, T, p)
    # one interior vertex, at the centre, at pressure E; every tet uses it
    assert np.count_nonzero(p) == 1 and np.all(V[p > 0] == 0)
    assert (T == np.argmax(p)).any(axis=1).all()
    np.testing.assert_allclose(np.linalg.norm(V[np.unique(B)], axis=1), r, rtol=1e-15)
    # octahedron refinement: 8 * 4^L boundary triangles, edges within the hint
    L = round(math.log(B.shape[0] / 8, 4))
    assert B.shape[0] == 8 * 4 ** L
    e = np.linalg.norm(V[B[:, 0]] - V[B[:, 1]], axis=1)
    assert e.max() <= hint * 1.0000001 or L == 7
    Vs, Fs = S.geom_mesh("sphere", [r], hint)
    _check_surface(Vs, Fs, v6_sum(V, T))


def v6_sum(V, T):
    return _vol6(V, T).sum() / 6


def test_ellipsoid_is_the_scaled_sphere():
    a, b, c = 0.04, 0.03, 0.02
    V, T, p = S.geom_mesh("ellipsoid", [a, b, c], 0.01, E)
    _check_soft(V, T, p)
    B = np.unique(_boundary(T))
    np.testing.assert_allclose(((V[B] / [a, b, c]) ** 2).sum(axis=1), 1.0, rtol=1e-14)
    Vs, T1, _ = S.geom_mesh("sphere", [1.0], 0.01 / 0.04, E)
    np.testing.assert_array_equal(T1, T)
    np.testing.assert_allclose(V, Vs * [a, b, c], rtol=1e-15, atol=1e-18)


def _phi_box(x, h):
    return np.min(h - np.abs(x), axis=-1)


@pytest.mark.parametrize("h", [(0.1, 0.05, 0.02), (0.03, 0.03, 0.03), (0.04, 0.04, 0.01), (0.02, 0.05, 0.05)])
def test_box_medial_axis_field_is_exact(h):
    h = np.array(h)
    V, T, p = S.geom_mesh("box", h, 0.0, E)
    B = _check_soft(V, T, p, analytic_volume=8 * h.prod())
    np.testing.assert_allclose(np.abs(V[np.unique(B)]).max(axis=0), h, rtol=0)
    m = h.min()
    np.testing.assert_allclose(p, E * _phi_box(V, h) / m, rtol=0, atol=1e-9 * E)
    # the interpolated field at random points equals E phi / min(h)
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, (400, 3)) * h
    a = V[T[:, 0]]
    M = np.stack([V[T[:, k]] - a for k in (1, 2, 3)], axis=2)
    Minv = np.linalg.inv(M)
    for x in X:
        lam = np.einsum("tij,tj->ti", Minv, x - a)
        bary = np.concatenate([1 - lam.sum(axis=1, keepdims=True), lam], axis=1)
        t = np.argmax(bary.min(axis=1))
        assert bary[t].min() > -1e-9
        val = bary[t] @ p[T[t]]
        assert val == pytest.approx(E * _phi_box(x, h) / m, rel=0, abs=1e-8 * E)


def test_box_grid_within_hint():
    h = np.array([0.1, 0.05, 0.02])
    V, T, p = S.geom_mesh("box", h, 0.01, E)
    _check_soft(V, T, p, analytic_volume=8 * h.prod())
    np.testing.assert_allclose(p, E * _phi_box(V, h) / h.min(), rtol=0, atol=1e-9 * E)
    for k in range(3):
        assert np.diff(np.unique(V[:, k])).max() <= 0.01 + 1e-15


def _phi_cyl(x, r, hl):
    rho = np.hypot(x[..., 0], x[..., 1])
    return np.minimum(r - rho, hl - np.abs(x[..., 2]))


@pytest.mark.parametrize("r,hl", [(0.03, 0.08), (0.05, 0.02), (0.04, 0.04)])
def test_cylinder_medial_axis(r, hl):
    V, T, p = S.geom_mesh("cylinder", [r, hl], 0.01, E)
    n = max(3, math.ceil(2 * math.pi * r / 0.01))
    prism = n / 2 * r * r * math.sin(2 * math.pi / n) * 2 * hl
    B = _check_soft(V, T, p, analytic_volume=prism)
    bv = V[np.unique(B)]
    on_side = np.isclose(np.hypot(bv[:, 0], bv[:, 1]), r, rtol=1e-12)
    on_cap = np.isclose(np.abs(bv[:, 2]), hl, rtol=1e-12)
    assert (on_side | on_cap).all()
    np.testing.assert_allclose(p, E * np.maximum(_phi_cyl(V, r, hl), 0) / min(r, hl), rtol=0, atol=1e-9 * E)
    Vs, Fs = S.geom_mesh("cylinder", [r, hl], 0.01)
    _check_surface(Vs, Fs, prism)


def test_rigid_box_surface():
    h = np.array([0.1, 0.05, 0.02])
    Vs, Fs = S.geom_mesh("box", h, 0.05)
    _check_surface(Vs, Fs, 8 * h.prod())


def test_unmeshed_types_and_soft_cylinder_without_hint_are_rejected():
    lib = S.A.lib()
    s = np.array([0.05, 0.1, 0.0])
    nv, ne = S.A.C.c_int32(), S.A.C.c_int32()
    for t in (0, 1, 3):  # plane, height field, capsule (plugin.cpp:634-647)
        assert lib.hcs_gen_geom_size(t, 1, S.A.dptr(s), 0.01, S.A.C.byref(nv), S.A.C.byref(ne)) == S.A.HCS_ERR_ARG
    assert lib.hcs_gen_geom_size(5, 1, S.A.dptr(s), 0.0, S.A.C.byref(nv), S.A.C.byref(ne)) == S.A.HCS_ERR_ARG


@pytest.mark.parametrize("depth", [0.004, 0.011])
def test_flat_indenter_on_a_soft_box_closed_form(depth):
    """A rigid flat grid pressed `depth` into the top of a soft box meshed
    around its medial axis.  On the section plane the field is
    E min(depth, hx - |x|, hy - |y|) / min(h), so over the whole section
    (2 hx x 2 hy) the polygons integrate to
    4 E / min(h) [d (hx - d)(hy - d) + d^2 (hx + hy - 2 d) / 2 + d^3 / 3];
    the top frustum's faces carry the uniform pressure E d / min(h) over the
    inner rectangle 2 (hx - d) x 2 (hy - d) with the full stiffness, while
    in the side frusta the gradient lies in the plane and g vanishes (to
    rounding), which the face filter of plugin.cpp:361-368 tests."""
    h = np.array([0.05, 0.04, 0.02])
    a, b, d, m = h[0], h[1], depth, h.min()
    V, T, p = S.geom_mesh("box", h, 0.0, E)
    z = h[2] - depth
    g = np.array([[x, y, z] for y in (-0.2, 0.0, 0.2) for x in (-0.2, 0.0, 0.2)])
    F = []
    for j in range(2):
        for i in range(2):
            q = j * 3 + i
            F += [[q, q + 4, q + 1], [q, q + 3, q + 4]]  # normal -z: out of the indenter, into the box
    res = O.soft_rigid(V, T, p, g, np.array(F, np.int32), pose(), pose(), np.zeros(6), np.zeros(6), E, 0.0, 0.3, 0.3,
                       representation=1)
    P = res["polys"]
    areas = []
    for rec in P:
        n = int(rec[0])
        v = rec[1:1 + 3 * n].reshape(n, 3)
        np.testing.assert_allclose(v[:, 2], z, rtol=0, atol=1e-15)
        areas.append(0.5 * abs(sum(np.cross(v[k] - v[0], v[k + 1] - v[0])[2] for k in range(1, n - 1))))
    areas = np.array(areas)
    assert areas.sum() == pytest.approx(4 * a * b, rel=1e-12)
    integral = 4 * E / m * (d * (a - d) * (b - d) + d * d * (a + b - 2 * d) / 2 + d ** 3 / 3)
    assert (areas * P[:, 36]).sum() == pytest.approx(integral, rel=1e-11)
    f = res["faces"]
    inner = np.abs(f[:, 10] - E * d / m) < 1e-9 * E * d / m
    assert f[inner, 9].sum() == pytest.approx(4 * (a - d) * (b - d), rel=1e-12)
    # stiffness g = grad p . n (the rigid side adds nothing): E / m in the
    # top frustum, rounding noise in the side frusta
    np.testing.assert_allclose(f[inner, 11], E / m, rtol=1e-10)
    assert np.all(f[~inner, 11] < 1e-6 * E / m)
