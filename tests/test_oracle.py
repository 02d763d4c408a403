"""Pins the CPU oracle against what the mathematics fixes (the reference has
no tests or golden vectors of its own for this path).  CPU only."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from meshes import cube_grid, pose, rot, signed_volumes

UNIT_TET_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
UNIT_TET_T = np.array([[0, 1, 2, 3]], np.int32)


def unit_rec(p=(0.0, 0.0, 0.0, 0.0)):
    return O.tet_records(UNIT_TET_V, UNIT_TET_T, np.asarray(p, float))[0]


def area_of(poly, n):
    a = 0.0
    for i in range(1, len(poly) - 1):
        a += np.dot(np.cross(poly[i] - poly[0], poly[i + 1] - poly[0]), n)
    return 0.5 * a


# ---------------------------------------------------------------- tet records


def test_tet_record_planes_and_field():
    r = unit_rec((1.0, 2.0, 3.0, 5.0))
    # outward unit normals of faces {1,2,3},{0,3,2},{0,1,3},{0,2,1}
    s3 = 1 / math.sqrt(3)
    np.testing.assert_allclose(r[0:3], [s3, s3, s3], atol=1e-15)
    np.testing.assert_allclose(r[3:6], [-1, 0, 0], atol=0)
    np.testing.assert_allclose(r[6:9], [0, -1, 0], atol=0)
    np.testing.assert_allclose(r[9:12], [0, 0, -1], atol=0)
    np.testing.assert_allclose(r[12:16], [-s3, 0, 0, 0], atol=1e-15)
    # p = 1 + x + 2y + 4z
    np.testing.assert_allclose(r[16:20], [1, 2, 4, 1], atol=1e-14)
    np.testing.assert_allclose(r[23:29], [0, 0, 0, 1, 1, 1])


# ------------------------------------------------------------------ clipping


def test_triangle_inside_tet_is_returned_unchanged():
    tri = np.array([[0.1, 0.1, 0.1], [0.3, 0.1, 0.1], [0.1, 0.3, 0.2]])
    poly = O.clip_tri_tet(tri, unit_rec())
    assert len(poly) == 3
    # same vertex set (cyclic order may start elsewhere)
    for v in tri:
        assert min(np.abs(poly - v).max(axis=1)) == 0.0
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    assert area_of(poly, n / np.linalg.norm(n)) == pytest.approx(0.5 * np.linalg.norm(n), rel=1e-15)


@pytest.mark.parametrize("h", [0.0, 0.1, 0.25, 0.5, 0.9])
def test_plane_section_of_unit_tet(h):
    big = np.array([[-5, -5, h], [10, -5, h], [-5, 10, h]], float)
    poly = O.clip_tri_tet(big, unit_rec())
    assert len(poly) == 3
    assert area_of(poly, np.array([0, 0, 1.0])) == pytest.approx(0.5 * (1 - h) ** 2, rel=1e-13)


def test_slanted_section_is_a_quad_with_closed_form_area():
    # plane x + y = 1/2 cuts the unit tet in a quadrilateral; its area is
    # the projection onto the xz plane divided by cos = 1/sqrt(2):
    # projected region {x in [0,1/2], 0 <= z <= 1/2 - ... } = triangle-trapezoid.
    big = np.array([[0.5 + 5, -5, -5], [0.5 - 5, 5, -5], [0.5, 0.0, 10]], float)
    big[2] = [0.25, 0.25, 10.0]
    poly = O.clip_tri_tet(big, unit_rec())
    assert len(poly) == 4
    n = np.array([1, 1, 0]) / math.sqrt(2)
    # exact: section of x+y=1/2, x,y,z>=0, x+y+z<=1 is {x in [0,1/2], z in [0,1/2]} sheared -> area sqrt(2)/4 * 1/2 * 2
    # computed independently from its 4 corners:
    corners = np.array([[0.5, 0, 0], [0, 0.5, 0], [0, 0.5, 0.5], [0.5, 0, 0.5]])
    ref = abs(area_of(corners, n))
    assert abs(area_of(poly, n)) == pytest.approx(ref, rel=1e-13)


def _exact_clip(tri, tet):
    """Sutherland-Hodgman in exact rationals with unnormalised face normals
    (inside: n.x + d <= 0); same walk order as the engine."""
    faces = [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]
    poly = [tuple(v) for v in tri]

    def sub(a, b):
        return tuple(x - y for x, y in zip(a, b))

    def cross(a, b):
        return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])

    def dot(a, b):
        return sum(x * y for x, y in zip(a, b))

    for f in faces:
        A, B, Cc = tet[f[0]], tet[f[1]], tet[f[2]]
        n = cross(sub(B, A), sub(Cc, A))
        d = -dot(n, A)
        out = []
        m = len(poly)
        for i in range(m):
            cur, prev = poly[i], poly[i - 1]
            sc, sp = dot(n, cur) + d, dot(n, prev) + d
            if (sc <= 0) != (sp <= 0):
                wa = sp / (sp - sc)
                out.append(tuple(wa * c + (1 - wa) * p for c, p in zip(cur, prev)))
            if sc <= 0:
                out.append(cur)
        poly = out
        if not poly:
            break
    # exact duplicates collapse
    res = []
    for v in poly:
        if not res or v != res[-1]:
            res.append(v)
    if len(res) >= 3 and res[-1] == res[0]:
        res.pop()
    return res if len(res) >= 3 else []


@pytest.mark.parametrize("seed", range(40))
def test_clip_matches_exact_rationals(seed):
    rng = np.random.default_rng(seed)
    while True:
        tet = [tuple(Fraction(int(v), 8) for v in rng.integers(-16, 17, 3)) for _ in range(4)]
        tv = np.array([[float(c) for c in v] for v in tet])
        if np.dot(tv[1] - tv[0], np.cross(tv[2] - tv[0], tv[3] - tv[0])) > 0.5:
            break
    tri = [tuple(Fraction(int(v), 8) for v in rng.integers(-16, 17, 3)) for _ in range(3)]
    exact = _exact_clip(tri, tet)
    rec = O.tet_records(tv, UNIT_TET_T, np.zeros(4))[0]
    got = O.clip_tri_tet(np.array([[float(c) for c in v] for v in tri]), rec)
    assert len(got) == len(exact)
    for g, e in zip(got, exact):
        np.testing.assert_allclose(g, [float(c) for c in e], rtol=0, atol=1e-13)


# ------------------------------------------------------- surface invariants


def _grid_scene(n=4):
    # soft cube [-1,1]^3 (m), pressure p = 1000 (z + 1): gradient +z everywhere
    V, T, p = cube_grid(n, field=lambda x: 1000.0 * (x[:, 2] + 1.0))
    return V, T, p


def test_partition_of_unity_triangle_inside_convex_mesh():
    V, T, p = _grid_scene()
    tri = np.array([[-0.7, -0.6, 0.13], [0.55, -0.3, 0.13], [0.1, 0.8, 0.13]])
    F = np.array([[0, 1, 2]], np.int32)
    res = O.soft_rigid(V, T, p, tri, F, pose(), pose(), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3,
                       representation=1)
    area = 0.5 * np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
    assert res["faces"][:, 9].sum() == pytest.approx(area, rel=1e-12)
    # linear field: fn0 = A p(centroid), and p = 1000 (z + 1) on the plane z = 0.13
    np.testing.assert_allclose(res["faces"][:, 10], 1130.0, rtol=1e-12)
    # kTriangle fans partition the same polygons
    res3 = O.soft_rigid(V, T, p, tri, F, pose(), pose(), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3,
                        representation=0)
    assert res3["faces"][:, 9].sum() == pytest.approx(area, rel=1e-12)
    assert res3["counts"][1] == res["counts"][1]


def test_plane_grid_cross_section_area_and_force():
    V, T, p = _grid_scene()
    # rigid 2x2 grid of squares spanning [-3,3]^2 at z = h: section of the cube = 4 m^2
    h = 0.37
    g = np.array([[x, y, h] for y in (-3.0, 0.0, 3.0) for x in (-3.0, 0.0, 3.0)])
    F = []
    for j in range(2):
        for i in range(2):
            a = j * 3 + i
            F += [[a, a + 1, a + 4], [a, a + 4, a + 3]]
    res = O.soft_rigid(V, T, p, g, np.array(F, np.int32), pose(), pose(), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3,
                       0.3, representation=0)
    assert res["faces"][:, 9].sum() == pytest.approx(4.0, rel=1e-12)
    # at rest: net force = integral of p over the section, along +z (into the soft body)
    np.testing.assert_allclose(res["wrench"][:3], [0, 0, 4.0 * 1000 * (h + 1)], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(res["wrench"][6:9], -res["wrench"][:3], rtol=0, atol=0)
    # torque about the soft origin vanishes by symmetry; about the rigid origin too (same origin)
    assert np.abs(res["wrench"][3:6]).max() < 1e-8


def test_gradient_filter_rejects_opposed_faces():
    V, T, p = _grid_scene()
    tri = np.array([[-0.7, -0.6, 0.13], [0.1, 0.8, 0.13], [0.55, -0.3, 0.13]])  # normal -z
    res = O.soft_rigid(V, T, p, tri, np.array([[0, 1, 2]], np.int32), pose(), pose(), np.zeros(6), np.zeros(6), 1e3,
                       0.0, 0.3, 0.3)
    assert res["counts"][0] > 0 and res["counts"][1] == 0 and res["counts"][2] == 0


def test_soft_is_M_swaps_normals_not_forces():
    V, T, p = _grid_scene()
    tri = np.array([[-0.7, -0.6, 0.13], [0.55, -0.3, 0.13], [0.1, 0.8, 0.13]])
    F = np.array([[0, 1, 2]], np.int32)
    VS = np.array([0.1, -0.2, 0.3, 0.01, 0.02, -0.03])
    a = O.soft_rigid(V, T, p, tri, F, pose(), pose(), VS, np.zeros(6), 1e3, 0.5, 0.3, 0.3, soft_is_M=True)
    b = O.soft_rigid(V, T, p, tri, F, pose(), pose(), VS, np.zeros(6), 1e3, 0.5, 0.3, 0.3, soft_is_M=False)
    np.testing.assert_array_equal(a["wrench"], b["wrench"])
    np.testing.assert_array_equal(a["faces"][:, 3:6], -b["faces"][:, 3:6])


def test_rigid_motion_invariance():
    V, T, p = _grid_scene()
    tri = np.array([[-0.7, -0.6, 0.13], [0.55, -0.3, 0.13], [0.1, 0.8, 0.13]])
    F = np.array([[0, 1, 2]], np.int32)
    base = O.soft_rigid(V, T, p, tri, F, pose(), pose(), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3)
    R = rot([1, 2, 3], 0.7)
    t = np.array([0.3, -1.2, 2.0])
    moved = O.soft_rigid(V, T, p, tri, F, pose(R, t), pose(R, t), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3)
    np.testing.assert_allclose(moved["wrench"][:3], R @ base["wrench"][:3], rtol=1e-10, atol=1e-9)
    assert moved["counts"][1] == base["counts"][1]


# ---------------------------------------------------------------- force law


def test_force_at_rest_is_fn0_along_normal():
    f, fn = O.point_force([0.1, 0, 0], [0, 0, 1], 2.5, 1e4, 3.0, [0, 0, 0], np.zeros(6), [1, 0, 0], np.zeros(6),
                          0.3)
    assert fn == 2.5
    np.testing.assert_allclose(f, [0, 0, 2.5], atol=1e-15)


def test_force_normal_approach_damps_and_stiffens():
    # A moves along -n at 0.1 m/s: vn = -0.1 -> fn = (1 + d 0.1)(fn0 + dt k 0.1)
    VA = np.array([0, 0, 0, 0, 0, -0.1])
    f, fn = O.point_force([0, 0, 0], [0, 0, 1], 2.0, 500.0, 3.0, [0, 0, 0], VA, [0, 0, 0], np.zeros(6), 0.3)
    assert fn == pytest.approx(1.3 * (2.0 + 0.001 * 500.0 * 0.1), rel=1e-15)


def test_force_separating_fast_is_zero():
    VA = np.array([0, 0, 0, 0, 0, 1.0])  # d vn = 3 > 1
    f, fn = O.point_force([0, 0, 0], [0, 0, 1], 2.0, 500.0, 3.0, [0, 0, 0], VA, [0, 0, 0], np.zeros(6), 0.3)
    assert fn == 0.0 and np.all(f == 0.0)


def test_adhesive_branch_is_not_clamped():
    # the second factor (fn0 - dt k vn) may be negative (plugin.cpp:458 quirk)
    VA = np.array([0, 0, 0, 0, 0, 0.2])
    f, fn = O.point_force([0, 0, 0], [0, 0, 1], 0.01, 1e5, 3.0, [0, 0, 0], VA, [0, 0, 0], np.zeros(6), 0.3)
    assert fn == pytest.approx((1 - 0.6) * (0.01 - 0.001 * 1e5 * 0.2), rel=1e-15)
    assert fn < 0


def test_sliding_friction_is_coulomb():
    VA = np.array([0, 0, 0, 0.5, 0, 0])  # fast slip along +x
    f, fn = O.point_force([0, 0, 0], [0, 0, 1], 2.0, 1e3, 3.0, [0, 0, 0], VA, [0, 0, 0], np.zeros(6), 0.4)
    assert fn == 2.0
    assert f[0] == pytest.approx(-0.4 * 2.0 * 0.5 / math.sqrt(0.25 + 1e-12), rel=1e-15)
    assert f[2] == 2.0


def test_stiction_regularisation():
    vt = 5e-5  # s = v_slip / 1e-4 ~ 0.5
    VA = np.array([0, 0, 0, vt, 0, 0])
    f, fn = O.point_force([0, 0, 0], [0, 0, 1], 2.0, 1e3, 3.0, [0, 0, 0], VA, [0, 0, 0], np.zeros(6), 0.4)
    vs = math.sqrt(vt * vt + 1e-12)
    s = vs / 1e-4
    assert f[0] == pytest.approx(-0.4 * s * (2 - s) * (vt / vs) * 2.0, rel=1e-14)


def test_rotational_velocity_enters_through_the_lever_arm():
    # B spins about z at 1 rad/s; point at (0.1,0,0) from B's origin moves +y at 0.1
    VB = np.array([0, 0, 1.0, 0, 0, 0])
    f, fn = O.point_force([0.1, 0, 0], [0, 0, 1], 2.0, 1e3, 0.0, [0.1, 0, 0], np.zeros(6), [0, 0, 0], VB, 0.5)
    assert f[1] == pytest.approx(0.5 * 2.0 * 0.1 / math.sqrt(0.01 + 1e-12), rel=1e-14)  # A slides -y rel. to B


def test_apply_forces_off_gives_zero_force():
    f, fn = O.point_force([0, 0, 0], [0, 0, 1], 2.0, 1e3, 3.0, [0, 0, 0], np.zeros(6), [0, 0, 0], np.zeros(6), 0.3,
                          apply=False)
    assert fn == 2.0 and np.all(f == 0)


def test_combined_material_parameters():
    inf = math.inf
    assert O.lib().orc_combined_modulus(inf, 3e4) == 3e4
    assert O.lib().orc_combined_modulus(2e4, inf) == 2e4
    assert O.lib().orc_combined_modulus(1e4, 3e4) == pytest.approx(7500.0, rel=1e-15)
    assert O.lib().orc_combined_dissipation(inf, 1.0, inf, 3.0) == 2.0
    assert O.lib().orc_combined_dissipation(5e4, 5.0, inf, 1.0) == 5.0
    # both soft: E*/E_A d_A + E*/E_B d_B with E* = 7500
    assert O.lib().orc_combined_dissipation(1e4, 2.0, 3e4, 4.0) == pytest.approx(0.75 * 2 + 0.25 * 4, rel=1e-15)
    assert O.lib().orc_combined_friction(0.3, 0.3) == pytest.approx(0.3, rel=1e-15)
    assert O.lib().orc_combined_friction(0.0, 0.0) == 0.0
    assert O.lib().orc_combined_friction(0.2, 0.6) == pytest.approx(0.3, rel=1e-15)


# --------------------------------------------------------------- candidates


def test_candidates_bvh_equals_numpy_brute_force():
    from mujoco_contact_surfaces_amd import scenes as S
    sc = S.sphere_box(3)
    s, r = sc["soft"], sc["rigid"]
    keys = O.candidates(s["V"], s["T"], r["V"], r["F"], sc["X_WA"], sc["X_WB"], brute=False)
    # independent brute force: numpy AABBs in the soft frame
    X = O.compose_SR(sc["X_WA"], sc["X_WB"])
    R, t = X[:9].reshape(3, 3), X[9:]
    VR = ((r["V"][:, 0:1] * R[:, 0] + r["V"][:, 1:2] * R[:, 1]) + r["V"][:, 2:3] * R[:, 2]) + t
    tl, th = s["V"][s["T"]].min(axis=1), s["V"][s["T"]].max(axis=1)
    fl, fh = VR[r["F"]].min(axis=1), VR[r["F"]].max(axis=1)
    ov = np.all((tl[:, None, :] <= fh[None, :, :]) & (fl[None, :, :] <= th[:, None, :]), axis=2)
    ti, fi = np.nonzero(ov)
    ref = np.sort(ti.astype(np.int64) * r["F"].shape[0] + fi)
    np.testing.assert_array_equal(keys, ref)


@pytest.mark.parametrize("seed", range(3))
def test_candidates_bvh_equals_oracle_brute_force(seed):
    from mujoco_contact_surfaces_amd import scenes as S
    sc = S.tactile_pad(seed, L=3, K=3, m=24)
    s, r = sc["soft"], sc["rigid"]
    a = O.candidates(s["V"], s["T"], r["V"], r["F"], sc["X_WA"], sc["X_WB"], brute=False)
    b = O.candidates(s["V"], s["T"], r["V"], r["F"], sc["X_WA"], sc["X_WB"], brute=True)
    np.testing.assert_array_equal(a, b)


# ------------------------------------------------------------------- taxels


def test_taxels_fully_covered_cells_equal_the_field_at_cell_centres():
    # Soft cube tilted about y; rigid plane z = h in the world; the field is
    # linear across the whole contact plane, so the mean of the S x S samples
    # (midpoint rule) equals the field at the taxel centre exactly.
    V, T, p = cube_grid(4, field=lambda x: 1000.0 * (x[:, 2] + 1.0))
    Rs = rot([0, 1, 0], 0.3)
    h = 0.1
    g = np.array([[x, y, h] for y in (-0.5, 0.5) for x in (-0.5, 0.5)])
    F = np.array([[0, 1, 3], [0, 3, 2]], np.int32)
    res = O.soft_rigid(V, T, p, g, F, pose(Rs), pose(), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3)
    # sensor frame: origin at z = h - zs so the top face is the rigid plane;
    # offset in x, y so no sample lies exactly on a polygon edge (there
    # Moeller-Trumbore's u, v >= 0 tests can reject the ray in both adjacent
    # triangles, as in the reference's bvh.cpp:49-74)
    zs = 0.002
    off = np.array([0.0123, -0.0071])
    Xs = pose(np.eye(3), [off[0], off[1], h - zs])
    tx = O.taxels(res["polys"], Xs, (0.4, 0.4, zs), 0.1, 8)
    cx = np.arange(8) * 0.1 - 0.4 + 0.05
    Xc, Yc = np.meshgrid(cx + off[0], cx + off[1])  # rows = y
    pts = np.stack([Xc, Yc, np.full_like(Xc, h)], -1)
    z_soft = pts @ Rs[:, 2]  # soft-frame z of world points (soft origin at 0)
    np.testing.assert_allclose(tx, 1000.0 * (z_soft + 1.0), rtol=1e-12)


def test_taxels_windows_weights():
    V, T, p = cube_grid(2, field=lambda x: 1000.0 * (x[:, 2] + 1.0))
    g = np.array([[x, y, 0.0] for y in (-0.5, 0.5) for x in (-0.5, 0.5)])
    F = np.array([[0, 1, 3], [0, 3, 2]], np.int32)
    res = O.soft_rigid(V, T, p, g, F, pose(), pose(), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3)
    Xs = pose(np.eye(3), [0, 0, -0.002])
    S = 10
    none = O.taxels(res["polys"], Xs, (0.2, 0.2, 0.002), 0.1, S, window=0)
    sq = O.taxels(res["polys"], Xs, (0.2, 0.2, 0.002), 0.1, S, window=3)
    gs = O.taxels(res["polys"], Xs, (0.2, 0.2, 0.002), 0.1, S, window=1, sigma=0.5)
    np.testing.assert_allclose(none, 1000.0, rtol=1e-13)  # constant field
    # square window: mean of (1 - r/r_max)^2 over the sample grid
    di = 0.1 / S
    off = np.arange(S) * di + di / 2 - 0.05
    r = np.hypot(off[:, None], off[None, :]) / (math.sqrt(2) * 0.05)
    np.testing.assert_allclose(sq, 1000.0 * ((1 - r) ** 2).mean(), rtol=1e-9)
    np.testing.assert_allclose(gs, 1000.0 * np.exp(-(r ** 2) * 0.5 / 0.25).mean(), rtol=1e-9)
    # Tukey window: the reference scales the sample distance n from the cell
    # edge by S^2 / sigma (flat_tactile_sensor.cpp:313, 367-379), so with
    # S = 10, sigma = 0.4 every n <= 2 gets 0.5 (1 - cos(2 pi n 250)) = 0 and
    # 5 of 10 samples per axis survive with weight 1: 1000 * (5/10)^2.  (The
    # textbook scaling n / (sigma S) would give per-axis weights 0, .5, 1, 1,
    # 1, 1, 1, 1, 1, .5: 640.)  With sigma = 0.3 only n <= 1 is tapered:
    # w(0) = 0, w(1) = 0.5 (1 - cos(2 pi / 3)) = 0.75, so (8.5 / 10)^2.
    tk = O.taxels(res["polys"], Xs, (0.2, 0.2, 0.002), 0.1, S, window=2, sigma=0.4)
    np.testing.assert_allclose(tk, 250.0, rtol=1e-12)
    tk3 = O.taxels(res["polys"], Xs, (0.2, 0.2, 0.002), 0.1, S, window=2, sigma=0.3)
    np.testing.assert_allclose(tk3, 722.5, rtol=1e-12)


def test_taxel_output_index_is_x_plus_cx_y():
    V, T, p = cube_grid(3, field=lambda x: 1000.0 * (x[:, 0] + 2.0))  # varies with x
    Rs = rot([0, 1, 0], math.pi / 2)  # soft x -> world -z ... use a tilt so the plane sees x
    Rs = rot([1, 0, 0], 0.0)
    g = np.array([[x, y, 0.0] for y in (-0.9, 0.9) for x in (-0.9, 0.9)])
    F = np.array([[0, 1, 3], [0, 3, 2]], np.int32)
    V2, T2, p2 = cube_grid(3, field=lambda x: 1000.0 * (x[:, 2] + 1.0) + 100.0 * x[:, 0])
    res = O.soft_rigid(V2, T2, p2, g, F, pose(Rs), pose(), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3)
    tx = O.taxels(res["polys"], pose(np.eye(3), [0, 0, -0.002]), (0.6, 0.2, 0.002), 0.2, 4)
    assert tx.shape == (2, 6)
    # increases along x (columns), constant along y (rows)
    assert np.all(np.diff(tx, axis=1) > 0)
    np.testing.assert_allclose(tx[0], tx[1], rtol=1e-12)


# --------------------------------------------------------------- generators


@pytest.mark.parametrize("L,K,shape", [(1, 1, 0), (2, 3, 0), (2, 2, 1), (3, 2, 2)])
def test_generator_invariants(L, K, shape):
    from mujoco_contact_surfaces_amd import scenes as S
    V, T, p = S.soft_mesh(L, K, shape, 0.05, 0.04, 0.03, 1e5)
    assert T.shape[0] == 20 * 4 ** L * (1 + 3 * (K - 1))
    vol = signed_volumes(V, T)
    assert vol.min() > 0
    # conformity: every face shared by <= 2 tets, boundary faces = 20 4^L
    faces = {}
    for t in T:
        for f in ((t[1], t[2], t[3]), (t[0], t[3], t[2]), (t[0], t[1], t[3]), (t[0], t[2], t[1])):
            k = tuple(sorted(f))
            faces[k] = faces.get(k, 0) + 1
    assert max(faces.values()) == 2
    assert sum(1 for v in faces.values() if v == 1) == 20 * 4 ** L
    if shape == 0:
        # volume of the inscribed polyhedron -> sphere volume as L grows
        assert vol.sum() < 4 / 3 * math.pi * 0.05 ** 3
        assert vol.sum() > 0.6 * 4 / 3 * math.pi * 0.05 ** 3
    assert p.max() == 1e5 and p.min() == 0.0


def test_soft_from_surface_centroid_fan():
    from mujoco_contact_surfaces_amd import scenes as S
    V, F = S.rigid_mesh(2, 2, 0.05)
    V = V + np.array([0.01, -0.02, 0.03])
    Vo, To, po = S.soft_from_surface(V, F, 2e4)
    # centroid of a centred icosphere is its centre
    np.testing.assert_allclose(Vo[-1], [0.01, -0.02, 0.03], atol=1e-15)
    assert signed_volumes(Vo, To).min() > 0
    assert po[-1] == 2e4 and np.all(po[:-1] == 0)


# ----------------------------------------------------------- rigid half-space


@pytest.mark.parametrize("h", [0.37, -0.5, 0.0])
def test_half_space_section_area_and_force(h):
    V, T, p = _grid_scene()
    # rigid half-space z <= h: its plane cuts the cube in a 2x2 square
    res = O.soft_plane(V, T, p, pose(), pose(t=(0, 0, h)), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3,
                       representation=1)
    assert res["counts"][1] > 0
    assert res["faces"][:, 9].sum() == pytest.approx(4.0, rel=1e-12)
    # kPolygon faces: p0 = p(centroid) = 1000 (h + 1) everywhere on the plane
    np.testing.assert_allclose(res["faces"][:, 10], 1000 * (h + 1), rtol=1e-12)
    np.testing.assert_allclose(res["wrench"][:3], [0, 0, 4.0 * 1000 * (h + 1)], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(res["wrench"][6:9], -res["wrench"][:3], rtol=0, atol=0)
    # keys: (tet, 0)
    assert (res["face_keys"][:, 1] == 0).all() and (res["poly_keys"][:, 1] == 0).all()


def test_half_space_tilted_plane_closed_form():
    V, T, p = _grid_scene()
    h, ang = 0.37, 0.2
    R = rot([1, 0, 0], ang)
    n = R[:, 2]
    res = O.soft_plane(V, T, p, pose(), pose(R, (0, 0, h)), np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3)
    # the section over the cube's (x, y) square: area 4 / n_z, force
    # 1000 * 4 (h + 1) / n_z along n (the odd terms of z(x, y) integrate to 0)
    assert res["faces"][:, 9].sum() == pytest.approx(4.0 / n[2], rel=1e-12)
    np.testing.assert_allclose(res["wrench"][:3], n * 4000.0 * (h + 1) / n[2], rtol=1e-12, atol=1e-9)


def test_half_space_normal_against_the_gradient_is_filtered():
    V, T, p = _grid_scene()
    # plane normal -z (rigid side above): opposed to the pressure gradient
    res = O.soft_plane(V, T, p, pose(), pose(rot([1, 0, 0], np.pi), (0, 0, 0.2)), np.zeros(6), np.zeros(6), 1e3,
                       0.0, 0.3, 0.3)
    assert res["counts"][0] > 0 and res["counts"][1] == 0 and res["counts"][2] == 0


def test_half_space_candidates_are_the_tets_straddling_the_plane():
    V, T, p = _grid_scene(n=3)
    R = rot([0.3, -0.2, 1.0], 0.4)
    X_WR = pose(R, (0.1, -0.05, 0.21))
    res = O.soft_plane(V, T, p, pose(), X_WR, np.zeros(6), np.zeros(6), 1e3, 0.0, 0.3, 0.3)
    # independent count: AABB of each tet against the plane, in numpy
    n, c = R[:, 2], X_WR[9:]
    lo, hi = V[T].min(axis=1), V[T].max(axis=1)
    mid, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    dist = (mid - c) @ n
    rad = np.abs(half) @ np.abs(n)
    assert res["counts"][0] == int((np.abs(dist) <= rad).sum())
    # polygons only where vertices lie on both sides
    s = (V[T] - c) @ n
    crossing = ((s > 0).sum(axis=1) % 4) != 0
    assert res["counts"][1] == int(crossing.sum())


def test_half_space_equals_a_plane_grid_mesh_on_force_and_area():
    from mujoco_contact_surfaces_amd import scenes as S
    V, T, p = S.soft_mesh(2, 3, 0, 0.05, 0.05, 0.05, 2e5)
    X_WS = pose(t=(0.0, 0.0, 0.045))
    VS = np.array([0, 0, 0, 0.01, -0.02, -0.05])  # translation only: vn is the same at every point
    # a large grid of rigid triangles in the plane z = 0 (normal +z)
    m, half = 9, 0.2
    xs = np.linspace(-half, half, m)
    g = np.array([[x, y, 0.0] for y in xs for x in xs])
    F = []
    for j in range(m - 1):
        for i in range(m - 1):
            a = j * m + i
            F += [[a, a + 1, a + m + 1], [a, a + m + 1, a + m]]
    F = np.array(F, np.int32)
    for repr_ in (0, 1):
        hs = O.soft_plane(V, T, p, X_WS, pose(), VS, np.zeros(6), 2e5, 3.0, 0.4, 0.3, representation=repr_)
        gr = O.soft_rigid(V, T, p, g, F, X_WS, pose(), VS, np.zeros(6), 2e5, 3.0, 0.4, 0.3, representation=repr_)
        assert hs["counts"][1] > 0
        assert hs["faces"][:, 9].sum() == pytest.approx(gr["faces"][:, 9].sum(), rel=1e-12)
        np.testing.assert_allclose(hs["wrench"][:3], gr["wrench"][:3], rtol=1e-11,
                                   atol=1e-12 * np.abs(gr["wrench"][:3]).max())


# -------------------------------------------------------------- taxel sensor


def test_area_importance_generator_is_libstdcxx_default_random_engine(tmp_path):
    """The restated generator against std::default_random_engine feeding
    std::uniform_real_distribution<double>(0, 1), compiled here."""
    src = tmp_path / "rng.cpp"
    src.write_text('#include <cstdio>\n#include <random>\nint main() { std::default_random_engine g; '
                   'std::uniform_real_distribution<double> u(0.0, 1.0); '
                   'for (int i = 0; i < 4000; ++i) std::printf("%a\\n", u(g)); }\n')
    exe = tmp_path / "rng"
    import subprocess
    subprocess.run(["g++", "-O0", str(src), "-o", str(exe)], check=True, timeout=120)
    ref = np.array([float.fromhex(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                                               text=True).stdout.split()])
    np.testing.assert_array_equal(O.minstd_uniform(4000), ref)


def _tri_rows(V, e):
    return np.concatenate([np.asarray(V, float).reshape(9), np.asarray(e, float)])[None, :]


def test_default_sample_count_follows_the_reference_loops():
    V = [[0, 0, 0], [0.013, 0, 0], [0, 0.0071, 0]]
    res = 0.001
    st0 = int(np.linalg.norm(np.subtract(V[1], V[0])) / res) + 1
    st2 = int(np.linalg.norm(np.subtract(V[2], V[0])) / res) + 1
    st = max(st0, st2)  # st1 is computed from v2 - v0 as well

    def count(step):
        n, a = 0, 0.0
        while a <= 1:
            n += 1
            a += 1.0 / step
        return n

    _, m = O.taxel_sensor(_tri_rows(V, [1, 1, 1]), [0], pose(), np.zeros((1, 3)), 0.01, res)
    assert m == count(st) * count(st2)


def test_closest_reads_the_vertex_pressure_only_when_visualizing():
    V = [[0, 0, 0], [0.01, 0, 0], [0, 0.01, 0]]
    tri = _tri_rows(V, [100.0, 200.0, 300.0])
    taxels = np.array([[0.01, 0, 0], [0, 0.01, 0], [0, 0, 0], [0.5, 0.5, 0.5]])
    prev = np.full(4, 7.0)
    vals, _ = O.taxel_sensor(tri, [0], pose(), taxels, 0.002, 0.001, "closest", visualize=True, values=prev)
    # a = b = 0 is always a sample, at v1; the loops a, b += 1/st stop short
    # of 1 here (11 steps of 1/11 exceed 1), so v0 and v2 are not sampled and
    # their taxels read the nearest interior sample
    assert vals[0] == 200.0
    assert 200.0 < vals[1] < 300.0 and 100.0 < vals[2] < 200.0
    assert vals[3] == 7.0  # out of the margin: the previous value stays
    vals, _ = O.taxel_sensor(tri, [0], pose(), taxels, 0.002, 0.001, "closest", visualize=False, values=prev)
    np.testing.assert_array_equal(vals, [0.0, 0.0, 0.0, 7.0])


def test_weighted_and_mean_fall_through_to_squared():
    rng = np.random.default_rng(3)
    V = rng.uniform(-0.01, 0.01, (3, 3))
    tri = _tri_rows(V, [50.0, 80.0, 120.0])
    taxels = rng.uniform(-0.01, 0.01, (20, 3))
    out = [O.taxel_sensor(tri, [0], pose(), taxels, 0.004, 0.0005, m, values=np.full(20, -1.0))[0]
           for m in ("weighted", "mean", "squared")]
    np.testing.assert_array_equal(out[0], out[2])
    np.testing.assert_array_equal(out[1], out[2])
    assert (out[2] > 0).any() and (out[2] == -1.0).any()


def test_squared_is_linear_in_pressure_and_rigidly_invariant():
    rng = np.random.default_rng(5)
    V = rng.uniform(-0.01, 0.01, (3, 3))
    taxels = rng.uniform(-0.008, 0.008, (16, 3))
    e = np.array([10.0, 40.0, 90.0])
    a, _ = O.taxel_sensor(_tri_rows(V, e), [0], pose(), taxels, 0.005, 0.0007)
    b, _ = O.taxel_sensor(_tri_rows(V, 2 * e), [0], pose(), taxels, 0.005, 0.0007)
    np.testing.assert_array_equal(b, 2 * a)  # doubling every pressure is exact
    # move the surface and the sensor together: same values up to rounding
    R, t = rot([0.3, 0.5, -0.2], 0.7), np.array([0.02, -0.01, 0.03])
    VW = np.asarray(V) @ R.T + t
    c, _ = O.taxel_sensor(_tri_rows(VW, e), [0], pose(R, t), taxels, 0.005, 0.0007)
    np.testing.assert_allclose(c, a, rtol=1e-9, atol=1e-12 * np.abs(a).max())


def test_area_importance_samples_about_one_per_resolution_step():
    rng = np.random.default_rng(7)
    tris = np.concatenate([_tri_rows(rng.uniform(-0.01, 0.01, (3, 3)), [1, 2, 3]) for _ in range(6)])
    surf = [0, 0, 0, 1, 1, 1]
    res = 0.05  # area_resolution = 5% of each surface's area
    _, m = O.taxel_sensor(tris, surf, pose(), np.zeros((1, 3)), 0.01, res, sample_method="area_importance")
    # per surface the accumulator steps through [0, total) in res * total steps
    assert 2 * 20 <= m <= 2 * 21


# -------------------------------------------------------------- curved sensor


def _csr(lists):
    ptr = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int32)
    idx = np.array([j for l in lists for j, _ in l], np.int32)
    w = np.array([x for l in lists for _, x in l], float)
    return ptr, idx, w


def test_curved_sensor_reads_a_linear_pressure_field_under_each_sample():
    # a large triangle at z = 0 carrying f(x, y) = 100 + 2000 x - 500 y at its
    # vertices: a ray along -z from above (x, y) reads f(x, y) exactly
    V = np.array([[-0.1, -0.1, 0.0], [0.2, -0.1, 0.0], [-0.1, 0.2, 0.0]])
    f = lambda x, y: 100 + 2000 * x - 500 * y
    tri = _tri_rows(V, [f(*v[:2]) for v in V])
    rng = np.random.default_rng(11)
    xy = rng.uniform(-0.03, 0.03, (6, 2))
    d = np.array([0.001, 0.0005, 0.0019, 0.003, 0.0001, 0.0025])  # margin 0.002: samples 3 and 5 miss
    P = np.c_[xy, d]
    N = np.tile([0.0, 0.0, 1.0], (6, 1))
    lists = [[(0, 0.3), (1, 0.7)], [(3, 2.0)], [], [(2, 1.0), (4, -0.5), (5, 4.0), (2, 0.25)]]
    ptr, idx, w = _csr(lists)
    vals, hits = O.curved_sensor(tri, pose(), P, N, ptr, idx, w, 0.002)
    assert hits == 4
    raw = [f(*xy[j]) if d[j] < 0.002 else 0.0 for j in range(6)]
    want = [sum(x * raw[j] for j, x in l) for l in lists]
    np.testing.assert_allclose(vals, want, rtol=1e-12, atol=1e-12)
    assert vals[2] == 0.0


def test_curved_sensor_takes_the_nearest_surface_along_the_inward_normal():
    big = np.array([[-1.0, -1.0, 0.0], [2.0, -1.0, 0.0], [-1.0, 2.0, 0.0]])
    upper = _tri_rows(big, [5.0, 5.0, 5.0])
    lower = _tri_rows(big - [0, 0, 0.001], [9.0, 9.0, 9.0])
    P = np.array([[0.1, 0.1, 0.0005], [0.1, 0.1, -0.0005], [0.1, 0.1, -0.0005], [0.2, 0.1, 0.0],
                  [0.1, 0.1, 0.0005]])
    N = np.array([[0, 0, 1.0], [0, 0, 1.0], [0, 0, -1.0], [0, 0, 1.0], [0, 0, -1.0]])
    ptr, idx, w = _csr([[(j, 1.0)] for j in range(5)])
    for tris in (np.concatenate([upper, lower]), np.concatenate([lower, upper])):
        vals, hits = O.curved_sensor(tris, pose(), P, N, ptr, idx, w, 0.01)
        # above both: the upper one; between, facing up: the lower one;
        # between, facing down (ray along +z): the upper one; on the upper
        # surface: the ray starts 1e-8 above it and hits it; above both,
        # facing down: the ray leaves both behind
        np.testing.assert_array_equal(vals, [5.0, 9.0, 5.0, 5.0, 0.0])
        assert hits == 4
    vals, _ = O.curved_sensor(lower, pose(), P, N, ptr, idx, w, 0.01)
    np.testing.assert_array_equal(vals, [9.0, 9.0, 0.0, 9.0, 0.0])


def test_curved_sensor_without_surfaces_reads_zero():
    ptr, idx, w = _csr([[(0, 1.0)], []])
    vals, hits = O.curved_sensor(np.zeros((0, 12)), pose(), np.zeros((1, 3)), np.array([[0, 0, 1.0]]), ptr, idx, w, 1.0)
    np.testing.assert_array_equal(vals, [0.0, 0.0])
    assert hits == 0


def test_curved_sensor_hits_agree_with_a_linear_solve():
    """Moeller-Trumbore against O + t D = v0 + u e1 + v e2 solved by LAPACK:
    e = (0, 1, 0) makes the reading u, e = (0, 0, 1) makes it v."""
    rng = np.random.default_rng(13)
    checked = 0
    for _ in range(600):
        V = rng.uniform(-1, 1, (3, 3))
        n = rng.normal(size=3)
        a, b = rng.uniform(-0.3, 1.0, 2)
        p = V[0] + a * (V[1] - V[0]) + b * (V[2] - V[0]) + rng.uniform(-0.5, 1.5) * n
        O_, D = p + n * 1e-8, -n
        t, u, v = np.linalg.solve(np.c_[-D, V[1] - V[0], V[2] - V[0]], O_ - V[0])
        if min(abs(u), abs(v), abs(1 - u - v)) < 1e-6 or abs(t) < 1e-6:
            continue  # too close to an edge or the origin to call
        inside = u > 0 and v > 0 and u + v < 1 and t > 0
        ptr, idx, w = _csr([[(0, 1.0)]])
        ru, h1 = O.curved_sensor(_tri_rows(V, [0, 1, 0]), pose(), p[None], n[None], ptr, idx, w, 10.0)
        rv, _ = O.curved_sensor(_tri_rows(V, [0, 0, 1]), pose(), p[None], n[None], ptr, idx, w, 10.0)
        assert h1 == int(inside and t < 10.0)
        if h1:
            np.testing.assert_allclose([ru[0], rv[0]], [u, v], rtol=1e-9, atol=1e-12)
            # the margin test is on t: just above t accepts, just below rejects
            assert O.curved_sensor(_tri_rows(V, [0, 1, 0]), pose(), p[None], n[None], ptr, idx, w, t * (1 + 1e-9))[1]
            assert not O.curved_sensor(_tri_rows(V, [0, 1, 0]), pose(), p[None], n[None], ptr, idx, w,
                                       t * (1 - 1e-9))[1]
            checked += 1
    assert checked > 100


def test_curved_sensor_is_rigidly_invariant():
    rng = np.random.default_rng(17)
    V = rng.uniform(-0.02, 0.02, (3, 3))
    V[:, 2] = 0.0
    e = np.array([30.0, 60.0, 90.0])
    P = np.c_[rng.uniform(-0.005, 0.005, (12, 2)), rng.uniform(0, 0.002, 12)]
    N = np.tile([0.0, 0.0, 1.0], (12, 1))
    ptr, idx, w = _csr([[(j, 1.0 + j)] for j in range(12)])
    a, ha = O.curved_sensor(_tri_rows(V, e), pose(), P, N, ptr, idx, w, 0.003)
    R, t = rot([0.3, -0.4, 0.8], 1.1), np.array([0.5, -0.2, 0.1])
    b, hb = O.curved_sensor(_tri_rows(V @ R.T + t, e), pose(R, t), P, N, ptr, idx, w, 0.003)
    assert ha == hb > 0
    np.testing.assert_allclose(b, a, rtol=1e-9, atol=1e-12 * np.abs(a).max())


# ------------------------------------------- per-step oracle (CPU baseline)


@pytest.mark.parametrize("env", range(6))
def test_prepared_step_equals_the_one_shot_query(env):
    """The split load-time / per-step oracle that the CPU baseline times
    (orc_soft_prepare + orc_soft_rigid_step) gives the one-shot query's
    counts and wrench bit for bit, and the brute-force candidate count."""
    from mujoco_contact_surfaces_amd import scenes as S
    sc = S.batched_env(7, env, nt_range=(2e3, 1.2e4), ntri_range=(12, 3e3))
    s, r = sc["soft"], sc["rigid"]
    E, d = s["props"]["hydroelastic_modulus"], s["props"]["dissipation"]
    mus, mur = s["props"]["dynamic_friction"], r["props"]["dynamic_friction"]
    one = O.soft_rigid(s["V"], s["T"], s["p"], r["V"], r["F"], sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"], E, d,
                       mus, mur, soft_is_M=s["geom_id"] < r["geom_id"], want=False)
    brute = O.soft_rigid(s["V"], s["T"], s["p"], r["V"], r["F"], sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"], E,
                         d, mus, mur, soft_is_M=s["geom_id"] < r["geom_id"], want=False, brute=True)
    prep = O.PreparedSoft(s["V"], s["T"], s["p"])
    pose_ = np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]])
    W, cnt = prep.step(r["V"], r["F"], pose_, [E, d, mus, mur, 0.001, 1, 0, int(s["geom_id"] < r["geom_id"])])
    prep.close()
    assert one["counts"][0] > 0
    np.testing.assert_array_equal(cnt, one["counts"])
    np.testing.assert_array_equal(cnt, brute["counts"])
    np.testing.assert_array_equal(W, one["wrench"])
    np.testing.assert_array_equal(W, brute["wrench"])
