"""Flat tactile sensors inside the batched step (hcs_batch_set_sensors):
per environment, the taxels of its pad sensor against the oracle's sampling
(orc_taxels) of the oracle's polygons of that environment, within 1e-10
relative; the sensor pose comes from the batch state."""
import sys

import numpy as np
import pytest

import oracle as O
from mujoco_contact_surfaces_amd import scenes as S
from meshes import rot

pytestmark = pytest.mark.gpu

RTOL = 1e-10
WINDOWS = ((0, 0.0), (1, 0.3), (2, 0.5), (3, 0.0))


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _compose(T, X):
    """World motion T (12: row-major R, p) applied to pose X: T X."""
    RT, pT = np.asarray(T[:9]).reshape(3, 3), np.asarray(T[9:])
    R, p = np.asarray(X[:9]).reshape(3, 3), np.asarray(X[9:])
    return np.concatenate([(RT @ R).reshape(9), RT @ p + pT])


def _moved(sc, k):
    """The scene under a rigid motion of the whole world (the pad's pose is
    then not the identity)."""
    rng = np.random.default_rng(100 + k)
    ax = rng.normal(size=3)
    T = np.concatenate([rot(ax / np.linalg.norm(ax), rng.uniform(0.1, 1.0)).reshape(9), rng.uniform(-0.1, 0.1, 3)])
    out = dict(sc)
    out["X_WA"], out["X_WB"] = _compose(T, sc["X_WA"]), _compose(T, sc["X_WB"])
    return out


def _oracle_taxels(sc, g):
    s, r = sc["soft"], sc["rigid"]
    ref = O.soft_rigid(s["V"], s["T"], s["p"], r["V"], r["F"], sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"],
                       s["props"]["hydroelastic_modulus"], s["props"]["dissipation"], s["props"]["dynamic_friction"],
                       r["props"]["dynamic_friction"], soft_is_M=s["geom_id"] < r["geom_id"])
    tx = O.taxels(ref["polys"], sc["X_WB"], g["half_size"], g["resolution"], g["sampling_resolution"], g["window"],
                  g["sigma"])
    return ref, tx


@pytest.mark.parametrize("n_env", [8, 48])
def test_batched_taxels_match_the_oracle(engine, n_env):
    """8 environments (config-4 pads and one config-2 pad, every window), and
    48 environments, whose polygons overflow the initial polygon buffer, so
    the step is re-run with larger buffers."""
    say("batched taxels", n_env)
    scs = [_moved(S.differentiation_env(1, e), e) for e in range(n_env - 1)] + [_moved(S.tactile_pad(1), 99)]
    handles = []
    for sc in scs:
        s, r = sc["soft"], sc["rigid"]
        handles.append((engine.add_soft_mesh(s["V"], s["T"], s["p"], s["props"], s["geom_id"]),
                        engine.add_rigid_mesh(r["V"], r["F"], r["props"], r["geom_id"])))
    b = engine.batch(handles)
    b.set_state(np.array([np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]]) for sc in scs]))
    grids = []
    for e, sc in enumerate(scs):
        w, sig = WINDOWS[e % len(WINDOWS)]
        g = dict(sc["grid"])
        g.update(window=w, sigma=sig, geom=handles[e][1])
        grids.append(g)
    # sensors listed in reverse pair order: the output follows sensor order
    order = list(range(n_env))[::-1]
    b.set_sensors(order, [grids[e] for e in order])
    b.step()
    W, cnt = b.get()
    tx = b.taxels()
    assert len(tx) == n_env
    worst = 0.0
    for k, e in enumerate(order):
        ref, txr = _oracle_taxels(scs[e], grids[e])
        np.testing.assert_array_equal(cnt[e], ref["counts"])
        assert tx[k].shape == txr.shape == (16, 16)
        assert np.abs(txr).max() > 0
        np.testing.assert_allclose(tx[k], txr, rtol=RTOL, atol=RTOL * np.abs(txr).max())
        worst = max(worst, np.abs(tx[k] - txr).max() / np.abs(txr).max())
    say("max taxel err (rel to max)", worst)
    # a second step on the same state gives the same taxels bit for bit
    b.step()
    tx2 = b.taxels()
    for a, c in zip(tx, tx2):
        np.testing.assert_array_equal(a, c)
    b.close()


def test_batch_sensor_arguments_are_checked(engine):
    say("batched sensor arguments")
    sc = S.differentiation_env(1, 0)
    s, r = sc["soft"], sc["rigid"]
    ha = engine.add_soft_mesh(s["V"], s["T"], s["p"], s["props"], s["geom_id"])
    hb = engine.add_rigid_mesh(r["V"], r["F"], r["props"], r["geom_id"])
    b = engine.batch([(ha, hb)])
    from mujoco_contact_surfaces_amd.engine import HcsError
    g = dict(sc["grid"], geom=hb)
    with pytest.raises(HcsError, match="pair out of range"):
        b.set_sensors([1], [g])
    with pytest.raises(HcsError, match="not in its pair"):
        b.set_sensors([0], [dict(g, geom=hb + 1000)])
    with pytest.raises(HcsError, match="no sensors"):
        b.taxels()
    # the batch still steps without sensors
    b.set_state(np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]])[None, :])
    b.step()
    W, cnt = b.get()
    assert cnt[0, 0] > 0
    b.close()
