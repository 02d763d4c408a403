"""GPU parity at the benchmark workloads' full sizes, per environment.

* Batched soft-rigid workload: 4096 environments (soft tets log-uniform over
  2k-100k, rigid triangles log-uniform over 12-20k), one batched graph step
  before and after the autotuner re-plans it; per env the candidate,
  polygon and face counts bit-exact and the wrench within 1e-10
  (norm-relative per 3-vector) of the oracle's per-step query; the
  canonical candidate keys bit-exact on 64 seeded envs spanning every soft
  shape x rigid kind plus the 8 heaviest envs.
* Soft-soft workload: pairs of 20k-200k-tet bodies, two through the plain
  per-call path and 256 through one batched step, each against the
  oracle's tet-tet query.

Reference path: collision_cb -> Drake query -> evaluateContactSurface ->
passiveCallback force loop (mujoco_contact_surfaces/src/
mujoco_contact_surfaces_plugin.cpp:255-318, 320-409, 411-483).
"""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from mujoco_contact_surfaces_amd import Engine, scenes as S

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def say(*a):
    print("[case]", *a, flush=True)


def _threads():
    # the GPU host's share of CPU threads for one GPU
    return max(1, min(16, len(os.sched_getaffinity(0))))


def wrench_err(got, ref):
    """Largest per-3-vector error of a 12-wrench relative to its own norm,
    with an absolute floor of 1e-14 of the reference force (torques about
    symmetry planes can cancel to nothing)."""
    got, ref = np.asarray(got), np.asarray(ref)
    scale = np.linalg.norm(ref[:3])
    e = 0.0
    for b in range(4):
        g, r = got[3 * b:3 * b + 3], ref[3 * b:3 * b + 3]
        d = np.linalg.norm(g - r)
        if d == 0.0:
            continue
        e = max(e, d / (np.linalg.norm(r) + 1e-4 * scale + 1e-300))
    return e


def oracle_soft_rigid_env(sc):
    s, r = sc["soft"], sc["rigid"]
    prep = O.PreparedSoft(s["V"], s["T"], s["p"])
    try:
        pose = np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]])
        params = [s["props"]["hydroelastic_modulus"], s["props"]["dissipation"], s["props"]["dynamic_friction"],
                  r["props"]["dynamic_friction"], 0.001, 1, 0, int(s["geom_id"] < r["geom_id"])]
        return prep.step(r["V"], r["F"], pose, params)
    finally:
        prep.close()


def oracle_soft_soft_env(sc):
    a, b = sc["soft"], sc["soft_b"]
    res = O.soft_soft(a["V"], a["T"], a["p"], b["V"], b["T"], b["p"], sc["X_WA"], sc["X_WB"], sc["V_WA"],
                      sc["V_WB"], a["props"]["hydroelastic_modulus"], a["props"]["dissipation"],
                      a["props"]["dynamic_friction"], b["props"]["hydroelastic_modulus"], b["props"]["dissipation"],
                      b["props"]["dynamic_friction"], a_is_M=a["geom_id"] < b["geom_id"], want=False)
    return res["wrench"], res["counts"]


def compare_envs(W, cnt, refs, what):
    """Per env: counts bit-exact, wrench within RTOL.  Returns the largest
    wrench error; asserts with the first failing env."""
    worst = 0.0
    bad = []
    for e, (Wr, cr) in enumerate(refs):
        if not np.array_equal(cnt[e], cr):
            bad.append((e, "counts", cnt[e].tolist(), cr.tolist()))
            continue
        err = wrench_err(W[e], Wr)
        worst = max(worst, err)
        if err > RTOL:
            bad.append((e, "wrench", err))
    say(what, "envs", len(refs), "mismatches", len(bad), "max wrench rel err %.2e" % worst)
    assert not bad, bad[:5]
    return worst


@pytest.mark.timeout(1500)
def test_batched_4096_env_workload_matches_the_oracle_per_env():
    n_env = 4096
    say("batched soft-rigid workload, full size:", n_env, "envs")
    t0 = time.time()
    eng = Engine(0)
    try:
        scs, handles, state = S.upload_batched(eng, n_env, seed=0, log=lambda m: say(m, "%.0f s" % (time.time() - t0)))
        say("total tets", sum(sc["soft"]["T"].shape[0] for sc in scs), "tris",
            sum(sc["rigid"]["F"].shape[0] for sc in scs))
        batch = eng.batch(handles)
        batch.set_state(state)
        batch.step()
        W0, cnt0 = batch.get()
        say("first step: clips", int(cnt0[:, 0].sum()), "polygons", int(cnt0[:, 1].sum()))
        ms = batch.tune(reps=1)
        say("tuned", batch.config(), "%.3f ms" % ms)
        batch.step()
        W1, cnt1 = batch.get()
        np.testing.assert_array_equal(cnt1, cnt0)
        with ThreadPoolExecutor(_threads()) as ex:
            refs = list(ex.map(oracle_soft_rigid_env, scs))
        say("oracle done %.0f s" % (time.time() - t0))
        compare_envs(W0, cnt0, refs, "untuned plan")
        compare_envs(W1, cnt1, refs, "tuned plan")
        # canonical candidate keys: 64 seeded envs over every shape x rigid
        # kind, plus the 8 heaviest
        rng = np.random.default_rng(2026)
        groups = {}
        for e, sc in enumerate(scs):
            groups.setdefault((sc["shape"], sc["rigid_kind"]), []).append(e)
        assert len(groups) == 9
        pick = []
        for k in sorted(groups):
            g = groups[k]
            pick += [int(e) for e in rng.choice(g, size=min(len(g), 7), replace=False)]
        pick += [int(e) for e in rng.choice(n_env, size=64 - len(pick), replace=False) if int(e) not in pick]
        pick = sorted(set(pick) | set(int(e) for e in np.argsort(-cnt0[:, 0])[:8]))
        assert len({(scs[e]["shape"], scs[e]["rigid_kind"]) for e in pick}) == 9

        def keys_ref(e):
            sc = scs[e]
            s, r = sc["soft"], sc["rigid"]
            return O.candidates(s["V"], s["T"], r["V"], r["F"], sc["X_WA"], sc["X_WB"])

        with ThreadPoolExecutor(_threads()) as ex:
            kref = list(ex.map(keys_ref, pick))
        for e, ref in zip(pick, kref):
            sc = scs[e]
            got = eng.candidates((int(handles[e, 0]), int(handles[e, 1]), sc["X_WA"], sc["X_WB"], sc["V_WA"],
                                  sc["V_WB"]))
            assert got.size == cnt0[e, 0]
            np.testing.assert_array_equal(got, ref, err_msg="env %d" % e)
        say("candidate keys bit-exact on", len(pick), "envs, largest", max(k.size for k in kref))
        batch.close()
    finally:
        eng.close()


@pytest.mark.timeout(900)
def test_soft_soft_pairs_at_full_size_match_the_oracle():
    say("soft-soft workload: 2 plain pairs, then 256 batched pairs of 20k-200k tets")
    t0 = time.time()
    eng = Engine(0)
    try:
        scs, handles, state = S.upload_soft_soft(eng, 256, seed=0, log=lambda m: say(m, "%.0f s" % (time.time() - t0)))
        sizes = np.array([[sc["soft"]["T"].shape[0], sc["soft_b"]["T"].shape[0]] for sc in scs])
        assert sizes.min() >= 2e4 * 0.7 and sizes.max() <= 2e5 * 1.5
        with ThreadPoolExecutor(_threads()) as ex:
            refs = list(ex.map(oracle_soft_soft_env, scs))
        say("oracle done %.0f s; clips %d" % (time.time() - t0, sum(int(c[0]) for _, c in refs)))
        # two plain per-call pairs: the two largest
        big = list(np.argsort(-sizes.sum(axis=1))[:2])
        for e in big:
            sc = scs[e]
            W, cnt = eng.contact_step([(int(handles[e, 0]), int(handles[e, 1]), sc["X_WA"], sc["X_WB"], sc["V_WA"],
                                        sc["V_WB"])])
            compare_envs(W, cnt, [refs[e]], "plain pair %d (%d + %d tets)" % (e, sizes[e, 0], sizes[e, 1]))
        batch = eng.batch(handles)
        batch.set_state(state)
        batch.step()
        W, cnt = batch.get()
        compare_envs(W, cnt, refs, "batched 256 pairs")
        assert int(cnt[:, 1].sum()) > 0
        batch.close()
    finally:
        eng.close()
