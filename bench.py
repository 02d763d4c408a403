"""Benchmark: batched FP64 hydroelastic contact steps on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--envs 4096]

Workload (synthetic, seeded): N independent environments, each a soft tet
body (tet count log-uniform over 2k-100k; sphere, ellipsoid or capsule)
against a rigid mesh (triangle count log-uniform over 12-20k; box, plane grid
or icosphere) at 0.5-8 mm penetration.  One step = poses and twists in (already
resident in HBM), per-environment net wrenches out: LBVH traversal, candidate
compaction, clipping, face evaluation, Hunt-Crossley/friction force law and
deterministic reductions, replayed as one hipGraph: k_cand (LBVH traversal or
tiled tet queries, exact cheap rejections, survivor compaction), a scan,
k_clip (clip, faces, force law per survivor) and a fixed-order reduction.

Metric: tet-tri clips per second, a clip being one element of the canonical
candidate set (tet and triangle AABBs overlap in the soft frame).  Printed on
one JSON line with the roofline of the dominant kernel and the CPU baseline
(the oracle's per-step work timed on this host on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tet–tri clips/s + µs per contact step (batched vs plain, FP64), ±adjoint"
HBM_PEAK = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(engine, n_env, seed):
    from mujoco_contact_surfaces_amd import scenes as S
    scs, handles = [], []
    t0 = time.time()
    for e in range(n_env):
        sc = S.batched_env(seed, e)
        s, r = sc["soft"], sc["rigid"]
        ha = engine.add_soft_mesh(s["V"], s["T"], s["p"], s["props"], s["geom_id"])
        hb = engine.add_rigid_mesh(r["V"], r["F"], r["props"], r["geom_id"])
        scs.append(sc)
        handles.append((ha, hb))
        if (e + 1) % 512 == 0:
            log("uploaded %d/%d envs (%.1f s)" % (e + 1, n_env, time.time() - t0))
    state = np.array([np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]]) for sc in scs])
    return scs, np.array(handles, dtype=np.int32), state


def algorithmic_bytes(engine, scs, handles, counts, survivors, nitems):
    """Algorithmic HBM bytes per step of the two pipeline kernels.

    k_cand (candidate generation + exact cheap rejections): every rigid mesh
    once (24 B per vertex, 12 B per triangle index triple, 24 B per triangle
    normal), each distinct candidate tet once (192 B record + 48 B box), the
    pair state (288 B) and work item (16 B) in, 16 B per survivor out.
    k_clip: 16 B per survivor in, one 80 B contribution out; element reads are
    left out, so this is a lower bound."""
    T_C = R_C = 0
    for sc, (ha, hb) in zip(scs, handles):
        keys = engine.candidates((ha, hb, sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]))
        ntri = sc["rigid"]["F"].shape[0]
        T_C += np.unique(keys // ntri).size
        R_C += np.unique(keys % ntri).size
    rigid = sum(24 * sc["rigid"]["V"].shape[0] + 36 * sc["rigid"]["F"].shape[0] for sc in scs)
    E = len(scs)
    b_cand = rigid + 240 * T_C + 288 * E + 16 * nitems + 16 * survivors
    b_clip = 96 * survivors
    return dict(k_cand=b_cand, k_clip=b_clip), dict(rigid_bytes=rigid, distinct_tets=T_C, distinct_tris=R_C,
                                                    survivors=survivors, items=nitems,
                                                    polygons=int(counts[:, 1].sum()))


def measured_traffic(kernel, envs, seed):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), when
    it covers this workload; else (None, None)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_batched_4096env.json")
    if envs != 4096 or seed != 0 or not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    v = d.get("hbm_bytes_per_launch", {}).get(kernel)
    return (v, d.get("source")) if v is not None else (None, None)


def cpu_baseline(scs, budget_s=12.0, chunk=128):
    """Oracle per-step work (load-time BVH excluded) on the environments in
    order, chunk by chunk, until about budget_s of serial step time: each
    chunk is prepared (untimed), stepped serially (timed), stepped again on
    all host threads (timed), then freed, so host memory stays bounded."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from concurrent.futures import ThreadPoolExecutor

    threads = min(16, os.cpu_count() or 1)
    clips = par_clips = n = 0
    serial = par = 0.0
    with ThreadPoolExecutor(threads) as ex:
        while n < len(scs) and serial < budget_s:
            preps, args = [], []
            for sc in scs[n:n + chunk]:
                s, r = sc["soft"], sc["rigid"]
                preps.append(O.PreparedSoft(s["V"], s["T"], s["p"]))
                pose = np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]])
                params = [s["props"]["hydroelastic_modulus"], s["props"]["dissipation"],
                          s["props"]["dynamic_friction"], r["props"]["dynamic_friction"], 0.001, 1, 0, 1]
                args.append((r["V"], r["F"], pose, params))
            t0 = time.perf_counter()
            for p, ar in zip(preps, args):
                clips += int(p.step(*ar)[1][0])
            serial += time.perf_counter() - t0
            t1 = time.perf_counter()
            par_clips += sum(ex.map(lambda k: int(preps[k].step(*args[k])[1][0]), range(len(preps))))
            par += time.perf_counter() - t1
            for p in preps:
                p.close()
            n += len(preps)
    return dict(value=clips / serial, unit="clips/s", cores=1, kind="port",
                sample="first %d of %d envs, %d clips, %.1f s serial, load-time BVH excluded" % (n, len(scs), clips,
                                                                                                    serial),
                parallel=dict(value=par_clips / par, cores=threads))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--envs", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--plain-envs", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--config", default="", help="fixed launch shape: blocks_per_cu,queries_per_lane,tet_max[,split_target[,queues]]")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (scale reference for the HBM "
                         "byte counters of a profiler) and exit")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = a.copy_calibration << 20
        src = torch.ones(n // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_bytes=n, copies=5)))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: the engine is single-device (multi-GPU sharding is design-only); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from mujoco_contact_surfaces_amd import Engine

    engine = Engine(0)
    log("building workload: %d envs" % a.envs)
    scs, handles, state = build_workload(engine, a.envs, a.seed)
    total_tets = int(sum(sc["soft"]["T"].shape[0] for sc in scs))
    total_tris = int(sum(sc["rigid"]["F"].shape[0] for sc in scs))
    batch = engine.batch(handles)
    batch.set_state(state)
    if a.config:
        batch.set_config(*[int(v) for v in a.config.split(",")])
        log("fixed config: %s" % batch.config())
    batch.step()
    W0, cnt = batch.get()
    clips = int(cnt[:, 0].sum())
    log("step counts: clips %d polygons %d faces %d" % tuple(cnt.sum(axis=0)))
    if not a.config and not a.no_tune:
        ms = batch.tune(reps=2)
        log("autotuned: %s (%.3f ms)" % (batch.config(), ms))

    for _ in range(a.warmup):
        batch.step()
    engine.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        batch.step()
    engine.synchronize()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    W1, cnt1 = batch.get()
    assert np.array_equal(cnt1, cnt)
    kt = batch.kernel_times(reps=max(3, a.steps // 4))
    log("kernel ms: %s" % kt)

    # plain per-pair GPU path: one synchronous hcs_contact_step per pair
    npl = min(a.plain_envs, a.envs)
    plain = None
    if npl > 0:
        pairs = [(ha, hb, sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]) for sc, (ha, hb) in zip(scs[:npl], handles)]
        for p in pairs[:8]:
            engine.contact_step([p])
        t1 = time.perf_counter()
        pclips = 0
        for p in pairs:
            _, c = engine.contact_step([p])
            pclips += int(c[0, 0])
        plain = pclips / (time.perf_counter() - t1)

    st = batch.item_stats()
    survivors = int(st[:, 6].sum())
    B, bparts = algorithmic_bytes(engine, scs, handles, cnt, survivors, st.shape[0])
    dom = "k_cand" if kt["cand"] >= kt["clip"] else "k_clip"
    k_ms = kt["cand"] if dom == "k_cand" else kt["clip"]
    achieved = B[dom] / (k_ms * 1e-3)
    traffic, tsrc = measured_traffic(dom, a.envs, a.seed)
    roof = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK,
                traffic=traffic, traffic_source=tsrc, kernel=dom, kernel_ms=k_ms, alg_bytes=B[dom],
                kernel_ms_all=kt, alg_bytes_all=B, **bparts)
    cpu = None if a.no_cpu else cpu_baseline(scs)
    value = clips / dt
    out = dict(metric=METRIC, value=value, unit="clips/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=dt * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
               data="synthetic",
               config=dict(workload="batched: %d envs, soft tets log-uniform 2k-100k, rigid tris log-uniform 12-20k, "
                                    "0.5-8 mm penetration, one hipGraph step" % a.envs,
                           envs=a.envs, total_tets=total_tets, total_tris=total_tris, clips_per_step=clips,
                           polygons_per_step=int(cnt[:, 1].sum()), faces_per_step=int(cnt[:, 2].sum()),
                           autotune=batch.config(), parallelism="single GPU"),
               us_per_step=dt * 1e6, plain_clips_per_s=plain, batched_over_plain=value / plain if plain else None,
               roofline=roof, cpu_baseline=cpu)
    if cpu:
        out["gpu_over_cpu_serial"] = value / cpu["value"]
        out["gpu_over_cpu_parallel"] = value / cpu["parallel"]["value"]
    batch.close()
    engine.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
