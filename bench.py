"""Benchmark: batched FP64 hydroelastic contact steps on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--envs 4096]

Headline workload (synthetic, seeded): N = 4096 independent environments,
each a soft tet body (tet count log-uniform over 2k-100k; sphere, ellipsoid
or capsule) against a rigid mesh (triangle count log-uniform over 12-20k;
box, plane grid or icosphere) at 0.5-8 mm penetration.  One step = poses and
twists in (already resident in HBM), per-environment net wrenches out: LBVH
traversal, candidate compaction, clipping, face evaluation,
Hunt-Crossley/friction force law and deterministic reductions, replayed as
one hipGraph.

Metric: tet-tri clips per second, a clip being one element of the canonical
candidate set (tet and triangle AABBs overlap in the soft frame), over the
median of >= 50 replays timed with HIP events (the wall clock over the K
timed steps is reported beside it).  The same JSON line carries:
  * roofline of the dominant kernel (the candidate kernel) from the
    algorithmic bytes of its units, and of the whole step;
  * the per-call ("plain") path on the same environments, and the other
    configurations: soft sphere on a box (us per step), tactile pad with
    16 x 16 taxels (us per step), 1024 batched pads each with its 16 x 16
    sensor (ms per step), 256 soft-soft pairs (tet-tet clips/s);
  * the CPU baseline: the oracle's per-step work on this host (serial and on
    the host's CPU share), whose results are compared env by env with the
    GPU's (counts bit-exact, wrenches within 1e-10).
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
MIN_REPLAYS = 50


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return dict(cpu_model=model, nproc=os.cpu_count(), affinity=aff)


def host_threads():
    # the CPU share of one GPU on the GPU host is 16 threads
    return max(1, min(16, host_info()["affinity"] or 1))


def algorithmic_bytes(engine, scs, handles, counts):
    """Algorithmic bytes of one step, each distinct element read once and
    each output written once:
      rigid meshes       24 B per vertex + 12 B per index triple
      candidate tets     144 B per distinct tet (16 B indices, 4 x 24 B
                         vertices, 4 x 8 B pressures)
      candidate tris     84 B per distinct triangle (12 B indices, 3 x 24 B
                         moved vertices)
      candidate keys     16 B per candidate (8 B written, 8 B read)
      polygons           48 B per non-empty polygon (area, centroid, p0, g)
      per env            384 B (two poses, two twists in, the wrench out)
    The candidate kernel's units: the rigid meshes, the candidate tets and
    triangles, the key writes (8 B per candidate) and the poses (192 B per
    env)."""
    T_C = R_C = 0
    for sc, (ha, hb) in zip(scs, handles):
        keys = engine.candidates((int(ha), int(hb), sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]))
        ntri = sc["rigid"]["F"].shape[0]
        T_C += np.unique(keys // ntri).size
        R_C += np.unique(keys % ntri).size
    rigid = sum(24 * sc["rigid"]["V"].shape[0] + 12 * sc["rigid"]["F"].shape[0] for sc in scs)
    E = len(scs)
    C = int(counts[:, 0].sum())
    P = int(counts[:, 1].sum())
    step = rigid + 144 * T_C + 84 * R_C + 16 * C + 48 * P + 384 * E
    cand = rigid + 144 * T_C + 84 * R_C + 8 * C + 192 * E
    return dict(step=step, k_cand=cand, rest=step - cand), dict(rigid_bytes=rigid, distinct_tets=T_C,
                                                                distinct_tris=R_C, candidates=C, polygons=P, envs=E)


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


def wrench_err(got, ref):
    """Largest error of the four 3-vectors of a wrench relative to the
    vector's own norm plus 1e-4 of the reference force (so the bar is
    1e-10 |r| + 1e-14 |F|: torques about symmetry planes can cancel)."""
    scale = np.linalg.norm(ref[:3])
    e = 0.0
    for b in range(4):
        d = np.linalg.norm(got[3 * b:3 * b + 3] - ref[3 * b:3 * b + 3])
        if d:
            e = max(e, d / (np.linalg.norm(ref[3 * b:3 * b + 3]) + 1e-4 * scale + 1e-300))
    return e


def cpu_baseline(scs, W_gpu, cnt_gpu, budget_s=12.0, chunk=128):
    """The oracle's per-step work (load-time tet records and BVH excluded)
    timed on this host, chunk by chunk over every environment: each chunk is
    prepared (untimed, on the host threads), stepped serially while the
    serial leg is under budget_s, stepped on the host threads (timed), every
    result compared with the GPU's, then freed, so host memory stays
    bounded."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from concurrent.futures import ThreadPoolExecutor

    threads = host_threads()
    clips = par_clips = n = ns = 0
    serial = par = 0.0
    mism, wfail, worst = [], 0, 0.0

    def params(sc):
        s, r = sc["soft"], sc["rigid"]
        pose = np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]])
        return (r["V"], r["F"], pose, [s["props"]["hydroelastic_modulus"], s["props"]["dissipation"],
                                      s["props"]["dynamic_friction"], r["props"]["dynamic_friction"], 0.001, 1, 0,
                                      int(s["geom_id"] < r["geom_id"])])

    with ThreadPoolExecutor(threads) as ex:
        while n < len(scs):
            part = scs[n:n + chunk]
            preps = list(ex.map(lambda sc: O.PreparedSoft(sc["soft"]["V"], sc["soft"]["T"], sc["soft"]["p"]), part))
            args = [params(sc) for sc in part]
            if serial < budget_s:
                t0 = time.perf_counter()
                for p, ar in zip(preps, args):
                    clips += int(p.step(*ar)[1][0])
                serial += time.perf_counter() - t0
                ns += len(preps)
            t1 = time.perf_counter()
            res = list(ex.map(lambda k: preps[k].step(*args[k]), range(len(preps))))
            par += time.perf_counter() - t1
            for k, (Wr, cr) in enumerate(res):
                e = n + k
                par_clips += int(cr[0])
                if not np.array_equal(cnt_gpu[e], cr):
                    mism.append(e)
                    continue
                err = wrench_err(W_gpu[e], Wr)
                worst = max(worst, err)
                wfail += err > 1e-10
            for p in preps:
                p.close()
            n += len(preps)
            if n % 1024 == 0 or n == len(scs):
                log("cpu baseline: %d/%d envs" % (n, len(scs)))
    info = host_info()
    out = dict(value=par_clips / par, unit="clips/s", cores=threads, kind="port",
               sample="every env of the workload (%d envs, %d clips) on %d host threads; load-time tet records "
                      "and BVH excluded" % (n, par_clips, threads),
               cpu_model=info["cpu_model"], nproc=info["nproc"], affinity=info["affinity"],
               threads_note="16 threads: the GPU host's CPU share for one GPU",
               serial=dict(value=clips / serial if serial else None, cores=1,
                           sample="first %d envs, %d clips, %.1f s" % (ns, clips, serial)))
    parity = dict(envs_checked=int(n), count_mismatches=len(mism), first_mismatches=[int(e) for e in mism[:8]],
                  wrench_over_1e10=int(wfail), max_wrench_rel_err=float(worst), ok=bool(not mism and wfail == 0))
    return out, parity


def median_ms(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts))


def extra_configs(engine, n_soft_soft=256, n_tactile=1024):
    """The other configurations: per-call latency of a soft sphere on a box
    and of the tactile pad with its 16 x 16 taxels, batched tactile pads with
    their sensors, and batched soft-soft throughput."""
    from mujoco_contact_surfaces_amd import scenes as S
    out = {}
    sc = S.sphere_box(0)
    s, r = sc["soft"], sc["rigid"]
    ha = engine.add_soft_mesh(s["V"], s["T"], s["p"], s["props"], s["geom_id"])
    hb = engine.add_rigid_mesh(r["V"], r["F"], r["props"], r["geom_id"])
    pair = [(ha, hb, sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"])]
    for _ in range(5):
        engine.contact_step(pair)
    W, cnt = engine.contact_step(pair)
    out["sphere_box"] = dict(us_per_step=1e3 * median_ms(lambda: engine.contact_step(pair), 200),
                             tets=int(s["T"].shape[0]), tris=int(r["F"].shape[0]), clips=int(cnt[0, 0]),
                             polygons=int(cnt[0, 1]), path="plain per-call hcs_contact_step, host planning included")
    sc = S.tactile_pad(0)
    s, r = sc["soft"], sc["rigid"]
    ha = engine.add_soft_mesh(s["V"], s["T"], s["p"], s["props"], s["geom_id"])
    hb = engine.add_rigid_mesh(r["V"], r["F"], r["props"], r["geom_id"])
    pair = [(ha, hb, sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"])]
    grid = dict(sc["grid"], geom=hb, X_W_sensor=sc["X_WB"])

    def pad_step():
        engine.contact_step(pair, keep_polygons=True)
        engine.taxels([grid])

    for _ in range(3):
        pad_step()
    W, cnt = engine.contact_step(pair, keep_polygons=True)
    out["tactile_pad"] = dict(us_per_step=1e3 * median_ms(pad_step, 50), tets=int(s["T"].shape[0]),
                              tris=int(r["F"].shape[0]), clips=int(cnt[0, 0]), polygons=int(cnt[0, 1]),
                              taxels=256, rays_per_taxel=400,
                              path="plain per-call hcs_contact_step + hcs_taxels, host planning included")
    # the differentiation workload's forward half: pads from its generator
    # (<= 20k tets, <= 2k triangles), each with a 16 x 16 sensor, contact
    # step and taxels of every env in one graph replay
    t0 = time.time()
    handles4, state4, grids4 = [], [], []
    for e in range(n_tactile):
        sc = S.differentiation_env(0, e)
        s, r = sc["soft"], sc["rigid"]
        ha = engine.add_soft_mesh(s["V"], s["T"], s["p"], s["props"], s["geom_id"])
        hb = engine.add_rigid_mesh(r["V"], r["F"], r["props"], r["geom_id"])
        handles4.append((ha, hb))
        state4.append(np.concatenate([sc["X_WA"], sc["X_WB"], sc["V_WA"], sc["V_WB"]]))
        grids4.append(dict(sc["grid"], geom=hb))
    b = engine.batch(handles4)
    b.set_state(np.array(state4))
    b.step()
    W, cnt = b.get()
    b.tune(reps=1)
    ms_contact = float(np.median(b.step_times(MIN_REPLAYS)))
    b.set_sensors(list(range(n_tactile)), grids4)
    b.step()
    W2, cnt2 = b.get()
    tx = b.taxels()
    assert np.array_equal(cnt, cnt2)
    ms = float(np.median(b.step_times(MIN_REPLAYS)))
    out["tactile_batch"] = dict(envs=n_tactile, taxels=int(sum(t.size for t in tx)), rays_per_taxel=400,
                                clips_per_step=int(cnt[:, 0].sum()), polygons_per_step=int(cnt[:, 1].sum()),
                                ms_per_step=ms, us_per_env_step=1e3 * ms / n_tactile, ms_contact_only=ms_contact,
                                autotune=b.config(), setup_s=time.time() - t0,
                                path="batched graph replay: contact step + flat sensors of every env")
    b.close()
    t0 = time.time()
    scs, handles, state = S.upload_soft_soft(engine, n_soft_soft, seed=0)
    b = engine.batch(handles)
    b.set_state(state)
    b.step()
    W, cnt = b.get()
    b.tune(reps=1)
    b.step()
    W2, cnt2 = b.get()
    assert np.array_equal(cnt, cnt2)
    ms = float(np.median(b.step_times(MIN_REPLAYS)))
    clips = int(cnt[:, 0].sum())
    out["soft_soft"] = dict(pairs=n_soft_soft, clips_per_step=clips, polygons_per_step=int(cnt[:, 1].sum()),
                            ms_per_step=ms, clips_per_s=clips / (ms * 1e-3),
                            tets=int(sum(sc["soft"]["T"].shape[0] + sc["soft_b"]["T"].shape[0] for sc in scs)),
                            autotune=b.config(), setup_s=time.time() - t0)
    b.close()
    out["differentiation"] = dict(value=None, reason="the tangent/adjoint path is switched off in this build")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--envs", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--plain-envs", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the sphere-box, tactile-pad and soft-soft lines")
    ap.add_argument("--config", default="", help="fixed launch shape: blocks_per_cu,queries_per_lane,tet_max[,split_target[,queues[,order]]]")
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
    from mujoco_contact_surfaces_amd import Engine, scenes as S

    engine = Engine(0)
    log("building workload: %d envs" % a.envs)
    t0 = time.time()
    scs, handles, state = S.upload_batched(engine, a.envs, a.seed, log=lambda m: log(m, "(%.0f s)" % (time.time() - t0)))
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
    wall = (time.perf_counter() - t0) / a.steps
    W1, cnt1 = batch.get()
    assert np.array_equal(cnt1, cnt)
    replays = batch.step_times(max(MIN_REPLAYS, a.steps))
    dt = float(np.median(replays)) * 1e-3
    kt = batch.kernel_times(reps=MIN_REPLAYS)
    log("median replay %.3f ms, wall mean %.3f ms, kernel ms: %s" % (dt * 1e3, wall * 1e3, kt))

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

    B, parts = algorithmic_bytes(engine, scs, handles, cnt)
    traffic, tsrc = measured_traffic("k_cand", a.envs, a.seed)
    k_ms = kt["cand"]
    roof = dict(bound="hbm", achieved=B["k_cand"] / (k_ms * 1e-3) / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                frac=B["k_cand"] / (k_ms * 1e-3) / HBM_PEAK, traffic=traffic, traffic_source=tsrc, kernel="k_cand",
                kernel_ms=k_ms, alg_bytes=B["k_cand"], kernel_ms_all=kt, **parts)
    step_roof = dict(achieved=B["step"] / dt / 1e9, unit="GB/s", frac=B["step"] / dt / HBM_PEAK, alg_bytes=B["step"],
                     ms=dt * 1e3)
    extra = None if a.no_extra else extra_configs(engine)
    cpu, parity = (None, None) if a.no_cpu else cpu_baseline(scs, W0, cnt)
    value = clips / dt
    out = dict(metric=METRIC, value=value, unit="clips/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=dt * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
               data="synthetic",
               config=dict(workload="batched: %d envs, soft tets log-uniform 2k-100k, rigid tris log-uniform 12-20k, "
                                    "0.5-8 mm penetration, one hipGraph step" % a.envs,
                           envs=a.envs, total_tets=total_tets, total_tris=total_tris, clips_per_step=clips,
                           polygons_per_step=int(cnt[:, 1].sum()), faces_per_step=int(cnt[:, 2].sum()),
                           autotune=batch.config(), parallelism="single GPU"),
               timing=dict(replays=len(replays), median_ms=dt * 1e3, min_ms=float(np.min(replays)),
                           max_ms=float(np.max(replays)), wall_mean_ms=wall * 1e3, timed_steps=a.steps),
               us_per_step=dt * 1e6, plain_clips_per_s=plain, batched_over_plain=value / plain if plain else None,
               roofline=roof, step_roofline=step_roof, configs=extra, cpu_baseline=cpu, parity=parity)
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
        if cpu["serial"]["value"]:
            out["gpu_over_cpu_serial"] = value / cpu["serial"]["value"]
    batch.close()
    engine.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
