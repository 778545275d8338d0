#!/usr/bin/env python3
"""SPPM pass throughput on one MI355X.

A step is one full SPPM pass (trace → emit → hash/scan/sort → gather) of the
workload, with the scene already resident in HBM.  The metric is
photons+rays/s per pass = (pixels·1 camera ray + photons per pass) / T_pass.

Default workload: the 300k-triangle atrium (jittered columns and arches, GGX
roughness 0.05-0.9, 2048x1024 sky envmap with a 5e4 sun disk plus a matching
directional light for photons; 1920x1080, 2^22 photons/pass, grid 64).
--workload caustic runs the caustic stress scene (Cornell box + two
20480-triangle GGX spheres, 1024², 2^21 photons/pass, r0 = 0.01).

--workload adjoint512 measures gradients: forward + backward of a 32-pass
render of the eight-material room (512², 2^20 photons per pass) with respect
to 40 material parameters, reported as gradient evaluations per second.

--workload batch64 measures multi-scene passes instead: 64 generated scenes
(2k-200k triangles, 256²-1024², 2^16-2^21 photons per pass) rendered together,
reported as scene-passes/s (scenes × passes / s) beside the same scenes
rendered one after another with the one-scene driver.

Prints one JSON line.  With --gpus > 1 it exits: the pass runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import json
import os
import platform
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
# oracle passes timed for the CPU baseline per workload: the first passes of the
# seed schedule after pass 1 (heavy workloads time a stated subset)
CPU_PASSES = {"atrium": 8, "caustic": 12, "cornell": 48}   # about 10-15 s of oracle work on 16 host threads


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def gather_bytes(n_pixels: int, n_photons: int, n_buckets: int) -> int:
    """Algorithmic HBM bytes of one gather launch: per pixel 8 float4 + 1 f32
    read and 2 float4 + 1 f32 written (168 B), every stored photon record read
    once (24 B), cell start/end read once (8 B per bucket)."""
    return 168 * n_pixels + 24 * n_photons + 8 * n_buckets


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    f = ROOT / "profiles" / f"{kernel}_{workload}_pmc.txt"
    if not f.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


# VALU issue peak of one MI355X: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64
# instruction per SIMD every 4 cycles
VALU_PEAK_G = 256 * 4 * 2.4 / 4  # G wave-instructions/s (614.4)


def read_valu(kernel: str, workload: str):
    """VALU wave-instructions per launch from the committed SQ counter summary."""
    f = ROOT / "profiles" / f"{kernel}_{workload}_sq.txt"
    if not f.exists():
        return None
    m = re.search(r"valu_instructions_per_launch\s*[:=]\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def map_bytes(n_photons: int, n_buckets: int) -> dict:
    """Algorithmic HBM bytes of the photon-map stages: hash 20 B/photon (position
    in, key out) + 4 B/bucket (counts); scan 8 B/bucket; scatter 52 B/photon
    (key and 24-B record in, 24-B record out) + 8 B/bucket."""
    return {"hash": 20 * n_photons + 4 * n_buckets, "scan": 8 * n_buckets, "sort": 52 * n_photons + 8 * n_buckets}


def valu_roofline(instructions, ms: float):
    """VALU issue rate of a kernel from its counted wave-instructions per
    launch (committed SQ summary) and its measured time: the bound the gather
    actually runs against."""
    if not instructions:
        return None
    g = instructions / (ms * 1e-3) / 1e9
    return {"bound": "valu_issue", "instructions_per_launch": instructions, "achieved": g, "peak": VALU_PEAK_G,
            "unit": "G wave-instructions/s", "frac": g / VALU_PEAK_G,
            "source": "SQ_INSTS_VALU from profiles/k_gather_flat_<workload>_sq.txt, time from this run"}


def _stage_sequence(cfg, scene, passes: int, fast: bool) -> dict:
    from phosphor_amd.stages import StagePass

    sp = StagePass(cfg, scene, fast=fast)
    sp.run_pass()
    ms, n = [], []
    for _ in range(passes):
        sp.run_pass()
        ms.append(sp.stage_ms.copy())
        n.append(sp.last_n)
    names, accelerated = sp.stages, sp.accelerated
    sp.close()
    ms = np.mean(ms, axis=0)
    n_ph = int(np.mean(n))
    buckets = cfg.grid_res ** 3 + 1
    by = map_bytes(n_ph, buckets)
    by["map"] = by["hash"] + by["scan"] + by["sort"]
    by["gather"] = gather_bytes(cfg.width * cfg.height, n_ph, buckets)
    out = {"passes": passes, "ms_per_pass": float(ms.sum()),
           "stage_ms": {k: float(v) for k, v in zip(names, ms)}, "hbm": {}}
    if fast:
        out["pair_tree"] = accelerated
    for k in names:
        if k in by:
            t = float(ms[names.index(k)])
            gbs = by[k] / (t * 1e-3) / 1e9
            out["hbm"][k] = {"bytes": by[k], "ms": t, "GB_s": gbs, "frac": gbs / HBM_PEAK_GBPS}
    return out


def stage_abi(cfg, scene, passes: int) -> dict:
    """The reference's pass through the stage-level C ABI (phosphor_amd.stages),
    timed per stage with HIP events, in two wirings: the reference's kernels
    one for one (trace, emission in batches with the blocking count readback,
    hash, scan, sort, gather on one stream, the reference's median tree), and
    the drop-in's faster forms of the same stages (pair tree attached to the
    median tree, octant-ordered map in one call, octant gather).  HBM GB/s of
    the photon-map stages and the gather from their algorithmic bytes."""
    return {"reference_kernels": _stage_sequence(cfg, scene, passes, False),
            "fast_forms": _stage_sequence(cfg, scene, passes, True)}


def run_cpu_baseline(cfg, scene, passes: int, threads: int):
    import oracle

    oracle.lib().ref_set_num_threads(threads)
    ref = oracle.RefRenderer(cfg, scene)
    ref.next_seed()  # time pass 2 of the seed schedule, as the GPU median skips pass 1
    t0 = time.perf_counter()
    for _ in range(passes):
        rc = ref.run_pass()
        if rc != 0:
            raise RuntimeError(f"oracle pass failed: {rc}")
    dt = time.perf_counter() - t0
    stages = ref.stage_ms()
    ref.close()
    work = (cfg.width * cfg.height + cfg.photons_per_round) * passes
    return work / dt, dt, stages


def oracle_pass(P, cfg, scene, threads: int):
    """One pass of the oracle with ray counting, against one pass of the GPU
    renderer with the same configuration: scene_intersect calls of the pass
    (trace, emission) and the relative RMSE ‖L − L_ref‖₂ / ‖L_ref‖₂ of the
    GPU's linear estimate against the oracle's (0 when bit-identical)."""
    import oracle

    from phosphor_amd.output import relative_rmse

    oracle.lib().ref_set_num_threads(threads)
    ref = oracle.RefRenderer(cfg, scene, count_visits=True)
    if ref.run_pass() != 0:
        raise RuntimeError("oracle pass failed")
    rays = ref.rays()
    est_ref = ref.estimate()
    ref.close()
    r = P.Renderer(cfg, scene)
    r.run_pass()
    est = r.estimate()
    r.close()
    same = bool(np.array_equal(np.ascontiguousarray(est).view(np.uint32), np.ascontiguousarray(est_ref).view(np.uint32)))
    return int(rays[0]), int(rays[1]), relative_rmse(est, est_ref), same


def noise_rmse(P, cfg, scene, passes: int, ref_passes: int):
    """Relative RMSE of the linear estimate after `passes` passes against an
    independent render (seed + 1) of `ref_passes` passes: the estimator's
    noise at that pass count, not an error of the implementation."""

    from phosphor_amd.output import relative_rmse

    r = P.Renderer(dataclasses.replace(cfg, graph=True), scene)
    r.run(passes)
    est = r.estimate()
    r.close()
    ref = P.Renderer(dataclasses.replace(cfg, seed=cfg.seed + 1, graph=True), scene)
    ref.run(ref_passes)
    est_ref = ref.estimate()
    ref.close()
    return relative_rmse(est, est_ref)


def bench_batch(args) -> int:
    """64 scenes × `steps` passes as one batch, then the same scenes rendered
    one by one, each way both with graph replay and with pipelined passes
    (a pass's gather beside the next pass's trace and emission); all timed
    host-side between stream synchronisations, scene upload excluded.  The
    batch's value and the one-by-one baseline are each the faster mode."""
    import phosphor_amd as P
    from phosphor_amd import scenes

    t0 = time.perf_counter()
    work = scenes.batch_workload(64, seed=1)
    cfgs = [c for c, _ in work]
    scs = [s for _, s in work]
    n_px = sum(c.width * c.height for c in cfgs)
    n_ph = sum(c.photons_per_round for c in cfgs)
    print(f"[bench] batch64: {sum(s.n_triangles for s in scs)} triangles, {n_px} pixels, {n_ph} photons per pass "
          f"(generated in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    batch_s = {}
    for mode in ("graph", "pipelined"):
        t0 = time.perf_counter()
        b = P.Batch(cfgs, scs, graph=mode == "graph")
        print(f"[bench] batch ({mode}) created in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        for _ in range(args.warmup):
            b.run_pass()
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run_pass()
        b.synchronize()
        batch_s[mode] = time.perf_counter() - t0
        b.close()
        print(f"[bench] batch ({mode}): {batch_s[mode] * 1e3 / args.steps:.1f} ms/pass", file=sys.stderr, flush=True)
    plain_s = {}
    # a 256-byte copy (phx::k_copy16) before and after each scene's timed
    # passes, outside the timed region: it delimits the scenes in a kernel
    # trace, from which the fraction of each window with a kernel running
    # (the room batching has to fill) is read
    mark_src, mark_dst = C.c_void_p(), C.c_void_p()
    P.check(P.lib().phx_device_malloc(C.byref(mark_src), 256), "marker")
    P.check(P.lib().phx_device_malloc(C.byref(mark_dst), 256), "marker")

    def mark():
        P.check(P.lib().phx_device_copy(mark_dst, mark_src, 256, None), "marker")
        P.check(P.lib().phx_device_synchronize(), "marker")

    for mode in ("graph", "pipelined"):
        dt = 0.0
        for i, (c, s) in enumerate(work):
            r = P.Renderer(dataclasses.replace(c, graph=mode == "graph"), s)
            for _ in range(args.warmup):
                r.run_pass()
            r.synchronize()
            mark()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                r.run_pass()
            r.synchronize()
            dt += time.perf_counter() - t0
            mark()
            r.close()
            if i % 16 == 15:
                print(f"[bench] one-by-one ({mode}): {i + 1} scenes, {dt:.2f} s so far", file=sys.stderr,
                      flush=True)
        plain_s[mode] = dt
    batch_mode = min(batch_s, key=batch_s.get)
    plain_mode = min(plain_s, key=plain_s.get)
    dt_batch = batch_s[batch_mode]
    dt_plain = plain_s[plain_mode]
    value = 64 * args.steps / dt_batch
    line = {
        "metric": "batched SPPM scene-passes/s",
        "value": value,
        "unit": "scene-passes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_batch * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch64", "scenes": 64, "triangles": sum(s.n_triangles for s in scs),
                   "pixels_per_pass": n_px, "photons_per_pass": n_ph, "parallelism": "single-gpu"},
        "photons_rays_per_s": (n_px + n_ph) * args.steps / dt_batch,
        "batch_mode": batch_mode,
        "one_by_one": {"value": 64 * args.steps / dt_plain, "unit": "scene-passes/s",
                       "ms_per_step": dt_plain * 1e3 / args.steps, "mode": plain_mode},
        "modes_ms_per_step": {**{f"batch_{k}": v * 1e3 / args.steps for k, v in batch_s.items()},
                              **{f"one_by_one_{k}": v * 1e3 / args.steps for k, v in plain_s.items()}},
        "speedup_vs_one_by_one": dt_plain / dt_batch,
    }
    print(json.dumps(line))
    return 0


def bench_adjoint(args) -> int:
    """Full gradient evaluations (forward over 32 passes + backward) per second."""
    from phosphor_amd import PassConfig, scenes
    from phosphor_amd.adjoint import Adjoint, l2_loss

    res, rounds = 512, 32
    cfg = PassConfig(width=res, height=res, photons_per_round=1 << 20, max_photons_per_batch=1 << 20,
                     max_photon_bounces=16, grid_res=64, seed=1)
    scene = scenes.adjoint_scene(res, res, seed=1)
    adj = Adjoint(cfg, scene, max_rounds=rounds)
    th_true = adj.theta0()
    target = adj.forward(th_true, rounds)
    th = th_true.copy()
    th[:, :4] *= np.random.default_rng(7).uniform(0.8, 1.2, size=(th.shape[0], 4)).astype(np.float32)
    evals = max(1, args.steps // 4)
    for _ in range(max(1, args.warmup // 2)):
        img = adj.forward(th, rounds)
        adj.backward(l2_loss(img, target)[1])
    t_fwd = t_bwd = 0.0
    for _ in range(evals):
        t0 = time.perf_counter()
        img = adj.forward(th, rounds)
        t1 = time.perf_counter()
        grad = adj.backward(l2_loss(img, target)[1])
        t2 = time.perf_counter()
        t_fwd += t1 - t0
        t_bwd += t2 - t1
    adj.close()
    line = {
        "metric": "SPPM gradient evaluations/s",
        "value": evals / (t_fwd + t_bwd),
        "unit": "grads/s",
        "n_gpus": 1,
        "steps": evals,
        "warmup": max(1, args.warmup // 2),
        "ms_per_step": (t_fwd + t_bwd) * 1e3 / evals,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "adjoint512", "materials": int(th.shape[0]), "parameters": int(th.size),
                   "width": res, "height": res, "photons_per_pass": cfg.photons_per_round, "passes": rounds,
                   "parallelism": "single-gpu"},
        "forward_ms": t_fwd * 1e3 / evals,
        "backward_ms": t_bwd * 1e3 / evals,
        "grad_norm": float(np.linalg.norm(grad)),
    }
    print(json.dumps(line))
    return 0


def copy_calibration(mib: int) -> int:
    """Runs phx_device_copy over `mib` MiB three times: the known-traffic kernel
    (reads and writes exactly that many bytes) that scales the FETCH_SIZE and
    WRITE_SIZE counters of the profiled kernels."""
    import ctypes as C

    import phosphor_amd as P

    L = P.lib()
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    P.check(L.phx_device_malloc(C.byref(a), n), "malloc")
    P.check(L.phx_device_malloc(C.byref(b), n), "malloc")
    P.check(L.phx_memset_device(a, 1, n), "memset")
    t0 = time.perf_counter()
    for _ in range(3):
        P.check(L.phx_device_copy(b, a, n, None), "copy")
    P.check(L.phx_device_synchronize(), "sync")
    dt = time.perf_counter() - t0
    L.phx_device_free(a)
    L.phx_device_free(b)
    print(json.dumps({"metric": "device copy", "bytes_per_launch": n, "launches": 3,
                      "GB_per_s_read_plus_write": 2 * 3 * n / dt / 1e9}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="atrium", choices=["caustic", "cornell", "atrium", "batch64", "adjoint512"])
    ap.add_argument("--cpu-passes", type=int, default=-1,
                    help="oracle passes timed for the CPU baseline (0 = skip; default: 2 for the atrium, "
                         "4 for the caustic scene, 16 for the Cornell box)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0 = all cores of this host)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the known-size copy kernel (counter calibration) over MIB MiB")
    ap.add_argument("--noise-passes", type=int, default=0,
                    help="report the relative RMSE of an estimate after this many passes against an independent "
                         "render (0 = skip)")
    ap.add_argument("--noise-ref-passes", type=int, default=1024, help="passes of that independent render")
    ap.add_argument("--bvh-leaf-size", type=int, default=None,
                    help="triangles per leaf of the SAH tree (1-4; same results, default: the workload's)")
    ap.add_argument("--serial", action="store_true",
                    help="each pass's gather ends before the next pass starts (default: pipelined; same results)")
    ap.add_argument("--stage-passes", type=int, default=4,
                    help="passes after the timed region, each waited for, that give the stage timings (at least 1)")
    ap.add_argument("--abi-passes", type=int, default=2,
                    help="passes of the reference's sequence through the stage-level C ABI, timed per stage "
                         "(0 = skip)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(one process drives one GPU; a multi-GPU run would split the photons of a pass across GPUs)", file=sys.stderr)
        return 2

    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "adjoint512":
        return bench_adjoint(args)

    import phosphor_amd as P

    cfg, scene, _rounds = P.config_for(args.workload)
    if args.bvh_leaf_size is not None:
        cfg.bvh_leaf_size = args.bvh_leaf_size
    cfg.serial = args.serial
    n_px = cfg.width * cfg.height
    print(f"[bench] workload={args.workload} tris={scene.n_triangles} {cfg.width}x{cfg.height} "
          f"photons/pass={cfg.photons_per_round} grid={cfg.grid_res}", file=sys.stderr, flush=True)
    r = P.Renderer(cfg, scene)
    for _ in range(args.warmup):
        r.run_pass()
    r.synchronize()
    # timed region: passes enqueued back to back, no host synchronisation
    r.reset_gather_ms()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.run_pass()
    r.synchronize()
    dt = time.perf_counter() - t0
    # the gather kernel's own events over the timed region (the last 64 passes at most)
    g_timed = r.gather_ms(min(args.steps, 64))
    # stage timings: further passes, each waited for before the next is enqueued
    stage = []
    counts = []
    for i in range(max(1, args.stage_passes)):
        r.run_pass()
        stage.append(r.stage_ms())
        counts.append(r.last_photon_count)
    r.synchronize()
    stage = np.array(stage)
    ms_per_step = dt * 1e3 / args.steps
    value = (n_px + cfg.photons_per_round) * args.steps / dt
    mean_stage = stage.mean(axis=0)
    n_photons = int(np.mean([min(c, cfg.photons_per_round * cfg.max_photon_bounces) for c in counts]))
    buckets = cfg.grid_res ** 3 + 1
    gbytes = gather_bytes(n_px, n_photons, buckets)
    g_ms = float(mean_stage[3])
    achieved = gbytes / (g_ms * 1e-3) / 1e9
    timed_gather = None
    if len(g_timed):
        t_ms = float(np.mean(g_timed))
        t_ach = gbytes / (t_ms * 1e-3) / 1e9
        timed_gather = {"passes": int(len(g_timed)), "ms": t_ms, "achieved": t_ach, "frac": t_ach / HBM_PEAK_GBPS,
                        "note": "HIP events around the gather on its stream over the timed passes, where it runs "
                                "beside the next pass's trace and emission"}
    print(f"[bench] {ms_per_step:.2f} ms/pass, stages trace/emit/map/gather = "
          f"{', '.join(f'{x:.2f}' for x in mean_stage)} ms, deposits/pass={n_photons}", file=sys.stderr, flush=True)
    r.close()

    quality = {}
    if args.noise_passes > 0:
        t1 = time.perf_counter()
        rm = noise_rmse(P, cfg, scene, args.noise_passes, args.noise_ref_passes)
        quality["noise"] = {"relative_rmse": rm, "passes": args.noise_passes,
                            "reference_passes": args.noise_ref_passes, "reference": "independent render, seed + 1"}
        print(f"[bench] relative RMSE after {args.noise_passes} passes vs an independent "
              f"{args.noise_ref_passes}-pass render: {rm:.4f} ({time.perf_counter() - t1:.1f} s)",
              file=sys.stderr, flush=True)

    abi = None
    if args.abi_passes > 0:
        try:
            abi = stage_abi(cfg, scene, args.abi_passes)
            for name, seq in abi.items():
                print(f"[bench] stage-level ABI pass, {name}: {seq['ms_per_pass']:.2f} ms ("
                      + ", ".join(f"{k} {v:.2f}" for k, v in seq["stage_ms"].items()) + " ms); "
                      + ", ".join(f"{k} {v['GB_s']:.0f} GB/s" for k, v in seq["hbm"].items()), file=sys.stderr,
                      flush=True)
        except Exception as e:
            print(f"[bench] stage-level ABI pass skipped: {e}", file=sys.stderr)

    occupancy = None
    try:
        occupancy = {k: {"vgprs": v["vgprs"], "waves_per_simd": v["waves_per_simd"], "scratch_bytes": v["scratch_bytes"]}
                     for k, v in P.kernel_info().items()}
    except Exception as e:
        print(f"[bench] kernel occupancy not available: {e}", file=sys.stderr)
    cpu = None
    rays = None
    if args.cpu_passes < 0:
        args.cpu_passes = CPU_PASSES.get(args.workload, 1)
    if args.cpu_passes > 0:
        threads = args.cpu_threads or cpu_cores()
        try:
            n_tr, n_em, rm_oracle, same = oracle_pass(P, cfg, scene, threads)
            quality["vs_oracle"] = {"relative_rmse": rm_oracle, "bit_identical": same, "passes": 1}
            print(f"[bench] one pass against the oracle: relative RMSE {rm_oracle:.3g}, bit-identical {same}",
                  file=sys.stderr, flush=True)
            rays = {"counted_by": "oracle: scene_intersect calls of pass 1 (camera, shadow and photon rays; "
                                  "the GPU traces the same paths)",
                    "trace_per_pass": n_tr, "emit_per_pass": n_em,
                    "note": "the trace and emission stages run concurrently on two streams (and, pipelined, "
                            "beside the previous pass's gather); each stage's rate uses its own stream's time, "
                            "measured on passes run one at a time; rays_per_s uses the timed pass rate less "
                            "the gather's stage time",
                    "trace_rays_per_s": n_tr / (float(mean_stage[0]) * 1e-3),
                    "emit_rays_per_s": n_em / (float(mean_stage[1]) * 1e-3),
                    "rays_per_s": (n_tr + n_em) / ((ms_per_step - float(mean_stage[3])) * 1e-3)}
            print(f"[bench] rays per pass: trace {n_tr}, emission {n_em} -> "
                  f"{rays['rays_per_s']:.3e} rays/s in the trace + emission stages", file=sys.stderr, flush=True)
        except Exception as e:
            print(f"[bench] ray count skipped: {e}", file=sys.stderr)
        try:
            cpu_value, cpu_dt, cpu_stages = run_cpu_baseline(cfg, scene, args.cpu_passes, threads)
            cpu = {"value": cpu_value, "unit": "photons+rays/s", "cores": threads, "kind": "port",
                   "sample": f"{args.cpu_passes} full pass(es) of the {args.workload} workload "
                             f"({cpu_dt:.1f} s, stages {', '.join(f'{x:.0f}' for x in cpu_stages)} ms) "
                             f"on {cpu_model()}"}
            print(f"[bench] cpu oracle: {cpu_value:.3e} photons+rays/s over {cpu_dt:.1f}s on {threads} threads",
                  file=sys.stderr, flush=True)
        except Exception as e:  # the CPU leg is a reported baseline, not the measurement
            print(f"[bench] cpu baseline skipped: {e}", file=sys.stderr)

    line = {
        "metric": "SPPM photons+rays/s per pass",
        "value": value,
        "unit": "photons+rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "triangles": scene.n_triangles, "width": cfg.width,
                   "height": cfg.height, "photons_per_pass": cfg.photons_per_round,
                   "max_bounces": cfg.max_photon_bounces, "grid_res": cfg.grid_res,
                   "initial_radius": cfg.initial_radius, "direct_samples": cfg.direct_samples,
                   "bvh_leaf_size": cfg.bvh_leaf_size or 3, "passes": "serial" if cfg.serial else "pipelined",
                   "parallelism": "single-gpu"},
        "stage_ms": {"trace": float(mean_stage[0]), "emit": float(mean_stage[1]),
                     "hash_scan_sort": float(mean_stage[2]), "gather": float(mean_stage[3])},
        "roofline": {"kernel": "k_gather_flat", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": read_traffic("k_gather_flat", args.workload),
                     "valu": valu_roofline(read_valu("k_gather_flat", args.workload), g_ms),
                     "duration": {"ms": g_ms, "source": "HIP events around the gather on its stream in the "
                                  "stage-timing passes (the kernel alone)"},
                     "timed_region": timed_gather},
        "map_stage": {"kernels": "hash (octant keys) of the 32-byte deposit records, LSD radix sort whose last pass "
                                 "moves the records and writes their decoded directions, cell and octant starts",
                      "ms": float(mean_stage[2]),
                      "GB_s": sum(map_bytes(n_photons, buckets).values()) / (float(mean_stage[2]) * 1e-3) / 1e9,
                      "note": "algorithmic bytes of hash + scan + scatter over the whole map stage's time"},
        "stage_abi": abi,
        "rays": rays,
        "occupancy": occupancy,
        "quality": quality or None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
