"""Benchmark: Mrays/s of the plain tracer at 1080p on the 2M-triangle scene.

One step = one frame of the reference's per-frame path on config 2
(64 TRS instances of 5 seeded meshes, 2.0M unique BLAS triangles,
1920x1080, maxRecursions 4): TLAS build + traceRaysKHR equivalent
(raygen, wavefront traversal and closest-hit/miss recursion) + rough
approximation + composite + FXAA. BLAS builds happen once at scene load and
are not timed, as in the reference.

value = rays traced per frame / frame time, with inputs resident in HBM.
Timing: hipEvents recorded on the library's stream around exactly K steps.
The JSON line also carries the traversal kernel's roofline (algorithmic bytes
per ray from the oracle's node/triangle counts on the same trees) and a CPU
baseline: the C++ oracle (same LBVH + shading restatement, std::thread over
the host cores) timed on a bounded band of rows of the same frame.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# wave64 VALU instructions the GPU can issue per second without packed FP32:
# 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per instruction
VALU_PEAK_PER_S = 256 * 4 * 2.4e9 / 4
NODE_BYTES = 64   # compressed 4-wide node (rgrt::Node4q)
INST_BYTES = 64   # TLAS instance leaf (rgrt::DevInstance)
TRI_BYTES = 48    # rgrt::Tri
RAY_BYTES = 40    # queue entry: origin, tmin, dir, tmax, pixel, flags
HIT_BYTES = 20    # t, u, v, instance, primitive
GATHER_BYTES = 184  # closest-hit gathers: 12 B indices + 96 B vertices + 64 B material + 12 B offsets
# config 5 (4 samples per pixel, depth 6): far longer per-pixel ray chains
CONFIG5_CANDIDATES = {"wavefrontPasses": [24, 32, 48, 64], "traceBlocksPerCU": [4, 6, 8],
                      "drainBlocksPerCU": [2, 4, 8]}
# launch shapes the autotuner tries for the headline pipeline (coordinate descent)
TUNE_CANDIDATES = {"wavefrontPasses": [6, 8, 10, 12, 14], "traceBlocksPerCU": [4, 6, 7, 8],
                   "drainBlocksPerCU": [2, 4, 8], "shadeBlocksPerCU": [4, 8, 12], "frameBands": [1, 2]}
# One frame at a time the recursion tail is not hidden by the next frame:
# the frame traced in bands on separate streams (so one band's sparse last
# passes overlap another's work), fewer wavefront passes and a larger drain
# grid pay off.
ONE_FRAME_CANDIDATES = {"frameBands": [1, 2, 3, 4], "wavefrontPasses": [3, 4, 5, 6, 8, 10],
                        "drainBlocksPerCU": [2, 3, 4], "traceBlocksPerCU": [4, 6, 7], "shadeBlocksPerCU": [4, 8, 12]}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_valu(kernel, rays_per_launch, launch_ms):
    """The VALU-issue side of a kernel's roofline from the committed counter
    summary (SQ_INSTS_VALU per traced ray) at this run's rays and launch time:
    instructions per launch / (launch time x the issue peak); None without
    one."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_valu.json")
    try:
        with open(path) as f:
            per_ray = json.load(f)["valu_instructions_per_ray"]
    except (OSError, KeyError, ValueError):
        return None
    per_launch = per_ray * rays_per_launch
    return {"instructions_per_launch": round(per_launch), "peak_per_s": VALU_PEAK_PER_S,
            "frac": round(per_launch / (launch_ms * 1e-3) / VALU_PEAK_PER_S, 4), "source": f"profiles/{kernel}_valu.json"}


def measured_traffic(kernel, rays_per_launch):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per traced ray, scaled by the copy-kernel
    calibration), at this run's rays per launch; (None, None) without one."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return round(t["hbm_bytes_per_ray"] * rays_per_launch), f"profiles/{kernel}_traffic.json"
    except (OSError, KeyError, ValueError):
        return None, None


def copy_probe(mib):
    """A plain elementwise copy of `mib` MiB (reads and writes that many
    bytes): the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4, "launches": 3}), flush=True)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def main():
    from raygun_amd.hiptime import HipEvents

    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--tris-per-mesh", type=int, default=400_000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=0, help="rows of the frame the CPU baseline traces (0 = auto)")
    ap.add_argument("--no-fxaa", action="store_true")
    ap.add_argument("--mode", choices=("eager", "graph"), default="graph",
                    help="frame submission: eager calls, or rgrt_frame as one hipGraph launch")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (HIP's default is 4): the batched workload runs 32 "
                         "scene streams side by side, which 4 hardware queues would serialise")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shape instead of autotuning it for the measured pipeline")
    ap.add_argument("--tuning", type=str, default="",
                    help="launch shape as JSON (rgrt_tuning fields), used instead of autotuning (for profiling runs)")
    ap.add_argument("--only-chosen", action="store_true",
                    help="time only the chosen submission mode (for profiling runs)")
    ap.add_argument("--in-flight", type=int, choices=(1, 2, 3, 4, 5, 6), default=4,
                    help="frames in flight: N > 1 rotates frames over N contexts sharing the scene "
                         "(rgrt_create_shared), so one frame's tail overlaps the next frame's start")
    ap.add_argument("--batch-scenes", type=int, default=32,
                    help="scenes of the batched workload measured after the headline (0 = skip)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned path (training on traced frames, then learned vs plain at 1080p)")
    ap.add_argument("--learned-train-count", type=int, default=2000)
    ap.add_argument("--learned-train-steps", type=int, default=6000)
    ap.add_argument("--no-config5", action="store_true",
                    help="skip the config-5 measurement (3840x2160, 16M unique triangles, 4 spp, depth 6) after the headline")
    ap.add_argument("--config5-in-flight", type=int, choices=(1, 2, 3, 4, 5, 6), default=4,
                    help="frames in flight for the config-5 measurement")
    ap.add_argument("--profile-workload", choices=("config5", "batched", "denoise"), default="",
                    help="run only that secondary workload and print its JSON object (for profiler runs; "
                         "no headline line)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run only the counter-calibration copy kernel over MIB MiB")
    args = ap.parse_args()
    # read by the HIP runtime at its first call: set before anything touches the GPU
    os.environ["GPU_MAX_HW_QUEUES"] = str(max(1, min(args.hw_queues, 32)))
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.profile_workload:
        if args.profile_workload == "config5":
            out = bench_config5(max(2, args.steps), args.config5_in_flight, learned=False)
        elif args.profile_workload == "denoise":
            out = bench_denoise_inference(max(2, args.steps))
        else:
            out = bench_batched(args.batch_scenes or 32, max(2, args.steps))
        print(json.dumps(out), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only (the frame is one GPU's share); "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)

    from raygun_amd import autotune, procgen, rgrt
    from raygun_amd.hiptime import frames_ms
    from raygun_amd.scene import pack_models

    t0 = time.time()
    setup = procgen.config2(args.width, args.height, seed=args.seed, tris_per_mesh=args.tris_per_mesh)
    W, H = setup.width, setup.height
    log(f"[bench] scene generated in {time.time() - t0:.1f}s: {setup.info}")

    ctx = rgrt.Context(W, H, 0)
    models = setup.scene.models()
    vertices, indices, materials = pack_models(models)
    ctx.upload_geometry(vertices, indices)
    ctx.upload_materials(materials)
    t1 = time.time()
    blas_of = {}
    for m in models:
        k = id(m.mesh)
        if k not in blas_of:
            blas_of[k] = ctx.build_blas(m.mesh.first_vertex, m.mesh.index_offset_bytes, m.mesh.num_faces(), len(m.mesh.vertices))
        m.blas = blas_of[k]
    log(f"[bench] {len(blas_of)} BLAS built in {time.time() - t1:.2f}s (scene load, untimed)")
    instances = setup.scene.collect_instances()
    ubo = setup.ubo()
    use_fxaa = not args.no_fxaa

    ev = HipEvents()
    # further frames in flight: contexts on the same scene, each with its own
    # TLAS, images, wavefront state and stream
    ctxs = [ctx] + [ctx.shared() for _ in range(max(args.in_flight, 3) - 1)]

    def eager_on(c):
        def step():
            c.build_tlas(instances)
            c.set_frame(ubo)
            c.render(use_fxaa)
        return step

    def graph_on(c):
        return lambda: c.frame(ubo, instances, fxaa=use_fxaa, graph=True)

    def timed(fns, steps, warmup):
        return frames_ms(ctxs, fns, steps, warmup, ev)

    def name_of(mode, n):
        return mode if n == 1 else f"{mode}_{n}_in_flight"

    variants = {}
    for mode, make in (("eager", eager_on), ("graph", graph_on)):
        for n in (1, 2) if mode == "eager" else (1, 2, 3):
            variants[name_of(mode, n)] = [make(c) for c in ctxs[:n]]
    chosen = name_of(args.mode, args.in_flight)
    if chosen not in variants:
        variants[chosen] = [(eager_on if args.mode == "eager" else graph_on)(c) for c in ctxs[:args.in_flight]]
    tuned = None
    if args.tuning:
        t = dict(ctx.tuning())
        t.update(json.loads(args.tuning))
        for c in ctxs:
            c.set_tuning(**t)
        tuned = {"tuning": t, "given": True}
    elif not args.no_autotune:
        # launch shape for the measured pipeline, picked before the timed region
        t_tune = time.time()
        best, best_ms, table = autotune.tune(ctx, None, candidates=TUNE_CANDIDATES, contexts=ctxs,
                                             timer=lambda: timed(variants[chosen], 16, 3), confirm=3)
        tuned = {"tuning": best, "ms": round(best_ms, 4), "candidates_timed": len(table),
                 "seconds": round(time.time() - t_tune, 1)}
        log(f"[bench] autotuned {chosen}: {tuned}")
    frame_ms = {}
    for name, fns in variants.items():
        if name != chosen and not args.only_chosen:
            frame_ms[name] = round(timed(fns, max(6, args.steps // 2), 3), 4)
    ms_per_step = timed(variants[chosen], args.steps, args.warmup)
    frame_ms[chosen] = round(ms_per_step, 4)
    log(f"[bench] frame ms by submission mode: {frame_ms}")
    for c in ctxs[1:]:
        c.close()
    st = ctx.stats()
    rays = st["rays"]
    value = rays / (ms_per_step * 1e-3) / 1e6
    log(f"[bench] {args.steps} frames: {ms_per_step:.3f} ms/frame, {rays} rays/frame, {value:.1f} Mrays/s")

    # one frame at a time (each frame's graph starts when the previous frame
    # has finished, as the reference waits for its fence), with a launch
    # shape picked for that submission; the headline context gets its own
    # shape back afterwards
    one = None
    if not args.no_autotune:
        t_main = dict(ctx.tuning())
        one_fn = [graph_on(ctx)]
        b1, b1_ms, _ = autotune.tune(ctx, None, candidates=ONE_FRAME_CANDIDATES, contexts=[ctx], rounds=2,
                                     timer=lambda: frames_ms([ctx], one_fn, 20, 3, ev), confirm=3)
        one_ms = frames_ms([ctx], one_fn, args.steps, args.warmup, ev)
        one = {"ms_per_frame": round(one_ms, 4), "mrays_per_s": round(rays / (one_ms * 1e-3) / 1e6, 1), "tuning": b1}
        ctx.set_tuning(**t_main)
        log(f"[bench] one frame at a time: {one}")

    # per-stage and per-kernel timing of two more frames, alone on the GPU,
    # on a context with the timing events switched on (sharing the scene)
    tctx = ctx.shared(flags=rgrt.CFG_TIMING)
    tctx.set_tuning(**ctx.tuning())
    tctx.build_tlas(instances)
    tctx.set_frame(ubo)
    tctx.render(use_fxaa)
    tctx.build_tlas(instances)
    tctx.render(use_fxaa)
    ts = tctx.stats()
    passes = tctx.pass_counts()
    tctx.close()
    log(f"[bench] stages: {json.dumps({k: round(v, 4) if isinstance(v, float) else v for k, v in ts.items()})}")
    log(f"[bench] wavefront pass ray counts: {passes.tolist()}")

    # roofline of the traversal kernel: algorithmic bytes per ray from the
    # oracle's visit counts on the same trees (a seeded band of rows)
    roof = None
    cpu = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        osc = oracle_py.OracleScene(vertices, indices, materials)
        ob = {}
        for m in models:
            k = id(m.mesh)
            if k not in ob:
                ob[k] = osc.build_blas(m.mesh.first_vertex, m.mesh.index_offset_bytes, m.mesh.num_faces())
        osc.build_tlas(instances)
        cores = host_cores()
        # visit counts: every 27th row (40 rows of 1080p), all ray types in their frame proportions
        rows = list(range(13, H, max(1, H // 40)))
        vis = osc.render_counts(ubo, W, H, rows, threads=cores)
        r = vis["rays"]
        e_nodes, e_tris, p_hit = vis["nodes"] / r, vis["tris"] / r, vis["hits"] / r
        e_inst = vis["inst_leaves"] / r
        # traversal bytes of one ray: queue entry + hit record + wide nodes
        # (TLAS and BLAS) + instance leaves + triangles
        trav_bytes = RAY_BYTES + HIT_BYTES + NODE_BYTES * e_nodes + INST_BYTES * e_inst + TRI_BYTES * e_tris
        wf_rays = int(passes.sum())  # rays the wavefront traversal passes traced
        drain_rays = rays - wf_rays
        kernels = {
            "k_trace_queue": {"ms": ts["traceKernelMs"], "launches": max(1, ts["traceLaunches"]), "rays": wf_rays,
                              "bytes_per_ray": trav_bytes},
            # the drain kernel also does closest-hit shading: its gathers count
            "k_drain": {"ms": ts["drainKernelMs"], "launches": 1 if ts.get("drained") else 0, "rays": drain_rays,
                        "bytes_per_ray": trav_bytes + GATHER_BYTES * p_hit},
        }
        dom = max(kernels, key=lambda k: kernels[k]["ms"] if kernels[k]["launches"] else -1.0)
        roof = None
        for name in (dom,) + tuple(k for k in kernels if k != dom):
            kk = kernels[name]
            if not kk["launches"] or kk["ms"] <= 0:
                continue
            avg_ms = kk["ms"] / kk["launches"]
            alg_per_launch = kk["rays"] * kk["bytes_per_ray"] / kk["launches"]
            achieved = alg_per_launch / (avg_ms * 1e-3) / 1e9
            traffic, tsrc = measured_traffic(name, kk["rays"] / kk["launches"])
            obj = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                   "kernel": name, "avg_launch_ms": round(avg_ms, 5), "launches_per_frame": kk["launches"],
                   "rays_per_launch": round(kk["rays"] / kk["launches"]), "bytes_per_ray": round(kk["bytes_per_ray"], 1),
                   "algorithmic_bytes_per_launch": round(alg_per_launch)}
            if tsrc:
                obj["traffic_source"] = tsrc
            valu = measured_valu(name, kk["rays"] / kk["launches"], avg_ms)
            if valu:
                obj["valu_issue"] = valu
            if roof is None:
                roof = obj
                roof.update({"node_bytes": NODE_BYTES, "inst_bytes": INST_BYTES, "tri_bytes": TRI_BYTES,
                             "E_nodes": round(e_nodes, 2), "E_inst_leaves": round(e_inst, 3), "E_tris": round(e_tris, 2),
                             "P_hit": round(p_hit, 4), "others": []})
            else:
                roof["others"].append(obj)
        log(f"[bench] roofline {roof}")
        if not args.no_cpu_baseline:
            # a band of rows sized for ~15 s of wall time on these cores,
            # from a quick 16-row probe of the oracle's rate
            y0 = H // 2 - 8
            tc = time.time()
            probe = osc.render(ubo, W, H, threads=cores, rows=(y0, y0 + 16))
            rate = probe["rays"] / max(time.time() - tc, 1e-3)
            band = args.cpu_rows or int(min(H, max(16, 16 * (15.0 * rate / max(probe["rays"], 1)))))
            y0 = max(0, H // 2 - band // 2)
            tc = time.time()
            out = osc.render(ubo, W, H, threads=cores, rows=(y0, y0 + band))
            dt = time.time() - tc
            cpu = {"value": round(out["rays"] / dt / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
                   "sample": f"rows {y0}..{y0 + band - 1} of the 1920x1080 frame ({out['rays']} rays, {dt:.1f}s wall), "
                             "CPU oracle (C++ restatement of the same trees and shading, not lavapipe)"}
            log(f"[bench] cpu baseline {cpu}")
    except Exception as e:  # the GPU number stands without the CPU side
        log(f"[bench] oracle side failed: {e!r}")

    result = {
        "metric": "Mrays/s @1080p 2M-tri",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural meshes)",
        "config": {"workload": "config2: 64 instances of 5 seeded meshes, 1920x1080, maxRecursions 4, 1 spp, FXAA",
                   "unique_triangles": setup.info["unique_triangles"],
                   "instanced_triangles": setup.info["instanced_triangles"],
                   "rays_per_frame": rays, "sky_lookups_per_frame": st["skyLookups"],
                   "width": W, "height": H, "seed": args.seed,
                   "stage_ms": {k: round(ts[k], 4) for k in ("asBuildMs", "rtOnlyMs", "roughMs", "postprocMs", "rtTotalMs",
                                                             "traceKernelMs", "shadeKernelMs", "drainKernelMs")},
                   "tuning": ctx.tuning(), "wavefront_pass_rays": passes.tolist(),
                   "submission": chosen, "frames_in_flight": args.in_flight, "frame_ms_by_submission": frame_ms,
                   "autotune": tuned, "one_frame_at_a_time": one},
        "roofline": roof,
        "cpu_baseline": cpu,
        "learned_vs_plain": {"value": None, "unit": "x"},
    }
    ctx.close()
    if not args.no_learned:
        try:
            result["learned_vs_plain"] = bench_learned(args.learned_train_count, args.learned_train_steps)
        except Exception as e:  # the headline stands without it
            log(f"[bench] learned path failed: {e!r}")
            result["learned_vs_plain"]["error"] = repr(e)
    if not args.no_config5:
        try:
            result["config5"] = bench_config5(max(4, args.steps // 6), args.config5_in_flight, learned=not args.no_learned)
        except Exception as e:  # the headline stands without it
            log(f"[bench] config5 failed: {e!r}")
    if args.batch_scenes:
        try:
            result["batched"] = bench_batched(args.batch_scenes, max(3, args.steps // 4))
        except Exception as e:  # the headline stands without the batched figures
            log(f"[bench] batched workload failed: {e!r}")
    print(json.dumps(result), flush=True)


LEARNED_SPP = (4, 8, 16)  # sample counts of the learned path's input
PSNR_TARGET_DB = 38.0


def bench_learned(train_count=2000, train_steps=6000, train_res=256, crop=128, eval_seeds=(60000, 60001, 60002)):
    """The metric's second half: frame ms of the learned path (stochastic
    frame + the denoiser) against the plain tracer's at equal PSNR, at
    1920x1080 on held-out scenes of the config-1 family, PSNR against each
    scene's 256-spp plain frame. Per scene the learned path uses the fewest
    samples of LEARNED_SPP whose denoised frame reaches PSNR_TARGET_DB (the
    most if none does; `psnr_target_met` says which); the plain tracer's ms
    at that PSNR is interpolated over its own sample sweep. The network is
    trained first on frames the tracer renders (train_count seeded scenes at
    train_res^2, inputs of 4, 8 and 16 samples in turn, 256-spp targets,
    random crop^2 windows), on the GPU in a child process (PyTorch fp32; the
    library's own backward kernels are switched off); inference runs on the
    library's bf16 MFMA kernels (rgrt_denoise)."""
    import math
    from raygun_amd import denoise as D
    from raygun_amd import rgrt
    t0 = time.time()
    ctx = rgrt.Context(train_res, train_res, 0)
    den = rgrt.Denoiser(ctx, seed=1)
    scenes = D.GlossyScenes(ctx, train_res, train_res)
    first, last, t_render, t_train = D.train_network(ctx, den, scenes, train_count, spp_in=LEARNED_SPP, steps=train_steps,
                                                     batch=16, lr=1e-3, crop=crop)
    params = den.get_params()
    den.close()
    ctx.close()
    log(f"[bench] learned: trained on {train_count} scenes ({t_render:.1f}s rendering, {t_train:.1f}s training, "
        f"loss {first:.5f} -> {last:.5f})")
    W, H = 1920, 1080
    ctx = rgrt.Context(W, H, 0)
    den = rgrt.Denoiser(ctx, seed=1)
    den.set_params(params)
    sc = D.GlossyScenes(ctx, W, H)
    res = D.learned_vs_plain(ctx, den, sc, list(eval_seeds), spp_learned=LEARNED_SPP[0], spp_more=LEARNED_SPP[1:])
    den.close()
    ctx.close()
    chosen = []
    for r in res:
        opts = [{"spp": LEARNED_SPP[0], "psnr_learned": r["psnr_learned"], "learned_ms": r["learned_ms"],
                 "ratio_at_equal_psnr": r["ratio_at_equal_psnr"]}] + r["learned_more"]
        pick = next((o for o in opts if o["psnr_learned"] >= PSNR_TARGET_DB), opts[-1])
        chosen.append({"seed": r["seed"], "spp": pick["spp"], "psnr_learned": pick["psnr_learned"],
                       "psnr_target_met": pick["psnr_learned"] >= PSNR_TARGET_DB, "learned_ms": pick["learned_ms"],
                       "ratio_at_equal_psnr": pick["ratio_at_equal_psnr"]})
    ratios = [c["ratio_at_equal_psnr"] for c in chosen]
    ratios4 = [r["ratio_at_equal_psnr"] for r in res]
    den_ms = float(np.mean([r["denoise_ms"] for r in res]))
    flops = 2.0 * D.MACS_PER_PIXEL * W * H
    out = {
        "value": round(math.exp(float(np.mean(np.log(ratios)))), 2),
        "unit": "x (plain-tracer ms at the learned frame's PSNR / learned ms, geometric mean over scenes; per scene "
                "the fewest learned samples reaching the PSNR target)",
        "workload": "config4: config-1 glossy scene family at 1920x1080, 4/8/16 spp stochastic + denoiser vs the plain "
                    "tracer; PSNR on clamped RGB against each scene's 256-spp plain frame",
        "psnr_target_db": PSNR_TARGET_DB,
        "chosen": chosen,
        "psnr_target_met": sum(c["psnr_target_met"] for c in chosen),
        "value_at_4spp": round(math.exp(float(np.mean(np.log(ratios4)))), 2),
        "psnr_learned_db": [r["psnr_learned"] for r in res],
        "psnr_noisy_4spp_db": [r["psnr_noisy"] for r in res],
        "learned_ms": [r["learned_ms"] for r in res],
        "reference_256spp_ms": [r["reference_ms"] for r in res],
        "ratio_vs_256spp": [r["ratio_vs_reference"] for r in res],
        "scenes": res,
        "inference": {"ms": round(den_ms, 3), "gflop": round(flops / 1e9, 1),
                      "achieved_tflops": round(flops / (den_ms * 1e-3) / 1e12, 1), "peak_tflops": 2500.0,
                      "frac": round(flops / (den_ms * 1e-3) / 2.5e15, 4), "dtype": "bf16 MFMA, fp32 accumulate",
                      "note": "whole rgrt_denoise call (pack, 11 convolutions, pooling, tile check, output)"},
        "training": {"scenes": train_count, "resolution": train_res, "crop": crop, "steps": train_steps, "batch": 16,
                     "input_spp": list(LEARNED_SPP),
                     "loss_first": round(first, 6), "loss_last": round(last, 6), "render_s": round(t_render, 1),
                     "train_s": round(t_train, 1),
                     "how": "PyTorch fp32 on the GPU in a child process; the library's own backward kernels are "
                            "switched off (their one GPU run faulted, no cause shown)"},
        "seconds": round(time.time() - t0, 1),
    }
    log(f"[bench] learned vs plain {out['value']}x ({out['psnr_target_met']} of {len(chosen)} scenes at >= {PSNR_TARGET_DB} dB: "
        f"{[(c['spp'], c['psnr_learned']) for c in chosen]}); at 4 spp {out['value_at_4spp']}x, PSNR {out['psnr_learned_db']} "
        f"(noisy {out['psnr_noisy_4spp_db']})")
    return out


def bench_denoise_inference(reps):
    """The denoiser's inference alone on a 1920x1080 4-spp stochastic frame
    of the config-1 family, `reps` times (for kernel profiles: the network's
    cost does not depend on its parameter values, so the library's own
    initialisation is loaded as the parameters)."""
    from raygun_amd import denoise as D
    from raygun_amd import rgrt
    from raygun_amd.hiptime import time_ms
    W, H = 1920, 1080
    ctx = rgrt.Context(W, H, 0)
    den = rgrt.Denoiser(ctx, seed=1)
    den.set_params(den.get_params())
    sc = D.GlossyScenes(ctx, W, H)
    ubo = sc.scene(60000, 4)
    for _ in D.trace_mean(ctx, ubo, 4, 1, 1):
        pass
    ms = time_ms(ctx, lambda: den.denoise(1e30, 1e30), reps=reps, warmup=1)
    flops = 2.0 * D.MACS_PER_PIXEL * W * H
    den.close()
    ctx.close()
    return {"workload": "denoiser inference, 1920x1080", "reps": reps, "ms": round(ms, 3), "gflop": round(flops / 1e9, 1),
            "achieved_tflops": round(flops / (ms * 1e-3) / 1e12, 1)}


def bench_config5_learned(ctx, setup, train_views=300, train_steps=2000, res=128):
    """Config 5's learned frame: 4 spp (stochastic sampling) + the denoiser
    on config 5's own 3840x2160 view, against the plain tracer at equal PSNR,
    PSNR against the 256-spp plain frame. The network is trained for this
    scene family: seeded views of another config-5 scene (seed 11, 150k
    triangles per mesh) at res^2, 4-spp inputs and 128-spp targets, on the
    GPU in a child process (PyTorch fp32); inference on the library's MFMA
    kernels."""
    from raygun_amd import denoise as D
    from raygun_amd import procgen, rgrt
    t0 = time.time()
    tctx = rgrt.Context(res, res, 0)
    tden = rgrt.Denoiser(tctx, seed=1)
    views = D.SceneViews(tctx, procgen.config5(res, res, seed=11, tris_per_mesh=150_000), res, res)
    first, last, t_render, t_train = D.train_network(tctx, tden, views, train_views, first_seed=5000, spp_ref=128,
                                                     steps=train_steps, batch=16, lr=1e-3)
    params = tden.get_params()
    tden.close()
    tctx.close()
    log(f"[bench] config5 learned: trained on {train_views} views ({t_render:.1f}s rendering, {t_train:.1f}s training, "
        f"loss {first:.5f} -> {last:.5f})")
    den = rgrt.Denoiser(ctx, seed=1)
    try:
        den.set_params(params)
        ev = D.SceneViews(ctx, setup, setup.width, setup.height, fixed_camera=True, upload=False)
        r = D.learned_vs_plain(ctx, den, ev, [5], sweep=(4, 8, 16, 32, 64))[0]
    finally:
        den.close()
        ctx.set_sampling(rgrt.SAMPLING_WHITTED)
    flops = 2.0 * D.MACS_PER_PIXEL * setup.width * setup.height
    r["inference"] = {"gflop": round(flops / 1e9, 1), "achieved_tflops": round(flops / (r["denoise_ms"] * 1e-3) / 1e12, 1),
                      "frac": round(flops / (r["denoise_ms"] * 1e-3) / 2.5e15, 4)}
    r["training"] = {"views": train_views, "resolution": res, "steps": train_steps, "target_spp": 128,
                     "loss_first": round(first, 6), "loss_last": round(last, 6), "render_s": round(t_render, 1),
                     "train_s": round(t_train, 1)}
    r["seconds"] = round(time.time() - t0, 1)
    log(f"[bench] config5 learned: PSNR {r['psnr_learned']} (noisy {r['psnr_noisy']}), {r['learned_ms']} ms, "
        f"{r['ratio_at_equal_psnr']}x at equal PSNR")
    return r


def bench_config5(frames, in_flight=4, learned=True):
    """Config 5's plain tracer: 3840x2160, 8 seeded displaced meshes of ~2M
    triangles (16M unique), 256 instances, one in five refractive,
    maxRecursions 6, 4 samples per pixel (the SSAA table), FXAA; frames as
    hipGraphs, `in_flight` frames in flight on contexts sharing the scene.
    The learned denoiser and the batch with 8 lower-resolution scenes that
    config 5 also names are not part of this figure."""
    from raygun_amd import autotune, procgen, rgrt
    from raygun_amd.hiptime import HipEvents, frames_ms
    from raygun_amd.scene import pack_models
    t0 = time.time()
    setup = procgen.config5()
    models = setup.scene.models()
    v, i, m = pack_models(models)
    log(f"[bench] config5 scene generated in {time.time() - t0:.1f}s: {setup.info}")
    ctx = rgrt.Context(setup.width, setup.height, 0)
    ctxs = [ctx]
    try:
        ctx.upload_geometry(v, i)
        ctx.upload_materials(m)
        t1 = time.time()
        blas_of = {}
        for mo in models:
            k = id(mo.mesh)
            if k not in blas_of:
                blas_of[k] = ctx.build_blas(mo.mesh.first_vertex, mo.mesh.index_offset_bytes, mo.mesh.num_faces(),
                                            len(mo.mesh.vertices))
            mo.blas = blas_of[k]
        build_s = time.time() - t1
        ctxs += [ctx.shared() for _ in range(in_flight - 1)]
        inst = setup.scene.collect_instances()
        ubo = setup.ubo()
        fns = [(lambda c=c: c.frame(ubo, inst, fxaa=True, graph=True)) for c in ctxs]
        ev = HipEvents()
        t2 = time.time()
        best, best_ms, table = autotune.tune(ctx, None, candidates=CONFIG5_CANDIDATES, contexts=ctxs,
                                             timer=lambda: frames_ms(ctxs, fns, 3, 1, ev), confirm=3)
        log(f"[bench] config5 autotuned in {time.time() - t2:.1f}s: {best} {best_ms:.2f} ms")
        ms = frames_ms(ctxs, fns, frames, 2, ev)
        rays = ctx.stats()["rays"]
        out = {"workload": "config5: 3840x2160, 8 displaced meshes (16M unique triangles), 256 instances, 20% refractive "
                           "(IOR U[1.3,1.7]), maxRecursions 6, 4 spp (SSAA table), FXAA",
               "unique_triangles": setup.info["unique_triangles"], "instanced_triangles": setup.info["instanced_triangles"],
               "blas_build_s": round(build_s, 3), "frames": frames, "frames_in_flight": in_flight, "tuning": best,
               "rays_per_frame": rays, "ms_per_frame": round(ms, 3), "mrays_per_s": round(rays / (ms * 1e-3) / 1e6, 1)}
        log(f"[bench] config5: {out}")
        try:
            out["with_8_scenes"] = bench_config5_batch(ctx, ubo, inst, max(3, frames // 2), ev)
        except Exception as e:  # the config-5 figure stands without it
            log(f"[bench] config5 batch failed: {e!r}")
        if learned:
            try:
                out["learned"] = bench_config5_learned(ctx, setup)
            except Exception as e:  # the config-5 figure stands without it
                log(f"[bench] config5 learned path failed: {e!r}")
        return out
    finally:
        for c in ctxs:
            c.close()


def bench_config5_batch(ctx, ubo, inst, frames, ev):
    """Config 5's frame submitted together with 8 lower-resolution scenes of
    the batched family (320x180 .. 960x540, 10k-1M triangles, animated, one
    deforming) in one rgrt_batch_frame (a hipGraph per scene on its own
    stream): ms per batch frame against the config-5 frame alone (one frame
    at a time), and the batch's rays per second."""
    from raygun_amd import batched, rgrt
    w = batched.BatchWorkload(num_scenes=8, seed=7, lo=10_000, hi=1_000_000,
                              resolutions=[(320, 180), (480, 270), (640, 360), (960, 540)])
    batch = rgrt.Batch([ctx] + [s.ctx for s in w.scenes], rgrt.BATCH_GRAPH_PER_SCENE)
    try:
        stream = batch.stream()

        def frame(f, t):
            for s in w.scenes:
                if s.deforms and f:
                    s.deform(f)
            batch.frame([(ubo, inst, True, True)] + [(s.ubo, s.instances(t), True, f > 0) for s in w.scenes])

        frame(0, 0.0)
        frame(1, 0.05)
        batch.synchronize()

        def timed(fn, n):
            a, b = ev.create(), ev.create()
            ctx.synchronize()
            ev.record(a, stream)
            for k in range(n):
                fn(k)
            ev.record(b, stream)
            ms = ev.elapsed_ms(a, b) / n
            batch.synchronize()
            ev.destroy(a)
            ev.destroy(b)
            return ms

        ms_batch = timed(lambda k: frame(2 + k, 0.05 * (2 + k)), frames)
        rays5 = ctx.stats()["rays"]
        rays8 = sum(s.ctx.stats()["rays"] for s in w.scenes)
        from raygun_amd.hiptime import time_ms
        ms_alone = time_ms(ctx, lambda: ctx.frame(ubo, inst, fxaa=True, refit=True, graph=True), reps=frames, warmup=1,
                           events=ev)
        out = {"scenes": [[s.width, s.height, s.mesh.num_faces()] for s in w.scenes],
               "rays_per_batch_frame": rays5 + rays8, "ms_per_batch_frame": round(ms_batch, 3),
               "ms_config5_frame_alone": round(ms_alone, 3),
               "mrays_per_s_batch": round((rays5 + rays8) / (ms_batch * 1e-3) / 1e6, 1)}
        log(f"[bench] config5 with 8 scenes: {out}")
        return out
    finally:
        batch.close()
        w.close()


def bench_batched(num_scenes, frames):
    """Config 3: num_scenes scenes of log-normal size (10k-4M triangles) and
    mixed resolution, animated; per-scene autotuned launch shapes; batched
    frames (a hipGraph per scene on its own stream, and one hipGraph with a
    branch per scene) against the same frames submitted scene by scene."""
    from raygun_amd import autotune, batched, rgrt
    from raygun_amd.hiptime import HipEvents
    t0 = time.time()
    w = batched.BatchWorkload(num_scenes=num_scenes, seed=3)
    log(f"[bench] batched workload: {num_scenes} scenes, {sum(w.info['triangles'])} triangles, "
        f"{w.info['pixels']} pixels, built in {time.time() - t0:.1f}s")
    try:
        cand = {"wavefrontPasses": [2, 4, 6, 8], "traceBlocksPerCU": [4, 6, 8]}
        t1 = time.time()
        tunings = []
        for sc in w.scenes:
            inst = sc.instances(0.0)
            best, ms, _ = autotune.tune(sc.ctx, lambda: sc.ctx.frame(sc.ubo, inst, fxaa=True), candidates=cand, reps=1)
            tunings.append(best)
        log(f"[bench] autotuned {num_scenes} scenes in {time.time() - t1:.1f}s")
        ev = HipEvents()
        bstream = w.batch.stream()
        w.frame_batched(0.0)  # first frame builds the TLASes
        w.synchronize()

        def timed_batched(mode, t_first):
            w.batch.set_mode(mode)
            w.frame_batched(t_first)  # warm: captures and instantiates the graphs of this mode
            w.synchronize()
            a, b = ev.create(), ev.create()
            ev.record(a, bstream)  # every scene's stream is idle here
            for f in range(frames):
                w.frame_batched(t_first + 0.05 * (f + 1))
            ev.record(b, bstream)  # the batch's stream is ordered after every scene
            ms = ev.elapsed_ms(a, b) / frames
            w.synchronize()
            ev.destroy(a)
            ev.destroy(b)
            return ms

        ms_per_scene = timed_batched(rgrt.BATCH_GRAPH_PER_SCENE, 0.0)
        rays = w.rays()
        ms_one_graph = timed_batched(rgrt.BATCH_ONE_GRAPH, 1.0)
        # one graph tracing all scenes in shared launches: the batch's launch
        # shape is the first scene's tuning, picked here for the batch
        s0 = w.scenes[0]
        own = s0.ctx.tuning()
        persistent = {}
        for passes in (4, 6, 8, 10):
            for blocks in (6, 8):
                s0.ctx.set_tuning(wavefrontPasses=passes, traceBlocksPerCU=blocks)
                persistent[(passes, blocks)] = timed_batched(rgrt.BATCH_PERSISTENT, 3.0 + 0.1 * passes + blocks)
        best_shape = min(persistent, key=persistent.get)
        s0.ctx.set_tuning(wavefrontPasses=best_shape[0], traceBlocksPerCU=best_shape[1])
        ms_persistent = timed_batched(rgrt.BATCH_PERSISTENT, 5.0)

        def latency(mode, t_first):
            """ms from submission to completion of one batch frame at a time."""
            w.batch.set_mode(mode)
            w.frame_batched(t_first)
            w.synchronize()
            a, b = ev.create(), ev.create()
            total = 0.0
            for f in range(frames):
                w.synchronize()
                ev.record(a, bstream)
                w.frame_batched(t_first + 0.05 * (f + 1))
                ev.record(b, bstream)
                total += ev.elapsed_ms(a, b)
            w.synchronize()
            ev.destroy(a)
            ev.destroy(b)
            return total / frames

        lat = {"persistent": latency(rgrt.BATCH_PERSISTENT, 6.0)}
        s0.ctx.set_tuning(**own)
        lat["graph_per_scene"] = latency(rgrt.BATCH_GRAPH_PER_SCENE, 7.0)
        lat["one_graph"] = latency(rgrt.BATCH_ONE_GRAPH, 8.0)
        w.batch.set_mode(rgrt.BATCH_GRAPH_PER_SCENE)
        t2 = time.perf_counter()
        for f in range(frames):
            w.frame_sequential(2.0 + 0.05 * (f + 1))
        w.synchronize()
        ms_seq_wall = (time.perf_counter() - t2) * 1e3 / frames
        log(f"[bench] batched ms per frame: graph per scene {ms_per_scene:.3f}, one graph {ms_one_graph:.3f}, "
            f"persistent {ms_persistent:.3f} (passes, blocks per CU {best_shape}), scene by scene (eager, wall) "
            f"{ms_seq_wall:.3f}; one batch frame at a time {lat}")
        return {"workload": f"config3: {num_scenes} scenes, triangles log-normal in [10k, 4M], resolutions 320x180.."
                            "1920x1080, animated instances (TLAS refit), one scene in four deforming (device-resident "
                            "poses, full BLAS rebuild)",
                "scenes": num_scenes, "triangles": w.info["triangles"], "resolutions": w.info["resolutions"],
                "pixels_per_frame": w.info["pixels"], "rays_per_frame": rays, "frames": frames,
                "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                "ms_per_batched_frame": round(ms_per_scene, 3),
                "mrays_per_s_batched": round(rays / (ms_per_scene * 1e-3) / 1e6, 1),
                "ms_per_batched_frame_one_graph": round(ms_one_graph, 3),
                "ms_per_batched_frame_persistent": round(ms_persistent, 3),
                "persistent_shape": {"wavefrontPasses": best_shape[0], "traceBlocksPerCU": best_shape[1]},
                "persistent_sweep_ms": {f"{p}x{b}": round(m, 3) for (p, b), m in persistent.items()},
                "ms_one_batch_frame_at_a_time": {k: round(v, 3) for k, v in lat.items()},
                "ms_per_sequential_frame_wall": round(ms_seq_wall, 3),
                "autotuned": tunings}
    finally:
        w.close()


if __name__ == "__main__":
    main()
