"""Throughput of the HIP DrawPath pipeline on one MI355X.

Workload: curves + strokes, 4,096 seeded paths (8-200 quad/cubic/conic
segments, control points uniform in 2048^2, conic weights 0.2-4, 50% even-odd,
20% stroked at width 1-16 px with miter/round/bevel joins and butt/round
caps) drawn src-over into a 2048x2048 premultiplied RGBA8 surface.

One step = clear the surface + draw the whole prepared scene (stroke
outlines, flatten, sort, analytic coverage walk, tile composite), with the
scene geometry already resident in device memory.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "paths/s + Mpix/s filled on 1 MI355X (batched+learned vs plain; % HBM roofline)"
HBM_PEAK_GBS = 8000.0
EDGE_STATE_BYTES = 32     # x, y, dx, dy, upper_x, upper_y, lower_y + winding/shift
PLANE_BYTES = 256         # one 16x16 tile of 8-bit coverage
WORKLOAD = 2


def algorithmic_bytes(stage, st, geom_bytes, canvas_px):
    """Algorithmic bytes per launch of one pipeline stage."""
    e, planes, u = st["edges"], st["tile_planes"], st["pixels"]
    return {
        "flatten_count": geom_bytes,
        "flatten": geom_bytes + EDGE_STATE_BYTES * e,
        "sort": 2 * EDGE_STATE_BYTES * e,
        "walk_count": EDGE_STATE_BYTES * e,
        "walk": EDGE_STATE_BYTES * e + PLANE_BYTES * planes,
        "composite": 8 * u + PLANE_BYTES * planes,
    }.get(stage)


def log(msg):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def geometry_bytes(sc):
    """G of the scene: 8 B per point, 1 B per verb, 4 B per conic weight, a
    32-byte draw record per path and 20 B per gradient stop (the scene has
    no canvas matrix, so no 24-byte matrix records)."""
    s = sc.scene.contents
    n = s.n_paths
    stops = 0
    for p in range(n):
        g = s.paints[p].gradient
        if g >= 0:
            stops += s.gradients[g].color_count
    return (8 * s.point_offsets[n] + s.verb_offsets[n] + 4 * s.weight_offsets[n] + 32 * n
            + 20 * stops)


def frame_alg_bytes(sc, st, blur_bytes=0):
    """Algorithmic bytes of one frame: G + 8 U + B (G the scene's geometry and
    paints, U the distinct target pixels the frame touches, each read and
    written once, B 16 bytes per pixel of every blurred layer)."""
    return geometry_bytes(sc) + 8 * st["pixels"] + blur_bytes


PMC_SUMMARY = os.path.join(HERE, "profiles", "pmc_curves_strokes_4096.json")
STAGE_KERNELS = {"walk": ("hipr::WalkKernel<4, 1>", "hipr::WalkKernel<8, 1>",
                          "hipr::WalkKernel<16, 1>"),
                 "walk_count": ("hipr::WalkKernel<4, 0>", "hipr::WalkKernel<8, 0>",
                                "hipr::WalkKernel<16, 0>"),
                 "composite": ("hipr::CompositeKernel",), "flatten": ("hipr::FlattenKernel<1>",),
                 "sort": ("hipr::SortGatherKernel",)}


def measured_traffic(stage, form):
    """HBM bytes of one launch of a stage (all its kernels in the frame of
    `form`), from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    this workload (counter values scaled by the copy-kernel calibration), or
    None."""
    if not os.path.exists(PMC_SUMMARY) or stage not in STAGE_KERNELS:
        return None
    with open(PMC_SUMMARY) as fh:
        js = json.load(fh)
    disp = js.get("frames", {}).get(form, [])
    hits = [d for d in disp if d["kernel"] in STAGE_KERNELS[stage]]
    if not hits:
        return None
    return int(sum(d["fetch_mib"] + d["write_mib"] for d in hits) * 1024 * 1024)


def _probe_scene(workload):
    """(scene, width, height, owner) of a --pmc-probe workload: a generator
    workload number, "tiger" (tiger.skp under Scale(4.1) on 4096^2, as the
    tiger leg draws it) or "text" (the text leg's 2048^2 page)."""
    if workload == "tiger":
        from skity_amd import picture
        rec = picture.Recording(open(TIGER_SKP, "rb").read(), 4096, 4096,
                                (4.1, 0, -560, 0, 4.1, 60))
        if rec.scene is None:
            raise RuntimeError(f"tiger.skp not drawable: {rec.rc} {rec.error}")
        return rec.scene, 4096, 4096, rec
    if workload == "text":
        sys.path.insert(0, os.path.join(HERE, "tests"))
        import text_kit
        made = text_kit.text_page(3)
        if made is None:
            raise RuntimeError("text probe: no TrueType font installed")
        return made[0].scene, 2048, 2048, made[0]
    from skity_amd import scenes
    sc = scenes.generate(int(workload))
    return sc.scene, sc.width, sc.height, sc


def pmc_probe(workload=WORKLOAD):
    """The program the committed counter summaries profile: one exact frame
    and one learned frame of the workload (curves + strokes by default; 3 =
    gradients + blur; "tiger"; "text") with the walk passes launched
    directly, so every kernel is its own dispatch, then a 256 MiB
    device-to-device copy whose counters calibrate FETCH_SIZE / WRITE_SIZE."""
    import torch
    from skity_amd import abi, device, presizer
    scene, width, height, owner = _probe_scene(workload)
    ctx = device.Context(0)
    prep = device.Prepared(ctx, scene)
    surf = device.Surface(ctx, width, height)
    prm, margin = presizer.load()
    presizer.install(ctx, prm, margin)
    surf.clear(0)
    surf.draw(prep, mode=abi.MODE_EXACT, flags=abi.FLAG_NO_GRAPH)
    surf.clear(0)
    surf.draw(prep, mode=abi.MODE_LEARNED, flags=abi.FLAG_NO_GRAPH)
    surf.finish()
    a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del owner
    print("pmc probe done", flush=True)


def cpu_baseline(sc, budget_s):
    """The oracle (a single-threaded CPU restatement of skity's software
    backend) on a bounded prefix of the same scene."""
    from oracle import oracle_py as O
    n_total = sc.n_paths
    chunk = 32
    done, t_used = 0, 0.0
    img = np.zeros((sc.height, sc.width, 4), np.uint8)
    while done < n_total and t_used < budget_s:
        part = sc.slice(done, min(done + chunk, n_total))
        t0 = time.perf_counter()
        img = O.render(part.scene, sc.width, sc.height, img)
        t_used += time.perf_counter() - t0
        done += part.n_paths
    return {"value": done / t_used, "unit": "paths/s", "cores": 1, "kind": "port",
            "sample": f"first {done} of the {n_total} paths drawn in order on the "
                      f"{sc.width}x{sc.height} target, {t_used:.1f} s, 1 thread"}


def device_info(torch):
    p = torch.cuda.get_device_properties(0)
    return {"name": p.name, "compute_units": p.multi_processor_count,
            "arch": getattr(p, "gcnArchName", "")}


def run_form(surf, prep, mode, steps, warmup, torch):
    """Times `steps` frames of one form; returns wall seconds, per-stage
    totals (HIP events) and the last frame's statistics."""
    for _ in range(warmup):
        surf.clear(0)
        surf.draw(prep, mode=mode)
    torch.cuda.synchronize()
    stage_tot = {}
    fallbacks = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        surf.clear(0)
        surf.draw(prep, mode=mode)
        for k, (ms, n) in surf.stage_times().items():
            a = stage_tot.setdefault(k, [0.0, 0])
            a[0] += ms
            a[1] += n
        fallbacks += surf.stats()["fallbacks"]
    surf.finish()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = surf.stats()
    st["fallbacks_per_frame"] = fallbacks / steps
    return wall, stage_tot, st


def blur_work(sc):
    """Per blurred draw of a scene (fills): filter-bitmap area A_f and radius
    r, as HandleFilter / ApproximateFilteredBounds size it; returns
    (algorithmic FLOPs = sum 2 passes x 4 ch x 2(2r+1) x A_f, bytes = sum 16 A_f)."""
    s = sc.scene.contents
    flops, nbytes = 0, 0
    for p in range(s.n_paths):
        pt = s.paints[p]
        if pt.blur_radius <= 0:
            continue
        p0, p1 = s.point_offsets[p], s.point_offsets[p + 1]
        v0, v1 = s.verb_offsets[p], s.verb_offsets[p + 1]
        if v1 - v0 > 1 and s.verbs[v1 - 1] == 0:
            p1 -= 1
        xs = np.array([s.points[2 * i] for i in range(p0, p1)], np.float32)
        ys = np.array([s.points[2 * i + 1] for i in range(p0, p1)], np.float32)
        r = np.float32(pt.blur_radius)
        w = int(np.ceil(xs.max() + r) - np.floor(xs.min() - r))
        h = int(np.ceil(ys.max() + r) - np.floor(ys.min() - r))
        rad = min(int(np.floor(float(r) + 0.5)), 254)  # std::round of a positive radius
        if w <= 0 or h <= 0 or rad <= 1:
            continue
        flops += 2 * 4 * 2 * (2 * rad + 1) * w * h
        nbytes += 16 * w * h
    return flops, nbytes


def blur_leg(ctx, torch, steps):
    """Gradients + blur: 1,024 cubic paths with linear/radial/sweep gradients
    and normal mask-filter blur (sigma 1-32) on 4096^2; the blur runs as
    MFMA banded products."""
    from skity_amd import abi, device, scenes
    sc = scenes.generate(3)
    prep = device.Prepared(ctx, sc.scene)
    surf = device.Surface(ctx, sc.width, sc.height)
    wall, stage_tot, st = run_form(surf, prep, abi.MODE_EXACT, steps, 1, torch)
    flops, nbytes = blur_work(sc)
    blur_ms = stage_tot.get("blur", [0.0, 0])[0] / steps
    out = {"workload": "gradients+blur: 1024 cubic paths, 2-16 stop gradients, blur sigma 1-32, "
                       "4096x4096", "paths": sc.n_paths,
           "paths_per_s": round(sc.n_paths / (wall / steps), 1),
           "ms_per_step": round(wall * 1e3 / steps, 3),
           "mpix_per_s": round(st["blends"] / (wall / steps) / 1e6, 1),
           "blur_ms_per_step": round(blur_ms, 3),
           "blur_alg_tflops": round(flops / (blur_ms * 1e-3) / 1e12, 3) if blur_ms else None,
           "blur_alg_gbs": round(nbytes / (blur_ms * 1e-3) / 1e9, 1) if blur_ms else None,
           "blur_alg_flops": flops, "blur_alg_bytes": nbytes}
    surf.close()
    prep.close()
    return out


TIGER_SKP = os.path.join(HERE, "tests", "golden", "skp", "tiger.skp")


def tiger_leg(ctx, torch, steps, side=4096, scale=4.1):
    """A real scene: the reference's tiger.skp (305 DrawPath fills and
    strokes of cubic outlines under two CONCAT44 matrices) read, played back
    into a display list under Scale(4.1) and drawn on a 4096^2 surface;
    exact and learned forms.  The timed region is the device frame of the
    recorded scene (reading and recording the SKP are host work done once)."""
    from skity_amd import abi, device, picture
    t0 = time.perf_counter()
    rec = picture.Recording(open(TIGER_SKP, "rb").read(), side, side,
                            (scale, 0, -560, 0, scale, 60))
    if rec.scene is None:
        raise RuntimeError(f"tiger.skp not drawable: {rec.rc} {rec.error}")
    record_ms = (time.perf_counter() - t0) * 1e3
    n = rec.scene.contents.n_paths
    prep = device.Prepared(ctx, rec.scene)
    surf = device.Surface(ctx, side, side)
    out = {"workload": f"tiger.skp (305 paths) played back through a display list, "
                       f"Scale({scale}) on {side}x{side} RGBA8",
           "paths": n, "read_and_record_ms": round(record_ms, 2)}
    imgs = {}
    for name, mode in (("exact", abi.MODE_EXACT), ("learned", abi.MODE_LEARNED)):
        wall, _, st = run_form(surf, prep, mode, steps, 2, torch)
        imgs[name] = surf.read()
        out[name] = {"paths_per_s": round(n / (wall / steps), 1),
                     "ms_per_step": round(wall * 1e3 / steps, 3),
                     "mpix_per_s": round(st["blends"] / (wall / steps) / 1e6, 1),
                     "fallbacks_per_frame": st["fallbacks_per_frame"]}
    out["learned_equals_exact"] = bool(np.array_equal(imgs["exact"], imgs["learned"]))
    out["covered_pixels"] = int((imgs["exact"][..., 3] > 0).sum())
    surf.close()
    prep.close()
    rec.close()
    return out


def text_leg(ctx, torch, steps):
    """Text as paths: a 2048^2 page of 14 px DejaVu Sans glyph runs (about
    17,000 glyphs, headings larger, rotated and stroked) drawn as glyph runs
    (DrawGlyphs); exact and learned forms.  Glyph outlines come from the
    TrueType file installed with the system; skipped without one."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import text_kit
    from skity_amd import abi, device
    made = text_kit.text_page(3)
    if made is None:
        return {"skipped": "no TrueType font installed"}
    sc, glyphs = made
    prep = device.Prepared(ctx, sc.scene)
    surf = device.Surface(ctx, 2048, 2048)
    out = {"workload": "text page: 14 px DejaVu Sans glyph runs with larger, rotated and "
                       "stroked headings, 2048x2048 RGBA8",
           "glyphs": glyphs, "draws": sc.scene.contents.n_paths}
    imgs = {}
    for name, mode in (("exact", abi.MODE_EXACT), ("learned", abi.MODE_LEARNED)):
        wall, _, st = run_form(surf, prep, mode, steps, 2, torch)
        imgs[name] = surf.read()
        out[name] = {"glyphs_per_s": round(glyphs / (wall / steps), 1),
                     "ms_per_step": round(wall * 1e3 / steps, 3),
                     "fallbacks_per_frame": st["fallbacks_per_frame"]}
    out["learned_equals_exact"] = bool(np.array_equal(imgs["exact"], imgs["learned"]))
    surf.close()
    prep.close()
    return out


def batched_leg(ctx, torch, steps, n_canvases=64, workload=4, tune_canvases=8):
    """Batched uneven scenes at the family's stated scale (workload 4): n
    canvases of side 256-4096, each with 100-50,000 closed outlines of
    4-10,000 line/quad/cubic segments (30% gradient paints), drawn per canvas
    (plain: one exact pipeline run per canvas) and as one batch (learned
    pre-sizing, work split autotuned on the first canvases)."""
    from skity_amd import abi, device, scenes
    t0 = time.perf_counter()
    scs = []
    for i in range(n_canvases):
        scs.append(scenes.generate(workload, seed=scenes.SEED_BASE + 400 + i))
        if (i + 1) % 8 == 0:
            log(f"batched leg: generated {i + 1} canvases, {time.perf_counter() - t0:.0f} s")
    t_gen = time.perf_counter() - t0
    preps = []
    for sc in scs:
        preps.append(device.Prepared(ctx, sc.scene))
    log(f"batched leg: {len(preps)} scenes prepared")
    surfs = [device.Surface(ctx, sc.width, sc.height) for sc in scs]
    n_paths = sum(sc.n_paths for sc in scs)
    n_segs = sum(sc.scene.contents.verb_offsets[sc.n_paths] for sc in scs)
    tune = device.Batch(ctx)
    for su, pr in list(zip(surfs, preps))[:tune_canvases]:
        tune.add(su, pr)
    tuned = tune.autotune(abi.MODE_LEARNED, repeats=1)
    tune.close()
    log(f"batched leg: autotuned on {tune_canvases} canvases: {tuned}")
    batch = device.Batch(ctx)
    for su, pr in zip(surfs, preps):
        batch.add(su, pr)
    batch.set_split(*tuned)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        for su in surfs:
            su.finish()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    def plain():
        for i, (su, pr) in enumerate(zip(surfs, preps)):
            su.clear(0)
            su.draw(pr, mode=abi.MODE_EXACT)
            if (i + 1) % 16 == 0:
                su.finish()
                log(f"batched leg: plain, {i + 1} canvases drawn")

    def batched():
        for su in surfs:
            su.clear(0)
        batch.draw(mode=abi.MODE_LEARNED)

    t_plain = timed(plain)
    log(f"batched leg: plain {t_plain:.2f} s per frame")
    plain_frames = [su.read() for su in surfs]
    t_batch = timed(batched)
    log(f"batched leg: batched + learned {t_batch:.2f} s per frame")
    st = batch.stats()
    same = all(np.array_equal(su.read(), f) for su, f in zip(surfs, plain_frames))
    if not same:
        print("bench: the batched frame differs from the per-canvas frames", file=sys.stderr,
              flush=True)
        sys.exit(3)
    out = {"workload": f"batched uneven: {n_canvases} canvases 256^2-4096^2, 100-50,000 "
                       "outlines of 4-10,000 line/quad/cubic segments each",
           "canvases": n_canvases, "paths": n_paths, "segments": int(n_segs),
           "generation_s": round(t_gen, 1),
           "plain_paths_per_s": round(n_paths / t_plain, 1),
           "plain_ms_per_step": round(t_plain * 1e3, 3),
           "batched_learned_paths_per_s": round(n_paths / t_batch, 1),
           "batched_learned_ms_per_step": round(t_batch * 1e3, 3),
           "speedup": round(t_plain / t_batch, 3),
           "identical_to_plain": same,
           "mpix_per_s": round(st["blends"] / t_batch / 1e6, 1),
           "fallbacks": st["fallbacks"], "underestimates": st["underestimates"],
           "edges": st["edges"], "tile_planes": st["tile_planes"],
           "autotuned": {"band_rows": tuned[0], "band_min_edges": tuned[1], "flags": tuned[2],
                         "on_canvases": tune_canvases}}
    batch.close()
    for su in surfs:
        su.close()
    for pr in preps:
        pr.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--paths", type=int, default=0, help="0 = the workload's 4096")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="budget of the CPU baseline sample (0 = skip)")
    ap.add_argument("--forms", default="exact,learned",
                    help="comma list of forms to time: exact, learned")
    ap.add_argument("--batch-steps", type=int, default=1,
                    help="timed frames of the batched uneven leg (each is ~650M segments)")
    ap.add_argument("--legs", default="batched,blur,tiger,text",
                    help="extra workloads reported beside the headline: batched, blur, tiger, "
                         "text")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run the short program the rocprofv3 counter summaries profile")
    ap.add_argument("--probe-workload", default=str(WORKLOAD),
                    choices=["2", "3", "tiger", "text"],
                    help="workload of --pmc-probe: 2 curves + strokes, 3 gradients + blur, "
                         "tiger (tiger.skp at 4096^2), text (the 2048^2 text page)")
    args = ap.parse_args()
    if args.pmc_probe:
        pmc_probe(args.probe_workload)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: every box of this build has a single MI355X "
              "and the pipeline has no multi-GPU path; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from skity_amd import abi, device, presizer, scenes

    sc = scenes.generate(WORKLOAD, n_paths=args.paths)
    ctx = device.Context(0)
    prep = device.Prepared(ctx, sc.scene)
    surf = device.Surface(ctx, sc.width, sc.height)
    forms = [f for f in args.forms.split(",") if f]
    if "learned" in forms:
        if os.path.exists(presizer.WEIGHTS):
            prm, margin = presizer.load()
            presizer.install(ctx, prm, margin)
        else:
            forms.remove("learned")
    results = {}
    for form in forms:
        mode = abi.MODE_EXACT if form == "exact" else abi.MODE_LEARNED
        results[form] = run_form(surf, prep, mode, args.steps, args.warmup, torch)
        log(f"{form} form: {results[form][0] * 1e3 / args.steps:.1f} ms per frame")
    geom = geometry_bytes(sc)
    out_forms = {}
    for form, (wall, stage_tot, st) in results.items():
        stages = {k: {"ms_per_step": round(v[0] / args.steps, 4),
                      "launches_per_step": round(v[1] / args.steps, 2)}
                  for k, v in stage_tot.items() if v[1]}
        out_forms[form] = {
            "paths_per_s": round(sc.n_paths / (wall / args.steps), 1),
            "ms_per_step": round(wall * 1e3 / args.steps, 3),
            "mpix_per_s": round(st["blends"] / (wall / args.steps) / 1e6, 1),
            "fallbacks_per_frame": st["fallbacks_per_frame"],
            "underestimates_per_frame": st["underestimates"],
            "stages": stages}
    best = max(out_forms, key=lambda f: out_forms[f]["paths_per_s"])
    wall, stage_tot, st = results[best]
    # The timed frames are checked: the learned frame must equal the exact
    # frame byte for byte.
    same = None
    if "exact" in results and "learned" in results:
        frames = {}
        for form, mode in (("exact", abi.MODE_EXACT), ("learned", abi.MODE_LEARNED)):
            surf.clear(0)
            surf.draw(prep, mode=mode)
            frames[form] = surf.read()
        same = bool(np.array_equal(frames["exact"], frames["learned"]))
        if not same:
            nd = int((frames["exact"] != frames["learned"]).any(-1).sum())
            print(f"bench: the learned frame differs from the exact frame at {nd} pixels",
                  file=sys.stderr, flush=True)
            sys.exit(3)
    # Roofline of the dominant stage (the walk pass): the frame's algorithmic
    # bytes G + 8 U + B over that stage's average launch time, against HBM.
    dom = max((k for k in stage_tot if stage_tot[k][1]), key=lambda k: stage_tot[k][0])
    avg_launch_ms = stage_tot[dom][0] / stage_tot[dom][1]
    alg = frame_alg_bytes(sc, st)
    model = algorithmic_bytes(dom, st, geom, sc.width * sc.height)
    achieved = alg / (avg_launch_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 7),
            "traffic": measured_traffic(dom, best),
            "traffic_source": "committed profile " + os.path.relpath(PMC_SUMMARY, HERE),
            "alg_bytes": int(alg), "alg_bytes_def": "G + 8 U + B per frame",
            "avg_launch_ms": round(avg_launch_ms, 4),
            "traffic_model_bytes": int(model) if model is not None else None,
            "traffic_model_def": "32 B per edge + 256 B per tile plane (walk)"}
    ref = out_forms.get("exact", out_forms[best])
    out = {
        "metric": METRIC, "value": out_forms[best]["paths_per_s"], "unit": "paths/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": out_forms[best]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "curves+strokes: 4096 paths, 8-200 quad/cubic/conic segments, "
                               "20% stroked, 2048x2048 RGBA8 src-over",
                   "paths": sc.n_paths, "canvas": [sc.width, sc.height],
                   "seed": scenes.SEED_BASE + WORKLOAD, "form": best},
        "mpix_per_s": out_forms[best]["mpix_per_s"],
        "speedup_vs_exact": round(out_forms[best]["paths_per_s"] / ref["paths_per_s"], 3),
        "learned_equals_exact": same,
        "frame": {"edges": st["edges"], "tile_planes": st["tile_planes"],
                  "blends": st["blends"], "pixels": st["pixels"],
                  "device_ms": round(st["ms_total"], 3)},
        "forms": out_forms,
        "roofline": roof,
        "device": device_info(torch),
    }
    legs = {}
    leg_names = [x for x in args.legs.split(",") if x]
    if "batched" in leg_names and "learned" in forms:
        legs["batched_uneven"] = batched_leg(ctx, torch, args.batch_steps)
    if "blur" in leg_names:
        legs["gradients_blur"] = blur_leg(ctx, torch, max(2, args.steps // 4))
        log("blur leg done")
    if "tiger" in leg_names:
        legs["tiger_skp"] = tiger_leg(ctx, torch, max(3, args.steps // 2))
        log("tiger leg done")
    if "text" in leg_names:
        legs["text_page"] = text_leg(ctx, torch, max(3, args.steps // 2))
        log("text leg done")
    out["legs"] = legs
    if args.cpu_seconds > 0:
        log("cpu baseline")
        cb = cpu_baseline(sc, args.cpu_seconds)
        cb["value"] = round(cb["value"], 2)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    surf.close()
    prep.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
