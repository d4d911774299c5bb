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


def geometry_bytes(sc):
    """G of the scene: 8 B per point, 1 B per verb, 4 B per conic weight and
    a 32-byte draw record per path."""
    s = sc.scene.contents
    n = s.n_paths
    return 8 * s.point_offsets[n] + s.verb_offsets[n] + 4 * s.weight_offsets[n] + 32 * n


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--paths", type=int, default=0, help="0 = the workload's 4096")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="budget of the CPU baseline sample (0 = skip)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: every box of this build has a single MI355X "
              "and the pipeline has no multi-GPU path; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from skity_amd import device, scenes

    sc = scenes.generate(WORKLOAD, n_paths=args.paths)
    ctx = device.Context(0)
    prep = device.Prepared(ctx, sc.scene)
    surf = device.Surface(ctx, sc.width, sc.height)
    for _ in range(args.warmup):
        surf.clear(0)
        surf.draw(prep)
    torch.cuda.synchronize()
    stage_tot = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        surf.clear(0)
        surf.draw(prep)
        for k, (ms, n) in surf.stage_times().items():
            a = stage_tot.setdefault(k, [0.0, 0])
            a[0] += ms
            a[1] += n
    surf.finish()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = surf.stats()
    ms_step = wall * 1e3 / args.steps
    paths_per_s = sc.n_paths / (wall / args.steps)
    mpix = st["blends"] / (wall / args.steps) / 1e6

    # dominant kernel and its roofline
    geom = geometry_bytes(sc)
    stages = {k: {"ms_per_step": v[0] / args.steps, "launches_per_step": v[1] / args.steps}
              for k, v in stage_tot.items() if v[1]}
    dom = max(stages, key=lambda k: stages[k]["ms_per_step"])
    avg_launch_ms = stage_tot[dom][0] / stage_tot[dom][1]
    alg = algorithmic_bytes(dom, st, geom, sc.width * sc.height)
    roof = None
    if alg is not None:
        achieved = alg / (avg_launch_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None, "alg_bytes_per_launch": int(alg),
                "avg_launch_ms": round(avg_launch_ms, 4)}
    out = {
        "metric": METRIC, "value": round(paths_per_s, 1), "unit": "paths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "curves+strokes: 4096 paths, 8-200 quad/cubic/conic segments, "
                               "20% stroked, 2048x2048 RGBA8 src-over",
                   "paths": sc.n_paths, "canvas": [sc.width, sc.height],
                   "seed": scenes.SEED_BASE + WORKLOAD, "form": "plain (exact two-pass)"},
        "mpix_per_s": round(mpix, 1),
        "frame": {"edges": st["edges"], "tile_planes": st["tile_planes"],
                  "blends": st["blends"], "pixels": st["pixels"],
                  "device_ms": round(st["ms_total"], 3)},
        "stages": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in stages.items()},
        "roofline": roof,
    }
    if args.cpu_seconds > 0:
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
