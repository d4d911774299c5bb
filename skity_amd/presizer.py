"""The learned pre-sizer: data collection, training on the device, margins.

Labels come from the exact two-pass draw (per raster job: edges built and
coverage tile planes written); features from the library's features kernel.
Training runs on the GPU through hipr_presizer_train (the library's own
MFMA forward/backward kernels and Adam).  The installed network lets
HIPR_MODE_LEARNED skip the counting passes; every under-prediction is
detected on the device and redone by the exact method.

    python -m skity_amd.presizer --out skity_amd/presizer_weights.npz
"""
import argparse
import ctypes
import os
import sys
import time

import numpy as np

from . import abi, device, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "presizer_weights.npz")
N_IN, N_TRAINED, N_TOTAL = abi.PRESIZER_FEATURES, abi.PRESIZER_TRAINED, abi.PRESIZER_PARAMS


def _lib():
    L = device.lib()
    if not getattr(L, "_presizer_declared", False):
        fp = ctypes.POINTER(ctypes.c_float)
        L.hipr_prepared_job_count.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
        L.hipr_prepared_features.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_void_p, ctypes.c_int32]
        L.hipr_surface_get_job_counts.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p, ctypes.c_int32]
        L.hipr_presizer_predict.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int64, ctypes.c_void_p]
        L.hipr_presizer_gradients.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                              fp]
        L.hipr_presizer_train.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_float, ctypes.c_uint64,
                                          ctypes.c_void_p]
        L.hipr_context_set_presizer.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                                ctypes.c_float]
        for f in ("hipr_prepared_job_count", "hipr_prepared_features",
                  "hipr_surface_get_job_counts", "hipr_presizer_predict",
                  "hipr_presizer_gradients", "hipr_presizer_train",
                  "hipr_context_set_presizer"):
            getattr(L, f).restype = ctypes.c_int
        L._presizer_declared = True
    return L


def job_features(prepared, width, height):
    L = _lib()
    n = ctypes.c_int32()
    device.check(L.hipr_prepared_job_count(prepared.handle, ctypes.byref(n)),
                 "hipr_prepared_job_count")
    f = np.zeros((n.value, N_IN), np.float32)
    device.check(L.hipr_prepared_features(prepared.handle, width, height, f.ctypes.data, n.value),
                 "hipr_prepared_features")
    return f


def job_counts(surface, n_jobs):
    e = np.zeros(n_jobs, np.int32)
    p = np.zeros(n_jobs, np.int32)
    device.check(_lib().hipr_surface_get_job_counts(surface._h, e.ctypes.data, p.ctypes.data,
                                                    n_jobs), "hipr_surface_get_job_counts")
    return e, p


def collect(ctx, scene_list):
    """Features and log1p(exact counts) of every raster job of the scenes."""
    feats, targets = [], []
    for sc in scene_list:
        prep = device.Prepared(ctx, sc.scene)
        f = job_features(prep, sc.width, sc.height)
        surf = device.Surface(ctx, sc.width, sc.height)
        surf.draw(prep, mode=abi.MODE_EXACT)
        e, p = job_counts(surf, len(f))
        keep = e >= 0
        feats.append(f[keep])
        targets.append(np.stack([np.log1p(e[keep]), np.log1p(p[keep])], 1).astype(np.float32))
        surf.close()
        prep.close()
    return np.concatenate(feats), np.concatenate(targets)


def init_params(feats, targets, seed=0):
    """He-initialised weights and the fixed feature/target normalisation."""
    rng = np.random.default_rng(seed)
    prm = np.zeros(N_TOTAL, np.float32)
    o = 0
    for fan_in, fan_out in ((16, 64), (64, 64), (64, 2)):
        w = rng.normal(0, np.sqrt(2.0 / fan_in), (fan_in, fan_out)).astype(np.float32)
        prm[o:o + fan_in * fan_out] = w.ravel()
        o += fan_in * fan_out + fan_out
    mu = feats.mean(0)
    sd = feats.std(0)
    prm[N_TRAINED:N_TRAINED + 16] = mu
    prm[N_TRAINED + 16:N_TRAINED + 32] = np.where(sd > 1e-6, 1.0 / np.maximum(sd, 1e-6), 0.0)
    prm[N_TRAINED + 32:N_TRAINED + 34] = targets.mean(0)
    prm[N_TRAINED + 34:N_TRAINED + 36] = np.maximum(targets.std(0), 1e-3)
    return prm


def predict(ctx, prm, feats):
    out = np.zeros((len(feats), 2), np.float32)
    f = np.ascontiguousarray(feats, np.float32)
    p = np.ascontiguousarray(prm, np.float32)
    device.check(_lib().hipr_presizer_predict(ctx.handle, p.ctypes.data, f.ctypes.data, len(f),
                                              out.ctypes.data), "hipr_presizer_predict")
    return out


def gradients(ctx, prm, feats, targets):
    g = np.zeros(N_TRAINED, np.float32)
    loss = ctypes.c_float()
    f = np.ascontiguousarray(feats, np.float32)
    t = np.ascontiguousarray(targets, np.float32)
    p = np.ascontiguousarray(prm, np.float32)
    device.check(_lib().hipr_presizer_gradients(ctx.handle, p.ctypes.data, f.ctypes.data,
                                                t.ctypes.data, len(f), g.ctypes.data,
                                                ctypes.byref(loss)), "hipr_presizer_gradients")
    return float(loss.value), g


def train(ctx, prm, feats, targets, steps=3000, batch=4096, lr=2e-3, seed=1):
    p = np.ascontiguousarray(prm, np.float32).copy()
    f = np.ascontiguousarray(feats, np.float32)
    t = np.ascontiguousarray(targets, np.float32)
    curve = np.zeros(steps, np.float32)
    device.check(_lib().hipr_presizer_train(ctx.handle, p.ctypes.data, f.ctypes.data,
                                            t.ctypes.data, len(f), steps, batch, lr, seed,
                                            curve.ctypes.data), "hipr_presizer_train")
    return p, curve


def margins(pred, targets, quantile=0.99):
    """Per-output additive margin (log1p units) so that `quantile` of the
    samples are not under-predicted."""
    return np.quantile(targets - pred, quantile, axis=0).astype(np.float32)


def install(ctx, prm, margin):
    p = np.ascontiguousarray(prm, np.float32)
    device.check(_lib().hipr_context_set_presizer(ctx.handle, p.ctypes.data, float(margin[0]),
                                                  float(margin[1])), "hipr_context_set_presizer")


def save(path, prm, margin, info):
    np.savez(path, params=np.asarray(prm, np.float32), margin=np.asarray(margin, np.float32),
             **{k: np.asarray(v) for k, v in info.items()})


def load(path=WEIGHTS):
    z = np.load(path, allow_pickle=False)
    return z["params"], z["margin"]


def training_scenes(n_scenes, seed0=0x7A11):
    """Curves + strokes scenes of 1024 paths at fresh seeds (never a bench
    seed), plain-fill scenes, and canvases of the batched uneven family
    (256^2-4096^2, lines / quads / cubics, gradients)."""
    out = []
    for i in range(n_scenes):
        out.append(scenes.generate(2, n_paths=1024, seed=seed0 + i))
        if i % 4 == 0:
            out.append(scenes.generate(1, seed=seed0 + 1000 + i))
        for k in range(4):
            out.append(scenes.generate(5, seed=seed0 + 2000 + 4 * i + k))
    return out


def uneven_training_canvases(min_paths, seed0=0xC4A10000):
    """Canvases of the batched uneven family at its stated scale (workload 4:
    100-50,000 outlines of 4-10,000 segments on 256^2-4096^2) at fresh
    seeds, yielded one at a time until min_paths paths were produced."""
    n, i = 0, 0
    while n < min_paths:
        sc = scenes.generate(4, seed=seed0 + i)
        n += sc.n_paths
        i += 1
        yield sc


def collect_stream(ctx, scene_iter, log_every=16):
    """collect() over an iterator, releasing each scene after use."""
    feats, targets = [], []
    t0 = time.time()
    n = 0
    for k, sc in enumerate(scene_iter):
        f, t = collect(ctx, [sc])
        feats.append(f)
        targets.append(t)
        n += len(f)
        if (k + 1) % log_every == 0:
            print(f"  {k + 1} canvases, {n} labelled jobs, {time.time() - t0:.0f} s", flush=True)
    return np.concatenate(feats), np.concatenate(targets)


def linear_estimates(sc, edges_per_segment=4, edges_const=8, planes_per_tile=2):
    """A pre-sizer without the network: edges bounded linearly in the segment
    count (each segment rarely yields more than a few edges after the quad
    conversion and the y-extremum chop) and planes in the bounding box's
    tiles.  Per path, as HIPR_MODE_ESTIMATES takes them."""
    s = sc.scene.contents
    n = s.n_paths
    vo = np.ctypeslib.as_array(s.verb_offsets, shape=(n + 1,))
    po = np.ctypeslib.as_array(s.point_offsets, shape=(n + 1,))
    pts = np.ctypeslib.as_array(s.points, shape=(2 * int(po[-1]),)).reshape(-1, 2)
    segs = np.diff(vo).astype(np.int64)
    # per-path bounding box of the points (paths own contiguous point ranges;
    # a path without points gets an empty box)
    has = np.diff(po) > 0
    starts = po[:-1][has].astype(np.int64)
    lx = np.zeros(n)
    hx = np.zeros(n)
    ly = np.zeros(n)
    hy = np.zeros(n)
    if starts.size:
        lx[has] = np.minimum.reduceat(pts[:, 0], starts)
        hx[has] = np.maximum.reduceat(pts[:, 0], starts)
        ly[has] = np.minimum.reduceat(pts[:, 1], starts)
        hy[has] = np.maximum.reduceat(pts[:, 1], starts)
    tiles = (np.floor((hx - lx) / 16) + 2) * (np.floor((hy - ly) / 16) + 2)
    e = np.minimum(edges_per_segment * segs + edges_const, 2 ** 30)
    t = np.minimum(planes_per_tile * tiles, 2 ** 30)
    return e.astype(np.int32), t.astype(np.int32)


def evaluate_large(ctx, n_paths=1 << 20, side=8192, slice_paths=1 << 17, seed=0xC5E7A10,
                   box=(-8.0, -3.0), log=print):
    """The pre-sizer's evaluation at its stated scale: n_paths outlines of the
    uneven family on a side x side canvas, drawn as consecutive scenes of
    slice_paths paths, by the exact two-pass method, by the learned method
    and by a linear estimate; the three frames must be identical.  Returns
    per-form device time, fallbacks, under-estimates and capacity totals."""
    forms = {"exact": abi.MODE_EXACT, "learned": abi.MODE_LEARNED,
             "linear_estimate": abi.MODE_ESTIMATES}
    surfs = {k: device.Surface(ctx, side, side) for k in forms}
    rep = {k: {"ms": 0.0, "fallbacks": 0, "underestimates": 0, "jobs": 0, "edges": 0,
               "tile_planes": 0, "blends": 0} for k in forms}
    t_gen = 0.0
    done = 0
    while done < n_paths:
        m = min(slice_paths, n_paths - done)
        t0 = time.time()
        sc = scenes.outlines(seed, side, done, m, box[0], box[1])
        t_gen += time.time() - t0
        prep = device.Prepared(ctx, sc.scene)
        ee, te = linear_estimates(sc)
        for k, mode in forms.items():
            kw = {"edge_estimates": ee, "tile_estimates": te} if k == "linear_estimate" else {}
            surfs[k].draw(prep, mode=mode, **kw)
            surfs[k].finish()
            st = surfs[k].stats()
            r = rep[k]
            r["ms"] += st["ms_total"]
            r["fallbacks"] += st["fallbacks"]
            r["underestimates"] += st["underestimates"]
            r["jobs"] += st["raster_paths"]
            r["edges"] += st["edges"]
            r["tile_planes"] += st["tile_planes"]
            r["blends"] += st["blends"]
            stages = r.setdefault("stage_ms", {})
            for name, (ms, _) in surfs[k].stage_times().items():
                if ms > 0:
                    stages[name] = stages.get(name, 0.0) + ms
        prep.close()
        done += m
        log(f"  {done} paths: " + ", ".join(f"{k} {rep[k]['ms']:.0f} ms" for k in forms))
        del sc
    frames = {k: surfs[k].read() for k in forms}
    for k in forms:
        surfs[k].close()
        r = rep[k]
        r["paths_per_s"] = n_paths / (r["ms"] * 1e-3)
        r["fallback_rate"] = r["fallbacks"] / max(r["jobs"], 1)
        r["underestimate_rate"] = r["underestimates"] / max(r["jobs"], 1)
        r["identical_to_exact"] = bool(np.array_equal(frames[k], frames["exact"]))
    rep["scene"] = {"paths": n_paths, "canvas": [side, side], "slices": -(-n_paths // slice_paths),
                    "box_log2": list(box), "seed": seed, "generation_s": round(t_gen, 1)}
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--scenes", type=int, default=24)
    ap.add_argument("--uneven-paths", type=int, default=1000000,
                    help="labelled paths of the batched uneven family (workload 4)")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--quantile", type=float, default=0.995)
    ap.add_argument("--evaluate", metavar="JSON",
                    help="instead of training: evaluate the installed weights on the 1M-path "
                         "8192^2 scene and write the report here")
    ap.add_argument("--eval-paths", type=int, default=1 << 20)
    args = ap.parse_args()
    ctx = device.Context(0)
    if args.evaluate:
        import json
        install(ctx, *load())
        rep = evaluate_large(ctx, n_paths=args.eval_paths,
                             log=lambda m: print(m, flush=True))
        with open(args.evaluate, "w") as fh:
            json.dump(rep, fh, indent=1)
        print(json.dumps(rep), flush=True)
        ctx.close()
        return 0
    t0 = time.time()
    sc = training_scenes(args.scenes)
    feats, targets = collect(ctx, sc)
    del sc
    if args.uneven_paths > 0:
        f4, t4 = collect_stream(ctx, uneven_training_canvases(args.uneven_paths))
        print(f"uneven family: {len(f4)} labelled jobs", flush=True)
        feats = np.concatenate([feats, f4])
        targets = np.concatenate([targets, t4])
    n = len(feats)
    rng = np.random.default_rng(5)
    perm = rng.permutation(n)
    n_val = max(n // 10, 1)
    val, tr = perm[:n_val], perm[n_val:]
    print(f"collected {n} jobs in {time.time() - t0:.1f} s", flush=True)
    prm = init_params(feats[tr], targets[tr])
    t1 = time.time()
    prm, curve = train(ctx, prm, feats[tr], targets[tr], steps=args.steps)
    print(f"trained {args.steps} steps in {time.time() - t1:.1f} s; loss {curve[0]:.4f} -> "
          f"{curve[-50:].mean():.4f}", flush=True)
    pv = predict(ctx, prm, feats[val])
    m = margins(predict(ctx, prm, feats[tr]), targets[tr], args.quantile)
    under = ((pv + m) < targets[val]).any(1).mean()
    rmse = np.sqrt(((pv - targets[val]) ** 2).mean(0))
    print(f"validation rmse (log1p) {rmse}, margin {m}, under-prediction rate {under:.4f}",
          flush=True)
    save(args.out, prm, m, {"loss_curve": curve, "val_rmse": rmse, "val_under": under,
                            "n_train": len(tr), "n_val": n_val})
    print(f"saved {args.out}", flush=True)
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
