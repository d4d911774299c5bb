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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--scenes", type=int, default=24)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--quantile", type=float, default=0.995)
    args = ap.parse_args()
    ctx = device.Context(0)
    t0 = time.time()
    sc = training_scenes(args.scenes)
    feats, targets = collect(ctx, sc)
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
