"""Float64 restatement of the pre-sizer MLP -- TEST INFRASTRUCTURE ONLY.

The device network (skity_amd/csrc/hipr_mlp.hpp) is 16-64-64-2 with ReLU,
FP16 weights and activations as MFMA operands and FP32 accumulation; its loss
is the mean squared error in normalised target space.  This module computes
the same forward pass and the exact gradient in float64 (optionally rounding
weights and activations to float16 the way the device does), so the device
gradients can be checked against it.  There is no reference counterpart: the
reference sizes its buffers by growing them.
"""
import numpy as np

IN, HID, OUT = 16, 64, 2
W1, B1 = 0, IN * HID
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID * OUT
TRAINED = B3 + OUT
FMU, FSIG = TRAINED, TRAINED + IN
YMU, YSIG = FSIG + IN, FSIG + IN + OUT
TOTAL = YSIG + OUT


def unpack(prm):
    prm = np.asarray(prm, np.float64)
    return dict(
        w1=prm[W1:B1].reshape(IN, HID), b1=prm[B1:W2], w2=prm[W2:B2].reshape(HID, HID),
        b2=prm[B2:W3], w3=prm[W3:B3].reshape(HID, OUT), b3=prm[B3:TRAINED],
        fmu=prm[FMU:FSIG], fsig=prm[FSIG:YMU], ymu=prm[YMU:YSIG], ysig=prm[YSIG:TOTAL])


def _h(a, half):
    return a.astype(np.float16).astype(np.float64) if half else a


def forward(prm, x, half=True):
    """Returns (normalised outputs z3, cache)."""
    p = unpack(prm)
    xn = _h((np.asarray(x, np.float64) - p["fmu"]) * p["fsig"], half)
    z1 = xn @ _h(p["w1"], half) + p["b1"]
    h1 = _h(np.maximum(z1, 0), half)
    z2 = h1 @ _h(p["w2"], half) + p["b2"]
    h2 = _h(np.maximum(z2, 0), half)
    z3 = h2 @ _h(p["w3"], half) + p["b3"]
    return z3, (p, xn, z1, h1, z2, h2)


def predict(prm, x, half=True):
    z3, (p, *_) = forward(prm, x, half)
    return z3 * p["ysig"] + p["ymu"]


def loss_and_grad(prm, x, y, half=True):
    """Mean squared error over samples (summed over the two outputs) in
    normalised target space, and its gradient w.r.t. the trained parameters."""
    z3, (p, xn, z1, h1, z2, h2) = forward(prm, x, half)
    n = len(x)
    yn = (np.asarray(y, np.float64) - p["ymu"]) / p["ysig"]
    e = z3 - yn
    loss = float((e ** 2).sum() / n)
    d3 = 2 * e / n
    g = np.zeros(TRAINED)
    g[W3:B3] = (h2.T @ d3).ravel()
    g[B3:TRAINED] = d3.sum(0)
    d2 = (d3 @ _h(p["w3"], half).T) * (z2 > 0)
    g[W2:B2] = (h1.T @ d2).ravel()
    g[B2:W3] = d2.sum(0)
    d1 = (d2 @ _h(p["w2"], half).T) * (z1 > 0)
    g[W1:B1] = (xn.T @ d1).ravel()
    g[B1:W2] = d1.sum(0)
    return loss, g
