"""FP64 restatement of the voice-tail MLP (TEST INFRASTRUCTURE ONLY: the
checker of claysampler_amd/csrc/hip/mlp.hip, never used by the product).

12 -> 64 -> 64 -> 1 with ReLU; parameters packed as W1[12][64], b1, W2[64][64],
b2, w3[64], b3 (include/clay_hip.h); loss = mean of (yhat - y)^2 over the
samples, inputs and target already standardised. No reference counterpart
(the reference has no predictor); pinned by central finite differences."""

import numpy as np

IN, HID = 12, 64


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + IN * HID].reshape(IN, HID); o += IN * HID
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    h1 = np.maximum(x @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    return h2 @ w3 + b3, h1, h2


def loss_and_gradient(p, x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    yhat, h1, h2 = forward(p, x)
    n = len(y)
    e = yhat - y
    dy = 2 * e / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, w3) * (h2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    gW1 = x.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return float((e * e).mean()), g


def standardise(model, x, y=None):
    """Inputs (and target) in the model's standardised units, FP64."""
    xs = (np.asarray(x, np.float64) - np.asarray(model.input_mean, np.float64)) * np.asarray(model.input_scale,
                                                                                              np.float64)
    if y is None:
        return xs
    return xs, (np.asarray(y, np.float64) - model.output_mean) * model.output_scale
