"""Learned voice-tail predictor, CPU side: the FP64 MLP oracle is pinned by
central finite differences; the library's host-side initialisation
(standardisation statistics, He-normal weights) is checked against numpy; the
plan features are checked for consistency on configuration-2 jobs."""

import numpy as np
import pytest

import claysampler_amd as ca
from claysampler_amd import abi, synth, tails
from oracle import mlp

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(ca.HIP_LIB), reason="libclayhip.so not built")


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, abi.MLP_PARAMS)
    x = rng.normal(0, 1, (37, abi.MLP_INPUTS))
    y = rng.normal(0, 1, 37)
    _, g = mlp.loss_and_gradient(p, x, y)
    for k in rng.choice(abi.MLP_PARAMS, 60, replace=False):
        e = np.zeros_like(p)
        e[k] = 1e-6
        num = (mlp.loss_and_gradient(p + e, x, y)[0] - mlp.loss_and_gradient(p - e, x, y)[0]) / 2e-6
        assert abs(num - g[k]) <= 1e-6 + 1e-5 * abs(g[k]), k


def test_oracle_forward_closed_form():
    """A network with one active path computes relu(relu(a x0 + c) b) w + d."""
    p = np.zeros(abi.MLP_PARAMS)
    o = 0
    p[o + 0 * 64 + 5] = 2.0            # W1[0][5]
    o += 12 * 64
    p[o + 5] = 0.5                     # b1[5]
    o += 64
    p[o + 5 * 64 + 9] = 3.0            # W2[5][9]
    o += 64 * 64 + 64
    p[o + 9] = -1.5                    # w3[9]
    p[-1] = 0.25                       # b3
    x = np.zeros((3, 12))
    x[:, 0] = [-1.0, 0.0, 2.0]
    y, _, _ = mlp.forward(p, x)
    want = np.maximum(np.maximum(2 * x[:, 0] + 0.5, 0) * 3, 0) * -1.5 + 0.25
    assert np.allclose(y, want)


def test_init_standardises_and_is_deterministic():
    rng = np.random.default_rng(5)
    x = (rng.normal(3, 2, (5000, 12)) * np.arange(1, 13)).astype(np.float32)
    x[:, 4] = 1.0  # a constant feature keeps scale 1
    y = rng.normal(7, 0.5, 5000).astype(np.float32)
    m = tails.init(x, y, seed=11)
    assert np.allclose(np.ctypeslib.as_array(m.input_mean), x.mean(0), rtol=1e-5, atol=1e-4)
    std = x.std(0)
    want = np.where(std > 0, 1 / np.where(std > 0, std, 1), 1.0)
    assert np.allclose(np.ctypeslib.as_array(m.input_scale), want, rtol=1e-4)
    assert abs(m.output_mean - y.mean()) < 1e-4 and abs(m.output_scale - 1 / y.std()) < 1e-3
    W1, b1, W2, b2, w3, b3 = mlp.unpack(tails.params(m))
    assert abs(W1.std() - np.sqrt(2 / 12)) < 0.05 and abs(W2.std() - np.sqrt(2 / 64)) < 0.01
    assert not b1.any() and not b2.any() and b3 == 0
    assert np.array_equal(tails.params(tails.init(x, y, seed=11)), tails.params(m))
    assert not np.array_equal(tails.params(tails.init(x, y, seed=12)), tails.params(m))


def test_chain_features_of_configuration_2(synth_root):
    sfz = synth.c2_instrument(f"{synth_root}/c2")
    s = ca.Session()
    for j in range(2):
        s.add_job(sfz, synth.c2_midi(f"{synth_root}/c2/short_{j}.mid", 500 + j, seconds=8.0))
    plans = s.plans()
    x, release, end, ok = tails.chain_records(plans, s.bank_array())
    assert x.shape == (sum(p.filter_tracks for p in plans), 12)
    assert ok.sum() > 0.9 * len(ok) and np.isfinite(x).all()
    assert (release[ok] < end[ok]).all()
    # lpf_2p cutoff=200 + CC1 up to 6386 cents: 200 Hz .. 8 kHz at release.
    cutoff = 1000 * 2 ** x[ok, 5]
    assert cutoff.min() >= 199 and cutoff.max() <= 8010
    assert np.allclose(x[ok, 4], 1.0)          # loop_continuous
    assert np.allclose(x[ok, 6], 6 / 12)       # resonance 6 dB
    # The analytic baseline never exceeds the plan and starts at the release.
    base = tails.envelope_baseline(x, release, end, -90.0)
    assert (base[ok] <= end[ok]).all() and (base[ok] >= release[ok]).all()
