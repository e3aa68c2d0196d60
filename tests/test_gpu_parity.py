"""HIP render core vs the FP64 oracle on the same plans (configurations 1-3,
the reference's tail/pedal/segment/group scenarios, and a feature mix).

Bar: statuses, frame counts, data frames, fades and offsets identical; PCM16
within +-1 LSB (rounding of the float mix at the 0.5 boundary)."""

import numpy as np
import pytest

import claysampler_amd as ca
from claysampler_amd import abi, synth

from helpers import (assert_parity, chain_shapes_case, gain_case, gpu_render, group_case, oracle_render, pedal_case,
                     rich_case, segment_case, tail_case)

pytestmark = pytest.mark.gpu


def both(session, policy, taps, **kw):
    return gpu_render(session, policy, taps, **kw), oracle_render(session, policy, taps, **kw)


@pytest.mark.parametrize("policy", [abi.POLICY_NATURAL, abi.POLICY_ORIG])
def test_config1(synth_root, policy):
    sfz, mids = synth.config1(f"{synth_root}/c1")
    s = ca.Session(policy=policy)
    s.add_job(sfz, mids[0])
    assert_parity(*both(s, policy, 16))


@pytest.mark.parametrize("taps", [16, 32, 48, 64])
def test_config1_taps(synth_root, taps):
    sfz, mids = synth.config1(f"{synth_root}/c1")
    s = ca.Session()
    s.add_job(sfz, mids[0])
    assert_parity(*both(s, abi.POLICY_NATURAL, taps))


def test_config2_short(synth_root):
    """Configuration 2's instrument and MIDI generator, 4 jobs x 8 s: bends,
    CC1-driven lpf_2p, loops with crossfade, 32 taps."""
    sfz = synth.c2_instrument(f"{synth_root}/c2")
    s = ca.Session()
    for j in range(4):
        s.add_job(sfz, synth.c2_midi(f"{synth_root}/c2/short_{j}.mid", 500 + j, seconds=8.0))
    assert_parity(*both(s, abi.POLICY_NATURAL, 32))


def test_config3_sample(synth_root):
    sfz = synth.c1_instrument(f"{synth_root}/c1")
    s = ca.Session()
    for j, poly in enumerate((4, 24, 64)):
        s.add_job(sfz, synth.c3_midi(f"{synth_root}/c1/c3s_{j}.mid", 900 + j, 6.0, poly))
    assert_parity(*both(s, abi.POLICY_NATURAL, 16))


@pytest.mark.parametrize("policy,extra", [
    (abi.POLICY_NATURAL, {}), (abi.POLICY_ORIG, {}),
    (abi.POLICY_CUSTOM, {"tail_max_seconds": 1.0, "tail_fade_ms": 100}),
    (abi.POLICY_CUSTOM, {"tail_max_seconds": 4.0})])
def test_tail_case(tmp_path, policy, extra):
    sfz, mid = tail_case(tmp_path)
    s = ca.Session(policy=policy, **extra)
    s.add_job(sfz, mid)
    got = assert_parity(*both(s, policy, 48))
    if policy == abi.POLICY_NATURAL:
        assert got[0][0].frames == round(.65 * 44100)  # tests/check_tail.py:55


def test_pedals(tmp_path):
    sfz, mids = pedal_case(tmp_path)
    s = ca.Session()
    for name in ("regular", "held", "later", "sostenuto"):
        s.add_job(sfz, mids[name])
    res, _ = assert_parity(*both(s, abi.POLICY_NATURAL, 48))
    frames = [r.frames for r in res]
    assert frames[1] % 1024 == 0  # tests/check_tail.py:95


@pytest.mark.parametrize("policy", [abi.POLICY_NATURAL, abi.POLICY_ORIG])
def test_segment_case(tmp_path, policy):
    sfz, long_mid, short_mid = segment_case(tmp_path)
    s = ca.Session(policy=policy)
    for m in (short_mid, long_mid, short_mid, long_mid):
        s.add_job(sfz, m)
    assert_parity(*both(s, policy, 48))


def test_rich_features(tmp_path):
    sfz, mid = rich_case(tmp_path)
    for policy in (abi.POLICY_NATURAL, abi.POLICY_ORIG):
        s = ca.Session(policy=policy)
        s.add_job(sfz, mid)
        assert_parity(*both(s, policy, 32))


def test_filter_chain_shapes(tmp_path):
    """Every compiled filter-chain shape (stereo, multi-section, one or more
    bindings, high-pass) against the oracle, in both output layouts."""
    sfz, mid = chain_shapes_case(tmp_path)
    for policy in (abi.POLICY_NATURAL, abi.POLICY_ORIG):
        s = ca.Session(policy=policy)
        s.add_job(sfz, mid)
        assert_parity(*both(s, policy, 32))


def test_job_gain_automation(tmp_path):
    """Volume and expression changing inside voice spans (the job gain read
    per frame) and between them (one point per span), wet and dry voices."""
    sfz, mid = gain_case(tmp_path)
    for policy in (abi.POLICY_NATURAL, abi.POLICY_ORIG):
        s = ca.Session(policy=policy)
        s.add_job(sfz, mid)
        assert_parity(*both(s, policy, 32))


def test_quality_and_groups(tmp_path):
    presets, melody = group_case(tmp_path)
    s = ca.Session(measure=True, keep_onsets=True)
    # Group members are consecutive, as the reference manifest orders them (src/main.cpp:230-245).
    rows = [("soft", 1), ("quiet", 1), ("loud", 1), ("loud", 0)]
    for preset, group in rows:
        s.add_job(presets[preset], melody, group)
    got = assert_parity(*both(s, abi.POLICY_NATURAL, 48, quality=True, group_minimum=2))
    res = got[0]
    assert abi.TAIL_STATUS[res[1].status] == "quietPerformance"
    assert res[0].audio.gain > 0 and res[2].audio.gain > 0 and res[3].audio.gain == 0


def test_batching_is_bit_exact(synth_root):
    """One batch of all jobs vs one job per batch: identical PCM16."""
    sfz = synth.c2_instrument(f"{synth_root}/c2")
    s = ca.Session()
    for j in range(3):
        s.add_job(sfz, synth.c2_midi(f"{synth_root}/c2/short_{j}.mid", 500 + j, seconds=8.0))
    whole = gpu_render(s, abi.POLICY_NATURAL, 32)
    bank_ptr, bank_count = s.bank()
    r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=32, batch_jobs=1)
    r.upload_bank(bank_ptr, bank_count)
    pieces = []
    r.render_window(s.plans(), lambda res, pcm, first, failed: pieces.append((res, pcm)))
    assert len(pieces) == 3
    joined = np.concatenate([p for _, p in pieces])
    assert np.array_equal(joined, whole[1])
    # Every batch prepared before any runs: each upload after the first
    # rewrites the renderer's pinned record image while the copy before it
    # may still be in flight.
    plans = s.plans()
    batches = [r.prepare((abi.PlanView * 1)(plans[j])) for j in range(3)]
    ahead = []
    for b in batches:
        b.run()
        ahead.append(b.results()[1])
        b.close()
    r.close()
    assert np.array_equal(np.concatenate(ahead), whole[1])
