"""Pins of the oracle's quality measures and verdicts (CPU only).

The oracle's PCM for each scenario (itself pinned by the render and tail
tests) is measured again here with numpy from the definitions of
src/cuda/quality.cu:6-170: body power over 250 ms windows inside 2 s chunks,
the per-onset maximum of five 50 ms windows, levels in dB of full-scale
mono power, linearly interpolated percentiles, silent fraction and longest
silent run, clipped fraction; then the verdict order of quality.cu:152-164
(duration, quiet, silence, peak, clipping) and the group rules of
quality.cu:172-209. Every verdict is reached by one scenario."""

import math

import numpy as np
import pytest

import oracle
from claysampler_amd import abi

from helpers import quality_cases, quality_measures, quality_session, quality_verdict

RATE = 44100


def render(case):
    policy, jobs, audio, group_minimum, _ = case
    s = quality_session(case)
    cfg = oracle.config(policy=policy, taps=32, audio=audio, group_minimum=group_minimum)
    bp, bc = s.bank()
    plans = s.plans()
    res, pcm = oracle.render(cfg, bp, bc, plans)
    return s, plans, res, pcm


@pytest.mark.parametrize("name", ["ok", "quietPerformance", "excessiveSilence", "silentAudio", "clipping",
                                  "audioDuration", "orig_metrics"])
def test_quality_measures_and_verdicts(tmp_path, name):
    case = quality_cases(tmp_path)[name]
    policy, _, a, _, want = case
    channels = 1 if policy == abi.POLICY_NATURAL else 2
    s, plans, res, pcm = render(case)
    assert [abi.TAIL_STATUS[r.status] for r in res] == want
    for i, r in enumerate(res):
        if r.status != 0:
            # Rejected jobs keep their measures; re-render to read the PCM they had.
            relaxed = abi.AudioConfig.from_buffer_copy(a)
            relaxed.max_frames, relaxed.min_median, relaxed.min_quartile = 1 << 30, -1e9, -1e9
            relaxed.max_silent_fraction, relaxed.max_silent_frames, relaxed.min_peak = 2.0, 1e18, 0
            relaxed.max_clipped_fraction = 2.0
            bp, bc = s.bank()
            rres, rpcm = oracle.render(oracle.config(policy=policy, taps=32, audio=relaxed), bp, bc, plans)
            data, frames, peak = rpcm[rres[i].offset * channels:(rres[i].offset + rres[i].data_frames) * channels], \
                rres[i].frames, rres[i].peak
        else:
            data = pcm[r.offset * channels:(r.offset + r.data_frames) * channels]
            frames, peak = r.frames, r.peak
        p = plans[i]
        m = quality_measures(data, channels, frames, p.body_frames, [p.onsets[k] for k in range(p.onset_count)], a)
        assert r.audio.measured == 1
        for key in ("median", "quartile", "silent_fraction", "clipped_fraction", "body_rms"):
            assert getattr(r.audio, key) == pytest.approx(m[key], rel=1e-12, abs=1e-12), key
        assert (r.audio.longest_silence, r.audio.clipped) == (m["longest_silence"], m["clipped"])
        assert quality_verdict(m, frames, peak, a) == abi.TAIL_STATUS[r.status]
        assert r.audio.verdict == r.status


def test_group_verdicts_and_gains(tmp_path):
    """quality.cu:172-209: a member with no body is silentBody; a group left
    with fewer than the minimum accepted members is incompleteGroup; a full
    group shares target = min(mean body RMS, the safe level of its loudest
    peak) and each member's gain is target / its own body RMS."""
    case = quality_cases(tmp_path)["silentBody"]
    _, _, a, _, want = case
    s, plans, res, pcm = render(case)
    assert [abi.TAIL_STATUS[r.status] for r in res] == want
    # Group 2 measured without the group: body RMS and peak of each member.
    policy, jobs, _, _, _ = case
    bp, bc = s.bank()
    alone = [oracle.render(oracle.config(policy=policy, taps=32, audio=a), bp, bc, (abi.PlanView * 1)(plans[i])) for i in (2, 3)]
    rms, peaks = [], []
    for i, (r, p) in zip((2, 3), alone):
        data = p[:r[0].data_frames]
        m = quality_measures(data, 1, r[0].frames, plans[i].body_frames, [], a)
        rms.append(m["body_rms"])
        peaks.append(int(np.abs(data.astype(np.int32)).max()))
    target = min(sum(rms) / 2, min(32767 * 10 ** (-1 / 20) * q / pk for q, pk in zip(rms, peaks)))
    for k, i in enumerate((2, 3)):
        assert res[i].audio.target_rms == pytest.approx(target, rel=1e-12)
        assert res[i].audio.gain == pytest.approx(target / rms[k], rel=1e-12)
        data = alone[k][1][:alone[k][0][0].data_frames].astype(np.float64)
        got = pcm[res[i].offset:res[i].offset + res[i].data_frames]
        assert np.array_equal(got.astype(np.int64), np.rint(data * (target / rms[k])).astype(np.int64))
