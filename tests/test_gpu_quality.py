"""Quality and group verdicts on the GPU (src/cuda/quality.cu:6-209).

One scenario per verdict (ok, quietPerformance, excessiveSilence, silentAudio,
clipping, audioDuration, silentBody, incompleteGroup). For each job:
- the status equals the oracle's, and PCM16 is within the parity bar;
- the GPU's AudioMetrics equal the numpy restatement of the quality measures
  applied to the GPU's own PCM (exact integer sums, rel 1e-12 for doubles);
- the metrics are close to the oracle's (the PCM they measure may differ by
  1 LSB), and the verdict follows the reference's precedence."""

import numpy as np
import pytest

import claysampler_amd as ca
import oracle
from claysampler_amd import abi

from helpers import assert_parity, quality_cases, quality_measures, quality_session, quality_verdict

pytestmark = pytest.mark.gpu

METRICS = ("median", "quartile", "silent_fraction", "clipped_fraction", "body_rms", "target_rms", "gain")


def render_both(case, taps=32):
    policy, jobs, audio, group_minimum, _ = case
    s = quality_session(case)
    bp, bc = s.bank()
    plans = s.plans()
    r = ca.Renderer(policy=policy, taps=taps, audio=audio, group_minimum=group_minimum, batch_jobs=16)
    r.upload_bank(bp, bc)
    b = r.prepare(plans)
    b.run()
    got = b.results()
    b.close()
    r.close()
    want = oracle.render(oracle.config(policy=policy, taps=taps, audio=audio, group_minimum=group_minimum), bp, bc,
                         plans)
    return s, plans, got, want  # the session owns the memory the plans point into


@pytest.mark.parametrize("name", ["ok", "quietPerformance", "excessiveSilence", "silentAudio", "clipping",
                                  "audioDuration", "silentBody", "orig_metrics"])
def test_quality_verdicts_match_the_oracle(tmp_path, name):
    case = quality_cases(tmp_path)[name]
    policy, _, a, group_minimum, want_status = case
    channels = 1 if policy == abi.POLICY_NATURAL else 2
    session, plans, (gres, gpcm), (wres, wpcm) = render_both(case)
    print(f"case {name}: {[abi.TAIL_STATUS[r.status] for r in gres]}", flush=True)
    assert [abi.TAIL_STATUS[r.status] for r in gres] == want_status
    lsb = 1 + max([int(np.ceil(r.audio.gain)) for r in wres] + [0])
    assert_parity((gres, gpcm), (wres, wpcm), lsb=lsb)
    for i, (g, w) in enumerate(zip(gres, wres)):
        assert (g.audio.measured, g.audio.verdict) == (w.audio.measured, w.audio.verdict)
        for key in METRICS:
            assert getattr(g.audio, key) == pytest.approx(getattr(w.audio, key), rel=2e-3, abs=1e-3), (i, key)
        assert abs(g.audio.clipped - w.audio.clipped) <= max(2, w.audio.clipped // 1000)
        if not g.audio.measured or group_minimum:
            continue
        # The GPU's measures of its own PCM, restated with numpy.
        data = gpcm[g.offset * channels:(g.offset + g.data_frames) * channels]
        if g.status != 0:
            continue  # rejected jobs publish no PCM; their measures are checked against the oracle above
        p = plans[i]
        m = quality_measures(data, channels, g.frames, p.body_frames, [p.onsets[k] for k in range(p.onset_count)], a)
        for key in ("median", "quartile", "silent_fraction", "clipped_fraction", "body_rms"):
            assert getattr(g.audio, key) == pytest.approx(m[key], rel=1e-12, abs=1e-12), (i, key)
        assert (g.audio.longest_silence, g.audio.clipped) == (m["longest_silence"], m["clipped"])
        assert quality_verdict(m, g.frames, g.peak, a) == abi.TAIL_STATUS[g.status]
