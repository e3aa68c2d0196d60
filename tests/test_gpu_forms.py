"""Execution forms of the HIP render core produce identical PCM16.

- Plain form, one segment per batch vs segments of 256 .. 65536 frames: the
  voice chains carry their FP64 state (filter sections, smoothed and target
  coefficients, pitch phase) across segment boundaries, so every segmentation
  gives the same bits (the reference checks the same property between its own
  configurations, tests/check_segments.py:56).
- Persistent form (resident units, rolling admission, rounds replayed from a
  captured hipGraph, tuned or fixed launch shapes) vs the plain form, with
  uneven job lengths, refills, volume groups and every tail policy.
"""

import numpy as np
import pytest

import claysampler_amd as ca
from claysampler_amd import abi, synth

from helpers import assert_parity, group_case, oracle_render, pedal_case, rich_case, segment_case

pytestmark = pytest.mark.gpu


def resident_render(session, policy, taps, segment_frames=0, quality=False, group_minimum=0):
    bank_ptr, bank_count = session.bank()
    r = ca.Renderer(policy=policy, taps=taps, rate=session.rate, segment_frames=segment_frames, quality=quality,
                    group_minimum=group_minimum)
    r.upload_bank(bank_ptr, bank_count)
    b = r.prepare(session.plans())
    b.run()
    out = b.results()
    b.close()
    stats = r.stats()
    r.close()
    return out, stats


def full_tuple(r):
    a = r.audio
    return (r.status, r.frames, r.data_frames, r.fade, r.peak, r.first, r.last, r.scale, a.median, a.quartile,
            a.silent_fraction, a.clipped_fraction, a.body_rms, a.target_rms, a.gain, a.longest_silence, a.clipped,
            a.verdict, a.measured)


def per_job(results, pcm, channels):
    return [(full_tuple(r), pcm[r.offset * channels:(r.offset + r.data_frames) * channels].copy()) for r in results]


def window_render(session, policy, taps, form, batch_jobs, segment_frames=0, flags=0, split=None, quality=False,
                  group_minimum=0):
    """clay_hip_render_window; jobs keyed by index. `split` hands the plans
    over in pieces through refill."""
    bank_ptr, bank_count = session.bank()
    r = ca.Renderer(policy=policy, taps=taps, rate=session.rate, form=form, batch_jobs=batch_jobs,
                    segment_frames=segment_frames, flags=flags, quality=quality, group_minimum=group_minimum)
    r.upload_bank(bank_ptr, bank_count)
    plans = session.plans()
    pieces = [plans] if not split else [
        (abi.PlanView * (b - a))(*plans[a:b]) for a, b in zip([0] + split, split + [len(plans)])]
    jobs = {}

    def on_batch(results, pcm, first, failed):
        assert not failed
        for k, item in enumerate(per_job(results, pcm, r.channels)):
            assert first + k not in jobs
            jobs[first + k] = item

    r.render_window(pieces[0], on_batch, refills=pieces[1:])
    stats = r.stats()
    r.close()
    assert sorted(jobs) == list(range(len(plans)))
    return [jobs[i] for i in range(len(plans))], stats


def assert_same_jobs(a, b):
    assert len(a) == len(b)
    for i, ((ra, pa), (rb, pb)) in enumerate(zip(a, b)):
        assert ra == rb, (i, ra, rb)
        assert np.array_equal(pa, pb), (i, int(np.abs(pa.astype(np.int32) - pb.astype(np.int32)).max()))


def c2_session(synth_root, jobs=3, seconds=8.0):
    sfz = synth.c2_instrument(f"{synth_root}/c2")
    s = ca.Session()
    for j in range(jobs):
        s.add_job(sfz, synth.c2_midi(f"{synth_root}/c2/short_{j}.mid", 500 + j, seconds=seconds))
    return s


def uneven_session(synth_root, policy=abi.POLICY_NATURAL):
    """Configuration 3 in miniature: the configuration-1 instrument, uneven
    lengths and polyphony, plus two filtered jobs of configuration 2."""
    s = ca.Session(policy=policy)
    sfz1 = synth.c1_instrument(f"{synth_root}/c1")
    for j, (seconds, poly) in enumerate(((2.0, 4), (9.0, 40), (1.0, 8), (5.0, 64), (3.5, 16), (0.5, 4))):
        s.add_job(sfz1, synth.c3_midi(f"{synth_root}/c1/uneven_{j}.mid", 700 + j, seconds, poly))
    sfz2 = synth.c2_instrument(f"{synth_root}/c2")
    for j, seconds in enumerate((6.0, 2.5)):
        s.add_job(sfz2, synth.c2_midi(f"{synth_root}/c2/uneven_{j}.mid", 800 + j, seconds=seconds))
    return s


@pytest.mark.parametrize("segment", [256, 4096, 65536])
def test_plain_segments_bit_exact(synth_root, segment):
    s = c2_session(synth_root)
    (whole, _), (pieces, stats) = (resident_render(s, abi.POLICY_NATURAL, 32),
                                   resident_render(s, abi.POLICY_NATURAL, 32, segment_frames=segment))
    assert stats.segments > 1
    assert_same_jobs(per_job(*whole, 1), per_job(*pieces, 1))


@pytest.mark.parametrize("case", ["rich", "segment", "pedal"])
def test_plain_segments_feature_cases(tmp_path, case):
    """Variable pitch, multi-section filters with envelopes, EQ, stereo, loops."""
    if case == "rich":
        sfz, mid = rich_case(tmp_path)
        mids = [mid]
    elif case == "segment":
        sfz, long_mid, short_mid = segment_case(tmp_path)
        mids = [long_mid, short_mid]
    else:
        sfz, by_name = pedal_case(tmp_path)
        mids = list(by_name.values())
    for policy, channels in ((abi.POLICY_NATURAL, 1), (abi.POLICY_ORIG, 2)):
        s = ca.Session(policy=policy)
        for m in mids:
            # CC120 ("held") and sostenuto are natural-policy only.
            if policy == abi.POLICY_NATURAL or not m.endswith(("held.mid", "sostenuto.mid")):
                s.add_job(sfz, m)
        (whole, _), (pieces, _) = (resident_render(s, policy, 48),
                                   resident_render(s, policy, 48, segment_frames=768))
        assert_same_jobs(per_job(*whole, channels), per_job(*pieces, channels))


@pytest.mark.parametrize("batch_jobs,segment,flags", [
    (1, 256, abi.FLAG_NO_TUNE),
    (3, 4096, abi.FLAG_NO_TUNE),
    (3, 4096, abi.FLAG_NO_TUNE | abi.FLAG_NO_GRAPH),
    (5, 0, 0),
    (64, 0, 0)])
def test_persistent_matches_plain(synth_root, batch_jobs, segment, flags):
    s = uneven_session(synth_root)
    plain, _ = window_render(s, abi.POLICY_NATURAL, 16, abi.FORM_PLAIN, 64)
    rolled, stats = window_render(s, abi.POLICY_NATURAL, 16, abi.FORM_PERSISTENT, batch_jobs, segment, flags)
    assert_same_jobs(plain, rolled)
    # Jobs are admitted in cohorts of 1, 2, 4, ... consecutive jobs, up to batch_jobs / 8.
    cohort, admitted, cohorts = max(1, batch_jobs // 8), 0, 0
    while admitted < len(plain):
        admitted += min(cohort, 1 << cohorts)
        cohorts += 1
    assert stats.admissions == cohorts
    assert stats.rounds >= 2
    if not flags & abi.FLAG_NO_GRAPH:
        assert stats.graph_launches == stats.rounds


@pytest.mark.parametrize("policy", [abi.POLICY_ORIG, abi.POLICY_CUSTOM])
def test_persistent_policies_and_refill(synth_root, policy):
    s = uneven_session(synth_root, policy)
    plain, _ = window_render(s, policy, 16, abi.FORM_PLAIN, 2)
    rolled, _ = window_render(s, policy, 16, abi.FORM_PERSISTENT, 3, 2048, abi.FLAG_NO_TUNE, split=[2, 5])
    assert_same_jobs(plain, rolled)


def test_persistent_groups_and_quality(tmp_path):
    presets, melody = group_case(tmp_path)
    s = ca.Session(measure=True, keep_onsets=True)
    for preset, group in [("soft", 1), ("quiet", 1), ("loud", 1), ("loud", 0), ("soft", 2), ("loud", 2)]:
        s.add_job(presets[preset], melody, group)
    plain, _ = window_render(s, abi.POLICY_NATURAL, 48, abi.FORM_PLAIN, 64, quality=True, group_minimum=2)
    rolled, _ = window_render(s, abi.POLICY_NATURAL, 48, abi.FORM_PERSISTENT, 2, 1024, abi.FLAG_NO_TUNE,
                              quality=True, group_minimum=2)
    assert_same_jobs(plain, rolled)


def test_persistent_against_oracle(synth_root):
    """The rolling form is also held to the oracle bar directly."""
    s = uneven_session(synth_root)
    rolled, _ = window_render(s, abi.POLICY_NATURAL, 16, abi.FORM_PERSISTENT, 3, 1024, abi.FLAG_NO_TUNE)
    want_results, want_pcm = oracle_render(s, abi.POLICY_NATURAL, 16)
    for r, (t, pcm) in zip(want_results, rolled):
        assert (t[0], t[1], t[2], t[3]) == (r.status, r.frames, r.data_frames, r.fade)
        want = want_pcm[r.offset:r.offset + r.data_frames]
        assert int(np.abs(pcm.astype(np.int32) - want.astype(np.int32)).max(initial=0)) <= 1


@pytest.mark.parametrize("hot", [False, True])
def test_straight_tail_statistics_fused_or_scanned(synth_root, tmp_path, hot):
    """A straight batch (dry plain voices, configuration 3's kind) rendered in
    one window stores the tail-block statistics and the unit-scale mono PCM16
    from the mix kernel; rendered in 1024- and 4096-frame windows it runs
    clayScanStereo and quantises every job from the float mix instead.
    Statuses, frame counts and PCM16 agree bit for bit, and with the FP64
    oracle. The hot batch (+3 dB) has jobs scaled below the unit scale (peak
    above full scale), which take the float path in either case."""
    if hot:
        sfz = synth.c1_instrument(str(tmp_path / "hot"), volume_db=3)
        plan = [(970 + j, 4.0, poly) for j, poly in enumerate((2, 4, 8, 16, 32, 64))]
    else:
        sfz = synth.c1_instrument(f"{synth_root}/c1")
        plan = [(950 + j, 5.0 + 2 * j, poly) for j, poly in enumerate((4, 24, 64, 9))]
    s = ca.Session()
    for j, (seed, seconds, poly) in enumerate(plan):
        s.add_job(sfz, synth.c3_midi(str(tmp_path / f"c3f_{j}.mid"), seed, seconds, poly))
    whole, _ = resident_render(s, abi.POLICY_NATURAL, 16)
    if hot:
        assert any(r.status == 0 and r.scale < 32768 for r in whole[0])  # status 0: ok
    for segment in (1024, 4096):
        split, _ = resident_render(s, abi.POLICY_NATURAL, 16, segment_frames=segment)
        assert [full_tuple(r) for r in split[0]] == [full_tuple(r) for r in whole[0]], segment
        assert np.array_equal(split[1], whole[1]), segment
    assert_parity(whole, oracle_render(s, abi.POLICY_NATURAL, 16))
