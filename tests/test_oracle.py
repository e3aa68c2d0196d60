"""Pins of the FP64 oracle (CPU only). Each test checks the oracle against
something the oracle's own formulas do not produce: an independent numpy /
scipy restatement of the mathematics (Kaiser-sinc, RBJ biquads, DF2T
recurrence), exact invariants, or the known answers of the reference's own
CLI tests (tests/check_tail.py, tests/check_group_audio.py)."""

import ctypes as C
import math

import numpy as np
import pytest
import scipy.signal

import claysampler_amd as ca
import oracle
from claysampler_amd import abi

from helpers import group_case, pedal_case, tail_case
from planbuild import Plan, bank_pointer, span, views

RATE = 44100


# ------------------------------------------------------------------ sinc table (renderer.cpp:347-376)

@pytest.mark.parametrize("taps", [16, 32, 48, 64])
def test_sinc_table_equals_numpy_kaiser_sinc(taps):
    table = oracle.sinc_table(taps)
    t = np.arange(taps)[:, None]
    p = np.arange(2049)[None, :]
    x = t - (taps / 2 - 1) - p / 2048.0
    f = x / (taps / 2)
    window = np.i0(8.5 * np.sqrt(np.clip(1 - f * f, 0, None))) / np.i0(8.5)
    want = window * np.sinc(x)
    diff = np.abs(table.astype(np.float64) - want)
    assert np.all(diff <= np.spacing(np.abs(want).astype(np.float32)).astype(np.float64) * 0.51 + 1e-12)


@pytest.mark.parametrize("taps", [16, 32, 48, 64])
def test_sinc_table_invariants(taps):
    table = oracle.sinc_table(taps)
    centre = taps // 2 - 1
    assert table[centre, 0] == 1.0
    # Integer position: one unit tap; the others are sin(pi k) rounding residue.
    assert np.max(np.abs(np.delete(table[:, 0], centre))) < 1e-15
    # Mirror symmetry w[t][p] = w[T-1-t][P-p], exact.
    assert np.array_equal(table, table[::-1, ::-1])
    # Unit DC gain of the interpolator at every phase.
    assert np.max(np.abs(table.sum(axis=0) - 1)) < 2e-3


# ------------------------------------------------------------------ interpolation (render.cu:21-61)

def interpolate(s, bank, taps, pos):
    out = (C.c_double * 2)()
    b, ptr, _ = bank_pointer(bank)
    oracle.lib().clay_oracle_interpolate(C.byref(s), ptr, taps, pos, out)
    return out[0], out[1]


def test_integer_positions_return_the_sample():
    rng = np.random.default_rng(1)
    bank = rng.integers(-32768, 32767, 4000).astype(np.int16)
    s = span(0, 100, source_end=4000)
    for pos in (0, 1, 17, 1999, 3999):
        assert interpolate(s, bank, 32, float(pos)) == (bank[pos] / 32768.0, bank[pos] / 32768.0)
    assert interpolate(s, bank, 32, 4000.0) == (0.0, 0.0)  # outside the source


def test_dc_source_gives_dc_times_tap_sum():
    bank = np.full(5000, 1000, np.int16)
    s = span(0, 100, source_end=5000)
    table = oracle.sinc_table(48)
    for pos in (100.25, 2500.5, 3000.0001):
        phase = int(np.rint((pos - math.floor(pos)) * 2048))
        want = 1000 / 32768.0 * float(np.sum(table[:, phase].astype(np.float64)))
        got = interpolate(s, bank, 48, pos)
        assert abs(got[0] - want) < 1e-14


@pytest.mark.parametrize("channels", [1, 2])
def test_interpolation_brute_force(channels):
    rng = np.random.default_rng(channels)
    frames = 300
    bank = rng.integers(-32768, 32767, frames * channels).astype(np.int16)
    s = span(0, 100, source_end=frames, channels=channels)
    for taps in (16, 64):
        table = oracle.sinc_table(taps).astype(np.float64)
        for pos in list(rng.uniform(-3, frames + 3, 60)) + [0.5, frames - 0.5, 23.5]:
            centre = math.floor(pos)
            if pos == centre:
                continue
            phase = int(np.rint((pos - centre) * 2048))
            want = np.zeros(2)
            for t in range(taps):
                i = centre - (taps // 2 - 1) + t
                if 0 <= i < frames:
                    sample = bank[i * channels: i * channels + channels] / 32768.0
                    want += table[t, phase] * (sample if channels == 2 else np.array([sample[0], sample[0]]))
            got = interpolate(s, bank, taps, pos)
            assert np.allclose(got, want, rtol=0, atol=1e-13), (pos, got, want)


# ------------------------------------------------------------------ envelope (render.cu:6-19)

def test_envelope_stages_closed_form():
    s = span(100, 50000, attack_start=150, hold_start=1150, decay_start=2150, attack_step=1 / 1000.0, start=0.2,
             decay_rate=-9 / (0.3 * RATE), sustain=0.4)
    env = oracle.lib().clay_oracle_envelope
    assert env(C.byref(s), 120) == pytest.approx(0.2, abs=1e-7)
    assert env(C.byref(s), 500) == pytest.approx(min(1, 0.2 + (500 - 150 + 1) / 1000.0), abs=1e-7)
    assert env(C.byref(s), 1500) == 1.0
    n = 3000
    assert env(C.byref(s), n) == pytest.approx(max(0.4, math.exp((n - 2150 + 1) * (-9 / (0.3 * RATE)))), rel=1e-6)
    assert env(C.byref(s), 40000) == pytest.approx(0.4, abs=1e-7)
    # Release then linear fade (plan.cpp:237-248 shape).
    s.release_start, s.release_value, s.release_rate = 20000, 0.4, -9 / (0.1 * RATE)
    s.fade_start, s.fade_value, s.fade_step = 22000, 1e-4, 1e-4 / 2205
    assert env(C.byref(s), 21000) == pytest.approx(0.4 * math.exp(1001 * (-9 / (0.1 * RATE))), rel=1e-6)
    assert env(C.byref(s), 22000 + 1102) == pytest.approx(1e-4 - 1103 * 1e-4 / 2205, rel=1e-5)
    assert env(C.byref(s), 22000 + 5000) == 0.0


# ------------------------------------------------------------------ biquads (render.cu:278-336)

def rbj(kind, f, res, fs=RATE, bw=1.0, gain=0.0):
    """Audio EQ Cookbook (R. Bristow-Johnson) biquads, normalised by a0."""
    w = 2 * math.pi * min(min(20000.0, fs * .499), max(1.0, f)) / fs
    cw, sw = math.cos(w), math.sin(w)
    if kind in ("low1", "high1"):
        p = math.exp(-w)
        return [1 - p, 0, 0, -p, 0] if kind == "low1" else [(1 + p) / 2, -(1 + p) / 2, 0, -p, 0]
    if kind == "peak":
        A = 10 ** (gain / 40)
        alpha = sw * math.sinh(math.log(2) / 2 * bw * w / sw)
        b, a = [1 + alpha * A, -2 * cw, 1 - alpha * A], [1 + alpha / A, -2 * cw, 1 - alpha / A]
    else:
        alpha = sw / (2 * 10 ** (res / 20))
        a = [1 + alpha, -2 * cw, 1 - alpha]
        b = {"low": [(1 - cw) / 2, 1 - cw, (1 - cw) / 2], "high": [(1 + cw) / 2, -(1 + cw), (1 + cw) / 2],
             "band": [alpha, 0, -alpha], "notch": [1, -2 * cw, 1]}[kind]
    return [b[0] / a[0], b[1] / a[0], b[2] / a[0], a[1] / a[0], a[2] / a[0]]


@pytest.mark.parametrize("ftype,kind", [(1, "low1"), (2, "high1"), (3, "low"), (4, "high"), (5, "band"), (6, "notch"),
                                        (7, "low"), (8, "high"), (9, "band"), (10, "low"), (11, "high"),
                                        (12, "band"), (13, "peak")])
def test_filter_coefficients_match_the_cookbook(ftype, kind):
    out = (C.c_double * 5)()
    for f, res, bw, g in ((200.0, 6.0, 1.0, 6.0), (1800.0, 3.0, 0.5, -12.0), (9000.0, -2.0, 2.0, 3.0)):
        oracle.lib().clay_oracle_filter_coefficients(ftype, f, res, RATE, bw, g, out)
        want = rbj(kind, f, res, bw=bw, gain=g)
        assert np.allclose(list(out), want, rtol=1e-12, atol=1e-14), (ftype, f)


@pytest.mark.parametrize("ftype,sections", [(3, 1), (7, 2), (10, 3)])
def test_static_filter_recurrence_equals_lfilter(ftype, sections):
    """A static filter on one mono voice: oracle wet mix == scipy lfilter of the dry mix."""
    rng = np.random.default_rng(ftype)
    bank = rng.integers(-20000, 20000, 6000).astype(np.int16)
    b, ptr, n = bank_pointer(bank)
    dry = span(0, 5000, source_end=6000, step=0.73, gain=0.8, pan=0.25)
    wet = span(0, 5000, source_end=6000, step=0.73, gain=0.8, pan=0.25, filter_track=0)
    wet.filters[0].type, wet.filters[0].cutoff, wet.filters[0].resonance = ftype, 700.0, 4.0
    cfg = oracle.config(taps=32)
    x = oracle.mix(cfg, ptr, n, Plan([dry], 5000).view)
    y = oracle.mix(cfg, ptr, n, Plan([wet], 5000, filter_tracks=1).view)
    c = rbj("low", 700.0, 4.0)
    want = x.copy()
    for _ in range(sections):
        want = scipy.signal.lfilter([c[0], c[1], c[2]], [1, c[3], c[4]], want, axis=0)
    assert np.max(np.abs(y - want)) < 1e-12


# ------------------------------------------------------------------ PCM16 rules (tail.cu:147-174, :46-51)

def test_mono_average_rounds_ties_to_even():
    mono = oracle.lib().clay_oracle_mono
    grid = range(-40, 41)
    for l in grid:
        for r in grid:
            assert mono(l, r) == round((l + r) / 2), (l, r)
    rng = np.random.default_rng(5)
    for l, r in rng.integers(-32768, 32768, (20000, 2)).tolist():
        assert mono(l, r) == round((l + r) / 2)


def test_mono_average_turns_two_channel_flips_into_two_steps():
    """When both channels of a frame sit at a .5 tie, rounding each the other
    way moves their sum by 2, and the ties-to-even average then moves by 2:
    the natural policy's mono PCM can differ by 2 where the float mix differs
    by far less than 1 LSB (the GPU parity check accepts exactly these)."""
    mono = oracle.lib().clay_oracle_mono
    assert mono(-4258, -4311) == -4284 and mono(-4259, -4312) == -4286
    assert mono(10, 11) == 10 and mono(11, 12) == 12


# ------------------------------------------------------------------ known answers of the reference tests

def session_render(policy, jobs, taps=48, **kw):
    quality = kw.pop("quality", False)
    group_minimum = kw.pop("group_minimum", 0)
    s = ca.Session(policy=policy, measure=quality or group_minimum > 0, keep_onsets=quality, **kw)
    for job in jobs:
        s.add_job(*job)
    bp, bc = s.bank()
    res, pcm = oracle.render(oracle.config(policy=policy, taps=taps, quality=quality, group_minimum=group_minimum),
                             bp, bc, s.plans())
    return res, pcm


def job_pcm(res, pcm, i, channels):
    r = res[i]
    return pcm[r.offset * channels:(r.offset + r.data_frames) * channels]


def test_natural_tail_known_answers(tmp_path):
    """tests/check_tail.py:50-61: 0.65 s, audible tail, silent last sample, OK-line extent."""
    sfz, mid = tail_case(tmp_path)
    res, pcm = session_render(abi.POLICY_NATURAL, [(sfz, mid)])
    r = res[0]
    assert (r.status, r.frames) == (0, round(.65 * RATE))
    data = job_pcm(res, pcm, 0, 1).astype(np.int32)
    assert data[r.frames - 5000] > 3 and data[-1] == 0
    audible = np.nonzero(np.abs(data) > 3)[0]
    assert (r.peak, r.first, r.last) == (int(np.abs(data).max()), int(audible[0]), int(audible[-1]))


def test_orig_and_custom_known_answers(tmp_path):
    """tests/check_tail.py:63-77."""
    sfz, mid = tail_case(tmp_path)
    res, pcm = session_render(abi.POLICY_ORIG, [(sfz, mid)])
    assert res[0].frames == round(5.1 * RATE)
    orig = job_pcm(res, pcm, 0, 2)
    res, pcm = session_render(abi.POLICY_CUSTOM, [(sfz, mid)], tail_max_seconds=1.0, tail_fade_ms=100)
    dres, dpcm = session_render(abi.POLICY_CUSTOM, [(sfz, mid)], tail_max_seconds=1.0, tail_fade_ms=0)
    assert res[0].frames == round(1.1 * RATE) + 1
    custom, dry = job_pcm(res, pcm, 0, 2), job_pcm(dres, dpcm, 0, 2)
    assert np.array_equal(custom[:-4410 * 2], dry[:-4410 * 2])
    assert np.all(custom[-2:] == 0) and np.any(dry[-2:] != 0)
    res, pcm = session_render(abi.POLICY_CUSTOM, [(sfz, mid)], tail_max_seconds=4.0)
    assert res[0].frames == round(3.4 * RATE)
    assert np.array_equal(job_pcm(res, pcm, 0, 2), orig[:res[0].frames * 2])


def test_max_audio_seconds_rejects(tmp_path):
    """tests/check_tail.py:79-81: 0.5 s limit -> duration rejection."""
    sfz, mid = tail_case(tmp_path)
    res, _ = session_render(abi.POLICY_NATURAL, [(sfz, mid)], max_audio_seconds=0.5)
    assert abi.TAIL_STATUS[res[0].status] == "duration" and res[0].data_frames == 0


def test_pedal_equivalences(tmp_path):
    """tests/check_tail.py:91-95."""
    sfz, mids = pedal_case(tmp_path)
    res, pcm = session_render(abi.POLICY_NATURAL, [(sfz, mids[k]) for k in ("regular", "held", "later", "sostenuto")])
    jobs = [job_pcm(res, pcm, i, 1) for i in range(4)]
    assert res[0].frames == res[1].frames and np.array_equal(jobs[0], jobs[1])
    assert res[2].frames == res[3].frames and np.array_equal(jobs[2], jobs[3])
    assert res[2].frames > res[1].frames + RATE // 4
    assert res[1].frames % 1024 == 0


def test_peak_repair_and_rejection(tmp_path):
    """tests/check_tail.py:103-112: repair to -1 dBFS (+-1), reject beyond 5 dB."""
    tail_case(tmp_path)
    base = "<region> sample=sample.wav key=60 loop_mode=one_shot amp_veltrack=0 pan_oncc10=0"
    (tmp_path / "hot.sfz").write_text(base + " volume=33\n")
    res, pcm = session_render(abi.POLICY_NATURAL, [(str(tmp_path / "hot.sfz"), str(tmp_path / "source.mid"))])
    assert abs(int(np.abs(job_pcm(res, pcm, 0, 1).astype(np.int32)).max()) - round(32768 * 10 ** (-1 / 20))) <= 1
    (tmp_path / "hot.sfz").write_text(base + " volume=42\n")
    res, _ = session_render(abi.POLICY_NATURAL, [(str(tmp_path / "hot.sfz"), str(tmp_path / "source.mid"))])
    assert abi.TAIL_STATUS[res[0].status] == "peak"


def test_silent_padding(tmp_path):
    """tests/check_tail.py:114-125: a long MIDI gap after a short sample pads silence to 65 x 1024 frames."""
    from helpers import smf, text, wav

    wav(tmp_path / "short.wav", [1600] * (RATE * 2 // 5))
    sfz = text(tmp_path / "short.sfz", "<region> sample=short.wav key=60 loop_mode=one_shot amp_veltrack=0\n")
    gap = smf(tmp_path / "gap.mid", bytes.fromhex("00 90 3c 64 8b 20 80 3c 00 00 ff 2f 00"))
    res, pcm = session_render(abi.POLICY_NATURAL, [(sfz, gap)])
    assert res[0].frames == 65 * 1024
    data = job_pcm(res, pcm, 0, 1)
    assert not np.any(data[RATE * 2 // 5:])


def test_group_gain_known_answer(tmp_path):
    """tests/check_group_audio.py:70-79: shared target = mean body RMS; x' = round(x * target / rms)."""
    presets, melody = group_case(tmp_path)
    off_res, off_pcm = session_render(abi.POLICY_NATURAL, [(presets["soft"], melody), (presets["loud"], melody)])
    soft, loud = (job_pcm(off_res, off_pcm, i, 1).astype(np.float64) for i in range(2))
    body = RATE // 2
    rms = [math.sqrt(np.sum(x[:body] ** 2) / body) for x in (soft, loud)]
    target = sum(rms) / 2
    on_res, on_pcm = session_render(abi.POLICY_NATURAL, [(presets["soft"], melody, 1), (presets["quiet"], melody, 1),
                                                         (presets["loud"], melody, 1), (presets["loud"], melody, 0)],
                                    quality=True, group_minimum=2)
    assert abi.TAIL_STATUS[on_res[1].status] == "quietPerformance"
    for i, x, level in ((0, soft, rms[0]), (2, loud, rms[1])):
        got = job_pcm(on_res, on_pcm, i, 1)
        want = np.array([round(v * (target / level)) for v in x], dtype=np.int64)
        assert np.array_equal(got.astype(np.int64), want)
    assert on_res[3].audio.gain == 0 and np.array_equal(job_pcm(on_res, on_pcm, 3, 1), job_pcm(off_res, off_pcm, 1, 1))


def test_group_peak_limit(tmp_path):
    """tests/check_group_audio.py:114-124: a spiky member limits the group to -1 dBFS of 32767."""
    from helpers import text, wav

    presets, melody = group_case(tmp_path)
    wav(tmp_path / "spiky.wav", [12000] * (RATE // 1000) + [100] * RATE)
    spiky = text(tmp_path / "spiky.sfz",
                 "<region> sample=spiky.wav key=60 loop_mode=one_shot amp_veltrack=0 volume=10\n")
    res, pcm = session_render(abi.POLICY_NATURAL, [(presets["loud"], melody, 1), (spiky, melody, 1)],
                              group_minimum=2)
    assert res[0].audio.target_rms == res[1].audio.target_rms
    assert res[0].audio.target_rms < (res[0].audio.body_rms + res[1].audio.body_rms) / 2
    assert int(np.abs(job_pcm(res, pcm, 1, 1).astype(np.int32)).max()) == round(32767 * 10 ** (-1 / 20))


def test_quality_verdicts(tmp_path):
    """tests/check_group_audio.py:85-104 verdict strings."""
    from helpers import smf

    presets, melody = group_case(tmp_path)
    res, _ = session_render(abi.POLICY_NATURAL, [(presets["quiet"], melody)], quality=True)
    assert abi.TAIL_MESSAGE[res[0].status] == "quiet_performance"
    gap = smf(tmp_path / "gap2.mid", bytes.fromhex("00 90 3c 64 60 80 3c 00 96 40 90 3c 64 60 80 3c 00 00 ff 2f 00"))
    audio = abi.default_audio_config(RATE, True)
    audio.max_silent_frames = 1.0 * RATE
    s = ca.Session(measure=True, keep_onsets=True)
    s.add_job(presets["soft"], gap)
    bp, bc = s.bank()
    res, _ = oracle.render(oracle.config(audio=audio), bp, bc, s.plans())
    assert abi.TAIL_MESSAGE[res[0].status] == "excessive_audio_silence"
