"""Shared test fixtures: generated inputs in the style of the reference's CLI
tests (tests/check_tail.py, check_group_audio.py, check_segments.py), and
helpers that render one set of plans through the HIP core and the oracle."""

import math
import os
import struct

import numpy as np

import claysampler_amd as ca
from claysampler_amd import abi, synth

RATE = 44100


def smf(path, data):
    """Format 0, PPQ 480 MIDI around raw track bytes (tests/check_tail.py:14-16)."""
    with open(path, "wb") as f:
        f.write(b"MThd" + struct.pack(">IHHH", 6, 0, 1, 480) + b"MTrk" + struct.pack(">I", len(data)) + data)
    return str(path)


def wav(path, samples, rate=RATE, channels=1):
    synth.write_wav(str(path), np.asarray(samples), rate, channels)
    return str(path)


def text(path, content):
    with open(path, "w") as f:
        f.write(content)
    return str(path)


def tail_case(root):
    """tests/check_tail.py:26-36: 0.4 s at 1600 then 3 s at 12, a one-shot region,
    a 0.1 s note followed by a 5 s end-of-track gap."""
    wav(root / "sample.wav", [1600] * (RATE * 2 // 5) + [12] * (RATE * 3))
    sfz = text(root / "sample.sfz", "<region> sample=sample.wav key=60 loop_mode=one_shot amp_veltrack=0\n")
    mid = smf(root / "source.mid", bytes.fromhex("00 90 3c 64 60 80 3c 00 a5 40 ff 2f 00"))
    return sfz, mid


def pedal_case(root):
    """tests/check_tail.py:83-90: looping region, pedal and sostenuto MIDIs."""
    if not os.path.exists(root / "sample.wav"):
        tail_case(root)
    loop = text(root / "loop.sfz", "<region> sample=sample.wav key=60 loop_mode=loop_continuous loop_start=0 "
                                   "loop_end=800 ampeg_release=0.1 amp_veltrack=0\n")
    mids = {
        "regular": "00 90 3c 64 60 80 3c 00 00 ff 2f 00",
        "held": "00 b0 40 7f 00 90 3c 64 60 80 3c 00 00 b0 78 00 a5 40 ff 2f 00",
        "later": "00 b0 40 7f 00 90 3c 64 60 80 3c 00 82 68 b0 40 00 00 ff 2f 00",
        "sostenuto": "00 90 3c 64 30 b0 42 7f 30 80 3c 00 82 68 b0 42 00 00 ff 2f 00",
    }
    return loop, {k: smf(root / f"{k}.mid", bytes.fromhex(v)) for k, v in mids.items()}


def segment_case(root):
    """tests/check_segments.py:17-28: 3000 sin(0.04 i) loop, filter + CC1 cutoff/pitch."""
    wav(root / "seg.wav", [round(3000 * math.sin(i * .04)) for i in range(RATE)])
    sfz = text(root / "voice.sfz", "<region> sample=seg.wav key=60 loop_mode=loop_continuous loop_start=0 "
                                   "loop_end=44099 amp_veltrack=0 ampeg_release=1.7 cutoff=1800 resonance=3 "
                                   "cutoff_oncc1=3000 pitch_oncc1=37\n")
    long_mid = smf(root / "long.mid", bytes.fromhex(
        "00 b0 40 7f 00 90 3c 64 87 40 b0 01 40 87 40 e0 00 50 87 40 80 3c 00 87 40 b0 40 00 00 ff 2f 00"))
    short_mid = smf(root / "short.mid", bytes.fromhex("00 90 3c 64 60 80 3c 00 00 ff 2f 00"))
    return sfz, long_mid, short_mid


def group_case(root):
    """tests/check_group_audio.py:23-31: constant-level one-shot presets and a 0.5 s melody."""
    melody = smf(root / "melody.mid", bytes.fromhex("00 90 3c 64 83 60 80 3c 00 00 ff 2f 00"))
    presets = {}
    for name, value in [("soft", 1600), ("loud", 6400), ("quiet", 100), ("tiny", 8)]:
        wav(root / f"{name}.wav", [value] * RATE)
        presets[name] = text(root / f"{name}.sfz",
                             f"<region> sample={name}.wav key=60 loop_mode=one_shot amp_veltrack=0\n")
    return presets, melody


def chain_shapes_case(root):
    """One voice chain of each compiled filter-chain shape: stereo with one
    two-pole section, mono with two sections (lpf_4p), mono two-pole with one
    binding (cutoff CC1) and with two (cutoff and resonance CC1), and a
    high-pass; CC1 moves in steps and ramps, and notes overlap and bend."""
    n = np.arange(RATE)
    left = np.rint(8000 * np.sin(2 * np.pi * 220 * n / RATE) + 2000 * np.sin(2 * np.pi * 3100 * n / RATE))
    right = np.rint(6000 * np.sin(2 * np.pi * 331 * n / RATE + 0.5) + 1500 * np.sin(2 * np.pi * 4700 * n / RATE))
    wav(root / "shape_st.wav", np.stack([left, right], 1).reshape(-1), channels=2)
    mono = np.rint(9000 * np.sin(2 * np.pi * 180 * n / RATE) + 2500 * np.sin(2 * np.pi * 2900 * n / RATE))
    wav(root / "shape_mono.wav", mono)
    sfz = text(root / "shapes.sfz", (
        "<region> sample=shape_st.wav lokey=40 hikey=49 pitch_keycenter=45 fil_type=lpf_2p cutoff=400 "
        "cutoff_oncc1=4000 resonance=4 loop_mode=loop_continuous loop_start=100 loop_end=40000 ampeg_release=0.2\n"
        "<region> sample=shape_mono.wav lokey=50 hikey=59 pitch_keycenter=55 fil_type=lpf_4p cutoff=700 "
        "cutoff_oncc1=3000 resonance=2 ampeg_release=0.15\n"
        "<region> sample=shape_mono.wav lokey=60 hikey=69 pitch_keycenter=64 fil_type=lpf_2p cutoff=250 "
        "cutoff_oncc1=5000 resonance=6 loop_mode=loop_continuous loop_start=0 loop_end=44000 ampeg_release=0.3\n"
        "<region> sample=shape_mono.wav lokey=70 hikey=79 pitch_keycenter=74 fil_type=lpf_2p cutoff=900 "
        "cutoff_oncc1=2400 resonance_oncc1=9 ampeg_release=0.1\n"
        "<region> sample=shape_mono.wav lokey=80 hikey=89 pitch_keycenter=84 fil_type=hpf_2p cutoff=1500 "
        "cutoff_oncc1=-1200 resonance=3 ampeg_release=0.1\n"))
    events = []
    rng = synth.SplitMix64(91)
    tick = 0
    for _ in range(20):
        key, vel = rng.integer(40, 89), rng.integer(40, 127)
        dur = rng.integer(150, 900)
        events.append((tick, 2, 0x90, key, vel))
        events.append((tick + dur, 1, 0x80, key, 0))
        events.append((tick + dur // 3, 0, 0xB0, 1, rng.integer(0, 127)))
        events.append((tick + dur // 2, 0, 0xE0, rng.integer(0, 127), rng.integer(0, 127)))
        tick += rng.integer(80, 300)
    mid = str(root / "shapes.mid")
    synth.write_midi(mid, events)
    return sfz, mid


def gain_case(root):
    """Volume (CC7) and expression (CC11) moving under held notes, so that the
    job gain changes inside voice spans, for filtered mono and stereo voices
    and a dry one; the controllers also step between notes."""
    sfz, _ = chain_shapes_case(root)
    text(root / "gain.sfz", open(sfz).read() + "<region> sample=shape_mono.wav lokey=90 hikey=99 pitch_keycenter=94 "
         "ampeg_release=0.2\n")
    events = []
    rng = synth.SplitMix64(123)
    tick = 0
    for _ in range(16):
        key, vel = rng.integer(40, 99), rng.integer(50, 127)
        dur = rng.integer(300, 1200)
        events.append((tick, 2, 0x90, key, vel))
        events.append((tick + dur, 1, 0x80, key, 0))
        for k in range(1, 4):
            events.append((tick + k * dur // 4, 0, 0xB0, 7 if k % 2 else 11, rng.integer(20, 127)))
        tick += rng.integer(100, 500)
    mid = str(root / "gain.mid")
    synth.write_midi(mid, events)
    return str(root / "gain.sfz"), mid


def rich_case(root):
    """Stereo sample with pan/width/position, loop crossfade, generators,
    velocity-modulated volume, a multi-section filter and a pitch-CC voice."""
    n = np.arange(RATE)
    left = np.rint(9000 * np.sin(2 * np.pi * 330 * n / RATE) * np.exp(-n / 30000))
    right = np.rint(7000 * np.sin(2 * np.pi * 495 * n / RATE + 1) * np.exp(-n / 20000))
    wav(root / "stereo.wav", np.stack([left, right], 1).reshape(-1), channels=2)
    sfz = text(root / "rich.sfz", (
        "<group> amp_veltrack=50\n"
        "<region> sample=stereo.wav lokey=40 hikey=59 pitch_keycenter=50 pan=-30 width=60 position=20 "
        "loop_mode=loop_continuous loop_start=1000 loop_end=30000 loop_crossfade=0.05 ampeg_release=0.25\n"
        "<region> sample=stereo.wav lokey=60 hikey=71 pitch_keycenter=64 fil_type=lpf_4p cutoff=900 "
        "resonance=2 fileg_depth=1200 fileg_attack=0.1 fileg_decay=0.3 fileg_sustain=20 ampeg_attack=0.02\n"
        "<region> sample=*sine lokey=72 hikey=80 pitch_keycenter=69 ampeg_release=0.1 volume=-6\n"
        "<region> sample=*noise lokey=81 hikey=84 volume=-30 ampeg_release=0.05\n"
        "<region> sample=stereo.wav lokey=85 hikey=96 pitch_keycenter=90 pitch_oncc1=300 eq1_freq=2000 "
        "eq1_gain=6 eq1_bw=1 loop_mode=one_shot\n"))
    events = []
    rng = synth.SplitMix64(77)
    tick = 0
    for _ in range(24):
        key, vel = rng.integer(40, 96), rng.integer(30, 127)
        dur = rng.integer(100, 700)
        events.append((tick, 2, 0x90, key, vel))
        events.append((tick + dur, 1, 0x80, key, 0))
        events.append((tick + dur // 2, 0, 0xB0, 1, rng.integer(0, 127)))
        events.append((tick + dur // 3, 0, 0xE0, rng.integer(0, 127), rng.integer(0, 127)))
        tick += rng.integer(60, 400)
    mid = str(root / "rich.mid")
    synth.write_midi(mid, events)
    return sfz, mid


def gpu_render(session, policy, taps, quality=False, group_minimum=0, batch_jobs=64):
    bank_ptr, bank_count = session.bank()
    plans = session.plans()
    r = ca.Renderer(policy=policy, taps=taps, quality=quality, group_minimum=group_minimum, batch_jobs=batch_jobs,
                    rate=session.rate)
    r.upload_bank(bank_ptr, bank_count)
    b = r.prepare(plans)
    b.run()
    results, pcm = b.results()
    b.close()
    r.close()
    return results, pcm


class OracleOutput(tuple):
    """(results, pcm) of the oracle, plus `mix_of(i)`: job i's FP64 float mix
    (frames x 2), computed on demand."""


def oracle_render(session, policy, taps, quality=False, group_minimum=0, threads=4):
    import oracle

    bank_ptr, bank_count = session.bank()
    plans = session.plans()
    cfg = oracle.config(policy=policy, taps=taps, quality=quality, group_minimum=group_minimum, threads=threads,
                        rate=session.rate)
    out = OracleOutput(oracle.render(cfg, bank_ptr, bank_count, plans))
    out.mix_of = lambda i: oracle.mix(cfg, bank_ptr, bank_count, plans[i])
    out.session = session  # keeps the plans' memory alive
    return out


def near_tie(x, tolerance=1e-3):
    """True where x lies within `tolerance` of a half-integer (a rounding tie)."""
    frac = x - np.floor(x)
    return np.abs(frac - 0.5) < tolerance


def mono_tie_flips(want, bad):
    """Natural-policy samples (indices into the packed mono PCM) whose 2-step
    difference is explained by rounding: the FP64 mix puts both channels
    within 1e-3 LSB of a .5 tie, so float32 may round each the other way,
    and the ties-to-even average of the two PCM16 channels (src/cuda/tail.cu:
    164-165) turns two 1-LSB channel flips into a 2-step mono difference."""
    wres = want[0]
    explained = np.zeros(len(bad), bool)
    for i, r in enumerate(wres):
        inside = (bad >= r.offset) & (bad < r.offset + r.data_frames)
        if not inside.any():
            continue
        mix = want.mix_of(i)
        frames = bad[inside] - r.offset
        explained[inside] = near_tie(mix[frames, 0] * r.scale) & near_tie(mix[frames, 1] * r.scale)
    return explained


def result_tuple(r):
    return (r.status, r.frames, r.data_frames, r.fade, r.offset)


def assert_parity(got, want, lsb=1, extent=True):
    """Frame counts and statuses exact; PCM within `lsb`; peak/first/last consistent.
    For the natural policy's mono PCM a difference of lsb + 1 is accepted where
    mono_tie_flips explains it (both channels of the FP64 mix at a rounding tie)."""
    (gres, gpcm), (wres, wpcm) = got, want
    assert [result_tuple(r) for r in gres] == [result_tuple(r) for r in wres]
    assert gpcm.shape == wpcm.shape
    if gpcm.size:
        diff = np.abs(gpcm.astype(np.int32) - wpcm.astype(np.int32))
        over = np.nonzero(diff > lsb)[0]
        mono = sum(r.data_frames for r in wres) == wpcm.size
        if over.size and mono and hasattr(want, "mix_of") and int(diff.max()) == lsb + 1 and all(
                r.audio.gain == 0 for r in wres):
            explained = mono_tie_flips(want, over)
            over = over[~explained]
        assert over.size == 0, f"max |GPU - oracle| = {int(diff[over].max())} LSB at {int(over[0])}"
    if extent:
        for g, w in zip(gres, wres):
            assert abs(g.peak - w.peak) <= lsb + 1
            assert abs(g.first - w.first) <= 64 and abs(g.last - w.last) <= 64
    return gres, gpcm


def quality_cases(root):
    """One scenario per quality / group verdict (src/cuda/quality.cu:107-209):
    name -> (policy, jobs [(sfz, midi, group)], audio config, group minimum,
    expected status of each job)."""
    presets, melody = group_case(root)
    wav(root / "zero.wav", [0] * RATE)
    presets["zero"] = text(root / "zero.sfz", "<region> sample=zero.wav key=60 loop_mode=one_shot amp_veltrack=0\n")
    wav(root / "full.wav", [32767] * RATE)
    presets["hot"] = text(root / "hot.sfz",
                          "<region> sample=full.wav key=60 loop_mode=one_shot amp_veltrack=0 volume=20\n")
    gap = smf(root / "gap2.mid", bytes.fromhex("00 90 3c 64 60 80 3c 00 96 40 90 3c 64 60 80 3c 00 00 ff 2f 00"))
    notes = bytearray()
    for k, vel in enumerate((100, 40, 127, 70, 90, 20)):
        notes += bytes([0x00, 0x90, 0x3C, vel, 0x81, 0x70, 0x80, 0x3C, 0x00])
        notes += bytes([0x60 if k % 2 else 0x20, 0xB0, 0x07, 127 - 10 * k])
    notes += bytes.fromhex("00 ff 2f 00")
    phrase = smf(root / "phrase.mid", bytes(notes))
    base = abi.default_audio_config(RATE, True)

    def audio(**fields):
        a = abi.AudioConfig.from_buffer_copy(base)
        for k, v in fields.items():
            setattr(a, k, v)
        return a

    return {
        "ok": (abi.POLICY_NATURAL, [(presets["soft"], phrase, 0)], base, 0, ["ok"]),
        "quietPerformance": (abi.POLICY_NATURAL, [(presets["quiet"], melody, 0)], base, 0, ["quietPerformance"]),
        "excessiveSilence": (abi.POLICY_NATURAL, [(presets["soft"], gap, 0)], audio(max_silent_frames=1.0 * RATE), 0,
                             ["excessiveSilence"]),
        "silentAudio": (abi.POLICY_NATURAL, [(presets["tiny"], melody, 0)],
                        audio(min_median=-200.0, min_quartile=-200.0, silence_power=0.0, max_silent_fraction=1.0),
                        0, ["silentAudio"]),
        "clipping": (abi.POLICY_ORIG, [(presets["hot"], melody, 0)], base, 0, ["clipping"]),
        "audioDuration": (abi.POLICY_NATURAL, [(presets["soft"], melody, 0)], audio(max_frames=RATE // 5), 0,
                          ["audioDuration"]),
        "silentBody": (abi.POLICY_NATURAL, [(presets["zero"], melody, 1), (presets["soft"], melody, 1),
                                            (presets["loud"], phrase, 2), (presets["soft"], phrase, 2)],
                       abi.default_audio_config(RATE, False), 2, ["silentBody", "incompleteGroup", "ok", "ok"]),
        "orig_metrics": (abi.POLICY_ORIG, [(presets["loud"], phrase, 0), (presets["soft"], gap, 0)], base, 0,
                         ["ok", "ok"]),
    }


def quality_session(case, rate=RATE):
    policy, jobs, _, group_minimum, _ = case
    s = ca.Session(policy=policy, measure=True, keep_onsets=True, rate=rate)
    for sfz, mid, group in jobs:
        s.add_job(sfz, mid, group)
    return s


def quality_measures(pcm, channels, frames, body, onsets, a):
    """Quality measures of one job's PCM, restated with numpy."""
    x = pcm.reshape(-1, channels).astype(np.int64)
    mono = 2 * x[:, 0] if channels == 1 else x[:, 0] + x[:, 1]
    power = mono * mono
    end = min(frames, body)
    span = 5 * a.window
    chunk = 2 * a.rate
    # Body windows: consecutive 250 ms windows restarting at every 2 s chunk.
    starts = [c + w for c in range(0, end, chunk) for w in range(0, chunk, span) if c + w < end]
    total, silent, run, longest = 0, 0, 0, 0
    for i, first in enumerate(starts):
        nxt = min(end, starts[i + 1] if i + 1 < len(starts) else first + min(span, chunk - (first % chunk)))
        stop = min(min(body, len(power)), nxt)
        wp = int(power[first:stop].sum())
        total += wp
        if wp < (nxt - first) * a.silence_power:
            silent += nxt - first
            run += nxt - first
            longest = max(longest, run)
        else:
            run = 0
    body_rms = math.sqrt(total / body) * 0.5 if body else 0.0
    levels = []
    for o in onsets:
        best = max(int(power[o + w * a.window:min(o + (w + 1) * a.window, len(power))].sum()) for w in range(5))
        levels.append(max(-120.0, 10 * math.log10(best / (a.window * 2.0 ** 32))) if best else -120.0)
    median = float(np.percentile(levels, 50)) if levels else -120.0
    quartile = float(np.percentile(levels, 25)) if levels else -120.0
    clipped = int((np.abs(x) >= 32767).sum())
    return dict(body_rms=body_rms, median=median, quartile=quartile, silent_fraction=silent / max(1, end),
                longest_silence=longest, clipped=clipped, clipped_fraction=clipped / (max(1, frames) * channels),
                levels=levels)


def quality_verdict(m, frames, peak, a):
    if frames > a.max_frames:
        return "audioDuration"
    if not m["levels"] or m["median"] < a.min_median or m["quartile"] < a.min_quartile:
        return "quietPerformance"
    if m["silent_fraction"] > a.max_silent_fraction or m["longest_silence"] > a.max_silent_frames:
        return "excessiveSilence"
    if peak < a.min_peak:
        return "silentAudio"
    if m["clipped_fraction"] > a.max_clipped_fraction:
        return "clipping"
    return "ok"
