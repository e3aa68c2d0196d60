"""Known answers for the host planner and SFZ loader (claysampler_amd/csrc/host),
worked from the reference's rules rather than from the planner's code:

- region selection by key, velocity and CC window (src/render/plan.cpp:326-356,
  :930-958): every (note, region) pair is brute-forced against the ranges;
- bend splits (:114-134, :396-435): a held note cut at the bend's 1024-frame
  block and at the bend, with the closed-form step and exponential ramp;
- automation (:136-216): an unsmoothed controller ramp, and the smoothed
  volume track approaching a step change linearly over rate/100 frames;
- loop crossfade length (:736-741): lround(min(loop_start, seconds x rate));
- release (:200-222, :846-870): release start, rate, fade start and span end
  in closed form, immediately and under the sustain pedal;
- the loader rejects the features this render core leaves out.
Generators (*sine) stand in for samples where the sample data do not matter."""

import math
import struct

import numpy as np
import pytest

import claysampler_amd as ca
from claysampler_amd import abi

from helpers import text, wav

RATE = 44100
FRAMES_PER_TICK = RATE * 500000 / 480 / 1e6  # PPQ 480 at 120 BPM: 45.9375


def varlen(v):
    out = [v & 0x7F]
    v >>= 7
    while v:
        out.append(0x80 | (v & 0x7F))
        v >>= 7
    return bytes(reversed(out))


def midi(path, events):
    """events: (tick, status, a, b) in tick order; b None for two-byte messages."""
    track, last = bytearray(), 0
    for tick, status, a, b in events:
        track += varlen(tick - last) + bytes([status, a] + ([] if b is None else [b]))
        last = tick
    track += b"\x00\xff\x2f\x00"
    with open(path, "wb") as f:
        f.write(b"MThd" + struct.pack(">IHHH", 6, 0, 1, 480) + b"MTrk" + struct.pack(">I", len(track)) + track)
    return str(path)


def frame(tick):
    return math.floor(tick * FRAMES_PER_TICK + 1e-7)  # orig policy rounding


def plan_of(sfz, mid, policy=abi.POLICY_ORIG):
    s = ca.Session(policy=policy)
    s.add_job(sfz, mid)
    p = s.plan(0)
    spans = np.ctypeslib.as_array(p.spans, shape=(p.span_count,)).copy()
    offsets = np.ctypeslib.as_array(p.track_offsets, shape=(p.track_count + 1,))
    points = np.ctypeslib.as_array(p.points, shape=(int(offsets[-1]),)).copy()
    tracks = [points[offsets[t]:offsets[t + 1]] for t in range(p.track_count)]
    mods = np.ctypeslib.as_array(p.modulation, shape=(p.modulation_count,)).copy() if p.modulation_count else None
    return spans, tracks, mods, s


def track_value(track, f):
    i = max(0, np.searchsorted(track["frame"], f, side="right") - 1)
    p = track[i]
    return float(p["value"]) + (f - int(p["frame"])) * float(p["step"])


def sine_step(key, root):
    return 440 * 2 ** ((root - 69) / 12) / RATE * 2 ** ((key - root) / 12)


def test_region_selection_by_key_velocity_and_cc_window(tmp_path):
    regions = [  # (lokey, hikey, lovel, hivel, cc20 window, root, volume)
        (40, 59, 1, 63, None, 50, -1), (40, 59, 64, 127, None, 50, -2), (55, 70, 0, 127, None, 60, -3),
        (72, 72, 0, 127, None, 72, -4), (80, 90, 0, 127, (64, 127), 85, -5)]
    lines = []
    for lo, hi, lv, hv, cc, root, vol in regions:
        window = f" locc20={cc[0]} hicc20={cc[1]}" if cc else ""
        lines.append(f"<region> sample=*sine lokey={lo} hikey={hi} lovel={lv} hivel={hv} pitch_keycenter={root} "
                     f"volume={vol} amp_veltrack=0 ampeg_release=0.01{window}\n")
    sfz = text(tmp_path / "select.sfz", "".join(lines))
    rng = np.random.default_rng(3)
    notes, events, cc20, tick = [], [], 0, 0
    for k in range(60):
        if k % 7 == 3:
            cc20 = int(rng.integers(0, 128))
            events.append((tick, 0xB0, 20, cc20))
        key, vel = int(rng.integers(36, 95)), int(rng.integers(1, 128))
        events += [(tick, 0x90, key, vel), (tick + 60, 0x80, key, 0)]
        notes.append((frame(tick), key, vel, cc20))
        tick += 64
    spans, _, _, _ = plan_of(sfz, midi(tmp_path / "select.mid", events))
    gain_of = {round(10 ** ((-7.35 + vol + 3) / 20), 5): i for i, (*_, vol) in enumerate(regions)}
    got = sorted((int(s["begin"]), gain_of[round(float(s["gain"]), 5)], round(float(s["step"]), 9)) for s in spans)
    want = []
    for f, key, vel, cc in notes:
        for i, (lo, hi, lv, hv, window, root, _) in enumerate(regions):
            if lo <= key <= hi and lv <= vel <= hv and (window is None or window[0] <= cc <= window[1]):
                want.append((f, i, round(sine_step(key, root), 9)))
    assert got == sorted(want) and len(want) > 40


def test_bend_splits_a_held_note_at_its_block_and_at_the_bend(tmp_path):
    sfz = text(tmp_path / "bend.sfz", "<region> sample=*sine key=60 amp_veltrack=0 ampeg_release=0.01\n")
    bend_tick, value = 100, 12000  # 14-bit bend; +3808 / 8191 of bend_up (200 cents)
    mid = midi(tmp_path / "bend.mid", [(0, 0x90, 60, 100), (bend_tick, 0xE0, value & 0x7F, value >> 7),
                                       (400, 0x80, 60, 0)])
    spans, _, _, _ = plan_of(sfz, mid)
    bend = frame(bend_tick)
    block = bend // 1024 * 1024
    assert [(int(s["begin"]), int(s["end"])) for s in spans][:2] == [(0, block), (block, bend)]
    assert int(spans[2]["begin"]) == bend and len(spans) == 3
    base = sine_step(60, 60)
    cents = (value - 8192) / 8191 * 200
    assert spans[0]["step"] == pytest.approx(base, rel=1e-12) and spans[0]["pitch_slope"] == 0
    assert spans[1]["step"] == pytest.approx(base, rel=1e-12)  # the ramp starts from no bend
    assert spans[1]["pitch_slope"] == pytest.approx(cents / (bend - block) * math.log(2) / 1200, rel=1e-12)
    assert spans[2]["step"] == pytest.approx(base * 2 ** (cents / 1200), rel=1e-12)
    assert spans[2]["pitch_slope"] == 0
    # Source position carries over each cut: a linear run, then the geometric ramp.
    n = bend - block
    ramp = base * math.expm1(spans[1]["pitch_slope"] * n) / math.expm1(spans[1]["pitch_slope"])
    assert spans[1]["position"] == pytest.approx(spans[0]["position"] + base * block, rel=1e-12)
    assert spans[2]["position"] == pytest.approx(spans[1]["position"] + ramp, rel=1e-12)


def test_unsmoothed_controller_ramp_and_smoothed_volume_step(tmp_path):
    sfz = text(tmp_path / "cc.sfz", "<region> sample=*sine key=60 amp_veltrack=0 cutoff=500 cutoff_oncc1=1200\n")
    cc1_tick, cc7_tick = 75, 16384  # CC1 inside a block; CC7 at 752640 frames, a block boundary
    mid = midi(tmp_path / "cc.mid", [(0, 0x90, 60, 100), (cc1_tick, 0xB0, 1, 90), (cc7_tick, 0xB0, 7, 50),
                                     (16700, 0x80, 60, 0)])
    spans, tracks, mods, _ = plan_of(sfz, mid)
    # CC1 binding (no smoothing): linear from the event's block start to the event.
    cc1_track = tracks[int(mods[0]["track"])]
    f1 = frame(cc1_tick)
    b1 = f1 // 1024 * 1024
    for f in (0, b1, (b1 + f1) // 2, f1, f1 + 5000):
        want = 0.0 if f <= b1 else 90 / 127 if f >= f1 else 90 / 127 * (f - b1) / (f1 - b1)
        assert track_value(cc1_track, f) == pytest.approx(want, abs=1e-6)
    # CC7 (volume, curve x^2, smoothed over rate/100 frames): a linear approach
    # to the new level that starts in the 16-frame group of the change.
    f7 = frame(cc7_tick)
    assert f7 % 1024 == 0
    start, end, smooth = (100 / 127) ** 2, (50 / 127) ** 2, RATE // 100
    for f in range(f7 - 32, f7 + smooth + 64, 7):
        want = start + (end - start) * min(1.0, max(0.0, (f - f7 + 1) / smooth))
        assert track_value(tracks[0], f) == pytest.approx(want, abs=2e-6), f


@pytest.mark.parametrize("loop_start,want", [(1000, 1000), (5000, 2205)])
def test_loop_crossfade_length(tmp_path, loop_start, want):
    wav(tmp_path / "tone.wav", np.rint(8000 * np.sin(np.arange(RATE) * 0.05)))
    sfz = text(tmp_path / "xf.sfz", f"<region> sample=tone.wav key=60 loop_mode=loop_continuous "
                                    f"loop_start={loop_start} loop_end=30000 loop_crossfade=0.05\n")
    spans, _, _, _ = plan_of(sfz, midi(tmp_path / "xf.mid", [(0, 0x90, 60, 100), (100, 0x80, 60, 0)]))
    assert {int(s["loop_crossfade"]) for s in spans} == {want}
    assert {(int(s["loop_first"]), int(s["loop_end"])) for s in spans} == {(loop_start, 30001)}


def release_answer(release_frame, seconds):
    rate = np.float32(-9) / (np.float32(seconds) * np.float32(RATE))
    to_silence = int(np.float32(math.log(np.float32(1e-4))) / rate)
    fade_start = release_frame + to_silence
    return rate, fade_start, fade_start + RATE // 20 + 1


def test_release_closed_form_and_sustain_pedal(tmp_path):
    sfz = text(tmp_path / "rel.sfz", "<region> sample=*sine key=60 amp_veltrack=0 ampeg_release=0.5\n")
    spans, _, _, _ = plan_of(sfz, midi(tmp_path / "rel.mid", [(0, 0x90, 60, 100), (200, 0x80, 60, 0)]))
    (s,) = spans
    r = frame(200)
    rate, fade_start, end = release_answer(r, 0.5)
    assert (int(s["release_start"]), float(s["release_value"])) == (r, 1.0)
    assert np.float32(s["release_rate"]) == rate
    assert (int(s["fade_start"]), int(s["end"])) == (fade_start, end)
    # Under the sustain pedal the release waits for the pedal.
    mid = midi(tmp_path / "ped.mid", [(0, 0xB0, 64, 127), (0, 0x90, 60, 100), (200, 0x80, 60, 0),
                                      (500, 0xB0, 64, 0)])
    (s,) = plan_of(sfz, mid)[0]
    assert int(s["release_start"]) == frame(500)
    assert int(s["end"]) == release_answer(frame(500), 0.5)[2]


@pytest.mark.parametrize("opcode", ["lfo1_freq=3", "pitchlfo_depth=20", "eg1_pitch=100", "xfin_lokey=40",
                                    "seq_length=2", "lorand=0.5", "sw_last=24", "off_by=2", "group=1",
                                    "note_polyphony=2", "trigger=release", "on_locc64=64", "amp_random=3",
                                    "ampeg_dynamic=1", "cutoff_oncc135=100"])
def test_loader_rejects_features_outside_the_render_core(tmp_path, opcode):
    sfz = text(tmp_path / "x.sfz", f"<region> sample=*sine key=60 {opcode}\n")
    s = ca.Session()
    with pytest.raises(ca.ClayError, match="unsupported"):
        s.add_job(sfz, midi(tmp_path / "x.mid", [(0, 0x90, 60, 100), (100, 0x80, 60, 0)]))


def test_loader_scopes_macros_and_includes(tmp_path):
    text(tmp_path / "inc.sfz", "<region> sample=*sine key=$K volume=-6\n")
    sfz = text(tmp_path / "top.sfz", "<global> amp_veltrack=0 ampeg_release=0.01 // comment\n"
                                     "<group> volume=-2\n#define $K 62\n#include \"inc.sfz\"\n"
                                     "<region> sample=*sine lokey=63 hikey=64 pitch_keycenter=D#4\n")
    mid = midi(tmp_path / "m.mid", [(0, 0x90, 62, 100), (0, 0x90, 64, 100), (100, 0x80, 62, 0),
                                    (100, 0x80, 64, 0)])
    spans, _, _, _ = plan_of(sfz, mid)
    by_step = sorted((round(float(s["step"]), 9), round(float(s["gain"]), 6)) for s in spans)
    # Volumes add across scopes (group -2, region -6); D#4 is key 63.
    assert by_step == sorted([(round(sine_step(62, 62), 9), round(10 ** ((-7.35 - 8 + 3) / 20), 6)),
                              (round(sine_step(64, 63), 9), round(10 ** ((-7.35 - 2 + 3) / 20), 6))])
