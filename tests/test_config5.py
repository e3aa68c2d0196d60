"""Configuration 5's generator on the CPU: seeded and deterministic, the
stated region layout, the dense pedal MIDI, and the planner reading it at
96 kHz into volume groups of 8."""

import os

import numpy as np

import claysampler_amd as ca
from claysampler_amd import synth


def test_instrument_layout_and_determinism(tmp_path):
    a = synth.c5_instrument(str(tmp_path / "a"), frames=4410, keys=(60, 62))
    b = synth.c5_instrument(str(tmp_path / "b"), frames=4410, keys=(60, 62))
    lines = open(a).read().splitlines()
    assert lines[0].startswith("<group>") and "ampeg_release=4" in lines[0] and "volume=-30" in lines[0]
    regions = [l for l in lines if l.startswith("<region>")]
    assert len(regions) == 3 * 4  # keys x velocity layers
    assert [l.split("lovel=")[1] for l in regions[:4]] == ["1 hivel=31", "32 hivel=63", "64 hivel=95", "96 hivel=127"]
    for name in ("k60_v0_4410.wav", "k62_v3_4410.wav"):
        x, y = open(os.path.join(tmp_path / "a", name), "rb").read(), open(os.path.join(tmp_path / "b", name), "rb").read()
        assert x == y and len(x) == 44 + 2 * 4410
    assert open(a).read().replace("/a", "") == open(b).read().replace("/b", "")


def test_midi_is_dense_with_pedal(tmp_path):
    p1 = synth.c5_midi(str(tmp_path / "m1.mid"), 7, seconds=20.0)
    p2 = synth.c5_midi(str(tmp_path / "m2.mid"), 7, seconds=20.0)
    assert open(p1, "rb").read() == open(p2, "rb").read()
    data = open(p1, "rb").read()
    assert data.count(bytes([0xB0, 64, 127])) >= 3 and data.count(bytes([0xB0, 64, 0])) >= 3
    assert data.count(bytes([0x90])) > 0.8 * 24 * 20  # Poisson onsets at 24/s, running status never used


def test_planner_reads_configuration_5_at_96k(tmp_path):
    sfz, mids = synth.config5(str(tmp_path), jobs=2, seconds=4.0, frames=22050, keys=(60, 71))
    s = ca.Session(rate=96000, measure=True)
    for j, m in enumerate(mids):
        s.add_job(sfz, m, group=1)
    plans = s.plans()
    assert len(plans) == 2 and all(p.frames > 4 * 96000 for p in plans)  # 4 s + pedal and 4 s releases
    v = np.ctypeslib.as_array(plans[0].spans, shape=(plans[0].span_count,))
    voices = (v["end"] - v["begin"]).sum() / plans[0].frames
    assert voices > 5  # overlapping voices (0.5 s regions end them early here)
    assert s.bank()[1] == 12 * 4 * 22050
