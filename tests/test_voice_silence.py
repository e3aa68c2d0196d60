"""Voice-silence tail policy of the oracle (the judge of the learned tail
predictor), pinned by closed forms: a DC source under an exponential release
crosses the level at a frame known analytically, so the first silent chunk
and the retirement frame (16 chunks later) follow; a source that is silent
from the start may only retire 16 chunks after its release begins."""

import math

import numpy as np

import oracle
from claysampler_amd import abi

from planbuild import Plan, bank_pointer, span

LEVEL_DB = -90.0
LEVEL = np.float32(10 ** (LEVEL_DB / 20))


def dc_chain(value, release, rate, frames=40000):
    s = span(0, frames, filter_track=0, release_start=release, release_value=1.0, release_rate=rate,
             fade_start=1 << 30)
    return Plan([s], frames, filter_tracks=1)


def tails(plan, bank, db=LEVEL_DB):
    bank, ptr, n = bank_pointer(bank)
    cfg = oracle.config(taps=16, voice_silence_db=db)
    return oracle.voice_tails(cfg, ptr, n, plan.view), oracle.mix(cfg, ptr, n, plan.view)


def test_dc_release_retires_sixteen_chunks_after_the_first_silent_chunk():
    release, rate = 2970, -0.001
    plan = dc_chain(16384, release, rate)
    f = np.arange(release, 40000)
    x = (0.5 * np.exp((f - release + 1) * float(np.float32(rate)))).astype(np.float32)
    first_silent = int(f[np.argmax(x < LEVEL)])
    assert 8 <= first_silent % 64 <= 56  # the crossing is well inside a chunk
    want = (first_silent // 64 + 1 + 16) * 64
    got, mix = tails(plan, np.full(1 << 20, 16384))
    assert got.tolist() == [want]
    assert np.all(mix[want:] == 0) and np.any(mix[want - 64:want] != 0)
    # The same plan without the policy keeps the tail.
    off, mix_off = tails(plan, np.full(1 << 20, 16384), db=0.0)
    assert off.tolist() == [40000] and np.array_equal(mix[:want], mix_off[:want]) and np.any(mix_off[want:] != 0)


def test_silent_voice_waits_for_its_release():
    release = 5000
    plan = dc_chain(0, release, -0.001)
    got, _ = tails(plan, np.zeros(1 << 20, np.int16))
    assert got.tolist() == [(math.ceil(release / 64) + 16) * 64]


def test_voice_that_never_falls_silent_keeps_its_span():
    plan = dc_chain(16384, 1000, -1e-6)
    got, _ = tails(plan, np.full(1 << 20, 16384))
    assert got.tolist() == [40000]


def test_chunk_split_across_spans_is_judged_once():
    """Bend-split spans of one voice share chunks: the same retirement frame as
    one span."""
    release, rate = 3000, -0.001
    one = dc_chain(16384, release, rate)
    a = span(0, 12345, filter_track=0, release_start=release, release_value=1.0, release_rate=rate,
             fade_start=1 << 30)
    b = span(12345, 40000, position=12345.0, filter_track=0, release_start=release, release_value=1.0,
             release_rate=rate, fade_start=1 << 30)
    two = Plan([a, b], 40000, filter_tracks=1)
    bank = np.full(1 << 20, 16384)
    assert tails(one, bank)[0].tolist() == tails(two, bank)[0].tolist()
