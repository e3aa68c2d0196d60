"""Pins of the oracle's remaining render functions (CPU only).

Each test drives the FP64 oracle through `clay_oracle_mix` with a hand-built
plan whose output has an independent description: a closed form derived from
what the function means (a geometric sum, a half-cosine blend, an equal-power
pan law with its special cases), a brute-force evaluation written differently
from the oracle (linear scans instead of binary search, a cumulative sum
instead of expm1), or the published steps of an algorithm. Reference lines:
src/cuda/render.cu:63-100 (loops, crossfades, generators, bent spans),
:107-126 (control tracks), :128-187 (pan law, width and position),
:222-274 (variable pitch), :321-328 and :376-407 (moving-filter smoothing).
"""

import math

import numpy as np
import pytest
import scipy.signal

import oracle
from claysampler_amd import abi

from planbuild import Plan, bank_pointer, span

RATE = 44100
Q = math.pi / 4


def mix(plan, bank, taps=32):
    _, ptr, n = bank_pointer(bank)
    return oracle.mix(oracle.config(taps=taps), ptr, n, plan.view)


def fir(bank, pos, taps, table, channels=1):
    """Windowed-sinc value at `pos` by direct summation over the taps (zero
    outside the bank), the exact sample at integer positions."""
    frames = len(bank) // channels
    centre = math.floor(pos)
    if pos == centre:
        if 0 <= centre < frames:
            return bank[centre * channels] / 32768.0
        return 0.0
    phase = int(np.rint((pos - centre) * 2048))
    total = 0.0
    for t in range(taps):
        i = centre - (taps // 2 - 1) + t
        if 0 <= i < frames:
            total += float(table[t, phase]) * bank[i * channels] / 32768.0
    return total


def f32(x):
    """The float32 value a plan field holds."""
    return float(np.float32(x))


def centre_gain(gain=1.0):
    return float(np.float32(gain * math.cos(Q)))


# ------------------------------------------------------------------ loop wrap (render.cu:85-87)

def test_loop_wrap_on_integer_steps_reads_the_wrapped_sample():
    """Step 1 from position 0 over a loop [500, 1500): output frame n is sample
    n up to the loop end, then cycles through the loop body."""
    rng = np.random.default_rng(11)
    bank = rng.integers(-30000, 30000, 2000).astype(np.int16)
    s = span(0, 5000, source_end=2000, loop_first=500, loop_end=1500)
    out = mix(Plan([s], 5000), bank)
    idx = np.arange(5000)
    wrapped = np.where(idx < 1500, idx, 500 + (idx - 500) % 1000)
    want = bank[wrapped] / 32768.0 * centre_gain()
    assert np.array_equal(out[:, 0], want) and np.array_equal(out[:, 1], want)


def test_loop_wrap_fractional_steps_against_brute_force():
    """Non-integer step: the wrapped position is loop_first + (pos - loop_first)
    mod period, and the window around it is read linearly (not wrapped)."""
    rng = np.random.default_rng(12)
    bank = rng.integers(-30000, 30000, 1200).astype(np.int16)
    step, taps = 1.37, 16
    s = span(0, 2000, source_end=1200, step=step, position=3.25, loop_first=300, loop_end=1000)
    out = mix(Plan([s], 2000), bank, taps)
    table = oracle.sinc_table(taps).astype(np.float64)
    for n in range(0, 2000, 7):
        pos = 3.25 + n * step
        # Subtract whole periods one at a time (an independent form of the fmod).
        while pos >= 1000:
            pos -= 700
        want = fir(bank, pos, taps, table) * centre_gain()
        assert abs(out[n, 0] - want) < 1e-12, n


# ------------------------------------------------------------------ loop crossfade (render.cu:88-98)

def test_loop_crossfade_is_a_raised_cosine_blend():
    """Inside the last `crossfade` frames before the loop end, the output moves
    from the sample at pos to the sample one period earlier along half a cosine:
    weight 0 at the fade start, 1/2 at its middle, approaching 1 at the loop end."""
    rng = np.random.default_rng(13)
    bank = rng.integers(-30000, 30000, 2000).astype(np.int16)
    xf = 100
    s = span(0, 4000, source_end=2000, loop_first=500, loop_end=1500, loop_crossfade=xf)
    out = mix(Plan([s], 4000), bank)[:, 0] / centre_gain()
    idx = np.arange(4000)
    pos = np.where(idx < 1500, idx, 500 + (idx - 500) % 1000)
    x = pos - (1500 - xf)
    blend = np.where(x >= 0, np.sin(np.pi * np.clip(x, 0, None) / (2 * xf)) ** 2, 0.0)  # = (1 - cos(pi x / xf)) / 2
    here, entry = bank[pos] / 32768.0, bank[np.clip(pos - 1000, 0, None)] / 32768.0
    want = (1 - blend) * here + blend * entry
    assert np.max(np.abs(out - want)) < 1e-12
    middle = 1500 - xf // 2
    assert abs(out[middle] - (bank[middle] + bank[middle - 1000]) / 65536.0) < 1e-12
    assert out[1500 - xf] == bank[1500 - xf] / 32768.0


# ------------------------------------------------------------------ generators (render.cu:72-84)

def test_sine_generator_is_minus_sine_of_the_phase():
    s = span(0, 3000, position=0.3, step=0.0137, generator=abi_generator("sine"), source_end=0)
    out = mix(Plan([s], 3000), np.zeros(4, np.int16))[:, 0] / centre_gain()
    pos = 0.3 + np.arange(3000) * 0.0137
    assert np.max(np.abs(out - -np.sin(2 * np.pi * (pos % 1.0)))) < 1e-12
    # Quarter phases: -1 at 0.25, 0 at 0.5, +1 at 0.75.
    s = span(0, 4, position=0.25, step=0.25, generator=abi_generator("sine"), source_end=0)
    out = mix(Plan([s], 4), np.zeros(4, np.int16))[:, 0] / centre_gain()
    assert np.allclose(out, [-1.0, 0.0, 1.0, 0.0], atol=1e-15)


def lowbias32(x):
    """Chris Wellons' lowbias32 integer hash (hash-prospector), step by step."""
    x = np.uint64(x) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    return int(x)


def abi_generator(name):
    return {"none": 0, "sine": 1, "silence": 2, "noise": 3}[name]


@pytest.mark.parametrize("seed", [0, 1, 12345, 0xFFFFFF00])
def test_noise_generator_follows_lowbias32(seed):
    """Frame f emits lowbias32(seed + f - noise_start) >> 8 mapped to [-1, 1)."""
    frames = 600
    s = span(100, 100 + frames, generator=abi_generator("noise"), noise_seed=seed, noise_start=40, source_end=0)
    out = mix(Plan([s], 100 + frames), np.zeros(4, np.int16))[100:, 0] / centre_gain()
    want = np.array([(lowbias32(seed + f - 40) >> 8) * 2.0 / 16777216.0 - 1.0 for f in range(100, 100 + frames)])
    assert np.max(np.abs(out - want)) < 1e-14
    assert np.all(out >= -1.0) and np.all(out < 1.0)


def test_lowbias32_is_a_bijection_with_the_published_fixed_point():
    assert lowbias32(0) == 0
    values = {lowbias32(x) for x in range(1 << 14)}
    assert len(values) == 1 << 14


# ------------------------------------------------------------------ bent spans (render.cu:63-70)

@pytest.mark.parametrize("slope", [3e-4, -2e-4])
def test_bent_travel_is_a_geometric_sum_of_increments(slope):
    """A bent span's first frame advances by exactly `step` and each later
    increment is e^slope times the previous one; the sine generator exposes the
    position's fractional part."""
    step, n = 0.0113, 4000
    s = span(0, n, position=0.1, step=step, pitch_slope=slope, generator=abi_generator("sine"), source_end=0)
    out = mix(Plan([s], n), np.zeros(4, np.int16))[:, 0] / centre_gain()
    increments = step * np.exp(slope * np.arange(n, dtype=np.float64))
    pos = 0.1 + np.concatenate([[0.0], np.cumsum(increments)[:-1]])
    assert np.max(np.abs(out - -np.sin(2 * np.pi * (pos % 1.0)))) < 1e-9


# ------------------------------------------------------------------ control tracks (render.cu:107-126, :189-193)

def track_value(points, f):
    """Linear scan: the last point at or before f (the first point when none)."""
    chosen = points[0]
    for p in points:
        if p[0] <= f:
            chosen = p
    return chosen[1] + chosen[2] * (f - chosen[0])


def random_track(rng, frames, lo, hi, count):
    frames_at = np.sort(rng.choice(np.arange(1, frames), count - 1, replace=False)).tolist()
    points = []
    for f in [int(rng.integers(-20, 20))] + frames_at:
        points.append((int(f), float(np.float32(rng.uniform(lo, hi))), float(np.float32(rng.uniform(-1e-3, 1e-3)))))
    return points


def test_job_gain_pan_and_modulation_tracks_brute_force():
    """DC source on integer steps, random piecewise-linear CC7, CC11, CC10 and two
    modulation tracks (amplitude with a shape track, volume in dB)."""
    rng = np.random.default_rng(21)
    frames = 2500
    cc7, cc11 = random_track(rng, frames, 0.2, 1.0, 9), random_track(rng, frames, 0.3, 1.0, 6)
    cc10 = random_track(rng, frames, -0.6, 0.6, 7)
    amp, shape, vol = random_track(rng, frames, 0.5, 1.5, 5), random_track(rng, frames, 0.2, 1.0, 4), \
        random_track(rng, frames, -12, 6, 5)
    mods = [(3, 0, 0.8, 4), (5, 1, 0.5, -1)]  # amplitude: depth 0.8 x track 3 x shape 4; volume: 0.5 x track 5 dB
    s = span(0, frames, source_end=4000, gain=0.9, pan=0.2, first_modulation=0, modulation_count=2)
    plan = Plan([s], frames, tracks=[cc7, cc11, cc10, amp, shape, vol], mods=mods)
    bank = np.full(4000, 8000, np.int16)
    out = mix(plan, bank)
    for f in range(0, frames, 3):
        level = f32(0.8) * track_value(amp, f) * track_value(shape, f) * 10 ** (0.5 * track_value(vol, f) / 20)
        offset = track_value(cc10, f)
        p = min(1.0, max(-1.0, f32(0.2) + offset))
        gl, gr = (f32(0.9) * math.cos((1 + p) * Q), f32(0.9) * math.cos((1 - p) * Q)) if offset != 0 else \
            (float(s.left), float(s.right))
        g = track_value(cc7, f) * track_value(cc11, f)
        want = 8000 / 32768.0 * level * g
        assert out[f, 0] == pytest.approx(want * gl, rel=1e-9, abs=1e-15), f
        assert out[f, 1] == pytest.approx(want * gr, rel=1e-9, abs=1e-15), f


# ------------------------------------------------------------------ pan law, width, position (render.cu:128-187)

@pytest.mark.parametrize("pan,offset", [(0.0, 0.5), (-0.3, 0.1), (0.7, 0.6), (-0.9, -0.4)])
def test_constant_pan_offset_follows_the_equal_power_law(pan, offset):
    frames = 64
    s = span(0, frames, source_end=100, gain=0.75, pan=pan)
    plan = Plan([s], frames, tracks=[[(0, 1.0, 0.0)], [(0, 1.0, 0.0)], [(0, offset, 0.0)]])
    out = mix(plan, np.full(100, 16384, np.int16))
    p = min(1.0, max(-1.0, f32(pan) + f32(offset)))
    # Equal power: L^2 + R^2 = gain^2 (at unit source), L = R at the centre,
    # one channel silent at either end.
    l, r = out[0] / 0.5
    assert l * l + r * r == pytest.approx(0.75 ** 2, rel=1e-12)
    assert l == pytest.approx(0.75 * math.cos((1 + p) * Q), abs=1e-14)
    assert r == pytest.approx(0.75 * math.cos((1 - p) * Q), abs=1e-14)
    if p == 1.0:
        assert abs(l) < 1e-12


def stereo_case(width, position, width_mod=0.0, position_mod=0.0):
    frames = 32
    mods, count = [], 0
    if width_mod:
        mods.append((-1, abi_target("width"), width_mod, -1))
    if position_mod:
        mods.append((-1, abi_target("position"), position_mod, -1))
    s = span(0, frames, source_end=100, channels=2, gain=1.0, stereo_matrix=1, width=width, stereo_position=position,
             first_modulation=0, modulation_count=len(mods))
    s.width_direct, s.width_cross = math.cos((1 - width) * Q), math.cos((1 + width) * Q)
    s.position_left, s.position_right = math.cos((1 + position) * Q), math.cos((1 - position) * Q)
    bank = np.tile(np.array([12000, -4000], np.int16), 100)
    return mix(Plan([s], frames, mods=mods), bank)[0] / math.cos(Q)  # unit pan gains


def abi_target(name):
    return {"amplitude": 0, "volume": 1, "pan": 2, "width": 3, "position": 4, "pitch": 5, "cutoff": 15}[name]


def test_stereo_width_special_cases():
    """Width 1 keeps the channels, 0 folds them to their mid (sum / sqrt 2 on
    both), -1 swaps them; position 0 leaves both at cos(pi/4)."""
    L, R = 12000 / 32768.0, -4000 / 32768.0
    c = math.cos(Q)
    assert np.allclose(stereo_case(1.0, 0.0), [L * c, R * c], atol=1e-7)
    assert np.allclose(stereo_case(0.0, 0.0), [(L + R) * c * c, (L + R) * c * c], atol=1e-7)
    assert np.allclose(stereo_case(-1.0, 0.0), [R * c, L * c], atol=1e-7)
    # Modulated width and position, evaluated from the same law at run time.
    got = stereo_case(0.3, -0.2, width_mod=0.25, position_mod=0.4)
    w, p = 0.55, 0.2
    direct, cross = math.cos((1 - w) * Q), math.cos((1 + w) * Q)
    want = [(L * direct + R * cross) * math.cos((1 + p) * Q), (R * direct + L * cross) * math.cos((1 - p) * Q)]
    assert np.allclose(got, want, atol=1e-12)
    # Clamped at the ends: width 0.9 + 0.5 behaves like width 1.
    assert np.allclose(stereo_case(0.9, 0.0, width_mod=0.5), [L * c, R * c], atol=1e-12)


# ------------------------------------------------------------------ moving filter (render.cu:321-328, :376-407)

def df2t(x, coefficients):
    """Direct form II transposed with per-sample coefficients (b0, b1, b2, a1, a2)."""
    y = np.zeros_like(x)
    z1 = z2 = 0.0
    for n in range(len(x)):
        b0, b1, b2, a1, a2 = coefficients[n]
        y[n] = b0 * x[n] + z1
        z1 = b1 * x[n] - a1 * y[n] + z2
        z2 = b2 * x[n] - a2 * y[n]
    return y


def lowpass(f, res, fs=RATE):
    """Audio EQ Cookbook low-pass, normalised by a0."""
    w = 2 * math.pi * f / fs
    alpha = math.sin(w) / (2 * 10 ** (res / 20))
    a0 = 1 + alpha
    return np.array([(1 - math.cos(w)) / 2, 1 - math.cos(w), (1 - math.cos(w)) / 2, -2 * math.cos(w),
                     1 - alpha]) / a0


def test_per_sample_df2t_equals_lfilter_when_static():
    rng = np.random.default_rng(31)
    x = rng.standard_normal(3000)
    c = lowpass(900.0, 3.0)
    want = scipy.signal.lfilter(c[:3], [1, c[3], c[4]], x)
    assert np.max(np.abs(df2t(x, np.tile(c, (3000, 1))) - want)) < 1e-12


def test_cutoff_step_smooths_coefficients_exponentially():
    """A cutoff modulation that steps from 0 to 2400 cents at frame k: the
    coefficients stay at the first target, then approach the second as
    t + (1 - a)^(n - k + 1) (c0 - t) with a = 1 - exp(-1000 / rate); the wet
    output is that time-varying DF2T of the dry voice."""
    rng = np.random.default_rng(32)
    bank = rng.integers(-20000, 20000, 5000).astype(np.int16)
    n, k, cutoff, res = 3000, 1200, 300.0, 4.0
    track = [(0, 0.0, 0.0), (k, 2400.0, 0.0)]
    dry = span(0, n, source_end=5000, step=0.81)
    wet = span(0, n, source_end=5000, step=0.81, filter_track=0, first_modulation=0, modulation_count=1)
    wet.filters[0].type, wet.filters[0].cutoff, wet.filters[0].resonance = 3, cutoff, res
    tracks = [[(0, 1.0, 0.0)], [(0, 1.0, 0.0)], [(0, 0.0, 0.0)], track]
    x = mix(Plan([dry], n, tracks=tracks), bank)[:, 0]
    y = mix(Plan([wet], n, tracks=tracks, mods=[(3, abi_target("cutoff"), 1.0, -1)], filter_tracks=1), bank)[:, 0]
    a = 1 - math.exp(-1000.0 / RATE)
    first, second = lowpass(cutoff, res), lowpass(cutoff * 4, res)
    frames = np.arange(n)
    decay = np.where(frames >= k, (1 - a) ** (frames - k + 1).astype(np.float64), 1.0)[:, None]
    coefficients = second + decay * (first - second)
    want = df2t(x, coefficients)
    assert np.max(np.abs(y - want)) < 1e-10
    assert np.allclose(coefficients[k - 1], first) and np.allclose(coefficients[-1], second, atol=1e-12)


# ------------------------------------------------------------------ variable pitch (render.cu:222-274)

@pytest.mark.parametrize("cents,slope", [(300.0, 0.0), (-700.0, 1e-4)])
def test_variable_pitch_position_from_constant_cents(cents, slope):
    """A variable-pitch chain under a constant pitch modulation of c cents:
    frame n sits at p0 + sum_{j<n} step e^(slope j) 2^(c/1200)."""
    n, step = 3000, 0.0071
    s = span(0, n, position=0.2, step=step, pitch_slope=slope, generator=abi_generator("sine"), source_end=0,
             filter_track=0, variable_pitch=1, first_modulation=0, modulation_count=1)
    out = mix(Plan([s], n, mods=[(-1, abi_target("pitch"), cents, -1)], filter_tracks=1),
              np.zeros(4, np.int16))[:, 0] / centre_gain()
    increments = step * np.exp(slope * np.arange(n)) * 2 ** (cents / 1200)
    pos = 0.2 + np.concatenate([[0.0], np.cumsum(increments)[:-1]])
    assert np.max(np.abs(out - -np.sin(2 * np.pi * (pos % 1.0)))) < 1e-8


def test_variable_pitch_continues_across_chain_spans():
    """Two spans of one chain: the second continues from where the first left
    off (its own `position` field is ignored), at its own step."""
    n1, n2 = 1000, 1500
    a = span(0, n1, position=0.0, step=0.01, generator=abi_generator("sine"), source_end=0, filter_track=0,
             variable_pitch=1)
    b = span(n1, n1 + n2, position=123.0, step=0.02, generator=abi_generator("sine"), source_end=0, filter_track=0,
             variable_pitch=1)
    out = mix(Plan([a, b], n1 + n2, filter_tracks=1), np.zeros(4, np.int16))[:, 0] / centre_gain()
    pos = np.concatenate([np.arange(n1) * 0.01, n1 * 0.01 + np.arange(n2) * 0.02])
    assert np.max(np.abs(out - -np.sin(2 * np.pi * (pos % 1.0)))) < 1e-9
