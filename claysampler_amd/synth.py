"""Seeded synthetic workloads (SFZ + PCM16 WAV + SMF) for the five benchmark
configurations. Every random draw comes from a SplitMix64 stream, so the files
are identical on any machine and numpy version.

Choices the configurations leave open are fixed here:
  C1  sine440.wav = rint(16384 sin(2 pi 440 n / 44100)), 44100 frames; 32
      back-to-back notes, key U{48..84}, velocity U{40..127}, duration U[0.1, 1.0] s.
  C2  keys 21..108 x 2 velocity layers (1-63, 64-127) = 176 regions of 2 s;
      8 harmonics with 1/n amplitude, random phase, exp decay (tau 1.2 s / n),
      harmonics >= Nyquist dropped, + -40 dB noise; loop_continuous over the
      last 0.5 s; lpf_2p cutoff 200 Hz + cutoff_oncc1=6386 (to 8 kHz),
      resonance 6 dB; ampeg attack 5 ms, decay 0.4 s, sustain 60 %, release 0.3 s.
      MIDI 60 s: Poisson onsets 8/s, key U{36..96}, velocity U{20..127},
      duration U[0.1, 1.5] s, at most 16 held notes; pitch bend and CC1 random
      walks at 20 Hz (bend step U{-1024..1024} clipped to +-8191, CC1 step U{-6..6}).
  C3  C1 instrument with volume=-18 (up to 64 overlapping sines at amplitude
      0.5 would otherwise exceed full scale and end in the `peak` status); job
      length lognormal(ln 30 s, 1) clipped to [5, 300] s; polyphony
      P ~ U{4..64} from overlapping notes of U[0.5, 2.0] s at rate P/1.25 s.
  C5  keys 21..108 x 4 velocity layers (1-31, 32-63, 64-95, 96-127) = 352
      regions of 2,840,000 frames at 44.1 kHz (a 2.0 GB int16 bank). Each region
      repeats one second of 8 harmonics (1/n, random phase, harmonics >= Nyquist
      dropped) under a slow exp decay (tau 20 s), + -40 dB noise over the whole
      region; no loop; volume -30 dB (hundreds of voices sum without clipping);
      ampeg attack 5 ms, decay 2 s, sustain 40 %, release 4 s.
      MIDI 240 s rendered at 96 kHz: Poisson onsets 24/s, key U{21..108},
      velocity U{1..127}, duration U[0.2, 2.0] s, at most 256 sounding notes
      (held, or sustained by the pedal); CC64 down for U[2, 6] s then up for
      U[0.5, 1.5] s in turn. Jobs form volume groups of 8; 64 taps.
"""

import math
import os
import struct

import numpy as np

RATE = 44100
PPQ = 480
TEMPO = 500000  # 120 BPM: 960 ticks per second


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self):
        """[0, 1) with 53 random bits."""
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi] (inclusive)."""
        return lo + self.next() % (hi - lo + 1)

    def exponential(self, rate):
        return -math.log(1.0 - self.uniform()) / rate

    def normal(self):
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2 * math.pi * u2)


def write_wav(path, pcm, rate=RATE, channels=1):
    data = np.asarray(pcm, dtype="<i2").tobytes()
    header = struct.pack("<4sI4s4sIHHIIHH4sI", b"RIFF", 36 + len(data), b"WAVE", b"fmt ", 16, 1, channels, rate,
                         rate * channels * 2, channels * 2, 16, b"data", len(data))
    with open(path, "wb") as f:
        f.write(header + data)


def _varlen(v):
    out = [v & 0x7F]
    v >>= 7
    while v:
        out.append(0x80 | (v & 0x7F))
        v >>= 7
    return bytes(reversed(out))


def write_midi(path, events, ppq=PPQ, tempo=TEMPO):
    """events: (tick, order, status, a, b) tuples; b < 0 for 1-data-byte messages."""
    track = bytearray(b"\x00\xff\x51\x03" + tempo.to_bytes(3, "big"))
    last = 0
    for tick, _, status, a, b in sorted(events, key=lambda e: (e[0], e[1])):
        track += _varlen(tick - last) + bytes([status, a] + ([] if b < 0 else [b]))
        last = tick
    track += b"\x00\xff\x2f\x00"
    with open(path, "wb") as f:
        f.write(b"MThd" + struct.pack(">IHHH", 6, 0, 1, ppq) + b"MTrk" + struct.pack(">I", len(track)) + track)


def seconds_to_ticks(s):
    return int(round(s * PPQ * 1e6 / TEMPO))


# ----------------------------------------------------------------- config 1

C1_SFZ = ("<region> sample=sine440.wav lokey=0 hikey=127 pitch_keycenter=69 loop_mode=no_loop "
          "ampeg_attack=0.005 ampeg_decay=0.1 ampeg_sustain=70 ampeg_release=0.2\n")


def c1_instrument(directory, volume_db=None):
    os.makedirs(directory, exist_ok=True)
    n = np.arange(RATE)
    write_wav(os.path.join(directory, "sine440.wav"), np.rint(16384 * np.sin(2 * np.pi * 440 * n / RATE)))
    sfz = os.path.join(directory, "c1.sfz" if volume_db is None else "c3.sfz")
    with open(sfz, "w") as f:
        f.write(C1_SFZ if volume_db is None else C1_SFZ.replace("<region> ", f"<region> volume={volume_db:g} "))
    return sfz


def c1_midi(path, seed=1, notes=32):
    rng = SplitMix64(seed)
    events, tick = [], 0
    for _ in range(notes):
        key, vel = rng.integer(48, 84), rng.integer(40, 127)
        dur = seconds_to_ticks(0.1 + 0.9 * rng.uniform())
        events.append((tick, 1, 0x90, key, vel))
        events.append((tick + dur, 0, 0x80, key, 0))
        tick += dur
    write_midi(path, events)
    return path


def config1(directory, seed=1):
    """(sfz, [midi]) for configuration 1."""
    sfz = c1_instrument(directory)
    return sfz, [c1_midi(os.path.join(directory, f"c1_{seed}.mid"), seed)]


# ----------------------------------------------------------------- config 2

def c2_instrument(directory, seed=2, seconds=2.0):
    os.makedirs(directory, exist_ok=True)
    frames = int(seconds * RATE)
    loop_start = frames - int(0.5 * RATE)
    t = np.arange(frames) / RATE
    lines = ["<group> fil_type=lpf_2p cutoff=200 cutoff_oncc1=6386 resonance=6 loop_mode=loop_continuous "
             f"loop_start={loop_start} loop_end={frames - 1} ampeg_attack=0.005 ampeg_decay=0.4 "
             "ampeg_sustain=60 ampeg_release=0.3\n"]
    for key in range(21, 109):
        f0 = 440.0 * 2 ** ((key - 69) / 12)
        for layer, (lo, hi, level) in enumerate(((1, 63, 0.35), (64, 127, 0.6))):
            name = f"k{key}_v{layer}.wav"
            path = os.path.join(directory, name)
            if not os.path.exists(path):
                rng = SplitMix64(seed * 100000 + key * 2 + layer)
                x = np.zeros(frames)
                for h in range(1, 9):
                    if f0 * h >= RATE / 2:
                        break
                    phase = 2 * np.pi * rng.uniform()
                    x += np.sin(2 * np.pi * f0 * h * t + phase) * np.exp(-t * h / 1.2) / h
                x *= level / max(1e-9, np.max(np.abs(x)))
                # -40 dB noise from a PCG64 stream seeded by this region's SplitMix64.
                noise = np.random.Generator(np.random.PCG64(rng.next())).uniform(-1.0, 1.0, frames)
                x += 0.01 * level * noise
                write_wav(path, np.clip(np.rint(x * 32767), -32768, 32767))
            lines.append(f"<region> sample={name} lokey={key} hikey={key} pitch_keycenter={key} lovel={lo} hivel={hi}\n")
    sfz = os.path.join(directory, "c2.sfz")
    with open(sfz, "w") as f:
        f.writelines(lines)
    return sfz


def c2_midi(path, seed, seconds=60.0, rate_hz=8.0, max_held=16, control_hz=20.0):
    rng = SplitMix64(seed)
    events = []
    held = []  # end ticks of held notes
    t = rng.exponential(rate_hz)
    while t < seconds:
        start = seconds_to_ticks(t)
        held = [e for e in held if e > start]
        if len(held) < max_held:
            key, vel = rng.integer(36, 96), rng.integer(20, 127)
            end = seconds_to_ticks(min(seconds, t + 0.1 + 1.4 * rng.uniform()))
            if end > start:
                events.append((start, 2, 0x90, key, vel))
                events.append((end, 1, 0x80, key, 0))
                held.append(end)
        t += rng.exponential(rate_hz)
    bend, cc1 = 0, 64
    step = 1.0 / control_hz
    k = 0
    while k * step < seconds:
        tick = seconds_to_ticks(k * step)
        bend = max(-8191, min(8191, bend + rng.integer(-1024, 1024)))
        cc1 = max(0, min(127, cc1 + rng.integer(-6, 6)))
        value = bend + 8192
        events.append((tick, 0, 0xE0, value & 0x7F, value >> 7))
        events.append((tick, 0, 0xB0, 1, cc1))
        k += 1
    write_midi(path, events)
    return path


def config2(directory, jobs=64, seconds=60.0):
    sfz = c2_instrument(directory)
    mids = []
    for j in range(jobs):
        path = os.path.join(directory, f"c2_{j}_{int(seconds)}.mid")
        if not os.path.exists(path):
            c2_midi(path, 1000 + j, seconds)
        mids.append(path)
    return sfz, mids


# ----------------------------------------------------------------- config 3

def c3_midi(path, seed, seconds, polyphony):
    rng = SplitMix64(seed)
    events = []
    rate = polyphony / 1.25
    t = rng.exponential(rate)
    while t < seconds:
        key, vel = rng.integer(48, 84), rng.integer(40, 127)
        start = seconds_to_ticks(t)
        end = seconds_to_ticks(min(seconds, t + 0.5 + 1.5 * rng.uniform()))
        if end > start:
            events.append((start, 2, 0x90, key, vel))
            events.append((end, 1, 0x80, key, 0))
        t += rng.exponential(rate)
    write_midi(path, events)
    return path


def config3(directory, jobs=512, max_seconds=300.0):
    sfz = c1_instrument(directory, volume_db=-18)
    mids = []
    rng = SplitMix64(3)
    for j in range(jobs):
        seconds = min(max_seconds, max(5.0, math.exp(math.log(30.0) + rng.normal())))
        polyphony = rng.integer(4, 64)
        path = os.path.join(directory, f"c3_{j}.mid")
        if not os.path.exists(path):
            c3_midi(path, 3000 + j, seconds, polyphony)
        mids.append(path)
    return sfz, mids


# ----------------------------------------------------------------- config 5

C5_LAYERS = ((1, 31, 0.25), (32, 63, 0.4), (64, 95, 0.6), (96, 127, 0.85))


def c5_instrument(directory, seed=5, frames=2_840_000, keys=(21, 108)):
    """The dense stress instrument; `frames` and `keys` shrink it for tests.
    One second of each region's tone is synthesised and repeated (phase jumps
    at the seams are harmless for a stress bank), so the 2 GB bank is written
    in seconds."""
    os.makedirs(directory, exist_ok=True)
    n = np.arange(RATE) / RATE
    decay = np.exp(-np.arange(frames) / (20.0 * RATE))
    lines = ["<group> loop_mode=no_loop volume=-30 ampeg_attack=0.005 ampeg_decay=2 ampeg_sustain=40 "
             "ampeg_release=4\n"]
    for key in range(keys[0], keys[1] + 1):
        f0 = 440.0 * 2 ** ((key - 69) / 12)
        for layer, (lo, hi, level) in enumerate(C5_LAYERS):
            name = f"k{key}_v{layer}_{frames}.wav"
            path = os.path.join(directory, name)
            if not os.path.exists(path):
                rng = SplitMix64(seed * 1000003 + key * 4 + layer)
                second = np.zeros(RATE)
                for h in range(1, 9):
                    if f0 * h >= RATE / 2:
                        break
                    second += np.sin(2 * np.pi * f0 * h * n + 2 * np.pi * rng.uniform()) / h
                second *= level / max(1e-9, np.max(np.abs(second)))
                x = np.resize(second, frames) * decay
                x += 0.01 * level * np.random.Generator(np.random.PCG64(rng.next())).uniform(-1.0, 1.0, frames)
                write_wav(path, np.clip(np.rint(x * 32767), -32768, 32767))
            lines.append(f"<region> sample={name} lokey={key} hikey={key} pitch_keycenter={key} lovel={lo} hivel={hi}\n")
    sfz = os.path.join(directory, f"c5_{frames}_{keys[0]}_{keys[1]}.sfz")
    with open(sfz, "w") as f:
        f.writelines(lines)
    return sfz


def c5_midi(path, seed, seconds=240.0, rate_hz=24.0, max_sounding=256, keys=(21, 108)):
    rng = SplitMix64(seed)
    events = []
    # Pedal: down for U[2, 6] s, then up for U[0.5, 1.5] s, in turn.
    pedal, t, down = [], 0.0, True
    while t < seconds:
        length = 2.0 + 4.0 * rng.uniform() if down else 0.5 + rng.uniform()
        pedal.append((t, min(seconds, t + length), down))
        events.append((seconds_to_ticks(t), 0, 0xB0, 64, 127 if down else 0))
        t += length
        down = not down

    def released(end):
        """When a note released at `end` stops sustaining (the next pedal-up)."""
        for a, b, d in pedal:
            if d and a <= end < b:
                return b
        return end

    sounding = []  # frames (seconds) when each note stops sounding
    t = rng.exponential(rate_hz)
    while t < seconds:
        sounding = [e for e in sounding if e > t]
        key, vel = rng.integer(keys[0], keys[1]), rng.integer(1, 127)
        end = min(seconds, t + 0.2 + 1.8 * rng.uniform())
        if len(sounding) < max_sounding and seconds_to_ticks(end) > seconds_to_ticks(t):
            events.append((seconds_to_ticks(t), 2, 0x90, key, vel))
            events.append((seconds_to_ticks(end), 1, 0x80, key, 0))
            sounding.append(released(end) + 4.0)
        t += rng.exponential(rate_hz)
    write_midi(path, events)
    return path


def config5(directory, jobs=32, seconds=240.0, frames=2_840_000, keys=(21, 108)):
    """(sfz, [midi]) for configuration 5; jobs j belong to volume group j // 8 + 1."""
    sfz = c5_instrument(directory, frames=frames, keys=keys)
    mids = []
    for j in range(jobs):
        path = os.path.join(directory, f"c5_{j}_{int(seconds)}_{keys[0]}_{keys[1]}.mid")
        if not os.path.exists(path):
            c5_midi(path, 5000 + j, seconds, keys=keys)
        mids.append(path)
    return sfz, mids
