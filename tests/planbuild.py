"""Hand-built plans (clay_plan_view) for unit tests of the oracle and the HIP core."""

import ctypes as C

import numpy as np

from claysampler_amd import abi


class Plan:
    """Owns the ctypes arrays a PlanView points into."""

    def __init__(self, spans, frames, tracks=None, mods=(), filter_tracks=0, music_end=None, limit=None,
                 onsets=(), body_frames=0, group=0, fade_frames=0):
        tracks = tracks or [[(0, 1.0, 0.0)], [(0, 1.0, 0.0)], [(0, 0.0, 0.0)]]
        points = [p for t in tracks for p in t]
        offsets = np.cumsum([0] + [len(t) for t in tracks]).astype(np.int32)
        self.spans = (abi.VoiceSpan * max(1, len(spans)))(*spans)
        self.points = (abi.ControlPoint * max(1, len(points)))(*[abi.ControlPoint(*p) for p in points])
        self.offsets = (C.c_int32 * len(offsets))(*offsets.tolist())
        self.mods = (abi.Modulation * max(1, len(mods)))(*[abi.Modulation(*m) for m in mods])
        self.onsets = (C.c_int32 * max(1, len(onsets)))(*onsets)
        self.view = abi.PlanView(self.spans, len(spans), self.points, self.offsets, len(tracks), self.mods, len(mods),
                                 filter_tracks, frames, frames if music_end is None else music_end,
                                 frames + 120 * 44100 if limit is None else limit, fade_frames, body_frames,
                                 self.onsets, len(onsets), group)


def span(begin, end, sample=0, source_end=1 << 20, position=0.0, step=1.0, channels=1, gain=1.0, pan=0.0,
         **fields):
    """A span with a flat unit envelope and centred pan unless fields say otherwise."""
    s = abi.VoiceSpan()
    s.sample, s.begin, s.end, s.source_end = sample, begin, end, source_end
    s.position, s.step, s.channels = position, step, channels
    s.filter_track = -1
    s.minimum_pitch_ratio = 1.0
    s.sample_rate = 44100
    s.attack_start = s.hold_start = s.decay_start = begin
    s.release_start = -1
    s.attack_step = 1.0
    s.sustain = 1.0
    s.start = 0.0
    s.gain, s.pan = gain, pan
    q = 0.7853981633974483
    s.left, s.right = gain * np.cos((1 + pan) * q), gain * np.cos((1 - pan) * q)
    s.width = 1.0
    s.width_direct, s.width_cross, s.position_left, s.position_right = 1.0, 0.0, 1.0, 1.0
    for k, v in fields.items():
        setattr(s, k, v)
    return s


def views(*plans):
    arr = (abi.PlanView * len(plans))()
    for i, p in enumerate(plans):
        arr[i] = p.view
    return arr


def bank_pointer(bank):
    bank = np.ascontiguousarray(bank, dtype=np.int16)
    return bank, bank.ctypes.data_as(C.POINTER(C.c_int16)), bank.size
