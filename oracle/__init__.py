"""Test infrastructure: ctypes binding of the FP64 oracle (oracle/_build).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline import this,
as the checker of the HIP render core; the product never does.
"""

import ctypes as C
import os

import numpy as np

from claysampler_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libclayoracle.so")
REF_MIDI = os.path.join(HERE, "_ref", "librefmidi.so")

_lib = None


class OracleConfig(C.Structure):
    _fields_ = [("tail_policy", C.c_int32), ("sinc_taps", C.c_int32), ("audio", abi.AudioConfig),
                ("group_minimum", C.c_int32), ("threads", C.c_int32), ("voice_silence_db", C.c_float)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (make -C oracle)")
        L = C.CDLL(LIB)
        L.clay_oracle_sinc_table.argtypes = [C.c_int, C.POINTER(C.c_float), C.c_size_t]
        L.clay_oracle_mix.argtypes = [C.POINTER(OracleConfig), C.POINTER(C.c_int16), C.c_uint64,
                                      C.POINTER(abi.PlanView), C.POINTER(C.c_double), C.c_uint64]
        L.clay_oracle_render.argtypes = [C.POINTER(OracleConfig), C.POINTER(C.c_int16), C.c_uint64,
                                         C.POINTER(abi.PlanView), C.c_size_t, C.POINTER(abi.TailResult),
                                         C.POINTER(C.c_int16), C.c_uint64, C.POINTER(C.c_uint64)]
        L.clay_oracle_envelope.restype = C.c_double
        L.clay_oracle_voice_tails.argtypes = [C.POINTER(OracleConfig), C.POINTER(C.c_int16), C.c_uint64,
                                              C.POINTER(abi.PlanView), C.POINTER(C.c_int32), C.c_size_t]
        L.clay_oracle_envelope.argtypes = [C.POINTER(abi.VoiceSpan), C.c_int]
        L.clay_oracle_filter_coefficients.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                                      C.c_double, C.POINTER(C.c_double)]
        L.clay_oracle_interpolate.argtypes = [C.POINTER(abi.VoiceSpan), C.POINTER(C.c_int16), C.c_int, C.c_double,
                                              C.POINTER(C.c_double)]
        L.clay_oracle_mono.argtypes = [C.c_int, C.c_int]
        L.clay_oracle_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def config(policy=abi.POLICY_NATURAL, taps=48, rate=44100, quality=False, group_minimum=0, threads=1, audio=None,
           voice_silence_db=0.0):
    audio = audio if audio is not None else abi.default_audio_config(rate, quality)
    return OracleConfig(policy, taps, audio, group_minimum, threads, voice_silence_db)


def voice_tails(cfg, bank_ptr, bank_count, plan):
    """Exact per-chain retirement frames of one job (chain end when it does not retire)."""
    out = np.zeros(max(1, plan.filter_tracks), np.int32)
    status = lib().clay_oracle_voice_tails(C.byref(cfg), bank_ptr, bank_count, C.byref(plan),
                                           out.ctypes.data_as(C.POINTER(C.c_int32)), out.size)
    assert status == 0
    return out[:plan.filter_tracks]


def sinc_table(taps):
    out = np.zeros(taps * (abi.SINC_PHASES + 1), np.float32)
    assert lib().clay_oracle_sinc_table(taps, out.ctypes.data_as(C.POINTER(C.c_float)), out.size) == 0
    return out.reshape(taps, abi.SINC_PHASES + 1)


def mix(cfg, bank_ptr, bank_count, plan):
    out = np.zeros((plan.frames, 2))
    status = lib().clay_oracle_mix(C.byref(cfg), bank_ptr, bank_count, C.byref(plan),
                                   out.ctypes.data_as(C.POINTER(C.c_double)), plan.frames)
    assert status == 0, lib().clay_oracle_last_error()
    return out


def render(cfg, bank_ptr, bank_count, plans):
    """Oracle results (list of TailResult) and packed PCM16 for a batch."""
    n = len(plans)
    res = (abi.TailResult * n)()
    samples = C.c_uint64()
    status = lib().clay_oracle_render(C.byref(cfg), bank_ptr, bank_count, plans, n, res, None, 0, C.byref(samples))
    assert status == 0, lib().clay_oracle_last_error()
    pcm = np.zeros(samples.value, np.int16)
    status = lib().clay_oracle_render(C.byref(cfg), bank_ptr, bank_count, plans, n, res,
                                      pcm.ctypes.data_as(C.POINTER(C.c_int16)), pcm.size, C.byref(samples))
    assert status == 0, lib().clay_oracle_last_error()
    return list(res), pcm


def ref_midi_built():
    return os.path.exists(REF_MIDI)


def ref_midi():
    """The reference's own MIDI loader (oracle/_ref), or None where it is not built."""
    if not os.path.exists(REF_MIDI):
        return None
    L = C.CDLL(REF_MIDI)
    L.clay_ref_load_midi.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_int]
    return L
