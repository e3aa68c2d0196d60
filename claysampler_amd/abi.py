"""ctypes mirror of include/clay_hip.h and include/clay_host.h.

Field order and types follow the C headers exactly; tests compare every
struct size with the size the C compiler reports (clay_abi_sizeof).
"""

import ctypes as C

OK, E_INVALID, E_DEVICE, E_MEMORY, E_CALLBACK = 0, -1, -2, -3, -4
POLICY_NATURAL, POLICY_ORIG, POLICY_CUSTOM = 0, 1, 2
FORM_PLAIN, FORM_PERSISTENT = 0, 1
FILTER_SLOTS = 5
SINC_PHASES = 2048

TAIL_STATUS = [
    "ok", "duration", "nonfinite", "peak", "audioDuration", "quietPerformance", "excessiveSilence",
    "silentAudio", "clipping", "silentBody", "incompleteGroup", "scratchMemory",
]

# Reject / verdict strings of the reference CLI (src/main.cpp:363-380).
TAIL_MESSAGE = {
    1: "natural tail exceeded the tail or audio duration limit",
    2: "natural render produced non-finite audio",
    3: "excessive_peak: natural render requires more than 5 dB attenuation",
    4: "rendered_audio_too_long",
    5: "quiet_performance",
    6: "excessive_audio_silence",
    7: "effectively_silent_audio",
    8: "excessive_audio_clipping",
    9: "silent_melody_body",
    10: "insufficient_qualified_group_members",
    11: "CUDA mapped scratch allocation: out of memory",
}


class VoiceFilter(C.Structure):
    _fields_ = [("type", C.c_int32), ("cutoff", C.c_float), ("resonance", C.c_float),
                ("bandwidth", C.c_float), ("gain", C.c_float)]


class EnvelopeShape(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("attack_start", "hold_start", "decay_start", "release_start", "fade_start")] + \
               [(n, C.c_float) for n in ("start", "attack_step", "decay_rate", "sustain", "release_value",
                                         "release_rate", "fade_value", "fade_step", "depth", "release_seconds")]


class VoiceSpan(C.Structure):
    _fields_ = [
        ("sample", C.c_uint64), ("position", C.c_double), ("step", C.c_double), ("pitch_slope", C.c_double),
        ("channels", C.c_int32), ("begin", C.c_int32), ("end", C.c_int32), ("source_end", C.c_int32),
        ("loop_first", C.c_int32), ("loop_end", C.c_int32), ("loop_crossfade", C.c_int32),
        ("stereo_matrix", C.c_int32), ("generator", C.c_int32), ("noise_seed", C.c_uint32), ("noise_start", C.c_int32),
        ("first_modulation", C.c_int32), ("modulation_count", C.c_int32), ("filter_track", C.c_int32),
        ("variable_pitch", C.c_int32), ("minimum_pitch_ratio", C.c_float), ("sample_rate", C.c_int32),
        ("filters", VoiceFilter * FILTER_SLOTS), ("extra_envelope", EnvelopeShape * 2),
        ("active_envelope", C.c_int32 * 2),
        ("attack_start", C.c_int32), ("hold_start", C.c_int32), ("decay_start", C.c_int32),
        ("release_start", C.c_int32), ("fade_start", C.c_int32), ("bend_end", C.c_int32),
    ] + [(n, C.c_float) for n in (
        "gain", "pan", "left", "right", "start", "attack_step", "decay_rate", "sustain", "release_value",
        "release_rate", "fade_value", "fade_step", "width_direct", "width_cross", "position_left",
        "position_right", "width", "stereo_position")]


class ControlPoint(C.Structure):
    _fields_ = [("frame", C.c_int32), ("value", C.c_float), ("step", C.c_float)]


class Modulation(C.Structure):
    _fields_ = [("track", C.c_int32), ("target", C.c_int32), ("depth", C.c_float), ("shape_track", C.c_int32)]


class PlanView(C.Structure):
    _fields_ = [
        ("spans", C.POINTER(VoiceSpan)), ("span_count", C.c_int32),
        ("points", C.POINTER(ControlPoint)), ("track_offsets", C.POINTER(C.c_int32)), ("track_count", C.c_int32),
        ("modulation", C.POINTER(Modulation)), ("modulation_count", C.c_int32),
        ("filter_tracks", C.c_int32), ("frames", C.c_int32), ("music_end", C.c_int32), ("limit", C.c_int32),
        ("fade_frames", C.c_int32), ("body_frames", C.c_int32),
        ("onsets", C.POINTER(C.c_int32)), ("onset_count", C.c_int32), ("volume_group", C.c_uint64),
    ]


class AudioConfig(C.Structure):
    _fields_ = [("rate", C.c_int32), ("window", C.c_int32), ("max_frames", C.c_int32), ("min_peak", C.c_int32),
                ("enabled", C.c_int32)] + \
               [(n, C.c_double) for n in ("min_median", "min_quartile", "silence_power", "max_silent_fraction",
                                          "max_silent_frames", "max_clipped_fraction")]


class AudioMetrics(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("median", "quartile", "silent_fraction", "clipped_fraction", "body_rms",
                                          "target_rms", "gain")] + \
               [(n, C.c_int32) for n in ("longest_silence", "clipped", "verdict", "measured")]


class TailResult(C.Structure):
    _fields_ = [("offset", C.c_uint64)] + \
               [(n, C.c_int32) for n in ("frames", "data_frames", "fade", "peak", "first", "last", "status")] + \
               [("scale", C.c_float), ("audio", AudioMetrics)]


class HipConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("tail_policy", C.c_int32), ("sinc_taps", C.c_int32), ("form", C.c_int32),
                ("audio", AudioConfig), ("group_minimum", C.c_int32), ("waiting_limit", C.c_uint64),
                ("batch_jobs", C.c_uint64), ("segment_frames", C.c_int32), ("flags", C.c_int32),
                ("voice_silence_db", C.c_float), ("pad0", C.c_int32)]


SILENT_CHUNKS = 16


FLAG_NO_GRAPH = 1
FLAG_NO_TUNE = 2


class RenderBatch(C.Structure):
    _fields_ = [("pcm", C.POINTER(C.c_int16)), ("pcm_handle", C.c_void_p), ("results", C.POINTER(TailResult)),
                ("result_count", C.c_size_t), ("kernel_seconds", C.c_double), ("failed_scratch_bytes", C.c_uint64),
                ("job_indices", C.POINTER(C.c_size_t))]


class HipStats(C.Structure):
    _fields_ = [("device", C.c_char * 256)] + [(n, C.c_uint64) for n in (
        "peak_bytes", "sample_cache_budget", "peak_waiting_bytes", "spilled_bytes", "downloaded_bytes",
        "rounds", "admissions", "segments", "launches", "graph_launches")] + [
        ("segment_frames", C.c_int32), ("mix_blocks", C.c_int32), ("chain_waves", C.c_int32),
        ("audio_unroll", C.c_int32), ("launch_tuning", C.c_int32), ("pad", C.c_int32)]


class SessionConfig(C.Structure):
    _fields_ = [("rate", C.c_int32), ("tail_policy", C.c_int32), ("measure", C.c_int32), ("keep_onsets", C.c_int32),
                ("tail_max_seconds", C.c_double), ("max_audio_seconds", C.c_double), ("tail_fade_ms", C.c_int32)]


PUBLISH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(RenderBatch), C.c_size_t)
REFILL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.POINTER(PlanView)), C.POINTER(C.c_size_t))

ABI_STRUCTS = {
    "clay_voice_filter": VoiceFilter, "clay_envelope_shape": EnvelopeShape, "clay_voice_span": VoiceSpan,
    "clay_control_point": ControlPoint, "clay_modulation": Modulation, "clay_plan_view": PlanView,
    "clay_audio_config": AudioConfig, "clay_audio_metrics": AudioMetrics, "clay_tail_result": TailResult,
    "clay_hip_config": HipConfig, "clay_render_batch": RenderBatch, "clay_hip_stats": HipStats,
    "clay_session_config": SessionConfig,
}


def default_audio_config(rate=44100, enabled=False):
    """QA defaults of the reference CLI (src/main.cpp:169-187)."""
    power = lambda dbfs: 4294967296.0 * 10 ** (dbfs / 10)
    return AudioConfig(rate, int(round(rate / 20)), 375 * rate, 32, 1 if enabled else 0,
                       -38.0, -45.0, power(-70), 0.75, 10.0 * rate, 0.0001)


# Learned voice-tail predictor (include/clay_hip.h clay_tail_model).
MLP_INPUTS, MLP_HIDDEN = 12, 64
MLP_PARAMS = MLP_INPUTS * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN + 1
MOD_CUTOFF = 15


class TailModel(C.Structure):
    _fields_ = [("input_mean", C.c_float * MLP_INPUTS), ("input_scale", C.c_float * MLP_INPUTS),
                ("output_mean", C.c_float), ("output_scale", C.c_float), ("params", C.c_float * MLP_PARAMS)]


class MlpTrainConfig(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64)]
