"""MI355X-native ClaySampler render core.

Python access to the two native libraries built in this package directory:

* ``libclayhost.so`` - SFZ/MIDI loading and the CPU planner (include/clay_host.h)
* ``libclayhip.so``  - the HIP render core for gfx950 (include/clay_hip.h)

Nothing here renders audio in Python: every render goes through the HIP
library, and a missing library is an error, never a fallback.
"""

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "libclayhost.so")
HIP_LIB = os.path.join(HERE, "libclayhip.so")

_host = None
_hip = None


def _bind(lib, name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


def host():
    """The host planner library (loaded once)."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(HOST_LIB)
        P = C.c_void_p
        _bind(lib, "clay_session_create", C.c_int, C.POINTER(abi.SessionConfig), C.POINTER(P))
        _bind(lib, "clay_session_destroy", C.c_int, P)
        _bind(lib, "clay_session_add_job", C.c_int, P, C.c_char_p, C.c_char_p, C.c_uint64, C.POINTER(C.c_size_t))
        _bind(lib, "clay_session_plan", C.POINTER(abi.PlanView), P, C.c_size_t)
        _bind(lib, "clay_session_job_count", C.c_size_t, P)
        _bind(lib, "clay_session_bank", C.POINTER(C.c_int16), P, C.POINTER(C.c_uint64))
        _bind(lib, "clay_session_reset", C.c_int, P)
        _bind(lib, "clay_session_last_error", C.c_char_p, P)
        _bind(lib, "clay_abi_sizeof", C.c_size_t, C.c_char_p)
        _bind(lib, "clay_write_wave", C.c_int, C.c_char_p, C.POINTER(C.c_int16), C.c_int32, C.c_int32, C.c_int32,
              C.c_int32)
        _bind(lib, "clay_write_flac", C.c_int, C.c_char_p, C.POINTER(C.c_int16), C.c_int32, C.c_int32, C.c_int32,
              C.c_int32, C.c_int32)
        _bind(lib, "clay_read_flac", C.c_int, C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
              C.POINTER(C.c_int16), C.c_uint64, C.POINTER(C.c_uint64))
        _host = lib
    return _host


HIP_SYMBOLS = [
    "clay_hip_create", "clay_hip_destroy", "clay_hip_upload_bank", "clay_hip_render_window",
    "clay_hip_release_pcm", "clay_hip_get_stats", "clay_hip_last_error", "clay_hip_sinc_table",
    "clay_hip_batch_prepare", "clay_hip_batch_run", "clay_hip_batch_results", "clay_hip_batch_kernel_time",
    "clay_hip_batch_destroy", "clay_hip_batch_voice_tails", "clay_mlp_init", "clay_hip_mlp_gradient",
    "clay_hip_mlp_train", "clay_hip_mlp_predict", "clay_hip_mlp_gradient_noalloc",
    "clay_hip_batch_prepare_bounded", "clay_hip_batch_overflows", "clay_hip_batch_scratch_bytes",
    "clay_hip_set_memory_limit", "clay_hip_batch_memory",
]


def hip():
    """The HIP render core library (loaded once). Raises if it is not built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError(f"{HIP_LIB} is not built; the HIP render core is required")
        lib = C.CDLL(HIP_LIB)
        P = C.c_void_p
        _bind(lib, "clay_hip_create", C.c_int, C.POINTER(abi.HipConfig), C.POINTER(P))
        _bind(lib, "clay_hip_destroy", C.c_int, P)
        _bind(lib, "clay_hip_upload_bank", C.c_int, P, C.POINTER(C.c_int16), C.c_uint64, C.c_int)
        _bind(lib, "clay_hip_render_window", C.c_int, P, C.POINTER(abi.PlanView), C.c_size_t, abi.PUBLISH_FN,
              abi.REFILL_FN, P)
        _bind(lib, "clay_hip_release_pcm", C.c_int, P, P)
        _bind(lib, "clay_hip_batch_voice_tails", C.c_int, P, P, C.POINTER(C.c_int32), C.c_size_t,
              C.POINTER(C.c_size_t))
        F = C.POINTER(C.c_float)
        _bind(lib, "clay_mlp_init", C.c_int, C.POINTER(abi.TailModel), F, F, C.c_size_t, C.c_uint64)
        _bind(lib, "clay_hip_mlp_gradient", C.c_int, P, C.POINTER(abi.TailModel), F, F, C.c_size_t, F, F)
        _bind(lib, "clay_hip_mlp_train", C.c_int, P, C.POINTER(abi.TailModel), F, F, C.c_size_t,
              C.POINTER(abi.MlpTrainConfig), F)
        _bind(lib, "clay_hip_mlp_predict", C.c_int, P, C.POINTER(abi.TailModel), F, C.c_size_t, F)
        _bind(lib, "clay_hip_get_stats", C.c_int, P, C.POINTER(abi.HipStats))
        _bind(lib, "clay_hip_last_error", C.c_char_p, P)
        _bind(lib, "clay_hip_sinc_table", C.c_int, C.c_int, C.POINTER(C.c_float), C.c_size_t)
        _bind(lib, "clay_hip_batch_prepare", C.c_int, P, C.POINTER(abi.PlanView), C.c_size_t, C.POINTER(P))
        _bind(lib, "clay_hip_batch_run", C.c_int, P, P, C.POINTER(C.c_double))
        _bind(lib, "clay_hip_batch_results", C.c_int, P, P, C.POINTER(abi.TailResult), C.c_size_t,
              C.POINTER(C.c_int16), C.c_uint64, C.POINTER(C.c_uint64))
        _bind(lib, "clay_hip_batch_kernel_time", C.c_int, P, P, C.c_char_p, C.POINTER(C.c_double),
              C.POINTER(C.c_int))
        _bind(lib, "clay_hip_batch_destroy", C.c_int, P, P)
        _bind(lib, "clay_hip_batch_prepare_bounded", C.c_int, P, C.POINTER(abi.PlanView), C.c_size_t,
              C.POINTER(C.c_int32), C.c_size_t, C.POINTER(P))
        _bind(lib, "clay_hip_batch_overflows", C.c_int, P, P, C.POINTER(C.c_int32), C.c_size_t,
              C.POINTER(C.c_size_t))
        _bind(lib, "clay_hip_batch_scratch_bytes", C.c_int, P, P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
        _bind(lib, "clay_hip_set_memory_limit", C.c_int, P, C.c_uint64)
        _bind(lib, "clay_hip_batch_memory", C.c_int, P, P, C.POINTER(C.c_int32), C.POINTER(C.c_int32))
        _hip = lib
    return _hip


class ClayError(RuntimeError):
    pass


class Session:
    """Planner session: one sample bank, instruments cached by path."""

    def __init__(self, rate=44100, policy=abi.POLICY_NATURAL, measure=False, keep_onsets=False,
                 tail_max_seconds=-1.0, max_audio_seconds=-1.0, tail_fade_ms=-1):
        self.lib = host()
        cfg = abi.SessionConfig(rate, policy, int(measure), int(keep_onsets), tail_max_seconds, max_audio_seconds,
                                tail_fade_ms)
        self.handle = C.c_void_p()
        if self.lib.clay_session_create(C.byref(cfg), C.byref(self.handle)) != abi.OK:
            raise ClayError("invalid session configuration")
        self.rate, self.policy = rate, policy

    def add_job(self, sfz, midi, group=0):
        index = C.c_size_t()
        status = self.lib.clay_session_add_job(self.handle, os.fsencode(sfz), os.fsencode(midi), group, C.byref(index))
        if status != abi.OK:
            raise ClayError(self.lib.clay_session_last_error(self.handle).decode())
        return index.value

    def plan(self, index):
        return self.lib.clay_session_plan(self.handle, index).contents

    def plans(self):
        n = self.lib.clay_session_job_count(self.handle)
        arr = (abi.PlanView * n)()
        for i in range(n):
            arr[i] = self.plan(i)
        return arr

    def bank(self):
        count = C.c_uint64()
        ptr = self.lib.clay_session_bank(self.handle, C.byref(count))
        return ptr, count.value

    def bank_array(self):
        ptr, count = self.bank()
        if not count:
            return np.zeros(0, np.int16)
        return np.ctypeslib.as_array(ptr, shape=(count,)).copy()

    def close(self):
        if self.handle:
            self.lib.clay_session_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Renderer:
    """The HIP render core (clay_hip handle)."""

    def __init__(self, policy=abi.POLICY_NATURAL, taps=48, rate=44100, quality=False, group_minimum=0,
                 batch_jobs=64, device=0, form=abi.FORM_PLAIN, audio=None, segment_frames=0, flags=0,
                 voice_silence_db=0.0):
        self.lib = hip()
        audio = audio if audio is not None else abi.default_audio_config(rate, quality)
        self.cfg = abi.HipConfig(device, policy, taps, form, audio, group_minimum, 256 << 20, batch_jobs,
                                 segment_frames, flags, voice_silence_db, 0)
        self.handle = C.c_void_p()
        if self.lib.clay_hip_create(C.byref(self.cfg), C.byref(self.handle)) != abi.OK:
            raise ClayError(self.lib.clay_hip_last_error(None).decode())
        self.channels = 1 if policy == abi.POLICY_NATURAL else 2

    def _check(self, status):
        if status != abi.OK:
            raise ClayError(f"status {status}: {self.lib.clay_hip_last_error(self.handle).decode()}")

    def upload_bank(self, ptr, count, fresh=False):
        self._check(self.lib.clay_hip_upload_bank(self.handle, ptr, count, int(fresh)))

    def prepare(self, plans, scratch_end=None):
        """A resident batch; `scratch_end` (one absolute frame per voice chain)
        bounds each chain's filtered-voice scratch (voice-silence policy)."""
        b = C.c_void_p()
        if scratch_end is None:
            self._check(self.lib.clay_hip_batch_prepare(self.handle, plans, len(plans), C.byref(b)))
        else:
            ends = np.ascontiguousarray(scratch_end, np.int32)
            self._check(self.lib.clay_hip_batch_prepare_bounded(
                self.handle, plans, len(plans), ends.ctypes.data_as(C.POINTER(C.c_int32)), ends.size, C.byref(b)))
        return Batch(self, b, len(plans))

    def set_memory_limit(self, bytes_):
        """Cap on one batch arena's device bytes (0 = none); arenas over it fail as out of memory."""
        self._check(self.lib.clay_hip_set_memory_limit(self.handle, int(bytes_)))

    def stats(self):
        s = abi.HipStats()
        self._check(self.lib.clay_hip_get_stats(self.handle, C.byref(s)))
        return s

    def render_window(self, plans, on_batch, refills=()):
        """Render through clay_hip_render_window; on_batch(results, pcm, first, failed) per published
        batch. `refills` are further PlanView arrays handed out one per refill call."""
        errors = []
        pending = list(refills)
        handed = []

        def refill(user, wait, out_plans, out_count):
            if not pending:
                return 0
            arr = pending.pop(0)
            handed.append(arr)
            out_plans[0] = C.cast(arr, C.POINTER(abi.PlanView))
            out_count[0] = len(arr)
            return 1

        def publish(user, batch_ptr, first):
            try:
                b = batch_ptr.contents
                # Copies: the results array belongs to the library and is reused after the call.
                results = [abi.TailResult.from_buffer_copy(b.results[i]) for i in range(b.result_count)]
                samples = 0
                for r in results:
                    samples = max(samples, (r.offset + r.data_frames) * self.channels)
                pcm = np.ctypeslib.as_array(b.pcm, shape=(samples,)).copy() if samples else np.zeros(0, np.int16)
                if b.pcm_handle:
                    self.lib.clay_hip_release_pcm(self.handle, b.pcm_handle)
                on_batch(results, pcm, first, b.failed_scratch_bytes)
                return 0
            except Exception as e:  # propagate after the C call returns
                errors.append(e)
                return 1

        cb = abi.PUBLISH_FN(publish)
        rf = abi.REFILL_FN(refill) if refills else abi.REFILL_FN()
        status = self.lib.clay_hip_render_window(self.handle, plans, len(plans), cb, rf, None)
        if errors:
            raise errors[0]
        self._check(status)

    def close(self):
        if self.handle:
            self.lib.clay_hip_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A batch resident on the device (clay_hip_batch)."""

    def __init__(self, renderer, handle, count):
        self.r, self.handle, self.count = renderer, handle, count

    def run(self):
        seconds = C.c_double()
        self.r._check(self.r.lib.clay_hip_batch_run(self.r.handle, self.handle, C.byref(seconds)))
        return seconds.value

    def results(self):
        res = (abi.TailResult * self.count)()
        samples = C.c_uint64()
        self.r._check(self.r.lib.clay_hip_batch_results(self.r.handle, self.handle, res, self.count, None, 0,
                                                        C.byref(samples)))
        pcm = np.zeros(samples.value, np.int16)
        self.r._check(self.r.lib.clay_hip_batch_results(
            self.r.handle, self.handle, res, self.count, pcm.ctypes.data_as(C.POINTER(C.c_int16)), pcm.size,
            C.byref(samples)))
        return list(res), pcm

    def voice_tails(self):
        """Exact per-chain retirement frames of the last run (chain end when not retired)."""
        n = C.c_size_t()
        self.r._check(self.r.lib.clay_hip_batch_voice_tails(self.r.handle, self.handle, None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value), np.int32)
        self.r._check(self.r.lib.clay_hip_batch_voice_tails(self.r.handle, self.handle,
                                                            out.ctypes.data_as(C.POINTER(C.c_int32)), out.size,
                                                            C.byref(n)))
        return out[:n.value]

    def overflows(self):
        """Jobs whose voice chains reached their scratch bound in the last run."""
        n = C.c_size_t()
        self.r._check(self.r.lib.clay_hip_batch_overflows(self.r.handle, self.handle, None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value), np.int32)
        self.r._check(self.r.lib.clay_hip_batch_overflows(self.r.handle, self.handle,
                                                          out.ctypes.data_as(C.POINTER(C.c_int32)), out.size,
                                                          C.byref(n)))
        return out[:n.value]

    def memory(self):
        """(render window frames, 0 = whole batch; filtered scratch in host memory) of the placed batch."""
        w, host = C.c_int32(), C.c_int32()
        self.r._check(self.r.lib.clay_hip_batch_memory(self.r.handle, self.handle, C.byref(w), C.byref(host)))
        return w.value, bool(host.value)

    def scratch_bytes(self):
        """(filtered-voice scratch bytes, whole device arena bytes) of the batch."""
        f, a = C.c_uint64(), C.c_uint64()
        self.r._check(self.r.lib.clay_hip_batch_scratch_bytes(self.r.handle, self.handle, C.byref(f), C.byref(a)))
        return f.value, a.value

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_int()
        self.r._check(self.r.lib.clay_hip_batch_kernel_time(self.r.handle, self.handle, name.encode(), C.byref(ms),
                                                            C.byref(n)))
        return ms.value, n.value

    def close(self):
        if self.handle:
            self.r.lib.clay_hip_batch_destroy(self.r.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sinc_table(taps):
    """The kernels' sinc table as (taps, phases + 1) float32."""
    out = np.zeros(taps * (abi.SINC_PHASES + 1), np.float32)
    if hip().clay_hip_sinc_table(taps, out.ctypes.data_as(C.POINTER(C.c_float)), out.size) != abi.OK:
        raise ClayError("invalid tap count")
    return out.reshape(taps, abi.SINC_PHASES + 1)
