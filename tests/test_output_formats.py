"""PCM16 output files (src/io/audio.cpp:236-285): WAV checked with Python's
own `wave` reader; FLAC (this project's RFC 9639 encoder, replacing the
reference's libsndfile/libFLAC) round-tripped through this project's decoder
at every level, mono and stereo, padded and partial frames, and a corrupted
frame is refused by its CRC. The STREAMINFO MD5 equals hashlib's digest of
the audio. No independent FLAC decoder is installed here (no libFLAC, flac or
sox), so the frame and subframe coding is pinned by this decoder, the CRCs
and the MD5 only."""

import ctypes as C
import os
import wave

import numpy as np
import pytest

import claysampler_amd as ca
from claysampler_amd import abi


def write(kind, path, x, data, frames, channels, level=5):
    lib = ca.host()
    ptr = x.ctypes.data_as(C.POINTER(C.c_int16))
    if kind == "wav":
        return lib.clay_write_wave(os.fsencode(str(path)), ptr, data, frames, 44100, channels)
    return lib.clay_write_flac(os.fsencode(str(path)), ptr, data, frames, 44100, channels, level)


def read_flac(path):
    lib = ca.host()
    rate, channels, samples = C.c_int32(), C.c_int32(), C.c_uint64()
    if lib.clay_read_flac(os.fsencode(str(path)), C.byref(rate), C.byref(channels), None, 0,
                          C.byref(samples)) != abi.OK:
        return None
    pcm = np.zeros(samples.value, np.int16)
    lib.clay_read_flac(os.fsencode(str(path)), C.byref(rate), C.byref(channels),
                       pcm.ctypes.data_as(C.POINTER(C.c_int16)), pcm.size, C.byref(samples))
    return rate.value, channels.value, pcm


def signal(frames, channels, data, seed=1):
    rng = np.random.default_rng(seed)
    t = np.arange(frames)
    left = np.sin(t * 0.013) * 9000 + rng.integers(-40, 40, frames)
    x = np.stack([left, left * 0.6 + rng.integers(-5, 5, frames)], 1)[:, :channels]
    x[data:] = 0
    return np.ascontiguousarray(x.astype(np.int16).reshape(-1))


@pytest.mark.parametrize("channels", [1, 2])
def test_wave_header_and_padding(tmp_path, channels):
    x = signal(3000, channels, 2500)
    assert write("wav", tmp_path / "a.wav", x, 2500, 3000, channels) == abi.OK
    with wave.open(str(tmp_path / "a.wav")) as w:
        assert (w.getframerate(), w.getnchannels(), w.getsampwidth(), w.getnframes()) == (44100, channels, 2, 3000)
        got = np.frombuffer(w.readframes(3000), "<i2")
    assert np.array_equal(got, x)
    assert os.path.getsize(tmp_path / "a.wav") == 44 + 3000 * 2 * channels


@pytest.mark.parametrize("level", range(9))
@pytest.mark.parametrize("channels", [1, 2])
def test_flac_round_trip(tmp_path, level, channels):
    for frames, data in ((100352, 90000), (4096, 4096), (5000, 4000), (1, 1), (8193, 0)):
        x = signal(frames, channels, data, seed=frames + level)
        assert write("flac", tmp_path / "a.flac", x, data, frames, channels, level) == abi.OK
        rate, ch, got = read_flac(tmp_path / "a.flac")
        assert (rate, ch) == (44100, channels)
        assert np.array_equal(got, x), (frames, data)


def test_flac_extremes_and_compression(tmp_path):
    rng = np.random.default_rng(5)
    noise = rng.integers(-32768, 32768, 2 * 9000).astype(np.int16)          # verbatim subframes
    assert write("flac", tmp_path / "n.flac", noise, 9000, 9000, 2) == abi.OK
    assert np.array_equal(read_flac(tmp_path / "n.flac")[2], noise)
    tone = signal(44100, 1, 44100)
    assert write("flac", tmp_path / "t.flac", tone, 44100, 44100, 1) == abi.OK
    assert os.path.getsize(tmp_path / "t.flac") < 0.6 * 2 * 44100         # predictors do compress


@pytest.mark.parametrize("channels,data,frames", [(1, 30000, 30000), (2, 12345, 20000), (2, 0, 4096)])
def test_flac_streaminfo_md5_is_the_audio_digest(tmp_path, channels, data, frames):
    """STREAMINFO carries the MD5 of the unencoded audio (RFC 9639 8.2):
    frames x channels little-endian int16, zeros past the data frames, as
    hashlib computes it independently of this encoder."""
    import hashlib

    rng = np.random.default_rng(channels + data)
    x = rng.integers(-32768, 32767, max(1, data) * channels).astype(np.int16)
    assert write("flac", tmp_path / "m.flac", x, data, frames, channels, 5) == abi.OK
    raw = (tmp_path / "m.flac").read_bytes()
    assert raw[:4] == b"fLaC" and raw[4] & 0x7F == 0 and int.from_bytes(raw[5:8], "big") == 34
    audio = np.zeros(frames * channels, "<i2")
    audio[:data * channels] = x[:data * channels]
    assert raw[26:42] == hashlib.md5(audio.tobytes()).digest()


def test_flac_crc_rejects_corruption(tmp_path):
    x = signal(10000, 1, 10000)
    assert write("flac", tmp_path / "a.flac", x, 10000, 10000, 1) == abi.OK
    data = bytearray((tmp_path / "a.flac").read_bytes())
    data[len(data) // 2] ^= 0x10
    (tmp_path / "b.flac").write_bytes(bytes(data))
    assert read_flac(tmp_path / "b.flac") is None
