"""This project's MIDI loader vs the reference's own loader (src/io/midi.cpp,
compiled from its source file into oracle/_ref by oracle/Makefile), on
generated files: every event frame, kind and value, the end marker, the QA
onsets/body frames, and the error message of rejected files must be equal."""

import ctypes as C
import os
import struct

import pytest

import claysampler_amd as ca
import oracle
from claysampler_amd import synth

from helpers import smf

needs_reference = pytest.mark.skipif(not oracle.ref_midi_built(),
                                     reason="oracle/_ref/librefmidi.so not built (reference tree absent)")

CAP = 200000


def load(fn, path, rate, tail, measure):
    events = (C.c_int * (4 * CAP))()
    onsets = (C.c_int * CAP)()
    count, end, body, onset_count = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    err = C.create_string_buffer(512)
    status = fn(os.fsencode(path), rate, tail, measure, events, CAP, C.byref(count), C.byref(end), C.byref(body),
                onsets, CAP, C.byref(onset_count), err, 512)
    if status != 0:
        return ("error", err.value.decode())
    return (list(events[:4 * count.value]), end.value, body.value, list(onsets[:onset_count.value]))


def mine():
    lib = ca.host()
    lib.clay_midi_load.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int,
                                   C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_int]
    return lib.clay_midi_load


def varlen(v):
    out = [v & 0x7F]
    v >>= 7
    while v:
        out.append(0x80 | (v & 0x7F))
        v >>= 7
    return bytes(reversed(out))


def random_track(rng, channel=0, events=120, running=True, tempo=True):
    data = bytearray()
    last_status = None
    for _ in range(events):
        kind = rng.integer(0, 9)
        delta = rng.integer(0, 700) if rng.integer(0, 3) else 0
        if kind <= 3:
            status, body = 0x90 | channel, [rng.integer(30, 90), rng.integer(0, 127)]
        elif kind <= 5:
            status, body = 0x80 | channel, [rng.integer(30, 90), rng.integer(0, 127)]
        elif kind == 6:
            status, body = 0xB0 | channel, [[1, 7, 10, 11, 64, 66, 69, 121, 0, 91][rng.integer(0, 9)], rng.integer(0, 127)]
        elif kind == 7:
            status, body = 0xE0 | channel, [rng.integer(0, 127), rng.integer(0, 127)]
        elif kind == 8:
            status, body = 0xC0 | channel, [rng.integer(0, 127)]
        else:
            if tempo:
                data += varlen(delta) + b"\xff\x51\x03" + rng.integer(200000, 900000).to_bytes(3, "big")
                last_status = None
                continue
            status, body = 0x90 | channel, [60, 100]
        data += varlen(delta)
        if running and status == last_status:
            data += bytes(body)
        else:
            data += bytes([status] + body)
        last_status = status
    data += varlen(rng.integer(0, 3000)) + b"\xff\x2f\x00"
    return bytes(data)


def files(tmp_path):
    out = []
    d = str(tmp_path)
    sfz, mids = synth.config1(d, seed=3)
    out += mids
    out.append(synth.c2_midi(os.path.join(d, "c2.mid"), 7, seconds=12.0))
    out.append(synth.c3_midi(os.path.join(d, "c3.mid"), 8, 9.0, 40))
    for name, hexdata in [
        ("tail", "00 90 3c 64 60 80 3c 00 a5 40 ff 2f 00"),
        ("held", "00 b0 40 7f 00 90 3c 64 60 80 3c 00 00 b0 78 00 a5 40 ff 2f 00"),
        ("later", "00 b0 40 7f 00 90 3c 64 60 80 3c 00 82 68 b0 40 00 00 ff 2f 00"),
        ("sost", "00 90 3c 64 30 b0 42 7f 30 80 3c 00 82 68 b0 42 00 00 ff 2f 00"),
        ("gap", "00 90 3c 64 8b 20 80 3c 00 00 ff 2f 00"),
        ("segments", "00 b0 40 7f 00 90 3c 64 87 40 b0 01 40 87 40 e0 00 50 87 40 80 3c 00 87 40 b0 40 00 00 ff 2f 00"),
        ("no_eot", "00 90 3c 64 60 80 3c 00"),
        ("sysex", "00 f0 01 02 f7 00 ff 2f 00"),
        ("empty", "00 ff 2f 00"),
        ("two_channels", "00 90 3c 64 10 91 3e 64 60 80 3c 00 00 81 3e 00 00 ff 2f 00"),
    ]:
        out.append(smf(tmp_path / f"{name}.mid", bytes.fromhex(hexdata)))
    rng = synth.SplitMix64(11)
    for i in range(25):
        tracks = [random_track(rng, 0, rng.integer(5, 200), running=bool(i % 2), tempo=i % 3 != 0)
                  for _ in range(1 + i % 3)]
        path = tmp_path / f"rand{i}.mid"
        fmt = 0 if len(tracks) == 1 else 1
        with open(path, "wb") as f:
            f.write(b"MThd" + struct.pack(">IHHH", 6, fmt, len(tracks), [480, 96, 960][i % 3]))
            for t in tracks:
                f.write(b"MTrk" + struct.pack(">I", len(t)) + t)
        out.append(str(path))
    (tmp_path / "bad.mid").write_bytes(b"not MIDI")
    out.append(str(tmp_path / "bad.mid"))
    return out


@needs_reference
def test_midi_loader_matches_reference(tmp_path):
    # The reference-built loader is loaded here, in this CPU test only.
    REF = oracle.ref_midi()
    ours = mine()
    checked = 0
    for path in files(tmp_path):
        for rate in (44100, 96000):
            for tail in (0, 1, 2):
                for measure in (0, 1):
                    want = load(REF.clay_ref_load_midi, path, rate, tail, measure)
                    got = load(ours, path, rate, tail, measure)
                    assert got == want, (path, rate, tail, measure)
                    checked += 1
    assert checked > 400
