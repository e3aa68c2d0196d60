"""The claysampler command line on the GPU: a jobs.tsv manifest rendered
through planning windows, refill and writer lanes. It must write the same
PCM16 as the library path, publish WAV and FLAC files through temporary
names, report OK/REJECT/PASS lines, and skip complete outputs on --resume."""

import ctypes as C
import os
import subprocess
import wave

import numpy as np
import pytest

import claysampler_amd as ca
from claysampler_amd import abi, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "..", "claysampler_amd", "claysampler")


def run(*args):
    p = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout, p.stderr


def read_wav(path):
    with wave.open(str(path)) as w:
        return w.getframerate(), w.getnchannels(), np.frombuffer(w.readframes(w.getnframes()), "<i2")


def read_flac(path):
    lib = ca.host()
    rate, channels, samples = C.c_int32(), C.c_int32(), C.c_uint64()
    assert lib.clay_read_flac(os.fsencode(str(path)), C.byref(rate), C.byref(channels), None, 0,
                              C.byref(samples)) == abi.OK
    pcm = np.zeros(samples.value, np.int16)
    assert lib.clay_read_flac(os.fsencode(str(path)), C.byref(rate), C.byref(channels),
                              pcm.ctypes.data_as(C.POINTER(C.c_int16)), pcm.size, C.byref(samples)) == abi.OK
    return rate.value, channels.value, pcm


def jobs_case(tmp_path, fmt="wav"):
    root = tmp_path / "in"
    root.mkdir(exist_ok=True)
    sfz = synth.c1_instrument(str(root))
    mids = [synth.c3_midi(str(root / f"j{j}.mid"), 40 + j, secs, poly)
            for j, (secs, poly) in enumerate(((2.0, 8), (1.0, 4), (3.0, 24), (0.7, 4)))]
    (root / "bad.mid").write_bytes(b"not MIDI")
    rows = [f"-\t{os.path.relpath(sfz, tmp_path)}\t{os.path.relpath(m, tmp_path)}\tout/j{j}.{fmt}"
            for j, m in enumerate(mids)]
    rows.insert(2, f"-\t{os.path.relpath(sfz, tmp_path)}\tin/bad.mid\tout/bad.{fmt}")
    (tmp_path / "jobs.tsv").write_text("".join(r + "\n" for r in rows))
    return sfz, mids


def library_render(sfz, mids, policy, taps):
    s = ca.Session(policy=policy)
    for m in mids:
        s.add_job(sfz, m)
    bank_ptr, bank_count = s.bank()
    r = ca.Renderer(policy=policy, taps=taps, batch_jobs=64)
    r.upload_bank(bank_ptr, bank_count)
    out = {}

    def on_batch(results, pcm, first, failed):
        for k, res in enumerate(results):
            ch = r.channels
            data = pcm[res.offset * ch:(res.offset + res.data_frames) * ch]
            out[first + k] = (res, np.concatenate([data, np.zeros((res.frames - res.data_frames) * ch, np.int16)]))

    r.render_window(s.plans(), on_batch)
    r.close()
    return [out[i] for i in range(len(mids))]


@pytest.mark.parametrize("policy,form", [("natural", "persistent"), ("natural", "plain"), ("orig", "persistent")])
def test_cli_manifest_matches_library(tmp_path, policy, form):
    sfz, mids = jobs_case(tmp_path)
    args = ["--jobs", tmp_path / "jobs.tsv", "--sinc-taps", 16, "--cpu-batch-jobs", 2, "--gpu-batch-jobs", 2,
            "--tail-policy", policy, "--form", form]
    rc, out, err = run(*args)
    assert rc == 2, (out, err)  # one job rejected
    lines = out.splitlines()
    assert lines[0].startswith("DEVICE\t") and lines[1].startswith("INIT\t")
    oks = [l.split("\t") for l in lines if l.startswith("OK\t")]
    assert sorted(o[1] for o in oks) == sorted(str(tmp_path / "out" / f"j{j}.wav") for j in range(4))
    assert err.count("REJECT\t") == 1 and "bad.wav" in err
    passes = [l.split("\t") for l in lines if l.startswith("PASS\t")]
    assert len(passes) == 1 and passes[0][1] == "1" and passes[0][3] == "4"
    policy_id = abi.POLICY_NATURAL if policy == "natural" else abi.POLICY_ORIG
    want = library_render(sfz, mids, policy_id, 16)
    by_path = {o[1]: o for o in oks}
    for j, (res, pcm) in enumerate(want):
        path = tmp_path / "out" / f"j{j}.wav"
        rate, channels, got = read_wav(path)
        assert rate == 44100 and channels == (1 if policy == "natural" else 2)
        assert np.array_equal(got, pcm), j
        ok = by_path[str(path)]
        assert [int(x) for x in ok[2:]] == [res.frames, res.peak, res.first, res.last]
    assert not [p for p in os.listdir(tmp_path / "out") if p.endswith(".part")]


def test_cli_resume_and_flac(tmp_path):
    sfz, mids = jobs_case(tmp_path)
    base = ["--jobs", tmp_path / "jobs.tsv", "--sinc-taps", 16]
    rc, out, _ = run(*base)
    assert rc == 2
    rc, out, err = run(*base, "--resume")
    skips = [l for l in out.splitlines() if l.startswith("SKIP\t")]
    assert len(skips) == 4 and rc == 2 and err.count("REJECT\t") == 1
    # FLAC manifest: same PCM as the WAV outputs.
    flac_root = tmp_path / "flac"
    flac_root.mkdir()
    rows = (tmp_path / "jobs.tsv").read_text().splitlines()
    (flac_root / "jobs.tsv").write_text("".join(
        "\t".join([c[0], os.path.join("..", c[1]), os.path.join("..", c[2]), c[3].replace(".wav", ".flac")]) + "\n"
        for c in (r.split("\t") for r in rows)))
    rc, out, err = run("--jobs", flac_root / "jobs.tsv", "--sinc-taps", 16, "--output-format", "flac",
                       "--flac-level", 5)
    assert rc == 2, err
    for j in range(4):
        _, _, wav_pcm = read_wav(tmp_path / "out" / f"j{j}.wav")
        rate, channels, flac_pcm = read_flac(flac_root / "out" / f"j{j}.flac")
        assert rate == 44100 and channels == 1
        assert np.array_equal(flac_pcm, wav_pcm)
    rc, out, _ = run("--jobs", flac_root / "jobs.tsv", "--sinc-taps", 16, "--output-format", "flac", "--resume")
    assert len([l for l in out.splitlines() if l.startswith("SKIP\t")]) == 4
