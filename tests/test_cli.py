"""The claysampler command line (src/main.cpp surface): options, manifest
rules, --check-sfz and error lines, on the CPU (no render)."""

import os
import subprocess

import pytest

from claysampler_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "..", "claysampler_amd", "claysampler")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="claysampler CLI not built")


def run(*args, cwd=None):
    p = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=120)
    return p.returncode, p.stdout, p.stderr


def test_help():
    rc, out, _ = run("--help")
    assert rc == 0 and out.startswith("Usage: claysampler --sfz preset.sfz --output directory")


@pytest.mark.parametrize("args,message", [
    ([], "provide --sfz, --output, and at least one MIDI path"),
    (["--jobs", "j.tsv", "--sfz", "a.sfz"], "--jobs supplies the SFZ, MIDI and output paths"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--tail-policy", "custom"], "custom tail requires --tail-max-seconds"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--tail-fade-ms", "5"],
     "tail length and fade options require --tail-policy custom"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--tail-policy", "orig", "--max-audio-seconds", "3"],
     "--max-audio-seconds requires --tail-policy natural"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--quality-gate", "strict"],
     "expected --quality-gate off, clayengine, or custom"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--qa-min-peak", "3"],
     "quality thresholds require --quality-gate custom"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--group-volume", "mean"], "--group-volume mean requires --jobs"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--group-min-members", "3"],
     "group options require --group-volume mean"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--flac-level", "9"], "invalid numeric argument 9"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--output-format", "mp3"], "expected --output-format wav or flac"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--bogus", "1"], "unknown option --bogus"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--device"], "missing value for --device"),
    (["--sfz", "a.sfz", "--output", "o", "a.mid", "--sinc-taps", "20"], "expected --sinc-taps 16, 32, 48 or 64"),
])
def test_option_errors(args, message):
    rc, out, err = run(*args)
    assert rc == 1
    assert err == f"ERROR\t{message}\n"


def manifest(tmp_path, lines, name="jobs.tsv"):
    path = tmp_path / name
    path.write_text("".join(line + "\n" for line in lines))
    return path


def test_manifest_rules(tmp_path):
    rc, _, err = run("--jobs", tmp_path / "missing.tsv")
    assert rc == 1 and err.startswith("ERROR\tcannot open job manifest")
    rc, _, err = run("--jobs", manifest(tmp_path, ["-\ta.sfz\ta.mid"]))
    assert err == "ERROR\texpected group, SFZ, MIDI and output on manifest line 1\n"
    rc, _, err = run("--jobs", manifest(tmp_path, ["-\ta.sfz\ta.mid\tout/a.flac"]))
    assert err == "ERROR\texpected a .wav output on manifest line 1\n"
    rc, _, err = run("--jobs", manifest(tmp_path, ["-\ta.sfz\ta.mid\tout/a.wav", "-\tb.sfz\tb.mid\tout/./a.wav"]))
    assert err.startswith("ERROR\tduplicate output path")
    rc, _, err = run("--jobs", manifest(tmp_path, ["-\ta.sfz\ta.mid\tout/a.wav\textra"]))
    assert err == "ERROR\tinvalid job manifest line 1\n"
    rc, _, err = run("--jobs", manifest(tmp_path, ["", "\r"]))
    assert err == "ERROR\tjob manifest is empty\n"


def test_check_sfz(tmp_path):
    sfz, _ = synth.config1(str(tmp_path / "c1"), seed=3)
    (tmp_path / "bad.sfz").write_text("<region> sample=missing.wav\n")
    rc, out, _ = run("--check-sfz", sfz, tmp_path / "bad.sfz")
    lines = out.splitlines()
    assert lines[0] == "SFZ\tOK\t1"
    assert lines[1].startswith("SFZ\tREJECT\t")
    assert rc == 2


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: the render runs instead")
def test_no_device_is_an_error(tmp_path):
    sfz, mids = synth.config1(str(tmp_path / "c1"), seed=3)
    rc, out, err = run("--sfz", sfz, "--output", tmp_path / "out", mids[0])
    assert rc == 1 and err.startswith("ERROR\t")
