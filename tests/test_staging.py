"""The staging rules of the render core's pinned upload buffers
(claysampler_amd/csrc/hip/staging.h), on the CPU: tests/staging_check.cpp
replays random uploads of uneven sizes through a model of the copy queue in
which queued copies run in stream order at random times, and checks that
every copy reads the bytes it was queued with and no buffer is freed under a
queued copy. The pool of the persistent form (several producers, copies in
random order, a budget that forces waits) and the single image of the plain
form pass; a pool that hands a buffer out again before its copy has run is
caught."""

import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("staging") / "staging_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-o", exe, os.path.join(HERE, "staging_check.cpp")],
                   check=True)
    return exe


def run(exe, kind, seed, count):
    out = subprocess.run([exe, kind, str(seed), str(count)], check=True, capture_output=True, text=True).stdout
    fields = out.split()
    return int(fields[2]), int(fields[4])


@pytest.mark.parametrize("kind,count", [("pool", 5000), ("image", 100000)])
def test_no_copy_reads_rewritten_bytes(checker, kind, count):
    for seed in (1, 2):
        executed, failures = run(checker, kind, seed, count)
        assert executed == count and failures == 0, (seed, failures)


def test_the_check_catches_a_pool_that_ignores_fences(checker):
    _, failures = run(checker, "planted", 1, 5000)
    assert failures > 0
