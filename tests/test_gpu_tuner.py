"""The plain form's launch tuner: the filter pass's voice chains per workgroup
(1, 2 or 4 waves) and the audio pass's frames per lane in flight (1 or 2) are
measured on a renderer's first launches and the cheapest kept. Every candidate
computes the same bits, so each run of a batch while the tuner tries them
matches a renderer with fixed launch shapes (CLAY_FLAG_NO_TUNE), and that one
matches the FP64 oracle.
"""

import pytest

import claysampler_amd as ca
from claysampler_amd import abi

from helpers import assert_parity, chain_shapes_case, oracle_render, rich_case

pytestmark = pytest.mark.gpu


def runs(session, taps, flags, count):
    bank_ptr, bank_count = session.bank()
    r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps, rate=session.rate, flags=flags)
    r.upload_bank(bank_ptr, bank_count)
    b = r.prepare(session.plans())
    out = []
    for _ in range(count):
        b.run()
        results, pcm = b.results()
        out.append(([(x.status, x.frames, x.data_frames, x.offset, x.peak) for x in results], pcm.tobytes()))
    first = b.results()
    stats = r.stats()
    b.close()
    r.close()
    return out, stats, first


@pytest.mark.parametrize("case", ["chain_shapes", "rich"])
def test_tuned_launch_shapes_keep_the_bits(tmp_path, case):
    sfz, mid = (chain_shapes_case if case == "chain_shapes" else rich_case)(tmp_path)
    mids = [mid, mid, mid]
    s = ca.Session()
    for m in mids:
        s.add_job(sfz, m)
    # 3 x 2 trials of the chain width, 2 x 2 of the unroll: 8 runs see them all.
    tuned, stats, _ = runs(s, 32, 0, 8)
    fixed, fixed_stats, (results, pcm) = runs(s, 32, abi.FLAG_NO_TUNE, 1)
    for i, got in enumerate(tuned):
        assert got == fixed[0], f"run {i} differs from the fixed launch shapes"
    assert stats.chain_waves in (1, 2, 4) and stats.audio_unroll in (1, 2)
    assert (fixed_stats.chain_waves, fixed_stats.audio_unroll) == (2, 1)
    assert_parity((results, pcm), oracle_render(s, abi.POLICY_NATURAL, 32))
