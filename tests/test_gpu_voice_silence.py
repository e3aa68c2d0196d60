"""Voice-silence tail policy on the GPU: the chain kernel's online judge must
retire every voice chain at exactly the oracle's frame, the mix must drop the
retired frames (PCM16 within +-1 LSB of the oracle), and plain, segmented and
persistent forms must stay bit-identical with the policy on."""

import numpy as np
import pytest

import claysampler_amd as ca
import oracle
from claysampler_amd import abi, synth

from helpers import assert_parity
from planbuild import Plan, bank_pointer, span, views

pytestmark = pytest.mark.gpu

DB = -90.0


def gpu(session_or_plans, bank, policy=abi.POLICY_NATURAL, taps=16, segment_frames=0, db=DB):
    plans, (ptr, count) = session_or_plans, bank
    r = ca.Renderer(policy=policy, taps=taps, segment_frames=segment_frames, voice_silence_db=db)
    r.upload_bank(ptr, count)
    b = r.prepare(plans)
    b.run()
    results, pcm = b.results()
    tails = b.voice_tails()
    b.close()
    r.close()
    return results, pcm, tails


def oracle_side(plans, bank, policy=abi.POLICY_NATURAL, taps=16, db=DB):
    ptr, count = bank
    cfg = oracle.config(policy=policy, taps=taps, threads=4, voice_silence_db=db)
    results, pcm = oracle.render(cfg, ptr, count, plans)
    tails = np.concatenate([oracle.voice_tails(cfg, ptr, count, p) for p in plans]).astype(np.int32)
    return results, pcm, tails


def test_closed_form_chains():
    release, rate = 2970, -0.001
    plans = [Plan([span(0, 40000, filter_track=0, release_start=release, release_value=1.0, release_rate=rate,
                        fade_start=1 << 30)], 40000, filter_tracks=1),
             Plan([span(0, 30000, filter_track=0, release_start=5000, release_value=1.0, release_rate=-0.001,
                        fade_start=1 << 30, sample=1 << 20)], 30000, filter_tracks=1)]
    bank_arr, ptr, n = bank_pointer(np.concatenate([np.full(1 << 20, 16384), np.zeros(1 << 20)]))
    arr = views(*plans)
    g = gpu(arr, (ptr, n))
    o = oracle_side(arr, (ptr, n))
    assert g[2].tolist() == o[2].tolist()
    assert_parity((g[0], g[1]), (o[0], o[1]))


def c2_short(root, jobs=4, seconds=8.0):
    sfz = synth.c2_instrument(f"{root}/c2")
    s = ca.Session()
    for j in range(jobs):
        s.add_job(sfz, synth.c2_midi(f"{root}/c2/short_{j}.mid", 500 + j, seconds=seconds))
    return s


def test_config2_tails_match_the_oracle(synth_root):
    s = c2_short(synth_root)
    plans, bank = s.plans(), s.bank()
    g = gpu(plans, bank, taps=32)
    o = oracle_side(plans, bank, taps=32)
    ends = oracle_side(plans, bank, taps=32, db=0.0)[2]  # chain ends without the policy
    assert int((o[2] < ends).sum()) > len(ends) // 4, "too few voices retire: the policy is not exercised"
    assert g[2].tolist() == o[2].tolist()
    assert_parity((g[0], g[1]), (o[0], o[1]))


def test_policy_forms_bit_identical(synth_root):
    s = c2_short(synth_root, jobs=3)
    plans, bank = s.plans(), s.bank()
    whole = gpu(plans, bank, taps=32)
    for seg in (256, 4096):
        part = gpu(plans, bank, taps=32, segment_frames=seg)
        assert np.array_equal(whole[1], part[1]) and whole[2].tolist() == part[2].tolist()
    r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=32, form=abi.FORM_PERSISTENT, batch_jobs=2,
                    segment_frames=2048, flags=abi.FLAG_NO_TUNE, voice_silence_db=DB)
    r.upload_bank(*bank)
    got = {}

    def on_batch(results, pcm, first, failed):
        for k, res in enumerate(results):
            got[first + k] = pcm[res.offset:res.offset + res.data_frames].copy()

    r.render_window(plans, on_batch)
    r.close()
    for k, res in enumerate(whole[0]):
        assert np.array_equal(got[k], whole[1][res.offset:res.offset + res.data_frames]), k
