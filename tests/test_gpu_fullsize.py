"""Full-size property checks: configuration 3 (512 uneven jobs, lognormal
5-300 s, polyphony 4-64, --gpu-batch-jobs 128) and configuration 5 (32 dense
jobs x 240 s at 96 kHz, 64-tap sinc, a ~2 GB resident bank, volume groups of 8)
render through clay_hip_render_window in the plain and the persistent form,
and every job's status, frame counts, group gain and PCM16 bytes must be
identical between the forms. Size-independent checks stand in for the FP64
oracle at these sizes (the oracle itself is held to the miniature cases)."""

import hashlib

import numpy as np
import pytest

import claysampler_amd as ca
from claysampler_amd import abi, synth

pytestmark = pytest.mark.gpu


def digests(plans, bank, form, taps, rate=44100, **renderer):
    r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps, form=form, rate=rate, **renderer)
    r.upload_bank(*bank)
    out = {}

    def on_batch(results, pcm, first, failed):
        assert not failed
        for k, res in enumerate(results):
            job = pcm[res.offset:res.offset + res.data_frames]
            out[first + k] = (res.status, res.frames, res.data_frames, res.fade, round(res.audio.gain, 9),
                              hashlib.blake2b(job.tobytes(), digest_size=16).hexdigest())

    r.render_window(plans, on_batch)
    r.close()
    assert sorted(out) == list(range(len(plans)))
    return out


def test_config3_forms_identical_at_full_size(synth_root):
    sfz, mids = synth.config3(f"{synth_root}/c3")
    s = ca.Session()
    for m in mids:
        s.add_job(sfz, m)
    plans, bank = s.plans(), s.bank()
    assert len(plans) == 512
    plain = digests(plans, bank, abi.FORM_PLAIN, 16, batch_jobs=128)
    rolled = digests(plans, bank, abi.FORM_PERSISTENT, 16, batch_jobs=128)
    assert plain == rolled
    assert {abi.TAIL_STATUS[v[0]] for v in plain.values()} == {"ok"}


def test_config5_forms_identical_at_full_size(synth_root):
    sfz, mids = synth.config5(f"{synth_root}/c5")
    s = ca.Session(rate=96000, measure=True)
    for j, m in enumerate(mids):
        s.add_job(sfz, m, group=j // 8 + 1)
    plans, bank = s.plans(), s.bank()
    assert len(plans) == 32 and bank[1] * 2 > 1.5e9
    kw = dict(batch_jobs=16, group_minimum=8, quality=True)
    plain = digests(plans, bank, abi.FORM_PLAIN, 64, rate=96000, **kw)
    rolled = digests(plans, bank, abi.FORM_PERSISTENT, 64, rate=96000, **kw)
    assert plain == rolled
