"""The plain form under memory pressure, on the GPU, driven by the renderer's
memory limit (clay_hip_set_memory_limit): each step of the allocation chain
(src/cuda/renderer.cpp:864-989) must give the same PCM16 as an unconstrained
render:
- a limit the whole scratch does not fit: the batch renders in windows whose
  filtered-voice scratch fits the budget;
- a limit below even one job's windowed scratch: that job's scratch moves to
  mapped host memory;
- a limit below a multi-job batch: the render window halves the batch down to
  jobs that fit;
- a limit below one job's other buffers: every job is rejected as the
  reference rejects it (REJECT scratch memory), and the renderer goes on."""

import numpy as np
import pytest

import claysampler_amd as ca
from claysampler_amd import abi, synth

pytestmark = pytest.mark.gpu

TAPS = 32


def c2_session(root, seeds):
    sfz = synth.c2_instrument(f"{root}/c2")
    s = ca.Session()
    for seed in seeds:
        s.add_job(sfz, synth.c2_midi(f"{root}/c2/mem_{seed}.mid", seed, seconds=8.0))
    return s


def per_job(results, pcm):
    return [(r.status, r.frames, r.data_frames, pcm[r.offset:r.offset + r.data_frames].tobytes()) for r in results]


def render(s, plans, limit=0):
    """(per-job output, (window, host scratch), (filtered bytes, arena bytes)) of one resident batch."""
    r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=TAPS)
    r.upload_bank(*s.bank())
    r.set_memory_limit(limit)
    b = r.prepare(plans)
    b.run()
    out = per_job(*b.results()), b.memory(), b.scratch_bytes()
    b.close()
    r.close()
    return out


def windowed_scratch(plans, window):
    """Filtered scratch bytes of the plans' filtered spans in windows of `window` frames."""
    total = 0
    for p in plans:
        sp = np.ctypeslib.as_array(p.spans, shape=(p.span_count,))
        wet = sp[sp["filter_track"] >= 0]
        total += int((np.minimum(wet["end"] - wet["begin"], window) * wet["channels"]).sum()) * 4
    return total


def test_windows_then_host_scratch_give_identical_pcm(synth_root):
    s = c2_session(synth_root, range(7000, 7006))
    plans = s.plans()
    want, memory, (filtered, arena) = render(s, plans)
    assert memory == (0, False) and filtered > arena - filtered  # scratch dominates: windows can help
    fixed = arena - filtered
    got, memory, (small, _) = render(s, plans, limit=fixed + filtered // 3)
    assert memory[0] > 0 and not memory[1] and small < filtered
    assert got == want

    one = (abi.PlanView * 1)(plans[0])
    want1, _, (filtered1, arena1) = render(s, one)
    limit = arena1 - filtered1 + windowed_scratch(one, 256) // 2  # no window fits
    got1, memory, _ = render(s, one, limit=limit)
    assert memory[1]
    assert got1 == want1


def window_render(s, plans, limit):
    r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=TAPS, form=abi.FORM_PLAIN, batch_jobs=64)
    r.upload_bank(*s.bank())
    r.set_memory_limit(limit)
    jobs, rejected = {}, []

    def on_batch(results, pcm, first, failed):
        if failed:
            rejected.append(first)
            return
        for k, x in enumerate(results):
            jobs[first + k] = (x.status, x.frames, x.data_frames, pcm[x.offset:x.offset + x.data_frames].tobytes())

    r.render_window(plans, on_batch)
    r.close()
    return jobs, rejected


def test_window_halves_batches_and_rejects_what_cannot_fit(synth_root):
    s = c2_session(synth_root, range(7000, 7006))
    plans = s.plans()
    want, _ = window_render(s, plans, 0)
    one = (abi.PlanView * 1)(plans[0])
    _, _, (filtered1, arena1) = render(s, one)
    # Room for about one job's buffers: the batch of six is halved until it fits.
    got, rejected = window_render(s, plans, int(1.5 * (arena1 - filtered1)) + windowed_scratch(one, 4096))
    assert rejected == [] and got == want
    # Not even one job's buffers without scratch: every job is rejected, in order.
    got, rejected = window_render(s, plans, (arena1 - filtered1) // 4)
    assert got == {} and rejected == list(range(len(plans)))
