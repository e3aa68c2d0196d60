"""Learned voice-tail predictor: features, labels and the model calls.

A voice chain (the spans of one filtered or pitch-modulated voice) retires
under the voice-silence policy at an exact frame the GPU judges while it
renders (clay_hip_batch_voice_tails). The predictor estimates that frame
before rendering, from 12 features of the plan:

 0 log2 playback step (pitch ratio)      6 resonance / 12 (dB)
 1 log10 voice gain                      7 envelope level at release
 2 log10 release time constant (s)       8 log10 peak of the playing source region
 3 sustain level                         9 log10 planned tail after release (s)
 4 loop flag                            10 log10 job gain (CC7 x CC11) at release
 5 log2 cutoff at release / 1 kHz       11 log10 release fade time (s)

Target: log1p((tail - release) / 64), in 64-frame chunks after the release.
Only released chains are used. The model is a 12-64-64-1 ReLU MLP trained and
evaluated by the library's HIP kernels (claysampler_amd/csrc/hip/mlp.hip).
"""

import ctypes as C

import numpy as np

from . import abi, hip

F = C.POINTER(C.c_float)


def _spans(plan):
    return np.ctypeslib.as_array(plan.spans, shape=(plan.span_count,)) if plan.span_count else None


def _track_value(plan, points, track, frames):
    """Control track value at frames (last point at or before, else the first)."""
    offsets = np.ctypeslib.as_array(plan.track_offsets, shape=(plan.track_count + 1,))
    pts = points[offsets[track]:offsets[track + 1]]
    i = np.clip(np.searchsorted(pts["frame"], frames, side="right") - 1, 0, len(pts) - 1)
    p = pts[i]
    return p["value"].astype(np.float64) + p["step"].astype(np.float64) * (frames - p["frame"])


def chain_records(plans, bank=None):
    """Per voice chain of the batch (chain id = filter tracks of the jobs in
    order): (features [n, 12] float32, release frame, chain end, usable mask)."""
    feats, releases, ends, usable = [], [], [], []
    peaks = {}
    for plan in plans:
        n = plan.filter_tracks
        if n <= 0:
            continue
        sp = _spans(plan)
        points = np.ctypeslib.as_array(plan.points, shape=(int(np.ctypeslib.as_array(
            plan.track_offsets, shape=(plan.track_count + 1,))[-1]),))
        mods = (np.ctypeslib.as_array(plan.modulation, shape=(plan.modulation_count,))
                if plan.modulation_count else None)
        track = sp["filter_track"]
        chained = np.nonzero(track >= 0)[0]
        head = np.full(n, -1, np.int64)
        last = np.full(n, -1, np.int64)
        end = np.zeros(n, np.int64)
        # Spans are in plan order; the last one of a chain carries its release.
        t = track[chained]
        np.maximum.at(end, t, sp["end"][chained])
        last[t] = chained                      # later spans overwrite earlier ones
        head[t[::-1]] = chained[::-1]          # earlier spans overwrite later ones
        ok = head >= 0
        h, l = sp[np.maximum(head, 0)], sp[np.maximum(last, 0)]
        release = l["release_start"].astype(np.int64)
        ok &= (release >= 0) & (release < end)
        rate = np.maximum(l["sample_rate"], 1).astype(np.float64)
        relf = np.clip(release, 0, None)
        # Cutoff at release: static cutoff x 2^(cents/1200) of the cutoff bindings;
        # every (chain, binding) pair at once, control tracks evaluated per track.
        cents = np.zeros(n)
        if mods is not None:
            counts = np.where(ok, l["modulation_count"], 0).astype(np.int64)
            chain_of = np.repeat(np.arange(n), counts)
            within = np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts)
            mod_of = np.repeat(l["first_modulation"].astype(np.int64), counts) + within
            sel = mods["target"][mod_of] == abi.MOD_CUTOFF
            chain_of, mod_of = chain_of[sel], mod_of[sel]
            value = mods["depth"][mod_of].astype(np.float64)
            for field in ("track", "shape_track"):
                tracks = mods[field][mod_of]
                for tr in np.unique(tracks[tracks >= 0]):
                    m = tracks == tr
                    value[m] *= _track_value(plan, points, int(tr), relf[chain_of[m]])
            np.add.at(cents, chain_of, value)
        filt = l["filters"][:, 0]
        cutoff = np.where(filt["type"] > 0, filt["cutoff"] * np.exp2(cents / 1200.0), 1000.0)
        tau = np.where(l["release_rate"] < 0, -1.0 / np.minimum(l["release_rate"], -1e-12), 1e7) / rate
        fade = np.where(l["fade_step"] > 0, l["fade_value"] / np.maximum(l["fade_step"], 1e-12), 1.0) / rate
        job_gain = _track_value(plan, points, 0, relf) * _track_value(plan, points, 1, relf)
        peak = np.ones(n)
        if bank is not None:
            looped = l["loop_end"] > 0
            region = np.stack([l["sample"].astype(np.int64), l["channels"].astype(np.int64),
                               np.where(looped, l["loop_first"], 0).astype(np.int64),
                               np.where(looped, l["loop_end"], l["source_end"]).astype(np.int64)], 1)
            keys, inverse = np.unique(region, axis=0, return_inverse=True)
            values = np.zeros(len(keys))
            for k, (a, chans, lo, hi) in enumerate(keys):
                key = (int(a), int(chans), int(lo), int(hi))
                if key not in peaks:
                    seg = bank[key[0] + key[2] * key[1]:key[0] + key[3] * key[1]]
                    peaks[key] = float(np.abs(seg.astype(np.int32)).max()) / 32768.0 if seg.size else 0.0
                values[k] = peaks[key]
            peak = values[inverse.reshape(-1)]
        x = np.stack([
            np.log2(np.maximum(h["step"], 1e-9)),
            np.log10(np.maximum(l["gain"], 1e-6)),
            np.log10(np.maximum(tau, 1e-6)),
            l["sustain"],
            (l["loop_end"] > 0).astype(np.float64),
            np.log2(np.maximum(cutoff, 1.0) / 1000.0),
            filt["resonance"] / 12.0,
            l["release_value"],
            np.log10(peak + 1e-6),
            np.log10((end - relf) / rate + 1e-3),
            np.log10(np.maximum(np.abs(job_gain), 1e-6)),
            np.log10(np.maximum(fade, 1e-6)),
        ], 1).astype(np.float32)
        feats.append(x)
        releases.append(release)
        ends.append(end)
        usable.append(ok)
    if not feats:
        return np.zeros((0, abi.MLP_INPUTS), np.float32), np.zeros(0, np.int64), np.zeros(0, np.int64), \
            np.zeros(0, bool)
    return np.concatenate(feats), np.concatenate(releases), np.concatenate(ends), np.concatenate(usable)


def to_target(tail, release):
    return np.log1p(np.maximum(tail - release, 0) / 64.0).astype(np.float32)


def from_target(y, release, end):
    return np.minimum(end, release + np.round(np.expm1(np.maximum(y, 0.0)) * 64.0).astype(np.int64))


def envelope_baseline(x, release, end, level_db):
    """The method without the network: frames until gain x job gain x source peak
    x release level x exp(-n / tau) falls below the level (filter ignored)."""
    amplitude = 10 ** x[:, 1] * 10 ** x[:, 10] * 10 ** x[:, 8] * np.maximum(x[:, 7], 1e-9)
    tau = 10 ** x[:, 2] * 44100.0
    level = 10 ** (level_db / 20)
    n = np.where(amplitude > level, tau * np.log(amplitude / level), 0.0)
    return np.minimum(end, release + np.ceil(n).astype(np.int64))


def init(x, y, seed=1):
    m = abi.TailModel()
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    status = hip().clay_mlp_init(C.byref(m), x.ctypes.data_as(F), y.ctypes.data_as(F), len(y), seed)
    assert status == abi.OK
    return m


def gradient(renderer, model, x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    g = np.zeros(abi.MLP_PARAMS, np.float32)
    loss = C.c_float()
    renderer._check(renderer.lib.clay_hip_mlp_gradient(renderer.handle, C.byref(model), x.ctypes.data_as(F),
                                                       y.ctypes.data_as(F), len(y), g.ctypes.data_as(F),
                                                       C.byref(loss)))
    return g, loss.value


def train(renderer, model, x, y, epochs=30, batch=1024, lr=3e-3, seed=7):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    cfg = abi.MlpTrainConfig(epochs, batch, lr, 0.9, 0.999, 1e-8, seed)
    losses = np.zeros(epochs, np.float32)
    renderer._check(renderer.lib.clay_hip_mlp_train(renderer.handle, C.byref(model), x.ctypes.data_as(F),
                                                    y.ctypes.data_as(F), len(y), C.byref(cfg),
                                                    losses.ctypes.data_as(F)))
    return losses


def predict(renderer, model, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(len(x), np.float32)
    renderer._check(renderer.lib.clay_hip_mlp_predict(renderer.handle, C.byref(model), x.ctypes.data_as(F), len(x),
                                                      out.ctypes.data_as(F)))
    return out


def params(model):
    return np.ctypeslib.as_array(model.params).astype(np.float64)


def scratch_for_coverage(pred, exact, release, end, channels=1, coverage=0.99):
    """Post-release scratch bytes (4 B per sample) when each chain is given
    (pred - release) x (1 + margin) frames after its release, capped by its
    plan, with the smallest margin that holds the exact tail of `coverage` of
    the chains (the rest would fall back to a full render). The planned
    alternative is sum(end - release). Returns (bytes, margin)."""
    need = np.maximum(exact - release, 1).astype(np.float64)
    have = np.maximum(pred - release, 1).astype(np.float64)
    ratio = np.sort(need / have)
    margin = max(0.0, float(ratio[min(len(ratio) - 1, int(np.ceil(coverage * len(ratio))) - 1)]) - 1.0)
    alloc = np.minimum(end - release, np.ceil(have * (1 + margin)).astype(np.int64))
    return float(4 * channels * alloc.sum()), margin


def scratch_bounds(pred, release, end, usable, margin, chunk=64):
    """Per-chain scratch bounds (absolute frames, batch chain order) for
    clay_hip_batch_prepare_bounded: a usable chain is backed up to its
    predicted tail plus `margin` (a fraction of the predicted post-release
    length), rounded up to whole 64-frame silence chunks plus one; any other
    chain, or a bound past the chain's end, keeps its whole span. `pred`,
    `release` and `end` are per chain (all chains, as chain_records gives)."""
    pred = np.asarray(pred, np.int64)
    have = np.maximum(pred - release, 1).astype(np.float64)
    bound = release + np.ceil(have * (1.0 + margin)).astype(np.int64)
    bound = (bound + chunk - 1) // chunk * chunk + chunk
    bound = np.where(usable, np.minimum(bound, end), end)
    return np.minimum(bound, np.iinfo(np.int32).max).astype(np.int32)


def render_bounded(renderer, plans, bounds):
    """Renders `plans` as one resident batch with per-chain scratch bounds,
    then renders the jobs that overflowed their bounds again without bounds.
    Returns (per-job (result, pcm) list, info) where info holds the bounded
    batch's filtered scratch bytes, the overflowed jobs and the seconds of
    both passes."""
    import time

    def per_job(results, pcm, channels):
        return [(r, pcm[r.offset * channels:(r.offset + r.data_frames) * channels].copy()) for r in results]

    t0 = time.perf_counter()
    b = renderer.prepare(plans, bounds)
    b.run()
    results, pcm = b.results()
    overflowed = [int(j) for j in b.overflows()]
    filtered, arena = b.scratch_bytes()
    b.close()
    jobs = per_job(results, pcm, renderer.channels)
    t1 = time.perf_counter()
    if overflowed:
        again = (abi.PlanView * len(overflowed))(*[plans[j] for j in overflowed])
        b = renderer.prepare(again)
        b.run()
        redo = per_job(*b.results(), renderer.channels)
        b.close()
        for k, j in enumerate(overflowed):
            jobs[j] = redo[k]
    t2 = time.perf_counter()
    return jobs, {"filtered_bytes": filtered, "arena_bytes": arena, "overflowed_jobs": overflowed,
                  "bounded_s": t1 - t0, "fallback_s": t2 - t1}
