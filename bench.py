"""Throughput of the HIP render core on one MI355X.

A step renders one batch of jobs completely on the device: every voice span
(resampling, envelopes, modulation, filters), the per-job mix, and the tail /
PCM16 finishing, with the plans, the int16 sample bank and the sinc table
already resident in HBM (no host transfers inside the timed region).

Default workload = configuration 3: 512 jobs in the uneven batched form on the
configuration-1 instrument (job lengths lognormal, clipped to 5-300 s,
polyphony 4-64), 16-tap sinc, natural tail policy, 44.1 kHz. A step renders
all 512 jobs resident on the device (one batch, plain form). The same line
reports the render windows of the plain and persistent forms at
--gpu-batch-jobs 128 (host transfers included) and whether their PCM16 is
identical, the roofline of the resample-mix kernel (clayMixStraight for batches of straight voices) on the
algorithmic bytes of its source windows, mix writes and span records, and
the FP64 oracle on the host cores as the CPU baseline. Prints one JSON line.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import claysampler_amd as ca  # noqa: E402
from claysampler_amd import abi, synth  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name, synth_dir, jobs, seconds):
    if name == "c2":
        sfz = synth.c2_instrument(os.path.join(synth_dir, "c2"))
        mids = [synth.c2_midi(os.path.join(synth_dir, "c2", f"bench_{j}_{int(seconds)}.mid"), 1000 + j, seconds)
                if not os.path.exists(os.path.join(synth_dir, "c2", f"bench_{j}_{int(seconds)}.mid"))
                else os.path.join(synth_dir, "c2", f"bench_{j}_{int(seconds)}.mid") for j in range(jobs)]
        return sfz, mids, 32
    if name == "c3":
        sfz, mids = synth.config3(os.path.join(synth_dir, "c3"), jobs)
        return sfz, mids, 16
    sfz, mids = synth.config1(os.path.join(synth_dir, "c1"))
    return sfz, mids * jobs, 16


def kernel_bytes(plans, taps, natural):
    """Algorithmic HBM bytes per launch of each render kernel.

    Source window bytes of a span = 2 B x channels x (|source travel| + taps);
    travel is step x frames (with the bend ramp factor for bent spans)."""
    src = {"dry": 0.0, "wet": 0.0}
    voice_frames = {"dry": 0, "wet": 0}
    wet_samples = 0
    spans = 0
    frames = 0
    span_bytes = 224  # device span record (the hot part the mix reads per voice)
    for p in plans:
        frames += p.frames
        spans += p.span_count
        v = np.ctypeslib.as_array(p.spans, shape=(p.span_count,))
        n = (v["end"] - v["begin"]).astype(np.float64)
        slope = v["pitch_slope"]
        bent = slope != 0
        travel = np.abs(v["step"]) * n
        travel[bent] = np.abs(v["step"][bent] * np.expm1(slope[bent] * n[bent]) / np.expm1(slope[bent]))
        sampled = v["generator"] == 0
        window = np.where(sampled, np.minimum(travel, v["source_end"]) + taps, 0.0) * 2 * v["channels"]
        wet = v["filter_track"] >= 0
        src["wet"] += float(window[wet].sum())
        src["dry"] += float(window[~wet].sum())
        voice_frames["wet"] += int(n[wet].sum())
        voice_frames["dry"] += int(n[~wet].sum())
        wet_samples += int((n * v["channels"])[wet].sum())
    mix_bytes = frames * (8 if natural else 4)
    return {
        # source windows of wet spans + one float write per wet sample
        "clayChainAudio": src["wet"] + 4 * wet_samples,
        # one float read and one float write per wet sample
        "clayVoiceChains": 8 * wet_samples,
        # dry source windows + wet reads + the mix write + span records
        "clayMixTiles": src["dry"] + 4 * wet_samples + mix_bytes + spans * span_bytes,
        "clayMixWaves": src["dry"] + 4 * wet_samples + mix_bytes + spans * span_bytes,
        # straight batches: dry source windows + the mix write + span records,
        # and (natural policy, the batch in one window) the unit-scale mono
        # PCM16 written beside the float mix
        "clayMixStraight": src["dry"] + mix_bytes + spans * span_bytes + (2 * frames if natural else 0),
    }, voice_frames, frames


def measured_traffic(config, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    this workload (profiles/traffic_<config>.txt: FETCH_SIZE and WRITE_SIZE
    passes scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{config}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if parts and parts[0] == kernel and not line.startswith("#"):
            return float(parts[-1])
    return None


def cpu_baseline(sfz, mids, taps, jobs, threads, voice_frame_budget=None):
    """FP64 oracle on a bounded sample of the same workload, on host cores:
    the first `jobs` jobs, or as many leading jobs as stay within
    `voice_frame_budget` voice-frames (about 10-30 s of CPU work)."""
    import oracle

    s = ca.Session()
    if voice_frame_budget is not None:
        total = 0
        for m in mids:
            index = s.add_job(sfz, m)
            v = np.ctypeslib.as_array(s.plan(index).spans, shape=(s.plan(index).span_count,))
            total += int((v["end"] - v["begin"]).astype(np.int64).sum())
            if total >= voice_frame_budget:
                break
        jobs = s.lib.clay_session_job_count(s.handle)
    else:
        for m in mids[:jobs]:
            s.add_job(sfz, m)
    plans = s.plans()
    threads = min(threads, len(plans))  # one job per thread at most
    bank_ptr, bank_count = s.bank()
    cfg = oracle.config(taps=taps, threads=threads)
    t0 = time.perf_counter()
    oracle.render(cfg, bank_ptr, bank_count, plans)
    dt = time.perf_counter() - t0
    audio = sum(p.frames for p in plans) / 44100.0
    return {"value": audio / dt, "unit": "audio-s/s", "cores": threads, "kind": "port",
            "sample": f"{jobs} of the workload's jobs ({audio:.1f} audio-s) through the FP64 oracle, "
                      f"{threads} threads, {dt:.1f} s"}


def window_pass(policy, taps, form, batch_jobs, plans, bank, checksums=False, flags=0, statuses=None, warm=None,
                **renderer):
    """clay_hip_render_window over every plan on one renderer: an untimed
    window over `warm` first (graph capture, pinned staging and PCM slots),
    then the timed window, then, with `checksums`, an untimed window that
    checksums every job's PCM16. Returns (wall s of the timed window, per-job
    checksums or None, stats of the timed window). The timed window's publish
    callback only releases the PCM, so the time is the render core's. Extra
    keyword arguments go to ca.Renderer; `statuses`, a dict, receives each
    job's tail status."""
    import ctypes as C

    r = ca.Renderer(policy=policy, taps=taps, form=form, batch_jobs=batch_jobs, flags=flags, **renderer)
    r.upload_bank(*bank)

    def run(plan_array, sums, statuses):
        published = [0]

        def publish(user, batch_ptr, first):
            b = batch_ptr.contents
            if sums is not None:
                for k in range(b.result_count):
                    res = b.results[k]
                    n = res.data_frames * r.channels
                    view = np.ctypeslib.as_array(C.cast(C.addressof(b.pcm.contents) + 2 * res.offset * r.channels,
                                                        C.POINTER(C.c_int16)), shape=(n,)) if n else np.zeros(0, np.int16)
                    sums[first + k] = (res.status, res.frames, res.data_frames, int(view.astype(np.int64).sum()),
                                       int((view.astype(np.int64) * np.arange(n, dtype=np.int64) % 65521).sum()))
            if statuses is not None:
                for k in range(b.result_count):
                    statuses[first + k] = abi.TAIL_STATUS[b.results[k].status]
            published[0] += max(1, b.result_count)
            if b.pcm_handle:
                r.lib.clay_hip_release_pcm(r.handle, b.pcm_handle)
            return 0

        cb = abi.PUBLISH_FN(publish)
        t0 = time.perf_counter()
        status = r.lib.clay_hip_render_window(r.handle, plan_array, len(plan_array), cb, abi.REFILL_FN(), None)
        wall = time.perf_counter() - t0
        r._check(status)
        assert published[0] == len(plan_array)
        return wall

    if warm is not None:
        run(warm, None, None)
    before = r.stats()
    wall = run(plans, None, statuses)
    stats = r.stats()
    for name in ("rounds", "admissions", "segments", "launches", "graph_launches", "downloaded_bytes"):
        setattr(stats, name, getattr(stats, name) - getattr(before, name))
    sums = None
    if checksums:
        sums = {}
        run(plans, sums, None)
    r.close()
    return wall, sums, stats


def c5_main(args):
    """Configuration 5: the dense stress workload (96 kHz, 64 taps, 2 GB bank,
    volume groups of 8) through render_window, persistent form; with --forms
    also the plain form and the PCM16 identity check."""
    jobs = args.jobs or 32
    seconds = args.seconds if args.seconds is not None else 240.0
    t0 = time.perf_counter()
    sfz, mids = synth.config5(os.path.join(args.synth_dir, "c5"), jobs, seconds)
    log(f"inputs ready in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    rate = 96000
    session = ca.Session(rate=rate, measure=True)
    for j, m in enumerate(mids[:jobs]):
        session.add_job(sfz, m, group=j // 8 + 1)
    plans = session.plans()
    bank = session.bank()
    frames = sum(p.frames for p in plans)
    voice_frames = 0
    for p in plans:
        v = np.ctypeslib.as_array(p.spans, shape=(p.span_count,))
        voice_frames += int((v["end"] - v["begin"]).astype(np.int64).sum())
    log(f"planned {jobs} jobs, {sum(p.span_count for p in plans)} spans, {bank[1] * 2 / 1e9:.2f} GB bank "
        f"in {time.perf_counter() - t0:.1f} s")
    kw = dict(rate=rate, group_minimum=8)
    warm = (abi.PlanView * min(8, len(plans)))(*plans[:8])
    statuses = {}
    rolled, got, stats = window_pass(abi.POLICY_NATURAL, 64, abi.FORM_PERSISTENT, args.batch_jobs, plans, bank,
                                     checksums=args.forms, statuses=statuses, warm=warm, **kw)
    log(f"persistent: {rolled:.3f} s, {stats.rounds} rounds, segment {stats.segment_frames}")
    line = {
        "metric": "rendered audio-sec/sec and voice-samples/sec on 1 GPU; % of HBM roofline",
        "value": frames / rate / rolled,
        "unit": "audio-s/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": rolled * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"config 5: {jobs} jobs x {seconds:g} s at 96 kHz, 64-tap sinc, CC64 pedal, 4 s "
                               f"releases, {bank[1] * 2 / 1e9:.2f} GB bank, volume groups of 8, persistent form, "
                               "whole window incl. finishing, group matching and PCM download",
                   "jobs": jobs, "sinc_taps": 64, "tail_policy": "natural", "rate": rate, "form": "persistent"},
        "audio_seconds": frames / rate,
        "voice_samples_per_s": voice_frames / rolled,
        "mean_voices": voice_frames / max(1, frames),
        "statuses": sorted(set(statuses.values())),
        "persistent": {"rounds": stats.rounds, "segment_frames": stats.segment_frames,
                       "mix_blocks": stats.mix_blocks, "graph_launches": stats.graph_launches},
    }
    if args.forms:
        plain, want, _ = window_pass(abi.POLICY_NATURAL, 64, abi.FORM_PLAIN, args.batch_jobs, plans, bank,
                                     checksums=True, warm=warm, **kw)
        line["plain_value"] = frames / rate / plain
        line["identical_pcm16"] = want == got
        log(f"plain: {plain:.3f} s, identical PCM16: {want == got}")
    print(json.dumps(line), flush=True)


def forms_main(args, jobs, sfz, mids, taps, plans, bank):
    """Configuration 3: the plain per-segment form vs the persistent rolling
    form on the same jobs, whole-window wall time, identical PCM16 required."""
    frames = sum(p.frames for p in plans)
    audio_seconds = frames / 44100.0
    warm = (abi.PlanView * len(plans))(*plans)
    plain, want, plain_stats = window_pass(abi.POLICY_NATURAL, taps, abi.FORM_PLAIN, args.batch_jobs, plans, bank,
                                           checksums=True, warm=warm)
    log(f"plain: {plain:.3f} s, {plain_stats.segments} segments")
    rolled, got, stats = window_pass(abi.POLICY_NATURAL, taps, abi.FORM_PERSISTENT, args.batch_jobs, plans, bank,
                                     checksums=True, warm=warm)
    log(f"persistent: {rolled:.3f} s, {stats.rounds} rounds, segment {stats.segment_frames}, "
        f"mix blocks {stats.mix_blocks}, graph launches {stats.graph_launches}")
    identical = want == got
    log(f"identical PCM16: {identical}")
    line = {
        "metric": "rendered audio-sec/sec and voice-samples/sec on 1 GPU; % of HBM roofline",
        "value": audio_seconds / rolled,
        "unit": "audio-s/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": rolled * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"config 3: {jobs} jobs, lognormal 5-300 s, polyphony 4-64, 16-tap sinc, "
                               f"--gpu-batch-jobs {args.batch_jobs}, whole window incl. host planning handoff, "
                               "finishing and PCM download",
                   "jobs": jobs, "sinc_taps": taps, "tail_policy": "natural", "rate": 44100,
                   "form": "persistent vs plain"},
        "audio_seconds": audio_seconds,
        "plain_value": audio_seconds / plain,
        "speedup_vs_plain": plain / rolled,
        "identical_pcm16": identical,
        "persistent": {"rounds": stats.rounds, "segment_frames": stats.segment_frames,
                       "mix_blocks": stats.mix_blocks, "graph_launches": stats.graph_launches},
    }
    print(json.dumps(line), flush=True)


def predictor_main(args, taps=32, db=-90.0):
    """Configuration 4: the learned voice-tail predictor on configuration 2's
    workload under the -90 dBFS voice-silence policy. Exact tails from the
    plain renderer (the chain kernel's judge) for args.tail_batches batches of
    64 jobs x 60 s, GPU training, then on a held-out batch: frame error against
    the analytic envelope estimate, post-release scratch for 99 % coverage,
    throughput with and without voice retirement, and plain vs persistent
    PCM16 identity with the policy on."""
    from claysampler_amd import tails

    sfz = synth.c2_instrument(os.path.join(args.synth_dir, "c2"))

    def batch(seeds):
        s = ca.Session()
        for seed in seeds:
            path = os.path.join(args.synth_dir, "c2", f"c4_{seed}.mid")
            if not os.path.exists(path):
                synth.c2_midi(path, seed, 60.0)
            s.add_job(sfz, path)
        return s

    def exact_tails(s, plans):
        r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps, voice_silence_db=db)
        r.upload_bank(*s.bank())
        b = r.prepare(plans)
        b.run()
        out = b.voice_tails().astype(np.int64)
        b.close()
        r.close()
        return out

    xs, ys = [], []
    t0 = time.perf_counter()
    for k in range(args.tail_batches):
        s = batch(range(100000 + 64 * k, 100000 + 64 * (k + 1)))
        plans = s.plans()
        exact = exact_tails(s, plans)
        x, rel, end, ok = tails.chain_records(plans, s.bank_array())
        xs.append(x[ok])
        ys.append(tails.to_target(exact[ok], rel[ok]))
        validation = (x[ok], rel[ok], end[ok], exact[ok])  # the last batch sets the scratch margin
        log(f"tail batch {k + 1}/{args.tail_batches}: {ok.sum()} voices, {time.perf_counter() - t0:.1f} s")
    x, y = np.concatenate(xs[:-1]), np.concatenate(ys[:-1])
    renderer = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps, voice_silence_db=db)
    model = tails.init(x, y, seed=1)
    t0 = time.perf_counter()
    losses = tails.train(renderer, model, x, y, epochs=args.epochs, batch=4096, lr=2e-3, seed=2)
    train_s = time.perf_counter() - t0
    log(f"trained on {len(y)} tails in {train_s:.1f} s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    held = batch(range(900000, 900064))
    plans = held.plans()
    bank = held.bank()
    exact = exact_tails(held, plans)
    xt, rel, end, ok = tails.chain_records(plans, held.bank_array())
    xt, rel, end, exact_ok = xt[ok], rel[ok], end[ok], exact[ok]
    pred = tails.from_target(tails.predict(renderer, model, xt), rel, end)
    base = tails.envelope_baseline(xt, rel, end, db)
    truth = tails.to_target(exact_ok, rel)
    frames_err = {name: float(np.median(np.abs(p - exact_ok))) for name, p in (("network", pred), ("envelope", base))}
    log_err = {name: float(np.abs(tails.to_target(p, rel) - truth).mean()) for name, p in (("network", pred),
                                                                                          ("envelope", base))}
    planned = 4.0 * float((end - rel).sum())
    scratch = {"planned": planned}
    for name, p in (("network", pred), ("envelope", base)):
        scratch[name], margin = tails.scratch_for_coverage(p, exact_ok, rel, end)
        scratch[name + "_margin"] = margin
    # Scratch sized from the predictions on the held-out batch. A job falls
    # back when any of its ~500 chains overflows, so the margin holds 99.99 %
    # of the validation batch's chain tails; jobs that overflow are rendered
    # again unbounded, and every job must equal the unbounded render.
    xv, relv, endv, exactv = validation
    _, margin = tails.scratch_for_coverage(tails.from_target(tails.predict(renderer, model, xv), relv, endv),
                                           exactv, relv, endv, coverage=0.9999)
    _, margin_env = tails.scratch_for_coverage(tails.envelope_baseline(xv, relv, endv, db), exactv, relv, endv,
                                               coverage=0.9999)
    xa, rela, enda, oka = tails.chain_records(plans, held.bank_array())
    pred_all = tails.from_target(tails.predict(renderer, model, xa), rela, enda)
    env_all = tails.envelope_baseline(xa, rela, enda, db)
    realised = {}
    r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps, voice_silence_db=db)
    r.upload_bank(*bank)
    b = r.prepare(plans)
    b.run()
    want_res, want_pcm = b.results()
    want = [(q.status, q.frames, q.data_frames, want_pcm[q.offset:q.offset + q.data_frames].tobytes()) for q in want_res]
    full_filtered, full_arena = b.scratch_bytes()
    b.close()
    # The same path (prepare, run, results) with every chain's bound at its end, for the time.
    tails.render_bounded(r, plans, enda.astype(np.int32))
    _, info = tails.render_bounded(r, plans, enda.astype(np.int32))
    realised["unbounded"] = {"filtered_bytes": full_filtered, "arena_bytes": full_arena,
                             "seconds": info["bounded_s"]}
    for name, p, m in (("network", pred_all, margin), ("envelope", env_all, margin_env)):
        bounds = tails.scratch_bounds(p, rela, enda, oka, m)
        tails.render_bounded(r, plans, bounds)  # warm
        jobs, info = tails.render_bounded(r, plans, bounds)
        got = [(q.status, q.frames, q.data_frames, pcm.tobytes()) for q, pcm in jobs]
        realised[name] = {"margin": m, "filtered_bytes": info["filtered_bytes"], "arena_bytes": info["arena_bytes"],
                          "overflowed_jobs": len(info["overflowed_jobs"]),
                          "seconds": info["bounded_s"] + info["fallback_s"], "fallback_s": info["fallback_s"],
                          "identical_pcm16": got == want}
        log(f"bounded scratch ({name}): {info['filtered_bytes'] / 1e9:.3f} GB vs {full_filtered / 1e9:.3f} GB, "
            f"{len(info['overflowed_jobs'])} jobs rendered again, identical {got == want}")
    r.close()
    renderer.close()
    # Throughput on the held-out batch, policy off and on (plain form, resident batch).
    rates = {}
    frames = sum(p.frames for p in plans)
    for name, level in (("all_frames", 0.0), ("voice_retirement", db)):
        r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps, voice_silence_db=level)
        r.upload_bank(*bank)
        b = r.prepare(plans)
        b.run()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run()
        rates[name] = frames / 44100.0 / ((time.perf_counter() - t0) / args.steps)
        b.close()
        r.close()
    # Plain vs persistent with the policy on: identical PCM16 per job.
    sums = []
    for form, flags in ((abi.FORM_PLAIN, 0), (abi.FORM_PERSISTENT, 0)):
        r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps, form=form, batch_jobs=16, flags=flags,
                        voice_silence_db=db)
        r.upload_bank(*bank)
        got = {}
        r.render_window(plans, lambda res, pcm, first, failed: got.update(
            {first + k: (x.status, x.frames, x.data_frames, int(pcm[x.offset:x.offset + x.data_frames].astype(
                np.int64).sum())) for k, x in enumerate(res)}))
        r.close()
        sums.append(got)
    line = {
        "metric": "rendered audio-sec/sec and voice-samples/sec on 1 GPU; % of HBM roofline",
        "value": rates["voice_retirement"],
        "unit": "audio-s/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": 1e3 * frames / 44100.0 / rates["voice_retirement"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "config 4: config-2 workload (64 jobs x 60 s, 32 taps) with the -90 dBFS voice-silence "
                               "policy; learned tail predictor trained on exact tails of the plain renderer",
                   "training_tails": int(len(y)), "epochs": args.epochs, "held_out_voices": int(len(truth))},
        "predictor": {"train_seconds": train_s, "loss_first_epoch": float(losses[0]),
                      "loss_last_epoch": float(losses[-1]), "median_abs_frame_error": frames_err,
                      "mean_abs_log_error": log_err, "post_release_scratch_bytes_99pct": scratch,
                      "realised_scratch": realised},
        "throughput_audio_s_per_s": rates,
        "identical_pcm16_plain_vs_persistent": sums[0] == sums[1],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--seconds", type=float, default=None, help="job length (configuration 2: 60, 5: 240)")
    ap.add_argument("--synth-dir", default=os.environ.get("CLAY_SYNTH_DIR", "/tmp/clay_synth"))
    ap.add_argument("--cpu-jobs", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-forms", action="store_true", help="configuration 3: skip the render-window comparison")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", action="store_true",
                    help="plain vs persistent form through render_window (configuration 3)")
    ap.add_argument("--batch-jobs", type=int, default=128)
    ap.add_argument("--launch-sizes", action="store_true",
                    help="configuration 3: mix time of the jobs split into resident batches of 16 to 512 jobs")
    ap.add_argument("--tail-batches", type=int, default=32, help="configuration 4: training batches of 64 jobs")
    ap.add_argument("--epochs", type=int, default=20, help="configuration 4: training epochs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the renderer has no multi-GPU path (jobs would shard by job, one process per GPU)",
              file=sys.stderr)
        sys.exit(2)
    if args.config == "c4":
        predictor_main(args)
        return
    if args.config == "c5":
        c5_main(args)
        return
    if args.seconds is None:
        args.seconds = 60.0
    default_jobs = {"c1": 1, "c2": 64, "c3": 512}[args.config]
    jobs = args.jobs or default_jobs

    t0 = time.perf_counter()
    sfz, mids, taps = workload(args.config, args.synth_dir, jobs, args.seconds)
    log(f"inputs ready in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    session = ca.Session()
    for m in mids[:jobs]:
        session.add_job(sfz, m)
    plans = session.plans()
    bank_ptr, bank_count = session.bank()
    log(f"planned {jobs} jobs, {sum(p.span_count for p in plans)} spans in {time.perf_counter() - t0:.1f} s")
    if args.forms:
        forms_main(args, jobs, sfz, mids, taps, plans, (bank_ptr, bank_count))
        return
    if args.launch_sizes:
        launch_sizes(plans, (bank_ptr, bank_count), taps)
        return

    renderer = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps)
    renderer.upload_bank(bank_ptr, bank_count)
    batch = renderer.prepare(plans)
    for _ in range(args.warmup):
        batch.run()
    # The launch tuner's trial runs (each candidate launch shape timed twice on
    # the audio and filter passes), untimed; a batch without voice chains has
    # nothing to tune and ends them at once.
    tuning_runs = 0
    while renderer.stats().launch_tuning and tuning_runs < 16 and any(p.filter_tracks for p in plans):
        batch.run()
        tuning_runs += 1
    log("warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"{args.steps} steps: {ms_per_step:.2f} ms/step")

    # Per-kernel launch times with HIP events on the render stream (separate, untimed run).
    os.environ["CLAY_KERNEL_TIMING"] = "1"
    batch.run()
    os.environ["CLAY_KERNEL_TIMING"] = "0"
    bytes_per, voice_frames, frames = kernel_bytes(plans, taps, True)
    kernels = {}
    for name in ("clayChainAudio", "clayVoiceChains", "clayMixTiles", "clayMixWaves", "clayMixStraight",
                 "clayScanStereo", "clayQuantizeMono", "clayFinalizePcm", "clayPitchChains", "claySelectNative",
                 "claySelectTrim"):
        ms, n = batch.kernel_time(name)
        if n:
            kernels[name] = {"ms": ms, "launches": n}
    # The roofline goes on the resample-mix kernel where it renders the voices
    # (configurations 1 and 3), else on the kernel that takes longest.
    mix_kernel = next((k for k in ("clayMixStraight", "clayMixWaves") if k in kernels), None)
    dominant = mix_kernel or max((k for k in kernels if k in bytes_per), key=lambda k: kernels[k]["ms"])
    avg_s = kernels[dominant]["ms"] / kernels[dominant]["launches"] / 1e3
    achieved = bytes_per[dominant] / kernels[dominant]["launches"] / avg_s / 1e9
    results, _ = batch.results()
    statuses = sorted({abi.TAIL_STATUS[r.status] for r in results})
    tuned = renderer.stats()

    audio_seconds = frames / 44100.0
    total_voice_frames = voice_frames["dry"] + voice_frames["wet"]
    full_size = jobs == default_jobs
    workload_name = {"c1": "config 1: 1 job, 32 notes, 16-tap sinc",
                     "c2": f"config 2: {jobs} jobs x {args.seconds:g} s, 32-tap sinc, lpf_2p+CC1, bends",
                     "c3": f"config 3: {jobs} jobs, lognormal 5-300 s, polyphony 4-64, 16-tap sinc"}[args.config]
    line = {
        "metric": "rendered audio-sec/sec and voice-samples/sec on 1 GPU; % of HBM roofline",
        "value": audio_seconds / (ms_per_step / 1e3),
        "unit": "audio-s/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload_name, "jobs": jobs, "sinc_taps": taps, "tail_policy": "natural",
                   "rate": 44100, "step": "every job rendered and finished resident on the device, plain form"},
        "voice_samples_per_s": total_voice_frames / (ms_per_step / 1e3),
        "audio_seconds": audio_seconds,
        "kernels_ms": {k: round(v["ms"], 3) for k, v in kernels.items()},
        "statuses": statuses,
        "launch_shapes": {"chain_waves": tuned.chain_waves, "audio_unroll": tuned.audio_unroll,
                          "tuning_runs": tuning_runs},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": bytes_per[dominant],
                     "traffic": measured_traffic(args.config, dominant) if full_size else None},
    }
    valu = measured_valu(args.config, dominant) if full_size else None
    if valu:
        line["roofline"]["valu"] = valu
    batch.close()
    renderer.close()
    if args.config == "c3" and not args.no_forms:
        line["forms"] = forms_compare(args, plans, (bank_ptr, bank_count), taps, frames)
    if not args.no_cpu:
        threads = max(1, args.cpu_threads)
        budget = None if args.config != "c3" else threads * 8e7  # about 15 s of FP64 oracle work
        line["cpu_baseline"] = cpu_baseline(sfz, mids, taps, min(args.cpu_jobs, jobs), min(threads, jobs), budget)
    print(json.dumps(line), flush=True)


def measured_valu(config, kernel):
    """VALU figures of `kernel` from the committed SQ counter summary of this
    workload (profiles/valu_<config>.txt: `key value` lines), or None."""
    path = os.path.join(ROOT, "profiles", f"valu_{config}.txt")
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == kernel and not line.startswith("#"):
            out[parts[1]] = float(parts[2])
    return out or None


def launch_sizes(plans, bank, taps):
    """Mix kernel time of the jobs rendered as resident batches of 16 to all
    jobs (HIP events on the render stream, second run of each batch): what a
    launch costs beyond its tiles."""
    r = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps)
    r.upload_bank(*bank)
    os.environ["CLAY_KERNEL_TIMING"] = "1"
    rows = []
    n = 16
    while n <= len(plans):
        total, launches = 0.0, 0
        for first in range(0, len(plans), n):
            part = plans[first:first + n]
            b = r.prepare((abi.PlanView * len(part))(*part))
            b.run()
            b.run()
            for name in ("clayMixStraight", "clayMixWaves", "clayMixTiles"):
                ms, k = b.kernel_time(name)
                total += ms
                launches += k
            b.close()
        rows.append({"batch_jobs": n, "launches": launches, "mix_ms": total})
        log(f"batches of {n:3d}: mix {total:.1f} ms over {launches} launches")
        n *= 2
    os.environ["CLAY_KERNEL_TIMING"] = "0"
    r.close()
    print(json.dumps({"launch_sizes": rows}), flush=True)


def forms_compare(args, plans, bank, taps, frames):
    """The plain and persistent forms through render_window on the same jobs
    (--gpu-batch-jobs as given), wall time each including host transfers,
    and per-job PCM16 checksums compared."""
    warm = (abi.PlanView * len(plans))(*plans)
    plain, want, plain_stats = window_pass(abi.POLICY_NATURAL, taps, abi.FORM_PLAIN, args.batch_jobs, plans, bank,
                                           checksums=True, warm=warm)
    rolled, got, stats = window_pass(abi.POLICY_NATURAL, taps, abi.FORM_PERSISTENT, args.batch_jobs, plans, bank,
                                     checksums=True, warm=warm)
    log(f"windows: plain {plain:.3f} s, persistent {rolled:.3f} s, identical PCM16 {want == got}")
    audio_seconds = frames / 44100.0
    return {"batch_jobs": args.batch_jobs, "plain_window_s": plain, "persistent_window_s": rolled,
            "plain_audio_s_per_s": audio_seconds / plain, "persistent_audio_s_per_s": audio_seconds / rolled,
            "persistent_vs_plain": plain / rolled, "identical_pcm16": want == got,
            "persistent": {"rounds": stats.rounds, "segment_frames": stats.segment_frames,
                           "graph_launches": stats.graph_launches},
            "note": "render windows include host planning handoff, bank-resident uploads of the plans and PCM downloads"}


if __name__ == "__main__":
    main()
