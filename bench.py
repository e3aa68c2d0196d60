"""Throughput of the HIP render core on one MI355X.

A step renders one batch of jobs completely on the device: every voice span
(resampling, envelopes, modulation, filters), the per-job mix, and the tail /
PCM16 finishing, with the plans, the int16 sample bank and the sinc table
already resident in HBM (no host transfers inside the timed region).

Default workload = configuration 2: 64 jobs x 60 s MIDI on the 176-region
synthetic instrument (lpf_2p + CC1, pitch-bend walk, loops), 32-tap sinc,
natural tail policy, 44.1 kHz. Prints one JSON line.
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
    span_bytes = 208  # device span record (hot part read per voice)
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
        "clayVoiceChains": src["wet"] + 4 * wet_samples,
        # dry source windows + wet reads + the mix write + span records
        "clayMixTiles": src["dry"] + 4 * wet_samples + mix_bytes + spans * span_bytes,
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


def cpu_baseline(sfz, mids, taps, jobs, threads):
    """FP64 oracle on a bounded sample of the same workload, on host cores."""
    import oracle

    s = ca.Session()
    for m in mids[:jobs]:
        s.add_job(sfz, m)
    plans = s.plans()
    bank_ptr, bank_count = s.bank()
    cfg = oracle.config(taps=taps, threads=threads)
    t0 = time.perf_counter()
    oracle.render(cfg, bank_ptr, bank_count, plans)
    dt = time.perf_counter() - t0
    audio = sum(p.frames for p in plans) / 44100.0
    return {"value": audio / dt, "unit": "audio-s/s", "cores": threads, "kind": "port",
            "sample": f"{jobs} of the workload's jobs ({audio:.1f} audio-s) through the FP64 oracle, "
                      f"{threads} threads, {dt:.1f} s"}


def window_pass(policy, taps, form, batch_jobs, plans, bank, checksums=False, flags=0):
    """One clay_hip_render_window over every plan. Returns (wall s, per-job
    checksums or None, stats). The publish callback only releases the PCM
    (and optionally checksums it), so the time is the render core's."""
    import ctypes as C

    r = ca.Renderer(policy=policy, taps=taps, form=form, batch_jobs=batch_jobs, flags=flags)
    r.upload_bank(*bank)
    sums = {} if checksums else None
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
        published[0] += max(1, b.result_count)
        if b.pcm_handle:
            r.lib.clay_hip_release_pcm(r.handle, b.pcm_handle)
        return 0

    cb = abi.PUBLISH_FN(publish)
    t0 = time.perf_counter()
    status = r.lib.clay_hip_render_window(r.handle, plans, len(plans), cb, abi.REFILL_FN(), None)
    wall = time.perf_counter() - t0
    r._check(status)
    stats = r.stats()
    r.close()
    assert published[0] == len(plans)
    return wall, sums, stats


def forms_main(args, jobs, sfz, mids, taps, plans, bank):
    """Configuration 3: the plain per-segment form vs the persistent rolling
    form on the same jobs, whole-window wall time, identical PCM16 required."""
    frames = sum(p.frames for p in plans)
    audio_seconds = frames / 44100.0
    warm = (abi.PlanView * min(8, len(plans)))(*plans[:8])
    for form in (abi.FORM_PLAIN, abi.FORM_PERSISTENT):
        window_pass(abi.POLICY_NATURAL, taps, form, args.batch_jobs, warm, bank)
    log("warmup done")
    plain, _, plain_stats = window_pass(abi.POLICY_NATURAL, taps, abi.FORM_PLAIN, args.batch_jobs, plans, bank)
    log(f"plain: {plain:.3f} s, {plain_stats.segments} segments")
    rolled, _, stats = window_pass(abi.POLICY_NATURAL, taps, abi.FORM_PERSISTENT, args.batch_jobs, plans, bank)
    log(f"persistent: {rolled:.3f} s, {stats.rounds} rounds, segment {stats.segment_frames}, "
        f"mix blocks {stats.mix_blocks}, graph launches {stats.graph_launches}")
    _, want, _ = window_pass(abi.POLICY_NATURAL, taps, abi.FORM_PLAIN, args.batch_jobs, plans, bank, True)
    _, got, _ = window_pass(abi.POLICY_NATURAL, taps, abi.FORM_PERSISTENT, args.batch_jobs, plans, bank, True)
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


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3"])
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--seconds", type=float, default=60.0)
    ap.add_argument("--synth-dir", default=os.environ.get("CLAY_SYNTH_DIR", "/tmp/clay_synth"))
    ap.add_argument("--cpu-jobs", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", action="store_true",
                    help="plain vs persistent form through render_window (configuration 3)")
    ap.add_argument("--batch-jobs", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (jobs shard by job; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
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

    renderer = ca.Renderer(policy=abi.POLICY_NATURAL, taps=taps)
    renderer.upload_bank(bank_ptr, bank_count)
    batch = renderer.prepare(plans)
    for _ in range(args.warmup):
        batch.run()
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
    for name in ("clayVoiceChains", "clayMixTiles", "clayScanStereo", "clayQuantizeMono", "clayFinalizePcm",
                 "clayPitchChains", "claySelectNative", "claySelectTrim"):
        ms, n = batch.kernel_time(name)
        if n:
            kernels[name] = {"ms": ms, "launches": n}
    dominant = max((k for k in kernels if k in bytes_per), key=lambda k: kernels[k]["ms"])
    avg_s = kernels[dominant]["ms"] / kernels[dominant]["launches"] / 1e3
    achieved = bytes_per[dominant] / kernels[dominant]["launches"] / avg_s / 1e9
    results, _ = batch.results()
    statuses = sorted({abi.TAIL_STATUS[r.status] for r in results})

    audio_seconds = frames / 44100.0
    total_voice_frames = voice_frames["dry"] + voice_frames["wet"]
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
        "config": {"workload": {"c1": "config 1: 1 job, 32 notes, 16-tap sinc",
                                "c2": f"config 2: {jobs} jobs x {args.seconds:g} s, 32-tap sinc, lpf_2p+CC1, bends",
                                "c3": f"config 3 (plain form): {jobs} jobs, 16-tap sinc"}[args.config],
                   "jobs": jobs, "sinc_taps": taps, "tail_policy": "natural", "rate": 44100, "form": "plain"},
        "voice_samples_per_s": total_voice_frames / (ms_per_step / 1e3),
        "audio_seconds": audio_seconds,
        "kernels_ms": {k: round(v["ms"], 3) for k, v in kernels.items()},
        "statuses": statuses,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.config, dominant) if jobs == default_jobs else None},
    }
    batch.close()
    renderer.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(sfz, mids, taps, min(args.cpu_jobs, jobs), min(args.cpu_jobs, jobs))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
