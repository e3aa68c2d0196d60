"""The documented C++ integration facade (the reference's `clay::Renderer`
over this core's C ABI, read from the repository's integration notes) compiles against the reference's own headers
(src/cuda/renderer.h, src/cuda/render_types.h, src/render/plan.h) and converts
every record field by field.

The check program sets every field of hand-built reference records to a
distinct value and compares each converted field with its source. The test
then requires the set of compared fields to be every field of the ABI record
(from the ctypes mirror in `claysampler_amd.abi`) and every field of the
reference record (parsed from render_types.h) except the planner- and
renderer-internal ones that have no ABI field. A field the facade drops, or
two fields it swaps, fails. Headers only: nothing of the reference is built
or linked. Skipped where the reference tree is absent."""

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from claysampler_amd import abi

REPO = Path(__file__).resolve().parents[1]
REFERENCE = Path("/root/reference/src")

pytestmark = pytest.mark.skipif(not (REFERENCE / "cuda" / "renderer.h").exists(),
                                reason="reference headers not present")

# VoiceSpan fields with no ABI counterpart: the planner's LFO range
# (plan.cpp:343,541,585) and the reference renderer's own links, offsets and
# control slots, which this core derives itself.
INTERNAL = {"firstLfo", "lfoCount", "nextFilteredSpan", "filteredOffset", "phaseOffset", "filterControls"}

CHECK_MAIN = r"""
#include "facade.cpp"

#include <cstdio>
#include <type_traits>

using namespace clay;

static int failures = 0;
static double counter = 100;

template<class T> void set(T& x) {
    counter += 1;
    if constexpr (std::is_enum_v<T>) x = T(int(counter) % 4);
    else x = T(counter);
}
template<class T> double num(T x) {
    if constexpr (std::is_enum_v<T>) return double(int(x));
    else return double(x);
}
#define CHECK(mine, theirs) do { const bool same = num(mine) == num(theirs); \
    std::printf("%s|%s|%s\n", #mine, #theirs, same ? "ok" : "MISMATCH"); failures += !same; } while (0)

#define SPAN(X) X(sample, sample) X(position, position) X(step, step) X(pitch_slope, pitchSlope) \
    X(channels, channels) X(begin, begin) X(end, end) X(source_end, sourceEnd) X(loop_first, loopFirst) \
    X(loop_end, loopEnd) X(loop_crossfade, loopCrossfade) X(stereo_matrix, stereoMatrix) X(generator, generator) \
    X(noise_seed, noiseSeed) X(noise_start, noiseStart) X(first_modulation, firstModulation) \
    X(modulation_count, modulationCount) X(filter_track, filterTrack) X(variable_pitch, variablePitch) \
    X(minimum_pitch_ratio, minimumPitchRatio) X(sample_rate, sampleRate) X(attack_start, attackStart) \
    X(hold_start, holdStart) X(decay_start, decayStart) X(release_start, releaseStart) X(fade_start, fadeStart) \
    X(bend_end, bendEnd) X(gain, gain) X(pan, pan) X(left, left) X(right, right) X(start, start) \
    X(attack_step, attackStep) X(decay_rate, decayRate) X(sustain, sustain) X(release_value, releaseValue) \
    X(release_rate, releaseRate) X(fade_value, fadeValue) X(fade_step, fadeStep) X(width_direct, widthDirect) \
    X(width_cross, widthCross) X(position_left, positionLeft) X(position_right, positionRight) X(width, width) \
    X(stereo_position, stereoPosition)
#define FILTER(X) X(type, type) X(cutoff, cutoff) X(resonance, resonance) X(bandwidth, bandwidth) X(gain, gain)
#define ENVELOPE(X) X(attack_start, attackStart) X(hold_start, holdStart) X(decay_start, decayStart) \
    X(release_start, releaseStart) X(fade_start, fadeStart) X(start, start) X(attack_step, attackStep) \
    X(decay_rate, decayRate) X(sustain, sustain) X(release_value, releaseValue) X(release_rate, releaseRate) \
    X(fade_value, fadeValue) X(fade_step, fadeStep) X(depth, depth) X(release_seconds, releaseSeconds)
#define AUDIO_CONFIG(X) X(rate, rate) X(window, window) X(max_frames, maxFrames) X(min_peak, minPeak) \
    X(enabled, enabled) X(min_median, minMedian) X(min_quartile, minQuartile) X(silence_power, silencePower) \
    X(max_silent_fraction, maxSilentFraction) X(max_silent_frames, maxSilentFrames) \
    X(max_clipped_fraction, maxClippedFraction)
#define TAIL(X) X(offset, offset) X(frames, frames) X(data_frames, dataFrames) X(fade, fade) X(peak, peak) \
    X(first, first) X(last, last) X(status, status) X(scale, scale)
#define METRICS(X) X(median, median) X(quartile, quartile) X(silent_fraction, silentFraction) \
    X(clipped_fraction, clippedFraction) X(body_rms, bodyRms) X(target_rms, targetRms) X(gain, gain) \
    X(longest_silence, longestSilence) X(clipped, clipped) X(verdict, verdict) X(measured, measured)

static VoiceSpan span() {
    VoiceSpan s{};
#define FILL(mine, theirs) set(s.theirs);
    SPAN(FILL)
    for (int f = 0; f < voiceFilterSlots; ++f) {
        VoiceFilter& filter = s.filters[f];
#define FILL_FILTER(mine, theirs) set(filter.theirs);
        FILTER(FILL_FILTER)
    }
    for (int e = 0; e < 2; ++e) {
        EnvelopeShape& envelope = s.extraEnvelope[e];
#define FILL_ENVELOPE(mine, theirs) set(envelope.theirs);
        ENVELOPE(FILL_ENVELOPE)
        set(s.activeEnvelope[e]);
    }
    // Internal fields carry values too; the conversion must not pick them up.
    set(s.firstLfo); set(s.lfoCount); set(s.nextFilteredSpan); set(s.filteredOffset); set(s.phaseOffset);
    for (int& c : s.filterControls) set(c);
    return s;
}

static void checkSpan(const clay_voice_span& sv, const VoiceSpan& s) {
#define CHECK_SPAN(mine, theirs) CHECK(sv.mine, s.theirs);
    SPAN(CHECK_SPAN)
    for (int f = 0; f < voiceFilterSlots; ++f) {
#define CHECK_FILTER(mine, theirs) CHECK(sv.filters[f].mine, s.filters[f].theirs);
        FILTER(CHECK_FILTER)
    }
    for (int e = 0; e < 2; ++e) {
#define CHECK_ENVELOPE(mine, theirs) CHECK(sv.extra_envelope[e].mine, s.extraEnvelope[e].theirs);
        ENVELOPE(CHECK_ENVELOPE)
        CHECK(sv.active_envelope[e], s.activeEnvelope[e]);
    }
}

int main() {
    static_assert(voiceFilterSlots == CLAY_FILTER_SLOTS, "filter slots");
    // Two plans: three control tracks of ragged lengths (one empty), modulation
    // bound to a track, to none (-1) and with a shape track.
    std::vector<RenderPlan> plans(2);
    for (int p = 0; p < 2; ++p) {
        RenderPlan& plan = plans[p];
        plan.spans = {span(), span(), span()};
        plan.controls = {{{0, 0.5f, 0.f}, {64, 0.75f, 0.01f}}, {}, {{3, 1.f, -0.5f}, {9, 2.f, 0.f}, {27, 3.f, 1.f}}};
        plan.modulation = {{2, 0, ModTarget::pitch, 1.5f}, {-1, 0, ModTarget::volume, -6.f},
                           {0, 0, ModTarget::amplitude, 1.f, 1, 0}};
        set(plan.filterTracks); set(plan.frames); set(plan.musicEnd); set(plan.limit); set(plan.fadeFrames);
        set(plan.bodyFrames); set(plan.volumeGroup);
        plan.onsets = {5, 17, 400 + p};
    }
    PlanViews views;
    for (const auto& plan : plans)
        views.add(plan);
    CHECK(views.views.size(), plans.size());
    for (size_t p = 0; p < plans.size(); ++p) {
        const clay_plan_view& v = views.views[p];
        const RenderPlan& plan = plans[p];
        CHECK(v.span_count, plan.spans.size());
        for (size_t i = 0; i < plan.spans.size(); ++i)
            checkSpan(v.spans[i], plan.spans[i]);
        CHECK(v.track_count, plan.controls.size());
        CHECK(v.track_offsets[0], 0);
        for (size_t t = 0; t < plan.controls.size(); ++t) {
            CHECK(v.track_offsets[t + 1] - v.track_offsets[t], plan.controls[t].size());
            for (size_t k = 0; k < plan.controls[t].size(); ++k) {
                const size_t at = size_t(v.track_offsets[t]) + k;
                CHECK(v.points[at].frame, plan.controls[t][k].frame);
                CHECK(v.points[at].value, plan.controls[t][k].value);
                CHECK(v.points[at].step, plan.controls[t][k].step);
            }
        }
        CHECK(v.modulation_count, plan.modulation.size());
        for (size_t m = 0; m < plan.modulation.size(); ++m) {
            CHECK(v.modulation[m].track, plan.modulation[m].first);
            CHECK(v.modulation[m].target, plan.modulation[m].target);
            CHECK(v.modulation[m].depth, plan.modulation[m].depth);
            CHECK(v.modulation[m].shape_track, plan.modulation[m].shapeFirst);
        }
        CHECK(v.filter_tracks, plan.filterTracks);
        CHECK(v.frames, plan.frames);
        CHECK(v.music_end, plan.musicEnd);
        CHECK(v.limit, plan.limit);
        CHECK(v.fade_frames, plan.fadeFrames);
        CHECK(v.body_frames, plan.bodyFrames);
        CHECK(v.onset_count, plan.onsets.size());
        for (size_t o = 0; o < plan.onsets.size(); ++o)
            CHECK(v.onsets[o], plan.onsets[o]);
        CHECK(v.volume_group, plan.volumeGroup);
    }

    AudioConfig audio{};
#define FILL_AUDIO(mine, theirs) set(audio.theirs);
    AUDIO_CONFIG(FILL_AUDIO)
    audio.enabled = true;
    const clay_audio_config a = toAbi(audio);
#define CHECK_AUDIO(mine, theirs) CHECK(a.mine, audio.theirs);
    AUDIO_CONFIG(CHECK_AUDIO)

    clay_tail_result r{};
#define FILL_TAIL(mine, theirs) set(r.mine);
    TAIL(FILL_TAIL)
#define FILL_METRICS(mine, theirs) set(r.audio.mine);
    METRICS(FILL_METRICS)
    r.status = int32_t(TailStatus::clipping);
    r.audio.verdict = int32_t(TailStatus::silentBody);
    r.audio.measured = 1;
    const TailResult t = toReference(r);
#define CHECK_TAIL(mine, theirs) CHECK(r.mine, t.theirs);
    TAIL(CHECK_TAIL)
#define CHECK_METRICS(mine, theirs) CHECK(r.audio.mine, t.audio.theirs);
    METRICS(CHECK_METRICS)

    // The bank: every PCM16 value as libsndfile's float decode gives it, then
    // values beyond full scale, appended in two uploads.
    std::vector<float> pcm;
    for (int k = -32768; k < 32768; ++k)
        pcm.push_back(float(k) / 32768.0f);
    std::vector<int16_t> bank;
    appendInt16(bank, pcm);
    pcm.push_back(1.0f);
    pcm.push_back(-1.5f);
    appendInt16(bank, pcm);
    int bankErrors = int(bank.size() != pcm.size());
    for (int k = -32768; k < 32768 && !bankErrors; ++k)
        bankErrors += bank[size_t(k + 32768)] != k;
    bankErrors += bank[65536] != 32767 || bank[65537] != -32768;
    std::printf("bank|%s\n", bankErrors ? "MISMATCH" : "ok");
    failures += bankErrors != 0;
    return failures ? 1 : 0;
}
"""


def facade_source():
    text = (REPO / "INTEGRATION.md").read_text()
    section = text.split("## Facade", 1)[1]
    return re.search(r"```cpp\n(.*?)```", section, re.S).group(1)


def build_and_run(tmp_path, facade):
    (tmp_path / "facade.cpp").write_text(facade)
    (tmp_path / "check.cpp").write_text(CHECK_MAIN)
    lib = REPO / "claysampler_amd"
    if not (lib / "libclayhip.so").exists():
        pytest.skip("libclayhip.so not built")
    exe = tmp_path / "check"
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-function", f"-I{REFERENCE / 'cuda'}",
           f"-I{REPO / 'include'}", str(tmp_path / "check.cpp"), "-o", str(exe), f"-L{lib}", "-lclayhip",
           f"-Wl,-rpath,{lib}"]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-4000:]
    return subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)


def flat_fields(struct, prefix=""):
    """Field names of a ctypes structure, nested structures (and arrays of
    them) flattened as `outer.inner`."""
    names = set()
    for name, kind in struct._fields_:
        element = getattr(kind, "_type_", None) if issubclass(kind, C.Array) else kind
        if isinstance(element, type) and issubclass(element, C.Structure):
            names |= flat_fields(element, prefix + name + ".")
        else:
            names.add(prefix + name)
    return names


def reference_fields(header, struct, nested):
    """Field names of `struct` in a reference header, nested structures
    expanded; a small declaration parser good for render_types.h."""
    body = re.search(r"struct %s \{(.*?)\};" % struct, header, re.S).group(1)
    names = set()
    for declaration in body.split(";"):
        declaration = re.sub(r"\s+", " ", declaration.split("=")[0]).strip()
        if not declaration:
            continue
        first, *rest = [part.strip() for part in declaration.split(",")]
        kind, first = first.rsplit(" ", 1)
        for name in [first] + rest:
            name = name.split("[")[0]
            if kind in nested:
                names |= {name + "." + inner for inner in reference_fields(header, kind, ())}
            else:
                names.add(name)
    return names


def compared(stdout, column, prefix):
    names = set()
    for line in stdout.splitlines():
        parts = line.split("|")
        if len(parts) == 3 and parts[column].startswith(prefix):
            name = re.sub(r"\[[^]]*\]", "", parts[column][len(prefix):])
            names.add(re.match(r"[\w.]+", name).group(0))
    return names


def test_facade_converts_every_field(tmp_path):
    run = build_and_run(tmp_path, facade_source())
    assert run.returncode == 0, "\n".join(l for l in run.stdout.splitlines() if "MISMATCH" in l)
    assert "bank|ok" in run.stdout

    header = (REFERENCE / "cuda" / "render_types.h").read_text()
    assert compared(run.stdout, 0, "sv.") == flat_fields(abi.VoiceSpan)
    assert compared(run.stdout, 1, "s.") == \
        reference_fields(header, "VoiceSpan", ("VoiceFilter", "EnvelopeShape")) - INTERNAL
    plan = compared(run.stdout, 0, "v.")
    assert {n.split(".")[0] for n in plan} == flat_fields(abi.PlanView) - {"spans"}
    for record, struct in (("points", abi.ControlPoint), ("modulation", abi.Modulation)):
        assert {n.split(".", 1)[1] for n in plan if n.startswith(record + ".")} == flat_fields(struct)
    assert compared(run.stdout, 0, "a.") == flat_fields(abi.AudioConfig)
    assert compared(run.stdout, 1, "audio.") == reference_fields(header, "AudioConfig", ())
    assert compared(run.stdout, 0, "r.") == flat_fields(abi.TailResult)
    assert compared(run.stdout, 1, "t.") == reference_fields(header, "TailResult", ("AudioMetrics",))

def test_facade_check_catches_a_swapped_field(tmp_path):
    """The check itself: the facade with two envelope fields exchanged fails."""
    planted = facade_source().replace("v.decay_start = s.decayStart;", "v.decay_start = s.releaseStart;")
    assert planted != facade_source()
    run = build_and_run(tmp_path, planted)
    assert run.returncode != 0
    assert "sv.decay_start|s.decayStart|MISMATCH" in run.stdout
