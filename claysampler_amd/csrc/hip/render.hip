// Render kernels of the ClaySampler core for gfx950 (MI355X). One round of
//   pitch chains -> voice chains -> mix tiles
// renders a frame window of one or more units (a batch of jobs):
//  - plain form: clayPitchChains, clayVoiceChains, clayMixTiles take one
//    unit's pointers as __restrict__ kernel arguments (scalar loads of the
//    uniform span, modulation and control records) and exact grids;
//  - persistent form: clayRound* walk the round's unit descriptors with fixed
//    grids, grid-stride, so a captured round replays unchanged.
// Both inline the same per-item functions; with expression-level FMA
// contraction (-ffp-contract=on) they compute bit-identical audio.
// Reference behaviour: src/cuda/render.cu. Written for wave64: per-voice
// parameters are wave-uniform (scalar loads), voice chains are wave-parallel
// FP64 scans, and the mix sums voices in span order.
#include "common.h"

// ====================================================================== render

// Variable-pitch recurrence of one voice chain: per-frame source phase
// (src/cuda/render.cu:222-266), in FP64.
__device__ void pitchChain(const Span* __restrict__ spans, const SpanCold* __restrict__ cold,
                           const clay_control_point* __restrict__ points, const Mod* __restrict__ mods,
                           double* __restrict__ phase, VoiceState* __restrict__ states, int segFirst, int segEnd,
                           int head)
{
    if (!spans[head].variablePitch)
        return;
    VoiceState& st = states[spans[head].filterTrack];  // one state per voice chain
    double position = st.pitchInitialized ? st.position : spans[head].position;
    int prevLoopEnd = st.previousLoopEnd, prevLoopStart = st.previousLoopStart;
    ModReader reader;
    for (int s = head; s >= 0; s = spans[s].next) {
        const Span& v = spans[s];
        if (v.end <= segFirst || v.begin >= segEnd)
            continue;
        const bool resumed = st.pitchInitialized && st.pitchSpan == s;
        if (!resumed && !v.loopEnd && prevLoopEnd > 0 && position >= prevLoopEnd)
            position = prevLoopStart + fmod(position - prevLoopStart, double(prevLoopEnd - prevLoopStart));
        prevLoopEnd = v.loopEnd;
        prevLoopStart = v.loopFirst;
        const double slope = exp(v.pitchSlope);
        double step = resumed ? st.step : v.step;
        const clay_envelope_shape& pitchEnv = cold[s].extra[1];
        const bool envActive = cold[s].active[1] != 0;
        const int first = max(v.begin, segFirst), stop = min(v.end, segEnd);
        reader.reset(mods, points, v.firstMod, v.modCount, first);
        for (int f = first; f < stop; ++f) {
            phase[v.phase + uint64_t(f - first)] = position;
            double cents = 0, depth = pitchEnv.depth;
            for (int j = 0; j < v.modCount; ++j) {
                const int target = mods[v.firstMod + j].target;
                if (target == CLAY_MOD_PITCH)
                    cents += reader.value(mods, points, j, f);
                else if (target == CLAY_MOD_PITCHEG_DEPTH)
                    depth += reader.value(mods, points, j, f);
            }
            if (envActive)
                cents += depth * envelopeShapeD(pitchEnv, f);
            position += step * exp2(fmin(24000.0, fmax(-24000.0, cents)) / 1200);
            step *= slope;
        }
        st.pitchInitialized = 1;
        st.pitchSpan = s;
        st.position = position;
        st.step = step;
        st.previousLoopEnd = prevLoopEnd;
        st.previousLoopStart = prevLoopStart;
    }
}

// One lane per variable-pitch chain of the round, grid-stride.
// Clears the tile counters of `count` mix queues (mix.h, TileClaims) before a
// mix launch: a kernel store, so that a captured round holds no memset node.
extern "C" __global__ void clayClearQueues(MixQueue* queues, int count)
{
    const int i = int(threadIdx.x);
    if (i < count * kXcds)
        queues[i / kXcds].xcd[i % kXcds].next = 0;
}

extern "C" __global__ void clayRoundPitchChains(UnitsPtr units, RoundPtr round)
{
    const RoundHeader header = *round;
    for (int item = int(blockIdx.x * blockDim.x + threadIdx.x); item < header.pitches;
         item += int(gridDim.x * blockDim.x)) {
        const int unit = unitOf(header, item, [&](int i) { return units[i].pitchBase; });
        const auto& u = constant(units)[unit];
        pitchChain(global(u.spans), global(u.cold), global(u.points), global(u.mods), global(u.phase),
                   global(u.states), u.segFirst, u.segEnd, global(u.pitchHeads)[item - u.pitchBase]);
    }
}

// Plain form: one unit, its pointers as kernel arguments, one lane per chain.
extern "C" __global__ void clayPitchChains(const Span* __restrict__ spans, const SpanCold* __restrict__ cold,
                                           const int* __restrict__ heads, const clay_control_point* __restrict__ points,
                                           const Mod* __restrict__ mods, double* __restrict__ phase,
                                           VoiceState* __restrict__ states, int segFirst, int segEnd, int count)
{
    const int item = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (item < count)
        pitchChain(spans, cold, points, mods, phase, states, segFirst, segEnd, heads[item]);
}

// ----------------------------------------------------------------- wave-parallel voice chains
namespace {

constexpr int kMaxSections = 9;   // 2 filters x up to 3 sections + 3 EQ bands
constexpr int kChainWaves = 2;    // voice chains per 128-thread filter block (LDS: 22 KB per block)
#ifndef CLAY_LANE_FRAMES
#define CLAY_LANE_FRAMES 4
#endif
constexpr int kLaneFrames = CLAY_LANE_FRAMES;  // consecutive frames per lane in the filter pass
constexpr int kPassFrames = 64 * kLaneFrames;  // frames per filter pass (segments are multiples of it)
// Waves per SIMD the filter kernels are compiled for: 3 (168 VGPRs), which
// the per-wave LDS carry and stage (11 KB) also allow.
constexpr int kChainOccupancy = 3;
// Audio pass: 16 chains (waves) per 1024-thread block; the sinc table sits in
// LDS for up to 32 taps.
constexpr int kAudioWaves = 16;
constexpr int kSharedTapsMax = 32;
// The mix kernels: 8 waves per SIMD (64 VGPRs) keep the wet path's scratch
// reads in flight (7.2 ms vs 12.9 ms at the unconstrained 5 on a 32-tap
// workload of filtered voices) at no cost to 16-tap dry voices. A 48- or
// 64-tap dry FIR spills heavily in 64 VGPRs: those mixes run at 5 waves per
// SIMD (93 VGPRs, no spills), which took a 64-tap workload of dry voices from
// 11.1 s to 7.0 s.
constexpr int kMixOccupancy = 8;
constexpr int mixOccupancy(int taps) { return taps >= 48 ? 5 : kMixOccupancy; }

// Per-wave carry between 64-frame chunks, in LDS (broadcast reads).
struct ChainCarry {
    double z[kMaxSections][2][2];        // [section][channel][z1, z2]
    double coef[CLAY_FILTER_SLOTS][5];   // current (smoothed) coefficients
    double target[CLAY_FILTER_SLOTS][5]; // last target
    double input[CLAY_FILTER_SLOTS][3];  // inputs of the last target
    double cents[2];                     // cutoff offsets (cents) behind input[0..1][0]
    double cutoffOf[2];                  // the static cutoff behind them (NaN: not known)
    int have[CLAY_FILTER_SLOTS];
    double stage[5][kLaneFrames][64];    // per-frame targets, then coefficients, after a target change
};

// Cross-lane moves on DPP (GFX9 data-parallel primitives: VALU moves, no LDS
// traffic). Lanes without a source lane (or outside `RowMask`) take `fill`.
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

// One 32-bit half. With every row enabled and a zero fill, the bound control
// writes 0 where the source lane is missing, so no fill value has to be moved
// into the destination first (the fills are constants at every call).
template<int Ctrl, int RowMask>
__device__ __forceinline__ int shiftHalf(int x, int fill)
{
    if (RowMask == 0xf && fill == 0)
        return __builtin_amdgcn_mov_dpp(x, Ctrl, 0xf, 0xf, true);
    return __builtin_amdgcn_update_dpp(fill, x, Ctrl, RowMask, 0xf, false);
}

template<int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ double shiftD(double x, double fill)
{
    const int lo = shiftHalf<Ctrl, RowMask>(__double2loint(x), __double2loint(fill));
    const int hi = shiftHalf<Ctrl, RowMask>(__double2hiint(x), __double2hiint(fill));
    return __hiloint2double(hi, lo);
}

template<int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ int shiftI(int x, int fill)
{
    return shiftHalf<Ctrl, RowMask>(x, fill);
}

template<int C, int M>
struct ScanStep {
    static constexpr int ctrl = C, rowMask = M;
};

// Inclusive scan of a wave (64 lanes) with an associative operator, as six
// steps: prefixes within each row of 16 lanes (shifts by 1, 2, 4, 8), then row
// 0's total into row 1 and row 2's into row 3 (row_bcast:15), then rows 0-1's
// total into rows 2 and 3 (row_bcast:31). `step` receives a ScanStep and
// combines each lane with the value moved to it (the identity where none).
template<class F>
__device__ __forceinline__ void scanWave(F&& step)
{
    step(ScanStep<kRowShr1, 0xf>{});
    step(ScanStep<kRowShr2, 0xf>{});
    step(ScanStep<kRowShr4, 0xf>{});
    step(ScanStep<kRowShr8, 0xf>{});
    step(ScanStep<kRowBcast15, 0xa>{});
    step(ScanStep<kRowBcast31, 0xc>{});
}

// Value of the previous lane (lane 0 gets `fill`).
__device__ __forceinline__ double previousLane(double x, double fill) { return shiftD<kWaveShr1>(x, fill); }

// Value of a uniform lane index, as a scalar.
__device__ __forceinline__ double laneValue(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

} // namespace

// Voice-silence policy state of one chain (wave-uniform values).
struct SilenceTrack {
    int retired, silence, silentChunks, openChunk;
    float openPeak;
    int openEligible;
};

// Voice audio of a chain's spans over the segment (before pan for mono, times
// the job gain), one wave per chain and one lane per frame, written to the float
// scratch that the filter pass then reads and overwrites. Frames carry no state
// here, so this pass runs at full occupancy apart from the serial filters.
// `U` frames per lane are rendered per loop step (their window and weight
// loads issued together); every frame's arithmetic is the same for any U.
template<int T, int U = 1, class Taps>
__device__ __forceinline__ void chainAudio(const int16_t* __restrict__ bank, const Taps& taps,
                                           const Span* __restrict__ spans, const Job* __restrict__ jobs,
                                           const clay_control_point* __restrict__ points, const Mod* __restrict__ mods,
                                           const double* __restrict__ phase, float* __restrict__ filtered,
                                           const VoiceState* __restrict__ states, int segFirst, int segEnd, int head,
                                           int lane)
{
    if (states[spans[head].filterTrack].retired)
        return;  // retired in an earlier segment: nothing more to render
    for (int s = head; s >= 0; s = spans[s].next) {
        const Span& v = spans[s];
        if (v.end <= segFirst || v.begin >= segEnd)
            continue;
        const Job& j = jobs[v.job];
        const int first = max(v.begin, segFirst), stop = min(min(v.end, segEnd), v.scratchEnd);
        if (stop <= first)
            continue;  // past the chain's scratch bound
        // Whether any modulation reaches the voice audio (filter targets do
        // not), so that frames skip the scan of the span's modulations.
        bool audioMods = false;
        for (int k = v.firstMod; k < v.firstMod + v.modCount; ++k)
            audioMods = audioMods || mods[k].target <= CLAY_MOD_POSITION;
        const double bend = bendDenominator(v);
        // The job gain's tracks (volume and expression curves) usually stay in
        // one segment over the span's frames: then every frame evaluates the
        // same two points, read once here instead of searched per frame.
        const int gainA = uniform(controlSegment(points, j.controls[0], j.controls[1], first));
        const int gainB = uniform(controlSegment(points, j.controls[1], j.controls[2], first));
        const bool fixedGain = gainA == controlSegment(points, j.controls[0], j.controls[1], stop - 1) &&
                               gainB == controlSegment(points, j.controls[1], j.controls[2], stop - 1);
        const clay_control_point pointA = points[gainA], pointB = points[gainB];
        // The frame loop compiled twice: for the common voice (voiceAudio's
        // `Plain`), and for any other.
        const auto frames = [&](auto plainTag) {
            constexpr bool plain = decltype(plainTag)::value;
            for (int step = first + lane; step < stop; step += 64 * U) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int frame = step + 64 * u;
                    if (frame >= stop)
                        break;
                    const int local = frame - first;
                    const double ph = !plain && v.variablePitch ? phase[v.phase + uint64_t(local)] : -1.0;
                    // A mono filter input is not panned: only stereo voices read the pan.
                    const float pan =
                        plain || v.channels == 1 ? 0.f : controlAt(points, j.controls[2], j.controls[3], frame);
                    PanCache panCache;  // unused: mono filter inputs are not panned here
                    const float2 a = voiceAudio<T, plain, true>(bank, taps, v, points, mods, pan, frame, true, ph,
                                                                panCache, bend, audioMods);
                    const float g =
                        fixedGain ? pointAt(pointA, frame) * pointAt(pointB, frame) : jobGain(points, j, frame);
                    float* dst = filtered + v.filtered + uint64_t(local) * uint64_t(plain ? 1 : v.channels);
                    dst[0] = a.x * g;
                    if (!plain && v.channels == 2)
                        dst[1] = a.y * g;
                }
            }
        };
        if (v.channels == 1 && v.generator == CLAY_GEN_NONE && !v.variablePitch && !audioMods)
            frames(std::true_type{});
        else
            frames(std::false_type{});
    }
}

// Voice-silence bookkeeping of one filter pass (kPassFrames frames, lane l
// holding frames kLaneFrames * l ...): each row of 16 lanes is one chunk of 64
// absolute frames. A chunk may take several passes when spans split it; it is
// judged when its last frame has been rendered. After CLAY_SILENT_CHUNKS
// consecutive silent chunks (peak |output| below the level, non-finite counting
// as loud, every active frame at or after the voice's release) the chain
// retires at the end of the chunk. Returns true when it retires.
__device__ __forceinline__ bool silencePass(SilenceTrack& t, float level, int f0, int lo, int hi,
                                            const bool (&act)[kLaneFrames], const float (&outL)[kLaneFrames],
                                            const float (&outR)[kLaneFrames], int base, int releaseStart)
{
    float a = 0.f;
    bool eligible = true;
#pragma unroll
    for (int k = 0; k < kLaneFrames; ++k)
        if (act[k]) {
            float x = fmaxf(fabsf(outL[k]), fabsf(outR[k]));
            if (!(x <= 3.402823466e38f))
                x = __builtin_inff();
            a = fmaxf(a, x);
            eligible = eligible && releaseStart >= 0 && f0 + base + k >= releaseStart;
        }
    for (int d = 8; d; d >>= 1)
        a = fmaxf(a, __shfl_xor(a, d, 64));  // peak of each row of 16 lanes
    const uint64_t loud = __ballot(!eligible);
    for (int j = 0; j < kLaneFrames; ++j) {
        const int c0 = 64 * j;
        if (hi <= c0 || lo >= c0 + 64)
            continue;  // no frame of this chunk in the pass
        const float peak = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a), 16 * j));
        const bool all = ((loud >> (16 * j)) & 0xffffu) == 0;
        const int chunk = ((f0 + c0) >> 6) + 1;
        if (t.openChunk != chunk) {
            t.openChunk = chunk;
            t.openPeak = 0.f;
            t.openEligible = 1;
        }
        t.openPeak = fmaxf(t.openPeak, peak);
        t.openEligible = t.openEligible && all;
        if (hi < c0 + 64)
            return false;  // the chunk continues in a later span
        const bool silent = t.openEligible && t.openPeak < level;
        t.openChunk = 0;
        t.silentChunks = silent ? t.silentChunks + 1 : 0;
        if (t.silentChunks >= CLAY_SILENT_CHUNKS) {
            t.retired = 1;
            t.silence = f0 + c0 + 64;
            return true;
        }
    }
    return false;
}

// One wave per voice chain (a voice's spans with filters and/or variable pitch),
// kPassFrames = 256 frames per pass, passes aligned to absolute frame multiples
// of 256, lane l holding the 4 consecutive frames 4l .. 4l + 3 of the pass:
//   1. the voice audio (chainAudio) read back;
//   2. filter inputs per frame; a frame whose inputs differ from the previous
//      frame's gets new targets, copied forward to the following frames
//      (src/cuda/render.cu:376-406);
//   3. coefficient smoothing c_n = c_{n-1} + a (t_n - c_{n-1}): with unchanged
//      targets in closed form c_n = t + (1 - a)^n (c_0 - t); otherwise the
//      lane composes the affine maps c -> M c + V of its frames, a scan over
//      the lanes (DPP moves) gives the map before each lane, and the lane
//      applies its frames in order;
//   4. each biquad section (direct form II transposed, render.cu:330-336) the
//      same way with 2x2 affine maps s -> A s + d u in FP64: compose the lane's
//      frames, scan over lanes, then run the lane's frames from its entry state;
//   5. the filtered frames written back over the audio for the mix kernel.
// Composing four frames per lane in registers leaves one wave scan per 256
// frames. State carries across passes in LDS and across segments in VoiceState.
// `Stereo` is the channel count of the chain's spans (all of one voice, so of
// one sample), fixed at compile time so that mono chains hold no registers
// for a right channel. `Single`: the chain's only filter is one biquad
// section in slot 0 (the common case), so the slot and section loops vanish.
// `Cursors`: modulation bindings per span that follow their tracks with
// cursors (any further one searches its track per frame).
template<bool Stereo, bool Single, int Cursors>
__device__ __forceinline__ void voiceChain(const Span* __restrict__ spans, const SpanCold* __restrict__ cold,
                                           const clay_control_point* __restrict__ points, const Mod* __restrict__ mods,
                                           float* __restrict__ filtered, VoiceState* __restrict__ states,
                                           int segFirst, int segEnd, int head, float silenceLevel, ChainCarry& carry,
                                           int lane)
{
    constexpr int K = kLaneFrames;
    VoiceState& st = states[spans[head].filterTrack];
    SilenceTrack track{st.retired, st.silence, st.silentChunks, st.openChunk, st.openPeak, st.openEligible};
    if (track.retired)
        return;  // retired in an earlier segment: nothing more to render
    int overflow = st.overflow;
    int sectionCount[CLAY_FILTER_SLOTS], total = 0;
    for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
        sectionCount[f] = sections(cold[head].filters[f].type);
        total += sectionCount[f];
    }
    const bool hasFilter = total > 0;
    if (!hasFilter && silenceLevel <= 0.f)
        return;  // pitch only: the audio pass wrote the final frames
    bool initialized = st.filterInitialized != 0;
    __builtin_amdgcn_wave_barrier();
    if (hasFilter && lane == 0) {
        int k = 0;
        for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
            for (int i = 0; i < 5; ++i) {
                carry.coef[f][i] = st.coefficients[f][i];
                carry.target[f][i] = st.target[f][i];
            }
            for (int i = 0; i < 3; ++i)
                carry.input[f][i] = st.input[f][i];
            carry.have[f] = st.have[f];
            for (int s = 0; s < sectionCount[f]; ++s, ++k)
                for (int c = 0; c < 2; ++c) {
                    carry.z[k][c][0] = st.z1[f][s][c];
                    carry.z[k][c][1] = st.z2[f][s][c];
                }
        }
        carry.cutoffOf[0] = carry.cutoffOf[1] = __builtin_nan("");
    }
    __builtin_amdgcn_wave_barrier();
    ModReaderN<Cursors> reader;
    for (int s = head; s >= 0; s = spans[s].next) {
        const Span& v = spans[s];
        if (v.end <= segFirst || v.begin >= segEnd)
            continue;
        const SpanCold& cv = cold[s];
        const int first = max(v.begin, segFirst), want = min(v.end, segEnd), stop = min(want, v.scratchEnd);
        if (stop <= first) {
            // The chain needs frames past its scratch bound: it is stopped at
            // the bound and flagged (its job is rendered again without one).
            track.retired = 1;
            track.silence = first;
            overflow = 1;
            break;
        }
        constexpr bool stereo = Stereo;
        const int rate = cv.sampleRate;
        const double amount = 1 - exp(-1000.0 / rate), keep = 1 - amount;
        bool moving = cv.active[0] != 0;
        for (int k = v.firstMod; k < v.firstMod + v.modCount; ++k) {
            const int t = mods[k].target;
            moving |= (t >= CLAY_MOD_CUTOFF && t <= CLAY_MOD_RESONANCE2) || t == CLAY_MOD_FILEG_DEPTH ||
                      t >= CLAY_MOD_EQ1_FREQUENCY;
        }
        const int base = K * lane;  // the lane's first frame within a pass
        reader.reset(mods, points, v.firstMod, v.modCount, max(first, (first & ~(kPassFrames - 1)) + base));
        // Where modulation target t feeds filter slot f: 0 nowhere, 1 cutoff
        // (cents), 2 resonance (dB), 3 filter-envelope depth, 4..6 EQ frequency,
        // bandwidth and gain.
        const auto routeOf = [](int f, int t) {
            if (f == 0 && t == CLAY_MOD_CUTOFF) return 1;
            if (f == 0 && t == CLAY_MOD_RESONANCE) return 2;
            if (f == 0 && t == CLAY_MOD_FILEG_DEPTH) return 3;
            if (f == 1 && t == CLAY_MOD_CUTOFF2) return 1;
            if (f == 1 && t == CLAY_MOD_RESONANCE2) return 2;
            if (f >= 2 && t >= CLAY_MOD_EQ1_FREQUENCY && (t - CLAY_MOD_EQ1_FREQUENCY) / 3 == f - 2)
                return 4 + (t - CLAY_MOD_EQ1_FREQUENCY) % 3;
            return 0;
        };
        // Filter inputs of one frame for slot f: cutoff offset (cents, turned
        // into Hz by the caller) and resonance, or for EQ slots frequency,
        // bandwidth and gain. `routes` holds routeOf for the cursor slots, 3
        // bits each, so the per-frame code branches on one uniform word.
        const auto slotInputs = [&](int f, int routes, int frame, auto&& value, double& in0, double& in1,
                                    double& in2) {
            double cents = 0, db = 0, depth = cv.extra[0].depth, eq0 = 0, eq1 = 0, eq2 = 0;
            // Every accumulator adds its share, 0 unless routed (choosing one of
            // them by reference kept them in a stack array). Adding +0 changes no
            // value: the sums start at +0 or a finite depth and never become -0.
            const auto add = [&](int route, double x) {
                cents += route == 1 ? x : 0.0;
                db += route == 2 ? x : 0.0;
                depth += route == 3 ? x : 0.0;
                eq0 += route == 4 ? x : 0.0;
                eq1 += route == 5 ? x : 0.0;
                eq2 += route == 6 ? x : 0.0;
            };
            // Unrolled over the cursor slots so the cursors stay in registers.
#pragma unroll
            for (int k = 0; k < Cursors; ++k) {
                const int route = routes >> (3 * k) & 7;
                if (route)
                    add(route, value(k));
            }
            for (int k = Cursors; k < v.modCount; ++k) {
                const int route = routeOf(f, mods[v.firstMod + k].target);
                if (route)
                    add(route, reader.value(mods, points, k, frame));
            }
            const clay_voice_filter& flt = cv.filters[f];
            if (f < 2) {
                if (f == 0 && cv.active[0])
                    cents += depth * envelopeShapeD(cv.extra[0], frame);
                in0 = cents;
                in1 = double(flt.resonance) + db;
                in2 = 0;
            } else {
                in0 = double(flt.cutoff) + eq0;
                in1 = double(flt.bandwidth) + eq1;
                in2 = double(flt.gain) + eq2;
            }
        };
        for (int f0 = first & ~(kPassFrames - 1); f0 < stop; f0 += kPassFrames) {
            const int lo = max(first, f0) - f0, hi = min(stop, f0 + kPassFrames) - f0;  // active frames [lo, hi)
            const int lastLane = (hi - 1) / K, lastSlot = (hi - 1) % K;                 // the last active frame
            bool act[K];
            int act_bits = 0;
            float outL[K], outR[K];
            float* dst = filtered + v.filtered;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int at = base + k;
                act[k] = at >= lo && at < hi;
                act_bits |= act[k] ? 1 << k : 0;
                outL[k] = outR[k] = 0.f;
                if (act[k]) {
                    const float* src = dst + uint64_t(f0 + at - first) * uint64_t(v.channels);
                    outL[k] = src[0];
                    if (stereo)
                        outR[k] = src[1];
                }
            }
            if (hasFilter) {
                double ul[K], ur[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    ul[k] = outL[k];
                    ur[k] = outR[k];
                }
                int sec = 0;
                for (int f = 0; f < (Single ? 1 : CLAY_FILTER_SLOTS); ++f) {
                    const int slotSections = Single ? 1 : sections(cold[head].filters[f].type);
                    if (!slotSections)
                        continue;
                    // Coefficients (b0, b1, b2, a1, a2) of the lane's frame k: t + power[k] d
                    // (the closed form of unchanged targets, or the carried coefficients with
                    // d = 0), or, after a target change, staged per frame in LDS.
                    double t[5], d[5], power[K];
                    for (int i = 0; i < 5; ++i) {
                        t[i] = carry.coef[f][i];
                        d[i] = 0;
                    }
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        power[k] = 0;
                    bool staged = false;
                    if (moving || !initialized) {
                        int routes = 0;
                        for (int k = 0; k < Cursors && k < v.modCount; ++k)
                            routes |= routeOf(f, mods[v.firstMod + k].target) << (3 * k);
                        routes = uniform(routes);
                        double nt[5], nc[5];  // target and coefficients at the last active frame
                        // 3a. Unchanged target t over the pass (a chain not yet initialized
                        // flags its first frame, so here it is initialized and moving): c after
                        // n smoothing steps is t + keep^n (c_0 - t), keep^n from the exclusive
                        // product over the lanes of keep^(active frames of a lane).
                        const auto closedForm = [&] {
                            for (int i = 0; i < 5; ++i) {
                                t[i] = nt[i] = carry.target[f][i];
                                d[i] = carry.coef[f][i] - t[i];
                            }
                            double m = 1;
#pragma unroll
                            for (int k = 0; k < K; ++k)
                                if (act[k])
                                    m *= keep;
                            scanWave([&](auto st) {
                                constexpr int C = decltype(st)::ctrl, R = decltype(st)::rowMask;
                                m *= shiftD<C, R>(m, 1.0);
                            });
                            double p = previousLane(m, 1.0), pl = 0;
#pragma unroll
                            for (int k = 0; k < K; ++k) {
                                if (act[k])
                                    p *= keep;
                                power[k] = p;
                                if (k == lastSlot)
                                    pl = p;
                            }
                            pl = laneValue(pl, lastLane);
                            for (int i = 0; i < 5; ++i)
                                nc[i] = fma(pl, d[i], t[i]);
                        };
                        // Steady pass: when every routed track stays constant over each lane's
                        // active frames and the first of them has the carried inputs, no frame
                        // changes targets: one input evaluation per lane instead of one per frame.
                        bool quick = false;
                        if (initialized && !cv.active[0] && v.modCount <= Cursors && carry.have[f]) {
                            const int fa = max(base, lo), la = min(base + K, hi) - 1;
                            bool same = true;
                            if (fa <= la) {
                                bool steady = true;
                                double x[3] = {0, 0, 0};
                                slotInputs(
                                    f, routes, f0 + fa,
                                    [&](int j) { return reader.steadyValue(mods, points, j, f0 + fa, f0 + la, steady); },
                                    x[0], x[1], x[2]);
                                if (f < 2) {
                                    if (carry.cutoffOf[f] == double(cv.filters[f].cutoff) && x[0] == carry.cents[f])
                                        x[0] = carry.input[f][0];
                                    else
                                        steady = false;
                                }
                                same = steady && x[0] == carry.input[f][0] && x[1] == carry.input[f][1] &&
                                       x[2] == carry.input[f][2];
                            }
                            quick = __ballot(!same) == 0;
                        }
                        if (quick) {
                            closedForm();
                            __builtin_amdgcn_wave_barrier();
                            if (lane == 0)
                                for (int i = 0; i < 5; ++i)
                                    carry.coef[f][i] = nc[i];
                            __builtin_amdgcn_wave_barrier();
                        } else {
                        // 2. Inputs of the lane's frames in order, flags where they differ
                        // from the previous frame's (the first active frame compares with the
                        // carried inputs), and the targets of flagged frames, staged. One loop
                        // iteration per frame slot; the last iteration judges slot 0 against the
                        // previous lane's last frame, which every lane has evaluated by then. The
                        // hot shape (single mono chains) unrolls it (0.7 ms less on a 64-job
                        // filtered workload); the other shapes keep one copy of the code.
                        double first3[3] = {0, 0, 0}, prev[3] = {0, 0, 0}, last4[4] = {0, 0, 0, 0};
                        int flags = 0;
                        // The resonance factor of the last flagged frame's resonance.
                        double resonanceOf = __builtin_nan(""), resonanceFactor = 0;
#pragma unroll(Single && !Stereo ? K + 1 : 1)
                        for (int k = 0; k <= K; ++k) {
                            const int slot = k == K ? 0 : k, at = base + slot;
                            const bool on = (act_bits >> slot) & 1;
                            double x[3], q[3];
                            bool valid = at > lo;
                            if (k < K) {
                                x[0] = x[1] = x[2] = 0;
                                if (on)
                                    slotInputs(
                                        f, routes, f0 + at, [&](int j) { return reader.value(mods, points, j, f0 + at); },
                                        x[0], x[1], x[2]);
                                const double cents = x[0];
                                if (f < 2) {
                                    // Cutoff = static cutoff x 2^(cents / 1200). When no active lane's
                                    // cents differ from those behind the carried cutoff (same span),
                                    // the carried cutoff is the value the formula gives: skip the exp2.
                                    if (carry.cutoffOf[f] == double(cv.filters[f].cutoff) &&
                                        __ballot(on && cents != carry.cents[f]) == 0)
                                        x[0] = carry.input[f][0];
                                    else if (on)
                                        x[0] = double(cv.filters[f].cutoff) * exp2(cents * (1.0 / 1200));
                                }
                                if (slot == lastSlot)
                                    last4[0] = x[0], last4[1] = x[1], last4[2] = x[2], last4[3] = cents;
                                for (int i = 0; i < 3; ++i) {
                                    q[i] = prev[i];
                                    prev[i] = x[i];
                                }
                                if (k == 0) {
                                    for (int i = 0; i < 3; ++i)
                                        first3[i] = x[i];
                                    continue;  // judged in the last iteration
                                }
                            } else {
                                for (int i = 0; i < 3; ++i) {
                                    x[i] = first3[i];
                                    q[i] = previousLane(prev[i], 0.0);
                                }
                            }
                            if (at == lo) {
                                for (int i = 0; i < 3; ++i)
                                    q[i] = carry.input[f][i];
                                valid = carry.have[f] != 0;
                            }
                            const bool flag = on && (!valid || x[0] != q[0] || x[1] != q[1] || x[2] != q[2]);
                            flags |= flag ? 1 << slot : 0;
                            if (__ballot(flag) == 0)
                                continue;
                            if (flag) {
                                if (f < 2 && x[1] != resonanceOf) {
                                    resonanceOf = x[1];
                                    resonanceFactor = resonanceGain(x[1]);
                                }
                                const CoeffD c = coefficientsD(cv.filters[f].type, x[0], resonanceFactor, rate,
                                                               f < 2 ? 0 : x[1], f < 2 ? 0 : x[2]);
                                carry.stage[0][slot][lane] = c.b0;
                                carry.stage[1][slot][lane] = c.b1;
                                carry.stage[2][slot][lane] = c.b2;
                                carry.stage[3][slot][lane] = c.a1;
                                carry.stage[4][slot][lane] = c.a2;
                            }
                        }
                        if (__ballot(flags != 0) == 0) {
                            closedForm();
                        } else {
                            staged = true;
                            // The nearest flagged frame at or before each frame: the lane's last
                            // flag, a scan over the lanes, then forward within the lane. When
                            // every active frame is flagged (a control ramp), each frame has its
                            // own target staged already.
                            double cur[5];
                            if (__ballot((flags & act_bits) != act_bits) != 0) {
                                int have = 0;
                                double lt[5] = {0, 0, 0, 0, 0};
#pragma unroll
                                for (int k = 0; k < K; ++k)
                                    if (flags >> k & 1) {
                                        have = 1;
                                        for (int i = 0; i < 5; ++i)
                                            lt[i] = carry.stage[i][k][lane];
                                    }
                                scanWave([&](auto st) {
                                    constexpr int C = decltype(st)::ctrl, R = decltype(st)::rowMask;
                                    const int hp = shiftI<C, R>(have, 0);
                                    double tp[5];
                                    for (int i = 0; i < 5; ++i)
                                        tp[i] = shiftD<C, R>(lt[i], 0.0);
                                    if (!have && hp) {
                                        for (int i = 0; i < 5; ++i)
                                            lt[i] = tp[i];
                                        have = 1;
                                    }
                                });
                                const int incoming = shiftI<kWaveShr1>(have, 0);
                                for (int i = 0; i < 5; ++i) {
                                    cur[i] = previousLane(lt[i], 0.0);
                                    if (!incoming)
                                        cur[i] = carry.target[f][i];
                                }
#pragma unroll
                                for (int k = 0; k < K; ++k)
                                    for (int i = 0; i < 5; ++i) {
                                        if (flags >> k & 1)
                                            cur[i] = carry.stage[i][k][lane];
                                        else
                                            carry.stage[i][k][lane] = cur[i];
                                    }
                            }
                            __builtin_amdgcn_wave_barrier();
                            for (int i = 0; i < 5; ++i)
                                nt[i] = carry.stage[i][lastSlot][lastLane];
                            __builtin_amdgcn_wave_barrier();
                            // 3b. Smoothing maps per frame: identity where inactive or for a
                            // filter without modulation after its first frame; (0, t) for the
                            // first frame of a chain; (keep, amount t) otherwise.
                            const auto kind = [&](int k) {
                                const int at = base + k;
                                if (!act[k] || (!moving && at != lo))
                                    return 0;
                                return !initialized && at == lo ? 1 : 2;
                            };
                            double M = 1, V[5] = {0, 0, 0, 0, 0};
#pragma unroll
                            for (int k = 0; k < K; ++k) {
                                const int how = kind(k);
                                if (how == 1) {
                                    M = 0;
                                    for (int i = 0; i < 5; ++i)
                                        V[i] = carry.stage[i][k][lane];
                                } else if (how == 2) {
                                    M *= keep;
                                    for (int i = 0; i < 5; ++i)
                                        V[i] = fma(keep, V[i], amount * carry.stage[i][k][lane]);
                                }
                            }
                            // (M, V) after (Mp, Vp) = (M Mp, M Vp + V); identity (1, 0).
                            scanWave([&](auto st) {
                                constexpr int C = decltype(st)::ctrl, R = decltype(st)::rowMask;
                                const double Mp = shiftD<C, R>(M, 1.0);
                                for (int i = 0; i < 5; ++i)
                                    V[i] = fma(M, shiftD<C, R>(V[i], 0.0), V[i]);
                                M *= Mp;
                            });
                            const double Me = previousLane(M, 1.0);
                            for (int i = 0; i < 5; ++i)
                                cur[i] = fma(Me, carry.coef[f][i], previousLane(V[i], 0.0));
#pragma unroll
                            for (int k = 0; k < K; ++k) {
                                const int how = kind(k);
                                for (int i = 0; i < 5; ++i) {
                                    if (how == 1)
                                        cur[i] = carry.stage[i][k][lane];
                                    else if (how == 2)
                                        cur[i] = fma(keep, cur[i], amount * carry.stage[i][k][lane]);
                                    carry.stage[i][k][lane] = cur[i];
                                }
                            }
                            __builtin_amdgcn_wave_barrier();
                            for (int i = 0; i < 5; ++i)
                                nc[i] = carry.stage[i][lastSlot][lastLane];
                        }
                        // Carry: the values of the last active frame.
                        double ni[4];
                        for (int i = 0; i < 4; ++i)
                            ni[i] = laneValue(last4[i], lastLane);
                        __builtin_amdgcn_wave_barrier();
                        if (lane == 0) {
                            for (int i = 0; i < 5; ++i) {
                                carry.coef[f][i] = nc[i];
                                carry.target[f][i] = nt[i];
                            }
                            for (int i = 0; i < 3; ++i)
                                carry.input[f][i] = ni[i];
                            if (f < 2) {
                                carry.cents[f] = ni[3];
                                carry.cutoffOf[f] = cv.filters[f].cutoff;
                            }
                            carry.have[f] = 1;
                        }
                        __builtin_amdgcn_wave_barrier();
                        }
                    }
                    // 4. Biquad sections of this slot: s -> A s + d u per frame, with
                    // A = [[-a1, 1], [-a2, 0]] and d = (b1 - a1 b0, b2 - a2 b0).
                    const auto runSections = [&](auto&& coef) {
                        for (int kk = 0; kk < slotSections; ++kk, ++sec) {
                            double p00 = 1, p01 = 0, p10 = 0, p11 = 1, wl0 = 0, wl1 = 0, wr0 = 0, wr1 = 0;
#pragma unroll
                            for (int k = 0; k < K; ++k)
                                if (act[k]) {
                                    const double b0 = coef(k, 0), a1 = coef(k, 3), a2 = coef(k, 4);
                                    const double d1 = fma(-a1, b0, coef(k, 1)), d2 = fma(-a2, b0, coef(k, 2));
                                    const double n00 = fma(-a1, p00, p10), n01 = fma(-a1, p01, p11);
                                    p10 = -a2 * p00;
                                    p11 = -a2 * p01;
                                    p00 = n00;
                                    p01 = n01;
                                    const double l0 = fma(d1, ul[k], fma(-a1, wl0, wl1));
                                    wl1 = fma(d2, ul[k], -a2 * wl0);
                                    wl0 = l0;
                                    if (stereo) {
                                        const double r0 = fma(d1, ur[k], fma(-a1, wr0, wr1));
                                        wr1 = fma(d2, ur[k], -a2 * wr0);
                                        wr0 = r0;
                                    }
                                }
                            // (A, w) after (B, p) = (A B, A p + w); identity (I, 0).
                            scanWave([&](auto st) {
                                constexpr int C = decltype(st)::ctrl, R = decltype(st)::rowMask;
                                const double b00 = shiftD<C, R>(p00, 1.0), b01 = shiftD<C, R>(p01, 0.0);
                                const double b10 = shiftD<C, R>(p10, 0.0), b11 = shiftD<C, R>(p11, 1.0);
                                const double ql0 = shiftD<C, R>(wl0, 0.0), ql1 = shiftD<C, R>(wl1, 0.0);
                                wl0 = fma(p00, ql0, fma(p01, ql1, wl0));
                                wl1 = fma(p10, ql0, fma(p11, ql1, wl1));
                                if (stereo) {
                                    const double qr0 = shiftD<C, R>(wr0, 0.0), qr1 = shiftD<C, R>(wr1, 0.0);
                                    wr0 = fma(p00, qr0, fma(p01, qr1, wr0));
                                    wr1 = fma(p10, qr0, fma(p11, qr1, wr1));
                                }
                                const double n00 = fma(p00, b00, p01 * b10), n01 = fma(p00, b01, p01 * b11);
                                const double n10 = fma(p10, b00, p11 * b10), n11 = fma(p10, b01, p11 * b11);
                                p00 = n00, p01 = n01, p10 = n10, p11 = n11;
                            });
                            // The state entering the lane: the previous lanes' map applied to the carry.
                            const double e00 = previousLane(p00, 1.0), e01 = previousLane(p01, 0.0);
                            const double e10 = previousLane(p10, 0.0), e11 = previousLane(p11, 1.0);
                            double zl1 = carry.z[sec][0][0], zl2 = carry.z[sec][0][1];
                            double zr1 = carry.z[sec][1][0], zr2 = carry.z[sec][1][1];
                            {
                                const double el0 = previousLane(wl0, 0.0), el1 = previousLane(wl1, 0.0);
                                const double s1 = fma(e00, zl1, fma(e01, zl2, el0));
                                const double s2 = fma(e10, zl1, fma(e11, zl2, el1));
                                zl1 = s1, zl2 = s2;
                            }
                            if (stereo) {
                                const double er0 = previousLane(wr0, 0.0), er1 = previousLane(wr1, 0.0);
                                const double s1 = fma(e00, zr1, fma(e01, zr2, er0));
                                const double s2 = fma(e10, zr1, fma(e11, zr2, er1));
                                zr1 = s1, zr2 = s2;
                            }
                            // The lane's frames in order: y = b0 u + z1.
#pragma unroll
                            for (int k = 0; k < K; ++k)
                                if (act[k]) {
                                    const double b0 = coef(k, 0), a1 = coef(k, 3), a2 = coef(k, 4);
                                    const double d1 = fma(-a1, b0, coef(k, 1)), d2 = fma(-a2, b0, coef(k, 2));
                                    const double yl = fma(b0, ul[k], zl1);
                                    const double l1 = fma(d1, ul[k], fma(-a1, zl1, zl2));
                                    zl2 = fma(d2, ul[k], -a2 * zl1);
                                    zl1 = l1;
                                    ul[k] = yl;
                                    if (stereo) {
                                        const double yr = fma(b0, ur[k], zr1);
                                        const double r1 = fma(d1, ur[k], fma(-a1, zr1, zr2));
                                        zr2 = fma(d2, ur[k], -a2 * zr1);
                                        zr1 = r1;
                                        ur[k] = yr;
                                    }
                                }
                            // Carry: the state after the last active frame.
                            const double ql1 = laneValue(zl1, lastLane), ql2 = laneValue(zl2, lastLane);
                            const double qr1 = laneValue(zr1, lastLane), qr2 = laneValue(zr2, lastLane);
                            __builtin_amdgcn_wave_barrier();
                            if (lane == 0) {
                                carry.z[sec][0][0] = ql1, carry.z[sec][0][1] = ql2;
                                carry.z[sec][1][0] = qr1, carry.z[sec][1][1] = qr2;
                            }
                            __builtin_amdgcn_wave_barrier();
                        }
                    };
                    if (staged)
                        runSections([&](int k, int i) { return carry.stage[i][k][lane]; });
                    else
                        runSections([&](int k, int i) { return fma(power[k], d[i], t[i]); });
                }
                initialized = true;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    outL[k] = float(ul[k]);
                    outR[k] = stereo ? float(ur[k]) : 0.f;
                    if (act[k]) {
                        float* o = dst + uint64_t(f0 + base + k - first) * uint64_t(v.channels);
                        o[0] = outL[k];
                        if (stereo)
                            o[1] = outR[k];
                    }
                }
            }
            if (silenceLevel > 0.f &&
                silencePass(track, silenceLevel, f0, lo, hi, act, outL, outR, base, v.releaseStart))
                break;
        }
        if (!track.retired && stop < want) {
            track.retired = 1;  // reached the scratch bound without retiring
            track.silence = stop;
            overflow = 1;
        }
        if (track.retired)
            break;
    }
    // Save the chain state for the next segment.
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        st.overflow = overflow;
        st.retired = track.retired;
        st.silence = track.silence;
        st.silentChunks = track.silentChunks;
        st.openChunk = track.openChunk;
        st.openPeak = track.openPeak;
        st.openEligible = track.openEligible;
    }
    if (hasFilter && lane == 0) {
        int k = 0;
        for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
            for (int i = 0; i < 5; ++i) {
                st.coefficients[f][i] = carry.coef[f][i];
                st.target[f][i] = carry.target[f][i];
            }
            for (int i = 0; i < 3; ++i)
                st.input[f][i] = carry.input[f][i];
            st.have[f] = carry.have[f];
            for (int s = 0; s < sectionCount[f]; ++s, ++k)
                for (int c = 0; c < 2; ++c) {
                    st.z1[f][s][c] = carry.z[k][c][0];
                    st.z2[f][s][c] = carry.z[k][c][1];
                }
        }
        st.filterInitialized = initialized;
    }
}

// Runs body(taps) with the sinc weights the audio pass reads: for T <= 32 the
// half table staged in this block's LDS (148 KB at 32 taps, so one block per
// CU), which turns the per-lane gathers of random 128-byte rows from L2 into
// LDS reads; for wider tables the global table.
template<int MaxShared>
struct SharedUpTo {
};

template<int T, class Body, int MaxShared = kSharedTapsMax>
__device__ __forceinline__ void withAudioTaps(const float* __restrict__ sinc, Body&& body,
                                              SharedUpTo<MaxShared> = {})
{
    if constexpr (T <= MaxShared) {
        using S = SharedTaps<T>;
        __shared__ float4 table[S::kRows * S::kPitch / 4];
        for (int i = int(threadIdx.x); i < S::kRows * (T / 4); i += int(blockDim.x)) {
            const int row = i / (T / 4), q = i % (T / 4);
            table[row * (S::kPitch / 4) + q] = reinterpret_cast<const float4*>(sinc + size_t(row) * T)[q];
        }
        __syncthreads();
        body(S{table});
    } else {
        body(GlobalTaps<T>{sinc});
    }
}

// Chain of wave w (of `waves`) in row `row` of a boustrophedon walk: rows run
// forward and backward in turn, so with the chains longest first every wave
// gets a long and a short one per pair of rows (makespan about 1.03 x the mean
// on configuration 2, against 1.12 x for plain grid-stride).
__device__ __forceinline__ int snake(int row, int w, int waves)
{
    return row * waves + ((row & 1) ? waves - 1 - w : w);
}

// Audio pass: one wave per voice chain of the round, grid-stride; 16 chains per block.
template<int T>
__global__ void __launch_bounds__(64 * kAudioWaves) clayRoundChainAudio(const int16_t* __restrict__ bank,
                                                                         const float* __restrict__ sinc,
                                                                         UnitsPtr units, RoundPtr round)
{
    const int wave = uniform(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const RoundHeader header = *round;
    withAudioTaps<T>(sinc, [&](const auto& taps) {
        // Uniform per wave: no block barriers below.
        for (int row = 0;; ++row) {
            const int item = snake(row, int(blockIdx.x) * kAudioWaves + wave, int(gridDim.x) * kAudioWaves);
            if (item >= header.chains)
                break;
            const int unit = unitOf(header, item, [&](int i) { return units[i].chainBase; });
            const auto& u = constant(units)[unit];
            chainAudio<T>(bank, taps, global(u.spans), global(u.jobs), global(u.points), global(u.mods),
                          global(u.phase), global(u.filtered), global(u.states), u.segFirst, u.segEnd,
                          uniform(global(u.chainHeads)[item - u.chainBase]), lane);
        }
    });
}

// Audio pass, plain form: one unit, its pointers as kernel arguments, one wave
// per chain, grid-stride (the grid is at most one block per CU when the taps
// sit in LDS); `U` frames per lane in flight (chosen by the launch tuner).
template<int T, int U>
__global__ void __launch_bounds__(64 * kAudioWaves) clayChainAudio(
    const int16_t* __restrict__ bank, const float* __restrict__ sinc, const Span* __restrict__ spans,
    const Job* __restrict__ jobs, const int* __restrict__ heads, const clay_control_point* __restrict__ points,
    const Mod* __restrict__ mods, const double* __restrict__ phase, float* __restrict__ filtered,
    const VoiceState* __restrict__ states, int segFirst, int segEnd, int count)
{
    const int wave = uniform(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    withAudioTaps<T>(sinc, [&](const auto& taps) {
        for (int row = 0;; ++row) {
            const int item = snake(row, int(blockIdx.x) * kAudioWaves + wave, int(gridDim.x) * kAudioWaves);
            if (item >= count)
                break;
            chainAudio<T, U>(bank, taps, spans, jobs, points, mods, phase, filtered, states, segFirst, segEnd,
                             uniform(heads[item]), lane);
        }
    });
}

// Calls body(stereo, single, cursors) with the compile-time variant of the chain that
// starts at span `head` (the filter slots are the head span's, as voiceChain
// reads them).
template<class Body>
__device__ __forceinline__ void chainVariant(const Span* __restrict__ spans, const SpanCold* __restrict__ cold, int head,
                                             Body&& body)
{
    bool single = sections(cold[head].filters[0].type) == 1;
    for (int f = 1; f < CLAY_FILTER_SLOTS; ++f)
        single = single && sections(cold[head].filters[f].type) == 0;
    using One = std::integral_constant<int, 1>;
    using All = std::integral_constant<int, kCursors>;
    if (spans[head].channels == 2) {
        if (single)
            body(std::true_type{}, std::true_type{}, All{});
        else
            body(std::true_type{}, std::false_type{}, All{});
    } else if (single) {
        // One binding (a single cutoff or resonance control) is the common voice.
        if (spans[head].modCount <= 1)
            body(std::false_type{}, std::true_type{}, One{});
        else
            body(std::false_type{}, std::true_type{}, All{});
    } else {
        body(std::false_type{}, std::false_type{}, All{});
    }
}

// Filter pass: one wave per voice chain of the round, grid-stride; 2 chains per block.
__global__ void __launch_bounds__(64 * kChainWaves) __attribute__((amdgpu_waves_per_eu(kChainOccupancy)))
clayRoundVoiceChains(UnitsPtr units, RoundPtr round)
{
    __shared__ ChainCarry carries[kChainWaves];
    const int wave = uniform(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const RoundHeader header = *round;
    // Uniform per wave: no block barriers below.
    for (int item = int(blockIdx.x) * kChainWaves + wave; item < header.chains; item += int(gridDim.x) * kChainWaves) {
        const int unit = unitOf(header, item, [&](int i) { return units[i].chainBase; });
        const auto& u = constant(units)[unit];
        const int head = uniform(global(u.chainHeads)[item - u.chainBase]);
        chainVariant(global(u.spans), global(u.cold), head, [&](auto stereo, auto single, auto cursors) {
            voiceChain<decltype(stereo)::value, decltype(single)::value, decltype(cursors)::value>(
                global(u.spans), global(u.cold), global(u.points), global(u.mods), global(u.filtered),
                global(u.states), u.segFirst, u.segEnd, head, header.silence, carries[wave], lane);
        });
    }
}

// Filter pass, plain form: one unit, its pointers as kernel arguments, one wave
// per chain, `W` chains (voices) per workgroup (chosen by the launch tuner).
template<int W>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(kChainOccupancy)))
clayVoiceChains(const Span* __restrict__ spans, const SpanCold* __restrict__ cold, const int* __restrict__ heads,
                const clay_control_point* __restrict__ points, const Mod* __restrict__ mods,
                float* __restrict__ filtered, VoiceState* __restrict__ states, int segFirst, int segEnd, int count,
                float silenceLevel)
{
    __shared__ ChainCarry carries[W];
    const int wave = uniform(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const int item = int(blockIdx.x) * W + wave;
    if (item >= count)
        return;  // uniform per wave; no block barriers below
    const int head = uniform(heads[item]);
    chainVariant(spans, cold, head, [&](auto stereo, auto single, auto cursors) {
        voiceChain<decltype(stereo)::value, decltype(single)::value, decltype(cursors)::value>(
            spans, cold, points, mods, filtered, states, segFirst, segEnd, head, silenceLevel, carries[wave], lane);
    });
}

// A tile voice as the mix loop reads it, staged in LDS once per tile so that
// the per-voice loop reads no span record or voice state from memory.
struct MixVoice {
    int span;            // span index (a dry voice renders from the full record)
    int begin, stop;     // frames the voice contributes: [begin, end), or up to a retired chain's silence
    int wet;             // 1: filtered scratch; 0: dry, rendered inline
    int channels, firstMod, modCount, panMods;  // panMods: 1 when a modulation binding targets pan
    int64_t base;        // scratch index of frame f: base + f * channels
    float left, right, pan, gain;
};
constexpr int kMixVoices = 64;  // voices staged per batch (more run in further batches)
constexpr int kWetAhead = 4;    // scratch frames in flight per thread in the wet-only mix

// Per output frame (one per thread of the tile's block): the dry voices of the
// tile rendered inline, times the job gain, plus the filtered voices from
// scratch (pan applied for mono). Voices are summed in span order, so every
// execution form yields the same bits (src/cuda/render.cu:458-523). Natural
// writes float2, orig/custom short2. Every thread of the block calls this
// (block barriers stage the voice records).
// `lane` is the thread's frame within the tile and `sync` the barrier of the
// threads that share `staged` (the block, or one wave). `WetOnly`: the unit
// has no dry voice, so the inline FIR is compiled out and kWetAhead voices'
// scratch frames are read before any of them is summed (same sums, same order).
template<int T, bool FloatOut, bool WetOnly = false, class Taps, class Sync>
__device__ __forceinline__ void mixTile(const int16_t* __restrict__ bank, const Taps& taps,
                                        const Span* __restrict__ spans, const Job* __restrict__ jobs,
                                        const int* __restrict__ voices, const clay_control_point* __restrict__ points,
                                        const Mod* __restrict__ mods, const float* __restrict__ filtered,
                                        const VoiceState* __restrict__ states, void* __restrict__ out, int segFirst,
                                        int segEnd, const Tile& tile, MixVoice* staged, int lane, int stager,
                                        Sync&& sync)
{
    const int frame = tile.frame + lane;
    const bool on = lane < tile.frames && frame >= segFirst && frame < segEnd;
    const Job& j = jobs[uniform(tile.job)];
    const float pan = on ? controlAt(points, j.controls[2], j.controls[3], frame) : 0.f;
    float left = 0, right = 0, wetLeft = 0, wetRight = 0;
    PanCache panCache;
    for (int b0 = tile.first; b0 < tile.first + tile.count; b0 += kMixVoices) {
        const int n = min(kMixVoices, tile.first + tile.count - b0);
        sync();  // the previous batch has been read
        if (stager < n) {
            const int s = voices[b0 + stager];
            const Span& v = spans[s];
            MixVoice m;
            m.span = s;
            m.begin = v.begin;
            m.stop = v.end;
            m.wet = v.filterTrack >= 0;
            if (m.wet) {
                const VoiceState& st = states[v.filterTrack];
                if (st.retired)
                    m.stop = min(m.stop, st.silence);  // retired under the voice-silence policy
            }
            m.channels = v.channels;
            m.firstMod = v.firstMod;
            m.modCount = v.modCount;
            int panMods = 0;
            for (int k = v.firstMod; k < v.firstMod + v.modCount; ++k)
                panMods |= mods[k].target == CLAY_MOD_PAN;
            m.panMods = panMods;
            m.base = int64_t(v.filtered) - int64_t(max(v.begin, segFirst)) * v.channels;
            m.left = v.left;
            m.right = v.right;
            m.pan = v.pan;
            m.gain = v.gain;
            staged[stager] = m;
        }
        sync();
        if (!on)
            continue;
        // A wet voice's scratch frame, panned when mono (render.cu:458-523).
        const auto wetFrame = [&](const MixVoice& m, float2 s) {
            if (m.channels == 1) {
                float offset = pan;
                if (m.panMods)
                    for (int k = m.firstMod; k < m.firstMod + m.modCount; ++k) {
                        const Mod md = mods[k];
                        if (md.target == CLAY_MOD_PAN)
                            offset += modulationAt(md, points, frame);
                    }
                // panGains (src/cuda/render.cu:130-136)
                float gl = m.left, gr = m.right;
                if (offset != 0) {
                    const float p = fminf(1.f, fmaxf(-1.f, m.pan + offset));
                    if (p != panCache.p) {
                        panCache.p = p;
                        panCache.a = cosf((1 + p) * kQuarterPi);
                        panCache.b = cosf((1 - p) * kQuarterPi);
                    }
                    gl = m.gain * panCache.a;
                    gr = m.gain * panCache.b;
                }
                s.x *= gl;
                s.y *= gr;
            }
            wetLeft += s.x;
            wetRight += s.y;
        };
        const auto read = [&](const MixVoice& m) {
            const float* src = filtered + (m.base + int64_t(frame) * m.channels);
            return make_float2(src[0], m.channels == 1 ? src[0] : src[1]);
        };
        if constexpr (WetOnly) {
            for (int i = 0; i < n; i += kWetAhead) {
                bool in[kWetAhead];
                float2 got[kWetAhead];
#pragma unroll
                for (int k = 0; k < kWetAhead; ++k) {
                    const MixVoice& m = staged[min(i + k, n - 1)];
                    in[k] = i + k < n && frame >= m.begin && frame < m.stop;
                    got[k] = in[k] ? read(m) : make_float2(0.f, 0.f);
                }
#pragma unroll
                for (int k = 0; k < kWetAhead; ++k)
                    if (in[k])
                        wetFrame(staged[i + k], got[k]);
            }
            continue;
        }
        for (int i = 0; i < n; ++i) {
            const MixVoice& m = staged[i];
            if (frame < m.begin || frame >= m.stop)
                continue;
            if (m.wet) {
                wetFrame(m, read(m));
            } else {
                const Span& v = spans[uniform(m.span)];
                const float2 s = voiceAudio<T>(bank, taps, v, points, mods, pan, frame, false, -1.0, panCache);
                left += s.x;
                right += s.y;
            }
        }
    }
    if (!on)
        return;
    const float amp = jobGain(points, j, frame);
    const float2 mixed = make_float2(fmaf(left, amp, wetLeft), fmaf(right, amp, wetRight));
    const uint64_t at = j.mix + uint64_t(frame);
    if (FloatOut) {
        static_cast<float2*>(out)[at] = mixed;
    } else {
        const short l = short(__float2int_rn(fminf(32767.f, fmaxf(-32768.f, mixed.x * 32768.f))));
        const short r = short(__float2int_rn(fminf(32767.f, fmaxf(-32768.f, mixed.y * 32768.f))));
        static_cast<short2*>(out)[at] = make_short2(l, r);
    }
}

// Plain form: one unit, its pointers as kernel arguments, one block per tile.
template<int T, bool FloatOut, bool WetOnly>
__global__ void __launch_bounds__(kTileFrames) __attribute__((amdgpu_waves_per_eu(mixOccupancy(T)))) clayMixTiles(
    const int16_t* __restrict__ bank, const float* __restrict__ sinc, const Span* __restrict__ spans,
    const Job* __restrict__ jobs, const Tile* __restrict__ tiles, const int* __restrict__ voices,
    const clay_control_point* __restrict__ points, const Mod* __restrict__ mods, const float* __restrict__ filtered,
    const VoiceState* __restrict__ states, void* __restrict__ out, int segFirst, int segEnd, int firstTile)
{
    __shared__ MixVoice staged[kMixVoices];
    const Tile tile = tiles[firstTile + int(blockIdx.x)];
    mixTile<T, FloatOut, WetOnly>(bank, GlobalTaps<T>{sinc}, spans, jobs, voices, points, mods, filtered, states, out, segFirst,
                         segEnd, tile, staged, int(threadIdx.x), int(threadIdx.x), [] { __syncthreads(); });
}

#define CLAY_INSTANTIATE_MIX(T, FloatOut, WetOnly)                                                                     \
    template __global__ void clayMixTiles<T, FloatOut, WetOnly>(const int16_t*, const float*, const Span*, const Job*, \
                                                                const Tile*, const int*, const clay_control_point*,    \
                                                                const Mod*, const float*, const VoiceState*, void*,    \
                                                                int, int, int);
#define CLAY_INSTANTIATE_AUDIO(T, U)                                                                                   \
    template __global__ void clayChainAudio<T, U>(const int16_t*, const float*, const Span*, const Job*, const int*,   \
                                                  const clay_control_point*, const Mod*, const double*, float*,        \
                                                  const VoiceState*, int, int, int);
#define CLAY_INSTANTIATE_TAPS(T)                                                                                       \
    template __global__ void clayRoundChainAudio<T>(const int16_t*, const float*, UnitsPtr, RoundPtr);                 \
    CLAY_INSTANTIATE_AUDIO(T, 1)                                                                                        \
    CLAY_INSTANTIATE_AUDIO(T, 2)                                                                                        \
    CLAY_INSTANTIATE_MIX(T, true, true)                                                                                 \
    CLAY_INSTANTIATE_MIX(T, false, true)
CLAY_INSTANTIATE_TAPS(16)
CLAY_INSTANTIATE_TAPS(32)
CLAY_INSTANTIATE_TAPS(48)
CLAY_INSTANTIATE_TAPS(64)

#define CLAY_INSTANTIATE_CHAINS(W)                                                                                     \
    template __global__ void clayVoiceChains<W>(const Span*, const SpanCold*, const int*, const clay_control_point*,   \
                                                const Mod*, float*, VoiceState*, int, int, int, float);
CLAY_INSTANTIATE_CHAINS(1)
CLAY_INSTANTIATE_CHAINS(2)
CLAY_INSTANTIATE_CHAINS(4)
