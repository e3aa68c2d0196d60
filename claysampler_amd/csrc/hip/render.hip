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

// The round descriptors are read through the constant address space: the
// compiler then takes the pointers stored in them for global memory (global
// and, for uniform addresses of read-only records, scalar loads instead of
// flat loads). The per-item functions take them as __restrict__ parameters,
// which is what licenses the scalar loads next to the stores of one kernel.
#define CLAY_CONST __attribute__((address_space(4)))
using UnitsPtr = const UnitView* __restrict__;
using RoundPtr = const RoundHeader* __restrict__;

__device__ __forceinline__ const CLAY_CONST UnitView* constant(const UnitView* p)
{
    return (const CLAY_CONST UnitView*)p;
}

template<class P>
__device__ __forceinline__ P* global(P* p)
{
    return p;
}

// Resident unit of work item `item` given the exclusive prefix sums of the
// round (units are few; the search runs on scalar registers).
template<class F>
__device__ __forceinline__ int unitOf(const RoundHeader& header, int item, F baseOf)
{
    int lo = 0, hi = header.units;
    while (lo + 1 < hi) {
        const int mid = (lo + hi) >> 1;
        if (baseOf(mid) <= item)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

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
constexpr int kChainWaves = 4;    // voice chains per 256-thread block
// Waves per SIMD the chain kernels are compiled for: 4 caps them at 128 VGPRs
// (88 B of spills), about 2.5 % faster on configuration 2 than 3 (154 VGPRs)
// and well ahead of the unconstrained 2.
constexpr int kChainOccupancy = 4;
// Audio pass: 16 chains (waves) per 1024-thread block; the sinc table sits in
// LDS for up to 32 taps.
constexpr int kAudioWaves = 16;
constexpr int kSharedTapsMax = 32;
// The mix kernels: 8 waves per SIMD (64 VGPRs) keep the wet path's scratch
// reads in flight (7.2 ms vs 12.9 ms at the unconstrained 5 on configuration
// 2) at no cost to the dry path (configuration 3 unchanged).
constexpr int kMixOccupancy = 8;

// Per-wave carry between 64-frame chunks, in LDS (broadcast reads).
struct ChainCarry {
    double z[kMaxSections][2][2];        // [section][channel][z1, z2]
    double coef[CLAY_FILTER_SLOTS][5];   // current (smoothed) coefficients
    double target[CLAY_FILTER_SLOTS][5]; // last target
    double input[CLAY_FILTER_SLOTS][3];  // inputs of the last target
    double cents[2];                     // cutoff offsets (cents) behind input[0..1][0]
    int have[CLAY_FILTER_SLOTS];
};

// Cross-lane moves on DPP (GFX9 data-parallel primitives: VALU moves, no LDS
// traffic). Lanes without a source lane (or outside `RowMask`) take `fill`.
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

template<int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ double shiftD(double x, double fill)
{
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(x), Ctrl, RowMask, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(x), Ctrl, RowMask, 0xf, false);
    return __hiloint2double(hi, lo);
}

template<int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ int shiftI(int x, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, x, Ctrl, RowMask, 0xf, false);
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

// Measures one chunk pass of a chain: the wave's peak |output| (non-finite
// counts as loud) and whether every active frame is at or after the voice's
// release. A chunk (64 absolute frames) may take several passes when spans
// split it; it is judged when its last frame has been rendered. After
// CLAY_SILENT_CHUNKS consecutive silent chunks the chain retires at the end
// of the current chunk. Returns true when it retires.
__device__ __forceinline__ bool silenceStep(SilenceTrack& t, float level, int f0, int hi, bool active, float outL,
                                            float outR, int frame, int releaseStart)
{
    float a = active ? fmaxf(fabsf(outL), fabsf(outR)) : 0.f;
    if (!(a <= 3.402823466e38f))
        a = __builtin_inff();
    for (int d = 32; d; d >>= 1)
        a = fmaxf(a, __shfl_xor(a, d, 64));
    const bool eligible = !active || (releaseStart >= 0 && frame >= releaseStart);
    const bool all = __ballot(!eligible) == 0;
    const int chunk = (f0 >> 6) + 1;
    if (t.openChunk != chunk) {
        t.openChunk = chunk;
        t.openPeak = 0.f;
        t.openEligible = 1;
    }
    t.openPeak = fmaxf(t.openPeak, a);
    t.openEligible = t.openEligible && all;
    if (hi < 64)
        return false;
    const bool silent = t.openEligible && t.openPeak < level;
    t.openChunk = 0;
    t.silentChunks = silent ? t.silentChunks + 1 : 0;
    if (t.silentChunks < CLAY_SILENT_CHUNKS)
        return false;
    t.retired = 1;
    t.silence = f0 + 64;
    return true;
}

// Voice audio of a chain's spans over the segment (before pan for mono, times
// the job gain), one wave per chain and one lane per frame, written to the float
// scratch that the filter pass then reads and overwrites. Frames carry no state
// here, so this pass runs at full occupancy apart from the serial filters.
template<int T, class Taps>
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
        const int first = max(v.begin, segFirst), stop = min(v.end, segEnd);
        for (int frame = first + lane; frame < stop; frame += 64) {
            const int local = frame - first;
            const double ph = v.variablePitch ? phase[v.phase + uint64_t(local)] : -1.0;
            const float pan = controlAt(points, j.controls[2], j.controls[3], frame);
            const float2 a = voiceAudio<T>(bank, taps, v, points, mods, pan, frame, true, ph);
            const float g = jobGain(points, j, frame);
            float* dst = filtered + v.filtered + uint64_t(local) * uint64_t(v.channels);
            dst[0] = a.x * g;
            if (v.channels == 2)
                dst[1] = a.y * g;
        }
    }
}

// One wave per voice chain (a voice's spans with filters and/or variable pitch),
// 64 output frames per step, chunks aligned to absolute frame multiples of 64:
//   1. the 64 frames of voice audio (chainAudio) read back, one per lane;
//   2. filter targets where a lane's inputs differ from the previous frame's,
//      copied forward to the following lanes (src/cuda/render.cu:376-406);
//   3. coefficient smoothing c_n = c_{n-1} + a (t_n - c_{n-1}) as an inclusive
//      scan of affine maps c -> M c + V over the 64 lanes (DPP moves);
//   4. each biquad section (direct form II transposed, render.cu:330-336) as a
//      scan of 2x2 affine maps s_n = A_n s_{n-1} + B_n u_n, in FP64;
//   5. the filtered frames written to the float scratch the mix kernel reads.
// State carries across chunks in LDS and across segments in VoiceState.
__device__ __forceinline__ void voiceChain(const Span* __restrict__ spans, const SpanCold* __restrict__ cold,
                                           const clay_control_point* __restrict__ points, const Mod* __restrict__ mods,
                                           float* __restrict__ filtered, VoiceState* __restrict__ states,
                                           int segFirst, int segEnd, int head, float silenceLevel, ChainCarry& carry,
                                           int lane)
{
    VoiceState& st = states[spans[head].filterTrack];
    SilenceTrack track{st.retired, st.silence, st.silentChunks, st.openChunk, st.openPeak, st.openEligible};
    if (track.retired)
        return;  // retired in an earlier segment: nothing more to render
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
    }
    __builtin_amdgcn_wave_barrier();
    for (int s = head; s >= 0; s = spans[s].next) {
        const Span& v = spans[s];
        if (v.end <= segFirst || v.begin >= segEnd)
            continue;
        const SpanCold& cv = cold[s];
        const int first = max(v.begin, segFirst), stop = min(v.end, segEnd);
        const bool stereo = v.channels == 2;
        const int rate = cv.sampleRate;
        const double amount = 1 - exp(-1000.0 / rate);
        bool moving = cv.active[0] != 0;
        for (int k = v.firstMod; k < v.firstMod + v.modCount; ++k) {
            const int t = mods[k].target;
            moving |= (t >= CLAY_MOD_CUTOFF && t <= CLAY_MOD_RESONANCE2) || t == CLAY_MOD_FILEG_DEPTH ||
                      t >= CLAY_MOD_EQ1_FREQUENCY;
        }
        ModReader reader;
        reader.reset(mods, points, v.firstMod, v.modCount, max(first, (first & ~63) + lane));
        // Bit f: carry.cents[f] holds the cents behind carry.input[f][0], computed in this span.
        int centsValid = 0;
        for (int f0 = first & ~63; f0 < stop; f0 += 64) {
            const int frame = f0 + lane;
            const bool active = frame >= first && frame < stop;
            const int lo = max(first, f0) - f0, hi = min(stop, f0 + 64) - f0;  // active lanes [lo, hi)
            float* dst = filtered + v.filtered + uint64_t(max(frame - first, 0)) * uint64_t(v.channels);
            float xl = 0.f, xr = 0.f;
            if (active) {
                xl = dst[0];
                if (stereo)
                    xr = dst[1];
            }
            float outL = xl, outR = xr;
            if (hasFilter) {
                // Filter inputs of this lane's frame for slot f: cutoff (Hz) and
                // resonance, or for EQ slots frequency, bandwidth and gain.
                const auto slotInputs = [&](int f, double& in0, double& in1, double& in2) {
                    double cents = 0, db = 0, depth = cv.extra[0].depth, eq0 = 0, eq1 = 0, eq2 = 0;
                    const auto add = [&](int k) {
                        const int t = mods[v.firstMod + k].target;
                        if (f == 0 && t == CLAY_MOD_CUTOFF) cents += reader.value(mods, points, k, frame);
                        else if (f == 0 && t == CLAY_MOD_RESONANCE) db += reader.value(mods, points, k, frame);
                        else if (f == 0 && t == CLAY_MOD_FILEG_DEPTH) depth += reader.value(mods, points, k, frame);
                        else if (f == 1 && t == CLAY_MOD_CUTOFF2) cents += reader.value(mods, points, k, frame);
                        else if (f == 1 && t == CLAY_MOD_RESONANCE2) db += reader.value(mods, points, k, frame);
                        else if (f >= 2 && t >= CLAY_MOD_EQ1_FREQUENCY && (t - CLAY_MOD_EQ1_FREQUENCY) / 3 == f - 2) {
                            const int field = (t - CLAY_MOD_EQ1_FREQUENCY) % 3;
                            const double x = reader.value(mods, points, k, frame);
                            if (field == 0) eq0 += x;
                            else if (field == 1) eq1 += x;
                            else eq2 += x;
                        }
                    };
                    // Unrolled over the cursor slots so the cursors stay in registers.
    #pragma unroll
                    for (int k = 0; k < kCursors; ++k)
                        if (k < v.modCount)
                            add(k);
                    for (int k = kCursors; k < v.modCount; ++k)
                        add(k);
                    const double eq[3] = {eq0, eq1, eq2};
                    const clay_voice_filter& flt = cv.filters[f];
                    if (f < 2) {
                        if (f == 0 && cv.active[0])
                            cents += depth * envelopeShapeD(cv.extra[0], frame);
                        in0 = cents;  // the caller turns it into the cutoff
                        in1 = double(flt.resonance) + db;
                        in2 = 0;
                    } else {
                        in0 = double(flt.cutoff) + eq[0];
                        in1 = double(flt.bandwidth) + eq[1];
                        in2 = double(flt.gain) + eq[2];
                    }
                };
                const bool recompute = moving || !initialized;
                int sec = 0;
                double ul = xl, ur = xr;
                for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
                    const int slotSections = sections(cold[head].filters[f].type);
                    if (!slotSections)
                        continue;
                    double c[5];
                    if (recompute) {
                        double in0 = 0, in1 = 0, in2 = 0;
                        if (active)
                            slotInputs(f, in0, in1, in2);
                        double cents = 0;
                        if (f < 2) {
                            // Cutoff = static cutoff x 2^(cents / 1200). When every active lane has
                            // the cents of the span's last evaluated frame, the carried cutoff is
                            // the value the formula gives: skip the FP64 exp2.
                            cents = in0;
                            if ((centsValid >> f & 1) && __ballot(active && cents != carry.cents[f]) == 0)
                                in0 = carry.input[f][0];
                            else if (active)
                                in0 = double(cv.filters[f].cutoff) * exp2(cents / 1200);
                        }
                        // 2. Targets: recompute where inputs change, copy forward otherwise.
                        double p0 = previousLane(in0, 0.0), p1 = previousLane(in1, 0.0), p2 = previousLane(in2, 0.0);
                        bool prevValid = lane > lo;
                        if (lane == lo) {
                            p0 = carry.input[f][0], p1 = carry.input[f][1], p2 = carry.input[f][2];
                            prevValid = carry.have[f] != 0;
                        }
                        const bool flag = active && (!prevValid || in0 != p0 || in1 != p1 || in2 != p2);
                        const bool anyFlag = __ballot(flag) != 0;
                        double t[5] = {0, 0, 0, 0, 0};
                        if (!anyFlag) {
                            for (int i = 0; i < 5; ++i)
                                t[i] = carry.target[f][i];
                        } else {
                            if (flag) {
                                const CoeffD q = f < 2 ? coefficientsD(cv.filters[f].type, in0, in1, rate, 0, 0)
                                                       : coefficientsD(cv.filters[f].type, in0, 0, rate, in1, in2);
                                t[0] = q.b0, t[1] = q.b1, t[2] = q.b2, t[3] = q.a1, t[4] = q.a2;
                            }
                            // Nearest flagged lane at or before each lane (the later one wins).
                            int have = flag ? 1 : 0;
                            scanWave([&](auto st) {
                                constexpr int C = decltype(st)::ctrl, R = decltype(st)::rowMask;
                                const int hp = shiftI<C, R>(have, 0);
                                double tp[5];
                                for (int i = 0; i < 5; ++i)
                                    tp[i] = shiftD<C, R>(t[i], 0.0);
                                if (!have && hp) {
                                    for (int i = 0; i < 5; ++i)
                                        t[i] = tp[i];
                                    have = 1;
                                }
                            });
                            if (!have)
                                for (int i = 0; i < 5; ++i)
                                    t[i] = carry.target[f][i];
                        }
                        if (!anyFlag && moving && initialized) {
                            // 3a. Constant target over the chunk: c_n = t + q^k (c - t), q = 1 - a,
                            // k = frames smoothed so far in this chunk.
                            const int k = lane - lo + 1;
                            double qk = 1, qp = 1 - amount;
                            for (int bit = 1; bit <= 64; bit <<= 1, qp *= qp)
                                if (k & bit)
                                    qk *= qp;
                            for (int i = 0; i < 5; ++i)
                                c[i] = fma(qk, carry.coef[f][i] - t[i], t[i]);
                        } else {
                            // 3b. Smoothing scan: c_n = M_n c_{n-1} + V_n.
                            // A filter without modulation takes its coefficients once, unsmoothed.
                            double M, V[5];
                            if (!active || (!moving && lane != lo)) {
                                M = 1;
                                for (int i = 0; i < 5; ++i)
                                    V[i] = 0;
                            } else if (!initialized && lane == lo) {
                                M = 0;
                                for (int i = 0; i < 5; ++i)
                                    V[i] = t[i];
                            } else {
                                M = 1 - amount;
                                for (int i = 0; i < 5; ++i)
                                    V[i] = amount * t[i];
                            }
                            // (M, V) o (Mp, Vp) = (M Mp, M Vp + V); identity (1, 0).
                            scanWave([&](auto st) {
                                constexpr int C = decltype(st)::ctrl, R = decltype(st)::rowMask;
                                const double Mp = shiftD<C, R>(M, 1.0);
                                for (int i = 0; i < 5; ++i)
                                    V[i] = fma(M, shiftD<C, R>(V[i], 0.0), V[i]);
                                M *= Mp;
                            });
                            for (int i = 0; i < 5; ++i)
                                c[i] = fma(M, carry.coef[f][i], V[i]);
                        }
                        // Carry: values of the last active lane.
                        const int last = hi - 1;
                        double nc[5], nt[5], ni[3];
                        for (int i = 0; i < 5; ++i) {
                            nc[i] = laneValue(c[i], last);
                            nt[i] = laneValue(t[i], last);
                        }
                        ni[0] = laneValue(in0, last);
                        ni[1] = laneValue(in1, last);
                        ni[2] = laneValue(in2, last);
                        const double nCents = laneValue(cents, last);
                        __builtin_amdgcn_wave_barrier();
                        if (lane == 0) {
                            if (f < 2)
                                carry.cents[f] = nCents;
                            for (int i = 0; i < 5; ++i) {
                                carry.coef[f][i] = nc[i];
                                carry.target[f][i] = nt[i];
                            }
                            for (int i = 0; i < 3; ++i)
                                carry.input[f][i] = ni[i];
                            carry.have[f] = 1;
                        }
                        if (f < 2)
                            centsValid |= 1 << f;
                        __builtin_amdgcn_wave_barrier();
                    } else {
                        for (int i = 0; i < 5; ++i)
                            c[i] = carry.coef[f][i];
                    }
                    // 4. Biquad sections of this slot: s_n = A_n s_{n-1} + B_n u_n.
                    for (int k = 0; k < slotSections; ++k, ++sec) {
                        // A = [[-a1, 1], [-a2, 0]]; inactive lanes are the identity.
                        double a00 = -c[3], a01 = 1, a10 = -c[4], a11 = 0;
                        double wl0 = (c[1] - c[3] * c[0]) * ul, wl1 = (c[2] - c[4] * c[0]) * ul;
                        double wr0 = (c[1] - c[3] * c[0]) * ur, wr1 = (c[2] - c[4] * c[0]) * ur;
                        if (!active) {
                            a00 = 1, a01 = 0, a10 = 0, a11 = 1;
                            wl0 = wl1 = wr0 = wr1 = 0;
                        }
                        // (A, w) o (B, p) = (A B, A p + w); identity (I, 0).
                        scanWave([&](auto st) {
                            constexpr int C = decltype(st)::ctrl, R = decltype(st)::rowMask;
                            const double b00 = shiftD<C, R>(a00, 1.0), b01 = shiftD<C, R>(a01, 0.0);
                            const double b10 = shiftD<C, R>(a10, 0.0), b11 = shiftD<C, R>(a11, 1.0);
                            const double pl0 = shiftD<C, R>(wl0, 0.0), pl1 = shiftD<C, R>(wl1, 0.0);
                            wl0 = fma(a00, pl0, fma(a01, pl1, wl0));
                            wl1 = fma(a10, pl0, fma(a11, pl1, wl1));
                            if (stereo) {
                                const double pr0 = shiftD<C, R>(wr0, 0.0), pr1 = shiftD<C, R>(wr1, 0.0);
                                wr0 = fma(a00, pr0, fma(a01, pr1, wr0));
                                wr1 = fma(a10, pr0, fma(a11, pr1, wr1));
                            }
                            const double n00 = fma(a00, b00, a01 * b10), n01 = fma(a00, b01, a01 * b11);
                            const double n10 = fma(a10, b00, a11 * b10), n11 = fma(a10, b01, a11 * b11);
                            a00 = n00, a01 = n01, a10 = n10, a11 = n11;
                        });
                        const double zl1 = carry.z[sec][0][0], zl2 = carry.z[sec][0][1];
                        const double zr1 = carry.z[sec][1][0], zr2 = carry.z[sec][1][1];
                        const double sl1 = fma(a00, zl1, fma(a01, zl2, wl0)), sl2 = fma(a10, zl1, fma(a11, zl2, wl1));
                        const double sr1 = fma(a00, zr1, fma(a01, zr2, wr0)), sr2 = fma(a10, zr1, fma(a11, zr2, wr1));
                        // y_n = b0 u_n + z1_{n-1}
                        double prevL = previousLane(sl1, zl1), prevR = previousLane(sr1, zr1);
                        if (lane == lo) {
                            prevL = zl1;
                            prevR = zr1;
                        }
                        ul = fma(c[0], ul, prevL);
                        ur = stereo ? fma(c[0], ur, prevR) : 0.0;
                        const int last = hi - 1;
                        const double ql1 = laneValue(sl1, last), ql2 = laneValue(sl2, last);
                        const double qr1 = laneValue(sr1, last), qr2 = laneValue(sr2, last);
                        __builtin_amdgcn_wave_barrier();
                        if (lane == 0) {
                            carry.z[sec][0][0] = ql1, carry.z[sec][0][1] = ql2;
                            carry.z[sec][1][0] = qr1, carry.z[sec][1][1] = qr2;
                        }
                        __builtin_amdgcn_wave_barrier();
                    }
                }
                initialized = true;
                outL = float(ul);
                outR = float(ur);
            }
            if (active && hasFilter) {
                dst[0] = outL;
                if (stereo)
                    dst[1] = outR;
            }
            if (silenceLevel > 0.f &&
                silenceStep(track, silenceLevel, f0, hi, active, outL, stereo ? outR : 0.f, frame, v.releaseStart))
                break;
        }
        if (track.retired)
            break;
    }
    // Save the chain state for the next segment.
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
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
template<int T, class Body>
__device__ __forceinline__ void withAudioTaps(const float* __restrict__ sinc, Body&& body)
{
    if constexpr (T <= kSharedTapsMax) {
        using S = SharedTaps<T>;
        __shared__ float4 table[S::kRows * S::kPitch / 4];
        for (int i = int(threadIdx.x); i < S::kRows * (T / 4); i += int(blockDim.x)) {
            const int row = i / (T / 4), q = i % (T / 4);
            table[row * (S::kPitch / 4) + q] = reinterpret_cast<const float4*>(sinc + size_t(row) * T)[q];
        }
        __syncthreads();
        body(S{reinterpret_cast<const float*>(table)});
    } else {
        body(GlobalTaps<T>{sinc});
    }
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
        for (int item = int(blockIdx.x) * kAudioWaves + wave; item < header.chains;
             item += int(gridDim.x) * kAudioWaves) {
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
// sit in LDS).
template<int T>
__global__ void __launch_bounds__(64 * kAudioWaves) clayChainAudio(
    const int16_t* __restrict__ bank, const float* __restrict__ sinc, const Span* __restrict__ spans,
    const Job* __restrict__ jobs, const int* __restrict__ heads, const clay_control_point* __restrict__ points,
    const Mod* __restrict__ mods, const double* __restrict__ phase, float* __restrict__ filtered,
    const VoiceState* __restrict__ states, int segFirst, int segEnd, int count)
{
    const int wave = uniform(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    withAudioTaps<T>(sinc, [&](const auto& taps) {
        for (int item = int(blockIdx.x) * kAudioWaves + wave; item < count; item += int(gridDim.x) * kAudioWaves)
            chainAudio<T>(bank, taps, spans, jobs, points, mods, phase, filtered, states, segFirst, segEnd,
                          uniform(heads[item]), lane);
    });
}

// Filter pass: one wave per voice chain of the round, grid-stride; 4 chains per block.
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
        voiceChain(global(u.spans), global(u.cold), global(u.points), global(u.mods), global(u.filtered),
                   global(u.states), u.segFirst, u.segEnd, uniform(global(u.chainHeads)[item - u.chainBase]),
                   header.silence, carries[wave], lane);
    }
}

// Filter pass, plain form: one unit, its pointers as kernel arguments, one wave per chain.
__global__ void __launch_bounds__(64 * kChainWaves) __attribute__((amdgpu_waves_per_eu(kChainOccupancy)))
clayVoiceChains(const Span* __restrict__ spans, const SpanCold* __restrict__ cold, const int* __restrict__ heads,
                const clay_control_point* __restrict__ points, const Mod* __restrict__ mods,
                float* __restrict__ filtered, VoiceState* __restrict__ states, int segFirst, int segEnd, int count,
                float silenceLevel)
{
    __shared__ ChainCarry carries[kChainWaves];
    const int wave = uniform(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const int item = int(blockIdx.x) * kChainWaves + wave;
    if (item >= count)
        return;  // uniform per wave; no block barriers below
    voiceChain(spans, cold, points, mods, filtered, states, segFirst, segEnd, uniform(heads[item]), silenceLevel,
               carries[wave], lane);
}

// Per output frame: the dry voices of the tile rendered inline, times the job
// gain, plus the filtered voices from scratch (pan applied for mono). Voices
// are summed in span order, so every execution form yields the same bits
// (src/cuda/render.cu:458-523). Natural writes float2, orig/custom short2.
template<int T, bool FloatOut>
__device__ __forceinline__ void mixFrame(const int16_t* __restrict__ bank, const float* __restrict__ sinc,
                                         const Span* __restrict__ spans, const Job* __restrict__ jobs,
                                         const int* __restrict__ voices, const clay_control_point* __restrict__ points,
                                         const Mod* __restrict__ mods, const float* __restrict__ filtered,
                                         const VoiceState* __restrict__ states, void* __restrict__ out, int segFirst,
                                         int segEnd, const Tile& tile, int lane)
{
    const int frame = tile.frame + lane;
    if (lane >= tile.frames || frame < segFirst || frame >= segEnd)
        return;
    const Job& j = jobs[uniform(tile.job)];
    const float pan = controlAt(points, j.controls[2], j.controls[3], frame);
    float left = 0, right = 0, wetLeft = 0, wetRight = 0;
    for (int i = tile.first; i < tile.first + tile.count; ++i) {
        const Span& v = spans[uniform(voices[i])];
        if (frame < v.begin || frame >= v.end)
            continue;
        if (v.filterTrack >= 0) {
            const VoiceState& st = states[v.filterTrack];
            if (st.retired && frame >= st.silence)
                continue;  // retired under the voice-silence policy
            const float* src = filtered + v.filtered + uint64_t(frame - max(v.begin, segFirst)) * uint64_t(v.channels);
            float2 s = make_float2(src[0], v.channels == 1 ? src[0] : src[1]);
            if (v.channels == 1) {
                float offset = pan;
                for (int k = v.firstMod; k < v.firstMod + v.modCount; ++k) {
                    const Mod m = mods[k];
                    if (m.target == CLAY_MOD_PAN)
                        offset += modulationAt(m, points, frame);
                }
                const float2 g = panGains(v, offset);
                s.x *= g.x;
                s.y *= g.y;
            }
            wetLeft += s.x;
            wetRight += s.y;
        } else {
            const float2 s = voiceAudio<T>(bank, GlobalTaps<T>{sinc}, v, points, mods, pan, frame, false, -1.0);
            left += s.x;
            right += s.y;
        }
    }
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

// One block per tile of the round (256 frames, one per thread), grid-stride.
template<int T, bool FloatOut>
__global__ void __launch_bounds__(kTileFrames) __attribute__((amdgpu_waves_per_eu(kMixOccupancy)))
clayRoundMixTiles(const int16_t* __restrict__ bank, const float* __restrict__ sinc, UnitsPtr units, RoundPtr round)
{
    const RoundHeader header = *round;
    for (int item = int(blockIdx.x); item < header.tiles; item += int(gridDim.x)) {
        const int unit = unitOf(header, item, [&](int i) { return units[i].tileBase; });
        const auto& u = constant(units)[unit];
        const Tile tile = global(u.tiles)[u.tileFirst + item - u.tileBase];
        mixFrame<T, FloatOut>(bank, sinc, global(u.spans), global(u.jobs), global(u.voices), global(u.points),
                              global(u.mods), global(u.filtered), global(u.states), global(u.mix), u.segFirst,
                              u.segEnd, tile, int(threadIdx.x));
    }
}

// Plain form: one unit, its pointers as kernel arguments, one block per tile.
template<int T, bool FloatOut>
__global__ void __launch_bounds__(kTileFrames) __attribute__((amdgpu_waves_per_eu(kMixOccupancy))) clayMixTiles(
    const int16_t* __restrict__ bank, const float* __restrict__ sinc, const Span* __restrict__ spans,
    const Job* __restrict__ jobs, const Tile* __restrict__ tiles, const int* __restrict__ voices,
    const clay_control_point* __restrict__ points, const Mod* __restrict__ mods, const float* __restrict__ filtered,
    const VoiceState* __restrict__ states, void* __restrict__ out, int segFirst, int segEnd, int firstTile)
{
    const Tile tile = tiles[firstTile + int(blockIdx.x)];
    mixFrame<T, FloatOut>(bank, sinc, spans, jobs, voices, points, mods, filtered, states, out, segFirst, segEnd,
                          tile, int(threadIdx.x));
}

#define CLAY_INSTANTIATE_TAPS(T)                                                                                       \
    template __global__ void clayRoundChainAudio<T>(const int16_t*, const float*, UnitsPtr, RoundPtr);                 \
    template __global__ void clayRoundMixTiles<T, true>(const int16_t*, const float*, UnitsPtr, RoundPtr);             \
    template __global__ void clayRoundMixTiles<T, false>(const int16_t*, const float*, UnitsPtr, RoundPtr);            \
    template __global__ void clayChainAudio<T>(const int16_t*, const float*, const Span*, const Job*, const int*,      \
                                               const clay_control_point*, const Mod*, const double*, float*,           \
                                               const VoiceState*, int, int, int);                                      \
    template __global__ void clayMixTiles<T, true>(const int16_t*, const float*, const Span*, const Job*, const Tile*, \
                                                   const int*, const clay_control_point*, const Mod*, const float*,     \
                                                   const VoiceState*, void*, int, int, int);                            \
    template __global__ void clayMixTiles<T, false>(const int16_t*, const float*, const Span*, const Job*,             \
                                                    const Tile*, const int*, const clay_control_point*, const Mod*,     \
                                                    const float*, const VoiceState*, void*, int, int, int);
CLAY_INSTANTIATE_TAPS(16)
CLAY_INSTANTIATE_TAPS(32)
CLAY_INSTANTIATE_TAPS(48)
CLAY_INSTANTIATE_TAPS(64)
