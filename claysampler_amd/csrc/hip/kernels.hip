// HIP kernels of the ClaySampler render core for gfx950 (MI355X).
//
// Render (per segment):  pitchChains -> filterSources -> filterChains -> mixTiles
// Natural finishing:     scanStereo -> selectNative -> quantizeMono -> selectTrim
// Every policy:          layoutPcm -> finalizePcm
// Quality / groups:      measureBody -> measureOnsets -> summarizeAudio
//                        -> matchVolume -> layoutAccepted -> packAccepted
//
// Reference behaviour: src/cuda/render.cu, tail.cu, quality.cu. The kernels
// here are written for wave64: reductions use 64-lane shuffles, block
// reductions combine at most 16 waves through LDS, and per-voice parameters
// are read through wave-uniform (scalar) loads.
#include <hip/hip_runtime.h>

#include "device_types.h"

using namespace clay_dev;

namespace {

constexpr float kQuarterPi = .7853981633974483f;

// ---------------------------------------------------------------- envelopes
// src/cuda/render.cu:6-19
__device__ __forceinline__ float envelopeSpan(const Span& v, int frame)
{
    if (v.releaseStart >= 0 && frame >= v.releaseStart) {
        if (frame >= v.fadeStart)
            return fmaxf(0.f, v.fadeValue - float(frame - v.fadeStart + 1) * v.fadeStep);
        return v.releaseValue * expf(float(frame - v.releaseStart + 1) * v.releaseRate);
    }
    if (frame < v.attackStart)
        return v.start;
    if (frame < v.holdStart)
        return fminf(1.f, v.start + float(frame - v.attackStart + 1) * v.attackStep);
    if (frame < v.decayStart)
        return 1.f;
    if (v.decayRate == 0)
        return v.sustain;
    return fmaxf(v.sustain, expf(float(frame - v.decayStart + 1) * v.decayRate));
}

// ---------------------------------------------------------------- controls
// Last point at or before `frame` (the first point when none), then linear
// (src/cuda/render.cu:107-118).
__device__ __forceinline__ float controlAt(const clay_control_point* __restrict__ points, int lo, int hi, int frame)
{
    while (lo + 1 < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (points[mid].frame <= frame)
            lo = mid;
        else
            hi = mid;
    }
    const clay_control_point p = points[lo];
    return fmaf(p.step, float(frame - p.frame), p.value);
}

__device__ __forceinline__ float modulationAt(const Mod& m, const clay_control_point* __restrict__ points, int frame)
{
    float v = m.first < 0 ? m.depth : m.depth * controlAt(points, m.first, m.last, frame);
    if (m.shapeFirst >= 0)
        v *= controlAt(points, m.shapeFirst, m.shapeLast, frame);
    return v;
}

__device__ __forceinline__ float jobGain(const clay_control_point* __restrict__ points, const Job& j, int frame)
{
    return controlAt(points, j.controls[0], j.controls[1], frame) * controlAt(points, j.controls[1], j.controls[2], frame);
}

// ---------------------------------------------------------------- resampling
// Source frame of a span, zero outside [0, sourceEnd) (src/cuda/render.cu:21-28).
// Values are raw int16 (scaled by 2^-15 by the caller).
__device__ __forceinline__ float2 fetchRaw(const int16_t* __restrict__ bank, const Span& v, int frame)
{
    if (frame < 0 || frame >= v.sourceEnd)
        return make_float2(0.f, 0.f);
    const uint64_t i = v.sample + uint64_t(frame) * uint64_t(v.channels);
    const float left = float(bank[i]);
    return make_float2(left, v.channels == 1 ? left : float(bank[i + 1]));
}

// Windowed-sinc tap sum at the nearest of 2048 phases (src/cuda/render.cu:30-61).
// Accumulating the raw int16 values and scaling once by 2^-15 gives the same
// float result as weighting the pre-scaled float bank of the reference.
template<int T>
__device__ __forceinline__ float2 interpolateAt(const int16_t* __restrict__ bank, const float* __restrict__ sinc,
                                                const Span& v, double pos)
{
    constexpr float scale = 1.f / 32768.f;
    const double floorPos = floor(pos);
    const int centre = int(floorPos);
    if (pos == floorPos) {
        const float2 s = fetchRaw(bank, v, centre);
        return make_float2(s.x * scale, s.y * scale);
    }
    const int phase = __double2int_rn((pos - floorPos) * kPhases);
    const float4* __restrict__ w4 = reinterpret_cast<const float4*>(sinc + size_t(phase) * T);
    constexpr int first = T / 2 - 1;
    float a0 = 0.f, a1 = 0.f, b0 = 0.f, b1 = 0.f;
    if (centre >= first && centre < v.sourceEnd - T / 2) {
        const int16_t* __restrict__ src = bank + v.sample + uint64_t(centre - first) * uint64_t(v.channels);
        if (v.channels == 1) {
#pragma unroll
            for (int q = 0; q < T / 4; ++q) {
                const float4 w = w4[q];
                a0 = fmaf(w.x, float(src[4 * q + 0]), a0);
                a1 = fmaf(w.y, float(src[4 * q + 1]), a1);
                a0 = fmaf(w.z, float(src[4 * q + 2]), a0);
                a1 = fmaf(w.w, float(src[4 * q + 3]), a1);
            }
            const float m = (a0 + a1) * scale;
            return make_float2(m, m);
        }
#pragma unroll
        for (int q = 0; q < T / 4; ++q) {
            const float4 w = w4[q];
            a0 = fmaf(w.x, float(src[8 * q + 0]), a0);
            b0 = fmaf(w.x, float(src[8 * q + 1]), b0);
            a1 = fmaf(w.y, float(src[8 * q + 2]), a1);
            b1 = fmaf(w.y, float(src[8 * q + 3]), b1);
            a0 = fmaf(w.z, float(src[8 * q + 4]), a0);
            b0 = fmaf(w.z, float(src[8 * q + 5]), b0);
            a1 = fmaf(w.w, float(src[8 * q + 6]), a1);
            b1 = fmaf(w.w, float(src[8 * q + 7]), b1);
        }
        return make_float2((a0 + a1) * scale, (b0 + b1) * scale);
    }
    for (int q = 0; q < T / 4; ++q) {
        const float4 w = w4[q];
        const float ws[4] = {w.x, w.y, w.z, w.w};
        for (int k = 0; k < 4; ++k) {
            const float2 s = fetchRaw(bank, v, centre - first + 4 * q + k);
            a0 = fmaf(ws[k], s.x, a0);
            b0 = fmaf(ws[k], s.y, b0);
        }
    }
    return make_float2(a0 * scale, b0 * scale);
}

// Source position, generators, loop wrap and loop crossfade (src/cuda/render.cu:63-100).
// The position arithmetic uses explicitly rounded double operations so that
// the phase index equals the one of a plain IEEE double evaluation.
template<int T>
__device__ float2 interpolate(const int16_t* __restrict__ bank, const float* __restrict__ sinc, const Span& v,
                              int frame, double pos)
{
    if (pos < 0) {
        const int elapsed = frame - v.begin;
        double distance = elapsed;
        if (v.pitchSlope != 0)
            distance = __ddiv_rn(expm1(__dmul_rn(v.pitchSlope, double(elapsed))), expm1(v.pitchSlope));
        pos = __dadd_rn(v.position, __dmul_rn(distance, v.step));
    }
    if (v.generator != CLAY_GEN_NONE) {
        if (v.generator == CLAY_GEN_NOISE) {
            unsigned bits = v.noiseSeed + unsigned(frame - v.noiseStart);
            bits = (bits ^ (bits >> 16)) * 0x7feb352du;
            bits = (bits ^ (bits >> 15)) * 0x846ca68bu;
            bits ^= bits >> 16;
            const float x = float(bits >> 8) * (2.f / 16777216.f) - 1.f;
            return make_float2(x, x);
        }
        const float x = v.generator == CLAY_GEN_SINE ? -sinf(6.28318530717958647692f * float(pos - floor(pos))) : 0.f;
        return make_float2(x, x);
    }
    const int period = v.loopEnd - v.loopFirst;
    if (v.loopEnd > 0 && pos >= v.loopEnd)
        pos = __dadd_rn(double(v.loopFirst), fmod(__dsub_rn(pos, double(v.loopFirst)), double(period)));
    if (pos >= v.sourceEnd)
        return make_float2(0.f, 0.f);
    float2 s = interpolateAt<T>(bank, sinc, v, pos);
    const int fadeStart = v.loopEnd - v.loopCrossfade;
    if (v.loopEnd > 0 && v.loopCrossfade > 0 && pos >= fadeStart) {
        const float2 entry = interpolateAt<T>(bank, sinc, v, __dsub_rn(pos, double(period)));
        const float phase = float((pos - fadeStart) / v.loopCrossfade);
        const float blend = .5f - .5f * cosf(3.14159265358979323846f * phase);
        s.x = fmaf(blend, entry.x - s.x, s.x);
        s.y = fmaf(blend, entry.y - s.y, s.y);
    }
    return s;
}

// src/cuda/render.cu:130-136
__device__ __forceinline__ float2 panGains(const Span& v, float offset)
{
    if (offset == 0)
        return make_float2(v.left, v.right);
    const float p = fminf(1.f, fmaxf(-1.f, v.pan + offset));
    return make_float2(v.gain * cosf((1 + p) * kQuarterPi), v.gain * cosf((1 - p) * kQuarterPi));
}

// Envelope, modulation, resampling, pan and stereo matrix of one voice frame
// (src/cuda/render.cu:138-187).
template<int T>
__device__ float2 voiceAudio(const int16_t* __restrict__ bank, const float* __restrict__ sinc, const Span& v,
                             const clay_control_point* __restrict__ points, const Mod* __restrict__ mods, float pan,
                             int frame, bool filterInput, double phase)
{
    float level = envelopeSpan(v, frame);
    if (level <= 0)
        return make_float2(0.f, 0.f);
    float db = 0, widthOffset = 0, positionOffset = 0;
    for (int j = v.firstMod; j < v.firstMod + v.modCount; ++j) {
        const Mod m = mods[j];
        if (m.target > CLAY_MOD_POSITION)
            continue;
        const float x = modulationAt(m, points, frame);
        switch (m.target) {
        case CLAY_MOD_AMPLITUDE: level *= x; break;
        case CLAY_MOD_VOLUME: db += x; break;
        case CLAY_MOD_PAN: pan += x; break;
        case CLAY_MOD_WIDTH: widthOffset += x; break;
        case CLAY_MOD_POSITION: positionOffset += x; break;
        default: break;
        }
    }
    if (db != 0)
        level *= exp2f(db * .166096404744368f);
    const float2 s = interpolate<T>(bank, sinc, v, frame, phase);
    if (filterInput && v.channels == 1)
        return make_float2(s.x * level, s.x * level);
    const float2 g = panGains(v, pan);
    const float l = s.x * g.x, r = s.y * g.y;
    if (v.stereoMatrix) {
        float direct = v.widthDirect, cross = v.widthCross;
        if (widthOffset != 0) {
            const float w = fminf(1.f, fmaxf(-1.f, v.width + widthOffset));
            direct = cosf((1 - w) * kQuarterPi);
            cross = cosf((1 + w) * kQuarterPi);
        }
        float pl = v.positionLeft, pr = v.positionRight;
        if (positionOffset != 0) {
            const float p = fminf(1.f, fmaxf(-1.f, v.stereoPosition + positionOffset));
            pl = cosf((1 + p) * kQuarterPi);
            pr = cosf((1 - p) * kQuarterPi);
        }
        return make_float2(fmaf(r, cross, l * direct) * (level * pl), fmaf(l, cross, r * direct) * (level * pr));
    }
    return make_float2(l * level, r * level);
}

// ---------------------------------------------------------------- filters
// Biquad sections per filter slot (src/cuda/render.cu:312-319).
__device__ __forceinline__ int sections(int type)
{
    if (type == CLAY_FILTER_NONE) return 0;
    if (type == CLAY_FILTER_PEAK) return 1;
    if (type >= CLAY_FILTER_LOW6) return 3;
    if (type >= CLAY_FILTER_LOW4) return 2;
    return 1;
}

// ---------------------------------------------------------------- reductions
template<class T, class Op>
__device__ __forceinline__ T waveReduce(T v, Op op)
{
#pragma unroll
    for (int d = 32; d; d >>= 1)
        v = op(v, __shfl_xor(v, d, 64));
    return v;
}

// Block reduction for up to 1024 threads; every thread gets the result.
template<class T, class Op>
__device__ T blockReduce(T v, T identity, Op op)
{
    __shared__ T partial[16];
    v = waveReduce(v, op);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        partial[wave] = v;
    __syncthreads();
    const int waves = (blockDim.x + 63) >> 6;
    T r = identity;
    for (int i = 0; i < waves; ++i)
        r = op(r, partial[i]);
    __syncthreads();
    return r;
}

struct Add {
    template<class T> __device__ T operator()(T a, T b) const { return a + b; }
};
struct Max {
    template<class T> __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};
struct Min {
    template<class T> __device__ T operator()(T a, T b) const { return a < b ? a : b; }
};

// Index of the last entry whose `first` is <= block (src/cuda/tail.cu:32-44).
template<class F>
__device__ __forceinline__ int ownerOf(int count, int block, F firstOf)
{
    int lo = 0, hi = count;
    while (lo + 1 < hi) {
        const int mid = (lo + hi) / 2;
        if (firstOf(mid) <= block)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

} // namespace

// ====================================================================== render

// ----------------------------------------------------------------- FP64 voice-chain recurrences
namespace {

// Control track value in FP64 (src/cuda/render.cu:107-118 evaluated in double).
__device__ __forceinline__ double pointValue(const clay_control_point& p, int frame)
{
    return fma(double(p.step), double(frame - p.frame), double(p.value));
}

// Sequential reader of one control track: frames only move forward inside a
// chain, so the segment is found by advancing instead of a binary search.
// It lands on the same point as controlAt's search (last point at or before
// the frame, the first point when none).
struct Cursor {
    int at, last;
    __device__ void reset(const clay_control_point* __restrict__ points, int first, int end, int frame)
    {
        at = first;
        last = end;
        int lo = first, hi = end;
        while (lo + 1 < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (points[mid].frame <= frame)
                lo = mid;
            else
                hi = mid;
        }
        at = lo;
    }
    __device__ __forceinline__ double value(const clay_control_point* __restrict__ points, int frame)
    {
        while (at + 1 < last && points[at + 1].frame <= frame)
            ++at;
        return pointValue(points[at], frame);
    }
};

// Up to kCursors modulation bindings of a span read through cursors; any
// further ones fall back to the binary search.
constexpr int kCursors = 6;

struct ModReader {
    Cursor track[kCursors], shape[kCursors];
    int first, count;

    __device__ void reset(const Mod* __restrict__ mods, const clay_control_point* __restrict__ points, int firstMod,
                          int modCount, int frame)
    {
        first = firstMod;
        count = modCount;
        for (int j = 0; j < min(count, kCursors); ++j) {
            const Mod m = mods[first + j];
            if (m.first >= 0)
                track[j].reset(points, m.first, m.last, frame);
            if (m.shapeFirst >= 0)
                shape[j].reset(points, m.shapeFirst, m.shapeLast, frame);
        }
    }

    // src/cuda/render.cu:120-126 in double.
    __device__ double value(const Mod* __restrict__ mods, const clay_control_point* __restrict__ points, int j, int frame)
    {
        const Mod m = mods[first + j];
        double v = m.depth;
        if (m.first >= 0) {
            const double x = j < kCursors ? track[j].value(points, frame) : [&] {
                Cursor c;
                c.reset(points, m.first, m.last, frame);
                return pointValue(points[c.at], frame);
            }();
            v = double(m.depth) * x;
        }
        if (m.shapeFirst >= 0) {
            const double x = j < kCursors ? shape[j].value(points, frame) : [&] {
                Cursor c;
                c.reset(points, m.shapeFirst, m.shapeLast, frame);
                return pointValue(points[c.at], frame);
            }();
            v *= x;
        }
        return v;
    }
};

__device__ __forceinline__ double envelopeShapeD(const clay_envelope_shape& v, int frame)
{
    if (v.release_start >= 0 && frame >= v.release_start) {
        if (frame >= v.fade_start)
            return fmax(0.0, double(v.fade_value) - double(frame - v.fade_start + 1) * double(v.fade_step));
        return double(v.release_value) * exp(double(frame - v.release_start + 1) * double(v.release_rate));
    }
    if (frame < v.attack_start)
        return v.start;
    if (frame < v.hold_start)
        return fmin(1.0, double(v.start) + double(frame - v.attack_start + 1) * double(v.attack_step));
    if (frame < v.decay_start)
        return 1.0;
    if (v.decay_rate == 0)
        return v.sustain;
    return fmax(double(v.sustain), exp(double(frame - v.decay_start + 1) * double(v.decay_rate)));
}

struct CoeffD {
    double b0, b1, b2, a1, a2;
};

// src/cuda/render.cu:278-310 in double.
__device__ CoeffD coefficientsD(int type, double cutoff, double resonance, int rate, double bandwidth, double gain)
{
    constexpr double pi = 3.14159265358979323846;
    const double w = 2 * pi * fmin(fmin(20000.0, rate * .499), fmax(1.0, cutoff)) / rate;
    if (type == CLAY_FILTER_LOW1 || type == CLAY_FILTER_HIGH1) {
        const double p = exp(-w);
        if (type == CLAY_FILTER_LOW1)
            return {1 - p, 0, 0, -p, 0};
        const double h = (1 + p) * .5;
        return {h, -h, 0, -p, 0};
    }
    double s, c;
    sincos(w, &s, &c);
    if (type == CLAY_FILTER_PEAK) {
        const double a = exp10(fmin(60.0, fmax(-96.0, gain)) / 40);
        const double width = fmin(12.0, fmax(.01, bandwidth));
        const double alpha = s * sinh(fmin(20.0, 0.34657359027997264 * width * w / s));
        const double inv = 1 / (1 + alpha / a);
        const double mid = -2 * c * inv;
        return {(1 + alpha * a) * inv, mid, (1 - alpha * a) * inv, mid, (1 - alpha / a) * inv};
    }
    const double alpha = .5 * s * exp10(-fmin(60.0, fmax(-60.0, resonance)) / 20);
    const double inv = 1 / (1 + alpha);
    const double a1 = -2 * c * inv, a2 = (1 - alpha) * inv;
    if (type == CLAY_FILTER_BAND2 || type == CLAY_FILTER_BAND4 || type == CLAY_FILTER_BAND6)
        return {alpha * inv, 0, -alpha * inv, a1, a2};
    if (type == CLAY_FILTER_NOTCH2)
        return {inv, a1, inv, a1, a2};
    const bool high = type == CLAY_FILTER_HIGH2 || type == CLAY_FILTER_HIGH4 || type == CLAY_FILTER_HIGH6;
    const double b = (1 + (high ? c : -c)) * .5 * inv;
    return {b, (high ? -2 : 2) * b, b, a1, a2};
}

} // namespace

// Variable-pitch recurrence per voice chain: per-frame source phase
// (src/cuda/render.cu:222-266), in FP64. One lane per chain.
extern "C" __global__ void clayPitchChains(const Span* __restrict__ spans, const SpanCold* __restrict__ cold,
                                           const int* __restrict__ heads, const clay_control_point* __restrict__ points,
                                           const Mod* __restrict__ mods, double* __restrict__ phase,
                                           VoiceState* __restrict__ states, int segFirst, int segEnd, int count)
{
    const int track = blockIdx.x * blockDim.x + threadIdx.x;
    if (track >= count)
        return;
    const int head = heads[track];
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

// ----------------------------------------------------------------- wave-parallel voice chains
namespace {

constexpr int kMaxSections = 9;   // 2 filters x up to 3 sections + 3 EQ bands
constexpr int kChainWaves = 4;    // voice chains per 256-thread block

// Per-wave carry between 64-frame chunks, in LDS (broadcast reads).
struct ChainCarry {
    double z[kMaxSections][2][2];        // [section][channel][z1, z2]
    double coef[CLAY_FILTER_SLOTS][5];   // current (smoothed) coefficients
    double target[CLAY_FILTER_SLOTS][5]; // last target
    double input[CLAY_FILTER_SLOTS][3];  // inputs of the last target
    int have[CLAY_FILTER_SLOTS];
};

__device__ __forceinline__ double up(double v, int d) { return __shfl_up(v, d, 64); }

__device__ __forceinline__ double laneValue(double v, int lane) { return __shfl(v, lane, 64); }

} // namespace

// One wave per voice chain (a voice's spans with filters and/or variable pitch),
// 64 output frames per step, chunks aligned to absolute frame multiples of 64:
//   1. the 64 frames of voice audio before pan (mono) times the job gain, one per lane;
//   2. filter targets where a lane's inputs differ from the previous frame's,
//      copied forward to the following lanes (src/cuda/render.cu:376-406);
//   3. coefficient smoothing c_n = c_{n-1} + a (t_n - c_{n-1}) as an inclusive
//      scan of affine maps c -> M c + V over the 64 lanes;
//   4. each biquad section (direct form II transposed, render.cu:330-336) as a
//      scan of 2x2 affine maps s_n = A_n s_{n-1} + B_n u_n, in FP64;
//   5. the filtered frames written to the float scratch the mix kernel reads.
// State carries across chunks in LDS and across segments in VoiceState.
template<int T>
__global__ void __launch_bounds__(64 * kChainWaves) clayVoiceChains(
    const int16_t* __restrict__ bank, const float* __restrict__ sinc, const Span* __restrict__ spans,
    const SpanCold* __restrict__ cold, const Job* __restrict__ jobs, const int* __restrict__ heads,
    const clay_control_point* __restrict__ points, const Mod* __restrict__ mods, const double* __restrict__ phase,
    float* __restrict__ filtered, VoiceState* __restrict__ states, int segFirst, int segEnd, int count)
{
    __shared__ ChainCarry carries[kChainWaves];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int chain = blockIdx.x * kChainWaves + wave;
    if (chain >= count)
        return;  // uniform per wave; no block barriers below
    ChainCarry& carry = carries[wave];
    const int head = uniform(heads[chain]);
    VoiceState& st = states[spans[head].filterTrack];
    int sectionCount[CLAY_FILTER_SLOTS], total = 0;
    for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
        sectionCount[f] = sections(cold[head].filters[f].type);
        total += sectionCount[f];
    }
    const bool hasFilter = total > 0;
    bool initialized = st.filterInitialized != 0;
    if (hasFilter && lane == 0) {
        int k = 0;
        for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
            for (int i = 0; i < 5; ++i)
                carry.coef[f][i] = st.coefficients[f][i];
            carry.have[f] = 0;
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
        const Job& j = jobs[v.job];
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
        for (int f0 = first & ~63; f0 < stop; f0 += 64) {
            const int frame = f0 + lane;
            const bool active = frame >= first && frame < stop;
            const int lo = max(first, f0) - f0, hi = min(stop, f0 + 64) - f0;  // active lanes [lo, hi)
            float xl = 0.f, xr = 0.f;
            if (active) {
                const int local = frame - first;
                const double ph = v.variablePitch ? phase[v.phase + uint64_t(local)] : -1.0;
                const float pan = controlAt(points, j.controls[2], j.controls[3], frame);
                const float2 a = voiceAudio<T>(bank, sinc, v, points, mods, pan, frame, true, ph);
                const float g = jobGain(points, j, frame);
                xl = a.x * g;
                xr = a.y * g;
            }
            float* dst = filtered + v.filtered + uint64_t(max(frame - first, 0)) * uint64_t(v.channels);
            if (!hasFilter) {
                if (active) {
                    dst[0] = xl;
                    if (stereo)
                        dst[1] = xr;
                }
                continue;
            }
            // Filter inputs of this lane's frame.
            const bool recompute = moving || !initialized;
            double in[CLAY_FILTER_SLOTS][3] = {};
            if (recompute && active) {
                double cents[2] = {0, 0}, db[2] = {0, 0}, eq[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                double depth = cv.extra[0].depth;
                for (int k = 0; k < v.modCount; ++k) {
                    const int t = mods[v.firstMod + k].target;
                    if (t == CLAY_MOD_CUTOFF) cents[0] += reader.value(mods, points, k, frame);
                    else if (t == CLAY_MOD_CUTOFF2) cents[1] += reader.value(mods, points, k, frame);
                    else if (t == CLAY_MOD_RESONANCE) db[0] += reader.value(mods, points, k, frame);
                    else if (t == CLAY_MOD_RESONANCE2) db[1] += reader.value(mods, points, k, frame);
                    else if (t == CLAY_MOD_FILEG_DEPTH) depth += reader.value(mods, points, k, frame);
                    else if (t >= CLAY_MOD_EQ1_FREQUENCY) {
                        const int id = t - CLAY_MOD_EQ1_FREQUENCY;
                        eq[id / 3][id % 3] += reader.value(mods, points, k, frame);
                    }
                }
                if (cv.active[0])
                    cents[0] += depth * envelopeShapeD(cv.extra[0], frame);
                for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
                    const clay_voice_filter& flt = cv.filters[f];
                    if (f < 2) {
                        in[f][0] = double(flt.cutoff) * exp2(cents[f] / 1200);
                        in[f][1] = double(flt.resonance) + db[f];
                    } else {
                        in[f][0] = double(flt.cutoff) + eq[f - 2][0];
                        in[f][1] = double(flt.bandwidth) + eq[f - 2][1];
                        in[f][2] = double(flt.gain) + eq[f - 2][2];
                    }
                }
            }
            int sec = 0;
            double ul = xl, ur = xr;
            for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
                if (!sectionCount[f])
                    continue;
                double c[5];
                if (recompute) {
                    // 2. Targets: recompute where inputs change, copy forward otherwise.
                    double p0 = up(in[f][0], 1), p1 = up(in[f][1], 1), p2 = up(in[f][2], 1);
                    bool prevValid = lane > lo;
                    if (lane == lo) {
                        p0 = carry.input[f][0], p1 = carry.input[f][1], p2 = carry.input[f][2];
                        prevValid = carry.have[f] != 0;
                    }
                    const bool flag = active && (!prevValid || in[f][0] != p0 || in[f][1] != p1 || in[f][2] != p2);
                    double t[5] = {0, 0, 0, 0, 0};
                    if (flag) {
                        const CoeffD q = f < 2 ? coefficientsD(cv.filters[f].type, in[f][0], in[f][1], rate, 0, 0)
                                               : coefficientsD(cv.filters[f].type, in[f][0], 0, rate, in[f][1], in[f][2]);
                        t[0] = q.b0, t[1] = q.b1, t[2] = q.b2, t[3] = q.a1, t[4] = q.a2;
                    }
                    int have = flag ? 1 : 0;
                    for (int d = 1; d < 64; d <<= 1) {
                        const int hp = __shfl_up(have, d, 64);
                        double tp[5];
                        for (int i = 0; i < 5; ++i)
                            tp[i] = up(t[i], d);
                        if (!have && lane >= d && hp) {
                            for (int i = 0; i < 5; ++i)
                                t[i] = tp[i];
                            have = 1;
                        }
                    }
                    if (!have)
                        for (int i = 0; i < 5; ++i)
                            t[i] = carry.target[f][i];
                    // 3. Smoothing scan: c_n = M_n c_{n-1} + V_n.
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
                    for (int d = 1; d < 64; d <<= 1) {
                        const double Mp = up(M, d);
                        double Vp[5];
                        for (int i = 0; i < 5; ++i)
                            Vp[i] = up(V[i], d);
                        if (lane >= d) {
                            for (int i = 0; i < 5; ++i)
                                V[i] = fma(M, Vp[i], V[i]);
                            M *= Mp;
                        }
                    }
                    for (int i = 0; i < 5; ++i)
                        c[i] = fma(M, carry.coef[f][i], V[i]);
                    // Carry: values of the last active lane.
                    const int last = hi - 1;
                    double nc[5], nt[5], ni[3];
                    for (int i = 0; i < 5; ++i) {
                        nc[i] = laneValue(c[i], last);
                        nt[i] = laneValue(t[i], last);
                    }
                    for (int i = 0; i < 3; ++i)
                        ni[i] = laneValue(in[f][i], last);
                    __builtin_amdgcn_wave_barrier();
                    if (lane == 0) {
                        for (int i = 0; i < 5; ++i) {
                            carry.coef[f][i] = nc[i];
                            carry.target[f][i] = nt[i];
                        }
                        for (int i = 0; i < 3; ++i)
                            carry.input[f][i] = ni[i];
                        carry.have[f] = 1;
                    }
                    __builtin_amdgcn_wave_barrier();
                } else {
                    for (int i = 0; i < 5; ++i)
                        c[i] = carry.coef[f][i];
                }
                // 4. Biquad sections of this slot: s_n = A_n s_{n-1} + B_n u_n.
                for (int k = 0; k < sectionCount[f]; ++k, ++sec) {
                    // A = [[-a1, 1], [-a2, 0]]; inactive lanes are the identity.
                    double a00 = -c[3], a01 = 1, a10 = -c[4], a11 = 0;
                    double wl0 = (c[1] - c[3] * c[0]) * ul, wl1 = (c[2] - c[4] * c[0]) * ul;
                    double wr0 = (c[1] - c[3] * c[0]) * ur, wr1 = (c[2] - c[4] * c[0]) * ur;
                    if (!active) {
                        a00 = 1, a01 = 0, a10 = 0, a11 = 1;
                        wl0 = wl1 = wr0 = wr1 = 0;
                    }
                    for (int d = 1; d < 64; d <<= 1) {
                        const double b00 = up(a00, d), b01 = up(a01, d), b10 = up(a10, d), b11 = up(a11, d);
                        const double pl0 = up(wl0, d), pl1 = up(wl1, d);
                        double pr0 = 0, pr1 = 0;
                        if (stereo) {
                            pr0 = up(wr0, d);
                            pr1 = up(wr1, d);
                        }
                        if (lane >= d) {
                            // (A, w) o (B, p) = (A B, A p + w)
                            wl0 = fma(a00, pl0, fma(a01, pl1, wl0));
                            wl1 = fma(a10, pl0, fma(a11, pl1, wl1));
                            if (stereo) {
                                wr0 = fma(a00, pr0, fma(a01, pr1, wr0));
                                wr1 = fma(a10, pr0, fma(a11, pr1, wr1));
                            }
                            const double n00 = fma(a00, b00, a01 * b10), n01 = fma(a00, b01, a01 * b11);
                            const double n10 = fma(a10, b00, a11 * b10), n11 = fma(a10, b01, a11 * b11);
                            a00 = n00, a01 = n01, a10 = n10, a11 = n11;
                        }
                    }
                    const double zl1 = carry.z[sec][0][0], zl2 = carry.z[sec][0][1];
                    const double zr1 = carry.z[sec][1][0], zr2 = carry.z[sec][1][1];
                    const double sl1 = fma(a00, zl1, fma(a01, zl2, wl0)), sl2 = fma(a10, zl1, fma(a11, zl2, wl1));
                    const double sr1 = fma(a00, zr1, fma(a01, zr2, wr0)), sr2 = fma(a10, zr1, fma(a11, zr2, wr1));
                    // y_n = b0 u_n + z1_{n-1}
                    double prevL = up(sl1, 1), prevR = up(sr1, 1);
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
            if (active) {
                dst[0] = float(ul);
                if (stereo)
                    dst[1] = float(ur);
            }
        }
    }
    // Save the chain state for the next segment.
    if (hasFilter && lane == 0) {
        int k = 0;
        for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
            for (int i = 0; i < 5; ++i)
                st.coefficients[f][i] = carry.coef[f][i];
            for (int s = 0; s < sectionCount[f]; ++s, ++k)
                for (int c = 0; c < 2; ++c) {
                    st.z1[f][s][c] = carry.z[k][c][0];
                    st.z2[f][s][c] = carry.z[k][c][1];
                }
        }
        st.filterInitialized = initialized;
    }
}

// Per output frame: the dry voices of the tile rendered inline, times the job
// gain, plus the filtered voices from scratch (pan applied for mono). Voices
// are summed in span order, so every execution form yields the same bits
// (src/cuda/render.cu:458-523). Natural writes float2, orig/custom short2.
template<int T, bool FloatOut>
__global__ void __launch_bounds__(kTileFrames) clayMixTiles(
    const int16_t* __restrict__ bank, const float* __restrict__ sinc, const Span* __restrict__ spans,
    const Job* __restrict__ jobs, const Tile* __restrict__ tiles, const int* __restrict__ voices,
    const clay_control_point* __restrict__ points, const Mod* __restrict__ mods, const float* __restrict__ filtered,
    void* __restrict__ out, int segFirst, int segEnd, int firstTile)
{
    const Tile tile = tiles[firstTile + blockIdx.x];
    const int lane = int(threadIdx.x);
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
            const float2 s = voiceAudio<T>(bank, sinc, v, points, mods, pan, frame, false, -1.0);
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

#define CLAY_INSTANTIATE_TAPS(T)                                                                                       \
    template __global__ void clayVoiceChains<T>(const int16_t*, const float*, const Span*, const SpanCold*, const Job*, \
                                                const int*, const clay_control_point*, const Mod*, const double*,      \
                                                float*, VoiceState*, int, int, int);                                   \
    template __global__ void clayMixTiles<T, true>(const int16_t*, const float*, const Span*, const Job*, const Tile*, \
                                                   const int*, const clay_control_point*, const Mod*, const float*,     \
                                                   void*, int, int, int);                                               \
    template __global__ void clayMixTiles<T, false>(const int16_t*, const float*, const Span*, const Job*,             \
                                                    const Tile*, const int*, const clay_control_point*, const Mod*,     \
                                                    const float*, void*, int, int, int);
CLAY_INSTANTIATE_TAPS(16)
CLAY_INSTANTIATE_TAPS(32)
CLAY_INSTANTIATE_TAPS(48)
CLAY_INSTANTIATE_TAPS(64)

// ====================================================================== tail

// Power and peak of each 1024-frame block of the float mix (src/cuda/tail.cu:54-81).
extern "C" __global__ void clayScanStereo(const Job* __restrict__ jobs, int count, const float2* __restrict__ mix,
                                          StereoStats* __restrict__ stats, int firstBlock)
{
    const int block = firstBlock + blockIdx.x;
    const Job j = jobs[ownerOf(count, block, [&](int i) { return jobs[i].blockFirst; })];
    const int first = (block - j.blockFirst) * kTailBlock;
    const int end = min(first + kTailBlock, j.frames);
    double power = 0;
    float peak = 0;
    for (int f = first + int(threadIdx.x); f < end; f += blockDim.x) {
        const float2 x = mix[j.mix + uint64_t(f)];
        if (!isfinite(x.x) || !isfinite(x.y)) {
            peak = __int_as_float(0x7f800000);
            continue;
        }
        peak = fmaxf(peak, fmaxf(fabsf(x.x), fabsf(x.y)));
        power += double(__fmul_rn(x.x, x.x)) + double(__fmul_rn(x.y, x.y));
    }
    power = blockReduce(power, 0.0, Add{});
    peak = blockReduce(peak, 0.f, Max{});
    if (threadIdx.x == 0)
        stats[block] = {power, peak, 0};
}

// Natural endpoint, silence-beyond-plan rule and peak repair, one lane per job
// (src/cuda/tail.cu:83-145).
extern "C" __global__ void claySelectNative(const Job* __restrict__ jobs, int count, const float2* __restrict__ mix,
                                            const StereoStats* __restrict__ stats, clay_tail_result* __restrict__ results)
{
    const int index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= count)
        return;
    const Job j = jobs[index];
    clay_tail_result r{};
    float peak = 0;
    int end = 0;
    for (int first = 0; first < j.frames; first += kTailBlock) {
        const StereoStats st = stats[j.blockFirst + first / kTailBlock];
        if (!isfinite(st.peak)) {
            r.status = CLAY_TAIL_NONFINITE;
            results[index] = r;
            return;
        }
        peak = fmaxf(peak, st.peak);
        const int blockEnd = min(j.limit, first + kTailBlock);
        if (blockEnd < j.musicEnd)
            continue;
        const double threshold = 1e-12 * (blockEnd - first) * 2;
        double power = st.power;
        // A serial re-sum decides near-ties in a fixed order.
        if (fabs(power - threshold) <= threshold * 1e-12) {
            power = 0;
            for (int f = first; f < min(blockEnd, j.frames); ++f) {
                const float2 x = mix[j.mix + uint64_t(f)];
                power += __fmul_rn(x.x, x.x);
                power += __fmul_rn(x.y, x.y);
            }
        }
        if (power <= threshold) {
            end = blockEnd;
            break;
        }
    }
    if (!end) {
        const int silent = (j.frames + kTailBlock - 1) / kTailBlock * kTailBlock;
        if (silent < j.limit)
            end = min(j.limit, max(silent + kTailBlock, (j.musicEnd + kTailBlock - 1) / kTailBlock * kTailBlock));
    }
    if (!end) {
        r.status = CLAY_TAIL_DURATION;
    } else {
        const double attenuation = peak >= 1 ? 20 * log10(double(peak)) + 1 : 0;
        if (attenuation > 5 + 1e-6) {
            r.status = CLAY_TAIL_PEAK;
        } else {
            r.frames = end;
            r.data_frames = min(end, j.frames);
            r.scale = float(32768 * pow(10., -attenuation / 20));
        }
    }
    results[index] = r;
}

// Mono PCM16 (channel rounding then ties-to-even average) and 50 ms window
// statistics (src/cuda/tail.cu:147-181).
extern "C" __global__ void clayQuantizeMono(const Job* __restrict__ jobs, int count, const float2* __restrict__ mix,
                                            const clay_tail_result* __restrict__ results, int16_t* __restrict__ mono,
                                            MonoStats* __restrict__ stats, int window, int firstBlock)
{
    const int block = firstBlock + blockIdx.x;
    const int index = ownerOf(count, block, [&](int i) { return jobs[i].windowFirst; });
    const Job j = jobs[index];
    const clay_tail_result r = results[index];
    const int first = (block - j.windowFirst) * window;
    const int end = min(first + window, r.data_frames);
    if (first >= end)
        return;
    double power = 0;
    int peak = 0;
    for (int f = first + int(threadIdx.x); f < end; f += blockDim.x) {
        const float2 x = mix[j.mix + uint64_t(f)];
        const int l = __float2int_rn(fminf(32767.f, fmaxf(-32768.f, x.x * r.scale)));
        const int rr = __float2int_rn(fminf(32767.f, fmaxf(-32768.f, x.y * r.scale)));
        const int sum = l + rr;
        const int v = (sum >> 1) + ((sum & 1) & ((sum >> 1) & 1));
        mono[j.mix + uint64_t(f)] = int16_t(v);
        power += double(v) * v;
        peak = max(peak, abs(v));
    }
    power = blockReduce(power, 0.0, Add{});
    peak = blockReduce(peak, 0, Max{});
    if (threadIdx.x == 0)
        stats[block] = {power, peak, 0};
}

// Weak-tail trim, one lane per job (src/cuda/tail.cu:183-213).
extern "C" __global__ void claySelectTrim(const Job* __restrict__ jobs, int count, const MonoStats* __restrict__ stats,
                                          clay_tail_result* __restrict__ results, TailConfig cfg)
{
    const int index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= count || results[index].status != CLAY_TAIL_OK)
        return;
    clay_tail_result& r = results[index];
    const MonoStats* w = stats + jobs[index].windowFirst;
    const int windows = (r.data_frames + cfg.window - 1) / cfg.window;
    double maxPower = 0;
    for (int i = 0; i < windows; ++i)
        maxPower = fmax(maxPower, w[i].power);
    const double threshold = fmin(cfg.ceiling, fmax(cfg.floor, sqrt(maxPower / cfg.window) * cfg.relative));
    for (int i = windows - 1; i >= 0; --i) {
        if (sqrt(w[i].power / cfg.window) > threshold || w[i].peak > threshold * 4) {
            const int lastSound = min(r.frames, (i + 1) * cfg.window);
            if (r.frames - lastSound >= cfg.quietFrames) {
                r.frames = lastSound + cfg.padding;
                r.data_frames = min(r.frames, r.data_frames);
                r.fade = min(r.frames, cfg.fade);
            }
            break;
        }
    }
}

// Output offsets: exclusive prefix of dataFrames (src/cuda/tail.cu:216-238).
// orig/custom results are initialised from the plan here.
extern "C" __global__ void clayLayoutPcm(const Job* __restrict__ jobs, int count, clay_tail_result* __restrict__ results,
                                         int fromPlan)
{
    if (threadIdx.x || blockIdx.x)
        return;
    uint64_t offset = 0;
    for (int i = 0; i < count; ++i) {
        clay_tail_result& r = results[i];
        if (fromPlan) {
            r = clay_tail_result{};
            r.frames = jobs[i].frames;
            r.data_frames = jobs[i].frames;
            r.fade = jobs[i].fade;
        }
        r.offset = offset;
        r.peak = 0;
        r.first = r.frames;
        r.last = -1;
        offset += uint64_t(r.data_frames);
    }
}

// Fade, pack, peak, audible extent and clip count per 1024-frame block
// (src/cuda/tail.cu:240-297). `channels` is 1 (natural) or 2.
extern "C" __global__ void clayFinalizePcm(const Job* __restrict__ jobs, int count, const int16_t* __restrict__ input,
                                           int16_t* __restrict__ output, clay_tail_result* __restrict__ results,
                                           int channels, int countClipped, int firstBlock)
{
    const int block = firstBlock + blockIdx.x;
    const int index = ownerOf(count, block, [&](int i) { return jobs[i].blockFirst; });
    const Job j = jobs[index];
    const uint64_t outAt = results[index].offset;
    const int frames = results[index].frames, fade = results[index].fade;
    const int first = (block - j.blockFirst) * kTailBlock;
    const int end = min(first + kTailBlock, results[index].data_frames);
    if (first >= end)
        return;
    int peak = 0, audibleFirst = frames, audibleLast = -1, clipped = 0;
    const int fadeStart = frames - fade;
    constexpr double pi = 3.14159265358979323846;
    for (int f = first + int(threadIdx.x); f < end; f += blockDim.x) {
        int framePeak = 0;
        for (int c = 0; c < channels; ++c) {
            int v = input[(j.mix + uint64_t(f)) * uint64_t(channels) + uint64_t(c)];
            if (f >= fadeStart) {
                const int i = f - fadeStart;
                const double gain = fade == 1 ? 0 : .5 * (1 + cos(pi * i / (fade - 1)));
                v = int16_t(__double2int_rn(v * gain));
            }
            output[(outAt + uint64_t(f)) * uint64_t(channels) + uint64_t(c)] = int16_t(v);
            framePeak = max(framePeak, abs(v));
            clipped += abs(v) >= 32767;
        }
        peak = max(peak, framePeak);
        if (framePeak > 3) {
            audibleFirst = min(audibleFirst, f);
            audibleLast = f;
        }
    }
    peak = blockReduce(peak, 0, Max{});
    audibleFirst = blockReduce(audibleFirst, frames, Min{});
    audibleLast = blockReduce(audibleLast, -1, Max{});
    if (countClipped) {
        clipped = blockReduce(clipped, 0, Add{});
        if (threadIdx.x == 0 && clipped)
            atomicAdd(&results[index].audio.clipped, clipped);
    }
    if (threadIdx.x == 0 && audibleLast >= 0) {
        atomicMax(&results[index].peak, peak);
        atomicMin(&results[index].first, audibleFirst);
        atomicMax(&results[index].last, audibleLast);
    }
}

// ====================================================================== quality

namespace {

__device__ __forceinline__ unsigned long long monoPower(const int16_t* pcm, int channels, int frame)
{
    const int v = channels == 1 ? 2 * int(pcm[frame]) : int(pcm[2 * frame]) + int(pcm[2 * frame + 1]);
    return (unsigned long long)((long long)v * v);
}

__device__ __forceinline__ int bodyBlockStart(int block, const clay_audio_config& a)
{
    const int chunk = 2 * a.rate, span = 5 * a.window;
    const int windows = (chunk + span - 1) / span;
    return (block / windows) * chunk + (block % windows) * span;
}

__device__ double levelOf(unsigned long long power, int window)
{
    return power ? fmax(-120., 10 * log10(double(power) / (window * 4294967296.))) : -120.;
}

} // namespace

// Sum of (L+R)^2 over 250 ms body windows (src/cuda/quality.cu:21-45).
extern "C" __global__ void clayMeasureBody(const AudioJobDesc* __restrict__ jobs, int count, const int16_t* __restrict__ pcm,
                                           const clay_tail_result* __restrict__ results,
                                           unsigned long long* __restrict__ powers, clay_audio_config cfg, int channels,
                                           int firstBlock)
{
    const int block = firstBlock + blockIdx.x;
    const int index = ownerOf(count, block, [&](int i) { return jobs[i].bodyFirst; });
    const clay_tail_result& r = results[index];
    if (r.status != CLAY_TAIL_OK)
        return;
    const AudioJobDesc j = jobs[index];
    const int w = block - j.bodyFirst;
    const int first = bodyBlockStart(w, cfg);
    const int end = min(min(j.bodyFrames, r.data_frames), bodyBlockStart(w + 1, cfg));
    const int16_t* audio = pcm + r.offset * uint64_t(channels);
    unsigned long long power = 0;
    for (int f = first + int(threadIdx.x); f < end; f += blockDim.x)
        power += monoPower(audio, channels, f);
    power = blockReduce(power, 0ull, Add{});
    if (threadIdx.x == 0)
        powers[block] = power;
}

// Loudest of five 50 ms windows after each onset; one wave per window
// (src/cuda/quality.cu:47-77). Launched with 5 waves (320 threads).
extern "C" __global__ void clayMeasureOnsets(const AudioJobDesc* __restrict__ jobs, int count, const int* __restrict__ onsets,
                                             const int16_t* __restrict__ pcm, const clay_tail_result* __restrict__ results,
                                             unsigned long long* __restrict__ powers, int window, int channels,
                                             int firstBlock)
{
    const int block = firstBlock + blockIdx.x;
    const int index = ownerOf(count, block, [&](int i) { return jobs[i].onsetFirst; });
    const clay_tail_result& r = results[index];
    if (r.status != CLAY_TAIL_OK)
        return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int first = onsets[block] + wave * window;
    const int end = min(first + window, r.data_frames);
    const int16_t* audio = pcm + r.offset * uint64_t(channels);
    unsigned long long power = 0;
    for (int f = first + lane; f < end; f += 64)
        power += monoPower(audio, channels, f);
    power = waveReduce(power, Add{});
    __shared__ unsigned long long maxima[5];
    if (lane == 0)
        maxima[wave] = power;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long best = maxima[0];
        for (int k = 1; k < 5; ++k)
            best = best > maxima[k] ? best : maxima[k];
        powers[block] = best;
    }
}

// Body RMS, silence, onset percentiles and the verdict, one lane per job
// (src/cuda/quality.cu:79-170). Onset powers are sorted in place.
extern "C" __global__ void claySummarizeAudio(const AudioJobDesc* __restrict__ jobs, int count,
                                              const unsigned long long* __restrict__ body,
                                              unsigned long long* __restrict__ onsets,
                                              clay_tail_result* __restrict__ results, clay_audio_config cfg, int channels)
{
    const int index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= count || results[index].status != CLAY_TAIL_OK)
        return;
    const AudioJobDesc j = jobs[index];
    clay_tail_result& r = results[index];
    clay_audio_metrics& st = r.audio;
    const int end = min(r.frames, j.bodyFrames);
    unsigned long long total = 0;
    int silent = 0, run = 0;
    for (int w = 0, first = 0; first < end; ++w) {
        const int next = min(end, bodyBlockStart(w + 1, cfg));
        const unsigned long long wp = body[j.bodyFirst + w];
        total += wp;
        if (cfg.enabled) {
            if (double(wp) < (next - first) * cfg.silence_power) {
                silent += next - first;
                run += next - first;
                st.longest_silence = max(run, st.longest_silence);
            } else {
                run = 0;
            }
        }
        first = next;
    }
    st.body_rms = j.bodyFrames ? sqrt(double(total) / j.bodyFrames) * .5 : 0;
    if (!cfg.enabled)
        return;
    st.measured = 1;
    st.median = st.quartile = -120;
    st.silent_fraction = double(silent) / max(1, end);
    st.clipped_fraction = double(st.clipped) / (double(max(1, r.frames)) * channels);
    if (j.onsetCount) {
        unsigned long long* v = onsets + j.onsetFirst;
        const int n = j.onsetCount;
        // Heap sort, ascending.
        const auto sift = [&](int root, int size) {
            const unsigned long long value = v[root];
            for (int child = 2 * root + 1; child < size; child = 2 * root + 1) {
                if (child + 1 < size && v[child] < v[child + 1])
                    ++child;
                if (value >= v[child])
                    break;
                v[root] = v[child];
                root = child;
            }
            v[root] = value;
        };
        for (int root = n / 2; root-- > 0;)
            sift(root, n);
        for (int last = n - 1; last > 0; --last) {
            const unsigned long long top = v[0];
            v[0] = v[last];
            v[last] = top;
            sift(0, last);
        }
        const auto percentile = [&](double fraction) {
            const double position = (n - 1) * fraction;
            const int lo = int(position);
            const double lower = levelOf(v[lo], cfg.window);
            return lo + 1 < n ? lower + (levelOf(v[lo + 1], cfg.window) - lower) * (position - lo) : lower;
        };
        st.median = percentile(.5);
        st.quartile = percentile(.25);
    }
    if (r.frames > cfg.max_frames) r.status = CLAY_TAIL_AUDIO_DURATION;
    else if (!j.onsetCount || st.median < cfg.min_median || st.quartile < cfg.min_quartile) r.status = CLAY_TAIL_QUIET_PERFORMANCE;
    else if (st.silent_fraction > cfg.max_silent_fraction || st.longest_silence > cfg.max_silent_frames) r.status = CLAY_TAIL_EXCESSIVE_SILENCE;
    else if (r.peak < cfg.min_peak) r.status = CLAY_TAIL_SILENT_AUDIO;
    else if (st.clipped_fraction > cfg.max_clipped_fraction) r.status = CLAY_TAIL_CLIPPING;
    st.verdict = r.status;
}

// Shared gain per volume group (src/cuda/quality.cu:172-209). `groups` holds
// [first, end) batch job ranges.
extern "C" __global__ void clayMatchVolume(const int2* __restrict__ groups, int count, int minimum,
                                           clay_tail_result* __restrict__ results)
{
    const int index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= count)
        return;
    const int2 g = groups[index];
    double desired = 0, safe = 1e30;
    int accepted = 0;
    for (int i = g.x; i < g.y; ++i) {
        clay_tail_result& r = results[i];
        if (r.status != CLAY_TAIL_OK)
            continue;
        const double rms = r.audio.body_rms;
        if (rms <= 0 || r.last < 0) {
            r.status = CLAY_TAIL_SILENT_BODY;
            continue;
        }
        desired += rms;
        safe = fmin(safe, 32767 * .8912509381337456 * rms / max(1, r.peak));
        ++accepted;
    }
    const double target = fmin(desired / max(1, accepted), safe);
    for (int i = g.x; i < g.y; ++i) {
        clay_tail_result& r = results[i];
        if (r.status != CLAY_TAIL_OK)
            continue;
        if (accepted < minimum) {
            r.status = CLAY_TAIL_INCOMPLETE_GROUP;
        } else {
            r.audio.target_rms = target;
            r.audio.gain = target / r.audio.body_rms;
        }
    }
}

// Compaction offsets of accepted jobs (src/cuda/quality.cu:211-232).
extern "C" __global__ void clayLayoutAccepted(const Job* __restrict__ jobs, int count, clay_tail_result* __restrict__ results,
                                              PcmJob* __restrict__ pcm)
{
    if (threadIdx.x || blockIdx.x)
        return;
    uint64_t offset = 0;
    for (int i = 0; i < count; ++i) {
        clay_tail_result& r = results[i];
        pcm[i] = {r.offset, jobs[i].blockFirst, 0};
        if (r.status != CLAY_TAIL_OK || r.last < 0)
            r.data_frames = 0;
        r.offset = offset;
        offset += uint64_t(r.data_frames);
        if (r.audio.gain > 0) {
            r.peak = 0;
            r.first = r.frames;
            r.last = -1;
        }
    }
}

// Group gain applied while packing accepted PCM (src/cuda/quality.cu:234-280).
extern "C" __global__ void clayPackAccepted(const PcmJob* __restrict__ jobs, int count, const int16_t* __restrict__ source,
                                            int16_t* __restrict__ output, clay_tail_result* __restrict__ results,
                                            int channels, int firstBlock)
{
    const int block = firstBlock + blockIdx.x;
    const int index = ownerOf(count, block, [&](int i) { return jobs[i].blockFirst; });
    const PcmJob j = jobs[index];
    const clay_tail_result& r = results[index];
    const int first = (block - j.blockFirst) * kTailBlock;
    const int end = min(first + kTailBlock, r.data_frames);
    if (first >= end)
        return;
    const double gain = r.audio.gain;
    const int16_t* in = source + j.sourceOffset * uint64_t(channels);
    int16_t* out = output + r.offset * uint64_t(channels);
    int peak = 0, audibleFirst = r.frames, audibleLast = -1;
    for (int f = first + int(threadIdx.x); f < end; f += blockDim.x) {
        int framePeak = 0;
        for (int c = 0; c < channels; ++c) {
            const int i = f * channels + c;
            const int16_t v = gain > 0 ? int16_t(__double2int_rn(in[i] * gain)) : in[i];
            out[i] = v;
            framePeak = max(framePeak, abs(int(v)));
        }
        peak = max(peak, framePeak);
        if (framePeak > 3) {
            audibleFirst = min(audibleFirst, f);
            audibleLast = f;
        }
    }
    if (gain > 0) {
        peak = blockReduce(peak, 0, Max{});
        audibleFirst = blockReduce(audibleFirst, r.frames, Min{});
        audibleLast = blockReduce(audibleLast, -1, Max{});
        if (threadIdx.x == 0) {
            atomicMax(&results[index].peak, peak);
            if (audibleLast >= 0) {
                atomicMin(&results[index].first, audibleFirst);
                atomicMax(&results[index].last, audibleLast);
            }
        }
    }
}
