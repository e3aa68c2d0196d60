// Device helpers shared by the render and finishing kernels (gfx950).
// Reference behaviour: src/cuda/render.cu:1-336, tail.cu:1-51, quality.cu:1-19.
#pragma once

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
__device__ __forceinline__ int controlSegment(const clay_control_point* __restrict__ points, int lo, int hi, int frame)
{
    while (lo + 1 < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (points[mid].frame <= frame)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ float pointAt(const clay_control_point& p, int frame)
{
    return fmaf(p.step, float(frame - p.frame), p.value);
}

__device__ __forceinline__ float controlAt(const clay_control_point* __restrict__ points, int lo, int hi, int frame)
{
    return pointAt(points[controlSegment(points, lo, hi, frame)], frame);
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

// Sinc weights of one phase, four taps at a time, from the phase-major table
// in global memory ((kPhases + 1) rows of T floats).
template<int T>
struct GlobalTaps {
    const float* __restrict__ sinc;
    struct Row {
        const float4* __restrict__ w;
        __device__ __forceinline__ float4 operator[](int q) const { return w[q]; }
        // The row as stored, and the sample pair at mono position pair q: every
        // row is stored, so no lane reads a mirror.
        __device__ __forceinline__ float4 stored(int q) const { return w[q]; }
        __device__ __forceinline__ uint32_t pair(uint32_t, uint32_t kept) const { return kept; }
    };
    __device__ __forceinline__ Row row(int phase) const
    {
        return {reinterpret_cast<const float4*>(sinc + size_t(phase) * T)};
    }
};

// The same weights from the half table in LDS. The table is mirror-symmetric
// bit for bit (weight(tap, p) == weight(T - 1 - tap, kPhases - p): the window
// and sinc are even functions of an exactly negated argument; the renderer
// checks it when it builds the table), so rows 0..kPhases/2 hold every phase
// and a row p > kPhases/2 is row kPhases - p read backwards. Rows are padded
// by four floats so that rows of different lanes start on different banks.
template<int T>
struct SharedTaps {
    static constexpr int kPitch = T + 4;
    static constexpr int kRows = kPhases / 2 + 1;
    const float4* table;  // 16-byte aligned rows of kPitch floats (ds_read_b128)
    // Weights in tap order (stereo and edge frames): a lane walks its row
    // forwards, or a mirrored row backwards from its last 16-byte group, and
    // reverses the elements of each group with v_perm_b32 (a byte select
    // between two registers): branch-free, where a conditional element order
    // compiled to a divergent branch per group. Mono frames read the stored
    // row forwards and reverse the samples instead, one v_perm_b32 per pair of
    // int16 samples rather than one per weight.
    struct Row {
        const float4* w;     // group 0 in this lane's tap order
        int dir;             // +1, or -1 for a mirrored row
        uint32_t sel;        // v_perm_b32 selector: 0x03020100 keeps an element, 0x07060504 takes its mirror
        const float4* base;  // the stored row (this lane's phase, or its mirror)
        bool mirror;
        __device__ __forceinline__ float pick(float mirrored, float kept) const
        {
            return __uint_as_float(__builtin_amdgcn_perm(__float_as_uint(mirrored), __float_as_uint(kept), sel));
        }
        // Group q of the weights in this lane's tap order.
        __device__ __forceinline__ float4 operator[](int q) const
        {
            const float4 r = w[q * dir];
            return make_float4(pick(r.w, r.x), pick(r.z, r.y), pick(r.y, r.z), pick(r.x, r.w));
        }
        // Group q of the row as stored. A mirrored lane pairs stored weight i
        // with sample T - 1 - i instead: the same products.
        __device__ __forceinline__ float4 stored(int q) const { return base[q]; }
        // The two int16 samples at mono positions 2q, 2q + 1 from the aligned
        // pairs `kept` (samples 2q, 2q + 1) and `mirrored` (samples T - 2 - 2q,
        // T - 1 - 2q), swapped into reverse order for a mirrored lane.
        __device__ __forceinline__ uint32_t pair(uint32_t mirrored, uint32_t kept) const
        {
            return __builtin_amdgcn_perm(mirrored, kept, mirror ? 0x05040706u : 0x03020100u);
        }
    };
    __device__ __forceinline__ Row row(int phase) const
    {
        const bool mirror = phase > kPhases / 2;
        const float4* start = table + (mirror ? kPhases - phase : phase) * (kPitch / 4);
        return {mirror ? start + (T / 4 - 1) : start, mirror ? -1 : 1, mirror ? 0x07060504u : 0x03020100u, start, mirror};
    }
};

// Windowed-sinc tap sum at the nearest of 2048 phases (src/cuda/render.cu:30-61).
// Accumulating the raw int16 values and scaling once by 2^-15 gives the same
// float result as weighting the pre-scaled float bank of the reference.
// `Plain`: the caller knows the voice is mono (see voiceAudio). `Pairs`: mono
// frames sum in the order that lets a mirrored LDS row reverse sample pairs
// instead of weights (the audio pass); the mix keeps the tap order, whose
// global-table form issues its weight and sample loads together. Either
// order gives the same bits with either table, so both execution forms agree.
template<int T, bool Plain = false, bool Pairs = false, class Taps>
__device__ __forceinline__ float2 interpolateAt(const int16_t* __restrict__ bank, const Taps& taps, const Span& v,
                                                double pos)
{
    constexpr float scale = 1.f / 32768.f;
    const double floorPos = floor(pos);
    const int centre = int(floorPos);
    if (pos == floorPos) {
        const float2 s = fetchRaw(bank, v, centre);
        return make_float2(s.x * scale, s.y * scale);
    }
    const int phase = __double2int_rn((pos - floorPos) * kPhases);
    const auto w4 = taps.row(phase);
    constexpr int first = T / 2 - 1;
    float a0 = 0.f, a1 = 0.f, b0 = 0.f, b1 = 0.f;
    if (centre >= first && centre < v.sourceEnd - T / 2) {
        const int16_t* __restrict__ src = bank + v.sample + uint64_t(centre - first) * uint64_t(v.channels);
        if (Plain || v.channels == 1) {
            // The window as dword-aligned loads (the bank allocation is padded for
            // the one dword past its end), realigned by 16 bits on an odd start.
            // (Stepping back from `src` keeps the pointer's global address
            // space, so these stay global loads rather than flat ones.)
            const int odd = int(reinterpret_cast<uintptr_t>(src) >> 1) & 1;
            const uint32_t* __restrict__ words = reinterpret_cast<const uint32_t*>(src - odd);
            const uint32_t shift = odd ? 16u : 0u;
            uint32_t dw[T / 2 + 1];
#pragma unroll
            for (int i = 0; i < T / 2 + 1; ++i)
                dw[i] = words[i];
            if constexpr (!Pairs) {
#pragma unroll
                for (int q = 0; q < T / 4; ++q) {
                    const float4 w = w4[q];
                    const uint32_t p0 = __builtin_amdgcn_alignbit(dw[2 * q + 1], dw[2 * q], shift);
                    const uint32_t p1 = __builtin_amdgcn_alignbit(dw[2 * q + 2], dw[2 * q + 1], shift);
                    a0 = fmaf(w.x, float(int16_t(p0 & 0xffffu)), a0);
                    a1 = fmaf(w.y, float(int16_t(p0 >> 16)), a1);
                    a0 = fmaf(w.z, float(int16_t(p1 & 0xffffu)), a0);
                    a1 = fmaf(w.w, float(int16_t(p1 >> 16)), a1);
                }
                const float m = (a0 + a1) * scale;
                return make_float2(m, m);
            }
            uint32_t p[T / 2], s[T / 2];
#pragma unroll
            for (int i = 0; i < T / 2; ++i)
                p[i] = __builtin_amdgcn_alignbit(dw[i + 1], dw[i], shift);  // samples 2i, 2i + 1
#pragma unroll
            for (int i = 0; i < T / 2; ++i)
                s[i] = w4.pair(p[T / 2 - 1 - i], p[i]);  // the samples at positions 2i, 2i + 1
            // Weights as stored, positions 0 .. T/2 - 1 ascending into a0 and
            // T - 1 .. T/2 descending into a1. A mirrored row holds the weights of
            // tap T - 1 - i at position i, so its a0 and a1 sum the terms that a
            // lane with the row unmirrored sums in a1 and a0, in the same order:
            // the result does not depend on which of the two rows a lane reads.
            const auto sample = [&](int i) {
                const uint32_t x = s[i >> 1];
                return float(int16_t(i & 1 ? x >> 16 : x & 0xffffu));
            };
#pragma unroll
            for (int q = 0; q < T / 8; ++q) {
                const float4 lw = w4.stored(q), hw = w4.stored(T / 4 - 1 - q);
                const int i = 4 * q, j = T - 1 - 4 * q;
                a0 = fmaf(lw.x, sample(i), a0);
                a1 = fmaf(hw.w, sample(j), a1);
                a0 = fmaf(lw.y, sample(i + 1), a0);
                a1 = fmaf(hw.z, sample(j - 1), a1);
                a0 = fmaf(lw.z, sample(i + 2), a0);
                a1 = fmaf(hw.y, sample(j - 2), a1);
                a0 = fmaf(lw.w, sample(i + 3), a0);
                a1 = fmaf(hw.x, sample(j - 3), a1);
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
// The denominator of a bent span's travel, expm1(slope), nonzero for a
// nonzero slope: the same for every frame of the span, so the audio pass
// evaluates it once per span and passes it as `bend`; a zero `bend` (the mix,
// short of registers) evaluates it per frame.
__device__ __forceinline__ double bendDenominator(const Span& v) { return v.pitchSlope != 0 ? expm1(v.pitchSlope) : 1.0; }

template<int T, bool Plain = false, bool Pairs = false, class Taps>
__device__ __forceinline__ float2 interpolate(const int16_t* __restrict__ bank, const Taps& taps, const Span& v,
                                              int frame, double pos, double bend)
{
    if (Plain || pos < 0) {
        const int elapsed = frame - v.begin;
        double distance = elapsed;
        if (v.pitchSlope != 0)
            distance = __ddiv_rn(expm1(__dmul_rn(v.pitchSlope, double(elapsed))),
                                 bend != 0 ? bend : expm1(v.pitchSlope));
        pos = __dadd_rn(v.position, __dmul_rn(distance, v.step));
    }
    if (!Plain && v.generator != CLAY_GEN_NONE) {
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
    float2 s = interpolateAt<T, Plain, Pairs>(bank, taps, v, pos);
    const int fadeStart = v.loopEnd - v.loopCrossfade;
    if (v.loopEnd > 0 && v.loopCrossfade > 0 && pos >= fadeStart) {
        const float2 entry = interpolateAt<T, Plain, Pairs>(bank, taps, v, __dsub_rn(pos, double(period)));
        const float phase = float((pos - fadeStart) / v.loopCrossfade);
        const float blend = .5f - .5f * cosf(3.14159265358979323846f * phase);
        s.x = fmaf(blend, entry.x - s.x, s.x);
        s.y = fmaf(blend, entry.y - s.y, s.y);
    }
    return s;
}

// The pan-law cosines of the last clamped pan position a lane evaluated:
// consecutive voices of one frame usually share their pan, so the two cosf
// calls run once per position instead of once per voice-frame.
struct PanCache {
    float p = __builtin_nanf(""), a = 0.f, b = 0.f;
};

// src/cuda/render.cu:130-136
__device__ __forceinline__ float2 panGains(const Span& v, float offset, PanCache& cache)
{
    if (offset == 0)
        return make_float2(v.left, v.right);
    const float p = fminf(1.f, fmaxf(-1.f, v.pan + offset));
    if (p != cache.p) {
        cache.p = p;
        cache.a = cosf((1 + p) * kQuarterPi);
        cache.b = cosf((1 - p) * kQuarterPi);
    }
    return make_float2(v.gain * cache.a, v.gain * cache.b);
}

// Envelope, modulation, resampling, pan and stereo matrix of one voice frame
// (src/cuda/render.cu:138-187). `Plain` compiles the common voice only: mono,
// sampled (no generator), its position from the span (no pitch chain) and no
// modulation that reaches the audio; the caller checks this per span.
template<int T, bool Plain = false, bool Pairs = false, class Taps>
__device__ __forceinline__ float2 voiceAudio(const int16_t* __restrict__ bank, const Taps& taps, const Span& v,
                                             const clay_control_point* __restrict__ points,
                                             const Mod* __restrict__ mods, float pan, int frame, bool filterInput,
                                             double phase, PanCache& panCache, double bend = 0.0,
                                             bool audioMods = true)
{
    float level = envelopeSpan(v, frame);
    if (level <= 0)
        return make_float2(0.f, 0.f);
    float db = 0, widthOffset = 0, positionOffset = 0;
    for (int j = v.firstMod; !Plain && audioMods && j < v.firstMod + v.modCount; ++j) {
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
    const float2 s = interpolate<T, Plain, Pairs>(bank, taps, v, frame, phase, bend);
    if (filterInput && (Plain || v.channels == 1))
        return make_float2(s.x * level, s.x * level);
    const float2 g = panGains(v, pan, panCache);
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
// Mono PCM16 of a stereo frame at `scale` (src/cuda/tail.cu:147-181): each
// channel clamped and rounded to nearest, then their average rounded half to
// even. The mix kernel and clayQuantizeMono share it, so both give the same bits.
__device__ __forceinline__ int monoSample(float2 x, float scale)
{
    const int l = __float2int_rn(fminf(32767.f, fmaxf(-32768.f, x.x * scale)));
    const int r = __float2int_rn(fminf(32767.f, fmaxf(-32768.f, x.y * scale)));
    const int sum = l + r;
    return (sum >> 1) + ((sum & 1) & ((sum >> 1) & 1));
}
constexpr float kUnitScale = 32768.f;  // the natural policy's scale of a job whose peak stays below 1

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
// Only the position is per lane; the track end comes from the (uniform) binding.
struct Cursor {
    int at;
    __device__ void reset(const clay_control_point* __restrict__ points, int first, int end, int frame)
    {
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
    // The current point and the next one are loaded together (the index of the
    // next is clamped into the track), so a frame that stays in its segment
    // costs one round trip.
    __device__ __forceinline__ double value(const clay_control_point* __restrict__ points, int last, int frame)
    {
        clay_control_point p = points[at], n = points[min(at + 1, last - 1)];
        while (at + 1 < last && n.frame <= frame) {
            ++at;
            p = n;
            n = points[min(at + 1, last - 1)];
        }
        return pointValue(p, frame);
    }
    // The value at `frame`; clears `steady` unless the track stays constant
    // (zero slope, no further point) through frame `until`.
    __device__ __forceinline__ double steadyValue(const clay_control_point* __restrict__ points, int last, int frame,
                                                  int until, bool& steady)
    {
        clay_control_point p = points[at], n = points[min(at + 1, last - 1)];
        while (at + 1 < last && n.frame <= frame) {
            ++at;
            p = n;
            n = points[min(at + 1, last - 1)];
        }
        steady = steady && p.step == 0.f && (at + 1 >= last || n.frame > until);
        return pointValue(p, frame);
    }
};

// A Cursor that keeps its current and next point in registers: a frame in the
// same segment as the previous one loads nothing (readers of one binding can
// afford the registers; the loads were a serial latency per frame).
struct CachedCursor {
    int at;
    clay_control_point p, n;
    __device__ void reset(const clay_control_point* __restrict__ points, int first, int end, int frame)
    {
        Cursor c;
        c.reset(points, first, end, frame);
        at = c.at;
        p = points[at];
        n = points[min(at + 1, end - 1)];
    }
    __device__ __forceinline__ void advance(const clay_control_point* __restrict__ points, int last, int frame)
    {
        while (at + 1 < last && n.frame <= frame) {
            ++at;
            p = n;
            n = points[min(at + 1, last - 1)];
        }
    }
    __device__ __forceinline__ double value(const clay_control_point* __restrict__ points, int last, int frame)
    {
        advance(points, last, frame);
        return pointValue(p, frame);
    }
    __device__ __forceinline__ double steadyValue(const clay_control_point* __restrict__ points, int last, int frame,
                                                  int until, bool& steady)
    {
        advance(points, last, frame);
        steady = steady && p.step == 0.f && (at + 1 >= last || n.frame > until);
        return pointValue(p, frame);
    }
};

// Up to kCursors modulation bindings of a span read through cursors; any
// further ones fall back to the binary search.
constexpr int kCursors = 4;

// The cursors persist across the spans of a chain: a binding to the track a
// cursor already follows keeps its position (frames only move forward), so
// only a new track costs a binary search.
template<int N>
struct ModReaderN {
    using C = std::conditional_t<N == 1, CachedCursor, Cursor>;
    C track[N], shape[N];
    int trackKey[N], shapeKey[N];  // first point of the followed track, -1: none
    int first, count;

    __device__ ModReaderN()
    {
#pragma unroll
        for (int j = 0; j < N; ++j)
            trackKey[j] = shapeKey[j] = -1;
    }

    __device__ void reset(const Mod* __restrict__ mods, const clay_control_point* __restrict__ points, int firstMod,
                          int modCount, int frame)
    {
        first = firstMod;
        count = modCount;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (j >= count)
                break;
            const Mod m = mods[first + j];
            if (m.first >= 0 && m.first != trackKey[j]) {
                track[j].reset(points, m.first, m.last, frame);
                trackKey[j] = m.first;
            }
            if (m.shapeFirst >= 0 && m.shapeFirst != shapeKey[j]) {
                shape[j].reset(points, m.shapeFirst, m.shapeLast, frame);
                shapeKey[j] = m.shapeFirst;
            }
        }
    }

    // src/cuda/render.cu:120-126 in double.
    __device__ double value(const Mod* __restrict__ mods, const clay_control_point* __restrict__ points, int j, int frame)
    {
        const Mod m = mods[first + j];
        double v = m.depth;
        if (m.first >= 0) {
            const double x = j < N ? track[j].value(points, m.last, frame) : [&] {
                Cursor c;
                c.reset(points, m.first, m.last, frame);
                return pointValue(points[c.at], frame);
            }();
            v = double(m.depth) * x;
        }
        if (m.shapeFirst >= 0) {
            const double x = j < N ? shape[j].value(points, m.shapeLast, frame) : [&] {
                Cursor c;
                c.reset(points, m.shapeFirst, m.shapeLast, frame);
                return pointValue(points[c.at], frame);
            }();
            v *= x;
        }
        return v;
    }

    // value() of cursor slot j < N at `frame`, clearing `steady` unless
    // it stays constant through frame `until`.
    __device__ double steadyValue(const Mod* __restrict__ mods, const clay_control_point* __restrict__ points, int j,
                                  int frame, int until, bool& steady)
    {
        const Mod m = mods[first + j];
        double v = m.depth;
        if (m.first >= 0)
            v = double(m.depth) * track[j].steadyValue(points, m.last, frame, until, steady);
        if (m.shapeFirst >= 0)
            v *= shape[j].steadyValue(points, m.shapeLast, frame, until, steady);
        return v;
    }
};
using ModReader = ModReaderN<kCursors>;

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
// The resonance factor 10^(-dB / 20) of the two-pole filters (dB clamped to
// +-60), kept apart so that callers reuse it while the resonance repeats.
__device__ __forceinline__ double resonanceGain(double resonance)
{
    return exp10(-fmin(60.0, fmax(-60.0, resonance)) / 20);
}

// 1 / x for x in [1, 2]: the hardware reciprocal refined by two Newton steps
// (within an ulp of the quotient; the coefficients tolerate far more).
__device__ __forceinline__ double reciprocalD(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// sin and cos of an angle w in [0, pi] (a filter's angular frequency below
// Nyquist): w = k pi/2 + r with k = 0, 1 or 2 and |r| <= pi/4 (pi/2 in two
// parts, so k times the leading part is exact), then the Taylor series of sin
// and cos in r to degree 17 and 18, whose truncation error is below 1e-19.
__device__ __forceinline__ void sincosQuadrants(double w, double& s, double& c)
{
    constexpr double halfPiHi = 1.5707963267948966, halfPiLo = 6.123233995736766e-17;
    const double k = rint(w * 0.63661977236758134);  // w / (pi/2), rounded
    const double r = fma(-k, halfPiLo, fma(-k, halfPiHi, w)), r2 = r * r;
    double sp = 1.0 / 355687428096000.0;  // 1/17!
    sp = fma(sp, r2, -1.0 / 1307674368000.0);
    sp = fma(sp, r2, 1.0 / 6227020800.0);
    sp = fma(sp, r2, -1.0 / 39916800.0);
    sp = fma(sp, r2, 1.0 / 362880.0);
    sp = fma(sp, r2, -1.0 / 5040.0);
    sp = fma(sp, r2, 1.0 / 120.0);
    sp = fma(sp, r2, -1.0 / 6.0);
    const double sr = fma(r * r2, sp, r);
    double cp = -1.0 / 6402373705728000.0;  // -1/18!
    cp = fma(cp, r2, 1.0 / 20922789888000.0);
    cp = fma(cp, r2, -1.0 / 87178291200.0);
    cp = fma(cp, r2, 1.0 / 479001600.0);
    cp = fma(cp, r2, -1.0 / 3628800.0);
    cp = fma(cp, r2, 1.0 / 40320.0);
    cp = fma(cp, r2, -1.0 / 720.0);
    cp = fma(cp, r2, 1.0 / 24.0);
    cp = fma(cp, r2, -0.5);
    const double cr = fma(r2, cp, 1.0);
    s = k == 1 ? cr : k == 2 ? -sr : sr;
    c = k == 1 ? -sr : k == 2 ? -cr : cr;
}

// `gainOfResonance` is resonanceGain(resonance) for the two-pole types. The
// angular frequency multiplies by 2 pi / rate (loop-invariant for callers)
// instead of dividing by the rate.
__device__ inline CoeffD coefficientsD(int type, double cutoff, double gainOfResonance, int rate, double bandwidth,
                                       double gain)
{
    constexpr double pi = 3.14159265358979323846;
    const double w = fmin(fmin(20000.0, rate * .499), fmax(1.0, cutoff)) * (2 * pi / rate);
    if (type == CLAY_FILTER_LOW1 || type == CLAY_FILTER_HIGH1) {
        const double p = exp(-w);
        if (type == CLAY_FILTER_LOW1)
            return {1 - p, 0, 0, -p, 0};
        const double h = (1 + p) * .5;
        return {h, -h, 0, -p, 0};
    }
    double s, c;
    sincosQuadrants(w, s, c);
    if (type == CLAY_FILTER_PEAK) {
        const double a = exp10(fmin(60.0, fmax(-96.0, gain)) / 40);
        const double width = fmin(12.0, fmax(.01, bandwidth));
        const double alpha = s * sinh(fmin(20.0, 0.34657359027997264 * width * w / s));
        const double inv = 1 / (1 + alpha / a);
        const double mid = -2 * c * inv;
        return {(1 + alpha * a) * inv, mid, (1 - alpha * a) * inv, mid, (1 - alpha / a) * inv};
    }
    const double alpha = .5 * s * gainOfResonance;
    const double inv = reciprocalD(1 + alpha);
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

// ----------------------------------------------------------------- round descriptors
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

// Wave-level barrier for a wave's own staging area (LDS operations of one wave
// complete in order; this keeps the compiler from moving them across).
__device__ __forceinline__ void waveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
