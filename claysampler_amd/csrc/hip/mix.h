// Output mix of the render core for gfx950 (src/cuda/render.cu:21-187, :458-523).
//
// A wave owns a tile of one job's output frames and walks the tile's voices
// in span order, with lane l holding the frames l, l + 64, l + 128, ... of the
// tile (one "row" of 64 consecutive frames per register), so the per-voice
// work (the span record in scalar registers, the envelope stage, the pan
// gains) is paid once per tile, not once per frame. Sixteen waves share a
// 1024-thread block, one block per CU, grid-stride over an XCD's share of the
// tiles. Two kernels:
//  - clayMixStraight: batches whose voices are all straight (see below), on
//    1024-frame tiles (16 rows per lane);
//  - clayMixWaves: every other batch with dry voices, on 256-frame tiles
//    (4 rows per lane), with filtered voices read from the chain scratch.
//
// Straight voices (mono, sampled, dry, no modulation of their audio, no loop,
// no bend) take an integer FIR: the sinc weights are stored as packed int16
// pairs,
//     w = hi * 2^-14 + lo * 2^-24,   hi = rint(w 2^14),  lo = rint((w - hi 2^-14) 2^24),
// exact to 2^-25, and the int16 sample pairs of the bank are dotted with them
// by v_dot2c_i32_i16, two taps per instruction, in exact int32 sums
// (|sum hi s| < 1.6 * 2^29 and |sum lo s| <= T 2^9 2^15 <= 2^30 for T <= 64).
// Both kernels compute a straight voice's frame with the same functions, so
// the PCM of every execution form is bit-identical.
#pragma once

#include "common.h"

namespace {

constexpr int kMixWaves = 16;                  // tiles in flight per 1024-thread block
constexpr int kStraightWaves = 12;             // tiles in flight per 768-thread straight block (168 VGPRs)
constexpr int kLaneFrames = kTileFrames / 64;  // rows of a 256-frame tile
constexpr int kStraightRows = kStraightFrames / 64;
constexpr int kIntSharedTaps = 16;             // widest integer table held in LDS
constexpr int kRowQuads = 5;                   // LDS row stride: 4 quads of weights and one pad quad (80 B)

template<bool B>
struct Flag {
    static constexpr bool value = B;
};

// Integer weights of the full table in global memory: (kPhases + 1) rows of T
// dwords, the T/2 hi pairs then the T/2 lo pairs.
template<int T>
struct IntTapsGlobal {
    const uint32_t* __restrict__ table;
    struct Row {
        const uint32_t* __restrict__ w;
        __device__ __forceinline__ uint4 quad(int q) const { return reinterpret_cast<const uint4*>(w)[q]; }
        __device__ __forceinline__ uint32_t word(int i) const { return w[i]; }
    };
    __device__ __forceinline__ Row row(int phase) const { return {table + size_t(phase) * T}; }
};

// Rows 0..kPhases-1 of the 16-tap table in LDS (160 KB), 80 bytes apart: a
// row starts on 16-byte bank group 5r mod 8 (of 32 banks) or 5r mod 16 (of
// 64), so the rows of consecutive frames, whose phases advance by a constant
// step, fall on distinct bank groups whenever that step is odd, and the
// address of a row is one multiply. Phase kPhases is never read: rows 0 and
// kPhases are unit impulses at taps T/2-1 and T/2 (the renderer checks it),
// so (centre, kPhases) is (centre + 1, 0).
template<int T>
struct IntTapsShared {
    static_assert(T == 16, "the LDS table is laid out for 16 taps");
    const uint4* table;
    struct Row {
        const uint4* w;
        __device__ __forceinline__ uint4 quad(int q) const { return w[q]; }
        __device__ __forceinline__ uint32_t word(int i) const { return reinterpret_cast<const uint32_t*>(w)[i]; }
    };
    __device__ __forceinline__ Row row(int phase) const { return {table + phase * kRowQuads}; }
};

__device__ __forceinline__ int dot2(uint32_t a, uint32_t b, int c)
{
    typedef short v2s __attribute__((ext_vector_type(2)));
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, a), __builtin_bit_cast(v2s, b), c, false);
}

// Source value from the two integer sums: (hi + lo 2^-10) 2^-14, scaled by 2^-15.
__device__ __forceinline__ float fromSums(int hi, int lo) { return fmaf(float(lo), 0x1p-10f, float(hi)) * 0x1p-29f; }

// Centre sample and table phase of a source position: the FIR at `pos` uses
// the window centred on floor(pos) with the phase rint(frac(pos) 2048), and a
// phase of 2048 is taken as phase 0 of the next centre (the same weights, see
// IntTapsShared), so phase < kPhases.
struct Locus {
    int centre, phase;
};

// Any position (loop crossfade entries may lie before the source).
__device__ __forceinline__ Locus locate(double pos)
{
    const double floorPos = floor(pos);
    const int phase = min(max(__double2int_rn((pos - floorPos) * kPhases), 0), kPhases);
    return {int(floorPos) + (phase >> 11), phase & (kPhases - 1)};
}

// FIR over the sample pairs p[i] = (s[2i], s[2i+1]) of the window (low half first).
template<int T, class Taps>
__device__ __forceinline__ float firPairs(const typename Taps::Row& row, const uint32_t (&p)[T / 2])
{
    int hi = 0, lo = 0;
#pragma unroll
    for (int g = 0; g < T / 8; ++g) {
        const uint4 h = row.quad(g), l = row.quad(T / 8 + g);
        const uint32_t hv[4] = {h.x, h.y, h.z, h.w}, lv[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            hi = dot2(p[4 * g + k], hv[k], hi);
            lo = dot2(p[4 * g + k], lv[k], lo);
        }
    }
    return fromSums(hi, lo);
}

// The two sums one tap at a time, zero outside [0, sourceEnd): windows at the
// edges of the source. Returned as hi + lo 2^-10 (units of 2^-29).
template<int T, class Row>
__device__ __forceinline__ float firEdgeUnscaled(const int16_t* __restrict__ src, int sourceEnd, int centre,
                                                 const Row& row)
{
    int hi = 0, lo = 0;
#pragma unroll 1
    for (int t = 0; t < T; ++t) {
        const int i = centre - (T / 2 - 1) + t;
        const int s = i >= 0 && i < sourceEnd ? int(src[i]) : 0;
        const uint32_t h = row.word(t >> 1), l = row.word(T / 2 + (t >> 1));
        const int wh = int(int16_t(t & 1 ? h >> 16 : h & 0xffffu));
        const int wl = int(int16_t(t & 1 ? l >> 16 : l & 0xffffu));
        hi += wh * s;
        lo += wl * s;
    }
    return fmaf(float(lo), 0x1p-10f, float(hi));
}

template<int T, class Taps>
__device__ __forceinline__ float firEdge(const int16_t* __restrict__ bank, uint64_t sample, int sourceEnd, int centre,
                                         const typename Taps::Row& row)
{
    return firEdgeUnscaled<T>(bank + sample, sourceEnd, centre, row) * 0x1p-29f;
}

// Source value of a plain voice at a position (src/cuda/render.cu:21-61): the
// exact sample at an integer position, else the FIR at the nearest of 2048
// phases. `active`: the lane renders this frame; inactive lanes load from a
// valid dummy address and take no edge path.
template<int T, class Taps>
__device__ __forceinline__ float sourceAt(const int16_t* __restrict__ bank, const Taps& taps, const Span& v, double pos,
                                          bool active, const uint32_t* __restrict__ dummy)
{
    const Locus at = locate(pos);
    const int centre = at.centre;
    const auto row = taps.row(at.phase);
    const bool interior = centre >= T / 2 - 1 && centre < v.sourceEnd - T / 2;
    const uint64_t first = v.sample + uint64_t(int64_t(centre - (T / 2 - 1)));
    // Dword-aligned window (the bank is padded for the dword past its end),
    // realigned by 16 bits on an odd start.
    const int odd = interior ? int(first & 1) : 0;
    const uint32_t* __restrict__ words = interior ? reinterpret_cast<const uint32_t*>(bank + (first - uint64_t(odd)))
                                                  : dummy;
    uint32_t dw[T / 2 + 1];
#pragma unroll
    for (int i = 0; i < T / 2 + 1; ++i)
        dw[i] = words[i];
    uint32_t p[T / 2];
#pragma unroll
    for (int i = 0; i < T / 2; ++i)
        p[i] = __builtin_amdgcn_alignbit(dw[i + 1], dw[i], odd ? 16u : 0u);
    float s = firPairs<T, Taps>(row, p);
    if (pos == floor(pos)) {
        // Exact sample (render.cu:33-34): window tap T/2 - 1 when interior.
        const uint32_t pair = p[(T / 2 - 1) >> 1];
        const int x = interior ? int(int16_t(((T / 2 - 1) & 1) ? pair >> 16 : pair & 0xffffu))
                               : (centre >= 0 && centre < v.sourceEnd ? int(bank[v.sample + uint64_t(centre)]) : 0);
        s = float(x) * 0x1p-15f;
    } else if (active && !interior) {
        s = firEdge<T, Taps>(bank, v.sample, v.sourceEnd, centre, row);
    }
    return s;
}

// Source position of a bent span at `frame` (src/cuda/render.cu:63-70), `bend` = expm1(pitchSlope).
__device__ __forceinline__ double bentPosition(const Span& v, int frame, double bend)
{
    const int elapsed = frame - v.begin;
    const double distance = __ddiv_rn(expm1(__dmul_rn(v.pitchSlope, double(elapsed))), bend);
    return __dadd_rn(v.position, __dmul_rn(distance, v.step));
}

// Loop wrap (src/cuda/render.cu:85-87).
__device__ __forceinline__ double wrapLoop(double pos, int loopFirst, int period)
{
    return __dadd_rn(double(loopFirst), fmod(__dsub_rn(pos, double(loopFirst)), double(period)));
}

// Position, loop wrap and loop crossfade of a plain voice at `frame`
// (src/cuda/render.cu:63-100 without generators).
template<int T, class Taps>
__device__ __forceinline__ float plainSource(const int16_t* __restrict__ bank, const Taps& taps, const Span& v,
                                             int frame, double bend, bool active, const uint32_t* __restrict__ dummy)
{
    double pos = v.pitchSlope != 0 ? bentPosition(v, frame, bend)
                                   : __dadd_rn(v.position, __dmul_rn(double(frame - v.begin), v.step));
    const int period = v.loopEnd - v.loopFirst;
    if (v.loopEnd > 0 && pos >= v.loopEnd)
        pos = wrapLoop(pos, v.loopFirst, period);
    if (pos >= v.sourceEnd)
        return 0.f;
    float s = sourceAt<T>(bank, taps, v, pos, active, dummy);
    const int fadeStart = v.loopEnd - v.loopCrossfade;
    if (v.loopEnd > 0 && v.loopCrossfade > 0 && pos >= fadeStart) {
        const float entry = sourceAt<T>(bank, taps, v, __dsub_rn(pos, double(period)), active, dummy);
        const float phase = float((pos - fadeStart) / v.loopCrossfade);
        const float blend = .5f - .5f * cosf(3.14159265358979323846f * phase);
        s = fmaf(blend, entry - s, s);
    }
    return s;
}

// Amplitude envelope of a plain voice (src/cuda/render.cu:6-19). Its six
// stages share one formula with stage parameters:
//     n = frame - origin + 1,
//     level = isExp ? max(lo, scale 2^(rate n)) : min(hi, max(lo, base + slope n)),
// rates in log2 units (one v_exp_f32). Before the attack: base = start;
// attack: start + n attackStep below 1; hold: 1; decay: max(sustain,
// 2^(rate n)), or sustain for a zero rate; release: releaseValue 2^(rate n);
// fade: max(0, fadeValue - n fadeStep). A release overrides the earlier stages.
struct EnvStage {
    int origin;
    float slope, base, rate, scale, lo, hi;
    bool isExp;
};

// The envelope fields of a span, read unconditionally (so the scalar loads
// are issued together rather than one behind each branch that needs them).
struct EnvFields {
    int attackStart, holdStart, decayStart, releaseStart, fadeStart;
    float start, attackStep, decayRate, sustain, releaseValue, releaseRate, fadeValue, fadeStep;
};

__device__ __forceinline__ EnvFields envFields(const Span& v)
{
    return {v.attackStart, v.holdStart, v.decayStart, v.releaseStart, v.fadeStart, v.start, v.attackStep,
            v.decayRate, v.sustain, v.releaseValue, v.releaseRate, v.fadeValue, v.fadeStep};
}

__device__ __forceinline__ int envStageIndex(const EnvFields& v, int frame)
{
    if (v.releaseStart >= 0 && frame >= v.releaseStart)
        return frame >= v.fadeStart ? 5 : 4;
    return frame < v.attackStart ? 0 : frame < v.holdStart ? 1 : frame < v.decayStart ? 2 : 3;
}

// Parameters of stage `k` (selects, no branches: `k` may differ per lane).
__device__ __forceinline__ EnvStage envStage(const EnvFields& v, float decay2, float release2, int k)
{
    constexpr float inf = __builtin_huge_valf();
    EnvStage s;
    s.isExp = k == 4 || (k == 3 && v.decayRate != 0);
    s.origin = k == 1 ? v.attackStart : k == 3 ? v.decayStart : k == 4 ? v.releaseStart : v.fadeStart;
    s.slope = k == 1 ? v.attackStep : k == 5 ? -v.fadeStep : 0.f;
    s.base = k <= 1 ? v.start : k == 2 ? 1.f : k == 3 ? v.sustain : v.fadeValue;
    s.rate = k == 4 ? release2 : decay2;
    s.scale = k == 4 ? v.releaseValue : 1.f;
    s.lo = k == 3 ? v.sustain : k == 5 ? 0.f : -inf;
    s.hi = k == 1 ? 1.f : inf;
    return s;
}

__device__ __forceinline__ float envLevel(const EnvStage& s, int frame)
{
    const float n = float(frame - s.origin + 1);
    return s.isExp ? fmaxf(s.lo, s.scale * __builtin_amdgcn_exp2f(n * s.rate))
                   : fminf(s.hi, fmaxf(s.lo, fmaf(n, s.slope, s.base)));
}

// Pan gains of a mono voice for a pan offset (src/cuda/render.cu:130-136),
// the two cosines cached per lane while the clamped position repeats.
__device__ __forceinline__ float2 monoGains(float left, float right, float gain, float pan, float offset,
                                           PanCache& cache)
{
    if (offset == 0)
        return make_float2(left, right);
    const float p = fminf(1.f, fmaxf(-1.f, pan + offset));
    if (p != cache.p) {
        cache.p = p;
        cache.a = cosf((1 + p) * kQuarterPi);
        cache.b = cosf((1 - p) * kQuarterPi);
    }
    return make_float2(gain * cache.a, gain * cache.b);
}

// The envelope of a plain voice over the frames [first, last] of a tile: its
// stage when one stage holds over all of them (the usual case), and the level
// itself when that stage is flat. A decay that has fallen below the sustain
// level at `first` stays below it (the exponential only falls), so its level
// is the sustain level; the margin covers the last-place error of v_exp_f32.
struct TileEnvelope {
    EnvFields v;
    float decay2, release2;
    EnvStage stage;
    bool stageUniform, constLevel;
    float level0;
    __device__ __forceinline__ TileEnvelope(const Span& span, int first, int last)
    {
        v = envFields(span);
        decay2 = v.decayRate * 1.44269504088896341f;
        release2 = v.releaseRate * 1.44269504088896341f;
        const int k0 = envStageIndex(v, first), k1 = envStageIndex(v, last);
        stageUniform = k0 == k1;
        stage = envStage(v, decay2, release2, k0);
        if (stageUniform && k0 == 3 && stage.isExp && v.decayRate < 0 && envLevel(stage, first) == stage.lo &&
            __builtin_amdgcn_exp2f(float(first - stage.origin + 1) * stage.rate) < stage.lo * (1.f - 0x1p-20f)) {
            stage.isExp = false;
            stage.slope = 0.f;
            stage.base = stage.lo;
        }
        constLevel = stageUniform && !stage.isExp && stage.slope == 0;
        level0 = fminf(stage.hi, fmaxf(stage.lo, stage.base));
    }
    __device__ __forceinline__ float level(int frame) const
    {
        return constLevel     ? level0
               : stageUniform ? envLevel(stage, frame)
                              : envLevel(envStage(v, decay2, release2, envStageIndex(v, frame)), frame);
    }
};

// Accumulates a plain voice with a loop or a bend into the lane's dry sums.
template<class Source>
__device__ __forceinline__ void plainFrames(const Span& span, const int (&frame)[kLaneFrames],
                                            const bool (&on)[kLaneFrames], const float (&pan)[kLaneFrames],
                                            bool panUniform, int tileFirst, int tileEnd, float (&left)[kLaneFrames],
                                            float (&right)[kLaneFrames], PanCache& panCache, Source&& source)
{
    const int begin = span.begin, end = span.end;
    const TileEnvelope env(span, max(begin, tileFirst), min(end, tileEnd) - 1);
    float2 g0 = make_float2(span.left, span.right);
    if (panUniform)
        g0 = monoGains(span.left, span.right, span.gain, span.pan, pan[0], panCache);
#pragma unroll
    for (int k = 0; k < kLaneFrames; ++k) {
        const bool act = on[k] && frame[k] >= begin && frame[k] < end;
        const float src = source(frame[k], act);
        const float level = env.level(frame[k]);
        const float2 g = panUniform ? g0 : monoGains(span.left, span.right, span.gain, span.pan, pan[k], panCache);
        if (act) {
            left[k] = fmaf(src, g.x * level, left[k]);
            right[k] = fmaf(src, g.y * level, right[k]);
        }
    }
}

// ---------------------------------------------------------------- straight voices

// A straight voice: mono, sampled, dry, no modulation of its audio, no loop
// and no bend, at source position `position + (frame - begin) step` with
// position >= 0 and step >= 0 (the host checks both), so positions never fall
// with the frame and never go below zero.
//
// Its windows are read with dword-aligned 16-byte loads (the texture
// addresser takes unaligned ones one lane per clock). The renderer keeps a
// copy of the bank shifted by one sample (copy[i] = bank[i + 1]) at a fixed
// byte distance D after it, so a window starting on an odd sample j is the
// aligned run at copy[j - 1]: byte offset 2 j + (D - 2) from the bank, where
// an even start is at 2 j. `odd` is D - 2.
template<int T>
struct Straight {
    const uint32_t* __restrict__ words;  // the bank from the voice's sample 0 rounded down to a dword
    const int16_t* __restrict__ src;     // the voice's sample 0
    uint32_t odd;
    int parity;                          // sample & 1: where sample 0 sits in its dword
    double position, step;               // 2048 times the span's position and step
    double magic;                        // 1.5 2^52 + (parity - (T/2 - 1)) 2048
    int begin, end, sourceEnd;
    int lastStart;                       // sourceEnd - T: the last window start inside the source
    __device__ __forceinline__ Straight(const int16_t* __restrict__ bank, uint32_t oddBytes, const Span& v)
        : words(reinterpret_cast<const uint32_t*>(bank + (v.sample & ~uint64_t(1)))), src(bank + v.sample),
          odd(oddBytes), parity(int(v.sample & 1)), position(__dmul_rn(v.position, double(kPhases))),
          step(__dmul_rn(v.step, double(kPhases))), begin(v.begin), end(v.end), sourceEnd(v.sourceEnd),
          lastStart(v.sourceEnd - T)
    {
        // The constant's mantissa holds the window offset: an integer add on its bits.
        magic = __longlong_as_double(0x4338000000000000ll + int64_t(parity - (T / 2 - 1)) * kPhases);
    }
    // The window of frame begin + n (src/cuda/render.cu:21-61 and :63-70 with
    // no bend): its first sample, counted from `words`, and its table phase.
    // 2048 pos is exact in FP64, as is the scaling of position and step (so
    // this is 2048 times the reference's pos), and adding the constant rounds
    // it to the nearest integer q, ties to even (the constant is even): q
    // lands in the low 51 bits, shifted by the window offset, so bits 11..42
    // are floor(q / 2048) - (T/2 - 1) + parity (the centre floor(pos), or
    // floor(pos) + 1 when the phase rounds up to 2048) and bits 0..10 are the
    // phase, q mod 2048.
    struct Place {
        int start, phase;
    };
    __device__ __forceinline__ Place place(int n) const
    {
        const double pos = __dadd_rn(position, __dmul_rn(double(n), step));
        const uint64_t bits = uint64_t(__double_as_longlong(__dadd_rn(pos, magic)));
        const uint32_t lo = uint32_t(bits), hi = uint32_t(bits >> 32);
        return {int(__builtin_amdgcn_alignbit(hi, lo, 11)), int(lo & uint32_t(kPhases - 1))};
    }
    // Window start relative to the voice's sample 0 (centre - (T/2 - 1)).
    __device__ __forceinline__ int window(const Place& p) const { return p.start - parity; }
};

typedef uint32_t WindowQuad __attribute__((ext_vector_type(4), aligned(4)));

template<int T>
struct Fetched {
    uint4 q[T / 8];
    int start, phase;
    bool inside;  // the window lies inside the source (always, in the interior form)
};

// Starts one frame: its window's loads. Interior: the caller has shown the
// window inside the source; otherwise the window is clamped into it and
// `inside` says whether it had to be.
template<int T, bool Interior>
__device__ __forceinline__ Fetched<T> fetchStraight(const Straight<T>& s, int n)
{
    Fetched<T> f;
    const auto at = s.place(n);
    f.start = at.start;
    f.phase = at.phase;
    int start = at.start;
    f.inside = true;
    if (!Interior) {
        const int w = s.window(at);
        f.inside = s.lastStart >= 0 && uint32_t(w) <= uint32_t(s.lastStart);
        start = min(max(w, 0), max(s.lastStart, 0)) + s.parity;
    }
    // 2 start, plus D - 2 for an odd start (the shifted copy).
    const uint32_t offset = (uint32_t(start) << 1) + (uint32_t(__builtin_amdgcn_sbfe(start, 0, 1)) & s.odd);
    const char* p = reinterpret_cast<const char*>(s.words) + offset;
#pragma unroll
    for (int i = 0; i < T / 8; ++i) {
        const WindowQuad w = *reinterpret_cast<const WindowQuad*>(p + 16 * i);
        f.q[i] = make_uint4(w.x, w.y, w.z, w.w);
    }
    return f;
}

// First term of a dot chain, with the zero accumulator as an operand (no
// register to clear first).
__device__ __forceinline__ int dot2First(uint32_t a, uint32_t b)
{
    int d;
    asm("v_dot2_i32_i16 %0, %1, %2, 0" : "=v"(d) : "v"(a), "v"(b));
    return d;
}

// The FIR of a window with a table row, as hi + lo 2^-10 (units of 2^-29 of
// full scale).
template<int T, class Weights>
__device__ __forceinline__ float firWindow(const uint4 (&q)[T / 8], Weights&& w)
{
    int hi = 0, lo = 0;
#pragma unroll
    for (int g = 0; g < T / 8; ++g) {
        const uint4 h = w(g), l = w(T / 8 + g), p = q[g];
        hi = dot2(p.x, h.x, hi);
        lo = dot2(p.x, l.x, lo);
        hi = dot2(p.y, h.y, hi);
        lo = dot2(p.y, l.y, lo);
        hi = dot2(p.z, h.z, hi);
        lo = dot2(p.z, l.z, lo);
        hi = dot2(p.w, h.w, hi);
        lo = dot2(p.w, l.w, lo);
    }
    return fmaf(float(lo), 0x1p-10f, float(hi));
}

// Completes one frame: the FIR as hi + lo 2^-10. Rows 0 and kPhases are unit
// impulses, so an integer position gives its sample exactly
// (src/cuda/render.cu:33-34); a window past either end of the source is
// summed tap by tap with zeros outside, and a centre at or past the end gives
// 0 (render.cu:88-89).
template<int T, bool Interior, class Taps>
__device__ __forceinline__ float finishStraight(const Taps& taps, const Straight<T>& s, const Fetched<T>& f, bool act)
{
    const auto row = taps.row(f.phase);
    float x = firWindow<T>(f.q, [&](int q) { return row.quad(q); });
    if (!Interior && act && !f.inside) {
        const int centre = f.start - s.parity + (T / 2 - 1);
        x = centre < s.sourceEnd ? firEdgeUnscaled<T>(s.src, s.sourceEnd, centre, row) : 0.f;
    }
    return x;
}

// Frames of fetch-ahead: as many windows in flight per lane as the registers
// allow.
template<int T>
constexpr int straightAhead()
{
    return T <= 16 ? 2 : T <= 32 ? 2 : 1;
}

// A frame started with its weights too: the window and the table row both
// in flight (for tables in LDS).
template<int T>
struct Staged {
    Fetched<T> f;
    uint4 w[T / 4];
};

template<int T, class Taps>
__device__ __forceinline__ void stageStraight(const Taps& taps, const Straight<T>& s, int n, Staged<T>& out)
{
    out.f = fetchStraight<T, true>(s, n);
    const auto row = taps.row(out.f.phase);
#pragma unroll
    for (int q = 0; q < T / 4; ++q)
        out.w[q] = row.quad(q);
}

// The FIR of a staged frame (finishStraight with the row in registers).
template<int T>
__device__ __forceinline__ float sumStaged(const Staged<T>& st)
{
    return firWindow<T>(st.f.q, [&](int q) { return st.w[q]; });
}

// The rows of a tile that one straight voice covers completely, its windows
// inside the source and its pan flat: no row is skipped or masked and nothing
// branches. Software pipeline: the windows (and, for a table in LDS, the
// weight rows) of group g + 1 are requested before group g is summed; the
// barrier keeps the scheduler from hoisting more. `Exp`: the envelope is one
// exponential stage over the tile, else a constant level.
template<int T, int Rows, bool Exp, class Taps>
__device__ __forceinline__ void straightRowsCovered(const Taps& taps, const Straight<T>& sv, const TileEnvelope& env,
                                                    float2 g0s, float2 lg0s, int rel, float (&left)[Rows],
                                                    float (&right)[Rows])
{
    constexpr int Ahead = T <= kIntSharedTaps ? 1 : straightAhead<T>();
    const auto accumulate = [&](int k, float x) {
        float2 lgs = lg0s;
        if (Exp) {
            // envLevel of an exponential stage, without its branch.
            const EnvStage& st = env.stage;
            const float n = float(sv.begin + rel + 64 * k - st.origin + 1);
            const float level = fmaxf(st.lo, st.scale * __builtin_amdgcn_exp2f(n * st.rate));
            lgs = make_float2(g0s.x * level, g0s.y * level);
        }
        left[k] = fmaf(x, lgs.x, left[k]);
        right[k] = fmaf(x, lgs.y, right[k]);
    };
    if constexpr (T <= kIntSharedTaps) {
        Staged<T> buf[2][Ahead];
#pragma unroll
        for (int e = 0; e < Ahead; ++e)
            stageStraight<T>(taps, sv, rel + 64 * e, buf[0][e]);
#pragma unroll
        for (int g = 0; g < Rows; g += Ahead) {
            const int c = (g / Ahead) & 1;
            if (g + Ahead < Rows) {
#pragma unroll
                for (int e = 0; e < Ahead; ++e)
                    stageStraight<T>(taps, sv, rel + 64 * (g + Ahead + e), buf[c ^ 1][e]);
            }
#pragma unroll
            for (int e = 0; e < Ahead; ++e)
                accumulate(g + e, sumStaged<T>(buf[c][e]));
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        Fetched<T> buf[2][Ahead];
#pragma unroll
        for (int e = 0; e < Ahead; ++e)
            buf[0][e] = fetchStraight<T, true>(sv, rel + 64 * e);
#pragma unroll
        for (int g = 0; g < Rows; g += Ahead) {
            const int c = (g / Ahead) & 1;
            if (g + Ahead < Rows) {
#pragma unroll
                for (int e = 0; e < Ahead; ++e)
                    buf[c ^ 1][e] = fetchStraight<T, true>(sv, rel + 64 * (g + Ahead + e));
            }
#pragma unroll
            for (int e = 0; e < Ahead; ++e)
                accumulate(g + e, finishStraight<T, true>(taps, sv, buf[c][e], true));
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Accumulates one straight voice over rows of the tile starting at
// `tileFrame` (row k = frames tileFrame + 64 k + lane), for the frames
// [lo, hi) that the tile produces: contributions x (gains level 2^-29) in
// span order, the same arithmetic for every tile size and path. `panOf(k)`:
// the pan offset of the lane's frame in row k when it varies over the tile.
template<int T, int Rows, class Taps, class PanOf>
__device__ __forceinline__ void straightRows(const int16_t* __restrict__ bank, uint32_t odd, const Taps& taps, const Span& v,
                                             int tileFrame, int lo, int hi, bool panUniform, float pan0,
                                             PanOf&& panOf, float (&left)[Rows], float (&right)[Rows],
                                             PanCache& panCache, int lane)
{
    constexpr int Ahead = straightAhead<T>();
    static_assert(Rows % Ahead == 0, "rows come in fetch groups");
    const int first = max(v.begin, lo), last = min(v.end, hi) - 1;
    if (first > last)
        return;
    const Straight<T> sv(bank, odd, v);
    float2 g0 = make_float2(v.left, v.right);
    if (panUniform)
        g0 = monoGains(v.left, v.right, v.gain, v.pan, pan0, panCache);
    const float2 g0s = make_float2(g0.x * 0x1p-29f, g0.y * 0x1p-29f);
    // Lane offset of row 0 from the voice's first frame: row k's frame is
    // begin + rel + 64 k (rows are never precomputed, so nothing is held
    // across the voice loop but the sums).
    const int rel = tileFrame - v.begin + lane, length = v.end - v.begin;
    if (panUniform && tileFrame >= v.begin && tileFrame + 64 * Rows <= v.end) {
        // The voice covers the whole tile.
        const int nHi = tileFrame + 64 * Rows - 1 - v.begin;
        const int wa = sv.window(sv.place(tileFrame - v.begin)), wb = sv.window(sv.place(nHi));
        const bool interior = uniform(int(wa >= 0 && wb <= sv.lastStart)) != 0;
        const TileEnvelope env(v, tileFrame, tileFrame + 64 * Rows - 1);
        if (interior && (env.constLevel || (env.stageUniform && env.stage.isExp))) {
            const float2 lg0s = make_float2(g0s.x * env.level0, g0s.y * env.level0);
            if (env.constLevel)
                straightRowsCovered<T, Rows, false>(taps, sv, env, g0s, lg0s, rel, left, right);
            else
                straightRowsCovered<T, Rows, true>(taps, sv, env, g0s, lg0s, rel, left, right);
            return;
        }
    }
    const int k0 = (first - tileFrame) >> 6, k1 = (last - tileFrame) >> 6;
    const TileEnvelope env(v, first, last);
    const float2 lg0s = make_float2(g0s.x * env.level0, g0s.y * env.level0);
    const bool constGain = env.constLevel && panUniform;
    // Frame offsets the rows evaluate, clamped to the voice's frames here:
    // lanes outside them compute a valid frame and add nothing.
    const int nLo = first - v.begin, nHi = last - v.begin;
    const int wa = sv.window(sv.place(nLo)), wb = sv.window(sv.place(nHi));
    const bool interior = uniform(int(wa >= 0 && wb <= sv.lastStart)) != 0;
    const auto rows = [&](auto form) {
        constexpr bool Interior = decltype(form)::value;
#pragma unroll
        for (int g = 0; g < Rows; g += Ahead) {
            if (g + Ahead - 1 < k0 || g > k1)
                continue;
            Fetched<T> f[Ahead];
#pragma unroll
            for (int e = 0; e < Ahead; ++e)
                f[e] = fetchStraight<T, Interior>(sv, min(max(rel + 64 * (g + e), nLo), nHi));
#pragma unroll
            for (int e = 0; e < Ahead; ++e) {
                const int k = g + e;
                if (k < k0 || k > k1)
                    continue;
                const int n = rel + 64 * k;
                const bool act = uint32_t(n) < uint32_t(length);
                float x = finishStraight<T, Interior>(taps, sv, f[e], act);
                const int rowFrame = tileFrame + 64 * k;
                if (rowFrame < v.begin || rowFrame + 63 >= v.end)
                    x = act ? x : 0.f;
                float2 lgs = lg0s;
                if (!constGain) {
                    const float level = env.level(v.begin + n);
                    float2 gs = g0s;
                    if (!panUniform) {
                        const float2 g = monoGains(v.left, v.right, v.gain, v.pan, panOf(k), panCache);
                        gs = make_float2(g.x * 0x1p-29f, g.y * 0x1p-29f);
                    }
                    lgs = make_float2(gs.x * level, gs.y * level);
                }
                left[k] = fmaf(x, lgs.x, left[k]);
                right[k] = fmaf(x, lgs.y, right[k]);
            }
        }
    };
    if (interior)
        rows(Flag<true>{});
    else
        rows(Flag<false>{});
}

// Writes the rows of a tile's mix: the frames [lo, hi) of rows starting at
// `tileFrame`, scaled by the job gain (CC7 x CC11, src/cuda/render.cu:189-193),
// plus the wet sums. With `stats` (a whole 1024-frame tail block in one wave)
// it also stores the block's power and peak as clayScanStereo would, and with
// `mono` the frames' mono PCM16 at the unit scale, which clayQuantizeMono
// keeps for every job whose scale turns out to be that one
// (src/cuda/tail.cu:54-81): the same per-frame terms, summed in another order,
// which moves the power by a few units in the last place; the natural-end test
// re-sums serially within 1e-12 of its threshold, so its decisions do not move.
template<bool FloatOut, int Rows>
__device__ __forceinline__ void writeRows(const clay_control_point* __restrict__ points, const Job& j, void* out,
                                          int tileFrame, int tileEnd, int lo, int hi, const float (&left)[Rows],
                                          const float (&right)[Rows], const float (&wetLeft)[Rows],
                                          const float (&wetRight)[Rows], int lane, StereoStats* __restrict__ stats = nullptr,
                                          int statsIndex = 0, int16_t* __restrict__ mono = nullptr)
{
    double power = 0;
    float peak = 0;
    // The gain is one value over the tile when both tracks are flat there.
    const int s0 = controlSegment(points, j.controls[0], j.controls[1], tileFrame);
    const int s1 = controlSegment(points, j.controls[1], j.controls[2], tileFrame);
    const bool flat = points[s0].step == 0 && (s0 + 1 >= j.controls[1] || points[s0 + 1].frame >= tileEnd) &&
                      points[s1].step == 0 && (s1 + 1 >= j.controls[2] || points[s1 + 1].frame >= tileEnd);
    const float amp0 = pointAt(points[s0], tileFrame) * pointAt(points[s1], tileFrame);
#pragma unroll
    for (int k = 0; k < Rows; ++k) {
        const int frame = tileFrame + 64 * k + lane;
        if (frame < lo || frame >= hi)
            continue;
        const float amp = flat ? amp0 : jobGain(points, j, frame);
        const float2 mixed = make_float2(fmaf(left[k], amp, wetLeft[k]), fmaf(right[k], amp, wetRight[k]));
        if (stats) {
            if (!isfinite(mixed.x) || !isfinite(mixed.y)) {
                peak = __int_as_float(0x7f800000);
            } else {
                peak = fmaxf(peak, fmaxf(fabsf(mixed.x), fabsf(mixed.y)));
                power += double(__fmul_rn(mixed.x, mixed.x)) + double(__fmul_rn(mixed.y, mixed.y));
            }
        }
        const uint64_t at = j.mix + uint64_t(frame);
        if (mono)
            mono[at] = int16_t(monoSample(mixed, kUnitScale));
        if (FloatOut) {
            static_cast<float2*>(out)[at] = mixed;
        } else {
            const short l = short(__float2int_rn(fminf(32767.f, fmaxf(-32768.f, mixed.x * 32768.f))));
            const short r = short(__float2int_rn(fminf(32767.f, fmaxf(-32768.f, mixed.y * 32768.f))));
            static_cast<short2*>(out)[at] = make_short2(l, r);
        }
    }
    if (stats) {
        power = waveReduce(power, Add{});
        peak = waveReduce(peak, Max{});
        if (lane == 0)
            stats[statsIndex] = {power, peak, 0};
    }
}

// One 256-frame tile by one wave, any voices. Every lane of the wave calls
// this; nothing in it synchronises across waves.
template<int T, bool FloatOut, bool Simple, class Taps>
__device__ __forceinline__ void mixTileWave(const int16_t* __restrict__ bank, uint32_t odd, const Taps& taps,
                                            const float* __restrict__ sinc, const uint32_t* __restrict__ dummy,
                                            const Span* __restrict__ spans, const Job* __restrict__ jobs,
                                            const int* __restrict__ voices,
                                            const clay_control_point* __restrict__ points,
                                            const Mod* __restrict__ mods, const float* __restrict__ filtered,
                                            const VoiceState* __restrict__ states, void* __restrict__ out,
                                            int segFirst, int segEnd, const Tile& tile, int lane)
{
    const Job& j = jobs[uniform(tile.job)];
    const int tileEnd = tile.frame + tile.frames;
    const int lo = max(tile.frame, segFirst), hi = min(tileEnd, segEnd);
    // The pan offset (CC10 track) is one value over the tile when its segment
    // is flat and ends after the tile.
    const int panAt = controlSegment(points, j.controls[2], j.controls[3], tile.frame);
    const bool panUniform = points[panAt].step == 0 &&
                            (panAt + 1 >= j.controls[3] || points[panAt + 1].frame >= tileEnd);
    const float pan0 = pointAt(points[panAt], tile.frame);
    int frame[kLaneFrames];
    bool on[kLaneFrames];
    float pan[kLaneFrames], left[kLaneFrames], right[kLaneFrames], wetLeft[kLaneFrames], wetRight[kLaneFrames];
#pragma unroll
    for (int k = 0; k < kLaneFrames; ++k) {
        frame[k] = tile.frame + lane + 64 * k;
        on[k] = frame[k] >= lo && frame[k] < hi;
        pan[k] = panUniform ? pan0 : on[k] ? controlAt(points, j.controls[2], j.controls[3], frame[k]) : 0.f;
        left[k] = right[k] = wetLeft[k] = wetRight[k] = 0.f;
    }
    PanCache panCache;
    const int end = tile.first + tile.count;
    for (int b0 = tile.first; b0 < end; b0 += 64) {
        // Span indices of the next 64 voices, one per lane, read out uniformly.
        const int held = b0 + lane < end ? voices[b0 + lane] : 0;
        const int n = min(64, end - b0);
        for (int i = 0; i < n; ++i) {
            const int s = __builtin_amdgcn_readlane(held, i);
            const Span& v = spans[s];
            if (v.filterTrack >= 0) {
                // A filtered voice from the chain scratch, panned when mono (render.cu:458-523).
                int stop = v.end;
                const VoiceState& st = states[v.filterTrack];
                if (st.retired)
                    stop = min(stop, st.silence);  // retired under the voice-silence policy
                const int64_t base = int64_t(v.filtered) - int64_t(max(v.begin, segFirst)) * v.channels;
#pragma unroll
                for (int k = 0; k < kLaneFrames; ++k) {
                    if (!on[k] || frame[k] < v.begin || frame[k] >= stop)
                        continue;
                    const float* src = filtered + (base + int64_t(frame[k]) * v.channels);
                    float2 x = make_float2(src[0], v.channels == 1 ? src[0] : src[1]);
                    if (v.channels == 1) {
                        float offset = pan[k];
                        for (int m = v.firstMod; m < v.firstMod + v.modCount; ++m) {
                            const Mod md = mods[m];
                            if (md.target == CLAY_MOD_PAN)
                                offset += modulationAt(md, points, frame[k]);
                        }
                        const float2 g = monoGains(v.left, v.right, v.gain, v.pan, offset, panCache);
                        x.x *= g.x;
                        x.y *= g.y;
                    }
                    wetLeft[k] += x.x;
                    wetRight[k] += x.y;
                }
            } else if (Simple || v.straight) {
                // A straight voice; in a Simple launch every dry voice is one
                // (checked by the host).
                straightRows<T, kLaneFrames>(bank, odd, taps, v, tile.frame, lo, hi, panUniform, pan0,
                                             [&](int k) { return pan[k]; }, left, right, panCache, lane);
            } else if (v.plain) {
                // A plain voice with a loop or a bend.
                const double bend = v.pitchSlope != 0 ? expm1(v.pitchSlope) : 1.0;
                plainFrames(v, frame, on, pan, panUniform, tile.frame, tileEnd, left, right, panCache,
                            [&](int f, bool act) { return plainSource<T>(bank, taps, v, f, bend, act, dummy); });
            } else {
                // Any other dry voice: the float FIR of voiceAudio, one frame at a time.
#pragma unroll
                for (int k = 0; k < kLaneFrames; ++k) {
                    if (!on[k] || frame[k] < v.begin || frame[k] >= v.end)
                        continue;
                    const float2 x = voiceAudio<T>(bank, GlobalTaps<T>{sinc}, v, points, mods, pan[k], frame[k], false,
                                                   -1.0, panCache);
                    left[k] += x.x;
                    right[k] += x.y;
                }
            }
        }
    }
    writeRows<FloatOut>(points, j, out, tile.frame, tileEnd, lo, hi, left, right, wetLeft, wetRight, lane);
}

// One 1024-frame tile of straight voices by one wave.
template<int T, bool FloatOut, class Taps>
__device__ __forceinline__ void straightTileWave(const int16_t* __restrict__ bank, uint32_t odd, const Taps& taps,
                                                 const Span* __restrict__ spans, const Job* __restrict__ jobs,
                                                 const int* __restrict__ voices,
                                                 const clay_control_point* __restrict__ points, void* __restrict__ out,
                                                 int segFirst, int segEnd, const Tile& tileIn, int lane,
                                                 StereoStats* __restrict__ stereo, int16_t* __restrict__ mono)
{
    static_assert(kStraightFrames == kTailBlock, "a straight tile is one tail block");
    const Tile tile{uniform(tileIn.job), uniform(tileIn.frame), uniform(tileIn.frames), uniform(tileIn.first),
                    uniform(tileIn.count)};
    const Job& j = jobs[tile.job];
    const int tileEnd = tile.frame + tile.frames;
    const int lo = max(tile.frame, segFirst), hi = min(tileEnd, segEnd);
    const int panAt = controlSegment(points, j.controls[2], j.controls[3], tile.frame);
    const bool panUniform = points[panAt].step == 0 &&
                            (panAt + 1 >= j.controls[3] || points[panAt + 1].frame >= tileEnd);
    const float pan0 = pointAt(points[panAt], tile.frame);
    float left[kStraightRows], right[kStraightRows], zero[kStraightRows];
#pragma unroll
    for (int k = 0; k < kStraightRows; ++k)
        left[k] = right[k] = zero[k] = 0.f;
    PanCache panCache;
    const auto panOf = [&](int k) {
        return controlAt(points, j.controls[2], j.controls[3], tile.frame + 64 * k + lane);
    };
    const int end = tile.first + tile.count;
    for (int b0 = tile.first; b0 < end; b0 += 64) {
        const int held = b0 + lane < end ? voices[b0 + lane] : 0;
        const int n = min(64, end - b0);
        for (int i = 0; i < n; ++i)
            straightRows<T, kStraightRows>(bank, odd, taps, spans[__builtin_amdgcn_readlane(held, i)], tile.frame, lo, hi,
                                           panUniform, pan0, panOf, left, right, panCache, lane);
    }
    // The tail block's statistics when the launch renders the whole tile.
    StereoStats* stats = stereo && lo == tile.frame && hi == tileEnd ? stereo : nullptr;
    writeRows<FloatOut>(points, j, out, tile.frame, tileEnd, lo, hi, left, right, zero, zero, lane, stats,
                        j.blockFirst + tile.frame / kTailBlock, stats ? mono : nullptr);
}

// Runs body(taps) with the integer weights: for T <= kIntSharedTaps rows
// 0..kPhases-1 staged in this block's LDS (one block per CU), else the global
// table.
template<int T, class Body>
__device__ __forceinline__ void withIntTaps(const uint32_t* __restrict__ table, Body&& body)
{
    if constexpr (T <= kIntSharedTaps) {
        using S = IntTapsShared<T>;
        __shared__ uint4 shared[kPhases * kRowQuads];
        const uint4* source = reinterpret_cast<const uint4*>(table);
        for (int i = int(threadIdx.x); i < kPhases * 4; i += int(blockDim.x))
            shared[(i >> 2) * kRowQuads + (i & 3)] = source[i];
        __syncthreads();
        body(S{shared});
    } else {
        body(IntTapsGlobal<T>{table});
    }
}

// Work queue of one mix launch: the waves claim tiles one at a time, so a
// launch ends when its work does, not when the wave dealt the costliest tiles
// finishes (tiles differ by their voice counts). Tiles go to XCDs in chunks of
// `Chunk` consecutive tiles, round-robin (tiles are sorted by frame: the same
// frames of consecutive jobs), and a block's waves claim from the counter of
// the XCD the block runs on (blocks are dispatched to the 8 XCDs in turn), so
// neighbouring tiles meet in one L2; a wave whose XCD has no tiles left takes
// them from the next XCD's counter. The runtime clears the counters
// (clayClearQueues) before every launch; the queue's layout is MixQueue
// (device_types.h).
template<int Chunk, int Waves>
struct TileClaims {
    MixQueue* queue;
    int xcd, lane, from = 0;
    __device__ TileClaims(MixQueue* q, int laneIn) : queue(q), xcd(int(blockIdx.x) % kXcds), lane(laneIn) {}
    __device__ __forceinline__ static int tileOf(int x, unsigned n)
    {
        return int(((n / Chunk) * kXcds + unsigned(x)) * Chunk + n % Chunk);
    }
    // The next tile, or -1 when every XCD's share is claimed (tileOf grows
    // with n, so an XCD whose claim falls past the end has none left).
    __device__ __forceinline__ int next(int tiles)
    {
        for (; from < kXcds; ++from) {
            const int x = (xcd + from) % kXcds;
            unsigned n = 0;
            if (lane == 0)
                n = atomicAdd(&queue->xcd[x].next, 1u);
            const int t = tileOf(x, unsigned(uniform(int(n))));
            if (t < tiles)
                return t;
        }
        return -1;
    }
};

} // namespace

// Plain form: one unit, pointers as kernel arguments; tiles [firstTile, firstTile + tileCount).
template<int T, bool FloatOut, bool Simple>
__global__ void __launch_bounds__(64 * kMixWaves) clayMixWaves(
    const int16_t* __restrict__ bank, uint32_t odd, const uint32_t* __restrict__ intTable, const float* __restrict__ sinc,
    const Span* __restrict__ spans, const Job* __restrict__ jobs, const Tile* __restrict__ tiles,
    const int* __restrict__ voices, const clay_control_point* __restrict__ points, const Mod* __restrict__ mods,
    const float* __restrict__ filtered, const VoiceState* __restrict__ states, void* __restrict__ out, int segFirst,
    int segEnd, int firstTile, int tileCount, MixQueue* __restrict__ queue)
{
    const int lane = int(threadIdx.x) & 63;
    TileClaims<64, kMixWaves> claims(queue, lane);
    withIntTaps<T>(intTable, [&](const auto& taps) {
        for (int t; (t = claims.next(tileCount)) >= 0;) {
            const Tile tile = tiles[firstTile + t];
            mixTileWave<T, FloatOut, Simple>(bank, odd, taps, sinc, intTable, spans, jobs, voices, points, mods, filtered, states,
                                     out, segFirst, segEnd, tile, lane);
        }
    });
}

// Persistent form: the 256-frame tiles of every unit of the round.
template<int T, bool FloatOut, bool Simple>
__global__ void __launch_bounds__(64 * kMixWaves) clayRoundMixWaves(const int16_t* __restrict__ bank, uint32_t odd,
                                                                    const uint32_t* __restrict__ intTable,
                                                                    const float* __restrict__ sinc, UnitsPtr units,
                                                                    RoundPtr round, MixQueue* __restrict__ queue)
{
    const int lane = int(threadIdx.x) & 63;
    TileClaims<64, kMixWaves> claims(queue, lane);
    const RoundHeader header = *round;
    if (header.tiles == 0)
        return;  // every unit of the round is straight (clayRoundMixStraight); nothing claimed
    withIntTaps<T>(intTable, [&](const auto& taps) {
        for (int item; (item = claims.next(header.tiles)) >= 0;) {
            const int unit = unitOf(header, item, [&](int i) { return units[i].tileBase; });
            const auto& u = constant(units)[unit];
            const Tile tile = global(u.tiles)[u.tileFirst + item - u.tileBase];
            mixTileWave<T, FloatOut, Simple>(bank, odd, taps, sinc, intTable, global(u.spans), global(u.jobs), global(u.voices),
                                     global(u.points), global(u.mods), global(u.filtered), global(u.states),
                                     global(u.mix), u.segFirst, u.segEnd, tile, lane);
        }
    });
}

// Plain form, straight batches: 1024-frame tiles [firstTile, firstTile + tileCount);
// `stereo` and `mono` (or null) receive the tail-block statistics and the
// unit-scale mono PCM16 of whole tiles.
template<int T, bool FloatOut>
__global__ void __launch_bounds__(64 * kStraightWaves) clayMixStraight(
    const int16_t* __restrict__ bank, uint32_t odd, const uint32_t* __restrict__ intTable, const Span* __restrict__ spans,
    const Job* __restrict__ jobs, const Tile* __restrict__ tiles, const int* __restrict__ voices,
    const clay_control_point* __restrict__ points, void* __restrict__ out, int segFirst, int segEnd, int firstTile,
    int tileCount, MixQueue* __restrict__ queue, StereoStats* __restrict__ stereo, int16_t* __restrict__ mono)
{
    const int lane = int(threadIdx.x) & 63;
    TileClaims<16, kStraightWaves> claims(queue, lane);
    withIntTaps<T>(intTable, [&](const auto& taps) {
        for (int t; (t = claims.next(tileCount)) >= 0;) {
            const Tile tile = tiles[firstTile + t];
            straightTileWave<T, FloatOut>(bank, odd, taps, spans, jobs, voices, points, out, segFirst, segEnd, tile, lane,
                                          stereo, mono);
        }
    });
}

// Persistent form: the 1024-frame tiles of every straight unit of the round.
template<int T, bool FloatOut>
__global__ void __launch_bounds__(64 * kStraightWaves) clayRoundMixStraight(const int16_t* __restrict__ bank, uint32_t odd,
                                                                       const uint32_t* __restrict__ intTable,
                                                                       UnitsPtr units, RoundPtr round,
                                                                       MixQueue* __restrict__ queue)
{
    const int lane = int(threadIdx.x) & 63;
    TileClaims<16, kStraightWaves> claims(queue, lane);
    const RoundHeader header = *round;
    if (header.straightTiles == 0)
        return;  // nothing claimed
    withIntTaps<T>(intTable, [&](const auto& taps) {
        for (int item; (item = claims.next(header.straightTiles)) >= 0;) {
            const int unit = unitOf(header, item, [&](int i) { return units[i].straightBase; });
            const auto& u = constant(units)[unit];
            const Tile tile = global(u.straightTiles)[u.straightFirst + item - u.straightBase];
            straightTileWave<T, FloatOut>(bank, odd, taps, global(u.spans), global(u.jobs), global(u.straightVoices),
                                          global(u.points), global(u.mix), u.segFirst, u.segEnd, tile, lane,
                                          global(u.stereo), global(u.mono));
        }
    });
}

#define CLAY_INSTANTIATE_MIX_WAVES(T, FloatOut, Simple)                                                                \
    template __global__ void clayMixWaves<T, FloatOut, Simple>(const int16_t*, uint32_t, const uint32_t*, const float*, \
                                                               const Span*, const Job*, const Tile*, const int*,      \
                                                               const clay_control_point*, const Mod*, const float*,   \
                                                               const VoiceState*, void*, int, int, int, int,          \
                                                               MixQueue*);                                            \
    template __global__ void clayRoundMixWaves<T, FloatOut, Simple>(const int16_t*, uint32_t, const uint32_t*,        \
                                                                    const float*,                                      \
                                                                    UnitsPtr, RoundPtr, MixQueue*);
#define CLAY_INSTANTIATE_MIX_STRAIGHT(T, FloatOut)                                                                     \
    template __global__ void clayMixStraight<T, FloatOut>(const int16_t*, uint32_t, const uint32_t*, const Span*,     \
                                                          const Job*,                                                 \
                                                          const Tile*, const int*, const clay_control_point*, void*,  \
                                                          int, int, int, int, MixQueue*, StereoStats*, int16_t*);     \
    template __global__ void clayRoundMixStraight<T, FloatOut>(const int16_t*, uint32_t, const uint32_t*, UnitsPtr,   \
                                                               RoundPtr, MixQueue*);
#define CLAY_INSTANTIATE_MIX_WAVES_TAPS(T)                                                                             \
    CLAY_INSTANTIATE_MIX_WAVES(T, true, true)                                                                          \
    CLAY_INSTANTIATE_MIX_WAVES(T, false, true)                                                                         \
    CLAY_INSTANTIATE_MIX_WAVES(T, true, false)                                                                         \
    CLAY_INSTANTIATE_MIX_WAVES(T, false, false)                                                                        \
    CLAY_INSTANTIATE_MIX_STRAIGHT(T, true)                                                                             \
    CLAY_INSTANTIATE_MIX_STRAIGHT(T, false)
