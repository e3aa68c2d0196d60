// Output mix of the render core for gfx950 (src/cuda/render.cu:21-187, :458-523).
//
// One wave renders one 256-frame tile of one job: lane l holds the frames
// l, l + 64, l + 128 and l + 192 of the tile, and the wave walks the tile's
// voices in span order, so the per-voice work (the span record in scalar
// registers, the voice's kind, its envelope constants) is paid once per 256
// frames instead of once per 64. Sixteen waves share a 1024-thread block, one
// block per CU, grid-stride over the tiles of an XCD's share of the window.
//
// Plain voices (mono, sampled, dry, no modulation of the audio) take an
// integer FIR: the sinc weights are stored as packed int16 pairs,
//     w = hi * 2^-14 + lo * 2^-24,   hi = rint(w 2^14),  lo = rint((w - hi 2^-14) 2^24),
// exact to 2^-25, and the int16 sample pairs of the bank are dotted with them
// by v_dot2c_i32_i16, two taps per instruction, in exact int32 sums
// (|sum hi s| < 1.6 * 2^29 and |sum lo s| <= T 2^9 2^15 <= 2^30 for T <= 64).
// A 16-tap FIR is 16 dot instructions where the float FIR takes 16 converts
// and 16 FMAs. For 16 taps the weight table lives in LDS as an XOR-swizzled
// table (128 KB). The envelope of a plain voice uses one exp2 per frame.
// Other dry voices use the float FIR of voiceAudio, and filtered voices are
// read from the chain scratch; every execution form inlines these same
// functions, so their PCM is bit-identical.
#include "common.h"

namespace {

constexpr int kMixWaves = 16;                  // tiles in flight per 1024-thread block
constexpr int kLaneFrames = kTileFrames / 64;  // frames per lane of a tile
constexpr int kXcds = 8;                       // MI355X: 8 XCDs, blocks dispatched round-robin over them
constexpr int kXcdChunk = 64;                  // consecutive tiles (16384 frames of a job) kept on one XCD
constexpr int kIntSharedTaps = 16;             // widest integer table held in LDS

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

// The same full table in LDS (128 KB at 16 taps), its 16-byte quads stored in
// XOR-swizzled order (quad q of row r at q ^ (r & 3)): rows are 64 bytes, so
// unswizzled the first quad of every row would fall on one of two bank
// groups; swizzled, the rows that the lanes of a wave read start on eight.
template<int T>
struct IntTapsShared {
    static constexpr int kQuads = T / 4;
    static_assert(kQuads == 4, "the LDS table is laid out for 16 taps");
    const uint4* table;
    struct Row {
        const uint4* w;
        int x;  // r & 3
        __device__ __forceinline__ uint4 quad(int q) const { return w[q ^ x]; }
        __device__ __forceinline__ uint32_t word(int i) const
        {
            return reinterpret_cast<const uint32_t*>(w)[(((i >> 2) ^ x) << 2) | (i & 3)];
        }
    };
    __device__ __forceinline__ Row row(int phase) const { return {table + phase * kQuads, phase & 3}; }
};

__device__ __forceinline__ int dot2(uint32_t a, uint32_t b, int c)
{
    typedef short v2s __attribute__((ext_vector_type(2)));
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, a), __builtin_bit_cast(v2s, b), c, false);
}

// Source value from the two integer sums: (hi + lo 2^-10) 2^-14, scaled by 2^-15.
__device__ __forceinline__ float fromSums(int hi, int lo) { return fmaf(float(lo), 0x1p-10f, float(hi)) * 0x1p-29f; }

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

// The same sums one tap at a time, zero outside [0, sourceEnd): windows at the
// edges of the source.
template<int T, class Taps>
__device__ __forceinline__ float firEdge(const int16_t* __restrict__ bank, uint64_t sample, int sourceEnd, int centre,
                                         const typename Taps::Row& row)
{
    int hi = 0, lo = 0;
#pragma unroll 1
    for (int t = 0; t < T; ++t) {
        const int i = centre - (T / 2 - 1) + t;
        const int s = i >= 0 && i < sourceEnd ? int(bank[sample + uint64_t(i)]) : 0;
        const uint32_t h = row.word(t >> 1), l = row.word(T / 2 + (t >> 1));
        const int wh = int(int16_t(t & 1 ? h >> 16 : h & 0xffffu));
        const int wl = int(int16_t(t & 1 ? l >> 16 : l & 0xffffu));
        hi += wh * s;
        lo += wl * s;
    }
    return fromSums(hi, lo);
}

// Source value of a plain voice at a position (src/cuda/render.cu:21-61): the
// exact sample at an integer position, else the FIR at the nearest of 2048
// phases. `active`: the lane renders this frame; inactive lanes load from a
// valid dummy address and take no edge path.
template<int T, class Taps>
__device__ __forceinline__ float sourceAt(const int16_t* __restrict__ bank, const Taps& taps, const Span& v, double pos,
                                          bool active, const uint32_t* __restrict__ dummy)
{
    const double floorPos = floor(pos);
    const int centre = int(floorPos);
    const int phase = __double2int_rn((pos - floorPos) * kPhases);
    const auto row = taps.row(min(max(phase, 0), kPhases));
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
    if (pos == floorPos) {
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

// Per-voice constants of a straight plain voice (no loop, no bend): the
// voice's samples as a base pointer with 32-bit offsets, and the last window
// start that stays inside the source.
struct Straight {
    const int16_t* __restrict__ base;  // bank + sample - parity: dword aligned
    uint64_t sample;                   // bank offset of frame 0
    int parity;                        // sample & 1
    int sourceEnd;                     // source frames
    int last;                          // last interior window start (centre - (T/2 - 1)), -1 for none
};

// Source value of a straight plain voice at `pos` (src/cuda/render.cu:21-61),
// as sourceAt but with every lane loading a window clamped into the source
// (the bank is padded by 2 x CLAY_MAX_SINC_TAPS samples), so no lane takes a
// different address path; edges and integer positions outside the interior
// take the rare branch. In two halves, so that a caller can issue the loads of
// several frames before it waits for any of them.
template<int T>
struct Window {
    uint32_t dw[T / 2 + 1];  // the window's dwords, from the dword-aligned start
    double frac;             // pos - floor(pos)
    int centre, phase;
    uint32_t shift;          // 16 when the window starts on the high half of its first dword
    bool interior;
};

template<int T>
__device__ __forceinline__ Window<T> fetchWindow(const Straight& sv, double pos)
{
    Window<T> w;
    const double floorPos = floor(pos);
    w.centre = int(floorPos);
    w.frac = pos - floorPos;
    w.phase = __double2int_rn(w.frac * kPhases);  // 0..kPhases: 0 <= frac < 1
    const int start = w.centre - (T / 2 - 1);
    w.interior = sv.last >= 0 && uint32_t(start) <= uint32_t(sv.last);
    const int at = min(max(start, 0), max(sv.last, 0)) + sv.parity;  // window start from `base`
    w.shift = uint32_t(at & 1) << 4;
    const uint32_t* __restrict__ words = reinterpret_cast<const uint32_t*>(sv.base + uint32_t(at & ~1));
#pragma unroll
    for (int i = 0; i < T / 2 + 1; ++i)
        w.dw[i] = words[i];
    return w;
}

template<int T, class Taps>
__device__ __forceinline__ float finishWindow(const int16_t* __restrict__ bank, const Taps& taps, const Straight& sv,
                                              const Window<T>& w, bool active)
{
    const auto row = taps.row(w.phase);
    uint32_t p[T / 2];
#pragma unroll
    for (int i = 0; i < T / 2; ++i)
        p[i] = __builtin_amdgcn_alignbit(w.dw[i + 1], w.dw[i], w.shift);
    float s = firPairs<T, Taps>(row, p);
    // Exact sample at an integer position (render.cu:33-34): window tap T/2 - 1.
    const uint32_t pair = p[(T / 2 - 1) >> 1];
    const int exact = int(int16_t(((T / 2 - 1) & 1) ? pair >> 16 : pair & 0xffffu));
    s = w.frac == 0 ? float(exact) * 0x1p-15f : s;
    if (active && !w.interior && w.centre < sv.sourceEnd) {
        // Window past either end of the source.
        s = w.frac == 0 ? (w.centre >= 0 ? float(bank[sv.sample + uint64_t(w.centre)]) * 0x1p-15f : 0.f)
                        : firEdge<T, Taps>(bank, sv.sample, sv.sourceEnd, w.centre, row);
    }
    return w.centre < sv.sourceEnd ? s : 0.f;
}

template<int T, class Taps>
__device__ __forceinline__ float straightSource(const int16_t* __restrict__ bank, const Taps& taps, const Straight& sv,
                                                double pos, bool active)
{
    return finishWindow<T>(bank, taps, sv, fetchWindow<T>(sv, pos), active);
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

template<int T, class Taps>
__device__ __forceinline__ float crossfadeEntry(const int16_t* __restrict__ bank, const Taps& taps, const Span& v,
                                             double pos, bool active, const uint32_t* __restrict__ dummy)
{
    return sourceAt<T>(bank, taps, v, pos, active, dummy);
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
        const float entry = crossfadeEntry<T>(bank, taps, v, __dsub_rn(pos, double(period)), active, dummy);
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

// Accumulates a plain voice's frames into the lane's dry sums: the envelope
// stage and the pan gains are taken once per voice when they hold over the
// whole tile (the usual case), per frame otherwise.
// `fetch(frame)` starts a frame's source value (its loads) and `finish(started,
// active)` completes it; `Ahead` frames of the lane are fetched before the
// first of them is finished, so their loads are in flight together (as many
// as the registers allow: four windows at 16 taps).
template<int Ahead, class Fetch, class Finish>
__device__ __forceinline__ void plainFrames(const Span& span, const int (&frame)[kLaneFrames],
                                            const bool (&on)[kLaneFrames], const float (&pan)[kLaneFrames],
                                            bool panUniform, int tileFirst, int tileEnd, float (&left)[kLaneFrames],
                                            float (&right)[kLaneFrames], PanCache& panCache, Fetch&& fetch,
                                            Finish&& finish)
{
    const EnvFields v = envFields(span);
    const int begin = span.begin, end = span.end;
    const float gl = span.left, gr = span.right, gain = span.gain, voicePan = span.pan;
    const float decay2 = v.decayRate * 1.44269504088896341f, release2 = v.releaseRate * 1.44269504088896341f;
    // The voice's frames in this tile are [first, last]; one stage over all of them?
    const int first = max(begin, tileFirst), last = min(end, tileEnd) - 1;
    const int k0 = envStageIndex(v, first), k1 = envStageIndex(v, last);
    const bool stageUniform = k0 == k1;
    EnvStage uniformStage = envStage(v, decay2, release2, k0);
    // A decay that has fallen below the sustain level at the tile's first frame
    // stays below it (the exponential only falls), so the level is the sustain
    // level itself; the margin covers the last-place error of v_exp_f32.
    if (stageUniform && k0 == 3 && uniformStage.isExp && v.decayRate < 0 &&
        envLevel(uniformStage, first) == uniformStage.lo &&
        __builtin_amdgcn_exp2f(float(first - uniformStage.origin + 1) * uniformStage.rate) <
            uniformStage.lo * (1.f - 0x1p-20f)) {
        uniformStage.isExp = false;
        uniformStage.slope = 0.f;
        uniformStage.base = uniformStage.lo;
    }
    const bool constLevel = stageUniform && !uniformStage.isExp && uniformStage.slope == 0;
    const float level0 = fminf(uniformStage.hi, fmaxf(uniformStage.lo, uniformStage.base));
    float2 g0 = make_float2(gl, gr);
    if (panUniform)
        g0 = monoGains(gl, gr, gain, voicePan, pan[0], panCache);
    const float2 lg0 = make_float2(g0.x * level0, g0.y * level0);
    decltype(fetch(0)) started[kLaneFrames];
#pragma unroll
    for (int k = 0; k < kLaneFrames; ++k) {
        if (k % Ahead == 0) {
#pragma unroll
            for (int a = k; a < k + Ahead; ++a)
                started[a] = fetch(frame[a]);
        }
        const bool act = on[k] && frame[k] >= begin && frame[k] < end;
        const float src = finish(started[k], act);
        float2 lg;
        if (constLevel && panUniform) {
            lg = lg0;
        } else {
            const float level = constLevel ? level0
                                : stageUniform ? envLevel(uniformStage, frame[k])
                                               : envLevel(envStage(v, decay2, release2, envStageIndex(v, frame[k])),
                                                          frame[k]);
            const float2 g = panUniform ? g0 : monoGains(gl, gr, gain, voicePan, pan[k], panCache);
            lg = make_float2(g.x * level, g.y * level);
        }
        if (act) {
            left[k] = fmaf(src, lg.x, left[k]);
            right[k] = fmaf(src, lg.y, right[k]);
        }
    }
}

// One tile by one wave. Every lane of the wave calls this; nothing in it
// synchronises across waves.
template<int T, bool FloatOut, bool Simple, class Taps>
__device__ __forceinline__ void mixTileWave(const int16_t* __restrict__ bank, const Taps& taps,
                                            const float* __restrict__ sinc, const uint32_t* __restrict__ dummy,
                                            const Span* __restrict__ spans, const Job* __restrict__ jobs,
                                            const int* __restrict__ voices,
                                            const clay_control_point* __restrict__ points,
                                            const Mod* __restrict__ mods, const float* __restrict__ filtered,
                                            const VoiceState* __restrict__ states, void* __restrict__ out,
                                            int segFirst, int segEnd, const Tile& tile, int lane)
{
    const Job& j = jobs[uniform(tile.job)];
    // The pan offset (CC10 track) is one value over the tile when its segment
    // is flat and ends after the tile.
    const int panAt = controlSegment(points, j.controls[2], j.controls[3], tile.frame);
    const bool panUniform = points[panAt].step == 0 &&
                            (panAt + 1 >= j.controls[3] || points[panAt + 1].frame >= tile.frame + tile.frames);
    int frame[kLaneFrames];
    bool on[kLaneFrames];
    float pan[kLaneFrames], left[kLaneFrames], right[kLaneFrames], wetLeft[kLaneFrames], wetRight[kLaneFrames];
#pragma unroll
    for (int k = 0; k < kLaneFrames; ++k) {
        frame[k] = tile.frame + lane + 64 * k;
        on[k] = lane + 64 * k < tile.frames && frame[k] >= segFirst && frame[k] < segEnd;
        pan[k] = panUniform ? pointAt(points[panAt], tile.frame)
                            : on[k] ? controlAt(points, j.controls[2], j.controls[3], frame[k]) : 0.f;
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
            } else if (Simple || (v.plain && v.loopEnd <= 0 && v.pitchSlope == 0)) {
                // A straight plain voice; in a Simple launch every dry voice is
                // one (checked by the host).
                const uint64_t sample = v.sample;
                const int sourceEnd = v.sourceEnd;
                const Straight sv{bank + (sample - (sample & 1)), sample, int(sample & 1), sourceEnd, sourceEnd - T};
                const double position = v.position, step = v.step;
                const int begin = v.begin;
                plainFrames<T <= 16 ? kLaneFrames : T <= 32 ? 2 : 1>(
                    v, frame, on, pan, panUniform, tile.frame, tile.frame + tile.frames, left, right, panCache,
                    [&](int f) { return fetchWindow<T>(sv, __dadd_rn(position, __dmul_rn(double(f - begin), step))); },
                    [&](const Window<T>& w, bool act) { return finishWindow<T>(bank, taps, sv, w, act); });
            } else if (v.plain) {
                // A plain voice with a loop or a bend.
                const double bend = v.pitchSlope != 0 ? expm1(v.pitchSlope) : 1.0;
                plainFrames<1>(v, frame, on, pan, panUniform, tile.frame, tile.frame + tile.frames, left, right, panCache,
                            [&](int f) { return f; },
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
#pragma unroll
    for (int k = 0; k < kLaneFrames; ++k) {
        if (!on[k])
            continue;
        const float amp = jobGain(points, j, frame[k]);
        const float2 mixed = make_float2(fmaf(left[k], amp, wetLeft[k]), fmaf(right[k], amp, wetRight[k]));
        const uint64_t at = j.mix + uint64_t(frame[k]);
        if (FloatOut) {
            static_cast<float2*>(out)[at] = mixed;
        } else {
            const short l = short(__float2int_rn(fminf(32767.f, fmaxf(-32768.f, mixed.x * 32768.f))));
            const short r = short(__float2int_rn(fminf(32767.f, fmaxf(-32768.f, mixed.y * 32768.f))));
            static_cast<short2*>(out)[at] = make_short2(l, r);
        }
    }
}

// Runs body(taps) with the integer weights: for T <= kIntSharedTaps the full
// table staged in this block's LDS (one block per CU), else the global table.
template<int T, class Body>
__device__ __forceinline__ void withIntTaps(const uint32_t* __restrict__ table, Body&& body)
{
    if constexpr (T <= kIntSharedTaps) {
        using S = IntTapsShared<T>;
        __shared__ uint4 shared[(kPhases + 1) * S::kQuads];
        const uint4* source = reinterpret_cast<const uint4*>(table);
        for (int i = int(threadIdx.x); i < (kPhases + 1) * S::kQuads; i += int(blockDim.x)) {
            const int row = i / S::kQuads, q = i % S::kQuads;
            shared[row * S::kQuads + (q ^ (row & 3))] = source[i];
        }
        __syncthreads();
        body(S{shared});
    } else {
        body(IntTapsGlobal<T>{table});
    }
}

// Tile n of this wave's XCD share: tiles go to XCDs in chunks of kXcdChunk
// consecutive tiles, round-robin, so the consecutive tiles of a job (the same
// voices, the same source windows) meet in one L2.
struct XcdWalk {
    int xcd, wave, waves;
    __device__ XcdWalk(int waveInBlock)
    {
        const int perXcd = max(1, int(gridDim.x) / kXcds);
        xcd = int(blockIdx.x) % kXcds;
        wave = (int(blockIdx.x) / kXcds) * kMixWaves + waveInBlock;
        waves = perXcd * kMixWaves;
        if (int(blockIdx.x) >= perXcd * kXcds)
            wave = -1;  // a block past the last full group of kXcds renders nothing
    }
    __device__ __forceinline__ int tile(int n) const { return ((n / kXcdChunk) * kXcds + xcd) * kXcdChunk + n % kXcdChunk; }
};

} // namespace

// Plain form: one unit, pointers as kernel arguments; tiles [firstTile, firstTile + tileCount).
template<int T, bool FloatOut, bool Simple>
__global__ void __launch_bounds__(64 * kMixWaves) clayMixWaves(
    const int16_t* __restrict__ bank, const uint32_t* __restrict__ intTable, const float* __restrict__ sinc,
    const Span* __restrict__ spans, const Job* __restrict__ jobs, const Tile* __restrict__ tiles,
    const int* __restrict__ voices, const clay_control_point* __restrict__ points, const Mod* __restrict__ mods,
    const float* __restrict__ filtered, const VoiceState* __restrict__ states, void* __restrict__ out, int segFirst,
    int segEnd, int firstTile, int tileCount)
{
    const int lane = int(threadIdx.x) & 63;
    const XcdWalk walk(uniform(int(threadIdx.x) >> 6));
    withIntTaps<T>(intTable, [&](const auto& taps) {
        if (walk.wave < 0)
            return;
        for (int n = walk.wave;; n += walk.waves) {
            const int t = walk.tile(n);
            if (t >= tileCount)
                break;
            const Tile tile = tiles[firstTile + t];
            mixTileWave<T, FloatOut, Simple>(bank, taps, sinc, intTable, spans, jobs, voices, points, mods, filtered, states,
                                     out, segFirst, segEnd, tile, lane);
        }
    });
}

// Persistent form: the tiles of every unit of the round.
template<int T, bool FloatOut, bool Simple>
__global__ void __launch_bounds__(64 * kMixWaves) clayRoundMixWaves(const int16_t* __restrict__ bank,
                                                                    const uint32_t* __restrict__ intTable,
                                                                    const float* __restrict__ sinc, UnitsPtr units,
                                                                    RoundPtr round)
{
    const int lane = int(threadIdx.x) & 63;
    const XcdWalk walk(uniform(int(threadIdx.x) >> 6));
    const RoundHeader header = *round;
    withIntTaps<T>(intTable, [&](const auto& taps) {
        if (walk.wave < 0)
            return;
        for (int n = walk.wave;; n += walk.waves) {
            const int item = walk.tile(n);
            if (item >= header.tiles)
                break;
            const int unit = unitOf(header, item, [&](int i) { return units[i].tileBase; });
            const auto& u = constant(units)[unit];
            const Tile tile = global(u.tiles)[u.tileFirst + item - u.tileBase];
            mixTileWave<T, FloatOut, Simple>(bank, taps, sinc, intTable, global(u.spans), global(u.jobs), global(u.voices),
                                     global(u.points), global(u.mods), global(u.filtered), global(u.states),
                                     global(u.mix), u.segFirst, u.segEnd, tile, lane);
        }
    });
}

#define CLAY_INSTANTIATE_MIX_WAVES(T, FloatOut, Simple)                                                                \
    template __global__ void clayMixWaves<T, FloatOut, Simple>(const int16_t*, const uint32_t*, const float*,         \
                                                               const Span*, const Job*, const Tile*, const int*,      \
                                                               const clay_control_point*, const Mod*, const float*,   \
                                                               const VoiceState*, void*, int, int, int, int);          \
    template __global__ void clayRoundMixWaves<T, FloatOut, Simple>(const int16_t*, const uint32_t*, const float*,    \
                                                                    UnitsPtr, RoundPtr);
#define CLAY_INSTANTIATE_MIX_WAVES_TAPS(T)                                                                             \
    CLAY_INSTANTIATE_MIX_WAVES(T, true, true)                                                                          \
    CLAY_INSTANTIATE_MIX_WAVES(T, false, true)                                                                         \
    CLAY_INSTANTIATE_MIX_WAVES(T, true, false)                                                                         \
    CLAY_INSTANTIATE_MIX_WAVES(T, false, false)
CLAY_INSTANTIATE_MIX_WAVES_TAPS(16)
CLAY_INSTANTIATE_MIX_WAVES_TAPS(32)
CLAY_INSTANTIATE_MIX_WAVES_TAPS(48)
CLAY_INSTANTIATE_MIX_WAVES_TAPS(64)
