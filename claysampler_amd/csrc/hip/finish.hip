// Finishing kernels (tail policies, PCM16, quality gate, group volume) for
// gfx950. Reference behaviour: src/cuda/tail.cu and src/cuda/quality.cu.
// Block reductions are wave64 shuffles combined across at most 16 waves.
#include "common.h"

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

// Natural endpoint, silence-beyond-plan rule and peak repair, one wave per job
// (src/cuda/tail.cu:83-145). The reference walks the 1024-frame blocks in
// order and stops at the first non-finite block or the first block that ends
// the tail; here the lanes test the blocks in parallel, the first of each kind
// is a wave minimum, and the peak is the maximum over the blocks the walk
// would have visited.
extern "C" __global__ void __launch_bounds__(64) claySelectNative(const Job* __restrict__ jobs, int count,
                                                                  const float2* __restrict__ mix,
                                                                  const StereoStats* __restrict__ stats,
                                                                  clay_tail_result* __restrict__ results)
{
    const int index = int(blockIdx.x), lane = int(threadIdx.x);
    if (index >= count)
        return;
    const Job j = jobs[index];
    const int blocks = (j.frames + kTailBlock - 1) / kTailBlock;
    constexpr int kNone = 0x7fffffff;
    int firstBad = kNone, firstEnd = kNone;
    for (int b = lane; b < blocks; b += 64) {
        const StereoStats st = stats[j.blockFirst + b];
        if (!isfinite(st.peak)) {
            firstBad = b;
            break;  // later blocks of this lane come after it
        }
        const int first = b * kTailBlock;
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
            firstEnd = b;
            break;
        }
    }
    firstBad = waveReduce(firstBad, Min{});
    firstEnd = waveReduce(firstEnd, Min{});
    clay_tail_result r{};
    if (firstBad < firstEnd) {
        r.status = CLAY_TAIL_NONFINITE;
        if (lane == 0)
            results[index] = r;
        return;
    }
    const int visited = firstEnd == kNone ? blocks : firstEnd + 1;
    float peak = 0;
    for (int b = lane; b < visited; b += 64)
        peak = fmaxf(peak, stats[j.blockFirst + b].peak);
    peak = waveReduce(peak, Max{});
    if (lane)
        return;
    int end = firstEnd == kNone ? 0 : min(j.limit, firstEnd * kTailBlock + kTailBlock);
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
// With `monoReady` the mix kernel has already written every frame's mono PCM16
// at the unit scale: a job at that scale (peak below 1) only reads it back for
// the window statistics, 2 bytes a frame instead of the 8-byte float mix.
// One wave per window, kQuantizeWaves windows per block, wave reductions only
// (the power is a sum of integers below 2^53, exact in any order).
extern "C" __global__ void __launch_bounds__(64 * kQuantizeWaves) clayQuantizeMono(
    const Job* __restrict__ jobs, int count, const float2* __restrict__ mix,
    const clay_tail_result* __restrict__ results, int16_t* __restrict__ mono, MonoStats* __restrict__ stats,
    int window, int firstBlock, int blocks, int monoReady)
{
    const int lane = int(threadIdx.x) & 63;
    const int slot = int(blockIdx.x) * kQuantizeWaves + int(threadIdx.x >> 6);
    if (slot >= blocks)
        return;  // uniform per wave; no block barriers
    const int block = firstBlock + slot;
    const int index = ownerOf(count, block, [&](int i) { return jobs[i].windowFirst; });
    const Job j = jobs[index];
    const clay_tail_result r = results[index];
    const int first = (block - j.windowFirst) * window;
    const int end = min(first + window, r.data_frames);
    if (first >= end)
        return;
    double power = 0;
    int peak = 0;
    // kAhead of the lane's frames are loaded before any is used (one
    // dependent load per frame left the kernel bound by latency).
    constexpr int kAhead = 4;
    const float2* src = mix + j.mix;
    int16_t* dst = mono + j.mix;
    const bool ready = monoReady && r.scale == kUnitScale;
    for (int f0 = first + lane; f0 < end; f0 += kAhead * 64) {
        if (ready) {
            int x[kAhead];
#pragma unroll
            for (int k = 0; k < kAhead; ++k) {
                const int f = f0 + k * 64;
                x[k] = f < end ? int(dst[f]) : 0;
            }
#pragma unroll
            for (int k = 0; k < kAhead; ++k) {
                power += double(x[k]) * x[k];
                peak = max(peak, abs(x[k]));
            }
        } else {
            float2 x[kAhead];
#pragma unroll
            for (int k = 0; k < kAhead; ++k) {
                const int f = f0 + k * 64;
                x[k] = f < end ? src[f] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int k = 0; k < kAhead; ++k) {
                const int f = f0 + k * 64;
                if (f >= end)
                    break;
                const int v = monoSample(x[k], r.scale);
                dst[f] = int16_t(v);
                power += double(v) * v;
                peak = max(peak, abs(v));
            }
        }
    }
    power = waveReduce(power, Add{});
    peak = waveReduce(peak, Max{});
    if (lane == 0)
        stats[block] = {power, peak, 0};
}

// Weak-tail trim, one wave per job (src/cuda/tail.cu:183-213): the loudest
// window and the last window above the threshold as wave reductions.
extern "C" __global__ void __launch_bounds__(64) claySelectTrim(const Job* __restrict__ jobs, int count,
                                                                const MonoStats* __restrict__ stats,
                                                                clay_tail_result* __restrict__ results, TailConfig cfg)
{
    const int index = int(blockIdx.x), lane = int(threadIdx.x);
    if (index >= count || results[index].status != CLAY_TAIL_OK)
        return;
    const clay_tail_result r0 = results[index];
    const MonoStats* w = stats + jobs[index].windowFirst;
    const int windows = (r0.data_frames + cfg.window - 1) / cfg.window;
    double maxPower = 0;
    for (int i = lane; i < windows; i += 64)
        maxPower = fmax(maxPower, w[i].power);
    maxPower = waveReduce(maxPower, Max{});
    const double threshold = fmin(cfg.ceiling, fmax(cfg.floor, sqrt(maxPower / cfg.window) * cfg.relative));
    int last = -1;
    for (int i = lane; i < windows; i += 64)
        if (sqrt(w[i].power / cfg.window) > threshold || w[i].peak > threshold * 4)
            last = i;
    last = waveReduce(last, Max{});
    if (lane || last < 0)
        return;
    clay_tail_result& r = results[index];
    const int lastSound = min(r.frames, (last + 1) * cfg.window);
    if (r.frames - lastSound >= cfg.quietFrames) {
        r.frames = lastSound + cfg.padding;
        r.data_frames = min(r.frames, r.data_frames);
        r.fade = min(r.frames, cfg.fade);
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
// (src/cuda/tail.cu:240-297). One wave takes kFinalizeBlocks consecutive
// blocks (`blocks` in all from `firstBlock`), so the owner search, the job and
// result loads and the atomics are paid once per run of blocks of one job; the
// reductions are wave shuffles, with no LDS round or block barrier. `channels`
// is 1 (natural) or 2. In place (orig/custom) the output offset of a job equals
// its mix offset, so every frame is read and written at one address.
// The kernel for a compile-time channel count: the job's input and output
// are addressed from per-job base pointers with 32-bit frame offsets, which
// keeps the sixteen loads of a lane in flight within the register budget of
// eight waves per SIMD.
template<int Ch>
__device__ __forceinline__ void finalizeRun(const Job* __restrict__ jobs, int count, const int16_t* __restrict__ input,
                                            int16_t* __restrict__ output, clay_tail_result* __restrict__ results,
                                            int countClipped, int begin, int stopBlock, int lane)
{
    int index = ownerOf(count, begin, [&](int i) { return jobs[i].blockFirst; });
    int peak = 0, audibleFirst = 0x7fffffff, audibleLast = -1, clipped = 0;
    // The run's reductions into the job's result, once per job.
    const auto flush = [&] {
        peak = waveReduce(peak, Max{});
        audibleFirst = waveReduce(audibleFirst, Min{});
        audibleLast = waveReduce(audibleLast, Max{});
        if (countClipped) {
            clipped = waveReduce(clipped, Add{});
            if (lane == 0 && clipped)
                atomicAdd(&results[index].audio.clipped, clipped);
        }
        if (lane == 0 && audibleLast >= 0) {
            atomicMax(&results[index].peak, peak);
            atomicMin(&results[index].first, audibleFirst);
            atomicMax(&results[index].last, audibleLast);
        }
        peak = 0, audibleFirst = 0x7fffffff, audibleLast = -1, clipped = 0;
    };
    constexpr double pi = 3.14159265358979323846;
    constexpr int kPer = kTailBlock / 64;
    for (int block = begin; block < stopBlock; ++block) {
        if (index + 1 < count && jobs[index + 1].blockFirst <= block) {
            flush();
            do
                ++index;
            while (index + 1 < count && jobs[index + 1].blockFirst <= block);
        }
        const Job j = jobs[index];
        const int16_t* src = input + uint64_t(j.mix) * Ch;
        int16_t* dst = output + uint64_t(results[index].offset) * Ch;
        const int frames = results[index].frames, fade = results[index].fade;
        const int first = (block - j.blockFirst) * kTailBlock;
        const int end = min(first + kTailBlock, results[index].data_frames);
        const int fadeStart = frames - fade;
        // Blocks before the fade: every load of the lane's frames is issued
        // before any is used (one 128-byte wave load at a time left the kernel
        // bound by latency), unrolled. Blocks that reach the fade take a plain
        // loop, so the FP64 cosine is compiled once rather than per frame slot
        // (unrolled, it made the kernel too large for the instruction cache).
        const auto frame = [&](int f, int c, int v) {
            dst[f * Ch + c] = int16_t(v);
            clipped += abs(v) >= 32767;
            return abs(v);
        };
        const auto audible = [&](int f, int framePeak) {
            peak = max(peak, framePeak);
            if (framePeak > 3) {
                audibleFirst = min(audibleFirst, f);
                audibleLast = f;
            }
        };
        if (end <= fadeStart) {
            int16_t in[kPer][Ch];
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const int f = first + lane + 64 * i;
#pragma unroll
                for (int c = 0; c < Ch; ++c)
                    in[i][c] = f < end ? src[f * Ch + c] : int16_t(0);
            }
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const int f = first + lane + 64 * i;
                if (f >= end)
                    break;
                int framePeak = 0;
#pragma unroll
                for (int c = 0; c < Ch; ++c)
                    framePeak = max(framePeak, frame(f, c, in[i][c]));
                audible(f, framePeak);
            }
        } else {
#pragma unroll 1
            for (int f = first + lane; f < end; f += 64) {
                int framePeak = 0;
                for (int c = 0; c < Ch; ++c) {
                    int v = src[f * Ch + c];
                    if (f >= fadeStart) {
                        const int k = f - fadeStart;
                        const double gain = fade == 1 ? 0 : .5 * (1 + cos(pi * k / (fade - 1)));
                        v = int16_t(__double2int_rn(v * gain));
                    }
                    framePeak = max(framePeak, frame(f, c, v));
                }
                audible(f, framePeak);
            }
        }
    }
    flush();
}

// Fade, pack, peak, audible extent and clip count per 1024-frame block
// (src/cuda/tail.cu:240-297). Each wave of a block takes kFinalizeBlocks
// consecutive tail blocks (`blocks` in all from `firstBlock`), so the owner
// search, the job and result loads and the atomics are paid once per run of
// blocks of one job; the reductions are wave shuffles, with no LDS round or
// block barrier. `channels` is 1 (natural) or 2. In place (orig/custom) the
// output offset of a job equals its mix offset, so every frame is read and
// written at one address.
extern "C" __global__ void __launch_bounds__(64 * kFinalizeWaves) clayFinalizePcm(
    const Job* __restrict__ jobs, int count, const int16_t* __restrict__ input, int16_t* __restrict__ output,
    clay_tail_result* __restrict__ results, int channels, int countClipped, int firstBlock, int blocks)
{
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    const int begin = firstBlock + (int(blockIdx.x) * kFinalizeWaves + wave) * kFinalizeBlocks;
    const int stopBlock = min(begin + kFinalizeBlocks, firstBlock + blocks);
    if (begin >= stopBlock)
        return;  // uniform per wave; no block barriers
    if (channels == 1)
        finalizeRun<1>(jobs, count, input, output, results, countClipped, begin, stopBlock, lane);
    else
        finalizeRun<2>(jobs, count, input, output, results, countClipped, begin, stopBlock, lane);
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
