// CPU checks of the plain form's memory policy
// (claysampler_amd/csrc/hip/memory_policy.h):
//  - scratch bytes by window against a frame-by-frame count of the bytes each
//    window of each span addresses (the largest over the span's windows);
//  - windowFor against a brute-force scan of every window length;
//  - the budget formula at its corners;
//  - the allocation-failure chain run against an allocator with a hard
//    limit: the scratch halves while above 2 MiB, then the batch halves, then
//    one job moves its scratch to the host, and only then is it rejected.
// Prints "failures N"; usage: memory_check SEED CASES
#include "../claysampler_amd/csrc/hip/memory_policy.h"

#include <cstdio>
#include <cstdlib>
#include <random>

using namespace clay_memory;

static long failures = 0;
#define EXPECT(c) do { if (!(c)) { ++failures; std::printf("fail line %d\n", __LINE__); } } while (0)

// Frames a span addresses in windows of `w` frames: its longest overlap with one window.
static int64_t longestOverlap(int begin, int end, int w, int last)
{
    if (w <= 0)
        return end - begin;
    int64_t best = 0;
    for (int first = 0; first < last; first += w)
        best = std::max<int64_t>(best, std::min(end, first + w) - std::max(begin, first));
    return best;
}

int main(int argc, char** argv)
{
    if (argc != 3)
        return 2;
    std::mt19937_64 rng(std::strtoull(argv[1], nullptr, 10));
    const long cases = std::strtol(argv[2], nullptr, 10);
    for (long c = 0; c < cases; ++c) {
        const int last = 256 * int(1 + rng() % 400);
        std::vector<ScratchUse> uses(1 + rng() % 40);
        for (auto& u : uses) {
            u.begin = int(rng() % last);
            u.end = u.begin + 1 + int(rng() % (last - u.begin));
            u.channels = 1 + int(rng() % 2);
            u.pitch = rng() % 3 == 0;
            u.backed = rng() % 4 == 0 ? int(rng() % (u.end - u.begin + 1)) : INT32_MAX;
        }
        // Scratch by window: min(length, backed, w) floats per channel, min(length, w) phases.
        for (int w : {0, 256, 512, 4096, last}) {
            uint64_t want = 0;
            for (const auto& u : uses) {
                const int64_t addressed = longestOverlap(u.begin, u.end, w, last);
                const int64_t cap = w > 0 ? w : u.end - u.begin;
                want += uint64_t(std::min<int64_t>(std::min<int64_t>(u.end - u.begin, u.backed), cap)) * u.channels * 4;
                want += u.pitch ? uint64_t(std::min<int64_t>(u.end - u.begin, cap)) * 8 : 0;
                EXPECT(addressed <= std::min<int64_t>(u.end - u.begin, cap));  // the scratch covers every window
            }
            EXPECT(scratchBytes(uses, w) == want);
        }
        // windowFor: 0 when the whole batch fits, else the longest fitting multiple of 256, else 256.
        const uint64_t full = scratchBytes(uses, 0);
        const uint64_t budget = rng() % (full + 2);
        int best = -1;
        for (int g = 1; g * 256 <= last + 255; ++g)
            if (scratchBytes(uses, g * 256) <= budget)
                best = g * 256;
        const int got = windowFor(uses, budget, last);
        if (full <= budget)
            EXPECT(got == 0);
        else
            EXPECT(got == (best < 0 ? 256 : best));
        for (int g = 1; g < 64; ++g)
            EXPECT(scratchBytes(uses, g * 256) <= scratchBytes(uses, (g + 1) * 256));
    }
    // Budget: a quarter of the device, three quarters of free memory, what the
    // fixed buffers leave, at least 1 MiB, and under the limit.
    const uint64_t MiB = 1 << 20, GiB = uint64_t(1) << 30;
    EXPECT(scratchBudget(100 * GiB, 288 * GiB, 0, UINT64_MAX) == 72 * GiB);
    EXPECT(scratchBudget(40 * GiB, 288 * GiB, 0, UINT64_MAX) == 30 * GiB);
    EXPECT(scratchBudget(40 * GiB, 288 * GiB, 34 * GiB, UINT64_MAX) == 1 * GiB);
    EXPECT(scratchBudget(40 * GiB, 288 * GiB, 39 * GiB, UINT64_MAX) == MiB);
    EXPECT(scratchBudget(40 * GiB, 288 * GiB, 0, 5 * MiB) == 5 * MiB);
    // The chain against an allocator of `limit` bytes: a batch of `jobs` jobs
    // needs jobs x fixed + scratch bytes, scratch halving with the window.
    for (long c = 0; c < cases; ++c) {
        size_t jobs = 1 + rng() % 16;
        const uint64_t fixed = MiB * (1 + rng() % 64), perJobScratch = MiB * (rng() % 512);
        const uint64_t limit = MiB * (1 + rng() % 4096);
        uint64_t scratch = jobs * perJobScratch;
        bool host = false, placed = false, rejected = false;
        int steps = 0;
        while (!placed && !rejected && ++steps < 1000) {
            const uint64_t device = jobs * fixed + (host ? 0 : scratch);
            if (device <= limit) {
                placed = true;
                break;
            }
            const Next next = afterOutOfMemory(host ? 0 : scratch, jobs, host);
            if (next == Next::halveScratch) {
                EXPECT(scratch > 2 * MiB && !host);
                scratch /= 2;
            } else if (next == Next::halveJobs) {
                EXPECT((host ? 0 : scratch) <= 2 * MiB && jobs > 1);
                jobs /= 2;
                scratch = std::min(scratch, jobs * perJobScratch);
            } else if (next == Next::hostScratch) {
                EXPECT(jobs == 1 && !host && scratch > 0);
                host = true;
            } else {
                EXPECT(jobs == 1);
                rejected = true;
            }
        }
        EXPECT(steps < 1000);
        // A single job is rejected only when its non-scratch buffers alone exceed the limit.
        if (rejected)
            EXPECT(fixed > limit);
        if (fixed <= limit)
            EXPECT(placed);
    }
    std::printf("failures %ld\n", failures);
    return 0;
}
