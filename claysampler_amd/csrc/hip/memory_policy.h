// Device memory policy of a plain batch: how much filtered-voice scratch it
// needs when rendered in windows of a given length, the longest window that
// fits a scratch budget, and what to try after an allocation fails (the role
// of src/cuda/renderer.cpp:864-989).
//
// A filtered span's scratch is addressed from its first frame in the current
// window, so a window of W frames needs min(length, W) frames of it (and of
// its pitch phases). Host-only, so the rules are checked on the CPU
// (tests/memory_check.cpp).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace clay_memory {

// One filtered span: its frames, channels, whether it carries pitch phases,
// and the frames its predicted-tail bound backs (INT32_MAX: all).
struct ScratchUse {
    int begin, end, channels;
    bool pitch;
    int backed = INT32_MAX;
};

constexpr int kGranule = 256;  // windows are whole tiles

// Filtered (float) and phase (double) scratch bytes in windows of `window`
// frames (0: the whole batch in one window).
inline uint64_t scratchBytes(const std::vector<ScratchUse>& uses, int window)
{
    uint64_t floats = 0, doubles = 0;
    for (const auto& u : uses) {
        int64_t frames = std::min<int64_t>(u.end - u.begin, u.backed);
        int64_t phases = u.end - u.begin;
        if (window > 0) {
            frames = std::min<int64_t>(frames, window);
            phases = std::min<int64_t>(phases, window);
        }
        floats += uint64_t(std::max<int64_t>(0, frames)) * uint64_t(u.channels);
        if (u.pitch)
            doubles += uint64_t(phases);
    }
    return floats * 4 + doubles * 8;
}

// The longest window (a multiple of kGranule, or 0 for the whole batch of
// `lastFrame` frames) whose scratch fits `budget`; kGranule when none does.
inline int windowFor(const std::vector<ScratchUse>& uses, uint64_t budget, int lastFrame)
{
    if (scratchBytes(uses, 0) <= budget)
        return 0;
    int low = 1, high = std::max(1, (lastFrame + kGranule - 1) / kGranule);  // in granules
    if (scratchBytes(uses, kGranule) > budget)
        return kGranule;
    while (low < high) {  // scratchBytes grows with the window: the last granule count that fits
        const int mid = low + (high - low + 1) / 2;
        if (scratchBytes(uses, mid * kGranule) <= budget)
            low = mid;
        else
            high = mid - 1;
    }
    return low * kGranule;
}

// Scratch budget of a batch (renderer.cpp:801-811): at most a quarter of the
// device and three quarters of what is free, at least 1 MiB, and no more than
// what is left once the batch's other buffers (`fixed`) are placed.
inline uint64_t scratchBudget(uint64_t freeBytes, uint64_t totalBytes, uint64_t fixed, uint64_t limit)
{
    const uint64_t usable = freeBytes - freeBytes / 8;
    const uint64_t room = usable > fixed ? std::min(std::min(freeBytes - freeBytes / 4, totalBytes / 4), usable - fixed) : 0;
    return std::min(limit, std::max<uint64_t>(uint64_t(1) << 20, room));
}

// After a failed arena allocation: halve the scratch while it is more than
// 2 MiB, then halve the batch, then (one job) keep the scratch in mapped host
// memory, then reject the job (renderer.cpp:948-989).
enum class Next { halveScratch, halveJobs, hostScratch, reject };

inline Next afterOutOfMemory(uint64_t scratch, size_t jobs, bool hostScratchTried)
{
    if (scratch > (uint64_t(2) << 20))
        return Next::halveScratch;
    if (jobs > 1)
        return Next::halveJobs;
    if (!hostScratchTried && scratch > 0)
        return Next::hostScratch;
    return Next::reject;
}

} // namespace clay_memory
