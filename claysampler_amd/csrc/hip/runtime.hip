// Host runtime of the render core: the C ABI of include/clay_hip.h.
//
// Replaces the reference renderer (src/cuda/renderer.cpp). A window of plans
// is split into batches; each batch is flattened into device records
// (the role of BatchIndex::append, renderer.cpp:378-462), uploaded once into
// one device arena, rendered by the kernels of render.hip on one HIP stream,
// finished on the device by the kernels of finish.hip (tail, quality, group
// gain, packing) and downloaded as packed PCM16. MI355X's 288 GB of HBM lets a
// batch keep its whole filter scratch resident, so a plain batch normally
// renders as one segment.
//
// Two forms (bit-identical PCM):
//  - plain: batches of up to batch_jobs jobs, one round of render launches per
//    segment, then the batch's finishing launches (renderer.cpp:787-1142);
//  - persistent: up to batch_jobs resident units (a job, or a whole volume
//    group) advance by segment_frames per round; a round is one replay of a
//    captured hipGraph whose kernels walk the round's unit descriptors with
//    fixed grids; finished units are finished, downloaded and published while
//    the others keep rendering, and queued jobs are admitted in their place
//    (the role of renderRolling, renderer.cpp:1351-1746).
#include <hip/hip_runtime.h>
#include <sched.h>
#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>

#include "device_types.h"
#include "memory_policy.h"
#include "staging.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

using namespace clay_dev;

// Render kernels (render.hip). Plain form: one unit, pointers as arguments.
extern "C" __global__ void clayPitchChains(const Span*, const SpanCold*, const int*, const clay_control_point*,
                                           const Mod*, double*, VoiceState*, int, int, int);
template<int T, int U>
__global__ void clayChainAudio(const int16_t*, const float*, const Span*, const Job*, const int*,
                               const clay_control_point*, const Mod*, const double*, float*, const VoiceState*, int,
                               int, int);
template<int W>
__global__ void clayVoiceChains(const Span*, const SpanCold*, const int*, const clay_control_point*, const Mod*, float*,
                                VoiceState*, int, int, int, float);
template<int T, bool FloatOut, bool WetOnly>
__global__ void clayMixTiles(const int16_t*, const float*, const Span*, const Job*, const Tile*, const int*,
                             const clay_control_point*, const Mod*, const float*, const VoiceState*, void*, int, int,
                             int);
// Persistent form: the work of every resident unit listed in the round descriptors.
extern "C" __global__ void clayRoundPitchChains(const UnitView*, const RoundHeader*);
extern "C" __global__ void clayClearQueues(MixQueue*, int);
template<int T>
__global__ void clayRoundChainAudio(const int16_t*, const float*, const UnitView*, const RoundHeader*);
__global__ void clayRoundVoiceChains(const UnitView*, const RoundHeader*);
// Mix (mix.h): one wave per tile, integer FIR for plain voices; plain and persistent forms.
// The bank and, for the straight voices' aligned window loads, the byte
// offset of its shifted copy minus 2 (clay_hip::bankOdd).
template<int T, bool FloatOut, bool Simple>
__global__ void clayMixWaves(const int16_t*, uint32_t, const uint32_t*, const float*, const Span*, const Job*,
                             const Tile*, const int*, const clay_control_point*, const Mod*, const float*,
                             const VoiceState*, void*, int, int, int, int, MixQueue*);
template<int T, bool FloatOut, bool Simple>
__global__ void clayRoundMixWaves(const int16_t*, uint32_t, const uint32_t*, const float*, const UnitView*,
                                  const RoundHeader*, MixQueue*);
// Mix of straight batches on kStraightFrames tiles (mix.h).
template<int T, bool FloatOut>
__global__ void clayMixStraight(const int16_t*, uint32_t, const uint32_t*, const Span*, const Job*, const Tile*,
                                const int*, const clay_control_point*, void*, int, int, int, int, MixQueue*,
                                StereoStats*, int16_t*);
template<int T, bool FloatOut>
__global__ void clayRoundMixStraight(const int16_t*, uint32_t, const uint32_t*, const UnitView*, const RoundHeader*,
                                     MixQueue*);
// Finishing kernels (finish.hip).
extern "C" __global__ void clayScanStereo(const Job*, int, const float2*, StereoStats*, int);
extern "C" __global__ void claySelectNative(const Job*, int, const float2*, const StereoStats*, clay_tail_result*);
extern "C" __global__ void clayQuantizeMono(const Job*, int, const float2*, const clay_tail_result*, int16_t*,
                                            MonoStats*, int, int, int, int);
extern "C" __global__ void claySelectTrim(const Job*, int, const MonoStats*, clay_tail_result*, TailConfig);
extern "C" __global__ void clayLayoutPcm(const Job*, int, clay_tail_result*, int);
extern "C" __global__ void clayFinalizePcm(const Job*, int, const int16_t*, int16_t*, clay_tail_result*, int, int, int,
                                            int);
extern "C" __global__ void clayMeasureBody(const AudioJobDesc*, int, const int16_t*, const clay_tail_result*,
                                           unsigned long long*, clay_audio_config, int, int);
extern "C" __global__ void clayMeasureOnsets(const AudioJobDesc*, int, const int*, const int16_t*,
                                             const clay_tail_result*, unsigned long long*, int, int, int);
extern "C" __global__ void claySummarizeAudio(const AudioJobDesc*, int, const unsigned long long*,
                                              unsigned long long*, clay_tail_result*, clay_audio_config, int);
extern "C" __global__ void clayMatchVolume(const int2*, int, int, clay_tail_result*);
extern "C" __global__ void clayLayoutAccepted(const Job*, int, clay_tail_result*, PcmJob*);
extern "C" __global__ void clayPackAccepted(const PcmJob*, int, const int16_t*, int16_t*, clay_tail_result*, int, int);

namespace {

struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char* what)
{
    if (e == hipErrorOutOfMemory)
        throw Failure(CLAY_E_MEMORY, std::string(what) + ": out of memory");
    if (e != hipSuccess)
        throw Failure(CLAY_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

thread_local std::string g_createError;

// Device-to-host copies on the SDMA engines that ROCr names as the fastest for
// that direction (hsa_amd_memory_get_preferred_copy_engine: two of MI355X's 16
// engines). Measured into the renderer's pinned PCM slots while the mix
// kernel runs, each of the two moves about 28 GB/s and engines 4-15 at most
// 13 GB/s; HIP's asynchronous device-to-host copies on a second stream ran
// at 13-29 GB/s. Copies go through kLanes lanes, one thread each, every lane
// with its own share of the engines and its own completion signal, so that a
// copy that stalls holds up one lane only. A copy has no dependencies: the
// caller waits for the kernels that wrote its source first.
struct Downloader {
    static constexpr int kLanes = 2;
    bool open = false, usable = false, initialised = false;
    hsa_agent_t gpu{}, cpu{};
    std::vector<hsa_amd_sdma_engine_id_t> engines;
    hsa_signal_t signals[kLanes]{};
    int signalCount = 0;

    static hsa_status_t collect(hsa_agent_t a, void* data)
    {
        static_cast<std::vector<hsa_agent_t>*>(data)->push_back(a);
        return HSA_STATUS_SUCCESS;
    }

    // Finds the device's agent by PCI location; leaves usable false (plain
    // synchronous HIP copies) when any step fails.
    void start(const hipDeviceProp_t& p)
    {
        if (open)
            return;
        open = true;
        if (hsa_init() != HSA_STATUS_SUCCESS)
            return;
        initialised = true;
        std::vector<hsa_agent_t> agents;
        if (hsa_iterate_agents(collect, &agents) != HSA_STATUS_SUCCESS)
            return;
        const uint32_t bdf = (uint32_t(p.pciBusID) << 8) | (uint32_t(p.pciDeviceID) << 3);
        bool found = false;
        for (hsa_agent_t a : agents) {
            hsa_device_type_t type;
            uint32_t id = 0, domain = 0;
            if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &type) != HSA_STATUS_SUCCESS || type != HSA_DEVICE_TYPE_GPU)
                continue;
            if (hsa_agent_get_info(a, hsa_agent_info_t(HSA_AMD_AGENT_INFO_BDFID), &id) != HSA_STATUS_SUCCESS ||
                hsa_agent_get_info(a, hsa_agent_info_t(HSA_AMD_AGENT_INFO_DOMAIN), &domain) != HSA_STATUS_SUCCESS)
                continue;
            if ((id & ~7u) == bdf && domain == uint32_t(p.pciDomainID)) {
                gpu = a;
                found = true;
                break;
            }
        }
        if (!found || hsa_agent_get_info(gpu, hsa_agent_info_t(HSA_AMD_AGENT_INFO_NEAREST_CPU), &cpu) != HSA_STATUS_SUCCESS)
            return;
        uint32_t preferred = 0;
        if (hsa_amd_memory_get_preferred_copy_engine(cpu, gpu, &preferred) != HSA_STATUS_SUCCESS || !preferred)
            return;
        for (uint32_t bit = 1; bit; bit <<= 1)
            if (preferred & bit)
                engines.push_back(hsa_amd_sdma_engine_id_t(bit));
        for (; signalCount < kLanes; ++signalCount)
            if (hsa_signal_create(1, 0, nullptr, &signals[signalCount]) != HSA_STATUS_SUCCESS)
                return;
        usable = true;
    }

    // Called by one thread per lane at a time.
    void copy(void* host, const void* device, size_t bytes, const char* what, int lane)
    {
        if (!bytes)
            return;
        if (!usable) {
            check(hipMemcpy(host, device, bytes, hipMemcpyDeviceToHost), what);
            return;
        }
        // The lane's engines: every kLanes-th preferred engine from `lane`, or
        // all of them when there are fewer engines than lanes.
        const bool shared = engines.size() < size_t(kLanes);
        const size_t mine = shared ? engines.size() : (engines.size() - size_t(lane) + kLanes - 1) / kLanes;
        const auto engineOf = [&](size_t k) { return shared ? engines[k] : engines[size_t(lane) + k * kLanes]; };
        hsa_signal_t signal = signals[lane];
        // One part per engine, 4 KiB aligned; copies under 1 MiB take one engine.
        const size_t parts = bytes < (size_t(1) << 20) ? 1 : mine;
        const size_t part = ((bytes + parts - 1) / parts + 4095) & ~size_t(4095);
        size_t issued = 0;
        hsa_signal_store_screlease(signal, hsa_signal_value_t(parts));
        for (size_t k = 0; k < parts; ++k) {
            const size_t at = std::min(bytes, k * part), n = std::min(bytes, at + part) - at;
            hsa_status_t st = HSA_STATUS_SUCCESS;
            if (n)
                st = hsa_amd_memory_async_copy_on_engine(static_cast<char*>(host) + at, cpu,
                                                         static_cast<const char*>(device) + at, gpu, n, 0, nullptr,
                                                         signal, engineOf(k), false);
            else
                hsa_signal_subtract_screlease(signal, 1);
            if (st != HSA_STATUS_SUCCESS) {
                // Wait for the parts already issued before reporting.
                hsa_signal_subtract_screlease(signal, hsa_signal_value_t(parts - issued));
                (void)hsa_signal_wait_scacquire(signal, HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX, HSA_WAIT_STATE_BLOCKED);
                throw Failure(CLAY_E_DEVICE, std::string(what) + ": SDMA copy not accepted (status " +
                                                 std::to_string(int(st)) + ")");
            }
            ++issued;
        }
        const hsa_signal_value_t v =
            hsa_signal_wait_scacquire(signal, HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX, HSA_WAIT_STATE_BLOCKED);
        if (v < 0)
            throw Failure(CLAY_E_DEVICE, std::string(what) + ": SDMA copy failed");
    }

    void stop()
    {
        for (int k = 0; k < signalCount; ++k)
            hsa_signal_destroy(signals[k]);
        signalCount = 0;
        if (initialised)
            hsa_shut_down();
        open = usable = initialised = false;
    }
};

// Device blocks for the persistent form's cohort arenas, allocated with
// hipMalloc in 64 MiB steps and kept for the next cohorts. Arenas from the
// stream-ordered pool (hipMallocAsync) slowed the SDMA downloads that run
// beside the mix kernel to 4-13 GB/s, against 40-44 GB/s over hipMalloc
// blocks, and hipFree would stall the round loop (it waits for the device).
// A block is given back only when no queued work uses it any more, or when
// all later users are ordered after that work on the render stream.
struct ArenaCache {
    struct Block {
        void* base = nullptr;
        size_t bytes = 0;
        bool busy = false;
    };
    std::vector<Block> blocks;

    void* take(size_t bytes)
    {
        Block* best = nullptr;
        for (auto& b : blocks)
            if (!b.busy && b.bytes >= bytes && (!best || b.bytes < best->bytes))
                best = &b;
        if (best) {
            best->busy = true;
            return best->base;
        }
        const size_t step = size_t(64) << 20;
        Block b;
        b.bytes = (bytes + step - 1) / step * step;
        hipError_t e = hipMalloc(&b.base, b.bytes);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            trim(0);  // idle blocks back to the device, then once more
            e = hipMalloc(&b.base, b.bytes);
        }
        check(e, "cohort arena");
        b.busy = true;
        blocks.push_back(b);
        return b.base;
    }

    void give(void* base)
    {
        for (auto& b : blocks)
            if (b.base == base) {
                b.busy = false;
                return;
            }
    }

    // Frees idle blocks until at most `keep` bytes of idle blocks remain
    // (hipFree waits for the device).
    void trim(size_t keep)
    {
        size_t idle = 0;
        for (auto& b : blocks)
            idle += b.busy ? 0 : b.bytes;
        for (size_t i = blocks.size(); i-- > 0 && idle > keep;)
            if (!blocks[i].busy) {
                (void)hipFree(blocks[i].base);
                idle -= blocks[i].bytes;
                blocks.erase(blocks.begin() + std::ptrdiff_t(i));
            }
    }
};

template<class T, class A>
size_t bytesOf(const std::vector<T, A>& v) { return v.size() * sizeof(T); }

// An allocator whose elements start uninitialised: the large record arrays of
// a batch are sized first and then written element by element (on several
// threads), so value-initialising them would only add a pass over hundreds of
// megabytes.
template<class T>
struct Uninitialised : std::allocator<T> {
    using value_type = T;
    template<class U>
    struct rebind {
        using other = Uninitialised<U>;
    };
    Uninitialised() = default;
    template<class U>
    Uninitialised(const Uninitialised<U>&) noexcept {}
    template<class U>
    void construct(U* p) noexcept
    {
        ::new (static_cast<void*>(p)) U;
    }
    template<class U, class... A>
    void construct(U* p, A&&... a)
    {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
template<class T>
using Records = std::vector<T, Uninitialised<T>>;

// The large record arrays of a batch, kept by the renderer once the batch is
// gone and handed to the next batch it flattens: their pages are mapped
// already, so a batch writes its records without first touching hundreds of
// megabytes of fresh memory. Shared by the renderer and its batches, so a
// batch may outlive the renderer.
struct RecordSet {
    Records<Span> spans;
    Records<SpanCold> cold;
    Records<Tile> tiles;
    Records<int> voices;
    size_t bytes() const
    {
        return spans.capacity() * sizeof(Span) + cold.capacity() * sizeof(SpanCold) +
               tiles.capacity() * sizeof(Tile) + voices.capacity() * sizeof(int);
    }
};

class RecordPool {
public:
    static constexpr size_t kMaxBytes = size_t(2) << 30;  // host memory kept at most
    RecordSet take()
    {
        std::lock_guard<std::mutex> lock(mutex_);
        RecordSet out;
        if (!spare_.empty()) {
            out = std::move(spare_.back());
            spare_.pop_back();
            held_ -= out.bytes();
        }
        return out;
    }
    void give(RecordSet&& set)
    {
        const size_t bytes = set.bytes();
        std::lock_guard<std::mutex> lock(mutex_);
        if (bytes && held_ + bytes <= kMaxBytes) {
            held_ += bytes;
            spare_.push_back(std::move(set));
        }
    }

private:
    std::mutex mutex_;
    std::vector<RecordSet> spare_;
    size_t held_ = 0;
};

uint64_t ceilDiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Staging policy of the render core (staging.h): a hipEvent behind each copy
// on the renderer's stream, waited with hipEventSynchronize; pinned buffers
// from hipHostMalloc. A failed wait leaves the error to the next HIP call.
struct HipFences {
    const hipStream_t* stream = nullptr;
    using Fence = hipEvent_t;
    Fence record()
    {
        hipEvent_t e = nullptr;
        check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        check(hipEventRecord(e, *stream), "hipEventRecord");
        return e;
    }
    void wait(Fence& e)
    {
        if (!e)
            return;
        (void)hipEventSynchronize(e);
        (void)hipEventDestroy(e);
        e = nullptr;
    }
    bool done(Fence& e)
    {
        if (e && hipEventQuery(e) == hipErrorNotReady)
            return false;
        if (e)
            (void)hipEventDestroy(e);
        e = nullptr;
        return true;
    }
    char* allocate(size_t bytes)
    {
        void* p = nullptr;
        check(hipHostMalloc(&p, bytes, hipHostMallocDefault), "pinned staging");
        return static_cast<char*>(p);
    }
    void free(char* p) { (void)hipHostFree(p); }
};

// std::vector of trivially copyable records -> device, in one arena.
struct ArenaBuilder {
    size_t bytes = 0;
    size_t add(size_t n)
    {
        const size_t at = bytes;
        bytes += (n + 255) & ~size_t(255);
        return at;
    }
};

// Sinc table, tap-major as src/cuda/renderer.cpp:347-376 generalised to `taps`.
std::vector<float> sincTable(int taps)
{
    const double beta = 8.5, pi = 3.14159265358979323846;
    const auto bessel = [](double v) {
        double sum = 1, term = 1;
        for (int i = 1; i < 30; ++i) {
            term *= v * v / (4 * i * i);
            sum += term;
            if (term < 1e-14)
                break;
        }
        return sum;
    };
    const double norm = bessel(beta);
    const int centre = taps / 2 - 1;
    const double half = taps / 2;
    std::vector<float> t(size_t(taps) * (kPhases + 1));
    for (int tap = 0; tap < taps; ++tap)
        for (int p = 0; p <= kPhases; ++p) {
            const double x = tap - centre - double(p) / kPhases;
            const double f = x / half;
            const double window = bessel(beta * std::sqrt(1 - f * f)) / norm;
            t[size_t(tap) * (kPhases + 1) + size_t(p)] = float(window * (x == 0 ? 1 : std::sin(pi * x) / (pi * x)));
        }
    return t;
}

// The phase-major table as the mix's integer weights: w = hi 2^-14 + lo 2^-24
// with hi = rint(w 2^14) and lo = rint((w - hi 2^-14) 2^24), exact to 2^-25.
// Row p holds T dwords: the pairs (hi[2j], hi[2j+1]) for j < T/2 (low half
// first), then the pairs of lo. Mirror-symmetric like the float table, since
// hi and lo are functions of w alone.
std::vector<uint32_t> intSincTable(const std::vector<float>& phaseMajor, int taps)
{
    std::vector<uint32_t> out(phaseMajor.size());
    for (int p = 0; p <= kPhases; ++p)
        for (int j = 0; j < taps / 2; ++j) {
            uint32_t hi = 0, lo = 0;
            for (int k = 0; k < 2; ++k) {
                const double w = phaseMajor[size_t(p) * size_t(taps) + size_t(2 * j + k)];
                const long h = std::lrint(w * 16384.0);
                const long l = std::lrint((w - double(h) / 16384.0) * 16777216.0);
                if (h < -32768 || h > 32767 || l < -32768 || l > 32767)
                    throw Failure(CLAY_E_DEVICE, "sinc weight outside the integer table's range");
                hi |= uint32_t(uint16_t(int16_t(h))) << (16 * k);
                lo |= uint32_t(uint16_t(int16_t(l))) << (16 * k);
            }
            out[size_t(p) * size_t(taps) + size_t(j)] = hi;
            out[size_t(p) * size_t(taps) + size_t(taps / 2 + j)] = lo;
        }
    return out;
}

struct Timer {
    std::string name;
    hipEvent_t begin, end;
};

} // namespace

struct clay_hip_batch {
    // The renderer's pool of record arrays: flatten takes from it, and the
    // arrays go back to it when the batch is destroyed.
    std::shared_ptr<RecordPool> pool;
    ~clay_hip_batch()
    {
        if (pool)
            pool->give({std::move(spans), std::move(cold), std::move(tiles), std::move(voices)});
    }
    // Host records.
    Records<Span> spans;
    Records<SpanCold> cold;  // empty for a batch without voice chains (nothing reads it)
    std::vector<Mod> mods;
    std::vector<clay_control_point> points;
    std::vector<Job> jobs;
    Records<Tile> tiles;
    Records<int> voices;
    std::vector<int> pitchHeads, chainHeads;  // variable-pitch chains; every voice chain (longest first)
    std::vector<int> chainByStart;            // voice chains by first frame (windowed rounds)
    std::vector<int2> chainExtent;            // [first, end) frames of chainByStart[i]
    std::vector<AudioJobDesc> audioJobs;
    std::vector<int> onsets;
    std::vector<int2> groups;
    uint64_t frames = 0, filteredSamples = 0, pitchFrames = 0;
    int blocks = 0, windows = 0, bodyBlocks = 0, chains = 0;
    // Optional per-chain scratch bounds (absolute frames, batch chain order),
    // read by flatten; empty = every filtered span backed in full.
    std::vector<int32_t> scratchEnd;
    // Plain form under memory pressure: frames per render window (0 = the
    // whole batch at once), sizing the filtered and phase scratch; and that
    // scratch in mapped host memory (a single job that does not fit otherwise).
    int window = 0;
    bool hostScratch = false;
    void* hostScratchBuffer = nullptr;
    bool assess = false;
    bool allWet = false;  // every span renders through a voice chain (no dry voice in the mix)
    bool simpleMix = false;  // every dry voice is plain with no loop and no bend (the mix's Simple variant)
    // Every voice is straight (dry, plain, no loop, no bend, position and step
    // >= 0): `tiles` are then kStraightFrames long and mixed by clayMixStraight.
    bool straight = false;
    // Device arena and views into it.
    void* arena = nullptr;
    size_t arenaBytes = 0;
    Span* dSpans = nullptr;
    SpanCold* dCold = nullptr;
    Mod* dMods = nullptr;
    clay_control_point* dPoints = nullptr;
    Job* dJobs = nullptr;
    Tile* dTiles = nullptr;
    int* dVoices = nullptr;
    int *dPitchHeads = nullptr, *dChainHeads = nullptr, *dChainByStart = nullptr;
    int lastFrame = 0;
    bool cachedArena = false;  // arena from the renderer's ArenaCache (persistent form)
    VoiceState* dStates = nullptr;
    AudioJobDesc* dAudioJobs = nullptr;
    int* dOnsets = nullptr;
    int2* dGroups = nullptr;
    void* dMix = nullptr;
    float* dFiltered = nullptr;
    double* dPhase = nullptr;
    int16_t* dMono = nullptr;
    StereoStats* dStereo = nullptr;  // null unless the natural policy is on
    bool stereoDone = false;         // the last mix stored the tail-block statistics
    MonoStats* dWindows = nullptr;
    unsigned long long *dBody = nullptr, *dOnsetPower = nullptr;
    PcmJob* dPcmJobs = nullptr;
    int16_t* dOutput = nullptr;   // packed PCM of the batch
    clay_tail_result* dResults = nullptr;
    // Per-kernel timing of the last run.
    std::map<std::string, std::pair<double, int>> kernelTimes;
};

// Launch shapes of the plain form's audio and filter passes, chosen by
// measurement: the filter pass's voice chains (voices) per workgroup, 1, 2 or
// 4 waves, and the audio pass's frames per lane in flight, its sinc-FIR
// unroll, 1 or 2. Each candidate is timed on kTrials launches (HIP events on
// the render stream, read once they have completed) in ms per wet sample, and
// the cheapest is kept for the renderer's later launches. Every candidate
// computes the same bits, so tuning never changes the audio.
struct LaunchTuner {
    static constexpr int kTrials = 2;
    struct Dimension {
        std::vector<int> values;
        int chosen = 0;  // value used once tuning is over (or off)
        bool tuning = false;
        std::vector<double> cost;
        std::vector<int> trials, issued;

        void init(std::vector<int> v, int fallback, bool tune)
        {
            values = std::move(v);
            chosen = fallback;
            tuning = tune;
            cost.assign(values.size(), 0);
            trials.assign(values.size(), 0);
            issued.assign(values.size(), 0);
        }
        // Candidate to measure on the next launch, or -1 (use `chosen`).
        int next() const
        {
            if (tuning)
                for (size_t i = 0; i < values.size(); ++i)
                    if (issued[i] < kTrials)
                        return int(i);
            return -1;
        }
        int value(int candidate) const { return candidate < 0 ? chosen : values[size_t(candidate)]; }
        void record(int candidate, double ms, double work)
        {
            if (!tuning || candidate < 0 || work <= 0)
                return;
            cost[size_t(candidate)] += ms / work;
            ++trials[size_t(candidate)];
            for (int t : trials)
                if (t < kTrials)
                    return;
            size_t best = 0;
            for (size_t i = 1; i < cost.size(); ++i)
                if (cost[i] < cost[best])
                    best = i;
            chosen = values[best];
            tuning = false;
        }
    };
    Dimension chainWaves, audioUnroll;
    struct Pending {
        Dimension* dimension;
        int candidate;
        hipEvent_t begin, end;
        double work;
    };
    std::vector<Pending> pending;
};

struct clay_hip {
    clay_hip_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    LaunchTuner tuner;      // plain form's audio and filter pass launch shapes
    Downloader downloader;  // persistent form's download thread
    ArenaCache arenas;      // persistent form's cohort arenas
    float* sinc = nullptr;       // phase-major taps x (phases+1)
    uint32_t* sincInt = nullptr; // the same weights as packed int16 hi/lo pairs (intSincTable)
    // The sample bank, and one allocation-stride after it the same samples one
    // place later (copy[i] = bank[i + 1]): a window that starts on an odd
    // sample is a dword-aligned run of the copy. bankOdd = stride - 2, the byte
    // offset from bank sample j to copy sample j - 1 less the 2 bytes the
    // window offset 2 j already counts.
    int16_t* bank = nullptr;
    uint64_t bankCount = 0, bankCapacity = 0;
    uint32_t bankOdd = 0;
    // The mix launches' tile queues (0: clayMixWaves, 1: clayMixStraight), zero
    // between launches.
    MixQueue* queues = nullptr;
    std::shared_ptr<RecordPool> records = std::make_shared<RecordPool>();
    uint64_t bankShifted = 0;  // samples of the bank the shifted copy reflects
    std::string error;
    clay_hip_stats stats{};
    bool natural = true;
    float silenceLevel = 0;  // voice-silence policy threshold (linear), 0 = off
    TailConfig tail{};
    bool timing = false;
    std::vector<Timer> timers;
    std::vector<hipEvent_t> eventPool;
    // Pinned PCM slots handed to publish callbacks.
    struct Slot {
        void* host = nullptr;
        size_t capacity = 0;
        bool busy = false;
    };
    std::vector<Slot> slots;
    std::mutex slotLock;  // slots are taken by download threads and released by callers' threads
    size_t residentBytes = 0;
    int tunedSegment = 0;  // persistent form: segment frames chosen by the round tuner, 0 = not yet
    uint64_t memoryLimit = 0;  // bytes a batch arena may take (clay_hip_set_memory_limit), 0 = no limit
    // Round descriptors: a header and up to unitCapacity unit views on the
    // device, written through a ring of pinned staging buffers.
    static constexpr int kStages = 16;
    struct Stage {
        void* host = nullptr;
        hipEvent_t copied = nullptr;
        bool used = false;
    };
    Stage stages[kStages];
    int stageNext = 0;
    char* dRound = nullptr;
    size_t unitCapacity = 0;
    // Captured persistent rounds by launch shape.
    struct Graph {
        hipGraphExec_t exec;
        int mixBlocks, chainBlocks;
        bool simpleMix;
        const void* bank;  // captured launch arguments: the bank and its copy's offset
        uint32_t bankOdd;
    };
    std::vector<Graph> graphs;
    // Pinned image of a plain batch's records (grown on demand): the copy runs
    // at DMA speed instead of through a pageable bounce; rewritten or freed only
    // after the copy that last read it (a caller may prepare a batch while
    // another's upload is still in flight).
    clay_staging::Image<HipFences> staging{HipFences{&stream}};
    // Pinned images of persistent-form records, packed by the threads that
    // flatten cohorts and reused once the copies out of them have completed.
    clay_staging::Pool<HipFences> images{HipFences{&stream}};
    // Small pinned buffers (per-unit results) kept for reuse.
    std::vector<std::pair<void*, size_t>> pinnedSpare;

    void* pinned(size_t bytes)
    {
        for (size_t i = 0; i < pinnedSpare.size(); ++i)
            if (pinnedSpare[i].second >= bytes) {
                void* p = pinnedSpare[i].first;
                pinnedSpare.erase(pinnedSpare.begin() + std::ptrdiff_t(i));
                return p;
            }
        void* p = nullptr;
        const size_t capacity = std::max<size_t>(bytes, 64 * sizeof(clay_tail_result));
        check(hipHostMalloc(&p, capacity, hipHostMallocDefault), "pinned results");
        pinnedSizes[p] = capacity;
        return p;
    }

    void unpin(void* p)
    {
        if (p)
            pinnedSpare.emplace_back(p, pinnedSizes[p]);
    }
    std::map<void*, size_t> pinnedSizes;

    int channels() const { return natural ? 1 : 2; }

    hipEvent_t event()
    {
        if (!eventPool.empty()) {
            auto e = eventPool.back();
            eventPool.pop_back();
            return e;
        }
        hipEvent_t e;
        check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }

    // Launch wrapper: chunked grids plus optional per-launch events.
    template<class F>
    void launch(const char* name, F&& body)
    {
        Timer t{name, nullptr, nullptr};
        if (timing) {
            t.begin = event();
            t.end = event();
            check(hipEventRecord(t.begin, stream), "hipEventRecord");
        }
        body();
        check(hipGetLastError(), name);
        ++stats.launches;
        if (timing) {
            check(hipEventRecord(t.end, stream), "hipEventRecord");
            timers.push_back(t);
        }
    }

    // Tuner measurements whose launches have completed (all of them with `wait`).
    void collectTuning(bool wait)
    {
        auto& pending = tuner.pending;
        size_t kept = 0;
        for (size_t i = 0; i < pending.size(); ++i) {
            auto& m = pending[i];
            const hipError_t q = wait ? hipEventSynchronize(m.end) : hipEventQuery(m.end);
            if (q == hipErrorNotReady) {
                pending[kept++] = m;
                continue;
            }
            check(q, "tuner timing");
            float ms = 0;
            check(hipEventElapsedTime(&ms, m.begin, m.end), "hipEventElapsedTime");
            m.dimension->record(m.candidate, ms, m.work);
            eventPool.push_back(m.begin);
            eventPool.push_back(m.end);
        }
        pending.resize(kept);
    }

    // Runs `launch(value)` with the dimension's next candidate, timing it when
    // it is a trial.
    template<class F>
    void tuned(LaunchTuner::Dimension& d, double work, F&& launch)
    {
        const int candidate = d.next();
        if (candidate < 0) {
            launch(d.value(candidate));
            return;
        }
        LaunchTuner::Pending m{&d, candidate, event(), event(), work};
        check(hipEventRecord(m.begin, stream), "hipEventRecord");
        launch(d.value(candidate));
        check(hipEventRecord(m.end, stream), "hipEventRecord");
        ++d.issued[size_t(candidate)];
        tuner.pending.push_back(m);
    }

    void collectTimes(std::map<std::string, std::pair<double, int>>& out)
    {
        for (auto& t : timers) {
            float ms = 0;
            check(hipEventElapsedTime(&ms, t.begin, t.end), "hipEventElapsedTime");
            auto& slot = out[t.name];
            slot.first += ms;
            slot.second += 1;
            eventPool.push_back(t.begin);
            eventPool.push_back(t.end);
        }
        timers.clear();
    }
};

namespace {

void validatePlan(const clay_hip* h, const clay_plan_view& p, size_t index)
{
    const auto bad = [&](const std::string& what) {
        throw Failure(CLAY_E_INVALID, "plan " + std::to_string(index) + ": " + what);
    };
    if (p.frames < 0 || p.span_count < 0 || p.track_count < 3 || !p.track_offsets || p.modulation_count < 0 ||
        p.onset_count < 0 || p.filter_tracks < 0)
        bad("negative count or fewer than three control tracks");
    if ((p.span_count && !p.spans) || (p.modulation_count && !p.modulation) || (p.onset_count && !p.onsets))
        bad("missing array");
    if (p.track_offsets[0] != 0)
        bad("track offsets must start at 0");
    for (int t = 0; t < p.track_count; ++t)
        if (p.track_offsets[t + 1] <= p.track_offsets[t])
            bad("empty or unordered control track");
    if (p.track_offsets[p.track_count] && !p.points)
        bad("missing control points");
    for (int m = 0; m < p.modulation_count; ++m) {
        const auto& b = p.modulation[m];
        if (b.track >= p.track_count || b.shape_track >= p.track_count || b.target < 0 || b.target > CLAY_MOD_EQ3_GAIN)
            bad("modulation binding out of range");
    }
    for (int s = 0; s < p.span_count; ++s) {
        const auto& v = p.spans[s];
        if (v.begin < 0 || v.end <= v.begin || v.end > p.frames)
            bad("span outside the job");
        if (v.channels != 1 && v.channels != 2)
            bad("span channels must be 1 or 2");
        if (v.first_modulation < 0 || v.modulation_count < 0 || v.first_modulation + v.modulation_count > p.modulation_count)
            bad("span modulation range");
        if (v.filter_track >= p.filter_tracks)
            bad("span filter track");
        if (v.variable_pitch && v.filter_track < 0)
            bad("variable pitch span without a voice chain");
        if (v.generator < CLAY_GEN_NONE || v.generator > CLAY_GEN_NOISE)
            bad("span generator");
        if (v.generator == CLAY_GEN_NONE) {
            if (v.source_end < 0 || v.sample + uint64_t(v.source_end) * uint64_t(v.channels) > h->bankCount)
                bad("span source outside the sample bank");
            if (v.loop_end > 0 && (v.loop_first < 0 || v.loop_first >= v.loop_end || v.loop_end > v.source_end))
                bad("span loop");
        }
        for (const auto& f : v.filters)
            if (f.type < CLAY_FILTER_NONE || f.type > CLAY_FILTER_PEAK)
                bad("filter type");
        if (v.filter_track >= 0 && v.sample_rate <= 0)
            bad("filter sample rate");
    }
}

// CPUs this process may run on (its affinity mask, which a container or
// scheduler may set well below the machine's count).
size_t hostThreads()
{
    static const size_t n = [] {
        cpu_set_t set;
        CPU_ZERO(&set);
        if (sched_getaffinity(0, sizeof(set), &set) == 0 && CPU_COUNT(&set) > 0)
            return size_t(CPU_COUNT(&set));
        return size_t(std::max(1u, std::thread::hardware_concurrency()));
    }();
    return n;
}

// Runs f(i) for i in [0, count) on up to `limit` host threads, at most 16 and
// at most the CPUs of the process (the plans of a batch are independent once
// their record offsets are known).
template<class F>
void parallelFor(size_t count, F&& f, size_t limit = 16)
{
    const size_t threads = std::min<size_t>({count, limit, 16, hostThreads()});
    if (threads <= 1) {
        for (size_t i = 0; i < count; ++i)
            f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    pool.reserve(threads);
    for (size_t t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < count;)
                f(i);
        });
    for (auto& thread : pool)
        thread.join();
}

// A span the straight mix renders (clayMixStraight, and the straight path of
// clayMixWaves): dry, mono, sampled, no variable pitch, stereo matrix or
// modulation of its audio, no loop and no bend, at positions from 0 to 2^30
// that never fall with the frame.
bool straightSpan(const clay_plan_view& p, const clay_voice_span& v)
{
    if (v.filter_track >= 0 || v.channels != 1 || v.generator != CLAY_GEN_NONE || v.variable_pitch ||
        v.stereo_matrix || v.loop_end > 0 || v.pitch_slope != 0)
        return false;
    for (int k = 0; k < v.modulation_count; ++k)
        if (p.modulation[v.first_modulation + k].target <= CLAY_MOD_POSITION)
            return false;
    return v.position >= 0 && v.step >= 0 && v.position + double(v.end - v.begin) * v.step < 0x1p30;
}

// Flatten plans into device records (renderer.cpp:378-462 semantics). Three
// passes: per plan (parallel) validation and the sizes that depend on its
// spans; record offsets by prefix sums (serial); per plan (parallel) the
// records at those offsets. Records land exactly where a serial append would
// put them, on up to `threads` host threads (callers that flatten several
// batches at once share the CPUs between them).
void flatten(clay_hip* h, const clay_plan_view* plans, size_t count, clay_hip_batch& b, size_t threads = 16)
{
    const auto forEachPlan = [&](auto&& f) { parallelFor(count, f, threads); };
    const bool assess = h->cfg.audio.enabled || h->cfg.group_minimum;
    b.assess = assess;
    const int window = h->tail.window;
    const auto bodyWindows = [&](int frames) {
        const int chunk = 2 * h->cfg.audio.rate, span = 5 * h->cfg.audio.window;
        return (frames / chunk) * int(ceilDiv(uint64_t(chunk), uint64_t(span))) + int(ceilDiv(uint64_t(frames % chunk), uint64_t(span)));
    };
    // 1. Validation and span-dependent sizes.
    struct Sizes {
        int voices = 0, straightVoices = 0;
        bool straight = true;
        uint64_t filtered = 0, pitch = 0;
        std::string error;
    };
    std::vector<Sizes> sizes(count);
    // Chain ids of each plan start after the earlier plans' chains.
    std::vector<int> chainBase(count, 0);
    for (size_t i = 1; i < count; ++i)
        chainBase[i] = chainBase[i - 1] + std::max(0, plans[i - 1].filter_tracks);
    const bool bounded = !b.scratchEnd.empty();
    if (bounded && (count == 0 ? 0 : size_t(chainBase[count - 1] + std::max(0, plans[count - 1].filter_tracks))) !=
                       b.scratchEnd.size())
        throw Failure(CLAY_E_INVALID, "scratch bounds do not match the voice chains of the batch");
    // Frames of a filtered span backed by scratch: up to its chain's bound.
    const auto scratchFrames = [&](size_t i, const clay_voice_span& v) {
        if (!bounded)
            return int64_t(v.end - v.begin);
        const int32_t bound = b.scratchEnd[size_t(chainBase[i] + v.filter_track)];
        return std::max<int64_t>(0, int64_t(std::min(v.end, bound)) - v.begin);
    };
    forEachPlan([&](size_t i) {
        const auto& p = plans[i];
        Sizes& z = sizes[i];
        try {
            validatePlan(h, p, i);
        } catch (const Failure& f) {
            z.error = f.what();
            return;
        }
        for (int s = 0; s < p.span_count; ++s) {
            const auto& v = p.spans[s];
            z.voices += (v.end - 1) / kTileFrames - v.begin / kTileFrames + 1;
            z.straightVoices += (v.end - 1) / kStraightFrames - v.begin / kStraightFrames + 1;
            z.straight = z.straight && straightSpan(p, v);
            if (v.filter_track >= 0) {
                z.filtered += uint64_t(scratchFrames(i, v)) * uint64_t(v.channels);
                if (v.variable_pitch)
                    z.pitch += uint64_t(v.end - v.begin);
            }
        }
    });
    b.straight = false;
    for (const auto& z : sizes) {
        if (!z.error.empty())
            throw Failure(CLAY_E_INVALID, z.error);
        b.straight = b.straight || z.voices > 0;
    }
    for (const auto& z : sizes)
        b.straight = b.straight && z.straight;
    const int tileFrames = b.straight ? kStraightFrames : kTileFrames;
    // 2. Offsets of each plan's records.
    struct Base {
        int point, mod, tile, voice, span, chain, onset, block, window, body;
        uint64_t frame, filtered, pitch;
    };
    std::vector<Base> bases(count);
    Base at{};
    for (size_t i = 0; i < count; ++i) {
        const auto& p = plans[i];
        bases[i] = at;
        at.point += p.track_offsets[p.track_count];
        at.mod += p.modulation_count;
        at.tile += int(ceilDiv(uint64_t(p.frames), uint64_t(tileFrames)));
        at.voice += b.straight ? sizes[i].straightVoices : sizes[i].voices;
        at.span += p.span_count;
        at.chain += p.filter_tracks;
        at.onset += assess && h->cfg.audio.enabled ? p.onset_count : 0;
        at.block += int(ceilDiv(uint64_t(p.frames), kTailBlock));
        at.window += h->natural ? int(ceilDiv(uint64_t(p.frames), uint64_t(window))) : 0;
        if (assess)
            at.body += bodyWindows(h->cfg.audio.enabled || p.volume_group ? p.body_frames : 0);
        at.frame += uint64_t(p.frames);
        at.filtered += sizes[i].filtered;
        at.pitch += sizes[i].pitch;
    }
    if (!b.pool) {
        b.pool = h->records;
        RecordSet kept = b.pool->take();
        b.spans = std::move(kept.spans);
        b.cold = std::move(kept.cold);
        b.tiles = std::move(kept.tiles);
        b.voices = std::move(kept.voices);
    }
    b.points.resize(size_t(at.point));
    b.mods.resize(size_t(at.mod));
    b.jobs.resize(count);
    b.tiles.resize(size_t(at.tile));
    b.voices.resize(size_t(at.voice));
    b.spans.resize(size_t(at.span));
    const bool needCold = at.chain > 0;  // filter and pitch recurrences read it
    b.cold.resize(needCold ? size_t(at.span) : 0);
    if (assess)
        b.audioJobs.resize(count);
    b.onsets.resize(size_t(at.onset));
    b.chains = at.chain;
    b.blocks = at.block;
    b.windows = at.window;
    b.bodyBlocks = at.body;
    b.frames = at.frame;
    b.filteredSamples = at.filtered;
    b.pitchFrames = at.pitch;
    // Tiles by frame (the tiles of a frame window form one contiguous range),
    // plans in order within a frame: tile t of plan i goes to slot
    // first[t] + (plans before i with more than t tiles).
    Records<int> tileSlot(size_t(at.tile));
    {
        int most = 0;
        for (size_t i = 0; i < count; ++i)
            most = std::max(most, int(ceilDiv(uint64_t(plans[i].frames), uint64_t(tileFrames))));
        std::vector<int> next(size_t(most) + 1, 0);
        for (size_t i = 0; i < count; ++i)
            ++next[ceilDiv(uint64_t(plans[i].frames), uint64_t(tileFrames))];
        // next[t]: plans with exactly t tiles -> first slot of frame t.
        int slot = 0, alive = int(count);
        for (int t = 0; t < most; ++t) {
            alive -= next[size_t(t)];
            next[size_t(t)] = slot;
            slot += alive;
        }
        for (size_t i = 0; i < count; ++i) {
            const int n = int(ceilDiv(uint64_t(plans[i].frames), uint64_t(tileFrames)));
            for (int t = 0; t < n; ++t)
                tileSlot[size_t(bases[i].tile + t)] = next[size_t(t)]++;
        }
    }
    // 3. Each plan's records, and each voice chain's head, frame count, first
    // and last frame (in head order), so that sorting them walks no span list.
    struct ChainInfo {
        int head;
        int64_t frames;
        int begin, end;
    };
    std::vector<std::vector<ChainInfo>> chainInfo(count);
    std::vector<std::vector<int>> pitchHeads(count);
    std::vector<char> mixedChannels(count, 0), dry(count, 0), complexDry(count, 0);
    std::vector<std::exception_ptr> failed(count);
    forEachPlan([&](size_t i) {
        try {
            const auto& p = plans[i];
            const Base& o = bases[i];
            const int jobIndex = int(i);
            // Controls and modulation bindings with absolute ranges.
            std::copy(p.points, p.points + p.track_offsets[p.track_count], b.points.begin() + o.point);
            const auto trackFirst = [&](int t) { return o.point + p.track_offsets[t]; };
            for (int m = 0; m < p.modulation_count; ++m) {
                const auto& src = p.modulation[m];
                Mod d{-1, 0, src.target, src.depth, -1, 0};
                if (src.track >= 0) {
                    d.first = trackFirst(src.track);
                    d.last = trackFirst(src.track + 1);
                }
                if (src.shape_track >= 0) {
                    d.shapeFirst = trackFirst(src.shape_track);
                    d.shapeLast = trackFirst(src.shape_track + 1);
                }
                b.mods[size_t(o.mod + m)] = d;
            }
            Job j{};
            j.mix = o.frame;
            j.frames = p.frames;
            for (int t = 0; t < 4; ++t)
                j.controls[t] = trackFirst(t);
            j.musicEnd = p.music_end;
            j.limit = p.limit;
            j.fade = p.fade_frames;
            j.blockFirst = o.block;
            j.windowFirst = o.window;
            const int body = h->cfg.audio.enabled || p.volume_group ? p.body_frames : 0;
            j.bodyFrames = body;
            j.bodyFirst = o.body;
            j.onsetFirst = o.onset;
            j.onsetCount = assess && h->cfg.audio.enabled ? p.onset_count : 0;
            b.jobs[i] = j;
            if (assess) {
                b.audioJobs[i] = {body, o.body, j.onsetFirst, j.onsetCount};
                if (h->cfg.audio.enabled)
                    std::copy(p.onsets, p.onsets + p.onset_count, b.onsets.begin() + o.onset);
            }
            // Tiles and their voice lists (difference array over tile indices).
            const int tileCount = int(ceilDiv(uint64_t(p.frames), uint64_t(tileFrames)));
            std::vector<int> perTile(size_t(tileCount) + 1, 0);
            for (int s = 0; s < p.span_count; ++s) {
                ++perTile[size_t(p.spans[s].begin / tileFrames)];
                --perTile[size_t((p.spans[s].end - 1) / tileFrames + 1)];
            }
            for (int t = 1; t <= tileCount; ++t)
                perTile[size_t(t)] += perTile[size_t(t - 1)];
            std::vector<int> cursor(static_cast<size_t>(tileCount), 0);
            int voiceAt = o.voice;
            for (int t = 0; t < tileCount; ++t) {
                const int frame = t * tileFrames;
                b.tiles[size_t(tileSlot[size_t(o.tile + t)])] = {jobIndex, frame, std::min(tileFrames, p.frames - frame),
                                                                 voiceAt, perTile[size_t(t)]};
                cursor[size_t(t)] = voiceAt;
                voiceAt += perTile[size_t(t)];
            }
            // Spans, voice chains, scratch offsets.
            std::vector<int> chainTail(size_t(p.filter_tracks), -1), chainSlot(size_t(p.filter_tracks), -1);
            uint64_t filtered = o.filtered, pitch = o.pitch;
            for (int s = 0; s < p.span_count; ++s) {
                const auto& v = p.spans[s];
                const int index = o.span + s;
                for (int t = v.begin / tileFrames; t <= (v.end - 1) / tileFrames; ++t)
                    b.voices[size_t(cursor[size_t(t)]++)] = index;
                Span d{};
                d.sample = v.sample;
                d.position = v.position;
                d.step = v.step;
                d.pitchSlope = v.pitch_slope;
                d.begin = v.begin;
                d.end = v.end;
                d.sourceEnd = v.source_end;
                d.loopFirst = v.loop_first;
                d.loopEnd = v.loop_end;
                d.loopCrossfade = v.loop_crossfade;
                d.channels = v.channels;
                d.generator = v.generator;
                d.stereoMatrix = v.stereo_matrix;
                d.variablePitch = v.variable_pitch;
                d.noiseSeed = v.noise_seed;
                d.noiseStart = v.noise_start;
                d.firstMod = v.first_modulation + o.mod;
                d.modCount = v.modulation_count;
                d.filterTrack = v.filter_track >= 0 ? o.chain + v.filter_track : -1;
                d.job = jobIndex;
                d.next = -1;
                d.scratchEnd = INT32_MAX;
                d.attackStart = v.attack_start;
                d.holdStart = v.hold_start;
                d.decayStart = v.decay_start;
                d.releaseStart = v.release_start;
                d.fadeStart = v.fade_start;
                d.start = v.start;
                d.attackStep = v.attack_step;
                d.decayRate = v.decay_rate;
                d.sustain = v.sustain;
                d.releaseValue = v.release_value;
                d.releaseRate = v.release_rate;
                d.fadeValue = v.fade_value;
                d.fadeStep = v.fade_step;
                d.gain = v.gain;
                d.pan = v.pan;
                d.left = v.left;
                d.right = v.right;
                d.widthDirect = v.width_direct;
                d.widthCross = v.width_cross;
                d.positionLeft = v.position_left;
                d.positionRight = v.position_right;
                d.width = v.width;
                d.stereoPosition = v.stereo_position;
                bool audioMods = false;
                for (int k = 0; k < v.modulation_count; ++k)
                    audioMods = audioMods || p.modulation[v.first_modulation + k].target <= CLAY_MOD_POSITION;
                d.plain = v.filter_track < 0 && v.channels == 1 && v.generator == CLAY_GEN_NONE && !v.variable_pitch &&
                          !v.stereo_matrix && !audioMods;
                d.straight = straightSpan(p, v);
                SpanCold c{};
                std::memcpy(c.filters, v.filters, sizeof c.filters);
                std::memcpy(c.extra, v.extra_envelope, sizeof c.extra);
                c.active[0] = v.active_envelope[0];
                c.active[1] = v.active_envelope[1];
                c.sampleRate = v.sample_rate;
                if (v.filter_track < 0)
                    dry[i] = 1;
                if (v.filter_track < 0 && !d.straight)
                    complexDry[i] = 1;
                if (v.filter_track >= 0) {
                    int& tail = chainTail[size_t(v.filter_track)];
                    int& slot = chainSlot[size_t(v.filter_track)];
                    if (tail < 0) {
                        if (v.variable_pitch)
                            pitchHeads[i].push_back(index);
                        slot = int(chainInfo[i].size());
                        chainInfo[i].push_back({index, 0, v.begin, 0});
                    } else {
                        b.spans[size_t(tail)].next = index;
                        // The filter pass compiles a chain for its head span's channel count.
                        if (v.channels != p.spans[s - (index - chainInfo[i][size_t(slot)].head)].channels)
                            mixedChannels[i] = 1;
                    }
                    ChainInfo& info = chainInfo[i][size_t(slot)];
                    info.frames += v.end - v.begin;
                    info.end = std::max(info.end, v.end);
                    tail = index;
                    d.filtered = filtered;
                    filtered += uint64_t(scratchFrames(i, v)) * uint64_t(v.channels);
                    if (bounded)
                        d.scratchEnd = std::min(v.end, b.scratchEnd[size_t(o.chain + v.filter_track)]);
                    if (v.variable_pitch) {
                        d.phase = pitch;
                        pitch += uint64_t(v.end - v.begin);
                    }
                }
                b.spans[size_t(index)] = d;
                if (needCold)
                    b.cold[size_t(index)] = c;
            }
        } catch (...) {
            failed[i] = std::current_exception();
        }
    });
    for (const auto& e : failed)
        if (e)
            std::rethrow_exception(e);
    for (size_t i = 0; i < count; ++i)
        if (mixedChannels[i])
            throw Failure(CLAY_E_INVALID, "a voice chain mixes mono and stereo spans");
    b.allWet = std::find(dry.begin(), dry.end(), 1) == dry.end();
    b.simpleMix = std::find(complexDry.begin(), complexDry.end(), 1) == complexDry.end();
    std::vector<ChainInfo> chains;
    for (size_t i = 0; i < count; ++i) {
        chains.insert(chains.end(), chainInfo[i].begin(), chainInfo[i].end());
        b.pitchHeads.insert(b.pitchHeads.end(), pitchHeads[i].begin(), pitchHeads[i].end());
    }
    // Longest voice chains first: each is one wave's serial work.
    std::vector<std::pair<int64_t, int>> order;
    order.reserve(chains.size());
    for (size_t c = 0; c < chains.size(); ++c)
        order.emplace_back(-chains[c].frames, int(c));
    std::stable_sort(order.begin(), order.end());
    b.chainHeads.resize(chains.size());
    for (size_t i = 0; i < order.size(); ++i)
        b.chainHeads[i] = chains[size_t(order[i].second)].head;
    // Windowed rounds: chains by first frame, with their extents.
    std::vector<std::pair<int2, int>> byStart;
    byStart.reserve(chains.size());
    for (size_t i = 0; i < order.size(); ++i) {
        const ChainInfo& c = chains[size_t(order[i].second)];
        byStart.push_back({make_int2(c.begin, c.end), c.head});
    }
    std::stable_sort(byStart.begin(), byStart.end(), [](const auto& x, const auto& y) { return x.first.x < y.first.x; });
    for (const auto& c : byStart) {
        b.chainByStart.push_back(c.second);
        b.chainExtent.push_back(c.first);
    }
    for (const auto& j : b.jobs)
        b.lastFrame = std::max(b.lastFrame, j.frames);
    // Volume groups: consecutive jobs sharing a nonzero id.
    if (h->cfg.group_minimum)
        for (size_t first = 0, end; first < count; first = end) {
            end = first + 1;
            if (!plans[first].volume_group)
                continue;
            while (end < count && plans[end].volume_group == plans[first].volume_group)
                ++end;
            b.groups.push_back(make_int2(int(first), int(end)));
        }
}

void release(clay_hip* h, clay_hip_batch& b)
{
    if (b.arena) {
        if (b.cachedArena && h)
            h->arenas.give(b.arena);
        else
            (void)hipFree(b.arena);
    }
    b.arena = nullptr;
    if (b.hostScratchBuffer) {
        if (h)
            (void)hipStreamSynchronize(h->stream);  // the kernels that used it have finished
        (void)hipHostFree(b.hostScratchBuffer);
    }
    b.hostScratchBuffer = nullptr;
}

// Filtered spans of a batch as the memory policy sees them.
std::vector<clay_memory::ScratchUse> scratchUses(const clay_hip_batch& b)
{
    std::vector<clay_memory::ScratchUse> uses;
    for (const auto& v : b.spans)
        if (v.filterTrack >= 0)
            uses.push_back({v.begin, v.end, v.channels, v.variablePitch != 0, v.scratchEnd == INT32_MAX
                                                                               ? INT32_MAX
                                                                               : std::max(0, v.scratchEnd - v.begin)});
    return uses;
}

// Scratch offsets of the filtered spans for render windows of `window`
// frames (0: whole spans), in span order.
void assignScratch(clay_hip_batch& b, int window)
{
    uint64_t floats = 0, doubles = 0;
    for (auto& v : b.spans) {
        if (v.filterTrack < 0)
            continue;
        int64_t frames = v.scratchEnd == INT32_MAX ? v.end - v.begin : std::max(0, std::min(v.end, v.scratchEnd) - v.begin);
        int64_t phases = v.end - v.begin;
        if (window > 0) {
            frames = std::min<int64_t>(frames, window);
            phases = std::min<int64_t>(phases, window);
        }
        v.filtered = floats;
        floats += uint64_t(frames) * uint64_t(v.channels);
        if (v.variablePitch) {
            v.phase = doubles;
            doubles += uint64_t(phases);
        }
    }
    b.filteredSamples = floats;
    b.pitchFrames = doubles;
    b.window = window;
}

// Where each record array and buffer of a batch sits in its one device arena:
// the records first ([0, records), copied from one host image), then the
// scratch and finishing buffers.
struct ArenaLayout {
    size_t spans, cold, mods, points, jobs, tiles, voices, pitch, chains, byStart, audio, onsets, groups;
    size_t records;
    size_t states, mix, filtered, phase, mono, stereo, windows, body, onsetPower, pcmJobs, output, results;
    size_t total;
};

ArenaLayout layoutOf(const clay_hip* h, const clay_hip_batch& b)
{
    ArenaBuilder a;
    ArenaLayout L{};
    const int ch = h->channels();
    L.spans = a.add(bytesOf(b.spans));
    L.cold = a.add(bytesOf(b.cold));
    L.mods = a.add(bytesOf(b.mods));
    L.points = a.add(bytesOf(b.points));
    L.jobs = a.add(bytesOf(b.jobs));
    L.tiles = a.add(bytesOf(b.tiles));
    L.voices = a.add(bytesOf(b.voices));
    L.pitch = a.add(bytesOf(b.pitchHeads));
    L.chains = a.add(bytesOf(b.chainHeads));
    L.byStart = a.add(bytesOf(b.chainByStart));
    L.audio = a.add(bytesOf(b.audioJobs));
    L.onsets = a.add(bytesOf(b.onsets));
    L.groups = a.add(bytesOf(b.groups));
    L.records = a.bytes;
    L.states = a.add(size_t(b.chains) * sizeof(VoiceState));
    L.mix = a.add(size_t(b.frames) * (h->natural ? sizeof(float2) : sizeof(short2)));
    L.filtered = a.add(b.hostScratch ? 0 : size_t(b.filteredSamples) * sizeof(float));
    L.phase = a.add(b.hostScratch ? 0 : size_t(b.pitchFrames) * sizeof(double));
    L.mono = a.add(h->natural ? size_t(b.frames) * sizeof(int16_t) : 0);
    L.stereo = a.add(h->natural ? size_t(b.blocks) * sizeof(StereoStats) : 0);
    L.windows = a.add(size_t(b.windows) * sizeof(MonoStats));
    L.body = a.add(size_t(b.bodyBlocks) * sizeof(unsigned long long));
    L.onsetPower = a.add(b.onsets.size() * sizeof(unsigned long long));
    L.pcmJobs = a.add(b.assess ? b.jobs.size() * sizeof(PcmJob) : 0);
    L.output = a.add(b.assess ? size_t(b.frames) * size_t(ch) * sizeof(int16_t) : 0);
    L.results = a.add(b.jobs.size() * sizeof(clay_tail_result));
    L.total = a.bytes;
    return L;
}

// The batch's records at their arena offsets in a host image of L.records
// bytes; large arrays are copied by up to `threads` threads.
void packRecords(const clay_hip_batch& b, const ArenaLayout& L, char* image, size_t threads)
{
    const auto put = [&](size_t offset, const auto& v) {
        const size_t bytes = bytesOf(v);
        const auto* src = reinterpret_cast<const char*>(v.data());
        constexpr size_t kPiece = size_t(8) << 20;
        if (threads > 1 && bytes > 2 * kPiece)
            parallelFor(ceilDiv(bytes, kPiece), [&](size_t k) {
                const size_t at = k * kPiece;
                std::memcpy(image + offset + at, src + at, std::min(kPiece, bytes - at));
            }, threads);
        else if (bytes)
            std::memcpy(image + offset, src, bytes);
    };
    put(L.spans, b.spans);
    put(L.cold, b.cold);
    put(L.mods, b.mods);
    put(L.points, b.points);
    put(L.jobs, b.jobs);
    put(L.tiles, b.tiles);
    put(L.voices, b.voices);
    put(L.pitch, b.pitchHeads);
    put(L.chains, b.chainHeads);
    put(L.byStart, b.chainByStart);
    put(L.audio, b.audioJobs);
    put(L.onsets, b.onsets);
    put(L.groups, b.groups);
}

// The arena: from the renderer's cache (persistent form) or its own block.
void allocateArena(clay_hip* h, clay_hip_batch& b, const ArenaLayout& L, bool cached)
{
    if (h->memoryLimit && L.total > h->memoryLimit)
        throw Failure(CLAY_E_MEMORY, "batch arena: over the renderer's memory limit");
    if (cached)
        b.arena = h->arenas.take(L.total);
    else
        check(hipMalloc(&b.arena, L.total), "batch arena");
    b.cachedArena = cached;
    b.arenaBytes = L.total;
}

// Queues the copy of the packed records into the arena, points the batch's
// device views into it and clears the filter states.
void placeRecords(clay_hip* h, clay_hip_batch& b, const ArenaLayout& L, const char* image)
{
    auto* base = static_cast<char*>(b.arena);
    check(hipMemcpyAsync(base, image, L.records, hipMemcpyHostToDevice, h->stream), "batch upload");
    b.dSpans = reinterpret_cast<Span*>(base + L.spans);
    b.dCold = reinterpret_cast<SpanCold*>(base + L.cold);
    b.dMods = reinterpret_cast<Mod*>(base + L.mods);
    b.dPoints = reinterpret_cast<clay_control_point*>(base + L.points);
    b.dJobs = reinterpret_cast<Job*>(base + L.jobs);
    b.dTiles = reinterpret_cast<Tile*>(base + L.tiles);
    b.dVoices = reinterpret_cast<int*>(base + L.voices);
    b.dPitchHeads = reinterpret_cast<int*>(base + L.pitch);
    b.dChainHeads = reinterpret_cast<int*>(base + L.chains);
    b.dChainByStart = reinterpret_cast<int*>(base + L.byStart);
    b.dAudioJobs = reinterpret_cast<AudioJobDesc*>(base + L.audio);
    b.dOnsets = reinterpret_cast<int*>(base + L.onsets);
    b.dGroups = reinterpret_cast<int2*>(base + L.groups);
    b.dStates = reinterpret_cast<VoiceState*>(base + L.states);
    b.dMix = base + L.mix;
    b.dFiltered = reinterpret_cast<float*>(base + L.filtered);
    b.dPhase = reinterpret_cast<double*>(base + L.phase);
    if (b.hostScratch) {
        // Mapped host memory the kernels reach over the bus (slow, but it fits).
        const size_t floatBytes = (size_t(b.filteredSamples) * sizeof(float) + 255) & ~size_t(255);
        const size_t bytes = std::max<size_t>(1, floatBytes + size_t(b.pitchFrames) * sizeof(double));
        check(hipHostMalloc(&b.hostScratchBuffer, bytes, hipHostMallocMapped), "host scratch");
        void* device = nullptr;
        check(hipHostGetDevicePointer(&device, b.hostScratchBuffer, 0), "host scratch");
        b.dFiltered = static_cast<float*>(device);
        b.dPhase = reinterpret_cast<double*>(static_cast<char*>(device) + floatBytes);
    }
    b.dMono = reinterpret_cast<int16_t*>(base + L.mono);
    b.dStereo = h->natural ? reinterpret_cast<StereoStats*>(base + L.stereo) : nullptr;
    b.dWindows = reinterpret_cast<MonoStats*>(base + L.windows);
    b.dBody = reinterpret_cast<unsigned long long*>(base + L.body);
    b.dOnsetPower = reinterpret_cast<unsigned long long*>(base + L.onsetPower);
    b.dPcmJobs = reinterpret_cast<PcmJob*>(base + L.pcmJobs);
    // Natural: the packed mono PCM reuses the float mix once it is quantised.
    // orig/custom: finalize works in place on the short2 mix.
    b.dOutput = b.assess ? reinterpret_cast<int16_t*>(base + L.output) : reinterpret_cast<int16_t*>(b.dMix);
    b.dResults = reinterpret_cast<clay_tail_result*>(base + L.results);
    if (b.chains)
        check(hipMemsetAsync(b.dStates, 0, size_t(b.chains) * sizeof(VoiceState), h->stream), "state reset");
    h->residentBytes = L.total;
    // The bank and its shifted copy, the float and integer sinc tables.
    h->stats.peak_bytes = std::max<uint64_t>(h->stats.peak_bytes, L.total + 2 * (uint64_t(h->bankOdd) + 2) +
                                                                      size_t(h->cfg.sinc_taps) * (kPhases + 1) * 8);
}

// Plain form: the records through the renderer's one pinned image, grown as
// needed (batches upload one at a time), and a wait for the copy.
void upload(clay_hip* h, clay_hip_batch& b)
{
    const ArenaLayout L = layoutOf(h, b);
    allocateArena(h, b, L, false);
    char* image = h->staging.take(L.records);
    packRecords(b, L, image, 16);
    placeRecords(h, b, L, image);
    h->staging.commit();
    check(hipStreamSynchronize(h->stream), "batch upload");
}

// A plain batch onto the device within the memory it can get
// (renderer.cpp:864-989): render windows short enough for the filtered and
// phase scratch to fit the scratch budget; after a failed allocation, halve
// that scratch while it exceeds 2 MiB, then give up with CLAY_E_MEMORY so
// that the caller halves the batch, and for a single job move the scratch to
// mapped host memory before giving up.
void placeWithinMemory(clay_hip* h, clay_hip_batch& b)
{
    const auto uses = scratchUses(b);
    size_t freeBytes = 0, totalBytes = 0;
    check(hipMemGetInfo(&freeBytes, &totalBytes), "hipMemGetInfo");
    if (h->memoryLimit) {
        freeBytes = std::min<size_t>(freeBytes, h->memoryLimit);
        totalBytes = std::min<size_t>(totalBytes, h->memoryLimit);
    }
    const ArenaLayout whole = layoutOf(h, b);
    const uint64_t fullScratch = clay_memory::scratchBytes(uses, 0);
    const uint64_t fixed = whole.total > fullScratch ? whole.total - fullScratch : 0;
    uint64_t budget = clay_memory::scratchBudget(freeBytes, totalBytes, fixed, UINT64_MAX);
    int window = clay_memory::windowFor(uses, budget, b.lastFrame);
    if (window)
        assignScratch(b, window);
    for (;;) {
        try {
            upload(h, b);
            return;
        } catch (const Failure& f) {
            if (f.code != CLAY_E_MEMORY)
                throw;
            release(h, b);
            const uint64_t scratch = clay_memory::scratchBytes(uses, b.window);
            auto next = clay_memory::afterOutOfMemory(scratch, b.jobs.size(), b.hostScratch);
            if (next == clay_memory::Next::halveScratch) {
                budget = scratch / 2;
                const int shorter = clay_memory::windowFor(uses, budget, b.lastFrame);
                if (clay_memory::scratchBytes(uses, shorter) < scratch) {
                    assignScratch(b, shorter);
                    continue;
                }
                next = b.jobs.size() > 1 ? clay_memory::Next::halveJobs
                                         : b.hostScratch ? clay_memory::Next::reject : clay_memory::Next::hostScratch;
            }
            if (next == clay_memory::Next::hostScratch) {
                b.hostScratch = true;
                continue;
            }
            throw;  // halve the batch, or reject its one job
        }
    }
}

constexpr int kMaxBlocksPerLaunch = 1 << 20;

template<class F>
void chunked(int blocks, F&& launchChunk)
{
    for (int first = 0; first < blocks; first += kMaxBlocksPerLaunch)
        launchChunk(first, std::min(kMaxBlocksPerLaunch, blocks - first));
}

// ---------------------------------------------------------------- rounds

// A resident unit and the frame window it renders in one round.
struct Window {
    clay_hip_batch* b;
    int first, end;
    int chainLo, chainHi;  // range of b->chainByStart (windowed); -1 = every chain, longest first
};

// Chains of a window [first, end) in start order: from the first chain that
// has not ended before `first` (every earlier one has) to the first chain that
// starts at or after `end`. `lo` only moves forward across a unit's rounds.
void windowChains(const clay_hip_batch& b, int first, int end, int& lo, int& hi)
{
    const int n = int(b.chainExtent.size());
    while (lo < n && b.chainExtent[size_t(lo)].y <= first)
        ++lo;
    hi = lo;
    while (hi < n && b.chainExtent[size_t(hi)].x < end)
        ++hi;
}

UnitView viewOf(const Window& w)
{
    const clay_hip_batch& b = *w.b;
    UnitView v{};
    v.spans = b.dSpans;
    v.cold = b.dCold;
    v.mods = b.dMods;
    v.points = b.dPoints;
    v.jobs = b.dJobs;
    v.tiles = b.dTiles;
    v.voices = b.dVoices;
    v.pitchHeads = b.dPitchHeads;
    v.phase = b.dPhase;
    v.filtered = b.dFiltered;
    v.states = b.dStates;
    v.mix = b.dMix;
    v.segFirst = w.first;
    v.segEnd = w.end;
    // Tiles overlapping the window: window bounds are multiples of 256 frames
    // (or the batch end), straight tiles start on multiples of 1024.
    const int tileFrames = b.straight ? kStraightFrames : kTileFrames;
    const int from = w.first / tileFrames * tileFrames;
    const int first = int(std::lower_bound(b.tiles.begin(), b.tiles.end(), from,
                                           [](const Tile& t, int f) { return t.frame < f; }) -
                          b.tiles.begin());
    const int count = int(std::lower_bound(b.tiles.begin() + first, b.tiles.end(), w.end,
                                           [](const Tile& t, int f) { return t.frame < f; }) -
                          b.tiles.begin()) - first;
    if (b.straight) {
        v.straightTiles = b.dTiles;
        v.straightVoices = b.dVoices;
        v.straightFirst = first;
        v.straightCount = count;
        // One window over the whole batch: the mix stores the tail-block
        // statistics (clayScanStereo is skipped).
        v.stereo = w.first == 0 && w.end >= b.lastFrame ? b.dStereo : nullptr;
        v.mono = v.stereo ? b.dMono : nullptr;
    } else {
        v.tileFirst = first;
        v.tileCount = count;
    }
    if (w.chainLo < 0) {
        v.chainHeads = b.dChainHeads;
        v.chainCount = int(b.chainHeads.size());
    } else {
        v.chainHeads = b.dChainByStart + w.chainLo;
        v.chainCount = w.chainHi - w.chainLo;
    }
    v.pitchCount = int(b.pitchHeads.size());
    return v;
}

// Launch shape of a round: grids of the three render kernels (0 = sized to the work).
struct Shape {
    int mixBlocks = 0, chainBlocks = 0, pitchBlocks = 0, audioBlocks = 0;
    bool simpleMix = false;  // every resident unit's dry voices are plain and straight
};

constexpr int kChainWavesPerBlock = 2;   // render.hip: voice chains per 128-thread filter block (round kernel)
constexpr int kAudioWavesPerBlock = 16;  // render.hip: voice chains per 1024-thread audio block
constexpr int kSharedTapsMax = 32;       // render.hip: widest sinc table the audio pass keeps in LDS
constexpr int kMixWaves = 16;            // mix.h: tiles (one per wave) in flight per 1024-thread block
constexpr int kStraightWaves = 12;       // mix.h: straight tiles in flight per 768-thread block

// Grid of the wave mix for `tiles` tiles: one 1024-thread block per CU at
// most, a multiple of the XCD count (the kernel maps blocks to XCDs).
int mixGrid(clay_hip* h, int tiles)
{
    const int cus = std::max(kXcds, h->props.multiProcessorCount / kXcds * kXcds);
    const int wanted = int(std::min<int64_t>(cus, (int64_t(tiles) + kMixWaves - 1) / kMixWaves));
    return std::max(kXcds, (wanted + kXcds - 1) / kXcds * kXcds);
}

// Grid of the straight mix: one 768-thread block per CU at most (each holds
// the 160 KB weight table in LDS), a multiple of the XCD count.
int straightGrid(clay_hip* h, int tiles)
{
    const int cus = std::max(kXcds, h->props.multiProcessorCount / kXcds * kXcds);
    const int wanted = int(std::min<int64_t>(cus, (int64_t(tiles) + kStraightWaves - 1) / kStraightWaves));
    return std::max(kXcds, (wanted + kXcds - 1) / kXcds * kXcds);
}

// Persistent form: the round kernels with the fixed grids of `shape`.
template<int T>
void launchRound(clay_hip* h, const Shape& shape, bool timed)
{
    hipStream_t s = h->stream;
    const auto* units = reinterpret_cast<const UnitView*>(h->dRound + 256);
    const auto* round = reinterpret_cast<const RoundHeader*>(h->dRound);
    const auto run = [&](const char* name, auto&& body) {
        if (timed)
            h->launch(name, body);
        else {
            body();
            check(hipGetLastError(), name);
            ++h->stats.launches;
        }
    };
    run("clayRoundPitchChains", [&] {
        hipLaunchKernelGGL(clayRoundPitchChains, dim3(shape.pitchBlocks), dim3(64), 0, s, units, round);
    });
    run("clayRoundChainAudio", [&] {
        hipLaunchKernelGGL(clayRoundChainAudio<T>, dim3(shape.audioBlocks), dim3(64 * kAudioWavesPerBlock), 0, s,
                           h->bank, h->sinc, units, round);
    });
    run("clayRoundVoiceChains", [&] {
        hipLaunchKernelGGL(clayRoundVoiceChains, dim3(shape.chainBlocks), dim3(64 * kChainWavesPerBlock), 0, s, units,
                           round);
    });
    run("clayClearQueues", [&] { hipLaunchKernelGGL(clayClearQueues, dim3(1), dim3(64), 0, s, h->queues, 2); });
    run("clayRoundMixWaves", [&] {
        const int blocks = mixGrid(h, 1 << 30);
        const auto mix = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * kMixWaves), 0, s, h->bank, h->bankOdd, h->sincInt,
                               h->sinc, units, round, h->queues + 0);
        };
        if (shape.simpleMix)
            h->natural ? mix(clayRoundMixWaves<T, true, true>) : mix(clayRoundMixWaves<T, false, true>);
        else
            h->natural ? mix(clayRoundMixWaves<T, true, false>) : mix(clayRoundMixWaves<T, false, false>);
    });
    run("clayRoundMixStraight", [&] {
        const int blocks = straightGrid(h, 1 << 30);
        const auto mix = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * kStraightWaves), 0, s, h->bank, h->bankOdd,
                               h->sincInt, units, round, h->queues + 1);
        };
        h->natural ? mix(clayRoundMixStraight<T, true>) : mix(clayRoundMixStraight<T, false>);
    });
}

// Plain form: one unit's window with the direct kernels and exact grids.
template<int T>
void launchWindow(clay_hip* h, const Window& w)
{
    hipStream_t s = h->stream;
    clay_hip_batch& b = *w.b;
    const UnitView v = viewOf(w);
    b.stereoDone = v.stereo != nullptr;
    const int pitchCount = v.pitchCount, chainCount = v.chainCount;
    if (pitchCount)
        h->launch("clayPitchChains", [&] {
            hipLaunchKernelGGL(clayPitchChains, dim3(ceilDiv(pitchCount, 64)), dim3(64), 0, s, b.dSpans, b.dCold,
                               b.dPitchHeads, b.dPoints, b.dMods, b.dPhase, b.dStates, w.first, w.end, pitchCount);
        });
    if (chainCount) {
        // With the taps in LDS one audio block fits a CU: at most one block per CU, grid-stride.
        int audioBlocks = ceilDiv(chainCount, kAudioWavesPerBlock);
        if (T <= kSharedTapsMax)
            audioBlocks = std::min(audioBlocks, std::max(1, h->props.multiProcessorCount));
        // Wet samples of the window (the batch's, pro rata): the tuner's unit of work.
        const double work = double(b.filteredSamples) * double(w.end - w.first) / double(std::max(1, b.lastFrame));
        h->collectTuning(false);
        h->tuned(h->tuner.audioUnroll, work, [&](int unroll) {
            h->launch("clayChainAudio", [&] {
                const auto audio = [&](auto kernel) {
                    hipLaunchKernelGGL(kernel, dim3(audioBlocks), dim3(64 * kAudioWavesPerBlock), 0, s, h->bank,
                                       h->sinc, b.dSpans, b.dJobs, v.chainHeads, b.dPoints, b.dMods, b.dPhase,
                                       b.dFiltered, b.dStates, w.first, w.end, chainCount);
                };
                unroll == 2 ? audio(clayChainAudio<T, 2>) : audio(clayChainAudio<T, 1>);
            });
        });
        h->tuned(h->tuner.chainWaves, work, [&](int waves) {
            h->launch("clayVoiceChains", [&] {
                const auto chains = [&](auto kernel) {
                    hipLaunchKernelGGL(kernel, dim3(ceilDiv(chainCount, waves)), dim3(64 * waves), 0, s, b.dSpans,
                                       b.dCold, v.chainHeads, b.dPoints, b.dMods, b.dFiltered, b.dStates, w.first,
                                       w.end, chainCount, h->silenceLevel);
                };
                waves == 1 ? chains(clayVoiceChains<1>) : waves == 4 ? chains(clayVoiceChains<4>) : chains(clayVoiceChains<2>);
            });
        });
    }
    if (b.allWet) {
        // No dry voice: the scratch-only mix, four voices' frames in flight per thread.
        chunked(v.tileCount, [&](int first, int n) {
            h->launch("clayMixTiles", [&] {
                const auto mix = [&](auto kernel) {
                    hipLaunchKernelGGL(kernel, dim3(n), dim3(kTileFrames), 0, s, h->bank, h->sinc, b.dSpans, b.dJobs,
                                       b.dTiles, b.dVoices, b.dPoints, b.dMods, b.dFiltered, b.dStates, b.dMix,
                                       w.first, w.end, v.tileFirst + first);
                };
                h->natural ? mix(clayMixTiles<T, true, true>) : mix(clayMixTiles<T, false, true>);
            });
        });
        return;
    }
    if (b.straight || v.tileCount) {
        hipLaunchKernelGGL(clayClearQueues, dim3(1), dim3(64), 0, s, h->queues, 2);
        check(hipGetLastError(), "clayClearQueues");
    }
    if (b.straight) {
        if (v.straightCount)
            h->launch("clayMixStraight", [&] {
                const int blocks = straightGrid(h, v.straightCount);
                const auto mix = [&](auto kernel) {
                    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * kStraightWaves), 0, s, h->bank, h->bankOdd,
                                       h->sincInt, b.dSpans, b.dJobs, b.dTiles, b.dVoices, b.dPoints, b.dMix, w.first, w.end,
                                       v.straightFirst, v.straightCount, h->queues + 1, v.stereo, v.mono);
                };
                h->natural ? mix(clayMixStraight<T, true>) : mix(clayMixStraight<T, false>);
            });
        return;
    }
    if (v.tileCount)
        h->launch("clayMixWaves", [&] {
            const int blocks = mixGrid(h, v.tileCount);
            const auto mix = [&](auto kernel) {
                hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * kMixWaves), 0, s, h->bank, h->bankOdd, h->sincInt, h->sinc,
                                   b.dSpans, b.dJobs, b.dTiles, b.dVoices, b.dPoints, b.dMods, b.dFiltered, b.dStates,
                                   b.dMix, w.first, w.end, v.tileFirst, v.tileCount, h->queues + 0);
            };
            if (b.simpleMix)
                h->natural ? mix(clayMixWaves<T, true, true>) : mix(clayMixWaves<T, false, true>);
            else
                h->natural ? mix(clayMixWaves<T, true, false>) : mix(clayMixWaves<T, false, false>);
        });
}

void launchRoundTaps(clay_hip* h, const Shape& shape, bool timed)
{
    switch (h->cfg.sinc_taps) {
    case 16: launchRound<16>(h, shape, timed); break;
    case 32: launchRound<32>(h, shape, timed); break;
    case 48: launchRound<48>(h, shape, timed); break;
    default: launchRound<64>(h, shape, timed); break;
    }
}

// Copies the round's descriptors to the device through the next pinned stage.
RoundHeader stageRound(clay_hip* h, const std::vector<Window>& windows)
{
    if (windows.size() > h->unitCapacity)
        throw Failure(CLAY_E_INVALID, "round holds more units than the renderer was created for");
    auto& stage = h->stages[h->stageNext];
    h->stageNext = (h->stageNext + 1) % clay_hip::kStages;
    if (stage.used)
        check(hipEventSynchronize(stage.copied), "round staging");
    RoundHeader header{int(windows.size()), 0, 0, 0, h->silenceLevel, 0, {0, 0}};
    auto* views = reinterpret_cast<UnitView*>(static_cast<char*>(stage.host) + 256);
    for (size_t i = 0; i < windows.size(); ++i) {
        UnitView v = viewOf(windows[i]);
        windows[i].b->stereoDone = v.stereo != nullptr;
        v.tileBase = header.tiles;
        v.straightBase = header.straightTiles;
        v.chainBase = header.chains;
        v.pitchBase = header.pitches;
        header.tiles += v.tileCount;
        header.straightTiles += v.straightCount;
        header.chains += v.chainCount;
        header.pitches += v.pitchCount;
        views[i] = v;
    }
    *static_cast<RoundHeader*>(stage.host) = header;
    check(hipMemcpyAsync(h->dRound, stage.host, 256 + windows.size() * sizeof(UnitView), hipMemcpyHostToDevice,
                         h->stream), "round descriptors");
    check(hipEventRecord(stage.copied, h->stream), "hipEventRecord");
    stage.used = true;
    return header;
}

// One segment of a plain batch.
void renderRound(clay_hip* h, const Window& w)
{
    switch (h->cfg.sinc_taps) {
    case 16: launchWindow<16>(h, w); break;
    case 32: launchWindow<32>(h, w); break;
    case 48: launchWindow<48>(h, w); break;
    default: launchWindow<64>(h, w); break;
    }
    ++h->stats.segments;
}

// The captured round for a launch shape, captured on first use. The round's
// windows are read from the device descriptors at replay, so one graph
// serves every round of that shape.
const clay_hip::Graph& roundGraph(clay_hip* h, const Shape& shape)
{
    const clay_hip::Graph* found = nullptr;
    for (const auto& g : h->graphs)
        if (g.mixBlocks == shape.mixBlocks && g.chainBlocks == shape.chainBlocks && g.simpleMix == shape.simpleMix &&
            g.bank == h->bank && g.bankOdd == h->bankOdd)
            found = &g;
    if (!found) {
        hipGraph_t graph = nullptr;
        check(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            launchRoundTaps(h, shape, false);
        } catch (...) {
            (void)hipStreamEndCapture(h->stream, &graph);
            if (graph)
                (void)hipGraphDestroy(graph);
            throw;
        }
        check(hipStreamEndCapture(h->stream, &graph), "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check(e, "hipGraphInstantiate");
        h->graphs.push_back({exec, shape.mixBlocks, shape.chainBlocks, shape.simpleMix, h->bank, h->bankOdd});
        found = &h->graphs.back();
    }
    return *found;
}

// One round with fixed grids (persistent form): a replay of the captured
// round for this launch shape.
void renderRoundFixed(clay_hip* h, const std::vector<Window>& windows, const Shape& shape, bool useGraph)
{
    stageRound(h, windows);
    ++h->stats.segments;
    ++h->stats.rounds;
    if (!useGraph) {
        launchRoundTaps(h, shape, true);
        return;
    }
    const clay_hip::Graph* found = &roundGraph(h, shape);
    if (h->timing) {
        h->launch("round", [&] { check(hipGraphLaunch(found->exec, h->stream), "hipGraphLaunch"); });
    } else {
        check(hipGraphLaunch(found->exec, h->stream), "hipGraphLaunch");
    }
    ++h->stats.graph_launches;
}

// Segments of a plain batch: the whole batch, or windows of segment_frames.
void renderPlain(clay_hip* h, clay_hip_batch& b)
{
    if (b.chains)
        check(hipMemsetAsync(b.dStates, 0, size_t(b.chains) * sizeof(VoiceState), h->stream), "state reset");
    int step = h->cfg.segment_frames;
    if (b.window > 0)  // the scratch holds windows of b.window frames at most
        step = step > 0 ? std::min(step, b.window) : b.window;
    if (step <= 0 || step >= b.lastFrame) {
        renderRound(h, Window{&b, 0, b.lastFrame, -1, -1});
        return;
    }
    int lo = 0, hi = 0;
    for (int first = 0; first < b.lastFrame; first += step) {
        const int end = std::min(b.lastFrame, first + step);
        windowChains(b, first, end, lo, hi);
        renderRound(h, Window{&b, first, end, lo, hi});
    }
}

// Tail, quality, group gain and packing of a rendered batch (renderer.cpp:1046-1142).
void finishBatch(clay_hip* h, clay_hip_batch& b)
{
    hipStream_t s = h->stream;
    const int count = int(b.jobs.size());
    const int ch = h->channels();
    const auto mix2 = static_cast<const float2*>(b.dMix);
    const int lanes = 64;
    if (h->natural) {
        if (!b.stereoDone)
            chunked(b.blocks, [&](int first, int n) {
                h->launch("clayScanStereo", [&] {
                    hipLaunchKernelGGL(clayScanStereo, dim3(n), dim3(256), 0, s, b.dJobs, count, mix2, b.dStereo, first);
                });
            });
        h->launch("claySelectNative", [&] {
            hipLaunchKernelGGL(claySelectNative, dim3(count), dim3(64), 0, s, b.dJobs, count, mix2, b.dStereo,
                               b.dResults);
        });
        chunked(b.windows, [&](int first, int n) {
            h->launch("clayQuantizeMono", [&] {
                hipLaunchKernelGGL(clayQuantizeMono, dim3(ceilDiv(n, kQuantizeWaves)), dim3(64 * kQuantizeWaves), 0, s,
                                   b.dJobs, count, mix2, b.dResults, b.dMono, b.dWindows, h->tail.window, first, n,
                                   b.stereoDone ? 1 : 0);
            });
        });
        h->launch("claySelectTrim", [&] {
            hipLaunchKernelGGL(claySelectTrim, dim3(count), dim3(64), 0, s, b.dJobs, count, b.dWindows, b.dResults,
                               h->tail);
        });
    }
    h->launch("clayLayoutPcm", [&] {
        hipLaunchKernelGGL(clayLayoutPcm, dim3(1), dim3(1), 0, s, b.dJobs, count, b.dResults, h->natural ? 0 : 1);
    });
    // Natural: mono scratch -> packed (mix region). orig/custom: in place on the short2 mix.
    const int16_t* finalIn = h->natural ? b.dMono : static_cast<const int16_t*>(b.dMix);
    int16_t* packed = static_cast<int16_t*>(b.dMix);
    chunked(b.blocks, [&](int first, int n) {
        h->launch("clayFinalizePcm", [&] {
            hipLaunchKernelGGL(clayFinalizePcm, dim3(ceilDiv(ceilDiv(n, kFinalizeBlocks), kFinalizeWaves)),
                               dim3(64 * kFinalizeWaves), 0, s, b.dJobs, count,
                               finalIn, packed, b.dResults, ch, h->cfg.audio.enabled ? 1 : 0, first, n);
        });
    });
    if (b.assess) {
        chunked(b.bodyBlocks, [&](int first, int n) {
            h->launch("clayMeasureBody", [&] {
                hipLaunchKernelGGL(clayMeasureBody, dim3(n), dim3(256), 0, s, b.dAudioJobs, count, packed, b.dResults,
                                   b.dBody, h->cfg.audio, ch, first);
            });
        });
        if (h->cfg.audio.enabled && !b.onsets.empty())
            chunked(int(b.onsets.size()), [&](int first, int n) {
                h->launch("clayMeasureOnsets", [&] {
                    hipLaunchKernelGGL(clayMeasureOnsets, dim3(n), dim3(320), 0, s, b.dAudioJobs, count, b.dOnsets, packed,
                                       b.dResults, b.dOnsetPower, h->cfg.audio.window, ch, first);
                });
            });
        h->launch("claySummarizeAudio", [&] {
            hipLaunchKernelGGL(claySummarizeAudio, dim3(ceilDiv(count, lanes)), dim3(lanes), 0, s, b.dAudioJobs, count,
                               b.dBody, b.dOnsetPower, b.dResults, h->cfg.audio, ch);
        });
        if (!b.groups.empty())
            h->launch("clayMatchVolume", [&] {
                hipLaunchKernelGGL(clayMatchVolume, dim3(ceilDiv(b.groups.size(), lanes)), dim3(lanes), 0, s, b.dGroups,
                                   int(b.groups.size()), h->cfg.group_minimum, b.dResults);
            });
        h->launch("clayLayoutAccepted", [&] {
            hipLaunchKernelGGL(clayLayoutAccepted, dim3(1), dim3(1), 0, s, b.dJobs, count, b.dResults, b.dPcmJobs);
        });
        chunked(b.blocks, [&](int first, int n) {
            h->launch("clayPackAccepted", [&] {
                hipLaunchKernelGGL(clayPackAccepted, dim3(n), dim3(256), 0, s, b.dPcmJobs, count, packed, b.dOutput,
                                   b.dResults, ch, first);
            });
        });
    }
}

// PCM16 of a finished batch into a pinned slot (renderer.cpp:672-690).
// The smallest free slot that fits is reused. A new slot is a power of two
// (at least 1 MiB) so that batches of varying size reuse slots; free slots
// too small are given back (hipHostFree is slow) only when the slots hold
// more than kSlotBudget.
void* pinnedSlot(clay_hip* h, size_t bytes)
{
    std::lock_guard<std::mutex> guard(h->slotLock);
    constexpr size_t kSlotBudget = size_t(16) << 30;
    clay_hip::Slot* best = nullptr;
    size_t held = 0;
    for (auto& slot : h->slots) {
        held += slot.capacity;
        if (!slot.busy && slot.capacity >= bytes && (!best || slot.capacity < best->capacity))
            best = &slot;
    }
    if (best) {
        best->busy = true;
        return best->host;
    }
    size_t capacity = size_t(1) << 20;
    while (capacity < bytes)
        capacity *= 2;
    for (size_t i = 0; i < h->slots.size() && held + capacity > kSlotBudget;)
        if (!h->slots[i].busy) {
            check(hipHostFree(h->slots[i].host), "hipHostFree");
            held -= h->slots[i].capacity;
            h->slots.erase(h->slots.begin() + std::ptrdiff_t(i));
        } else {
            ++i;
        }
    clay_hip::Slot slot;
    check(hipHostMalloc(&slot.host, capacity, hipHostMallocDefault), "pinned PCM");
    slot.capacity = capacity;
    slot.busy = true;
    h->slots.push_back(slot);
    return slot.host;
}

template<class F>
int guarded(clay_hip* h, F&& body)
{
    try {
        body();
        return CLAY_OK;
    } catch (const Failure& f) {
        if (h)
            h->error = f.what();
        return f.code;
    } catch (const std::exception& e) {
        if (h)
            h->error = e.what();
        return CLAY_E_INVALID;
    }
}

void initTail(clay_hip* h)
{
    const int rate = h->cfg.audio.rate;
    h->tail = {h->cfg.audio.window, 2 * rate, int(std::nearbyint(rate / 4.)), int(std::nearbyint(rate / 10.)),
               std::pow(10., -1.5), 32768 * std::pow(10., -3.9), 32768 * std::pow(10., -3.)};
}

} // namespace

extern "C" {

int clay_hip_sinc_table(int taps, float* out, size_t capacity)
{
    if (taps < 4 || taps > CLAY_MAX_SINC_TAPS || taps % 4 || !out || capacity < size_t(taps) * (kPhases + 1))
        return CLAY_E_INVALID;
    const auto t = sincTable(taps);
    std::memcpy(out, t.data(), t.size() * sizeof(float));
    return CLAY_OK;
}

int clay_hip_create(const clay_hip_config* config, clay_hip** out)
{
    if (!config || !out) {
        g_createError = "null argument";
        return CLAY_E_INVALID;
    }
    auto h = std::make_unique<clay_hip>();
    const int status = guarded(h.get(), [&] {
        h->cfg = *config;
        const int taps = config->sinc_taps;
        if (taps != 16 && taps != 32 && taps != 48 && taps != 64)
            throw Failure(CLAY_E_INVALID, "sinc_taps must be 16, 32, 48 or 64");
        if (config->tail_policy < CLAY_POLICY_NATURAL || config->tail_policy > CLAY_POLICY_CUSTOM)
            throw Failure(CLAY_E_INVALID, "unknown tail policy");
        if (config->audio.rate <= 0 || config->audio.window <= 0)
            throw Failure(CLAY_E_INVALID, "audio rate and window must be positive");
        if (config->batch_jobs == 0)
            h->cfg.batch_jobs = 64;
        if (config->segment_frames < 0 || config->segment_frames % kTileFrames)
            throw Failure(CLAY_E_INVALID, "segment_frames must be a non-negative multiple of 256");
        if (config->form != CLAY_FORM_PLAIN && config->form != CLAY_FORM_PERSISTENT)
            throw Failure(CLAY_E_INVALID, "unknown execution form");
        h->natural = config->tail_policy == CLAY_POLICY_NATURAL;
        if (config->voice_silence_db > 0 || !(config->voice_silence_db > -400 || config->voice_silence_db == 0))
            throw Failure(CLAY_E_INVALID, "voice_silence_db must be 0 (off) or a negative level in dBFS");
        h->silenceLevel = config->voice_silence_db < 0 ? float(std::pow(10.0, config->voice_silence_db / 20.0)) : 0.f;
        initTail(h.get());
        const bool tune = !(config->flags & CLAY_FLAG_NO_TUNE);
        h->tuner.chainWaves.init({1, 2, 4}, 2, tune);
        h->tuner.audioUnroll.init({1, 2}, 1, tune);
        int devices = 0;
        check(hipGetDeviceCount(&devices), "hipGetDeviceCount");
        if (config->device < 0 || config->device >= devices)
            throw Failure(CLAY_E_DEVICE, "no HIP device " + std::to_string(config->device));
        h->device = config->device;
        check(hipSetDevice(h->device), "hipSetDevice");
        check(hipGetDeviceProperties(&h->props, h->device), "hipGetDeviceProperties");
        std::snprintf(h->stats.device, sizeof h->stats.device, "%s", h->props.name);
        check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
        // Phase-major table: each lane reads its taps contiguously.
        const auto tapMajor = sincTable(taps);
        // The audio pass reads phases above kPhases / 2 as mirrored rows of the lower half.
        for (int t = 0; t < taps; ++t)
            for (int p = 0; p <= kPhases; ++p)
                if (tapMajor[size_t(t) * (kPhases + 1) + size_t(p)] !=
                    tapMajor[size_t(taps - 1 - t) * (kPhases + 1) + size_t(kPhases - p)])
                    throw Failure(CLAY_E_DEVICE, "sinc table is not mirror-symmetric");
        std::vector<float> phaseMajor(tapMajor.size());
        for (int t = 0; t < taps; ++t)
            for (int p = 0; p <= kPhases; ++p)
                phaseMajor[size_t(p) * size_t(taps) + size_t(t)] = tapMajor[size_t(t) * (kPhases + 1) + size_t(p)];
        check(hipMalloc(&h->sinc, bytesOf(phaseMajor)), "sinc table");
        check(hipMemcpy(h->sinc, phaseMajor.data(), bytesOf(phaseMajor), hipMemcpyHostToDevice), "sinc upload");
        const auto packed = intSincTable(phaseMajor, taps);
        check(hipMalloc(&h->sincInt, bytesOf(packed)), "integer sinc table");
        check(hipMemcpy(h->sincInt, packed.data(), bytesOf(packed), hipMemcpyHostToDevice), "integer sinc upload");
        // Round descriptors: header + one view per resident unit.
        h->unitCapacity = size_t(std::min<uint64_t>(h->cfg.batch_jobs, 1u << 16)) + 1;
        const size_t roundBytes = 256 + h->unitCapacity * sizeof(UnitView);
        check(hipMalloc(&h->dRound, roundBytes), "round descriptors");
        check(hipMalloc(&h->queues, 2 * sizeof(MixQueue)), "mix queues");
        check(hipMemset(h->queues, 0, 2 * sizeof(MixQueue)), "mix queues");
        for (auto& stage : h->stages) {
            check(hipHostMalloc(&stage.host, roundBytes, hipHostMallocDefault), "round staging");
            check(hipEventCreateWithFlags(&stage.copied, hipEventDisableTiming), "hipEventCreate");
        }
        // Persistent units allocate stream-ordered; keep freed blocks in the pool.
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, h->device) == hipSuccess && pool) {
            uint64_t threshold = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
        }
        size_t freeBytes = 0, totalBytes = 0;
        check(hipMemGetInfo(&freeBytes, &totalBytes), "hipMemGetInfo");
        h->stats.sample_cache_budget = std::min<uint64_t>(freeBytes - freeBytes / 3, totalBytes / 2);
    });
    if (status != CLAY_OK) {
        g_createError = h->error;
        clay_hip_destroy(h.release());
        return status;
    }
    *out = h.release();
    return CLAY_OK;
}

int clay_hip_destroy(clay_hip* h)
{
    if (!h)
        return CLAY_OK;
    (void)hipSetDevice(h->device);
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    h->downloader.stop();
    for (auto& m : h->tuner.pending) {
        h->eventPool.push_back(m.begin);
        h->eventPool.push_back(m.end);
    }
    for (auto& b : h->arenas.blocks)
        (void)hipFree(b.base);
    for (auto& slot : h->slots)
        if (slot.host)
            (void)hipHostFree(slot.host);
    for (auto e : h->eventPool)
        (void)hipEventDestroy(e);
    for (auto& g : h->graphs)
        (void)hipGraphExecDestroy(g.exec);
    h->images.release();
    h->staging.release();
    for (auto& p : h->pinnedSizes)
        (void)hipHostFree(p.first);
    for (auto& stage : h->stages) {
        if (stage.host)
            (void)hipHostFree(stage.host);
        if (stage.copied)
            (void)hipEventDestroy(stage.copied);
    }
    if (h->dRound)
        (void)hipFree(h->dRound);
    if (h->queues)
        (void)hipFree(h->queues);
    if (h->bank)
        (void)hipFree(h->bank);
    if (h->sinc)
        (void)hipFree(h->sinc);
    if (h->sincInt)
        (void)hipFree(h->sincInt);
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return CLAY_OK;
}

int clay_hip_upload_bank(clay_hip* h, const int16_t* pcm, uint64_t count, int fresh)
{
    if (!h)
        return CLAY_E_INVALID;
    return guarded(h, [&] {
        check(hipSetDevice(h->device), "hipSetDevice");
        if (fresh) {
            check(hipStreamSynchronize(h->stream), "bank reset");
            if (h->bank)
                check(hipFree(h->bank), "hipFree");
            h->bank = nullptr;
            h->bankCount = h->bankCapacity = 0;
        }
        if (count <= h->bankCount && h->bank)
            return;
        if (count && !pcm)
            throw Failure(CLAY_E_INVALID, "null sample bank");
        constexpr uint64_t kPad = 2 * CLAY_MAX_SINC_TAPS;  // samples past the capacity (see below)
        if (count > h->bankCapacity || !h->bank) {
            // Append-only growth with slack (renderer.cpp:553-603).
            const uint64_t slack = std::min<uint64_t>(count / 2, 64ull << 20);
            const uint64_t capacity = std::max<uint64_t>(count + slack, 1);
            // Padding: a window clamped into a source shorter than the taps reads
            // up to taps samples from the source's start. The bank and its shifted
            // copy share one allocation, `stride` bytes apart; the mix addresses
            // the copy with 32-bit offsets from the bank.
            const uint64_t stride = ((capacity + kPad) * sizeof(int16_t) + 255) & ~uint64_t(255);
            if (stride > 0xF0000000ull)
                throw Failure(CLAY_E_MEMORY, "sample bank over 3.75 GB (the mix addresses it with 32-bit offsets)");
            int16_t* grown = nullptr;
            check(hipMalloc(&grown, 2 * stride), "sample bank");
            if (h->bankCount)
                check(hipMemcpyAsync(grown, h->bank, h->bankCount * sizeof(int16_t), hipMemcpyDeviceToDevice, h->stream),
                      "sample bank growth");
            check(hipStreamSynchronize(h->stream), "sample bank growth");
            if (h->bank)
                check(hipFree(h->bank), "hipFree");
            // Captured rounds hold the old bank's address and offset.
            for (auto& g : h->graphs)
                (void)hipGraphExecDestroy(g.exec);
            h->graphs.clear();
            h->bank = grown;
            h->bankCapacity = capacity;
            h->bankOdd = uint32_t(stride - 2);
            h->bankShifted = 0;  // the copy holds nothing yet
        }
        if (count > h->bankCount)
            check(hipMemcpy(h->bank + h->bankCount, pcm + h->bankCount, (count - h->bankCount) * sizeof(int16_t),
                            hipMemcpyHostToDevice), "sample upload");
        // The shifted copy from the first sample it lacks: copy[i] = bank[i + 1].
        const uint64_t first = h->bankShifted ? h->bankShifted - 1 : 0;
        char* copy = reinterpret_cast<char*>(h->bank) + (uint64_t(h->bankOdd) + 2);
        check(hipMemcpy(copy + 2 * first, h->bank + first + 1, (count + kPad - 1 - first) * sizeof(int16_t),
                        hipMemcpyDeviceToDevice), "shifted sample copy");
        h->bankShifted = count;
        h->bankCount = count;
    });
}

int clay_hip_batch_prepare(clay_hip* h, const clay_plan_view* plans, size_t count, clay_hip_batch** out)
{
    if (!h || !out || (count && !plans))
        return CLAY_E_INVALID;
    auto b = std::make_unique<clay_hip_batch>();
    const int status = guarded(h, [&] {
        if (!count)
            throw Failure(CLAY_E_INVALID, "empty batch");
        check(hipSetDevice(h->device), "hipSetDevice");
        flatten(h, plans, count, *b);
        placeWithinMemory(h, *b);
    });
    if (status != CLAY_OK) {
        release(h, *b);
        return status;
    }
    *out = b.release();
    return CLAY_OK;
}

int clay_hip_batch_run(clay_hip* h, clay_hip_batch* b, double* kernel_seconds)
{
    if (!h || !b)
        return CLAY_E_INVALID;
    return guarded(h, [&] {
        check(hipSetDevice(h->device), "hipSetDevice");
        hipEvent_t begin = h->event(), end = h->event();
        check(hipEventRecord(begin, h->stream), "hipEventRecord");
        const char* timing = std::getenv("CLAY_KERNEL_TIMING");
        h->timing = timing && *timing == '1';
        renderPlain(h, *b);
        finishBatch(h, *b);
        check(hipEventRecord(end, h->stream), "hipEventRecord");
        check(hipStreamSynchronize(h->stream), "batch run");
        h->collectTuning(false);
        float ms = 0;
        check(hipEventElapsedTime(&ms, begin, end), "hipEventElapsedTime");
        h->eventPool.push_back(begin);
        h->eventPool.push_back(end);
        b->kernelTimes.clear();
        h->collectTimes(b->kernelTimes);
        h->timing = false;
        if (kernel_seconds)
            *kernel_seconds = ms / 1000.0;
    });
}

int clay_hip_batch_kernel_time(clay_hip* h, clay_hip_batch* b, const char* kernel, double* ms, int* launches)
{
    if (!h || !b || !kernel)
        return CLAY_E_INVALID;
    const auto it = b->kernelTimes.find(kernel);
    if (ms)
        *ms = it == b->kernelTimes.end() ? 0 : it->second.first;
    if (launches)
        *launches = it == b->kernelTimes.end() ? 0 : it->second.second;
    return CLAY_OK;
}

int clay_hip_batch_results(clay_hip* h, clay_hip_batch* b, clay_tail_result* results, size_t result_capacity,
                           int16_t* pcm, uint64_t pcm_capacity, uint64_t* pcm_samples)
{
    if (!h || !b)
        return CLAY_E_INVALID;
    return guarded(h, [&] {
        check(hipSetDevice(h->device), "hipSetDevice");
        std::vector<clay_tail_result> r(b->jobs.size());
        check(hipMemcpy(r.data(), b->dResults, bytesOf(r), hipMemcpyDeviceToHost), "results download");
        uint64_t samples = 0;
        for (const auto& x : r)
            samples = std::max<uint64_t>(samples, (x.offset + uint64_t(x.data_frames)) * uint64_t(h->channels()));
        if (pcm_samples)
            *pcm_samples = samples;
        if (results) {
            if (result_capacity < r.size())
                throw Failure(CLAY_E_INVALID, "result capacity too small");
            std::memcpy(results, r.data(), bytesOf(r));
        }
        if (pcm) {
            if (pcm_capacity < samples)
                throw Failure(CLAY_E_INVALID, "pcm capacity too small");
            if (samples)
                check(hipMemcpy(pcm, b->dOutput, samples * sizeof(int16_t), hipMemcpyDeviceToHost), "pcm download");
            h->stats.downloaded_bytes += samples * sizeof(int16_t);
        }
    });
}

int clay_hip_batch_voice_tails(clay_hip* h, clay_hip_batch* b, int32_t* silence, size_t capacity, size_t* count)
{
    if (!h || !b)
        return CLAY_E_INVALID;
    return guarded(h, [&] {
        check(hipSetDevice(h->device), "hipSetDevice");
        const size_t n = size_t(b->chains);
        if (count)
            *count = n;
        if (!silence)
            return;
        if (capacity < n)
            throw Failure(CLAY_E_INVALID, "voice tail capacity too small");
        std::vector<VoiceState> states(n);
        if (n)
            check(hipMemcpy(states.data(), b->dStates, n * sizeof(VoiceState), hipMemcpyDeviceToHost), "voice tails");
        // Chain ends from the host records (chain id = span filter track).
        std::vector<int32_t> end(n, 0);
        for (const auto& sp : b->spans)
            if (sp.filterTrack >= 0)
                end[size_t(sp.filterTrack)] = std::max(end[size_t(sp.filterTrack)], sp.end);
        for (size_t i = 0; i < n; ++i)
            silence[i] = states[i].retired ? states[i].silence : end[i];
    });
}

int clay_hip_batch_prepare_bounded(clay_hip* h, const clay_plan_view* plans, size_t count, const int32_t* scratch_end,
                                   size_t chains, clay_hip_batch** out)
{
    if (!h || !out || (count && !plans) || (chains && !scratch_end))
        return CLAY_E_INVALID;
    auto b = std::make_unique<clay_hip_batch>();
    const int status = guarded(h, [&] {
        if (!count)
            throw Failure(CLAY_E_INVALID, "empty batch");
        if (chains && h->silenceLevel <= 0.f)
            throw Failure(CLAY_E_INVALID, "scratch bounds need the voice-silence policy");
        check(hipSetDevice(h->device), "hipSetDevice");
        b->scratchEnd.assign(scratch_end, scratch_end + chains);
        flatten(h, plans, count, *b);
        placeWithinMemory(h, *b);
    });
    if (status != CLAY_OK) {
        release(h, *b);
        return status;
    }
    *out = b.release();
    return CLAY_OK;
}

int clay_hip_batch_overflows(clay_hip* h, clay_hip_batch* b, int32_t* jobs, size_t capacity, size_t* count)
{
    if (!h || !b)
        return CLAY_E_INVALID;
    return guarded(h, [&] {
        check(hipSetDevice(h->device), "hipSetDevice");
        const size_t n = size_t(b->chains);
        std::vector<VoiceState> states(n);
        if (n)
            check(hipMemcpy(states.data(), b->dStates, n * sizeof(VoiceState), hipMemcpyDeviceToHost), "overflows");
        std::vector<char> flagged(b->jobs.size(), 0);
        for (const auto& sp : b->spans)
            if (sp.filterTrack >= 0 && states[size_t(sp.filterTrack)].overflow)
                flagged[size_t(sp.job)] = 1;
        size_t found = 0;
        for (size_t j = 0; j < flagged.size(); ++j)
            if (flagged[j]) {
                if (jobs && found < capacity)
                    jobs[found] = int32_t(j);
                ++found;
            }
        if (count)
            *count = found;
        if (jobs && found > capacity)
            throw Failure(CLAY_E_INVALID, "overflow capacity too small");
    });
}

int clay_hip_set_memory_limit(clay_hip* h, uint64_t bytes)
{
    if (!h)
        return CLAY_E_INVALID;
    h->memoryLimit = bytes;
    return CLAY_OK;
}

int clay_hip_batch_memory(clay_hip* h, clay_hip_batch* b, int32_t* window, int32_t* host_scratch)
{
    if (!h || !b)
        return CLAY_E_INVALID;
    if (window)
        *window = b->window;
    if (host_scratch)
        *host_scratch = b->hostScratch ? 1 : 0;
    return CLAY_OK;
}

int clay_hip_batch_scratch_bytes(clay_hip* h, clay_hip_batch* b, uint64_t* filtered, uint64_t* arena)
{
    if (!h || !b)
        return CLAY_E_INVALID;
    if (filtered)
        *filtered = b->filteredSamples * sizeof(float);
    if (arena)
        *arena = b->arenaBytes;
    return CLAY_OK;
}

int clay_hip_batch_destroy(clay_hip* h, clay_hip_batch* b)
{
    if (!b)
        return CLAY_OK;
    if (h) {
        (void)hipSetDevice(h->device);
        (void)hipStreamSynchronize(h->stream);
    }
    release(h, *b);
    delete b;
    return CLAY_OK;
}

} // extern "C"

namespace {

// Plain form: batches of up to batch_jobs jobs, never splitting a volume
// group, halving a batch that does not fit (renderer.cpp:1748-1820, 958-968).
int renderWindowPlain(clay_hip* h, const clay_plan_view* plans, size_t count, clay_publish_fn publish,
                      clay_refill_fn refill, void* user)
{
    // The next batch's records are flattened on a worker thread while the
    // current batch renders (flattening only reads the plans and the renderer's
    // configuration); its upload waits for its turn. A range that changes (a
    // halved batch) or a failed flatten falls back to the plain preparation.
    struct Ahead {
        size_t first = 0, end = 0;
        // The views of the batch being flattened, copied: the window's queue
        // may grow (refill) and move while the worker runs.
        std::vector<clay_plan_view> views;
        clay_hip_batch* b = nullptr;
        std::thread worker;
        bool ok = false;
        ~Ahead() { finish(); }
        void finish()
        {
            if (worker.joinable())
                worker.join();
        }
    } ahead;
    const auto dropAhead = [&] {
        ahead.finish();
        if (ahead.b) {
            release(h, *ahead.b);
            delete ahead.b;
            ahead.b = nullptr;
        }
    };
    return guarded(h, [&] {
        std::vector<clay_plan_view> queue(plans, plans + count);
        size_t first = 0;
        // A batch holds up to batch_jobs jobs and never splits a volume group.
        const auto batchEnd = [&](size_t from) {
            size_t end = std::min(queue.size(), from + size_t(h->cfg.batch_jobs));
            if (h->cfg.group_minimum)
                while (end < queue.size() && queue[end].volume_group && queue[end].volume_group == queue[end - 1].volume_group)
                    ++end;
            return end;
        };
        struct Cleanup {
            std::function<void()> f;
            ~Cleanup() { f(); }
        } cleanup{dropAhead};
        for (;;) {
            if (first == queue.size()) {
                const clay_plan_view* more = nullptr;
                size_t n = 0;
                if (!refill || !refill(user, 1, &more, &n) || !n)
                    break;
                queue.insert(queue.end(), more, more + n);
            }
            size_t end = batchEnd(first);
            clay_hip_batch* b = nullptr;
            int status = CLAY_OK;
            ahead.finish();
            if (ahead.b && ahead.ok && ahead.first == first && ahead.end == end) {
                b = ahead.b;
                ahead.b = nullptr;
                status = guarded(h, [&] {
                    check(hipSetDevice(h->device), "hipSetDevice");
                    placeWithinMemory(h, *b);
                });
                if (status != CLAY_OK) {
                    release(h, *b);
                    delete b;
                    b = nullptr;
                }
            }
            dropAhead();
            for (; !b;) {
                status = clay_hip_batch_prepare(h, queue.data() + first, end - first, &b);
                if (status != CLAY_E_MEMORY || end - first == 1)
                    break;
                size_t half = first + std::max<size_t>(1, (end - first) / 2);
                end = half;
            }
            if (status == CLAY_E_MEMORY) {
                clay_render_batch rejected{};
                rejected.failed_scratch_bytes = 1;
                if (publish(user, &rejected, first))
                    throw Failure(CLAY_E_CALLBACK, "publish aborted the window");
                first = end;
                continue;
            }
            if (status != CLAY_OK)
                throw Failure(status, h->error);
            // Flatten the next batch while this one renders.
            if (end < queue.size()) {
                ahead.first = end;
                ahead.end = batchEnd(end);
                ahead.b = new clay_hip_batch();
                ahead.ok = false;
                ahead.views.assign(queue.begin() + std::ptrdiff_t(ahead.first),
                                   queue.begin() + std::ptrdiff_t(ahead.end));
                ahead.worker = std::thread([h, &ahead] {
                    try {
                        flatten(h, ahead.views.data(), ahead.views.size(), *ahead.b);
                        ahead.ok = true;
                    } catch (...) {
                        ahead.ok = false;
                    }
                });
            }
            double seconds = 0;
            status = clay_hip_batch_run(h, b, &seconds);
            std::vector<clay_tail_result> results(end - first);
            uint64_t samples = 0;
            if (status == CLAY_OK)
                status = clay_hip_batch_results(h, b, results.data(), results.size(), nullptr, 0, &samples);
            void* slot = nullptr;
            if (status == CLAY_OK) {
                slot = pinnedSlot(h, samples * sizeof(int16_t));
                if (samples)
                    status = guarded(h, [&] {
                        check(hipMemcpy(slot, b->dOutput, samples * sizeof(int16_t), hipMemcpyDeviceToHost), "pcm download");
                    });
                h->stats.downloaded_bytes += samples * sizeof(int16_t);
            }
            const std::string error = h->error;
            clay_hip_batch_destroy(h, b);
            if (status != CLAY_OK)
                throw Failure(status, error);
            clay_render_batch batch{};
            batch.pcm = static_cast<const int16_t*>(slot);
            batch.pcm_handle = slot;
            batch.results = results.data();
            batch.result_count = results.size();
            batch.kernel_seconds = seconds;
            if (publish(user, &batch, first))
                throw Failure(CLAY_E_CALLBACK, "publish aborted the window");
            first = end;
        }
    });
}


// Persistent form (the role of renderRolling, renderer.cpp:1351-1746).
// Jobs are admitted in cohorts of consecutive jobs (1, 2, 4, ... up to
// batch_jobs / 8, never splitting a volume group), flattened on worker threads ahead of need. Up to
// batch_jobs jobs are rendering at once (a job whose frames are all rendered
// no longer counts, though its cohort stays resident until its last job is
// done); each round advances every resident cohort by
// `segment` frames with one replay of the captured round. A cohort whose last
// window was rendered is finished on the device right behind that round; its
// results, then its PCM, come back on a second stream while later rounds run,
// and it is published from the host loop. Its place is filled at the next
// round boundary. The host waits only for a staging buffer still in use, for
// a flatten when nothing is resident, or for the oldest completion when
// nothing is resident.
struct RoundTuner {
    std::vector<std::pair<int, int>> candidates;  // (segment frames, mix blocks)
    std::vector<double> cost;                     // ms per rendered frame, summed over trials
    std::vector<int> trials, issued;              // measured rounds completed / launched per candidate
    int segment = 0, mixBlocks = 0;
    bool tuning = false;
    static constexpr int kTrials = 2;

    void init(const clay_hip* h)
    {
        const int cus = std::max(1, h->props.multiProcessorCount);
        const int fixedSegment = h->cfg.segment_frames;
        segment = fixedSegment > 0 ? fixedSegment : 1 << 17;
        mixBlocks = cus * 16;
        tuning = !(h->cfg.flags & CLAY_FLAG_NO_TUNE);
        if (h->tunedSegment > 0) {  // tuned by an earlier window of this renderer
            segment = h->tunedSegment;
            tuning = false;
        }
        if (!tuning)
            return;
        for (int seg : {1 << 16, 1 << 17, 1 << 18, 1 << 19})
            if (fixedSegment <= 0 || seg == fixedSegment)
                candidates.emplace_back(seg, mixBlocks);
        if (fixedSegment > 0 && candidates.empty())
            candidates.emplace_back(fixedSegment, mixBlocks);
        cost.assign(candidates.size(), 0);
        trials.assign(candidates.size(), 0);
        issued.assign(candidates.size(), 0);
    }

    // Launch shape of the next round and the candidate it measures (-1: none).
    // Measurements complete asynchronously: until every candidate's trials
    // have come back, rounds past the last trial use the current default.
    std::pair<int, std::pair<int, int>> pick() const
    {
        if (tuning)
            for (size_t i = 0; i < candidates.size(); ++i)
                if (issued[i] < kTrials)
                    return {int(i), candidates[i]};
        return {-1, {segment, mixBlocks}};
    }

    void launched(int index) { ++issued[size_t(index)]; }

    void record(int index, double ms, uint64_t frames)
    {
        if (!tuning || index < 0 || !frames)
            return;
        cost[size_t(index)] += ms / double(frames);
        ++trials[size_t(index)];
        for (int t : trials)
            if (t < kTrials)
                return;
        size_t best = 0;
        for (size_t i = 1; i < cost.size(); ++i)
            if (cost[i] < cost[best])
                best = i;
        segment = candidates[best].first;
        mixBlocks = candidates[best].second;
        tuning = false;
    }

    // Keeps the choice for the renderer's later windows.
    void keep(clay_hip* h) const
    {
        if (!tuning && !(h->cfg.flags & CLAY_FLAG_NO_TUNE))
            h->tunedSegment = segment;
    }
};

int renderWindowPersistent(clay_hip* h, const clay_plan_view* plans, size_t count, clay_publish_fn publish,
                           clay_refill_fn refill, void* user)
{
    // Consecutive queued jobs (never splitting a volume group) flattened into
    // one batch: one arena, one record copy and one finishing sequence per
    // cohort instead of per job.
    struct Cohort {
        clay_hip_batch* b = nullptr;
        size_t first = 0, jobs = 0;
        int cursor = 0, chainLo = 0;
        hipEvent_t finished = nullptr, begin = nullptr;
        clay_tail_result* results = nullptr;  // pinned
        void* slot = nullptr;                 // pinned, taken by the download thread
        uint64_t samples = 0;
        // Set by the download thread under Downloads::lock.
        bool downloaded = false;
        int code = CLAY_OK;
        std::string message;
        std::vector<int> ends;  // frames of each job, ascending
        // Jobs still rendering: a finished job frees its place for admission
        // while the rest of its cohort renders on.
        size_t active() const { return size_t(ends.end() - std::upper_bound(ends.begin(), ends.end(), cursor)); }
    };
    // A cohort whose records a worker thread is flattening. The views are
    // copied: the window's queue may grow (refill) and move meanwhile.
    struct Flattening {
        size_t first = 0, end = 0;
        std::vector<clay_plan_view> views;
        clay_hip_batch* b = nullptr;
        ArenaLayout layout{};
        char* image = nullptr;  // packed records, pinned (h->images)
        std::thread worker;
        std::atomic<bool> done{false};
        int status = CLAY_OK;
        std::string message;
        void join()
        {
            if (worker.joinable())
                worker.join();
        }
    };
    // Finished cohorts come back through one thread that waits for each
    // cohort's finishing and copies its results, then its PCM, on the fastest
    // device-to-host SDMA engine (Downloader), so that downloads overlap the
    // rounds that follow at the full PCIe rate.
    struct Downloads {
        std::mutex lock;
        std::condition_variable changed;
        std::deque<Cohort*> queue;
        bool stop = false;
        std::vector<std::thread> workers;  // one per Downloader lane
        bool started() const { return !workers.empty(); }
        void start(clay_hip* h)
        {
            h->downloader.start(h->props);
            for (int lane = 0; lane < Downloader::kLanes; ++lane)
                workers.emplace_back([this, h, lane] { run(h, lane); });
        }
        void run(clay_hip* h, int lane)
        {
            (void)hipSetDevice(h->device);
            for (;;) {
                Cohort* u = nullptr;
                {
                    std::unique_lock<std::mutex> l(lock);
                    changed.wait(l, [&] { return stop || !queue.empty(); });
                    if (stop)
                        return;
                    u = queue.front();
                    queue.pop_front();
                }
                int code = CLAY_OK;
                std::string message;
                uint64_t samples = 0;
                try {
                    check(hipEventSynchronize(u->finished), "completion");
                    h->downloader.copy(u->results, u->b->dResults, u->jobs * sizeof(clay_tail_result),
                                       "results download", lane);
                    for (size_t k = 0; k < u->jobs; ++k)
                        samples = std::max<uint64_t>(
                            samples, (u->results[k].offset + uint64_t(u->results[k].data_frames)) * uint64_t(h->channels()));
                    void* slot = pinnedSlot(h, samples * sizeof(int16_t));
                    {
                        std::lock_guard<std::mutex> l(lock);
                        u->slot = slot;
                    }
                    h->downloader.copy(slot, u->b->dOutput, samples * sizeof(int16_t), "pcm download", lane);
                } catch (const Failure& e) {
                    code = e.code;
                    message = e.what();
                } catch (const std::exception& e) {
                    code = CLAY_E_INVALID;
                    message = e.what();
                }
                {
                    std::lock_guard<std::mutex> l(lock);
                    u->samples = samples;
                    u->code = code;
                    u->message = message;
                    u->downloaded = true;
                }
                changed.notify_all();
            }
        }
        void push(Cohort* u)
        {
            {
                std::lock_guard<std::mutex> l(lock);
                queue.push_back(u);
            }
            changed.notify_all();
        }
        // Stops after the cohorts being copied, if any; the rest stay queued.
        void finish()
        {
            {
                std::lock_guard<std::mutex> l(lock);
                stop = true;
            }
            changed.notify_all();
            for (auto& w : workers)
                if (w.joinable())
                    w.join();
        }
        ~Downloads() { finish(); }
    };
    constexpr size_t kAhead = 6;            // cohorts flattened ahead of admission
    constexpr size_t kCohortsPerWindow = 8; // resident cohorts when every one is full
    const size_t cohortJobs = std::max<size_t>(1, size_t(h->cfg.batch_jobs) / kCohortsPerWindow);

    std::vector<clay_plan_view> queue(plans, plans + count);
    std::vector<std::unique_ptr<Cohort>> resident, completing;
    std::deque<std::unique_ptr<Flattening>> ahead;
    Downloads downloads;  // declared after the cohorts: stopped before they are freed
    const auto drop = [&](Cohort& u) {
        if (u.b) {
            release(h, *u.b);
            delete u.b;
            u.b = nullptr;
        }
        h->unpin(u.results);
        u.results = nullptr;
        for (hipEvent_t* e : {&u.finished, &u.begin})
            if (*e) {
                h->eventPool.push_back(*e);
                *e = nullptr;
            }
    };
    const auto dropFlattening = [&](Flattening& f) {
        f.join();
        if (f.image)
            h->images.discard(f.image);
        f.image = nullptr;
        if (f.b) {
            release(h, *f.b);
            delete f.b;
            f.b = nullptr;
        }
    };
    const auto startFlatten = [&](size_t first, size_t end) {
        auto f = std::make_unique<Flattening>();
        f->first = first;
        f->end = end;
        f->views.assign(queue.begin() + std::ptrdiff_t(first), queue.begin() + std::ptrdiff_t(end));
        f->b = new clay_hip_batch();
        Flattening* raw = f.get();
        // The cohorts flattening at once share the CPUs, leaving two for this
        // loop and the download threads; a cohort started while fewer are in
        // flight (the first of a window, which rendering waits for) gets a
        // larger share.
        const size_t threads = std::max<size_t>(1, (std::max<size_t>(hostThreads(), 3) - 2) / (ahead.size() + 1));
        raw->worker = std::thread([h, raw, threads] {
            try {
                flatten(h, raw->views.data(), raw->views.size(), *raw->b, threads);
                raw->layout = layoutOf(h, *raw->b);
                check(hipSetDevice(h->device), "hipSetDevice");
                raw->image = h->images.take(raw->layout.records);
                packRecords(*raw->b, raw->layout, raw->image, threads);
            } catch (const Failure& e) {
                raw->status = e.code;
                raw->message = e.what();
            } catch (const std::exception& e) {
                raw->status = CLAY_E_INVALID;
                raw->message = e.what();
            }
            raw->done = true;
        });
        return f;
    };
    const int status = guarded(h, [&] {
        check(hipSetDevice(h->device), "hipSetDevice");
        RoundTuner tuner;
        tuner.init(h);
        const int cus = std::max(1, h->props.multiProcessorCount);
        const bool useGraph = !(h->cfg.flags & CLAY_FLAG_NO_GRAPH);
        const auto shapeOf = [&](int mixBlocks, bool simpleMix) {
            Shape fixed;
            fixed.mixBlocks = mixBlocks;
            fixed.chainBlocks = cus * 6;
            fixed.pitchBlocks = cus;
            fixed.audioBlocks = cus;
            fixed.simpleMix = simpleMix;
            return fixed;
        };
        // Both round graphs are captured before any worker thread starts: no
        // HIP call runs on another thread while a capture is open.
        if (useGraph)
            for (bool simple : {false, true})
                roundGraph(h, shapeOf(tuner.pick().second.second, simple));
        size_t next = 0;  // first queued job not yet handed to a cohort
        size_t cohorts = 0;
        bool exhausted = !refill;
        std::vector<Window> windows;
        // Rounds timed for the tuner, read back once their events complete.
        struct Measurement {
            int index;
            hipEvent_t begin, end;
            uint64_t frames;
        };
        std::deque<Measurement> measurements;
        const auto collect = [&](bool wait) {
            while (!measurements.empty()) {
                Measurement& m = measurements.front();
                if (wait)
                    check(hipEventSynchronize(m.end), "round timing");
                else if (hipEventQuery(m.end) == hipErrorNotReady)
                    return;
                float ms = 0;
                check(hipEventElapsedTime(&ms, m.begin, m.end), "hipEventElapsedTime");
                tuner.record(m.index, ms, m.frames);
                h->eventPool.push_back(m.begin);
                h->eventPool.push_back(m.end);
                measurements.pop_front();
            }
        };
        // Keeps kAhead cohorts flattening, refilling the queue when it runs
        // dry (waiting for the caller only when nothing else is in flight).
        const auto topUp = [&] {
            while (ahead.size() < kAhead) {
                if (next == queue.size()) {
                    if (exhausted)
                        return;
                    const clay_plan_view* more = nullptr;
                    size_t n = 0;
                    const int wait = resident.empty() && completing.empty() && ahead.empty();
                    if (!refill(user, wait, &more, &n) || !n) {
                        if (wait)
                            exhausted = true;
                        return;
                    }
                    queue.insert(queue.end(), more, more + n);
                }
                // The first cohorts grow from one job so that rendering starts
                // after one job's flatten.
                const size_t size = std::min(cohortJobs, size_t(1) << std::min<size_t>(cohorts, 20));
                ++cohorts;
                size_t end = std::min(queue.size(), next + size);
                if (h->cfg.group_minimum)
                    while (end < queue.size() && queue[end].volume_group &&
                           queue[end].volume_group == queue[end - 1].volume_group)
                        ++end;
                ahead.push_back(startFlatten(next, end));
                next = end;
            }
        };
        const auto residentJobs = [&] {
            size_t n = 0;
            for (auto& u : resident)
                n += u->active();
            return n;
        };
        try {
            for (;;) {
                topUp();
                // 1. Admission at the round boundary: cohorts whose records are
                // ready, while their jobs fit; wait for a flatten only when
                // nothing is resident.
                while (!ahead.empty()) {
                    Flattening& f = *ahead.front();
                    const size_t jobs = f.end - f.first;
                    if (!resident.empty() && (residentJobs() + jobs > h->cfg.batch_jobs || !f.done))
                        break;
                    f.join();
                    if (f.status != CLAY_OK) {
                        h->error = f.message;
                        throw Failure(f.status, f.message);
                    }
                    const int st = guarded(h, [&] {
                        allocateArena(h, *f.b, f.layout, true);
                        placeRecords(h, *f.b, f.layout, f.image);
                        h->images.give(f.image);
                        f.image = nullptr;
                    });
                    if (st == CLAY_E_MEMORY) {
                        release(h, *f.b);
                        if (!resident.empty() || !completing.empty())
                            break;  // retried once memory is returned
                        auto failed = std::move(ahead.front());
                        ahead.pop_front();
                        dropFlattening(*failed);
                        if (jobs > 1) {
                            // Halve the cohort (at a volume-group boundary when one exists).
                            size_t half = failed->first + jobs / 2;
                            if (h->cfg.group_minimum)
                                while (half < failed->end && queue[half].volume_group &&
                                       queue[half].volume_group == queue[half - 1].volume_group)
                                    ++half;
                            if (half == failed->end)
                                half = failed->first + jobs / 2;
                            ahead.push_front(startFlatten(half, failed->end));
                            ahead.push_front(startFlatten(failed->first, half));
                            continue;
                        }
                        clay_render_batch rejected{};
                        rejected.failed_scratch_bytes = 1;
                        if (publish(user, &rejected, failed->first))
                            throw Failure(CLAY_E_CALLBACK, "publish aborted the window");
                        continue;
                    }
                    if (st != CLAY_OK)
                        throw Failure(st, h->error);
                    auto u = std::make_unique<Cohort>();
                    u->b = f.b;
                    f.b = nullptr;
                    u->first = f.first;
                    u->jobs = jobs;
                    for (const Job& j : u->b->jobs)
                        u->ends.push_back(j.frames);
                    std::sort(u->ends.begin(), u->ends.end());
                    u->begin = h->event();
                    check(hipEventRecord(u->begin, h->stream), "hipEventRecord");
                    ++h->stats.admissions;
                    resident.push_back(std::move(u));
                    ahead.pop_front();
                    topUp();
                }
                if (resident.empty() && completing.empty() && ahead.empty() && (exhausted || next == queue.size()))
                    break;
                // 2. One round over every resident cohort.
                if (!resident.empty()) {
                    collect(false);
                    const auto picked = tuner.pick();
                    const auto shape = picked.second;
                    windows.clear();
                    uint64_t frames = 0;
                    bool simpleMix = true;
                    for (auto& u : resident) {
                        simpleMix = simpleMix && u->b->simpleMix;
                        // A straight cohort carries no state from frame to frame
                        // (no filter or pitch recurrence): all of it in one round.
                        const int first = u->cursor,
                                  end = u->b->straight ? u->b->lastFrame : std::min(u->b->lastFrame, first + shape.first);
                        int hi = 0;
                        windowChains(*u->b, first, end, u->chainLo, hi);
                        windows.push_back({u->b, first, end, u->chainLo, hi});
                        frames += uint64_t(end - first) * u->b->jobs.size();
                        u->cursor = end;
                    }
                    const Shape fixed = shapeOf(shape.second, simpleMix);
                    const bool measure = picked.first >= 0;
                    Measurement m{picked.first, nullptr, nullptr, frames};
                    if (measure) {
                        m.begin = h->event();
                        check(hipEventRecord(m.begin, h->stream), "hipEventRecord");
                    }
                    renderRoundFixed(h, windows, fixed, useGraph);
                    if (measure) {
                        m.end = h->event();
                        check(hipEventRecord(m.end, h->stream), "hipEventRecord");
                        measurements.push_back(m);
                        tuner.launched(m.index);
                    }
                    h->stats.segment_frames = shape.first;
                    h->stats.mix_blocks = shape.second;
                    // Cohorts whose last window was rendered: finish them behind
                    // this round and hand them to the download threads.
                    for (size_t i = 0; i < resident.size();) {
                        auto& u = resident[i];
                        if (u->cursor < u->b->lastFrame) {
                            ++i;
                            continue;
                        }
                        finishBatch(h, *u->b);
                        u->finished = h->event();
                        check(hipEventRecord(u->finished, h->stream), "hipEventRecord");
                        u->results = static_cast<clay_tail_result*>(h->pinned(u->jobs * sizeof(clay_tail_result)));
                        if (!downloads.started())
                            downloads.start(h);
                        downloads.push(u.get());
                        completing.push_back(std::move(u));
                        resident.erase(resident.begin() + std::ptrdiff_t(i));
                    }
                }
                // 3. Completions, oldest first; block only when nothing else can
                // run (no cohort resident, none being flattened or ready to admit).
                for (size_t i = 0; i < completing.size();) {
                    auto& u = completing[i];
                    bool downloaded;
                    {
                        std::unique_lock<std::mutex> l(downloads.lock);
                        if (!u->downloaded && resident.empty() && ahead.empty() && i == 0)
                            downloads.changed.wait(l, [&] { return u->downloaded; });
                        downloaded = u->downloaded;
                    }
                    if (!downloaded) {
                        ++i;
                        continue;
                    }
                    if (u->code != CLAY_OK)
                        throw Failure(u->code, u->message);
                    h->stats.downloaded_bytes += u->samples * sizeof(int16_t);
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, u->begin, u->finished);
                    clay_render_batch batch{};
                    batch.pcm = static_cast<const int16_t*>(u->slot);
                    batch.pcm_handle = u->slot;
                    batch.results = u->results;
                    batch.result_count = u->jobs;
                    batch.kernel_seconds = ms / 1000.0;
                    const size_t first = u->first;
                    const int aborted = publish(user, &batch, first);
                    drop(*u);
                    completing.erase(completing.begin() + std::ptrdiff_t(i));
                    if (aborted)
                        throw Failure(CLAY_E_CALLBACK, "publish aborted the window");
                }
            }
        } catch (...) {
            (void)hipStreamSynchronize(h->stream);
            for (auto& m : measurements) {
                h->eventPool.push_back(m.begin);
                h->eventPool.push_back(m.end);
            }
            throw;
        }
        collect(true);
        tuner.keep(h);
    });
    downloads.finish();
    if (status != CLAY_OK)
        (void)hipStreamSynchronize(h->stream);
    for (auto& f : ahead)
        dropFlattening(*f);
    if (status != CLAY_OK) {
        for (auto& u : resident)
            drop(*u);
        for (auto& u : completing) {
            if (u->slot)
                (void)clay_hip_release_pcm(h, u->slot);
            drop(*u);
        }
    }
    // Idle arenas beyond 16 GiB go back to the device; the rest serve the
    // next window.
    h->arenas.trim(size_t(16) << 30);
    return status;
}

} // namespace

extern "C" {

int clay_hip_render_window(clay_hip* h, const clay_plan_view* plans, size_t count, clay_publish_fn publish,
                           clay_refill_fn refill, void* user)
{
    if (!h || !publish || (count && !plans))
        return CLAY_E_INVALID;
    if (h->cfg.form == CLAY_FORM_PERSISTENT)
        return renderWindowPersistent(h, plans, count, publish, refill, user);
    return renderWindowPlain(h, plans, count, publish, refill, user);
}

int clay_hip_release_pcm(clay_hip* h, void* pcm_handle)
{
    if (!h)
        return CLAY_E_INVALID;
    {
        std::lock_guard<std::mutex> guard(h->slotLock);
        for (auto& slot : h->slots)
            if (slot.host == pcm_handle) {
                slot.busy = false;
                return CLAY_OK;
            }
    }
    h->error = "unknown PCM handle";
    return CLAY_E_INVALID;
}

int clay_hip_get_stats(clay_hip* h, clay_hip_stats* out)
{
    if (!h || !out)
        return CLAY_E_INVALID;
    *out = h->stats;
    out->chain_waves = h->tuner.chainWaves.chosen;
    out->audio_unroll = h->tuner.audioUnroll.chosen;
    out->launch_tuning = h->tuner.chainWaves.tuning || h->tuner.audioUnroll.tuning;
    return CLAY_OK;
}

const char* clay_hip_last_error(const clay_hip* h) { return h ? h->error.c_str() : g_createError.c_str(); }

} // extern "C"

// Accessors for the predictor's host code (predictor.hip).
hipStream_t clayStreamOf(clay_hip* h) { return h->stream; }
int clayDeviceOf(clay_hip* h) { return h->device; }
void claySetError(clay_hip* h, const char* message) { h->error = message; }
