// CPU model of the copy queue behind clay_staging::Ring and ::Image
// (claysampler_amd/csrc/hip/staging.h). Every copy is queued with a stamp
// written into its whole region; the model's "device" executes queued copies
// in order, at random times or when the host waits, and checks that each
// copy still reads its own stamp. A region rewritten before its copy ran, or
// a buffer freed under a queued copy, fails the check.
//
// usage: staging_check ring|image|planted SEED COUNT
//   planted: a ring that waits only for the oldest overlapping region (a
//   plausible mistake); the check must catch it.
#include "../claysampler_amd/csrc/hip/staging.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <random>
#include <map>
#include <string>

namespace {

struct Copy {
    int fence;
    const char* region;
    size_t bytes;
    uint8_t stamp;
};

struct Model {
    std::deque<Copy> queue;  // queued copies, in stream order
    std::map<const char*, size_t> live;  // allocated buffers
    int nextFence = 0;
    long failures = 0, executed = 0;
    std::mt19937_64 rng;

    void execute(const Copy& c)
    {
        ++executed;
        for (size_t i = 0; i < c.bytes; ++i)
            if (uint8_t(c.region[i]) != c.stamp) {
                ++failures;
                return;
            }
    }
    void progress()  // the device runs a random number of queued copies
    {
        size_t n = rng() % 3;
        while (n-- && !queue.empty()) {
            execute(queue.front());
            queue.pop_front();
        }
    }
    void waitFor(int fence)
    {
        while (!queue.empty() && queue.front().fence <= fence) {
            execute(queue.front());
            queue.pop_front();
        }
    }
};

Model* g_model = nullptr;

struct ModelFences {
    using Fence = int;
    Fence record() { return g_model->queue.empty() ? g_model->nextFence - 1 : g_model->queue.back().fence; }
    void wait(Fence& f) { g_model->waitFor(f); }
    char* allocate(size_t bytes)
    {
        char* p = static_cast<char*>(std::malloc(bytes));
        g_model->live[p] = bytes;
        return p;
    }
    void free(char* p)
    {
        // Freeing under a queued copy out of this buffer is a failure (the
        // copy is dropped rather than run on freed memory).
        const size_t size = g_model->live[p];
        auto& q = g_model->queue;
        for (auto it = q.begin(); it != q.end();) {
            if (it->region >= p && it->region < p + size) {
                ++g_model->failures;
                it = q.erase(it);
            } else {
                ++it;
            }
        }
        g_model->live.erase(p);
        std::free(p);
    }
};

// The mistake the ring once had: after a wrap-around, wait only for the
// oldest region that overlaps the new one.
class PlantedRing {
public:
    char* take(size_t bytes)
    {
        if (bytes > capacity_) {
            for (auto& u : uses_)
                fences_.wait(u.fence);
            uses_.clear();
            if (base_)
                fences_.free(base_);
            capacity_ = std::max<size_t>(bytes, 1 << 16);
            base_ = fences_.allocate(capacity_);
            head_ = 0;
        }
        if (head_ + bytes > capacity_)
            head_ = 0;
        const size_t begin = head_, end = head_ + bytes;
        for (auto it = uses_.begin(); it != uses_.end(); ++it)
            if (it->begin < end && begin < it->end) {
                fences_.wait(it->fence);
                uses_.erase(it);
                break;
            }
        pending_ = {begin, end, 0};
        head_ = end;
        return base_ + begin;
    }
    void commit()
    {
        pending_.fence = fences_.record();
        uses_.push_back(pending_);
    }
    void release()
    {
        for (auto& u : uses_)
            fences_.wait(u.fence);
        uses_.clear();
        if (base_)
            fences_.free(base_);
        base_ = nullptr;
    }

private:
    struct Use {
        size_t begin, end;
        int fence;
    };
    ModelFences fences_;
    char* base_ = nullptr;
    size_t capacity_ = 0, head_ = 0;
    Use pending_{};
    std::deque<Use> uses_;
};

template<class Buffer>
void run(Buffer& buffer, long count, size_t maxBytes)
{
    Model& m = *g_model;
    for (long k = 0; k < count; ++k) {
        const size_t bytes = 1 + m.rng() % maxBytes;
        char* p = buffer.take(bytes);
        const uint8_t stamp = uint8_t(1 + k % 251);
        std::memset(p, stamp, bytes);
        // The copy is queued behind any earlier work; the fence marks it.
        m.queue.push_back({m.nextFence++, p, bytes, stamp});
        buffer.commit();
        m.progress();
    }
    buffer.release();
    m.waitFor(m.nextFence);
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 4)
        return 2;
    const std::string kind = argv[1];
    Model model;
    model.rng.seed(std::strtoull(argv[2], nullptr, 10));
    g_model = &model;
    const long count = std::strtol(argv[3], nullptr, 10);
    if (kind == "ring") {
        clay_staging::Ring<ModelFences> ring;
        ring.setMinimum(1 << 16);
        run(ring, count, 24000);
    } else if (kind == "image") {
        clay_staging::Image<ModelFences> image;
        run(image, count, 100000);
    } else if (kind == "planted") {
        PlantedRing ring;
        run(ring, count, 24000);
    } else {
        return 2;
    }
    std::printf("%s executed %ld failures %ld\n", kind.c_str(), model.executed, model.failures);
    return 0;
}
