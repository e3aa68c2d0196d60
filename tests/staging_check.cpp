// CPU model of the copy queue behind clay_staging::Image and ::Pool
// (claysampler_amd/csrc/hip/staging.h). Every copy is queued with a stamp
// written into its whole buffer; the model's "device" executes queued copies
// in order, at random times or when the host waits, and checks that each
// copy still reads its own stamp. A buffer rewritten before its copy ran, or
// freed under a queued copy, fails the check.
//
// usage: staging_check image|pool|planted SEED COUNT
//   pool: up to four producers hold buffers at once (threads packing records)
//   and queue their copies in random order.
//   planted: a pool that hands a buffer out again as soon as it comes back,
//   before its copy has run (a plausible mistake); the check must catch it.
#include "../claysampler_amd/csrc/hip/staging.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <random>
#include <string>
#include <vector>

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
    bool ran(int fence) const { return queue.empty() || queue.front().fence > fence; }
};

Model* g_model = nullptr;

struct ModelFences {
    using Fence = int;
    Fence record() { return g_model->queue.empty() ? g_model->nextFence - 1 : g_model->queue.back().fence; }
    void wait(Fence& f) { g_model->waitFor(f); }
    bool done(Fence& f) { return g_model->ran(f); }
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

// The mistake: a returned buffer is free again at once, its copy still queued.
class PlantedPool {
public:
    char* take(size_t bytes)
    {
        for (auto& b : buffers_)
            if (!b.busy && b.capacity >= bytes) {
                b.busy = true;
                return b.host;
            }
        buffers_.push_back({fences_.allocate(bytes), bytes, true});
        return buffers_.back().host;
    }
    void give(char* host)
    {
        for (auto& b : buffers_)
            if (b.host == host)
                b.busy = false;
    }
    void release()
    {
        g_model->waitFor(g_model->nextFence);
        for (auto& b : buffers_)
            fences_.free(b.host);
        buffers_.clear();
    }

private:
    struct Buffer {
        char* host;
        size_t capacity;
        bool busy;
    };
    ModelFences fences_;
    std::vector<Buffer> buffers_;
};

void queueCopy(char* p, size_t bytes, uint8_t stamp)
{
    // The copy is queued behind any earlier work; the fence marks it.
    g_model->queue.push_back({g_model->nextFence++, p, bytes, stamp});
}

void runImage(long count)
{
    Model& m = *g_model;
    clay_staging::Image<ModelFences> image;
    for (long k = 0; k < count; ++k) {
        const size_t bytes = 1 + m.rng() % 100000;
        char* p = image.take(bytes);
        const uint8_t stamp = uint8_t(1 + k % 251);
        std::memset(p, stamp, bytes);
        queueCopy(p, bytes, stamp);
        image.commit();
        m.progress();
    }
    image.release();
    m.waitFor(m.nextFence);
}

template<class P>
void runPool(P& pool, long count)
{
    Model& m = *g_model;
    struct Held {
        char* p;
        size_t bytes;
        uint8_t stamp;
    };
    std::vector<Held> held;
    for (long k = 0; k < count;) {
        if (held.size() < 4 && (held.empty() || m.rng() % 2)) {
            // A producer packs a new image (sizes spanning several powers of two).
            const size_t bytes = 1 + m.rng() % (size_t(3) << 20);
            const uint8_t stamp = uint8_t(1 + k % 251);
            char* p = pool.take(bytes);
            std::memset(p, stamp, bytes);
            held.push_back({p, bytes, stamp});
            ++k;
        } else {
            // One of the packed images is copied, in random order.
            const size_t i = m.rng() % held.size();
            queueCopy(held[i].p, held[i].bytes, held[i].stamp);
            pool.give(held[i].p);
            held.erase(held.begin() + std::ptrdiff_t(i));
        }
        m.progress();
    }
    for (auto& h : held) {
        queueCopy(h.p, h.bytes, h.stamp);
        pool.give(h.p);
    }
    pool.release();
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
    if (kind == "image") {
        runImage(count);
    } else if (kind == "pool") {
        clay_staging::Pool<ModelFences> pool{ModelFences{}, size_t(24) << 20};
        runPool(pool, count);
    } else if (kind == "planted") {
        PlantedPool pool;
        runPool(pool, count);
    } else {
        return 2;
    }
    std::printf("%s executed %ld failures %ld\n", kind.c_str(), model.executed, model.failures);
    return 0;
}
