// Pinned host staging for asynchronous host-to-device copies: which bytes may
// be rewritten while copies out of them are still queued. Pool serves the
// persistent form, whose records are packed on worker threads; Image the
// plain form.
//
// Host-only and independent of HIP: the fence type, its wait and the buffer
// allocation are a policy, so the rules run on the CPU against a model of the
// copy queue (tests/staging_check.cpp). The render core's policy records a
// hipEvent behind each copy and waits with hipEventSynchronize.
//
// Policy interface:
//   using Fence = ...;                 a mark behind one queued copy
//   Fence record();                    a fence behind the copy just queued
//   void wait(Fence&);                 returns once that copy has read its bytes; releases the fence
//   bool done(Fence&);                 (Pool) whether it has, without waiting; releases the fence if so
//   char* allocate(size_t);            a pinned buffer
//   void free(char*);
#pragma once

#include <algorithm>
#include <cstddef>
#include <mutex>
#include <vector>

namespace clay_staging {

// One buffer reused by consecutive copies (the plain form's record image):
// rewritten or freed only after the copy that last read it.
template<class Policy>
class Image {
public:
    explicit Image(Policy policy = Policy()) : policy_(policy) {}
    Image(const Image&) = delete;
    Image& operator=(const Image&) = delete;
    ~Image() { release(); }

    char* take(size_t bytes)
    {
        bytes = std::max<size_t>(bytes, 1);
        if (pending_) {
            policy_.wait(fence_);
            pending_ = false;
        }
        if (bytes > capacity_) {
            if (base_)
                policy_.free(base_);
            base_ = nullptr;
            capacity_ = 0;
            base_ = policy_.allocate(bytes);
            capacity_ = bytes;
        }
        return base_;
    }

    void commit()
    {
        fence_ = policy_.record();
        pending_ = true;
    }

    void release()
    {
        if (pending_)
            policy_.wait(fence_);
        pending_ = false;
        if (base_)
            policy_.free(base_);
        base_ = nullptr;
        capacity_ = 0;
    }

private:
    Policy policy_;
    char* base_ = nullptr;
    size_t capacity_ = 0;
    typename Policy::Fence fence_{};
    bool pending_ = false;
};

// Buffers for several producers (threads that pack records while earlier
// copies are queued). A buffer comes back with the fence of the copy that
// reads it and is handed out again only once that copy has completed: a
// producer never waits for the device unless the buffers would exceed
// `budget` bytes. New buffers are powers of two, at least 1 MiB.
template<class Policy>
class Pool {
public:
    explicit Pool(Policy policy = Policy(), size_t budget = size_t(4) << 30) : policy_(policy), budget_(budget) {}
    Pool(const Pool&) = delete;
    Pool& operator=(const Pool&) = delete;
    ~Pool() { release(); }

    // A buffer of at least `bytes` bytes, for the caller alone until give().
    char* take(size_t bytes)
    {
        std::lock_guard<std::mutex> lock(mutex_);
        for (;;) {
            Buffer* best = nullptr;
            size_t held = 0;
            for (auto& b : buffers_) {
                held += b.capacity;
                if (b.fenced && policy_.done(b.fence))
                    b.fenced = false;
                if (!b.busy && !b.fenced && b.capacity >= bytes && (!best || b.capacity < best->capacity))
                    best = &b;
            }
            if (best) {
                best->busy = true;
                return best->host;
            }
            size_t capacity = size_t(1) << 20;
            while (capacity < bytes)
                capacity *= 2;
            if (held + capacity > budget_) {
                // Over budget: wait for the oldest fenced buffers, free the idle ones that are too small.
                bool waited = false;
                for (auto& b : buffers_)
                    if (b.fenced) {
                        policy_.wait(b.fence);
                        b.fenced = false;
                        waited = true;
                    }
                for (auto it = buffers_.begin(); it != buffers_.end();)
                    if (!it->busy && it->capacity < bytes) {
                        policy_.free(it->host);
                        it = buffers_.erase(it);
                    } else {
                        ++it;
                    }
                if (waited)
                    continue;
            }
            buffers_.push_back({policy_.allocate(capacity), capacity, true, false, {}});
            return buffers_.back().host;
        }
    }

    // The copy out of `host` has been queued (on the queue the policy records on).
    void give(char* host)
    {
        std::lock_guard<std::mutex> lock(mutex_);
        for (auto& b : buffers_)
            if (b.host == host) {
                b.busy = false;
                b.fenced = true;
                b.fence = policy_.record();
                return;
            }
    }

    // A buffer taken but never copied from.
    void discard(char* host)
    {
        std::lock_guard<std::mutex> lock(mutex_);
        for (auto& b : buffers_)
            if (b.host == host)
                b.busy = false;
    }

    // Waits for every fenced copy and frees every buffer (none may be taken).
    void release()
    {
        std::lock_guard<std::mutex> lock(mutex_);
        for (auto& b : buffers_) {
            if (b.fenced)
                policy_.wait(b.fence);
            policy_.free(b.host);
        }
        buffers_.clear();
    }

private:
    struct Buffer {
        char* host;
        size_t capacity;
        bool busy, fenced;
        typename Policy::Fence fence;
    };
    Policy policy_;
    size_t budget_;
    std::mutex mutex_;
    std::vector<Buffer> buffers_;
};

} // namespace clay_staging
