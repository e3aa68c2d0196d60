// Pinned host staging for asynchronous host-to-device copies: which bytes may
// be rewritten while copies out of them are still queued.
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
//   char* allocate(size_t);            a pinned buffer
//   void free(char*);
#pragma once

#include <algorithm>
#include <cstddef>
#include <deque>

namespace clay_staging {

// A ring of regions, one per queued copy (the persistent form admits units
// while earlier uploads are still queued behind render rounds). A region is
// handed out only after every earlier region overlapping it has been read:
// after wrap-arounds those need not be the oldest ones.
template<class Policy>
class Ring {
public:
    explicit Ring(Policy policy = Policy()) : policy_(policy) {}
    Ring(const Ring&) = delete;
    Ring& operator=(const Ring&) = delete;
    ~Ring() { release(); }

    // Space for the next copy, of at least `bytes` bytes.
    char* take(size_t bytes)
    {
        bytes = std::max<size_t>(bytes, 1);
        if (bytes > capacity_) {
            drain();
            if (base_)
                policy_.free(base_);
            base_ = nullptr;
            capacity_ = std::max(bytes, minimum_);
            base_ = policy_.allocate(capacity_);
            head_ = 0;
        }
        if (head_ + bytes > capacity_)
            head_ = 0;
        const size_t begin = head_, end = head_ + bytes;
        for (auto it = uses_.begin(); it != uses_.end();) {
            if (it->begin < end && begin < it->end) {
                policy_.wait(it->fence);
                it = uses_.erase(it);
            } else {
                ++it;
            }
        }
        pendingBegin_ = begin;
        pendingEnd_ = end;
        head_ = end;
        return base_ + begin;
    }

    // The copy out of the region last taken has been queued.
    void commit() { uses_.push_back({pendingBegin_, pendingEnd_, policy_.record()}); }

    // Waits for every queued copy and frees the buffer.
    void release()
    {
        drain();
        if (base_)
            policy_.free(base_);
        base_ = nullptr;
        capacity_ = head_ = 0;
    }

    void setMinimum(size_t bytes) { minimum_ = bytes; }
    size_t capacity() const { return capacity_; }

private:
    struct Use {
        size_t begin, end;
        typename Policy::Fence fence;
    };
    void drain()
    {
        for (auto& u : uses_)
            policy_.wait(u.fence);
        uses_.clear();
    }
    Policy policy_;
    char* base_ = nullptr;
    size_t capacity_ = 0, head_ = 0, minimum_ = size_t(64) << 20;
    size_t pendingBegin_ = 0, pendingEnd_ = 0;
    std::deque<Use> uses_;
};

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

} // namespace clay_staging
