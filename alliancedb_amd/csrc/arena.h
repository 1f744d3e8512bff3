// arena.h -- window memory of the serving loop: a first-fit allocator and the
// admission controller that makes requests wait for it.  Host-side
// bookkeeping only (offsets and lengths in tuples of one device buffer); no
// HIP dependency, so the admission logic is testable on a CPU.
//
// Free space is a map offset -> length of disjoint, non-adjacent free ranges;
// release() merges a range with its free neighbours, so a fully released
// arena is one range again.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <iterator>
#include <map>
#include <mutex>

namespace aj {

class Arena {
public:
    explicit Arena(uint64_t capacity = 0) { reset(capacity); }

    void reset(uint64_t capacity) {
        cap_ = capacity;
        used_ = 0;
        peak_ = 0;
        free_.clear();
        if (capacity) free_[0] = capacity;
    }

    // First fit: the lowest free range of at least n tuples.  n = 0 takes no
    // space and always succeeds (offset 0).
    bool alloc(uint64_t n, uint64_t *off) {
        if (n == 0) {
            *off = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n, roff = it->first + n;
            free_.erase(it);
            if (rest) free_[roff] = rest;
            used_ += n;
            if (used_ > peak_) peak_ = used_;
            return true;
        }
        return false;
    }

    // Returns [off, off+n) to the free space.  The range must have come from
    // alloc() and not have been released since; returns false (and changes
    // nothing) for a range that overlaps free space or lies outside.
    bool release(uint64_t off, uint64_t n) {
        if (n == 0) return true;
        if (off + n > cap_ || off + n < off) return false;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first < off + n) return false;
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second > off) return false;
        }
        uint64_t b = off, len = n;
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                b = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && next->first == off + n) {
            len += next->second;
            free_.erase(next);
        }
        free_[b] = len;
        used_ -= n;
        return true;
    }

    uint64_t capacity() const { return cap_; }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }
    uint64_t largest_free() const {
        uint64_t m = 0;
        for (const auto &f : free_)
            if (f.second > m) m = f.second;
        return m;
    }
    size_t free_ranges() const { return free_.size(); }

private:
    uint64_t cap_ = 0, used_ = 0, peak_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

// Admission by memory budget: acquire() reserves two ranges (a window's R and
// S tuples, both or neither) and waits while they do not fit; release()
// returns a range and wakes the waiters.  Requests queue, they do not fail,
// unless a wait limit is given or the request can never fit.
class Admission {
public:
    enum { OK = 0, TIMEOUT = 1, NEVER = 2, CLOSED = 3 };

    explicit Admission(uint64_t capacity = 0) : arena_(capacity) {}

    void reset(uint64_t capacity) {
        std::lock_guard<std::mutex> lk(mu_);
        arena_.reset(capacity);
        closed_ = false;
        waits_ = 0;
    }

    // timeout_us < 0: wait without limit; 0: do not wait.
    int acquire(uint64_t a, uint64_t b, int64_t timeout_us, uint64_t *oa, uint64_t *ob) {
        std::unique_lock<std::mutex> lk(mu_);
        if (a + b > arena_.capacity() || a + b < a) return NEVER;
        const auto deadline = std::chrono::steady_clock::now() +
                              std::chrono::microseconds(timeout_us > 0 ? timeout_us : 0);
        bool waited = false;
        for (;;) {
            if (closed_) return CLOSED;
            if (try_both(a, b, oa, ob)) {
                if (waited) ++waits_;
                return OK;
            }
            if (timeout_us == 0) return TIMEOUT;
            waited = true;
            if (timeout_us < 0) {
                cv_.wait(lk);
            } else if (cv_.wait_until(lk, deadline) == std::cv_status::timeout) {
                if (!closed_ && try_both(a, b, oa, ob)) {
                    ++waits_;
                    return OK;
                }
                return closed_ ? CLOSED : TIMEOUT;
            }
        }
    }

    bool release(uint64_t off, uint64_t n) {
        bool ok;
        {
            std::lock_guard<std::mutex> lk(mu_);
            ok = arena_.release(off, n);
        }
        cv_.notify_all();
        return ok;
    }

    // Wakes every waiter; later acquires return CLOSED.
    void close() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            closed_ = true;
        }
        cv_.notify_all();
    }

    uint64_t capacity() {
        std::lock_guard<std::mutex> lk(mu_);
        return arena_.capacity();
    }
    uint64_t used() {
        std::lock_guard<std::mutex> lk(mu_);
        return arena_.used();
    }
    uint64_t peak() {
        std::lock_guard<std::mutex> lk(mu_);
        return arena_.peak();
    }
    uint64_t waits() {
        std::lock_guard<std::mutex> lk(mu_);
        return waits_;
    }
    size_t free_ranges() {
        std::lock_guard<std::mutex> lk(mu_);
        return arena_.free_ranges();
    }

private:
    bool try_both(uint64_t a, uint64_t b, uint64_t *oa, uint64_t *ob) {
        if (!arena_.alloc(a, oa)) return false;
        if (arena_.alloc(b, ob)) return true;
        arena_.release(*oa, a);
        return false;
    }

    std::mutex mu_;
    std::condition_variable cv_;
    Arena arena_;
    bool closed_ = false;
    uint64_t waits_ = 0;
};

}  // namespace aj
