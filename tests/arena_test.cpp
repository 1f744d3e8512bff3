// CPU test of the serving loop's window memory (alliancedb_amd/csrc/arena.h):
// first-fit allocation against a brute-force occupancy model, coalescing, and
// admission that queues requests while the budget is taken.  Built and run by
// tests/test_admission.py (with ThreadSanitizer when available).
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "arena.h"

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "FAILED %s at line %d\n", #c, __LINE__); \
            std::exit(1);                                               \
        }                                                               \
    } while (0)

using aj::Admission;
using aj::Arena;

// first fit on a byte map: lowest offset with n free cells
static bool model_alloc(std::vector<char> &m, uint64_t n, uint64_t *off) {
    if (n == 0) {
        *off = 0;
        return true;
    }
    uint64_t run = 0;
    for (uint64_t i = 0; i < m.size(); ++i) {
        run = m[i] ? 0 : run + 1;
        if (run == n) {
            *off = i + 1 - n;
            for (uint64_t j = *off; j <= i; ++j) m[j] = 1;
            return true;
        }
    }
    return false;
}

static void test_against_model() {
    const uint64_t cap = 1000;
    Arena a(cap);
    std::vector<char> m(cap, 0);
    std::vector<std::pair<uint64_t, uint64_t>> live;
    std::mt19937_64 rng(7);
    for (int step = 0; step < 20000; ++step) {
        if (live.empty() || rng() % 3) {
            const uint64_t n = rng() % 120;
            uint64_t o1 = 0, o2 = 0;
            const bool r1 = a.alloc(n, &o1), r2 = model_alloc(m, n, &o2);
            CHECK(r1 == r2);
            if (r1) {
                CHECK(o1 == o2);
                if (n) live.push_back({o1, n});
            }
        } else {
            const size_t k = rng() % live.size();
            const auto r = live[k];
            live.erase(live.begin() + k);
            CHECK(a.release(r.first, r.second));
            for (uint64_t j = r.first; j < r.first + r.second; ++j) m[j] = 0;
        }
        uint64_t used = 0;
        for (char c : m) used += c;
        CHECK(a.used() == used);
        CHECK(a.peak() >= used);
    }
    for (const auto &r : live) CHECK(a.release(r.first, r.second));
    CHECK(a.used() == 0);
    CHECK(a.free_ranges() == 1);
    CHECK(a.largest_free() == cap);
}

static void test_bad_release() {
    Arena a(100);
    uint64_t o = 0;
    CHECK(a.alloc(10, &o) && o == 0);
    CHECK(!a.release(50, 10));   // free space
    CHECK(!a.release(95, 10));   // past the end
    CHECK(!a.release(5, 10));    // overlaps free space
    CHECK(a.release(0, 10));
    CHECK(!a.release(0, 10));    // released twice
    CHECK(a.used() == 0 && a.free_ranges() == 1);
}

static void test_admission_queues() {
    Admission adm(100);
    uint64_t a = 0, b = 0;
    CHECK(adm.acquire(40, 40, 0, &a, &b) == Admission::OK);
    uint64_t c = 0, d = 0;
    // 20 left: a 30-tuple request cannot go in without waiting
    CHECK(adm.acquire(30, 0, 0, &c, &d) == Admission::TIMEOUT);
    // can never fit
    CHECK(adm.acquire(60, 50, -1, &c, &d) == Admission::NEVER);
    // a bounded wait times out
    const auto t0 = std::chrono::steady_clock::now();
    CHECK(adm.acquire(30, 0, 20000, &c, &d) == Admission::TIMEOUT);
    CHECK(std::chrono::steady_clock::now() - t0 >= std::chrono::milliseconds(15));
    // an unbounded wait is served once memory is returned
    std::atomic<int> got{-1};
    uint64_t e = 0, f = 0;
    std::thread waiter([&] { got = adm.acquire(30, 10, -1, &e, &f); });
    std::this_thread::sleep_for(std::chrono::milliseconds(30));
    CHECK(got.load() == -1);  // still queued
    CHECK(adm.release(a, 40));
    waiter.join();
    CHECK(got.load() == Admission::OK);
    CHECK(adm.waits() >= 1);
    CHECK(adm.used() == 80);
    // close() wakes a waiter, which then fails instead of hanging
    std::atomic<int> got2{-1};
    std::thread waiter2([&] {
        uint64_t g = 0, h = 0;
        got2 = adm.acquire(50, 0, -1, &g, &h);
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    adm.close();
    waiter2.join();
    CHECK(got2.load() == Admission::CLOSED);
}

static void test_admission_many_threads() {
    // 8 producers take and return windows of random sizes; the budget is never
    // exceeded and every request is eventually served.
    Admission adm(1000);
    std::atomic<uint64_t> inside{0}, served{0};
    std::atomic<bool> over{false};
    std::vector<std::thread> th;
    for (int t = 0; t < 8; ++t) {
        th.emplace_back([&, t] {
            std::mt19937_64 rng(100 + t);
            for (int i = 0; i < 300; ++i) {
                const uint64_t a = 1 + rng() % 200, b = 1 + rng() % 200;
                uint64_t oa = 0, ob = 0;
                if (adm.acquire(a, b, -1, &oa, &ob) != Admission::OK) {
                    over = true;
                    return;
                }
                if (inside.fetch_add(a + b) + a + b > 1000) over = true;
                std::this_thread::yield();
                inside.fetch_sub(a + b);
                if (!adm.release(oa, a) || !adm.release(ob, b)) over = true;
                ++served;
            }
        });
    }
    for (auto &x : th) x.join();
    CHECK(!over.load());
    CHECK(served.load() == 8 * 300);
    CHECK(adm.used() == 0);
    CHECK(adm.free_ranges() == 1);
}

int main() {
    test_against_model();
    test_bad_release();
    test_admission_queues();
    test_admission_many_threads();
    std::printf("arena_test OK\n");
    return 0;
}
