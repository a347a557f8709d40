// Stress check of the engine's host planning pool (lc_pool.h): every item of
// every job runs exactly once, exceptions reach the caller, jobs that start
// while the workers sleep, spin or are still leaving the previous job, and
// nested use from a second thread (which runs its loop inline).
#include <atomic>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

#include "lc_pool.h"

using lcpool::parallel_for;

static int check_jobs(int njob, unsigned seed) {
    std::mt19937 rng(seed);
    for (int it = 0; it < njob; it++) {
        const size_t n = 1 + rng() % 5000;
        std::vector<std::atomic<int>> hit(n);
        for (auto &h : hit) h.store(0);
        const bool slow = it % 61 == 0;
        if (it % 400 == 0) std::this_thread::sleep_for(std::chrono::microseconds(300 + rng() % 1500));
        parallel_for(n, [&](size_t i) {
            hit[i].fetch_add(1);
            if (slow && i % 97 == 0) std::this_thread::sleep_for(std::chrono::microseconds(30));
        });
        for (size_t i = 0; i < n; i++)
            if (hit[i].load() != 1) {
                std::printf("FAIL: job %d item %zu ran %d times\n", it, i, hit[i].load());
                return 1;
            }
    }
    return 0;
}

int main() {
    if (check_jobs(20000, 1)) return 1;
    int caught = 0;
    for (int it = 0; it < 300; it++) {
        try {
            parallel_for(2000, [&](size_t i) {
                if (i == (size_t)(it * 7 % 2000)) throw std::runtime_error("item");
            });
        } catch (const std::runtime_error &) {
            caught++;
        }
    }
    if (caught != 300) {
        std::printf("FAIL: %d of 300 exceptions reached the caller\n", caught);
        return 1;
    }
    // a second thread using the pool while the first does: the job that
    // finds it busy runs inline, both complete exactly
    int r1 = 0, r2 = 0;
    std::thread a([&] { r1 = check_jobs(3000, 2); });
    std::thread b([&] { r2 = check_jobs(3000, 3); });
    a.join();
    b.join();
    if (r1 || r2) return 1;
    std::printf("pool ok\n");
    return 0;
}
