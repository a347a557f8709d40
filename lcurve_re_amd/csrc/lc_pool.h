// lc_pool.h -- the engine's host planning pool (plain C++, no HIP): one
// process-wide pool of worker threads and parallel_for over it.
#pragma once

#include <immintrin.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace lcpool {

// Host planning workers: one process-wide pool of up to 15 threads.
// A job runs f(0..n-1) on the pool and the calling thread; one job at a
// time, so a call that finds the pool busy (another context planning on
// another thread) runs its loop on its own thread instead.
class HostPool {
  public:
    static HostPool &get() {
        static HostPool pool;
        return pool;
    }
    // Runs f(0..n-1) on the caller and the workers.  Items are claimed
    // through a counter tagged with the job's generation, so a worker that
    // wakes after the job has ended (or after the next one has started)
    // cannot claim an item of it; the caller waits only for workers inside
    // the claim loop, not for every worker to wake.
    bool try_run(size_t n, const std::function<void(size_t)> &f) {
        // a forked child has the pool object but not its threads
        if (workers_.empty() || getpid() != pid_ || n >= (1ull << 32)) return false;
        std::unique_lock<std::mutex> job(submit_, std::try_to_lock);
        if (!job.owns_lock()) return false;
        uint64_t g;
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &f;
            n_ = n;
            err_ = nullptr;
            g = ++gen_;
            claim_.store(g << 32);
            pub_.store(g);
        }
        wake_.notify_all();
        run_items(g, &f, n);
        while (busy_.load() != 0) _mm_pause();
        std::exception_ptr e;
        {
            std::lock_guard<std::mutex> lk(mu_);
            e = err_;
            err_ = nullptr;
        }
        if (e) std::rethrow_exception(e);
        return true;
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        wake_.notify_all();
        for (auto &t : workers_) t.join();
    }

  private:
    HostPool() : pid_(getpid()) {
        const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
        const unsigned nw = std::min(16u, hc) - 1;
        for (unsigned k = 0; k < nw; k++) workers_.emplace_back([this] { loop(); });
    }
    // claims runs of about n/64 items of job g until none is left (fn is
    // called only after a successful claim, so it is never touched once job
    // g has ended)
    void run_items(uint64_t g, const std::function<void(size_t)> *fn, size_t n) {
        const uint64_t step = std::max<uint64_t>(1, n / 64);
        for (;;) {
            uint64_t c = claim_.load();
            do {
                if ((c >> 32) != g || (c & 0xffffffffull) >= n) return;
            } while (!claim_.compare_exchange_weak(c, c + step));
            const size_t i0 = (size_t)(c & 0xffffffffull), i1 = std::min<size_t>(n, i0 + step);
            try {
                for (size_t i = i0; i < i1; i++) (*fn)(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu_);
                if (!err_) err_ = std::current_exception();
            }
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            // the planning of one call runs several loops within a
            // millisecond: spin briefly for the next job before sleeping
            if (pub_.load() == seen) {
                const auto t0 = std::chrono::steady_clock::now();
                while (pub_.load() == seen && std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(200))
                    _mm_pause();
            }
            const std::function<void(size_t)> *fn;
            size_t n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                wake_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                fn = fn_;
                n = n_;
                busy_.fetch_add(1);  // under mu_, before any claim of job `seen`
            }
            run_items(seen, fn, n);
            busy_.fetch_sub(1);
        }
    }
    const pid_t pid_;
    std::mutex submit_, mu_;
    std::condition_variable wake_;
    std::vector<std::thread> workers_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0;
    std::atomic<uint64_t> claim_{0};  // (generation << 32) | next item
    std::atomic<uint64_t> pub_{0};    // generation of the last job, for spinning workers
    std::atomic<int> busy_{0};        // workers inside run_items
    std::exception_ptr err_;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Loops of fewer than min_n items run inline: waking the pool costs tens
// of microseconds, more than a few items of planning.
template <class F> void parallel_for(size_t n, F f, size_t min_n = 32) {
    if (n >= std::max<size_t>(min_n, 2)) {
        const std::function<void(size_t)> fn(std::ref(f));
        if (HostPool::get().try_run(n, fn)) return;
    }
    for (size_t i = 0; i < n; i++) f(i);
}

}  // namespace lcpool
