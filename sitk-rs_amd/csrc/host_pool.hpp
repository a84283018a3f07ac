// A fixed set of host threads for the per-pair or per-start optimizer
// updates between batched evaluations. Host code only.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace sitkhip {

// run(n, f) calls f(i) for every i < n across the threads (the
// calling one included) and returns when all are done, rethrowing the first
// exception. SITKHIP_HOST_THREADS caps the count (default: up to 16).
class HostPool {
public:
    explicit HostPool(int want) {
        int t = static_cast<int>(std::thread::hardware_concurrency());
        if (const char *e = std::getenv("SITKHIP_HOST_THREADS")) t = std::atoi(e);
        t = std::max(1, std::min({t, 16, want}));
        for (int i = 1; i < t; ++i) th_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            quit_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    template <typename F>
    void run(int n, F f) {
        std::function<void(int)> job = f;
        {
            std::lock_guard<std::mutex> g(mu_);
            job_ = &job;
            n_ = n;
            next_ = 0;
            busy_ = static_cast<int>(th_.size());
            err_ = nullptr;
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [this] { return busy_ == 0; });
        job_ = nullptr;
        if (err_) std::rethrow_exception(err_);
    }

private:
    void work() {
        for (;;) {
            const int i = next_.fetch_add(1);
            if (i >= n_) break;
            try {
                (*job_)(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(mu_);
                if (!err_) err_ = std::current_exception();
            }
        }
    }
    void loop() {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
            }
            work();
            std::lock_guard<std::mutex> g(mu_);
            if (--busy_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::function<void(int)> *job_ = nullptr;
    std::atomic<int> next_{0};
    int n_ = 0, busy_ = 0;
    bool quit_ = false;
    unsigned long long gen_ = 0;
    std::exception_ptr err_;
};

}  // namespace sitkhip
