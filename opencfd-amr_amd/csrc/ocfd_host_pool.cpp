// Host worker threads (see ocfd_host_pool.hpp): a fixed set of threads started
// on first use; a job is an index range handed out one index at a time.
#include "ocfd_host_pool.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <mutex>
#include <thread>
#include <vector>

namespace ocfd {
namespace {

// set on the pool's workers, and on a caller while its job runs: a nested
// parallel_for from either runs serially instead of waiting for the pool
thread_local bool t_worker = false;

// spin time of an idle worker before it sleeps, and of a caller before it
// waits for the job's end: jobs come in bursts (the chops and passes of one
// clustering), and waking a sleeping thread costs tens of microseconds
constexpr auto kSpin = std::chrono::microseconds(100);

class Pool {
public:
    explicit Pool(int nthreads)
    {
        for (int t = 1; t < nthreads; ++t) workers_.emplace_back([this] { loop(); });
    }
    ~Pool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_.store(true);
        }
        cv_.notify_all();
        for (std::thread& t : workers_) t.join();
    }
    int size() const { return (int)workers_.size() + 1; }

    void run(int n, const std::function<void(int)>& fn)
    {
        std::lock_guard<std::mutex> one(job_mu_);   // one job at a time
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            left_.store(n);
            gen_.fetch_add(1);
        }
        cv_.notify_all();
        t_worker = true;
        work(fn, n);
        t_worker = false;
        // the job ends when every index ran and no worker is still inside it
        const auto until = std::chrono::steady_clock::now() + kSpin;
        while ((left_.load() != 0 || active_.load() != 0) && std::chrono::steady_clock::now() < until)
            std::this_thread::yield();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return left_.load() == 0 && active_.load() == 0; });
        fn_ = nullptr;
    }

private:
    void work(const std::function<void(int)>& fn, int n)
    {
        for (;;) {
            const int i = next_.fetch_add(1);
            if (i >= n) return;
            fn(i);
            if (left_.fetch_sub(1) == 1) {
                std::lock_guard<std::mutex> lk(mu_);
                done_cv_.notify_all();
            }
        }
    }
    void loop()
    {
        t_worker = true;
        unsigned long seen = 0;
        for (;;) {
            // a new job is looked for by spinning first, then by sleeping;
            // it is taken under the lock, so the job's function and size
            // are read together with the worker's entry into it
            const auto until = std::chrono::steady_clock::now() + kSpin;
            while (gen_.load() == seen && !stop_.load() && std::chrono::steady_clock::now() < until)
                std::this_thread::yield();
            const std::function<void(int)>* fn = nullptr;
            int n = 0;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_.load() || gen_.load() != seen; });
                if (stop_.load()) return;
                seen = gen_.load();
                if (!fn_) continue;
                fn = fn_;
                n = n_;
                active_.fetch_add(1);
            }
            work(*fn, n);
            std::lock_guard<std::mutex> lk(mu_);
            active_.fetch_sub(1);
            done_cv_.notify_all();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex mu_, job_mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)>* fn_ = nullptr;
    int n_ = 0;
    std::atomic<int> next_{0}, left_{0}, active_{0};
    std::atomic<unsigned long> gen_{0};
    std::atomic<bool> stop_{false};
};

int configured_threads()
{
    if (const char* e = std::getenv("OCFD_HOST_THREADS")) {
        const int v = std::atoi(e);
        if (v >= 1) return std::min(v, 64);
    }
    const int hw = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(8, hw));
}

Pool& pool()
{
    static Pool p(configured_threads());
    return p;
}

}  // namespace

int host_threads() { return t_worker ? 1 : pool().size(); }

void parallel_for(int n, const std::function<void(int)>& fn)
{
    if (n <= 0) return;
    if (n == 1 || t_worker || pool().size() == 1) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    pool().run(n, fn);
}

}  // namespace ocfd
