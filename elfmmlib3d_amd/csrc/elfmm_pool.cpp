// Host worker pool for the planner: a fixed set of threads created on first
// use, one parallel region at a time. std::thread rather than OpenMP so the
// library brings no second OpenMP runtime into processes (PyTorch, the
// Fortran runtime) that already carry one.
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "elfmm_internal.h"

namespace elfmm {
namespace {

class Pool {
  public:
    explicit Pool(int n) : nthreads_(n)
    {
        for (int t = 1; t < n; t++) workers_.emplace_back([this, t] { loop(t); });
    }
    // The pool lives until process exit; workers are detached so that exit
    // does not wait on them.
    ~Pool()
    {
        for (auto &w : workers_) w.detach();
    }
    int size() const { return nthreads_; }

    void run(int nchunks, const std::function<void(int, int)> &fn)
    {
        std::lock_guard<std::mutex> region(region_mu_);
        if (nthreads_ == 1 || nchunks <= 1) {
            for (int c = 0; c < nchunks; c++) fn(c, 0);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            nchunks_ = nchunks;
            next_.store(0);
            busy_ = nthreads_ - 1;
            gen_++;
        }
        cv_.notify_all();
        work(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return busy_ == 0; });
        fn_ = nullptr;
    }

  private:
    void work(int tid)
    {
        for (;;) {
            int c = next_.fetch_add(1);
            if (c >= nchunks_) break;
            (*fn_)(c, tid);
        }
    }
    void loop(int tid)
    {
        unsigned long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
            }
            work(tid);
            std::lock_guard<std::mutex> lk(mu_);
            if (--busy_ == 0) done_cv_.notify_one();
        }
    }

    int nthreads_;
    std::vector<std::thread> workers_;
    std::mutex region_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int, int)> *fn_ = nullptr;
    int nchunks_ = 0, busy_ = 0;
    unsigned long gen_ = 0;
    std::atomic<int> next_{0};
};

// CPUs this process may use: its affinity mask, capped by a cgroup v2 CPU
// quota (cpu.max "quota period") when one is set.
int cpus_available()
{
    int n = (int)std::thread::hardware_concurrency();
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    if (FILE *f = std::fopen("/sys/fs/cgroup/cpu.max", "r")) {
        char q[32] = {0};
        long period = 0;
        if (std::fscanf(f, "%31s %ld", q, &period) == 2 && period > 0 && std::strcmp(q, "max") != 0) {
            long quota = std::atol(q);
            if (quota > 0) n = std::min(n, (int)((quota + period - 1) / period));
        }
        std::fclose(f);
    }
    return std::max(1, n);
}

// Default size: half the available CPUs, at most 8. The rest stay free for
// the GPU runtime's threads and for the requests in flight beside the one
// being planned; with every CPU of a quota in the pool, served throughput
// fell by a quarter (CFS throttling of the whole process).
Pool &pool()
{
    static Pool *p = [] {
        int n = std::min(8, std::max(1, cpus_available() / 2));
        if (const char *e = std::getenv("ELFMM_HOST_THREADS")) n = std::atoi(e);
        n = std::max(1, std::min(n, 16));
        return new Pool(n);
    }();
    return *p;
}

}  // namespace

int host_threads() { return pool().size(); }

void parallel_chunks(int nchunks, const std::function<void(int, int)> &fn) { pool().run(nchunks, fn); }

}  // namespace elfmm
