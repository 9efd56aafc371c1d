// hostpool.hip -- see hostpool.hpp.  Workers take chunk indices from an
// atomic counter; the caller takes chunks too and then waits for the
// workers that joined the job.
#include "hostpool.hpp"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace lsdtt_gpu {

namespace {

class Pool {
public:
    Pool()
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int workers = static_cast<int>(std::min(hw, 16u)) - 1;
        for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { work(); });
    }
    // never destroyed: the pool lives until the process ends
    Pool(const Pool&) = delete;
    Pool& operator=(const Pool&) = delete;

    void run(long long n, long long grain, const std::function<void(long long, long long)>& fn)
    {
        std::lock_guard<std::mutex> one(call_);
        const long long max_chunks = static_cast<long long>(threads_.size()) + 1;
        const long long chunks = std::max(1ll, std::min(max_chunks * 4, n / std::max(1ll, grain)));
        if (chunks < 2 || threads_.empty()) {
            fn(0, n);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            fn_ = &fn;
            n_ = n;
            chunks_ = chunks;
            next_.store(0);
            active_ = static_cast<int>(threads_.size());
            ++job_;
        }
        cv_.notify_all();
        take_chunks();
        std::unique_lock<std::mutex> g(m_);
        done_cv_.wait(g, [this] { return active_ == 0; });
        fn_ = nullptr;
    }

private:
    void take_chunks()
    {
        for (;;) {
            const long long c = next_.fetch_add(1);
            if (c >= chunks_) return;
            const long long b = n_ * c / chunks_, e = n_ * (c + 1) / chunks_;
            (*fn_)(b, e);
        }
    }
    void work()
    {
        long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return job_ != seen; });
                seen = job_;
            }
            take_chunks();
            std::lock_guard<std::mutex> g(m_);
            if (--active_ == 0) done_cv_.notify_one();
        }
    }

    std::vector<std::thread> threads_;
    std::mutex call_, m_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(long long, long long)>* fn_ = nullptr;
    long long n_ = 0, chunks_ = 0, job_ = 0;
    std::atomic<long long> next_{0};
    int active_ = 0;
};

Pool& pool()
{
    static Pool* p = new Pool();
    return *p;
}

}  // namespace

void host_parallel_for(long long n, long long grain, const std::function<void(long long, long long)>& fn)
{
    if (n <= 0) return;
    if (n < 2 * grain) {
        fn(0, n);
        return;
    }
    pool().run(n, grain, fn);
}

}  // namespace lsdtt_gpu
