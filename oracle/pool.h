// Fork/join thread pool standing in for tbb::parallel_for
// (Parallelization/CPUParallelization.h:19-28). One parallel_for call is one
// colour sweep: the calling thread joins the work and returns only after every
// index has run, which is the per-colour barrier of the reference.
// TEST INFRASTRUCTURE ONLY (CPU oracle / CPU baseline).
#pragma once
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace oracle {

class ThreadPool {
  public:
    explicit ThreadPool(int n) : n_(n < 1 ? 1 : n) {
        for (int i = 1; i < n_; ++i) workers_.emplace_back([this] { loop(); });
    }
    ~ThreadPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    int size() const { return n_; }

    template <class F> void parallel_for(int begin, int end, F&& f) {
        if (end <= begin) return;
        std::function<void(int)> fn = f;
        {
            std::lock_guard<std::mutex> lk(m_);
            fn_ = &fn;
            next_.store(begin);
            end_ = end;
            pending_ = (int)workers_.size();
            ++gen_;
        }
        cv_.notify_all();
        run_chunks();
        std::unique_lock<std::mutex> lk(m_);
        done_cv_.wait(lk, [this] { return pending_ == 0; });
        fn_ = nullptr;
    }

  private:
    void run_chunks() {
        const int chunk = 16;
        for (;;) {
            int i = next_.fetch_add(chunk);
            if (i >= end_) break;
            int e = i + chunk < end_ ? i + chunk : end_;
            for (int k = i; k < e; ++k) (*fn_)(k);
        }
    }
    void loop() {
        unsigned long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
            }
            run_chunks();
            {
                std::lock_guard<std::mutex> lk(m_);
                if (--pending_ == 0) done_cv_.notify_one();
            }
        }
    }

    int n_;
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    std::function<void(int)>* fn_ = nullptr;
    std::atomic<int> next_{0};
    int end_ = 0;
    int pending_ = 0;
    unsigned long gen_ = 0;
    bool stop_ = false;
};

}  // namespace oracle
