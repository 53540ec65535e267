// gdp_copy_pool.h - host threads that split a large memcpy between them.
//
// The staging of bulk host<->device copies (gdp_engine.hip) fills and drains
// pinned buffers with several threads at once.  Plain C++, no HIP: the
// threads' lifecycle is tested on the CPU.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace gdp {

class CopyPool {
  public:
    static constexpr size_t kBlock = 256u << 10;  // blocks dealt out in turn

    ~CopyPool() { stop(); }

    // Copies `bytes` from src to dst on the calling thread and up to
    // max_threads - 1 pool threads (started on first use).  One copy at a
    // time: callers serialise.  Every copy is a new job generation, and each
    // pool thread takes each generation once.
    void copy(void *dst, const void *src, size_t bytes, int max_threads) {
        start(max_threads);
        int parts;
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<char *>(dst);
            src_ = static_cast<const char *>(src);
            bytes_ = bytes;
            parts_ = parts = static_cast<int>(workers_.size()) + 1;
            left_ = parts - 1;
            gen_++;
        }
        job_cv_.notify_all();
        copy_part(static_cast<char *>(dst), static_cast<const char *>(src), bytes, 0, parts);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return left_ == 0; });
    }

    // Joins the pool threads (idle between copies).
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        job_cv_.notify_all();
        for (auto &w : workers_) w.join();
        workers_.clear();
    }

    int threads() const { return static_cast<int>(workers_.size()); }

  private:
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable job_cv_, done_cv_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0;
    int parts_ = 1;
    int left_ = 0;
    unsigned long long gen_ = 0;
    bool quit_ = false;

    static void copy_part(char *d, const char *s, size_t bytes, int part, int parts) {
        const size_t blocks = (bytes + kBlock - 1) / kBlock;
        for (size_t b = static_cast<size_t>(part); b < blocks; b += static_cast<size_t>(parts)) {
            const size_t off = b * kBlock;
            std::memcpy(d + off, s + off, std::min(kBlock, bytes - off));
        }
    }

    // Threads start at the current generation: a job issued before they
    // existed is not theirs (its count was complete without them).
    void start(int max_threads) {
        if (!workers_.empty()) return;
        const int n = std::max(0, max_threads - 1);
        unsigned long long gen;
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = false;
            gen = gen_;
        }
        for (int t = 0; t < n; t++) workers_.emplace_back([this, t, gen] { work(t + 1, gen); });
    }

    void work(int part, unsigned long long seen) {
        for (;;) {
            char *d;
            const char *s;
            size_t bytes;
            int parts;
            {
                std::unique_lock<std::mutex> lk(mu_);
                job_cv_.wait(lk, [&] { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
                d = dst_;
                s = src_;
                bytes = bytes_;
                parts = parts_;
            }
            copy_part(d, s, bytes, part, parts);
            std::lock_guard<std::mutex> lk(mu_);
            if (--left_ == 0) done_cv_.notify_all();
        }
    }
};

}  // namespace gdp
