// Host worker threads for large memory copies (no HIP calls; used by the host
// runtime to fill and empty the pinned staging buffers).
#pragma once

#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>

namespace mmsk {

// Host threads that split large copies between pageable caller memory and the
// pinned bounce buffers. One thread copies about 31 GB/s on an MI355X host,
// below the 53 GB/s of a pinned DMA; four reach about 64 GB/s. The calling
// thread copies one part itself; concurrent callers take turns. The workers
// are detached and wait on a condition variable for the life of the process,
// so a pool must never be destroyed (the runtime creates one on first use and
// keeps it). A child process of a fork has no workers and copies alone.
class CopyWorkers {
  public:
    explicit CopyWorkers(int workers) : pid_(getpid()) {
        int started = 0;
        try {
            for (int w = 0; w < workers; ++w, ++started) std::thread([this, w] { loop(w + 1); }).detach();
        } catch (...) {
        }
        parts_ = started + 1;
    }

    void copy(void* dst, const void* src, size_t n) {
        if (n < kParallelMin || parts_ == 1 || getpid() != pid_) {
            std::memcpy(dst, src, n);
            return;
        }
        std::lock_guard<std::mutex> call(call_mu_);  // one copy at a time
        {
            std::lock_guard<std::mutex> g(mu_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            n_ = n;
            pending_ = parts_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> l(mu_);
        done_.wait(l, [&] { return pending_ == 0; });
    }

  private:
    static constexpr size_t kParallelMin = 1u << 20;

    // Slice i of the current copy (4 KiB-aligned boundaries).
    void part(int i) const {
        // ceil(n / parts) rounded up to 4 KiB: the slices cover all n bytes
        const size_t per = ((n_ + parts_ - 1) / parts_ + 4095) & ~size_t(4095);
        const size_t b = std::min(n_, per * (size_t)i), e = std::min(n_, b + per);
        if (e > b) std::memcpy(dst_ + b, src_ + b, e - b);
    }

    void loop(int idx) {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(mu_);
                cv_.wait(l, [&] { return gen_ != seen; });
                seen = gen_;
            }
            part(idx);
            std::lock_guard<std::mutex> g(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }

    const pid_t pid_;
    int parts_ = 1;
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    unsigned long long gen_ = 0;
    int pending_ = 0;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t n_ = 0;
};

}  // namespace mmsk
