// Stress test of the host copy workers (maxmulsk_amd/csrc/copy_workers.hpp):
// several caller threads copy buffers of random sizes (below and above the
// parallel threshold, not multiples of the slice alignment) through one pool
// and check every byte. Built with g++ directly (and once with ThreadSanitizer)
// by tests/test_copy_workers_cpu.py. Usage: copy_workers_stress <iterations>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "copy_workers.hpp"

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 100;
    mmsk::CopyWorkers& cw = *new mmsk::CopyWorkers(3);  // never destroyed, as in the runtime
    std::mutex bad_mu;
    long bad = 0;
    std::vector<std::thread> callers;
    for (int t = 0; t < 3; ++t) {
        callers.emplace_back([&, t] {
            std::mt19937_64 rng(1000 + t);
            for (int it = 0; it < iters; ++it) {
                // every fourth copy: a multiple of 16 KiB (above the parallel
                // threshold) plus 1..3 bytes, where slices of n / parts rounded
                // down to 4 KiB would miss the tail
                const size_t n = (it % 4 == 3) ? (size_t)(64 + rng() % 256) * 16384 + 1 + it % 3
                                               : 1 + rng() % (5u << 20);
                std::vector<unsigned char> src(n), dst(n, 0);
                for (size_t i = 0; i < n; i += 509) src[i] = (unsigned char)(rng() | 1);
                src[n - 1] = 0x5a;
                cw.copy(dst.data(), src.data(), n);
                if (std::memcmp(src.data(), dst.data(), n) != 0) {
                    std::lock_guard<std::mutex> g(bad_mu);
                    ++bad;
                }
            }
        });
    }
    for (auto& c : callers) c.join();
    std::printf("copies %d bad %ld\n", 3 * iters, bad);
    return bad == 0 ? 0 : 1;
}
