// ibim_hoststage.hpp -- the host side of the operator's staged transfers
// (ibim_hostio.hip): rows split over up to kMaxHostThreads threads and
// (de)interleaved between the reference's split real / imaginary rows and
// interleaved complex rows. Plain C++ (no HIP), so that the CPU sanitizers
// can run it.
#pragma once

#include <algorithm>
#include <cstddef>
#include <thread>
#include <vector>

namespace ibim {
namespace hoststage {

constexpr int kMaxHostThreads = 16;
constexpr size_t kBytesPerThread = size_t(4) << 20;  // below this a block is (de)interleaved by fewer threads

inline int host_threads(size_t block_bytes) {
    const unsigned hw = std::thread::hardware_concurrency();
    int t = (int)std::min<unsigned>(hw ? hw : 1, kMaxHostThreads);
    const int by_size = (int)(block_bytes / kBytesPerThread);
    return std::max(1, std::min(t, by_size));
}

// f(r) for r in [0, nr), rows split into contiguous ranges over the threads
template <class F>
void for_rows(int nr, size_t block_bytes, F f) {
    const int T = std::min(host_threads(block_bytes), nr);
    if (T <= 1) {
        for (int r = 0; r < nr; ++r) f(r);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T - 1);
    auto range = [&](int t) {
        const int a = (int)((long)nr * t / T), b = (int)((long)nr * (t + 1) / T);
        for (int r = a; r < b; ++r) f(r);
    };
    for (int t = 1; t < T; ++t) th.emplace_back(range, t);
    range(0);
    for (auto& x : th) x.join();
}

template <class T>
void interleave_row(T* d, const double* a, const double* b, int count, size_t ld) {
    for (int j = 0; j < count; ++j) {
        d[2 * (size_t)j] = (T)a[j];
        d[2 * (size_t)j + 1] = (T)b[j];
    }
    for (size_t j = count; j < ld; ++j) d[2 * j] = d[2 * j + 1] = T(0);
}

template <class T>
void split_row(const T* s, double* a, double* b, int count) {
    for (int j = 0; j < count; ++j) {
        a[j] = (double)s[2 * (size_t)j];
        b[j] = (double)s[2 * (size_t)j + 1];
    }
}


}  // namespace hoststage
}  // namespace ibim
