// Host-side row parallelism for the symbolic phase: rows 0 .. n-1 are dealt to up to 16 threads
// in chunks from a shared counter.  make() runs once per thread and returns that thread's row
// function (so per-thread scratch, such as a marker array, lives in its closure).  Results must
// not depend on which thread handles a row.  An exception in a row stops that thread; the first
// one is rethrown after all threads have joined.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace sankhya {
namespace host {

template <typename Make>
void parallel_rows(int64_t n, Make&& make, int64_t chunk = 256) {
    const int nt = std::max(1, std::min<int>(16, static_cast<int>(std::thread::hardware_concurrency())));
    std::atomic<int64_t> next{0};
    auto run = [&] {
        auto row = make();
        for (;;) {
            const int64_t b = next.fetch_add(chunk);
            if (b >= n) break;
            const int64_t e = std::min<int64_t>(n, b + chunk);
            for (int64_t i = b; i < e; ++i) row(i);
        }
    };
    if (n < 4 * chunk || nt == 1) {
        run();
        return;
    }
    std::exception_ptr first;
    std::mutex mu;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            try {
                run();
            } catch (...) {
                std::lock_guard<std::mutex> lock(mu);
                if (!first) first = std::current_exception();
                next.store(n);                      // the other threads stop at their next chunk
            }
        });
    for (auto& x : th) x.join();
    if (first) std::rethrow_exception(first);
}

}  // namespace host
}  // namespace sankhya
