// Host-side gather of per-target rows into a pinned staging buffer (no
// device code).  A batch upload copies thousands of separate host arrays
// (each target's flux, ivar and resolution diagonals) into one contiguous
// page-locked buffer before the DMA; Python-level copies of that many small
// arrays contend for the interpreter lock beyond 2 threads, so the copies run
// here on plain threads.
#include <cstring>
#include <thread>
#include <vector>

#include "common.hpp"

extern "C" int rr_gather_rows(const unsigned long long *src, long long n, long long row_bytes, void *dst,
                              int nthreads) {
    if (n < 0 || row_bytes < 0 || (n > 0 && (!src || !dst))) {
        rr::set_error("rr_gather_rows: bad arguments (n %lld, row_bytes %lld)", n, row_bytes);
        return RR_EINVAL;
    }
    if (n == 0 || row_bytes == 0) return RR_OK;
    for (long long i = 0; i < n; ++i)
        if (!src[i]) {
            rr::set_error("rr_gather_rows: row %lld has no data", i);
            return RR_EINVAL;
        }
    int nt = nthreads < 1 ? 1 : (nthreads > 64 ? 64 : nthreads);
    if (nt > n) nt = static_cast<int>(n);
    char *out = static_cast<char *>(dst);
    auto copy = [&](long long a, long long b) {
        for (long long i = a; i < b; ++i)
            std::memcpy(out + i * row_bytes, reinterpret_cast<const void *>(src[i]), row_bytes);
    };
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (int k = 1; k < nt; ++k) pool.emplace_back(copy, n * k / nt, n * (k + 1) / nt);
    copy(0, n / nt);
    for (auto &t : pool) t.join();
    return RR_OK;
}
