// Exercises the host copy pool of libminirecast_hip (csrc/mr_hostpool.cpp) under the
// sanitizers: random sizes and chunk counts, the bytes copied, every chunk handed on once
// and in order with its true offset and length, a callback error returned while the
// remaining chunks are still copied, many copies in a row on one pool (generations), and
// pools of 0 and 5 workers.  Exit status 0: every check held.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "mr_hostpool.h"

using namespace mr;

static int failures = 0;
#define CHECK(cond, ...)                                             \
    do {                                                             \
        if (!(cond)) {                                               \
            ++failures;                                              \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);                       \
            std::fprintf(stderr, "\n");                              \
        }                                                            \
    } while (0)

int main() {
    std::mt19937_64 rng(7);
    std::uniform_int_distribution<size_t> size(1, 3u << 20);
    std::uniform_int_distribution<int> chunks(1, 12);
    for (int workers : {0, 3, 5}) {
        HostCopyPool pool(workers);
        CHECK(pool.workers() == workers, "worker count");
        for (int r = 0; r < 60; ++r) {
            const size_t n = r % 7 == 0 ? 1 : size(rng);
            std::vector<unsigned char> src(n), dst(n, 0);
            for (size_t i = 0; i < n; ++i) src[i] = static_cast<unsigned char>(rng());
            const int nch = chunks(rng);
            std::vector<std::pair<size_t, size_t>> seen;
            const int fail_at = r % 5 == 0 ? static_cast<int>(rng() % 3) : -1;
            const int rc = pool.copy(dst.data(), src.data(), n, nch, [&](size_t o, size_t len) {
                // a chunk handed on is complete
                CHECK(std::memcmp(dst.data() + o, src.data() + o, len) == 0, "chunk not copied yet");
                seen.push_back({o, len});
                return static_cast<int>(seen.size()) - 1 == fail_at ? 7 : 0;
            });
            CHECK(std::memcmp(dst.data(), src.data(), n) == 0, "bytes differ (run %d)", r);
            size_t next = 0;
            const size_t want = fail_at >= 0 ? std::min<size_t>(fail_at + 1, seen.size()) : seen.size();
            for (size_t i = 0; i < want && i < seen.size(); ++i) {
                CHECK(seen[i].first == next && seen[i].second > 0, "chunk %zu out of order", i);
                next += seen[i].second;
            }
            if (fail_at < 0) {
                CHECK(next == n, "chunks cover %zu of %zu bytes", next, n);
                CHECK(rc == 0, "error without a failing callback");
                CHECK(static_cast<int>(seen.size()) <= nch, "more chunks than asked");
            } else if (fail_at < static_cast<int>(seen.size())) {
                CHECK(rc == 7, "callback error not returned");
                CHECK(static_cast<int>(seen.size()) == fail_at + 1, "chunks handed on after an error");
            }
        }
    }
    if (failures) std::fprintf(stderr, "%d check(s) failed\n", failures);
    else std::printf("hostpool_check: all checks passed\n");
    return failures ? 1 : 0;
}
