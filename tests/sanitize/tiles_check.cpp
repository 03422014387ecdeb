// Exercises the host-only request preparation of libminirecast_hip (csrc/mr_tiles.cpp)
// under AddressSanitizer and UndefinedBehaviorSanitizer: task validation (including
// offsets that overflow int64), tile enumeration against a brute-force box test, the
// height order, and the request cache against fresh enumerations over a sequence of
// more distinct requests than the cache holds.  Exit status 0: every check held.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "mr_tiles.h"

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

static mr_task random_task(std::mt19937_64& rng, int64_t* out_len) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::uniform_int_distribution<int> size(1, 100), planes(1, 5), coin(0, 1);
    mr_task T;
    std::memset(&T, 0, sizeof(T));
    const bool volume = coin(rng);
    T.nq = volume ? planes(rng) : 1;
    T.np = size(rng);
    T.ns = size(rng);
    T.mask_box = volume ? 0 : 1;
    for (int d = 0; d < 3; ++d) {
        T.o[d] = 12.0 * u(rng);
        T.a[d] = 0.3 * u(rng);
        T.b[d] = 0.3 * u(rng);
        T.c[d] = volume ? 0.2 * u(rng) : 0.0;
        T.box[d] = 2.0 + 6.0 * std::fabs(u(rng));
    }
    T.out_offset = *out_len;
    T.stride_q = int64_t(T.np) * T.ns;
    T.stride_p = T.ns;
    T.stride_s = 1;
    *out_len += int64_t(T.nq) * T.np * T.ns;
    return T;
}

static bool point_in_box(const mr_task& T, int q, int p, int s) {
    for (int d = 0; d < 3; ++d) {
        const double x = T.o[d] + q * T.c[d] + p * T.a[d] + s * T.b[d];
        if (std::fabs(x) > T.box[d]) return false;
    }
    return true;
}

static void check_request(const std::vector<mr_task>& tasks, int64_t out_len) {
    const int nt = int(tasks.size());
    CHECK(validate_tasks_msg(tasks.data(), nt, out_len).empty(), "valid request refused");
    CHECK(!validate_tasks_msg(tasks.data(), nt, out_len - 1).empty(), "overlong request accepted");
    std::vector<Tile> tiles, empty;
    build_tiles(tasks.data(), nt, tiles, &empty);
    // every candidate tile exactly once, in tiles or in empty
    std::vector<std::vector<int>> seen(nt);
    for (int i = 0; i < nt; ++i) {
        const mr_task& T = tasks[i];
        seen[i].assign(size_t(T.nq) * ((T.np + BP_TP - 1) / BP_TP) * ((T.ns + BP_TS - 1) / BP_TS), 0);
    }
    auto mark = [&](const Tile& t) {
        const mr_task& T = tasks[t.task];
        const int ntp = (T.np + BP_TP - 1) / BP_TP, nts = (T.ns + BP_TS - 1) / BP_TS;
        CHECK(t.task >= 0 && t.task < nt && t.q >= 0 && t.q < T.nq && t.tp >= 0 && t.tp < ntp &&
                  t.ts >= 0 && t.ts < nts, "tile out of its task");
        seen[t.task][(size_t(t.q) * ntp + t.tp) * nts + t.ts] += 1;
    };
    for (const Tile& t : tiles) mark(t);
    for (const Tile& t : empty) mark(t);
    for (int i = 0; i < nt; ++i)
        for (int v : seen[i]) CHECK(v == 1, "candidate tile covered %d times", v);
    // culled tiles hold no in-box point (the test is conservative)
    for (const Tile& t : empty) {
        const mr_task& T = tasks[t.task];
        for (int p = t.tp * BP_TP; p < std::min(T.np, (t.tp + 1) * BP_TP); ++p)
            for (int s = t.ts * BP_TS; s < std::min(T.ns, (t.ts + 1) * BP_TS); ++s)
                CHECK(!point_in_box(T, t.q, p, s), "culled tile has an in-box point");
    }
    // height order: a permutation of the tiles, non-decreasing centre height
    std::vector<Tile> ordered = tiles;
    order_tiles_by_height(tasks.data(), ordered);
    CHECK(ordered.size() == tiles.size(), "order changed the tile count");
    double prev = -1e300;
    for (const Tile& t : ordered) {
        const mr_task& T = tasks[t.task];
        const double p = t.tp * BP_TP + 0.5 * (BP_TP - 1), s = t.ts * BP_TS + 0.5 * (BP_TS - 1);
        const double z = T.o[2] + t.q * T.c[2] + p * T.a[2] + s * T.b[2];
        CHECK(z >= prev, "height order violated");
        prev = z;
    }
}

int main() {
    std::mt19937_64 rng(20241204);
    // validation corner cases
    {
        int64_t len = 0;
        mr_task T = random_task(rng, &len);
        CHECK(validate_tasks_msg(&T, 1, len).empty(), "last element exactly at out_len - 1");
        mr_task bad = T;
        bad.stride_p = -1;
        CHECK(!validate_tasks_msg(&bad, 1, len).empty(), "negative stride accepted");
        bad = T;
        bad.np = -3;
        CHECK(!validate_tasks_msg(&bad, 1, len).empty(), "negative size accepted");
        bad = T;
        bad.stride_q = INT64_MAX / 2;
        bad.nq = 5;
        CHECK(!validate_tasks_msg(&bad, 1, INT64_MAX).empty(), "overflowing offsets accepted");
        bad = T;
        bad.out_offset = INT64_MAX - 2;
        CHECK(!validate_tasks_msg(&bad, 1, INT64_MAX).empty(), "overflowing offset accepted");
        bad = T;
        bad.nq = 0;
        bad.stride_s = -5;                         // empty tasks are skipped
        CHECK(validate_tasks_msg(&bad, 1, 0).empty(), "empty task refused");
    }
    // requests: enumeration properties, then the cache against fresh enumerations
    std::vector<std::vector<mr_task>> reqs;
    std::vector<int64_t> lens;
    for (int r = 0; r < 7; ++r) {
        int64_t len = 0;
        std::vector<mr_task> tasks;
        const int nt = 1 + int(rng() % 4);
        for (int i = 0; i < nt; ++i) tasks.push_back(random_task(rng, &len));
        check_request(tasks, len);
        reqs.push_back(tasks);
        lens.push_back(len);
    }
    std::vector<TileCache> cache;
    for (int k = 0; k < 60; ++k) {
        const int r = int(rng() % reqs.size());
        const int flags = int(rng() % 4);
        const bool keep_empty = (flags & 1) == 0, by_height = (flags & 2) != 0;
        const TileCache& got = cached_tiles(cache, reqs[r].data(), int(reqs[r].size()), flags,
                                            keep_empty, by_height);
        std::vector<Tile> want, empty;
        build_tiles(reqs[r].data(), int(reqs[r].size()), want, keep_empty ? &empty : nullptr);
        if (by_height) order_tiles_by_height(reqs[r].data(), want);
        const size_t ne = empty.size();
        want.insert(want.end(), empty.begin(), empty.end());
        CHECK(got.ne == ne && got.tiles.size() == want.size(), "cache entry size (request %d)", r);
        for (size_t i = 0; i < want.size() && i < got.tiles.size(); ++i)
            CHECK(std::memcmp(&got.tiles[i], &want[i], sizeof(Tile)) == 0, "cache entry tile %zu", i);
        CHECK(cache.size() <= kTileCacheEntries, "cache grew beyond its entries");
        CHECK(&got == &cache.front(), "entry not moved to the front");
    }
    {
        // the same planes at another shift and output place: the same entry, not a new one
        std::vector<mr_task> a = reqs[0], b = reqs[0];
        for (mr_task& T : b) {
            T.cor_shift += 0.75;
            T.out_offset += 12345;
            T.stride_p += 1;
        }
        const TileCache& ea = cached_tiles(cache, a.data(), int(a.size()), 0, true, true);
        const std::vector<Tile> tiles_a = ea.tiles;
        const size_t n_before = cache.size();
        const TileCache& eb = cached_tiles(cache, b.data(), int(b.size()), 0, true, true);
        CHECK(&eb == &cache.front() && cache.size() == n_before, "shifted request missed the cache");
        CHECK(eb.tiles.size() == tiles_a.size() &&
                  std::memcmp(eb.tiles.data(), tiles_a.data(), sizeof(Tile) * tiles_a.size()) == 0,
              "shifted request changed the tiles");
    }
    // an empty request
    const TileCache& none = cached_tiles(cache, nullptr, 0, 0, true, true);
    CHECK(none.tiles.empty() && none.ne == 0, "empty request has tiles");

    // shift groups of rotation-centre searches: random axial planes, each with shifts on a
    // grid of 1/q px (random gaps), every task in exactly one group slot at its shift,
    // every tile of every group once, culled tiles without in-box points
    std::uniform_int_distribution<int> nset(1, 3), qpick(0, 2), span(1, 60), keep(0, 3);
    for (int r = 0; r < 200; ++r) {
        std::vector<mr_task> tasks;
        int64_t off = 0;
        int qmax = 1;
        const int ns_ = nset(rng);
        for (int g = 0; g < ns_; ++g) {
            int64_t dummy = 0;
            mr_task base = random_task(rng, &dummy);
            base.nq = 1;
            base.a[2] = 0.0;
            base.b[2] = 0.0;
            base.c[0] = base.c[1] = base.c[2] = 0.0;
            const int q = 1 << qpick(rng);
            qmax = std::max(qmax, q);
            const double s_lo = std::uniform_int_distribution<int>(-40, 40)(rng) / 4.0;
            const int m = span(rng);
            for (int k = 0; k < m; ++k) {
                if (k > 0 && keep(rng) == 0) continue;       // a gap
                mr_task T = base;
                T.cor_shift = s_lo + static_cast<double>(k) / q;
                T.out_offset = off;
                off += static_cast<int64_t>(T.np) * T.ns;
                tasks.push_back(T);
            }
        }
        // shuffle the task order
        std::shuffle(tasks.begin(), tasks.end(), rng);
        CorPlan P;
        const bool ok = cor_shift_groups(tasks.data(), static_cast<int>(tasks.size()), 1.0, P);
        if (tasks.size() < 2) {
            CHECK(!ok, "a single task grouped");
            continue;
        }
        CHECK(ok, "grid shifts not grouped (request %d)", r);
        if (!ok) continue;
        CHECK(P.q >= 1 && P.q <= qmax, "grid %d beyond the finest used %d", P.q, qmax);
        std::vector<int> used(tasks.size(), 0);
        for (const CorGroup& G : P.groups) {
            CHECK(G.set >= 0 && G.set < static_cast<int>(P.sets.size()), "group set index");
            CHECK(G.ns >= 1 && G.ns <= 6 * P.q + 1, "group size %d for q %d", G.ns, P.q);
            const mr_task& B = tasks[P.sets[G.set].task];
            for (int k = 0; k < kCorMaxShifts; ++k) {
                if (k >= G.ns) {
                    CHECK(G.out[k] == -1, "slot past the group size used");
                    continue;
                }
                if (G.out[k] < 0) continue;
                int hit = -1;
                for (size_t i = 0; i < tasks.size(); ++i)
                    if (tasks[i].out_offset == G.out[k]) hit = static_cast<int>(i);
                CHECK(hit >= 0, "group slot names no task");
                if (hit < 0) continue;
                ++used[hit];
                const mr_task& T = tasks[hit];
                CHECK(std::fabs(T.cor_shift - (G.s0 + static_cast<double>(k) / P.q)) < 1e-12,
                      "slot %d holds shift %g, group s0 %g q %d", k, T.cor_shift, G.s0, P.q);
                mr_task a = T, b = B;
                a.cor_shift = b.cor_shift = 0.0;
                a.out_offset = b.out_offset = 0;
                CHECK(std::memcmp(&a, &b, sizeof(a)) == 0, "task grouped with another plane");
            }
        }
        for (size_t i = 0; i < tasks.size(); ++i) CHECK(used[i] == 1, "task %zu in %d slots", i, used[i]);
        std::vector<std::vector<int>> cover(P.groups.size());
        for (size_t gi = 0; gi < P.groups.size(); ++gi) {
            const mr_task& T = tasks[P.sets[P.groups[gi].set].task];
            cover[gi].assign(static_cast<size_t>(((T.np + BP_TP - 1) / BP_TP) * ((T.ns + BP_TS - 1) / BP_TS)), 0);
        }
        size_t zeros = 0;
        for (size_t i = 0; i < P.tiles.size(); ++i) {
            const CorTile& t = P.tiles[i];
            CHECK(t.group >= 0 && t.group < static_cast<int>(P.groups.size()), "tile group index");
            const mr_task& T = tasks[P.sets[P.groups[t.group].set].task];
            const int nts = (T.ns + BP_TS - 1) / BP_TS;
            cover[t.group][static_cast<size_t>(t.tp * nts + t.ts)]++;
            CHECK((t.zero != 0) == (i >= P.tiles.size() - P.ne), "culled tiles not at the end");
            if (t.zero) {
                ++zeros;
                for (int p = t.tp * BP_TP; p < std::min((t.tp + 1) * BP_TP, T.np); ++p)
                    for (int s = t.ts * BP_TS; s < std::min((t.ts + 1) * BP_TS, T.ns); ++s)
                        CHECK(!point_in_box(T, 0, p, s), "culled group tile has an in-box point");
            }
        }
        CHECK(zeros == P.ne, "culled tile count");
        for (const auto& cv : cover)
            for (int v : cv) CHECK(v == 1, "group tile covered %d times", v);
    }
    {
        // refusals: off-grid shifts, a shift twice, an oblique or volume task, too few per group
        int64_t dummy = 0;
        mr_task B = random_task(rng, &dummy);
        B.nq = 1;
        B.a[2] = B.b[2] = 0.0;
        std::vector<mr_task> t(5, B);
        const double grid[5] = {0.0, 0.25, 0.5, 0.75, 1.0};
        for (int i = 0; i < 5; ++i) {
            t[i].cor_shift = grid[i];
            t[i].out_offset = i * 1000000;
        }
        CorPlan P;
        CHECK(cor_shift_groups(t.data(), 5, 4.0, P) && P.q == 4 && P.groups.size() == 1, "grid refused");
        std::vector<mr_task> u = t;
        u[2].cor_shift = 0.3;
        CHECK(!cor_shift_groups(u.data(), 5, 1.0, P) && P.groups.empty(), "off-grid shift grouped");
        u = t;
        u[2].cor_shift = 0.25;
        CHECK(!cor_shift_groups(u.data(), 5, 1.0, P), "a shift twice grouped");
        u = t;
        u[3].a[2] = 0.01;
        CHECK(!cor_shift_groups(u.data(), 5, 1.0, P), "oblique plane grouped");
        u = t;
        u[0].nq = 2;
        CHECK(!cor_shift_groups(u.data(), 5, 1.0, P), "volume task grouped");
        CHECK(!cor_shift_groups(t.data(), 5, 6.0, P), "fewer tasks per group than asked accepted");
        u = t;
        u[4].cor_shift = 40.0;                               // 160 grid steps: a second group
        CHECK(cor_shift_groups(u.data(), 5, 1.0, P) && P.groups.size() == 2, "far shift not split");
    }
    if (failures) std::fprintf(stderr, "%d check(s) failed\n", failures);
    else std::printf("tiles_check: all checks passed\n");
    return failures ? 1 : 0;
}
