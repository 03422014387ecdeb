// Host-only request preparation (see mr_tiles.h).
#include "mr_tiles.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>

namespace mr {

namespace {
// a + b * c for b >= 0, c >= 0, false when it would leave [0, 2^63)
bool mul_add_ok(int64_t a, int64_t b, int64_t c, int64_t* out) {
    if (b != 0 && c > (INT64_MAX - a) / b) return false;
    *out = a + b * c;
    return true;
}
}  // namespace

std::string validate_tasks_msg(const mr_task* tasks, int ntasks, int64_t out_len) {
    for (int i = 0; i < ntasks; ++i) {
        const mr_task& T = tasks[i];
        if (T.nq < 0 || T.np < 0 || T.ns < 0)
            return "task " + std::to_string(i) + ": negative grid size";
        if (T.nq == 0 || T.np == 0 || T.ns == 0) continue;
        if (T.out_offset < 0 || T.stride_q < 0 || T.stride_p < 0 || T.stride_s < 0)
            return "task " + std::to_string(i) + ": negative offset or stride";
        int64_t last = T.out_offset;
        if (!mul_add_ok(last, T.nq - 1, T.stride_q, &last) ||
            !mul_add_ok(last, T.np - 1, T.stride_p, &last) ||
            !mul_add_ok(last, T.ns - 1, T.stride_s, &last))
            return "task " + std::to_string(i) + ": offsets overflow";
        if (last >= out_len)
            return "task " + std::to_string(i) + " writes beyond the output (" +
                   std::to_string(last) + " >= " + std::to_string(out_len) + ")";
    }
    return std::string();
}

bool tile_may_hit_box(const mr_task& T, int q, int p0, int p1, int s0, int s1) {
    if (!T.mask_box) return true;
    for (int d = 0; d < 3; ++d) {
        const double lim = T.box[d] * (1.0 + 1e-9) + 1e-9;
        bool all_hi = true, all_lo = true;
        for (int cp = 0; cp < 2; ++cp)
            for (int cs = 0; cs < 2; ++cs) {
                const int p = cp ? p1 : p0, s = cs ? s1 : s0;
                const double x = T.o[d] + q * T.c[d] + p * T.a[d] + s * T.b[d];
                all_hi = all_hi && (x > lim);
                all_lo = all_lo && (x < -lim);
            }
        if (all_hi || all_lo) return false;
    }
    return true;
}

void build_tiles(const mr_task* tasks, int ntasks, std::vector<Tile>& tiles,
                 std::vector<Tile>* empty) {
    tiles.clear();
    if (empty) empty->clear();
    for (int i = 0; i < ntasks; ++i) {
        const mr_task& T = tasks[i];
        if (T.nq <= 0 || T.np <= 0 || T.ns <= 0) continue;
        const int ntp = (T.np + BP_TP - 1) / BP_TP, nts = (T.ns + BP_TS - 1) / BP_TS;
        for (int q = 0; q < T.nq; ++q)
            for (int tp = 0; tp < ntp; ++tp)
                for (int ts = 0; ts < nts; ++ts) {
                    const int p0 = tp * BP_TP, s0 = ts * BP_TS;
                    const int p1 = std::min(p0 + BP_TP, T.np) - 1;
                    const int s1 = std::min(s0 + BP_TS, T.ns) - 1;
                    if (tile_may_hit_box(T, q, p0, p1, s0, s1))
                        tiles.push_back({i, q, tp, ts});
                    else if (empty)
                        empty->push_back({i, q, tp, ts});
                }
    }
}

// With the XCD-contiguous block map each XCD takes a band of heights, i.e. (for a
// circular orbit about z) a band of detector rows in every view, so a 128-byte detector
// line is fetched into few of the eight L2 caches instead of all of them.
void order_tiles_by_height(const mr_task* tasks, std::vector<Tile>& tiles) {
    std::vector<std::pair<double, int>> key(tiles.size());
    for (size_t i = 0; i < tiles.size(); ++i) {
        const Tile& t = tiles[i];
        const mr_task& T = tasks[t.task];
        const double p = t.tp * BP_TP + 0.5 * (BP_TP - 1), s = t.ts * BP_TS + 0.5 * (BP_TS - 1);
        key[i] = {T.o[2] + t.q * T.c[2] + p * T.a[2] + s * T.b[2], static_cast<int>(i)};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                         return a.first < b.first;
                     });
    std::vector<Tile> out(tiles.size());
    for (size_t i = 0; i < tiles.size(); ++i) out[i] = tiles[key[i].second];
    tiles.swap(out);
}

const TileCache& cached_tiles(std::vector<TileCache>& cache, const mr_task* tasks, int ntasks,
                              int flags, bool keep_empty, bool by_height) {
    const size_t tb = sizeof(mr_task) * static_cast<size_t>(ntasks > 0 ? ntasks : 0);
    // the key: the task bytes without the fields the tiles do not depend on (the shift,
    // the output offset and strides), so that requests of the same planes at another
    // rotation-centre shift or output place reuse the entry
    std::vector<mr_task> keyed(tasks, tasks + (ntasks > 0 ? ntasks : 0));
    for (mr_task& T : keyed) {
        T.cor_shift = 0.0;
        T.out_offset = 0;
        T.stride_q = T.stride_p = T.stride_s = 0;
    }
    const uint8_t* tk = reinterpret_cast<const uint8_t*>(keyed.data());
    int hit = -1;
    for (int i = 0; i < static_cast<int>(cache.size()) && hit < 0; ++i)
        if (cache[i].flags == flags && cache[i].tasks.size() == tb &&
            (tb == 0 || std::memcmp(cache[i].tasks.data(), tk, tb) == 0))
            hit = i;
    if (hit < 0) {
        TileCache e;
        std::vector<Tile> empty;
        build_tiles(tasks, ntasks, e.tiles, keep_empty ? &empty : nullptr);
        if (by_height) order_tiles_by_height(tasks, e.tiles);
        e.ne = empty.size();
        e.tiles.insert(e.tiles.end(), empty.begin(), empty.end());
        if (tb) e.tasks.assign(tk, tk + tb);
        e.flags = flags;
        if (cache.size() >= kTileCacheEntries) cache.pop_back();
        cache.insert(cache.begin(), std::move(e));
    } else if (hit > 0) {
        std::rotate(cache.begin(), cache.begin() + hit, cache.begin() + hit + 1);
    }
    return cache.front();
}

void build_zcol_tiles(const mr_task* tasks, int ntasks, int zq, std::vector<Tile>& tiles) {
    tiles.clear();
    for (int i = 0; i < ntasks; ++i) {
        const mr_task& T = tasks[i];
        for (int q = 0; q < T.nq; q += zq)
            for (int tp = 0; tp * ZT < T.np; ++tp)
                for (int ts = 0; ts * ZT < T.ns; ++ts) tiles.push_back(Tile{i, q, tp, ts});
    }
}

bool cor_shift_groups(const mr_task* tasks, int ntasks, double min_avg, CorPlan& plan) {
    plan = CorPlan{};
    if (ntasks < 2 || !tasks) return false;
    // sets: the task bytes with the shift and the output offset cleared
    std::map<std::vector<uint8_t>, int> key_set;
    std::vector<std::vector<int>> members;
    for (int i = 0; i < ntasks; ++i) {
        const mr_task& T = tasks[i];
        if (T.nq != 1 || T.np < 1 || T.ns < 1 || T.a[2] != 0.0 || T.b[2] != 0.0 ||
            !std::isfinite(T.cor_shift))
            return false;
        mr_task K = T;
        K.cor_shift = 0.0;
        K.out_offset = 0;
        const uint8_t* kb = reinterpret_cast<const uint8_t*>(&K);
        std::vector<uint8_t> key(kb, kb + sizeof(K));
        auto it = key_set.find(key);
        if (it == key_set.end()) {
            it = key_set.emplace(std::move(key), static_cast<int>(members.size())).first;
            members.emplace_back();
        }
        members[it->second].push_back(i);
    }
    // per set: shifts in increasing order and their grid steps m for each q
    std::vector<std::vector<std::pair<double, int>>> sorted(members.size());
    for (size_t g = 0; g < members.size(); ++g) {
        for (int i : members[g]) sorted[g].push_back({tasks[i].cor_shift, i});
        std::sort(sorted[g].begin(), sorted[g].end());
    }
    auto steps = [&](size_t g, int q, std::vector<int64_t>& m) {
        m.clear();
        const double s_min = sorted[g].front().first;
        for (const auto& e : sorted[g]) {
            const double x = (e.first - s_min) * q;
            const double r = std::nearbyint(x);
            if (std::fabs(x - r) > 1e-9 * std::max(1.0, std::fabs(x)) || r > 1e9) return false;
            const int64_t mi = static_cast<int64_t>(r);
            if (!m.empty() && mi <= m.back()) return false;      // the same shift twice
            m.push_back(mi);
        }
        return true;
    };
    int q = 0;
    std::vector<int64_t> m;
    for (int cand : {1, 2, 4}) {
        bool ok = true;
        for (size_t g = 0; g < sorted.size() && ok; ++g) ok = steps(g, cand, m);
        if (ok) {
            q = cand;
            break;
        }
    }
    if (!q) return false;
    CorPlan P;
    P.q = q;
    for (size_t g = 0; g < sorted.size(); ++g) {
        steps(g, q, m);
        const double s_min = sorted[g].front().first;
        P.sets.push_back(CorRowSet{sorted[g].front().second, {0, 0, 0}});
        for (size_t j = 0; j < m.size();) {
            CorGroup G;
            std::memset(&G, 0, sizeof(G));
            G.set = static_cast<int32_t>(g);
            G.s0 = s_min + static_cast<double>(m[j]) / q;
            for (int k = 0; k < kCorMaxShifts; ++k) G.out[k] = -1;
            const int64_t m0 = m[j];
            for (; j < m.size() && m[j] - m0 <= 6 * q; ++j) {
                const int k = static_cast<int>(m[j] - m0);
                G.out[k] = tasks[sorted[g][j].second].out_offset;
                G.ns = k + 1;
            }
            P.groups.push_back(G);
        }
    }
    if (static_cast<double>(ntasks) < min_avg * static_cast<double>(P.groups.size())) return false;
    std::vector<CorTile> empty;
    for (size_t gi = 0; gi < P.groups.size(); ++gi) {
        const mr_task& T = tasks[P.sets[P.groups[gi].set].task];
        const int ntp = (T.np + BP_TP - 1) / BP_TP, nts = (T.ns + BP_TS - 1) / BP_TS;
        for (int tp = 0; tp < ntp; ++tp)
            for (int ts = 0; ts < nts; ++ts) {
                const int p0 = tp * BP_TP, s0 = ts * BP_TS;
                const int p1 = std::min(p0 + BP_TP, T.np) - 1;
                const int s1 = std::min(s0 + BP_TS, T.ns) - 1;
                const bool hit = tile_may_hit_box(T, 0, p0, p1, s0, s1);
                (hit ? P.tiles : empty).push_back(CorTile{static_cast<int32_t>(gi), tp, ts, hit ? 0 : 1});
            }
    }
    P.ne = empty.size();
    P.tiles.insert(P.tiles.end(), empty.begin(), empty.end());
    plan = std::move(P);
    return true;
}

}  // namespace mr
