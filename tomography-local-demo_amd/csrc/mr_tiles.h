// Host-only request preparation of the back-projection entry points: task validation,
// tile enumeration with the conservative box test, the height order of the tiles, and
// the per-context cache of recent tile lists.  Plain C++ (no HIP), so that it also
// builds with g++ and the sanitizers on its own.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/minirecast_hip.h"

namespace mr {

// BP tile shape: TP x TS points per 256-thread workgroup.
constexpr int BP_TS = 32;
constexpr int BP_TP = 32;

// Z-column tiles of volume requests: ZT x ZT point columns (one per thread) of ZQ planes
// (gathers from memory) or ZQL planes (taps staged through LDS: a box of fewer rows).
constexpr int ZT = 16;
constexpr int ZQ = 16;
constexpr int ZQL = 16;

// Tile work item of a back-projection launch: task, z-plane q, tile row tp, tile column ts
// (z-column tiles: the chunk's first plane q and ZT x ZT column tile (tp, ts)).
struct Tile {
    int32_t task, q, tp, ts;
};

// Tile list of a recent request of a context (keyed by the tasks' bytes and the
// operation / block-order flags), reused when a request repeats it.
constexpr size_t kTileCacheEntries = 4;
struct TileCache {
    int flags = 0;
    std::vector<uint8_t> tasks;
    std::vector<Tile> tiles;         // ordered tiles, then the culled ones
    size_t ne = 0;                   // culled tiles at the end
};

// Empty string when every task's writes stay inside [0, out_len), else the reason.
std::string validate_tasks_msg(const mr_task* tasks, int ntasks, int64_t out_len);

// Conservative test: can any point of the tile lie inside the box?
bool tile_may_hit_box(const mr_task& T, int q, int p0, int p1, int s0, int s1);

// tiles: those that may hit the box; empty (optional): the culled rest, whose points
// are all zero.
void build_tiles(const mr_task* tasks, int ntasks, std::vector<Tile>& tiles,
                 std::vector<Tile>* empty = nullptr);

// Tiles of the z-column kernel: every chunk of zq planes (from plane 0) x every ZT x ZT
// column tile of every task, in task, chunk, row, column order.
void build_zcol_tiles(const mr_task* tasks, int ntasks, int zq, std::vector<Tile>& tiles);

// ---- shift groups of rotation-centre searches (parallel beam, axial planes)
// A group: tasks that differ only in cor_shift and out_offset, with shifts s0 + k / q
// (q = 1, 2 or 4, k < 6 q + 1): their columns lie within 6 px of the first one, so the
// kCorWin detector values from floor(col + s0) on hold every tap of every shift.
constexpr int kCorWin = 8;              // detector values per point and view
constexpr int kCorMaxShifts = 25;       // 6 q + 1 for q = 4
constexpr int kCorPad = 16;             // zero columns on each side of a set's detector rows

// A set of tasks equal but for cor_shift and out_offset: one detector row per view.
struct CorRowSet {
    int32_t task;                       // a task of the set (plane, mask, strides)
    int32_t pad[3];
};
struct CorGroup {
    int32_t set;
    int32_t ns;                         // outputs k < ns: shift s0 + k / q
    int32_t pad[2];
    double s0;
    int64_t out[kCorMaxShifts];         // out_offset of the task with shift s0 + k / q; -1 none
};
// 32 x 32 tile (tp, ts) of a group; zero: no point of the tile lies in the box.
struct CorTile {
    int32_t group, tp, ts, zero;
};
struct CorPlan {
    int q = 0;
    std::vector<CorRowSet> sets;
    std::vector<CorGroup> groups;
    std::vector<CorTile> tiles;          // tiles that may hit the box, then the others
    size_t ne = 0;                       // tiles with zero = 1 (at the end)
};

// Fills `plan` and returns true when every task is an axial plane (nq = 1, a[2] = b[2] =
// 0, finite shift) and the tasks fall into sets whose shifts lie on one grid s0 + m / q
// with q = 1, 2 or 4 (the smallest that fits every set; no shift twice in a set), split
// into groups spanning at most 6 q grid steps, with at least min_avg tasks per group on
// average.  Otherwise returns false and leaves `plan` empty.
bool cor_shift_groups(const mr_task* tasks, int ntasks, double min_avg, CorPlan& plan);

// Orders tiles by the height z of their centre (stable).
void order_tiles_by_height(const mr_task* tasks, std::vector<Tile>& tiles);

// The tile list of a request from the cache (most recent first, at most
// kTileCacheEntries; keyed by the tasks without their cor_shift, out_offset and strides,
// which the tiles do not depend on): a hit moves the entry to the front; a miss enumerates the tiles
// (keeping the culled ones after them when keep_empty), orders them by height when
// by_height, and puts the new entry in front, dropping the oldest.  flags is part of
// the key.  The returned entry stays valid until the next call on the same cache.
const TileCache& cached_tiles(std::vector<TileCache>& cache, const mr_task* tasks, int ntasks,
                              int flags, bool keep_empty, bool by_height);

}  // namespace mr
