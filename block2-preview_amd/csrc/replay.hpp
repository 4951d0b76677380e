// Internal: the device replay plan shared by the H*psi and gradient
// launchers (runtime.hip, grad.hip).
#pragma once

#include "common.hpp"
#include "launch.hpp"

#include <vector>

namespace b2g {

struct GradPlan; // grad.hip
void grad_plan_destroy(GradPlan *g);
struct PkPlan; // pkplan.hip
void pk_destroy(PkPlan *k);

// one phase-2 output block (or window) and its segments [sb, se)
struct BlockJob {
    uint64_t d_off;
    uint32_t ldd, M, N, sb, se, beta;
};

struct Chunk {
    // host images; tasks are bucketed by tile configuration
    std::vector<Task> t1[kTileConfigs], t2[kTileConfigs];
    std::vector<Seg> s1, s2;
    std::vector<RedTask> red; // split-K reductions of phase 2
    std::vector<BlockJob> jobs2; // phase-2 blocks (re-tiled by the packed plan)
    uint64_t w_len = 0;       // W elements per right-hand side
    uint64_t p_len = 0;       // partial-tile elements per right-hand side
    // device copies (inside plan->dev)
    Task *d_t1[kTileConfigs] = {}, *d_t2[kTileConfigs] = {};
    Seg *d_s1 = nullptr, *d_s2 = nullptr;
    RedTask *d_red = nullptr;
    // algorithmic FLOPs per RHS of each tile configuration's launch
    double f1[kTileConfigs] = {}, f2[kTileConfigs] = {};
    // per-XCD task ranges of each launch
    XcdSplit x1[kTileConfigs] = {}, x2[kTileConfigs] = {};
};

} // namespace b2g

struct b2g_plan_s {
    b2g_ctx ctx = nullptr;
    b2g_buf lbuf = nullptr, rbuf = nullptr;
    size_t x_len = 0, y_len = 0;
    uint64_t macs = 0;
    std::vector<b2g::Chunk> chunks;
    uint64_t w_max = 0;
    void *dev = nullptr; // one allocation for all task / segment arrays
    float *l32 = nullptr, *r32 = nullptr;
    void *w = nullptr; // W scratch
    size_t w_bytes = 0;
    uint64_t p_max = 0;   // partial elements per RHS (max over chunks)
    void *pbuf = nullptr; // split-K partial tiles (FP64)
    size_t p_bytes = 0;
    unsigned flags = 0;
    // optional per-launch timing: (kernel id, start, stop) event records
    // resolved when queried; kernel ids as in b2g.h (B2G_NKERNELS)
    bool timing = false;
    struct Stamp {
        int kernel;
        hipEvent_t start, stop;
    };
    std::vector<Stamp> stamps;
    // algorithmic FLOPs of one H*psi per right-hand side, per kernel id
    double kflops[B2G_NKERNELS] = {};
    // the uploaded pair list (the gradient plan is built from it on first
    // use) and that plan
    std::vector<b2g_pair> pairs;
    b2g::GradPlan *grad = nullptr;
    // the split-BF16 replay on packed operand planes (built on first use)
    b2g::PkPlan *pk = nullptr;
    // bumped whenever scratch that launches address (W, partials, packed
    // planes) is reallocated: captured graphs of older generations are stale
    uint64_t scratch_gen = 0;
};


namespace b2g {

// one TM x TN tile of an output block
struct Tile {
    uint32_t row0, col0, tm, tn;
};
// reads the tiling settings of the environment (B2G_TILE, B2G_KSPLIT) for
// the plans this host thread builds next; every plan builder calls it first
void tile_env_refresh();
// tiling of an M x N output block for tasks of depth K (measured tile
// rates, or B2G_TILE); the reference stays valid until the next call
const std::vector<Tile> &tile_cover(uint32_t M, uint32_t N, uint64_t K,
                                    int table);
Task make_task(uint64_t d_off, uint32_t ldd, uint32_t M, uint32_t N,
               const Tile &t, uint32_t sb, uint32_t se, uint32_t beta);
int cfg_of(const Tile &t);
// one task per tile of the block, all over segments [sb, se)
void add_tiles(std::vector<Task> *tasks, uint64_t d_off, uint32_t M,
               uint32_t N, uint64_t K, uint32_t sb, uint32_t se,
               uint32_t beta);
// phase-2 tiles of one block job into t2 (tile rates `table`: 0 FP64 /
// FP32 kernels, 1 split-BF16 kernels); deep segment lists are cut into
// split-K groups with partial slots appended at p_len and a reduction task
void add_split_tiles(const std::vector<Seg> &s2, const BlockJob &j, int table,
                     std::vector<Task> *t2, std::vector<RedTask> &red,
                     uint64_t &p_len);
double task_flops(const std::vector<Task> &tasks, const std::vector<Seg> &segs,
                  int cfg);
void sort_tasks(std::vector<Task> &tasks, const std::vector<Seg> &segs);
// order one launch's tasks: grouped onto the 8 XCDs (the returned per-XCD
// ranges) by segment list, or (by_b) by the B operand of their first
// segment; or sorted by cost (empty ranges)
XcdSplit xcd_order(std::vector<Task> &tasks, const std::vector<Seg> &segs,
                   bool grouped, bool by_b = false);
// scratch elements per chunk and right-hand side (B2G_W_CAP)
uint64_t w_cap_elems();
// allocate the H*psi scratch for n_rhs at precision ahead of a capture
int plan_reserve(b2g_plan p, int n_rhs, int precision);
// the context's pinned staging buffer, at least `bytes` long (host uploads
// are written there, copied on the context stream and waited for)
int ctx_stage(b2g_ctx ctx, size_t bytes, char **out);

// Split-BF16 replay on packed operand planes (pkplan.hip).  Derived from the
// plan's chunks: phase 1 reads the packed wavefunction windows and packed
// right-operator blocks and writes packed W^T; phase 2 reads packed
// left-operator blocks and packed W^T and keeps the FP64 tasks, split-K
// reductions and partial buffer of the plan.
struct PkChunk {
    // tasks tiled with the split-BF16 tile rates, XCD-grouped; phase-1 tasks
    // address packed W^T; phase-2 split-K partials and reductions of its own
    Task *d_t1[kTileConfigs] = {}, *d_t2[kTileConfigs] = {};
    Seg *d_s1 = nullptr, *d_s2 = nullptr;
    RedTask *d_red = nullptr;
    int n_red = 0;
    size_t n1[kTileConfigs] = {}, n2[kTileConfigs] = {};
    double f1[kTileConfigs] = {}, f2[kTileConfigs] = {};
    XcdSplit x1[kTileConfigs] = {}, x2[kTileConfigs] = {};
};
struct PkPlan {
    std::vector<PkChunk> chunks;
    void *dev = nullptr;     // task, segment and pack-tile arrays
    uint16_t *op = nullptr;  // packed operator planes: L hi, L lo, R hi, R lo
    uint64_t l_len = 0, r_len = 0;
    // pack jobs: operator blocks (L, R) and wavefunction windows (X)
    PackMat *d_lmats = nullptr, *d_rmats = nullptr, *d_xmats = nullptr;
    int n_lmats = 0, n_rmats = 0, n_xmats = 0;
    uint64_t n_lgroups = 0, n_rgroups = 0, n_xgroups = 0; // 8-element groups
    uint64_t x_len = 0; // packed wavefunction elements per right-hand side
    uint64_t w_max = 0; // packed W^T elements per right-hand side
    uint64_t p_max = 0; // split-K partial elements per right-hand side
    // packed wavefunction and W^T planes for scratch_rhs right-hand sides:
    // x hi, x lo (x_len * rhs each), W hi, W lo (w_max * rhs each)
    uint16_t *scratch = nullptr;
    int scratch_rhs = 0;
    const uint16_t *lh() const { return op; }
    const uint16_t *ll() const { return op + l_len; }
    const uint16_t *rh() const { return op + 2 * l_len; }
    const uint16_t *rl() const { return op + 2 * l_len + r_len; }
    uint16_t *xh() const { return scratch; }
    uint16_t *xl() const { return scratch + x_len * scratch_rhs; }
    uint16_t *wh() const { return scratch + 2 * x_len * scratch_rhs; }
    uint16_t *wl() const {
        return scratch + (2 * x_len + w_max) * scratch_rhs;
    }
};
// build p->pk from the plan (packs the operator planes); no-op if built
int pk_build(b2g_plan p);
// packed scratch for n_rhs right-hand sides (grows only)
int pk_reserve(b2g_plan p, int n_rhs);
// re-pack the operator planes after the operator buffers changed
int pk_refresh(b2g_plan p);

} // namespace b2g
