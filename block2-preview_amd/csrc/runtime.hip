// C ABI of the thin device layer: contexts, buffers, replay plans, H*psi,
// diagonal and vector kernels.
#include "common.hpp"
#include "launch.hpp"
#include "replay.hpp"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <unistd.h>

namespace b2g {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// Scratch of the allocation-free H*psi (b2g_hpsi_noalloc): the W windows
// and split-K partials of one product.  Each device that loads the library
// holds its own copy until the process ends.  In host code `scratch` names
// a host-side stand-in, never the device memory: kernels get the device
// address taken by hipGetSymbolAddress at context creation
// (g_scratch_ptr[device]).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double
    scratch[kNoAllocScratchBytes / sizeof(double)];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Takes the device address of `scratch` on `device` (current) once.
static int scratch_setup(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) {
        set_error("device index beyond the scratch table");
        return B2G_ERR_ARG;
    }
    if (g_scratch_ptr[device] != nullptr)
        return B2G_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
        set_error(std::string("hipGetSymbolAddress(scratch) failed: ") +
                  hipGetErrorString(e));
        return B2G_ERR_DEVICE;
    }
    g_scratch_ptr[device] = ptr;
    return B2G_OK;
}

static void *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device]
                                               : nullptr;
}

} // namespace b2g

using namespace b2g;

extern "C" const char *b2g_last_error(void) { return g_err.c_str(); }

extern "C" int b2g_abi_sizes(size_t *out, int n) {
    const size_t sz[] = {sizeof(b2g_pair),        sizeof(b2g_block),
                         sizeof(b2g_diag_term),   sizeof(b2g_sparse_info),
                         sizeof(b2g_operator),    sizeof(b2g_term),
                         sizeof(b2g_tp_conn),     sizeof(b2g_problem),
                         sizeof(b2g_davidson_opts)};
    const int k = (int)(sizeof(sz) / sizeof(sz[0]));
    for (int i = 0; i < n && i < k && out; i++)
        out[i] = sz[i];
    return k;
}

// ------------------------------------------------------------------ context
extern "C" int b2g_ctx_create(int device, unsigned flags, b2g_ctx *out) {
    (void)flags;
    B2G_REQUIRE(out != nullptr, B2G_ERR_ARG, "b2g_ctx_create: out is NULL");
    int ndev = 0;
    B2G_HIP(hipGetDeviceCount(&ndev));
    B2G_REQUIRE(device >= 0 && device < ndev, B2G_ERR_ARG,
                "b2g_ctx_create: no such device");
    B2G_HIP(hipSetDevice(device));
    {
        const int r = scratch_setup(device);
        if (r != B2G_OK)
            return r;
    }
    b2g_ctx c = new b2g_ctx_s();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        set_error("hipGetDeviceProperties failed");
        return B2G_ERR_DEVICE;
    }
    c->num_cu = prop.multiProcessorCount;
    // a blocking stream (default flags and priority): work on the legacy
    // default stream, where callers such as PyTorch enqueue by default, is
    // ordered against it
    if (hipStreamCreateWithFlags(&c->stream, hipStreamDefault) != hipSuccess) {
        delete c;
        set_error("hipStreamCreate failed");
        return B2G_ERR_DEVICE;
    }
    // B2G_STREAMS: streams per H*psi phase, default 3 (with one right-hand
    // side the launches of the smaller tile shapes fill the compute units
    // that the largest launch leaves idle in its tail)
    int nstreams = 3;
    if (const char *e = getenv("B2G_STREAMS"))
        nstreams = std::max(1, std::min(4, atoi(e)));
    bool ok = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) ==
              hipSuccess;
    for (int i = 1; ok && i < nstreams; i++) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
        if (ok)
            c->aux.push_back(s);
        ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) ==
                       hipSuccess;
        if (ok)
            c->join.push_back(e);
    }
    if (!ok) {
        for (hipStream_t s : c->aux)
            hipStreamDestroy(s);
        for (hipEvent_t e : c->join)
            hipEventDestroy(e);
        if (c->fork)
            hipEventDestroy(c->fork);
        hipStreamDestroy(c->stream);
        delete c;
        set_error("creation of auxiliary streams failed");
        return B2G_ERR_DEVICE;
    }
    // reduction partials: red_grid() blocks x up to 64*64 values, plus room
    c->red_len = (size_t)red_grid() * 64 * 64 + 8192;
    if (hipMalloc(&c->red, c->red_len * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&c->red_host, 8192 * sizeof(double)) != hipSuccess) {
        if (c->red)
            hipFree(c->red);
        for (hipStream_t s : c->aux)
            hipStreamDestroy(s);
        for (hipEvent_t e : c->join)
            hipEventDestroy(e);
        hipEventDestroy(c->fork);
        hipStreamDestroy(c->stream);
        delete c;
        set_error("allocation of reduction scratch failed");
        return B2G_ERR_MEMORY;
    }
    *out = c;
    return B2G_OK;
}

extern "C" int b2g_ctx_destroy(b2g_ctx ctx) {
    if (ctx == nullptr)
        return B2G_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (hipStream_t s : ctx->aux) {
        hipStreamSynchronize(s);
        hipStreamDestroy(s);
    }
    for (hipEvent_t e : ctx->join)
        hipEventDestroy(e);
    hipEventDestroy(ctx->fork);
    hipFree(ctx->red);
    hipHostFree(ctx->red_host);
    if (ctx->stage)
        hipHostFree(ctx->stage);
    for (auto &b : ctx->win_pool)
        hipFree(b.second);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return B2G_OK;
}

extern "C" void *b2g_ctx_stream(b2g_ctx ctx) {
    return ctx ? (void *)ctx->stream : nullptr;
}

extern "C" int b2g_ctx_synchronize(b2g_ctx ctx) {
    B2G_REQUIRE(ctx != nullptr, B2G_ERR_ARG, "ctx is NULL");
    B2G_HIP(hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

extern "C" int b2g_ctx_num_cu(b2g_ctx ctx, int *out) {
    B2G_REQUIRE(ctx != nullptr && out != nullptr, B2G_ERR_ARG, "NULL arg");
    *out = ctx->num_cu;
    return B2G_OK;
}

// ------------------------------------------------------------------ buffers
static size_t dtype_size(int dtype) {
    return dtype == B2G_F64 ? 8 : dtype == B2G_F32 ? 4 : 0;
}

extern "C" int b2g_buf_create(b2g_ctx ctx, size_t n, int dtype, b2g_buf *out) {
    B2G_REQUIRE(ctx != nullptr && out != nullptr, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(dtype_size(dtype) != 0, B2G_ERR_ARG, "bad dtype");
    B2G_HIP(hipSetDevice(ctx->device));
    b2g_buf b = new b2g_buf_s();
    b->ctx = ctx;
    b->n = n;
    b->dtype = dtype;
    if (n != 0 && hipMalloc(&b->ptr, n * dtype_size(dtype)) != hipSuccess) {
        delete b;
        set_error("b2g_buf_create: hipMalloc failed");
        return B2G_ERR_MEMORY;
    }
    *out = b;
    return B2G_OK;
}

extern "C" int b2g_buf_destroy(b2g_buf buf) {
    if (buf == nullptr)
        return B2G_OK;
    hipSetDevice(buf->ctx->device);
    hipStreamSynchronize(buf->ctx->stream);
    if (buf->ptr && buf->owned)
        hipFree(buf->ptr);
    delete buf;
    return B2G_OK;
}

extern "C" int b2g_buf_wrap(b2g_ctx ctx, void *ptr, size_t n, int dtype,
                            b2g_buf *out) {
    B2G_REQUIRE(ctx != nullptr && out != nullptr && (ptr || n == 0),
                B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(dtype_size(dtype) != 0, B2G_ERR_ARG, "bad dtype");
    b2g_buf b = new b2g_buf_s();
    b->ctx = ctx;
    b->ptr = ptr;
    b->n = n;
    b->dtype = dtype;
    b->owned = false;
    *out = b;
    return B2G_OK;
}

extern "C" int b2g_buf_write(b2g_buf buf, const void *host, size_t offset,
                             size_t n) {
    B2G_REQUIRE(buf != nullptr && (host != nullptr || n == 0), B2G_ERR_ARG,
                "NULL arg");
    B2G_REQUIRE(offset + n <= buf->n, B2G_ERR_ARG, "b2g_buf_write: range");
    if (n == 0)
        return B2G_OK;
    const size_t es = dtype_size(buf->dtype);
    B2G_HIP(hipMemcpyAsync((char *)buf->ptr + offset * es, host, n * es,
                           hipMemcpyHostToDevice, buf->ctx->stream));
    B2G_HIP(hipStreamSynchronize(buf->ctx->stream));
    return B2G_OK;
}

extern "C" int b2g_buf_read(b2g_buf buf, void *host, size_t offset, size_t n) {
    B2G_REQUIRE(buf != nullptr && (host != nullptr || n == 0), B2G_ERR_ARG,
                "NULL arg");
    B2G_REQUIRE(offset + n <= buf->n, B2G_ERR_ARG, "b2g_buf_read: range");
    if (n == 0)
        return B2G_OK;
    const size_t es = dtype_size(buf->dtype);
    B2G_HIP(hipMemcpyAsync(host, (char *)buf->ptr + offset * es, n * es,
                           hipMemcpyDeviceToHost, buf->ctx->stream));
    B2G_HIP(hipStreamSynchronize(buf->ctx->stream));
    return B2G_OK;
}

extern "C" int b2g_upload_blocks(b2g_ctx ctx, const void *host, size_t n,
                                 int dtype, b2g_buf *out) {
    int r = b2g_buf_create(ctx, n, dtype, out);
    if (r != B2G_OK)
        return r;
    r = b2g_buf_write(*out, host, 0, n);
    if (r != B2G_OK) {
        b2g_buf_destroy(*out);
        *out = nullptr;
    }
    return r;
}

extern "C" void *b2g_buf_device_ptr(b2g_buf buf) {
    return buf ? buf->ptr : nullptr;
}
extern "C" size_t b2g_buf_size(b2g_buf buf) { return buf ? buf->n : 0; }

// -------------------------------------------------------------------- plans
namespace b2g {

// depth of each task (sum of its segments' K, + 1), from a prefix sum over
// the segment list
static std::vector<uint64_t> task_depths(const std::vector<Task> &tasks,
                                         const std::vector<Seg> &segs) {
    std::vector<uint64_t> pre(segs.size() + 1, 0);
    for (size_t s = 0; s < segs.size(); s++)
        pre[s + 1] = pre[s] + segs[s].K;
    std::vector<uint64_t> d(tasks.size());
    for (size_t i = 0; i < tasks.size(); i++)
        d[i] = pre[tasks[i].seg_end] - pre[tasks[i].seg_begin] + 1;
    return d;
}

// Tiling policy.  Every output block is covered by workgroup tiles whose
// row and column sizes are chosen from {32, 64, 128}: full 128 tiles for the
// bulk and one remainder tile per dimension, picked jointly to minimise the
// estimated time  sum(TM * TN * K / rate(TM, TN))  with measured per-config
// rates (b2g_set_tile_rates, filled by the autotuner).  B2G_TILE=32|64|128
// forces square tiles of one size.
// Two tables: [0] tiles of the FP64 / FP32 kernels, [1] tiles of the
// split-BF16 kernels (packed plan).  The split-BF16 default is the
// per-shape throughput those kernels reach on a heavy-tailed sector
// histogram (n=40, M=3000 SU(2) middle site, one right-hand side), where the
// small tiles, fed by shallow segments, run far below their rate on uniform
// blocks: 4.40 ms per product against 5.1 ms with rates measured on uniform
// 436 x 436 blocks.
static double g_rates[2][kTileConfigs] = {
    {30, 34, 36, 34, 42, 46, 36, 46, 50},
    {4, 15, 25, 20, 45, 80, 30, 90, 200}};

static int size_index(uint32_t t) { return t == 32 ? 0 : t == 64 ? 1 : 2; }

// candidate row (or column) covers of a dimension of length d
static std::vector<std::vector<uint32_t>> covers(uint32_t d) {
    std::vector<uint32_t> base(d / 128, 128u);
    const uint32_t r = d % 128;
    std::vector<std::vector<uint32_t>> out;
    if (r == 0) {
        out.push_back(base);
        return out;
    }
    for (const std::vector<uint32_t> &tail :
         std::vector<std::vector<uint32_t>>{{128}, {64}, {64, 32}, {32},
                                            {32, 32}}) {
        uint32_t cov = 0;
        for (uint32_t t : tail)
            cov += t;
        if (cov < r)
            continue;
        std::vector<uint32_t> v = base;
        v.insert(v.end(), tail.begin(), tail.end());
        out.push_back(v);
    }
    return out;
}

// tiles covering an M x N output block whose tasks have depth K
// Tilings are memoised per (M, N, table, forced shape): the estimated time
// of every cover scales with K + 1 alike, so the choice does not depend on
// K.  The memo is per host thread and cleared when rates change.
static std::atomic<uint64_t> g_rates_gen{1};
struct TileMemo {
    uint64_t gen = 0;
    std::unordered_map<uint64_t, std::vector<Tile>> map;
    std::vector<Tile> big; // blocks too large for the memo key
};
static thread_local TileMemo t_memo;

// tiling settings of the environment, read once per plan build
struct TileEnv {
    uint32_t fm = 0, fn = 0; // forced tile shape (0: measured rates)
    uint64_t ksplit = 2048;  // split-K depth threshold of phase 2
};
static thread_local TileEnv t_env;

void tile_env_refresh() {
    TileEnv v;
    if (const char *e = getenv("B2G_TILE")) {
        // "T" (square) or "TMxTN", sizes from {32, 64, 128}
        unsigned a = 0, b = 0;
        const int nr = sscanf(e, "%ux%u", &a, &b);
        if (nr == 1)
            b = a;
        auto okv = [](unsigned t) { return t == 32 || t == 64 || t == 128; };
        if (nr >= 1 && okv(a) && okv(b))
            v.fm = a, v.fn = b;
    }
    if (const char *e = getenv("B2G_KSPLIT"))
        if (atoll(e) > 0)
            v.ksplit = (uint64_t)atoll(e);
    t_env = v;
}

const std::vector<Tile> &tile_cover(uint32_t M, uint32_t N, uint64_t K,
                                    int table) {
    (void)K;
    const double *rates = g_rates[table == 1 ? 1 : 0];
    std::vector<uint32_t> rows, cols;
    const uint32_t fm = t_env.fm, fn = t_env.fn;
    if (t_memo.gen != g_rates_gen) {
        t_memo.map.clear();
        t_memo.gen = g_rates_gen;
    }
    const bool memo = M < (1u << 24) && N < (1u << 24);
    const uint64_t key = ((uint64_t)M << 36) | ((uint64_t)N << 12) |
                         ((uint64_t)(table == 1) << 11) |
                         ((uint64_t)(fm >> 5) << 8) | ((uint64_t)(fn >> 5) << 4);
    if (memo) {
        auto it = t_memo.map.find(key);
        if (it != t_memo.map.end())
            return it->second;
    }
    if (fm != 0) {
        rows.assign((M + fm - 1) / fm, fm);
        cols.assign((N + fn - 1) / fn, fn);
    } else {
        double best = 1e300;
        for (const auto &rc : covers(M))
            for (const auto &cc : covers(N)) {
                double c = 0;
                for (uint32_t a : rc)
                    for (uint32_t b : cc)
                        c += (double)a * b /
                             rates[3 * size_index(a) + size_index(b)];
                if (c < best)
                    best = c, rows = rc, cols = cc;
            }
    }
    std::vector<Tile> out;
    uint32_t r0 = 0;
    for (uint32_t a : rows) {
        uint32_t c0 = 0;
        for (uint32_t b : cols) {
            out.push_back(Tile{r0, c0, a, b});
            c0 += b;
        }
        r0 += a;
    }
    if (memo)
        return t_memo.map.emplace(key, std::move(out)).first->second;
    t_memo.big.swap(out);
    return t_memo.big;
}

Task make_task(uint64_t d_off, uint32_t ldd, uint32_t M, uint32_t N,
                      const Tile &t, uint32_t sb, uint32_t se, uint32_t beta) {
    Task k;
    k.d_off = d_off;
    k.ldd = ldd;
    k.M = M;
    k.N = N;
    k.row0 = t.row0;
    k.col0 = t.col0;
    k.seg_begin = sb;
    k.seg_end = se;
    k.beta = beta;
    return k;
}

int cfg_of(const Tile &t) {
    return 3 * size_index(t.tm) + size_index(t.tn);
}

// one task per tile (phase 1 and unsplit phase-2 blocks)
void add_tiles(std::vector<Task> *tasks, uint64_t d_off, uint32_t M,
                      uint32_t N, uint64_t K, uint32_t sb, uint32_t se,
                      uint32_t beta) {
    for (const Tile &t : tile_cover(M, N, K, 0))
        tasks[cfg_of(t)].push_back(make_task(d_off, N, M, N, t, sb, se, beta));
}

// Phase-2 tiles of an output block whose segments [sb, se) are deep: the
// segment list is cut at segment boundaries into groups of depth about
// ksplit(); each (tile, group) becomes a partial task writing a dense
// TM x TN slot of the partial buffer, and a reduction task sums the slots in
// group order (deterministic) into the output.
void add_split_tiles(const std::vector<Seg> &s2, const BlockJob &j, int table,
                     std::vector<Task> *t2, std::vector<RedTask> &red,
                     uint64_t &p_len) {
    std::vector<std::pair<uint32_t, uint32_t>> groups;
    // maximum depth of one phase-2 task before its segment list is split
    const uint64_t kmax = t_env.ksplit;
    // depth as the kernels walk it: every segment takes whole k-steps
    // (16 deep for the FP64 / FP32 kernels, 32 for the split-BF16 ones), so
    // that a long list of shallow segments is split like a deep one
    const uint64_t step = table == 1 ? 32 : 16;
    uint64_t K = 0, acc = 0;
    uint32_t g0 = j.sb;
    for (uint32_t k = j.sb; k < j.se; k++) {
        const uint64_t kk = (s2[k].K + step - 1) / step * step;
        K += kk;
        acc += kk;
        if (acc >= kmax && k + 1 < j.se) {
            groups.emplace_back(g0, k + 1);
            g0 = k + 1, acc = 0;
        }
    }
    groups.emplace_back(g0, j.se);
    const std::vector<Tile> &tiles =
        tile_cover(j.M, j.N, K / groups.size(), table);
    if (groups.size() == 1) {
        for (const Tile &t : tiles)
            t2[cfg_of(t)].push_back(
                make_task(j.d_off, j.ldd, j.M, j.N, t, j.sb, j.se, j.beta));
        return;
    }
    for (const Tile &t : tiles) {
        const uint64_t tsz = (uint64_t)t.tm * t.tn;
        const uint64_t p0 = p_len;
        for (size_t g = 0; g < groups.size(); g++) {
            // slot of group g: element (row, col) of the block lands at
            // p0 + g*tsz + (row - row0)*tn + (col - col0)  (mod 2^64)
            const uint64_t doff = p0 + g * tsz - (uint64_t)t.row0 * t.tn -
                                  (uint64_t)t.col0;
            t2[cfg_of(t)].push_back(make_task(doff, t.tn, j.M, j.N, t,
                                              groups[g].first,
                                              groups[g].second, 2u));
        }
        RedTask rt;
        rt.d_off = j.d_off;
        rt.p_off = p0;
        rt.ldd = j.ldd;
        rt.M = j.M;
        rt.N = j.N;
        rt.row0 = t.row0;
        rt.col0 = t.col0;
        rt.tm = t.tm;
        rt.tn = t.tn;
        rt.nparts = (uint32_t)groups.size();
        rt.beta = j.beta;
        rt.pad = 0;
        red.push_back(rt);
        p_len += groups.size() * tsz;
    }
}

// multiply-adds x 2 of the tasks restricted to their tiles (the edge tiles'
// masked rows and columns do not count)
double task_flops(const std::vector<Task> &tasks,
                         const std::vector<Seg> &segs, int cfg) {
    double f = 0;
    const uint32_t tm = 32u << (cfg / 3), tn = 32u << (cfg % 3);
    const std::vector<uint64_t> d = task_depths(tasks, segs);
    for (size_t i = 0; i < tasks.size(); i++) {
        const Task &t = tasks[i];
        const uint32_t m = std::min(tm, t.M - t.row0);
        const uint32_t n = std::min(tn, t.N - t.col0);
        f += 2.0 * m * n * (double)(d[i] - 1);
    }
    return f;
}

// Sorts (key, index) pairs given in ascending index order by key, equal
// keys in index order: LSD radix sort over 16-bit digits for keys below
// 2^32 (two linear passes), std::sort otherwise.
void sort_keyed(std::vector<std::pair<uint64_t, size_t>> &v) {
    uint64_t mx = 0;
    for (const auto &e : v)
        mx = std::max(mx, e.first);
    if (v.size() < 4096 || mx >= ((uint64_t)1 << 32)) {
        std::sort(v.begin(), v.end());
        return;
    }
    std::vector<std::pair<uint64_t, size_t>> tmp(v.size());
    std::vector<size_t> cnt(65537);
    for (int sh = 0; sh < 32 && (mx >> sh) != 0; sh += 16) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (const auto &e : v)
            cnt[((e.first >> sh) & 0xFFFF) + 1]++;
        for (size_t d = 0; d < 65536; d++)
            cnt[d + 1] += cnt[d];
        for (const auto &e : v)
            tmp[cnt[(e.first >> sh) & 0xFFFF]++] = e;
        v.swap(tmp);
    }
}

// deepest first, creation order among equal depths
void sort_tasks(std::vector<Task> &tasks, const std::vector<Seg> &segs) {
    const std::vector<uint64_t> d = task_depths(tasks, segs);
    const uint64_t dmax =
        d.empty() ? 0 : *std::max_element(d.begin(), d.end());
    std::vector<std::pair<uint64_t, size_t>> order(tasks.size());
    for (size_t i = 0; i < order.size(); i++)
        order[i] = std::make_pair(dmax - d[i], i);
    sort_keyed(order);
    std::vector<Task> out(tasks.size());
    for (size_t i = 0; i < order.size(); i++)
        out[i] = tasks[order[i].second];
    tasks.swap(out);
}

// B2G_XCD=1 turns the per-XCD task ranges on for the FP64 / FP32 kernels
// (measured on the n=40 M=3000 middle-site workload: no gain over the
// cost-sorted list, so off; the split-BF16 kernels, which need 5x the
// operand bandwidth per flop, always use them)
static bool xcd_enabled() {
    const char *e = getenv("B2G_XCD");
    return e != nullptr && atoi(e) != 0;
}

// Orders one launch's tasks for the 8 XCDs and returns their ranges (not
// grouped: sorted by cost, one range for all XCDs).  Tasks
// with the same segment list (the tiles of one output block, or of one
// split-K group) re-read the same operator and work panels, so they form a
// group that stays on one XCD and runs back to back there (shared L2).
// Groups are dealt longest first to the least-loaded XCD; each XCD runs its
// groups in that order, tiles of a group in creation (row-major) order.
XcdSplit xcd_order(std::vector<Task> &tasks, const std::vector<Seg> &segs,
                   bool grouped, bool by_b) {
    XcdSplit xs{};
    if (tasks.empty())
        return xs;
    if (!grouped) {
        sort_tasks(tasks, segs);
        return xs;
    }
    // groups: runs of equal key after sorting (key, index) pairs, so that
    // creation order is kept inside a group.  Group key: the segment list,
    // or (by_b) the B operand of the first segment, so that every task
    // reading one B block runs on one XCD
    std::vector<std::pair<uint64_t, size_t>> ki(tasks.size());
    bool sorted = true;
    for (size_t i = 0; i < tasks.size(); i++) {
        ki[i] = std::make_pair(
            by_b ? segs[tasks[i].seg_begin].b_off : tasks[i].seg_begin, i);
        sorted = sorted && (i == 0 || ki[i - 1].first <= ki[i].first);
    }
    if (!sorted)
        sort_keyed(ki);
    const std::vector<uint64_t> depth = task_depths(tasks, segs);
    std::vector<size_t> gstart;
    std::vector<uint64_t> gcost;
    for (size_t k = 0; k < ki.size(); k++) {
        if (k == 0 || ki[k].first != ki[k - 1].first) {
            gstart.push_back(k);
            gcost.push_back(0);
        }
        gcost.back() += depth[ki[k].second];
    }
    gstart.push_back(ki.size());
    const size_t ng = gcost.size();
    // costliest group first, lower group index among equal costs
    const uint64_t cmax = *std::max_element(gcost.begin(), gcost.end());
    std::vector<std::pair<uint64_t, size_t>> gord(ng);
    for (size_t g = 0; g < ng; g++)
        gord[g] = std::make_pair(cmax - gcost[g], g);
    sort_keyed(gord);
    std::vector<size_t> order(ng);
    for (size_t g = 0; g < ng; g++)
        order[g] = gord[g].second;
    // longest group first to the least-loaded XCD
    double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<int> gx(ng);
    size_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t g : order) {
        int x = 0;
        for (int y = 1; y < 8; y++)
            if (load[y] < load[x])
                x = y;
        load[x] += (double)gcost[g];
        gx[g] = x;
        cnt[x] += gstart[g + 1] - gstart[g];
    }
    size_t pos[8];
    size_t acc = 0;
    for (int x = 0; x < 8; x++) {
        xs.off[x] = (uint32_t)acc;
        pos[x] = acc;
        acc += cnt[x];
    }
    std::vector<Task> out(tasks.size());
    for (size_t g : order)
        for (size_t k = gstart[g]; k < gstart[g + 1]; k++)
            out[pos[gx[g]]++] = tasks[ki[k].second];
    xs.off[8] = (uint32_t)acc;
    tasks.swap(out);
    return xs;
}

// host workers fed from one queue; never destroyed (the threads are
// detached and wait on the queue until the process ends)
namespace {
class HostPool {
  public:
    explicit HostPool(int n) {
        for (int i = 0; i < n; i++)
            std::thread([this] { loop(); }).detach();
    }
    void submit(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> g(m_);
            q_.push_back(std::move(f));
        }
        cv_.notify_one();
    }

  private:
    void loop() {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [this] { return !q_.empty(); });
                f = std::move(q_.front());
                q_.pop_front();
            }
            f();
        }
    }
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
};

// one pool per process: a child created by fork has none of its parent's
// threads, so it starts a pool of its own
HostPool *host_pool() {
    static std::mutex m;
    static HostPool *pool = nullptr;
    static pid_t owner = 0;
    std::lock_guard<std::mutex> g(m);
    if (pool == nullptr || owner != getpid()) {
        const unsigned hw = std::thread::hardware_concurrency();
        const int n = (int)std::min(7u, std::max(2u, hw) - 1);
        pool = new HostPool(n);
        owner = getpid();
    }
    return pool;
}
} // namespace

void run_host_jobs(std::vector<std::function<void()>> &jobs) {
    if (jobs.empty())
        return;
    struct Wait {
        std::mutex m;
        std::condition_variable cv;
        size_t left = 0;
    } w;
    w.left = jobs.size() - 1;
    if (w.left > 0) {
        HostPool *pool = host_pool();
        for (size_t i = 1; i < jobs.size(); i++)
            pool->submit([&w, &jobs, i] {
                jobs[i]();
                std::lock_guard<std::mutex> g(w.m);
                if (--w.left == 0)
                    w.cv.notify_one();
            });
    }
    jobs[0]();
    std::unique_lock<std::mutex> g(w.m);
    w.cv.wait(g, [&w] { return w.left == 0; });
}

size_t plan_jobs(uint64_t items, uint64_t per_job) {
    if (const char *e = getenv("B2G_PLAN_JOB_ITEMS")) {
        const long long v = atoll(e);
        if (v > 0)
            per_job = std::min<uint64_t>(per_job, (uint64_t)v);
    }
    return (size_t)std::min<uint64_t>(8, items / std::max<uint64_t>(per_job, 1) + 1);
}

void parallel_memcpy(void *dst, const void *src, size_t bytes) {
    const size_t piece = (size_t)64 << 20;
    const size_t n = std::min<size_t>(8, (bytes + piece - 1) / piece);
    if (n <= 1) {
        if (bytes)
            std::memcpy(dst, src, bytes);
        return;
    }
    const size_t step = (bytes / n + 4095) & ~(size_t)4095;
    std::vector<std::function<void()>> jobs;
    for (size_t o = 0; o < bytes; o += step) {
        const size_t len = std::min(step, bytes - o);
        jobs.emplace_back([=]() {
            std::memcpy((char *)dst + o, (const char *)src + o, len);
        });
    }
    run_host_jobs(jobs);
}

uint64_t w_cap_elems() {
    const char *e = getenv("B2G_W_CAP");
    if (e != nullptr) {
        long long v = atoll(e);
        if (v > 0)
            return (uint64_t)v;
    }
    return (uint64_t)1 << 30; // 1 Gi elements (8 GiB FP64) per right-hand side
}

int ctx_stage(b2g_ctx ctx, size_t bytes, char **out) {
    if (bytes > ctx->stage_len) {
        B2G_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->stage) {
            B2G_HIP(hipHostFree(ctx->stage));
            ctx->stage = nullptr;
            ctx->stage_len = 0;
        }
        const size_t mib = (size_t)1 << 20;
        const size_t len =
            (std::max(bytes, 2 * ctx->stage_len) + mib - 1) / mib * mib;
        B2G_HIP(hipHostMalloc((void **)&ctx->stage, len));
        ctx->stage_len = len;
    }
    *out = ctx->stage;
    return B2G_OK;
}

// upload all chunks' task / segment / reduction arrays in one allocation
static int upload_chunks(b2g_plan p) {
    size_t bytes = 0;
    for (auto &ch : p->chunks) {
        for (int c = 0; c < kTileConfigs; c++)
            bytes += (ch.t1[c].size() + ch.t2[c].size()) * sizeof(Task);
        bytes += (ch.s1.size() + ch.s2.size()) * sizeof(Seg);
        bytes += ch.red.size() * sizeof(RedTask);
    }
    char *host = nullptr;
    const int rs = ctx_stage(p->ctx, bytes + 16, &host);
    if (rs != B2G_OK)
        return rs;
    B2G_HIP(hipMalloc(&p->dev, bytes + 16));
    size_t off = 0;
    auto put = [&](const void *src, size_t nb) -> size_t {
        size_t o = off;
        if (nb)
            memcpy(host + o, src, nb);
        off += nb;
        return o;
    };
    struct Offs {
        size_t t1[kTileConfigs], t2[kTileConfigs], s1, s2, red;
    };
    std::vector<Offs> offs;
    for (auto &ch : p->chunks) {
        Offs o;
        for (int c = 0; c < kTileConfigs; c++) {
            o.t1[c] = put(ch.t1[c].data(), ch.t1[c].size() * sizeof(Task));
            o.t2[c] = put(ch.t2[c].data(), ch.t2[c].size() * sizeof(Task));
        }
        o.s1 = put(ch.s1.data(), ch.s1.size() * sizeof(Seg));
        o.s2 = put(ch.s2.data(), ch.s2.size() * sizeof(Seg));
        o.red = put(ch.red.data(), ch.red.size() * sizeof(RedTask));
        offs.push_back(o);
    }
    B2G_HIP(hipMemcpyAsync(p->dev, host, bytes, hipMemcpyHostToDevice,
                           p->ctx->stream));
    B2G_HIP(hipStreamSynchronize(p->ctx->stream));
    char *base = (char *)p->dev;
    for (size_t i = 0; i < p->chunks.size(); i++) {
        Chunk &ch = p->chunks[i];
        for (int c = 0; c < kTileConfigs; c++) {
            ch.d_t1[c] = (Task *)(base + offs[i].t1[c]);
            ch.d_t2[c] = (Task *)(base + offs[i].t2[c]);
        }
        ch.d_s1 = (Seg *)(base + offs[i].s1);
        ch.d_s2 = (Seg *)(base + offs[i].s2);
        ch.d_red = (RedTask *)(base + offs[i].red);
    }
    return B2G_OK;
}

static int compile_plan(b2g_plan p, const b2g_pair *pairs, size_t n_pairs,
                        const b2g_block *yb, size_t n_yb) {
    SetupTimer tm("plan");
    tile_env_refresh();
    const size_t llen = p->lbuf->n, rlen = p->rbuf->n;
    FlatMap<uint64_t, size_t, U64Hash> yidx(n_yb);
    std::vector<std::pair<uint64_t, size_t>> ysorted; // (offset, block)
    for (size_t i = 0; i < n_yb; i++) {
        B2G_REQUIRE(yb[i].off + (uint64_t)yb[i].m * yb[i].n <= p->y_len,
                    B2G_ERR_ARG, "output block outside the output vector");
        *yidx.emplace(yb[i].off, i).first = i;
        ysorted.emplace_back(yb[i].off, i);
    }
    std::sort(ysorted.begin(), ysorted.end());
    p->macs = 0;
    std::vector<size_t> whole, windows; // pair indices by kind
    // output block of each whole-block pair (by pair index)
    std::vector<uint32_t> yof(n_pairs, 0);
    for (size_t i = 0; i < n_pairs; i++) {
        const b2g_pair &q = pairs[i];
        // the message is only built for a failing check
        auto msg = [i](const char *what) {
            return "pair " + std::to_string(i) + ": " + what;
        };
        B2G_REQUIRE(q.kind == 0 || q.kind == 1, B2G_ERR_ARG,
                    msg("unsupported kind"));
        B2G_REQUIRE(q.a_off + (uint64_t)q.c_m * q.v_m <= llen, B2G_ERR_ARG,
                    msg("A block outside the L buffer"));
        B2G_REQUIRE(q.b_off + (uint64_t)q.c_n * q.v_n <= rlen, B2G_ERR_ARG,
                    msg("B block outside the R buffer"));
        const uint64_t cld = q.c_ld ? q.c_ld : q.c_n;
        const uint64_t vld = q.v_ld ? q.v_ld : q.v_n;
        B2G_REQUIRE(cld >= q.c_n && vld >= q.v_n, B2G_ERR_ARG,
                    msg("row stride below the window width"));
        // the kernels address each operand block with 32-bit byte offsets
        const uint64_t lim = ((uint64_t)1 << 31) / 8 - 2;
        B2G_REQUIRE((uint64_t)q.c_m * cld < lim && (uint64_t)q.v_m * vld < lim &&
                        (uint64_t)q.c_m * q.v_m < lim &&
                        (uint64_t)q.c_n * q.v_n < lim &&
                        (uint64_t)q.c_m * q.v_n < lim,
                    B2G_ERR_ARG, msg("block of 2^28 elements or more"));
        B2G_REQUIRE(q.c_m == 0 || q.c_n == 0 ||
                        q.c_off + (q.c_m - 1) * cld + q.c_n <= p->x_len,
                    B2G_ERR_ARG, msg("C block outside the input vector"));
        if (q.kind == 0) {
            B2G_REQUIRE(q.c_ld == 0 && q.v_ld == 0, B2G_ERR_ARG,
                        msg("whole-block pair with a row stride"));
            const size_t *yi = yidx.find(q.v_off);
            B2G_REQUIRE(yi != nullptr, B2G_ERR_ARG,
                        msg("V block is not an output block"));
            B2G_REQUIRE(yb[*yi].m == q.v_m && yb[*yi].n == q.v_n, B2G_ERR_ARG,
                        msg("V block shape mismatch"));
            yof[i] = (uint32_t)*yi;
            whole.push_back(i);
        } else {
            // the window lies inside one output block, rows of that block
            auto it = std::upper_bound(
                ysorted.begin(), ysorted.end(),
                std::make_pair(q.v_off, std::numeric_limits<size_t>::max()));
            B2G_REQUIRE(it != ysorted.begin(), B2G_ERR_ARG,
                        msg("V window before the first output block"));
            const b2g_block &bl = yb[(it - 1)->second];
            const uint64_t rel = q.v_off - bl.off;
            B2G_REQUIRE(bl.n > 0 && vld == bl.n && rel % bl.n + q.v_n <= bl.n &&
                            rel / bl.n + q.v_m <= bl.m,
                        B2G_ERR_ARG, msg("V window outside its block"));
            windows.push_back(i);
        }
        p->macs += (uint64_t)q.c_m * q.c_n * q.v_n +
                   (uint64_t)q.v_m * q.c_m * q.v_n;
    }
    const uint64_t cap = w_cap_elems();
    // Chunks of the pairs in `idx` whose W scratch fits the cap.  Whole-block
    // pairs: the first chunk writes every output block (zero where no pair
    // lands), later chunks add.  Windowed pairs: their chunks follow and add;
    // tasks are grouped by identical windows (windows of one block are
    // identical or disjoint), so no two tasks of a launch overlap.
    auto build = [&](const std::vector<size_t> &idx, bool win) {
        size_t j0 = 0;
        bool first = !win;
        while (j0 < idx.size() || first) {
            Chunk ch;
            typedef std::tuple<uint64_t, uint64_t, uint32_t, uint32_t, uint32_t,
                               uint32_t>
                WKey;
            struct WHash {
                size_t operator()(const WKey &k) const {
                    uint64_t h = std::get<0>(k) * 0x9E3779B97F4A7C15ull;
                    h ^= std::get<1>(k) + 0x632BE59BD9B4E019ull + (h << 6) +
                         (h >> 2);
                    h ^= ((uint64_t)std::get<2>(k) << 48 |
                          (uint64_t)std::get<3>(k) << 24 | std::get<4>(k)) +
                         (h << 6) + (h >> 2);
                    h ^= std::get<5>(k) + (h << 6) + (h >> 2);
                    return (size_t)h;
                }
            };
            FlatMap<WKey, uint64_t, WHash> wmap(idx.size() - j0); // -> W offset
            // W offset of each pair of the chunk (by position in idx; none
            // for a pair that contributes nothing)
            const uint64_t kNone = ~(uint64_t)0;
            std::vector<uint64_t> pw;
            pw.reserve(idx.size() - j0);
            size_t j1 = j0;
            for (; j1 < idx.size(); j1++) {
                const b2g_pair &q = pairs[idx[j1]];
                if (q.c_m == 0 || q.c_n == 0 || q.v_m == 0 || q.v_n == 0) {
                    pw.push_back(kNone); // contributes nothing
                    continue;
                }
                WKey key(q.b_off, q.c_off, q.conj & 2, q.c_m, q.c_n, q.c_ld);
                const uint64_t *wo = wmap.find(key);
                uint64_t w_off;
                if (wo == nullptr) {
                    const uint64_t sz = (uint64_t)q.c_m * q.v_n;
                    if (ch.w_len + sz > cap && j1 > j0)
                        break;
                    w_off = ch.w_len;
                    wmap.emplace(key, w_off);
                    // phase 1: W = C * op(B)
                    Seg s;
                    s.a_off = q.c_off;
                    s.lda = q.c_ld ? q.c_ld : q.c_n;
                    s.b_off = q.b_off;
                    const bool tb = !(q.conj & 2);
                    s.ldb = tb ? q.c_n : q.v_n;
                    s.K = q.c_n;
                    s.flags = tb ? 2u : 0u;
                    s.alpha = 1.0;
                    const uint32_t si = (uint32_t)ch.s1.size();
                    ch.s1.push_back(s);
                    add_tiles(ch.t1, ch.w_len, q.c_m, q.v_n, q.c_n, si, si + 1,
                              0);
                    ch.w_len += sz;
                } else
                    w_off = *wo;
                pw.push_back(w_off);
            }
            auto emit = [&](uint64_t off, uint32_t m, uint32_t n, uint32_t ldd,
                            const size_t *lst, size_t nl, uint32_t beta) {
                const uint32_t sb = (uint32_t)ch.s2.size();
                for (size_t t = 0; t < nl; t++) {
                    const size_t j = lst[t];
                    const b2g_pair &q = pairs[idx[j]];
                    Seg s;
                    s.a_off = q.a_off;
                    const bool ta = q.conj & 1;
                    s.lda = ta ? q.v_m : q.c_m;
                    s.b_off = pw[j - j0];
                    s.ldb = q.v_n;
                    s.K = q.c_m;
                    s.flags = ta ? 1u : 0u;
                    s.alpha = q.alpha;
                    ch.s2.push_back(s);
                }
                const BlockJob job{off, ldd, m, n, sb,
                                   (uint32_t)ch.s2.size(), beta};
                ch.jobs2.push_back(job);
                add_split_tiles(ch.s2, job, 0, ch.t2, ch.red, ch.p_len);
            };
            // phase 2: this chunk's pairs grouped by output block or window,
            // record order within a group
            std::vector<size_t> js; // positions in idx, grouped
            js.reserve(j1 - j0);
            if (first) {
                // whole blocks: counting sort by output block, then one job
                // per output block (zero where no pair lands)
                std::vector<size_t> start(n_yb + 1, 0);
                for (size_t j = j0; j < j1; j++)
                    if (pw[j - j0] != kNone)
                        start[yof[idx[j]] + 1]++;
                for (size_t b = 0; b < n_yb; b++)
                    start[b + 1] += start[b];
                js.resize(start[n_yb]);
                std::vector<size_t> fill(start.begin(), start.end() - 1);
                for (size_t j = j0; j < j1; j++)
                    if (pw[j - j0] != kNone)
                        js[fill[yof[idx[j]]]++] = j;
                for (size_t b = 0; b < n_yb; b++)
                    emit(yb[b].off, yb[b].m, yb[b].n, yb[b].n,
                         js.data() + start[b], start[b + 1] - start[b], 0);
            } else {
                // windows: ordered by (offset, rows, columns, stride)
                typedef std::tuple<uint64_t, uint32_t, uint32_t, uint32_t> VKey;
                auto vkey = [&](size_t j) {
                    const b2g_pair &q = pairs[idx[j]];
                    return VKey(q.v_off, q.v_m, q.v_n, q.v_ld ? q.v_ld : q.v_n);
                };
                std::vector<std::pair<VKey, size_t>> kv;
                kv.reserve(j1 - j0);
                for (size_t j = j0; j < j1; j++)
                    if (pw[j - j0] != kNone)
                        kv.emplace_back(vkey(j), j);
                std::sort(kv.begin(), kv.end());
                for (auto &e : kv)
                    js.push_back(e.second);
                for (size_t g0 = 0; g0 < kv.size();) {
                    size_t g1 = g0 + 1;
                    while (g1 < kv.size() && kv[g1].first == kv[g0].first)
                        g1++;
                    const VKey &k = kv[g0].first;
                    emit(std::get<0>(k), std::get<1>(k), std::get<2>(k),
                         std::get<3>(k), js.data() + g0, g1 - g0, 1);
                    g0 = g1;
                }
            }
            tm.mark("chunk tasks");
            for (int c = 0; c < kTileConfigs; c++) {
                ch.x1[c] = xcd_order(ch.t1[c], ch.s1, xcd_enabled());
                ch.x2[c] = xcd_order(ch.t2[c], ch.s2, xcd_enabled());
                ch.f1[c] = task_flops(ch.t1[c], ch.s1, c);
                ch.f2[c] = task_flops(ch.t2[c], ch.s2, c);
                p->kflops[B2G_K_PHASE1 + c] += ch.f1[c];
                p->kflops[B2G_K_PHASE2 + c] += ch.f2[c];
            }
            tm.mark("chunk order and flops");
            p->w_max = std::max(p->w_max, ch.w_len);
            p->p_max = std::max(p->p_max, ch.p_len);
            p->chunks.push_back(std::move(ch));
            first = false;
            j0 = j1;
        }
    };
    build(whole, false);
    build(windows, true);
    const int r = upload_chunks(p);
    tm.mark("upload");
    return r;
}

} // namespace b2g

extern "C" int b2g_upload_plan(b2g_ctx ctx, b2g_buf lbuf, b2g_buf rbuf,
                               const b2g_pair *pairs, size_t n_pairs,
                               const b2g_block *yblocks, size_t n_yblocks,
                               size_t x_len, size_t y_len, unsigned flags,
                               b2g_plan *out) {
    B2G_REQUIRE(ctx && lbuf && rbuf && out, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(pairs != nullptr || n_pairs == 0, B2G_ERR_ARG, "NULL pairs");
    B2G_REQUIRE(yblocks != nullptr || n_yblocks == 0, B2G_ERR_ARG,
                "NULL yblocks");
    B2G_REQUIRE(lbuf->dtype == B2G_F64 && rbuf->dtype == B2G_F64, B2G_ERR_ARG,
                "operator buffers must be FP64");
    B2G_HIP(hipSetDevice(ctx->device));
    b2g_plan p = new b2g_plan_s();
    p->ctx = ctx;
    p->lbuf = lbuf;
    p->rbuf = rbuf;
    p->x_len = x_len;
    p->y_len = y_len;
    p->flags = flags;
    int r = compile_plan(p, pairs, n_pairs, yblocks, n_yblocks);
    if (r == B2G_OK)
        p->pairs.assign(pairs, pairs + n_pairs);
    if (r == B2G_OK && (flags & 3)) {
        if (hipMalloc(&p->l32, std::max<size_t>(lbuf->n, 1) * 4) !=
                hipSuccess ||
            hipMalloc(&p->r32, std::max<size_t>(rbuf->n, 1) * 4) !=
                hipSuccess) {
            set_error("b2g_upload_plan: FP32 shadow allocation failed");
            r = B2G_ERR_MEMORY;
        } else if (launch_f64_to_f32((const double *)lbuf->ptr, p->l32,
                                     lbuf->n, ctx->stream) != hipSuccess ||
                   launch_f64_to_f32((const double *)rbuf->ptr, p->r32,
                                     rbuf->n, ctx->stream) != hipSuccess ||
                   hipStreamSynchronize(ctx->stream) != hipSuccess) {
            set_error("b2g_upload_plan: FP32 conversion failed");
            r = B2G_ERR_DEVICE;
        }
    }
    if (r != B2G_OK) {
        b2g_plan_destroy(p);
        return r;
    }
    *out = p;
    return B2G_OK;
}

extern "C" int b2g_plan_refresh(b2g_plan p) {
    B2G_REQUIRE(p, B2G_ERR_ARG, "NULL plan");
    B2G_HIP(hipSetDevice(p->ctx->device));
    if (p->l32) {
        B2G_HIP(launch_f64_to_f32((const double *)p->lbuf->ptr, p->l32,
                                  p->lbuf->n, p->ctx->stream));
        B2G_HIP(launch_f64_to_f32((const double *)p->rbuf->ptr, p->r32,
                                  p->rbuf->n, p->ctx->stream));
    }
    return pk_refresh(p);
}

extern "C" int b2g_plan_destroy(b2g_plan p) {
    if (p == nullptr)
        return B2G_OK;
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->stream);
    for (auto &st : p->stamps) {
        hipEventDestroy(st.start);
        hipEventDestroy(st.stop);
    }
    if (p->dev)
        hipFree(p->dev);
    if (p->l32)
        hipFree(p->l32);
    if (p->r32)
        hipFree(p->r32);
    if (p->w)
        hipFree(p->w);
    if (p->pbuf)
        hipFree(p->pbuf);
    grad_plan_destroy(p->grad);
    pk_destroy(p->pk);
    delete p;
    return B2G_OK;
}

extern "C" int b2g_get_tile_rates_for(int precision, double *rates, int n) {
    B2G_REQUIRE(rates != nullptr && n == kTileConfigs, B2G_ERR_ARG,
                "b2g_get_tile_rates_for: expects room for 9 rates");
    B2G_REQUIRE(precision == B2G_PREC_F64 || precision == B2G_PREC_F32 ||
                    precision == B2G_PREC_BF16X3,
                B2G_ERR_ARG, "unsupported precision");
    const int t = precision == B2G_PREC_BF16X3 ? 1 : 0;
    for (int i = 0; i < n; i++)
        rates[i] = g_rates[t][i];
    return B2G_OK;
}

extern "C" int b2g_set_tile_rates_for(int precision, const double *rates,
                                      int n) {
    B2G_REQUIRE(rates != nullptr && n == kTileConfigs, B2G_ERR_ARG,
                "b2g_set_tile_rates: expects 9 rates");
    B2G_REQUIRE(precision == B2G_PREC_F64 || precision == B2G_PREC_F32 ||
                    precision == B2G_PREC_BF16X3,
                B2G_ERR_ARG, "unsupported precision");
    for (int i = 0; i < n; i++)
        B2G_REQUIRE(rates[i] > 0, B2G_ERR_ARG, "rates must be positive");
    const int t = precision == B2G_PREC_BF16X3 ? 1 : 0;
    bool same = true;
    for (int i = 0; i < n; i++)
        same = same && g_rates[t][i] == rates[i];
    if (same) // the tilings memoised for this table stay valid
        return B2G_OK;
    for (int i = 0; i < n; i++)
        g_rates[t][i] = rates[i];
    g_rates_gen++;
    return B2G_OK;
}

extern "C" int b2g_set_tile_rates(const double *rates, int n) {
    return b2g_set_tile_rates_for(B2G_PREC_F64, rates, n);
}

extern "C" int b2g_plan_set_timing(b2g_plan p, int enable) {
    B2G_REQUIRE(p, B2G_ERR_ARG, "NULL plan");
    p->timing = enable != 0;
    return B2G_OK;
}

extern "C" int b2g_plan_timing(b2g_plan p, double *ms, uint64_t *n_launch) {
    B2G_REQUIRE(p && ms && n_launch, B2G_ERR_ARG, "NULL arg");
    for (int k = 0; k < B2G_NKERNELS; k++)
        ms[k] = 0, n_launch[k] = 0;
    int rc = B2G_OK;
    for (auto &st : p->stamps) {
        float t = 0;
        if (rc == B2G_OK) {
            hipError_t e = hipEventSynchronize(st.stop);
            if (e == hipSuccess)
                e = hipEventElapsedTime(&t, st.start, st.stop);
            if (e != hipSuccess) {
                b2g::set_error(std::string("b2g_plan_timing: ") +
                               hipGetErrorString(e));
                rc = B2G_ERR_DEVICE;
            }
        }
        ms[st.kernel] += t;
        n_launch[st.kernel]++;
        hipEventDestroy(st.start);
        hipEventDestroy(st.stop);
    }
    p->stamps.clear();
    return rc;
}

extern "C" int b2g_plan_kernel_flops(b2g_plan p, double *flops) {
    B2G_REQUIRE(p && flops, B2G_ERR_ARG, "NULL arg");
    for (int k = 0; k < B2G_NKERNELS; k++)
        flops[k] = p->kflops[k];
    return B2G_OK;
}

extern "C" int b2g_plan_kernel_flops_for(b2g_plan p, int precision,
                                         double *flops) {
    B2G_REQUIRE(p && flops, B2G_ERR_ARG, "NULL arg");
    if (precision != B2G_PREC_BF16X3)
        return b2g_plan_kernel_flops(p, flops);
    B2G_HIP(hipSetDevice(p->ctx->device));
    const int r = pk_build(p);
    if (r != B2G_OK)
        return r;
    for (int k = 0; k < B2G_NKERNELS; k++)
        flops[k] = 0;
    for (const PkChunk &kc : p->pk->chunks)
        for (int c = 0; c < kTileConfigs; c++) {
            flops[B2G_K_PHASE1 + c] += kc.f1[c];
            flops[B2G_K_PHASE2 + c] += kc.f2[c];
        }
    return B2G_OK;
}

extern "C" int b2g_plan_macs(b2g_plan p, uint64_t *out) {
    B2G_REQUIRE(p && out, B2G_ERR_ARG, "NULL arg");
    *out = p->macs;
    return B2G_OK;
}

extern "C" int b2g_plan_stats(b2g_plan p, uint64_t *n1, uint64_t *n2,
                              uint64_t *scratch) {
    B2G_REQUIRE(p, B2G_ERR_ARG, "NULL plan");
    uint64_t a = 0, b = 0;
    for (auto &ch : p->chunks)
        for (int c = 0; c < kTileConfigs; c++)
            a += ch.t1[c].size(), b += ch.t2[c].size();
    if (n1)
        *n1 = a;
    if (n2)
        *n2 = b;
    if (scratch)
        *scratch = p->w_max;
    return B2G_OK;
}

namespace b2g {

// scratch (W, split-K partials) for n_rhs right-hand sides at `precision`;
// reallocates only when growing (not allowed while a stream is captured)
int plan_reserve(b2g_plan p, int n_rhs, int precision) {
    b2g_ctx ctx = p->ctx;
    const size_t es = precision == B2G_PREC_F64 ? 8 : 4;
    const size_t need = std::max<uint64_t>(p->w_max, 1) * n_rhs * es;
    if (precision == B2G_PREC_BF16X3) {
        // packed planes replace the W scratch
        int r = pk_build(p);
        if (r == B2G_OK)
            r = pk_reserve(p, n_rhs);
        if (r != B2G_OK)
            return r;
    } else if (need > p->w_bytes) {
        if (p->w) {
            B2G_HIP(hipStreamSynchronize(ctx->stream));
            B2G_HIP(hipFree(p->w));
            p->w = nullptr;
            p->w_bytes = 0;
        }
        B2G_HIP(hipMalloc(&p->w, need));
        p->w_bytes = need;
        p->scratch_gen++;
    }
    const uint64_t pmax =
        std::max<uint64_t>(p->p_max, p->pk != nullptr ? p->pk->p_max : 0);
    const size_t pneed = std::max<uint64_t>(pmax, 1) * n_rhs * 8;
    if (pneed > p->p_bytes) {
        if (p->pbuf) {
            B2G_HIP(hipStreamSynchronize(ctx->stream));
            B2G_HIP(hipFree(p->pbuf));
            p->pbuf = nullptr;
            p->p_bytes = 0;
        }
        B2G_HIP(hipMalloc(&p->pbuf, pneed));
        p->p_bytes = pneed;
        p->scratch_gen++;
    }
    return B2G_OK;
}

} // namespace b2g

extern "C" int b2g_plan_reserve(b2g_plan p, int n_rhs, int precision) {
    B2G_REQUIRE(p, B2G_ERR_ARG, "NULL plan");
    B2G_REQUIRE(n_rhs >= 1 && n_rhs <= 65535, B2G_ERR_ARG, "bad n_rhs");
    B2G_REQUIRE(precision == B2G_PREC_F64 || precision == B2G_PREC_F32 ||
                    precision == B2G_PREC_BF16X3,
                B2G_ERR_ARG, "unsupported precision");
    B2G_HIP(hipSetDevice(p->ctx->device));
    return plan_reserve(p, n_rhs, precision);
}

namespace b2g {

// One H*psi with the FP64 / FP32 W scratch at `W` and the split-K partials
// at `P` (for split-BF16 the packed planes of the plan); allocates nothing
// itself.
static int hpsi_run(b2g_ctx ctx, b2g_plan p, const double *x, double *y,
                    int n_rhs, double scale, int accumulate, int precision,
                    void *W, double *P) {
    const int grid = ctx->num_cu * 8;
    // brackets one launch on stream `st` with timing events when enabled
    auto timed = [&](int kernel, hipStream_t st,
                     auto &&launch) -> hipError_t {
        if (!p->timing)
            return launch(st);
        b2g_plan_s::Stamp sp{kernel, nullptr, nullptr};
        hipError_t e = hipEventCreate(&sp.start);
        if (e == hipSuccess)
            e = hipEventCreate(&sp.stop);
        if (e == hipSuccess)
            e = hipEventRecord(sp.start, st);
        if (e == hipSuccess)
            e = launch(st);
        if (e == hipSuccess)
            e = hipEventRecord(sp.stop, st);
        if (e == hipSuccess)
            p->stamps.push_back(sp);
        else {
            if (sp.start)
                hipEventDestroy(sp.start);
            if (sp.stop)
                hipEventDestroy(sp.stop);
        }
        return e;
    };
    // One phase: a launch per tile configuration with tasks.  The tasks of
    // a phase write disjoint tiles, so the launches are independent: the
    // largest runs on the context stream, the others on the auxiliary
    // streams (forked from and joined back to the context stream), where
    // they fill the large launch's idle compute units.
    auto phase = [&](const size_t *cnt, const double *fl, int kbase,
                     auto &&launch) -> hipError_t {
        int order[kTileConfigs], n = 0;
        for (int c = 0; c < kTileConfigs; c++)
            if (cnt[c] != 0)
                order[n++] = c;
        std::sort(order, order + n,
                  [fl](int a, int b) { return fl[a] > fl[b]; });
        // overlap only for one right-hand side (several fill the GPU by
        // themselves: -2% with overlap at 4 RHS, +8% at 1 RHS on the n=40
        // M=3000 middle site); with per-kernel timing the launches run back
        // to back, so that each event pair brackets one kernel alone.  A
        // product captured into a graph (the Davidson iteration) runs on one
        // stream: each instantiated graph with forked branches kept about
        // 0.27 GB of device memory after its destruction (measured on the
        // n=26 M=1000 synthetic site, every effective Hamiltonian of a sweep
        // instantiating one).
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(ctx->stream, &cap) != hipSuccess)
            cap = hipStreamCaptureStatusNone;
        const int naux = (p->timing || n_rhs > 1 ||
                          cap != hipStreamCaptureStatusNone)
                             ? 0
                             : std::min<int>((int)ctx->aux.size(), n - 1);
        hipError_t e = hipSuccess;
        if (naux > 0) {
            e = hipEventRecord(ctx->fork, ctx->stream);
            for (int i = 0; e == hipSuccess && i < naux; i++)
                e = hipStreamWaitEvent(ctx->aux[i], ctx->fork, 0);
        }
        for (int i = 0; e == hipSuccess && i < n; i++) {
            const int c = order[i];
            hipStream_t st = (i == 0 || naux == 0)
                                 ? ctx->stream
                                 : ctx->aux[(i - 1) % naux];
            e = timed(kbase + c, st,
                      [&](hipStream_t s) { return launch(c, s); });
        }
        for (int i = 0; e == hipSuccess && i < naux; i++) {
            e = hipEventRecord(ctx->join[i], ctx->aux[i]);
            if (e == hipSuccess)
                e = hipStreamWaitEvent(ctx->stream, ctx->join[i], 0);
        }
        return e;
    };
    const PkPlan *pk = precision == B2G_PREC_BF16X3 ? p->pk : nullptr;
    // split-K partials per right-hand side of the packed plan's tiling
    const uint64_t pstride = pk != nullptr ? pk->p_max : 0;
    typedef const __bf16 *cbf;
    if (pk != nullptr)
        B2G_HIP(timed(B2G_K_PACK, ctx->stream, [&](hipStream_t st) {
            return launch_pack_bf16(x, (__bf16 *)pk->xh(), (__bf16 *)pk->xl(),
                                    pk->d_xmats, pk->n_xmats, pk->n_xgroups,
                                    p->x_len,
                                    pk->x_len, n_rhs, grid, st);
        }));
    for (size_t ic = 0; ic < p->chunks.size(); ic++) {
        const Chunk &ch = p->chunks[ic];
        // chunk 0 overwrites unless accumulating; later chunks' tasks add
        const int beta0 = ic == 0 ? (accumulate ? 1 : 0) : 1;
        if (pk != nullptr) {
            const PkChunk &kc = pk->chunks[ic];
            B2G_HIP(phase(kc.n1, kc.f1, B2G_K_PHASE1,
                          [&](int c, hipStream_t st) {
                return launch_seg_gemm_pk<0>(
                    c, (cbf)pk->xh(), (cbf)pk->xl(), (cbf)pk->rh(),
                    (cbf)pk->rl(), pk->wh(), pk->wl(), kc.d_t1[c],
                    (int)kc.n1[c], kc.d_s1, pk->x_len, 0, pk->w_max, 0, 1.0,
                    0, n_rhs, grid, st, kc.x1[c]);
            }));
            B2G_HIP(phase(kc.n2, kc.f2, B2G_K_PHASE2,
                          [&](int c, hipStream_t st) {
                return launch_seg_gemm_pk<1>(
                    c, (cbf)pk->lh(), (cbf)pk->ll(), (cbf)pk->wh(),
                    (cbf)pk->wl(), y, P, kc.d_t2[c], (int)kc.n2[c], kc.d_s2,
                    0, pk->w_max, p->y_len, pstride, scale, beta0, n_rhs,
                    grid, st, kc.x2[c]);
            }));
            if (kc.n_red > 0)
                B2G_HIP(timed(B2G_K_SPLITK_REDUCE, ctx->stream,
                              [&](hipStream_t st) {
                    return launch_splitk_reduce<double>(
                        y, P, kc.d_red, kc.n_red, p->y_len, pstride, beta0,
                        n_rhs, grid, st);
                }));
            continue;
        }
        // phase 1, phase 2, then the split-K reduction
        size_t n1[kTileConfigs], n2[kTileConfigs];
        for (int c = 0; c < kTileConfigs; c++)
            n1[c] = ch.t1[c].size(), n2[c] = ch.t2[c].size();
        B2G_HIP(phase(n1, ch.f1, B2G_K_PHASE1, [&](int c, hipStream_t st) {
            const int nt = (int)ch.t1[c].size();
            if (precision == B2G_PREC_F64)
                return launch_seg_gemm<double, double, double, double, 1>(
                    c, x, (const double *)p->rbuf->ptr, (double *)W,
                    (double *)W, ch.d_t1[c], nt, ch.d_s1, p->x_len, 0,
                    p->w_max, 0, 1.0, 0, n_rhs, grid, st, ch.x1[c]);
            return launch_seg_gemm<double, float, float, float, 1>(
                c, x, p->r32, (float *)W, (float *)W, ch.d_t1[c], nt,
                ch.d_s1, p->x_len, 0, p->w_max, 0, 1.0, 0, n_rhs, grid, st, ch.x1[c]);
        }));
        B2G_HIP(phase(n2, ch.f2, B2G_K_PHASE2, [&](int c, hipStream_t st) {
            const int nt = (int)ch.t2[c].size();
            if (precision == B2G_PREC_F64)
                return launch_seg_gemm<double, double, double, double, 2>(
                    c, (const double *)p->lbuf->ptr, (const double *)W, y,
                    P, ch.d_t2[c], nt, ch.d_s2, 0, p->w_max, p->y_len,
                    p->p_max, scale, beta0, n_rhs, grid, st, ch.x2[c]);
            return launch_seg_gemm<float, float, double, float, 2>(
                c, p->l32, (const float *)W, y, P, ch.d_t2[c], nt,
                ch.d_s2, 0, p->w_max, p->y_len, p->p_max, scale, beta0, n_rhs,
                grid, st, ch.x2[c]);
        }));
        const int nr = (int)ch.red.size();
        if (nr > 0)
            B2G_HIP(timed(B2G_K_SPLITK_REDUCE, ctx->stream, [&](hipStream_t st) {
                return launch_splitk_reduce<double>(y, P, ch.d_red, nr,
                                                    p->y_len, p->p_max, beta0,
                                                    n_rhs, grid, st);
            }));
    }
    return B2G_OK;
}

} // namespace b2g

extern "C" int b2g_hpsi(b2g_ctx ctx, b2g_plan p, const double *x, double *y,
                        int n_rhs, double scale, int accumulate,
                        int precision) {
    B2G_REQUIRE(ctx && p && x && y, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(n_rhs >= 1 && n_rhs <= 65535, B2G_ERR_ARG, "bad n_rhs");
    B2G_REQUIRE(precision == B2G_PREC_F64 || precision == B2G_PREC_F32 ||
                    precision == B2G_PREC_BF16X3,
                B2G_ERR_ARG, "unsupported precision");
    B2G_REQUIRE(precision != B2G_PREC_F32 || p->l32 != nullptr, B2G_ERR_STATE,
                "plan was uploaded without FP32 shadows (flags bit 0 or 1)");
    B2G_HIP(hipSetDevice(ctx->device));
    {
        const int r = plan_reserve(p, n_rhs, precision);
        if (r != B2G_OK)
            return r;
    }
    return hpsi_run(ctx, p, x, y, n_rhs, scale, accumulate, precision, p->w,
                    (double *)p->pbuf);
}

extern "C" int b2g_hpsi_noalloc(b2g_ctx ctx, b2g_plan p, const double *x,
                                double *y, int n_rhs, double scale,
                                int accumulate, int precision) {
    B2G_REQUIRE(ctx && p && x && y, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(n_rhs >= 1 && n_rhs <= 65535, B2G_ERR_ARG, "bad n_rhs");
    B2G_REQUIRE(precision == B2G_PREC_F64 || precision == B2G_PREC_F32,
                B2G_ERR_ARG,
                "b2g_hpsi_noalloc: FP64 or FP32 only (split-BF16 reads packed "
                "planes the plan allocates)");
    B2G_REQUIRE(precision != B2G_PREC_F32 || p->l32 != nullptr, B2G_ERR_STATE,
                "plan was uploaded without FP32 shadows (flags bit 0 or 1)");
    B2G_REQUIRE(p->ctx == ctx, B2G_ERR_ARG, "plan belongs to another context");
    const size_t es = precision == B2G_PREC_F64 ? 8 : 4;
    // W windows first, the split-K partials 256-byte aligned after them
    const uint64_t w_bytes =
        (std::max<uint64_t>(p->w_max, 1) * n_rhs * es + 255) & ~(uint64_t)255;
    const uint64_t p_bytes = std::max<uint64_t>(p->p_max, 1) * n_rhs * 8;
    if (w_bytes + p_bytes > kNoAllocScratchBytes) {
        set_error("b2g_hpsi_noalloc: the product needs " +
                  std::to_string(w_bytes + p_bytes) +
                  " bytes of scratch; the allocation-free entry holds " +
                  std::to_string(kNoAllocScratchBytes));
        return B2G_ERR_ARG;
    }
    char *base = (char *)scratch_ptr(ctx->device);
    B2G_REQUIRE(base != nullptr, B2G_ERR_STATE,
                "no scratch address for this device");
    B2G_HIP(hipSetDevice(ctx->device));
    // nothing in the scratch needs to start at zero: phase 1 writes every
    // W element phase 2 reads, and each split-K partial is written before
    // the reduction reads it
    return hpsi_run(ctx, p, x, y, n_rhs, scale, accumulate, precision, base,
                    (double *)(base + w_bytes));
}

// ----------------------------------------------------------------- diagonal
extern "C" int b2g_diag(b2g_ctx ctx, b2g_buf lbuf, b2g_buf rbuf,
                        const b2g_diag_term *terms, size_t n_terms,
                        const b2g_block *dblocks, size_t n_dblocks, double *d,
                        size_t d_len) {
    B2G_REQUIRE(ctx && lbuf && rbuf && d, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(terms != nullptr || n_terms == 0, B2G_ERR_ARG, "NULL terms");
    B2G_REQUIRE(lbuf->dtype == B2G_F64 && rbuf->dtype == B2G_F64, B2G_ERR_ARG,
                "operator buffers must be FP64");
    FlatMap<uint64_t, size_t, U64Hash> didx(n_dblocks);
    for (size_t i = 0; i < n_dblocks; i++) {
        B2G_REQUIRE(dblocks[i].off + (uint64_t)dblocks[i].m * dblocks[i].n <=
                        d_len,
                    B2G_ERR_ARG, "diagonal block outside the output");
        *didx.emplace(dblocks[i].off, i).first = i;
    }
    // Work pieces: each block is cut into runs of `rows` rows (about 2048
    // elements), so that large blocks spread over many workgroups; a piece
    // lists the block's terms whose row window meets it, in term order (the
    // sum of every element runs over the same terms in the same order as
    // for the whole block).
    std::vector<uint32_t> rows(n_dblocks), pfirst(n_dblocks + 1, 0);
    for (size_t b = 0; b < n_dblocks; b++) {
        const uint32_t n = std::max<uint32_t>(dblocks[b].n, 1);
        rows[b] = std::max<uint32_t>(1, 2048 / n);
        pfirst[b + 1] = pfirst[b] + (dblocks[b].m + rows[b] - 1) / rows[b];
    }
    const size_t npieces = pfirst[n_dblocks];
    std::vector<uint32_t> tblock(n_terms);
    std::vector<size_t> cnt(npieces + 1, 0);
    for (size_t i = 0; i < n_terms; i++) {
        const b2g_diag_term &t = terms[i];
        const size_t *bi = didx.find(t.d_off);
        B2G_REQUIRE(bi != nullptr, B2G_ERR_ARG,
                    "diagonal term targets no output block");
        const b2g_block &blk = dblocks[*bi];
        B2G_REQUIRE((uint64_t)t.r0 + t.a_n <= blk.m &&
                        (uint64_t)t.c0 + t.b_n <= blk.n,
                    B2G_ERR_ARG, "diagonal term window outside its block");
        B2G_REQUIRE(t.a_off + (uint64_t)t.a_n * t.a_n <= lbuf->n &&
                        t.b_off + (uint64_t)t.b_n * t.b_n <= rbuf->n,
                    B2G_ERR_ARG, "diagonal term outside operator buffers");
        tblock[i] = (uint32_t)*bi;
        if (t.a_n == 0 || t.b_n == 0)
            continue;
        const uint32_t rb = rows[*bi];
        for (uint32_t q = t.r0 / rb; q <= (t.r0 + t.a_n - 1) / rb; q++)
            cnt[pfirst[*bi] + q + 1]++;
    }
    for (size_t q = 0; q < npieces; q++)
        cnt[q + 1] += cnt[q];
    std::vector<DiagBlock> blocks(npieces);
    for (size_t b = 0; b < n_dblocks; b++)
        for (uint32_t q = 0; q < pfirst[b + 1] - pfirst[b]; q++) {
            DiagBlock &pc = blocks[pfirst[b] + q];
            pc.row0 = q * rows[b];
            pc.m = std::min<uint32_t>(rows[b], dblocks[b].m - pc.row0);
            pc.n = dblocks[b].n;
            pc.d_off = dblocks[b].off + (uint64_t)pc.row0 * dblocks[b].n;
            pc.t_begin = (uint32_t)cnt[pfirst[b] + q];
            pc.t_end = (uint32_t)cnt[pfirst[b] + q + 1];
            pc.pad = 0;
        }
    B2G_REQUIRE(cnt[npieces] < ((uint64_t)1 << 32), B2G_ERR_ARG,
                "diagonal: too many term pieces");
    std::vector<DiagTermDev> dterms(cnt[npieces]);
    std::vector<size_t> fill(cnt.begin(), cnt.end() - 1);
    for (size_t i = 0; i < n_terms; i++) {
        const b2g_diag_term &t = terms[i];
        if (t.a_n == 0 || t.b_n == 0)
            continue;
        DiagTermDev d;
        d.a_off = t.a_off;
        d.b_off = t.b_off;
        d.a_n = t.a_n;
        d.b_n = t.b_n;
        d.r0 = t.r0;
        d.c0 = t.c0;
        d.alpha = t.alpha;
        const uint32_t b = tblock[i], rb = rows[b];
        for (uint32_t q = t.r0 / rb; q <= (t.r0 + t.a_n - 1) / rb; q++)
            dterms[fill[pfirst[b] + q]++] = d;
    }
    // costliest pieces first for the workgroups' grid-stride loop
    std::vector<std::pair<uint64_t, size_t>> cost(npieces);
    for (size_t q = 0; q < npieces; q++)
        cost[q] = std::make_pair(
            ~((uint64_t)blocks[q].m * blocks[q].n *
              (blocks[q].t_end - blocks[q].t_begin + 1)),
            q);
    std::sort(cost.begin(), cost.end());
    {
        std::vector<DiagBlock> sorted(npieces);
        for (size_t q = 0; q < npieces; q++)
            sorted[q] = blocks[cost[q].second];
        blocks.swap(sorted);
    }
    B2G_HIP(hipSetDevice(ctx->device));
    const size_t nb1 = blocks.size() * sizeof(DiagBlock),
                 nb2 = dterms.size() * sizeof(DiagTermDev);
    char *dev = nullptr;
    B2G_HIP(hipMalloc(&dev, nb1 + nb2 + 16));
    if (nb1)
        B2G_HIP(hipMemcpyAsync(dev, blocks.data(), nb1, hipMemcpyHostToDevice,
                               ctx->stream));
    if (nb2)
        B2G_HIP(hipMemcpyAsync(dev + nb1, dterms.data(), nb2,
                               hipMemcpyHostToDevice, ctx->stream));
    hipError_t e = launch_fill(0.0, d, d_len, ctx->stream);
    if (e == hipSuccess)
        e = launch_diag((const double *)lbuf->ptr, (const double *)rbuf->ptr,
                        d, (const DiagBlock *)dev, (int)blocks.size(),
                        (const DiagTermDev *)(dev + nb1), ctx->num_cu * 4,
                        ctx->stream);
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    hipFree(dev);
    B2G_HIP(e);
    B2G_HIP(e2);
    return B2G_OK;
}

// ------------------------------------------------------------- vector ops
extern "C" int b2g_vec_dot(b2g_ctx ctx, const double *x, const double *y,
                           size_t n, double *out_host) {
    B2G_REQUIRE(ctx && x && y && out_host, B2G_ERR_ARG, "NULL arg");
    B2G_HIP(launch_dot(x, y, n, ctx->red, ctx->red + ctx->red_len - 8,
                       ctx->stream));
    B2G_HIP(hipMemcpyAsync(ctx->red_host, ctx->red + ctx->red_len - 8,
                           sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    B2G_HIP(hipStreamSynchronize(ctx->stream));
    *out_host = ctx->red_host[0];
    return B2G_OK;
}

extern "C" int b2g_vec_axpy(b2g_ctx ctx, double a, const double *x, double *y,
                            size_t n) {
    B2G_REQUIRE(ctx && x && y, B2G_ERR_ARG, "NULL arg");
    B2G_HIP(launch_axpy(a, x, y, n, ctx->stream));
    return B2G_OK;
}

extern "C" int b2g_vec_scale(b2g_ctx ctx, double a, double *x, size_t n) {
    B2G_REQUIRE(ctx && x, B2G_ERR_ARG, "NULL arg");
    B2G_HIP(launch_scale(a, x, n, ctx->stream));
    return B2G_OK;
}

extern "C" int b2g_vec_copy(b2g_ctx ctx, const double *x, double *y,
                            size_t n) {
    B2G_REQUIRE(ctx && x && y, B2G_ERR_ARG, "NULL arg");
    B2G_HIP(launch_copy(x, y, n, ctx->stream));
    return B2G_OK;
}
