// Shared host-side definitions of libb2g: error state, context, buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>
#include <functional>
#include <vector>

#include "../../include/b2g.h"

namespace b2g {

void set_error(const std::string &msg);

#define B2G_HIP(call)                                                          \
    do {                                                                       \
        hipError_t _e = (call);                                                \
        if (_e != hipSuccess) {                                                \
            ::b2g::set_error(std::string(#call) + ": " +                       \
                             hipGetErrorString(_e));                           \
            return B2G_ERR_DEVICE;                                             \
        }                                                                      \
    } while (0)

#define B2G_REQUIRE(cond, code, msg)                                           \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::b2g::set_error(msg);                                             \
            return code;                                                       \
        }                                                                      \
    } while (0)

// Open-addressing hash map of the plan builders: linear probing in a
// power-of-two table kept at most half full, no erase.  One plan performs
// 10^4-10^5 lookups; a flat table avoids the per-node allocation and pointer
// chasing of std::unordered_map.  Pointers returned by find() / emplace()
// are invalidated by the next emplace().
template <typename K, typename V, typename H> class FlatMap {
  public:
    explicit FlatMap(size_t expect = 0) { grow(expect); }
    size_t size() const { return n_; }
    void reserve(size_t expect) { grow(expect); }
    V *find(const K &k) {
        for (size_t i = slot(k);; i = (i + 1) & mask_) {
            if (!used_[i])
                return nullptr;
            if (keys_[i] == k)
                return &vals_[i];
        }
    }
    // the value of k, inserted as v when absent; second: inserted
    std::pair<V *, bool> emplace(const K &k, const V &v) {
        if (2 * (n_ + 1) > used_.size())
            grow(n_ + 1);
        size_t i = slot(k);
        for (; used_[i]; i = (i + 1) & mask_)
            if (keys_[i] == k)
                return std::make_pair(&vals_[i], false);
        used_[i] = 1;
        keys_[i] = k;
        vals_[i] = v;
        n_++;
        return std::make_pair(&vals_[i], true);
    }

  private:
    size_t slot(const K &k) const {
        uint64_t h = (uint64_t)H()(k); // murmur3 finaliser over any hash
        h ^= h >> 33;
        h *= 0xFF51AFD7ED558CCDull;
        h ^= h >> 33;
        h *= 0xC4CEB9FE1A85EC53ull;
        h ^= h >> 33;
        return (size_t)h & mask_;
    }
    void grow(size_t expect) {
        size_t cap = 16;
        while (cap < 2 * expect)
            cap <<= 1;
        if (cap <= used_.size())
            return;
        std::vector<K> ok(cap);
        std::vector<V> ov(cap);
        std::vector<uint8_t> ou(cap, 0);
        ok.swap(keys_);
        ov.swap(vals_);
        ou.swap(used_);
        mask_ = cap - 1;
        n_ = 0;
        for (size_t i = 0; i < ou.size(); i++)
            if (ou[i])
                emplace(ok[i], ov[i]);
    }
    std::vector<K> keys_;
    std::vector<V> vals_;
    std::vector<uint8_t> used_;
    size_t mask_ = 0, n_ = 0;
};

struct U64Hash {
    size_t operator()(uint64_t v) const { return (size_t)v; }
};

// Wall time of the stages of a host-side setup (effective Hamiltonian,
// plan, packed plan): with B2G_TIME_SETUP set, mark() prints the time since
// the previous mark to stderr.
// Runs jobs[0] on the calling thread and the other jobs on persistent host
// worker threads (created on first use; their thread-local tiling memos stay
// warm from one plan build to the next) and returns when all have finished.
// Jobs report errors through their captures and must not throw, and must
// not call run_host_jobs themselves (waiting jobs could hold every worker).
void run_host_jobs(std::vector<std::function<void()>> &jobs);
// memcpy of a large host range split over the host workers (staging copies
// of plan uploads: one thread moves about 6-10 GB/s)
void parallel_memcpy(void *dst, const void *src, size_t bytes);
// number of host jobs (at most 8) for `items` units of plan-building work
// of about `per_job` units each (B2G_PLAN_JOB_ITEMS divides per_job down to
// that many items, for small plans split on purpose)
size_t plan_jobs(uint64_t items, uint64_t per_job);

class SetupTimer {
  public:
    explicit SetupTimer(const char *what)
        : what_(what), on_(getenv("B2G_TIME_SETUP") != nullptr),
          t_(std::chrono::steady_clock::now()) {}
    void mark(const char *stage) {
        if (!on_)
            return;
        const auto t = std::chrono::steady_clock::now();
        fprintf(stderr, "b2g setup %s: %s %.2f ms\n", what_, stage,
                std::chrono::duration<double, std::milli>(t - t_).count());
        t_ = t;
    }

  private:
    const char *what_;
    bool on_;
    std::chrono::steady_clock::time_point t_;
};

} // namespace b2g

struct b2g_ctx_s {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    // auxiliary streams (default priority) for the independent launches of
    // one H*psi phase, forked from and joined back to `stream` with events
    std::vector<hipStream_t> aux;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    // reusable device scratch for reductions (deterministic two-stage)
    double *red = nullptr;
    size_t red_len = 0;
    double *red_host = nullptr; // pinned
    // pinned staging buffer of the plan uploads (grown on demand; every
    // upload waits for its copy before returning)
    char *stage = nullptr;
    size_t stage_len = 0;
    // device blocks of destroyed window plans, kept for the next ones
    // (sweeps create and destroy several plans per step); freed with the
    // context
    std::vector<std::pair<size_t, void *>> win_pool;
};

struct b2g_buf_s {
    b2g_ctx ctx = nullptr;
    void *ptr = nullptr;
    size_t n = 0;
    int dtype = B2G_F64;
    bool owned = true; // false: a view of memory it must not free
};
