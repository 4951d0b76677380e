// Shared host-side definitions of libb2g: error state, context, buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
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

// Wall time of the stages of a host-side setup (effective Hamiltonian,
// plan, packed plan): with B2G_TIME_SETUP set, mark() prints the time since
// the previous mark to stderr.
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
};

struct b2g_buf_s {
    b2g_ctx ctx = nullptr;
    void *ptr = nullptr;
    size_t n = 0;
    int dtype = B2G_F64;
};
