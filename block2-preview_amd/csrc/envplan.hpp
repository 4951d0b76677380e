// Host-side plans of environment blocking and renormalisation: fused bases,
// tensor-product connection lists, the window records of the contraction
// kernel and the block lists of rotations, two-site contractions and
// fused-form swaps.
#pragma once

#include "plan.hpp"

#include <cstdint>
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace b2g {

// StateInfo (src/core/state_info.hpp:59-): quanta ascending, state counts.
struct StateSet {
    std::vector<uint64_t> q;
    std::vector<uint32_t> n;
    int size() const { return (int)q.size(); }
    int find(uint64_t x) const;
    uint64_t total() const;
};

// StateInfo::ConnectionInfo of a fused basis c = a (x) b: the (a, b) state
// pairs of fused state ic are ij[acc[ic] .. acc[ic + 1]), in (a, b) order.
struct StateConn {
    std::vector<uint32_t> acc;
    std::vector<std::pair<uint32_t, uint32_t>> ij;
};

// StateInfo::get_connection_info (state_info.hpp:283-312); every quantum of
// c must arise from some pair.  Returns false (and sets err) otherwise.
bool state_connections(Symm s, const StateSet &a, const StateSet &b,
                       const StateSet &c, StateConn &out, std::string &err);

// StateInfo::tensor_product(a, b, cref) (state_info.hpp:231-254): the
// quanta of cref, each with the summed dimensions of the (a, b) pairs
// coupling to it, empty ones removed.
StateSet state_tensor_product(Symm s, const StateSet &a, const StateSet &b,
                              const StateSet &cref);

// ConnectionInfo::initialize_tp (src/core/sparse_matrix.hpp:294-422): the
// tensor-product connection list of an operator block-set `cinfo` (delta
// quantum cdq) on the fused bases bra = bra_a (x) bra_b, ket = ket_a (x)
// ket_b, for the sub-label classes subdq, with factor operator infos
// ainfos / binfos (sorted by delta quantum).  An entry (ia, ib, ic, stride,
// factor) adds factor * op(A[ia]) (x) op(B[ib]) to block ic at element
// offset stride.
void build_tp_connections(Connections &ci, Symm s, const SU2Coupling &cg,
                          uint64_t cdq, const std::vector<SubLabel> &subdq,
                          const StateSet &bra, const StateSet &ket,
                          const StateSet &bra_a, const StateSet &bra_b,
                          const StateSet &ket_a, const StateSet &ket_b,
                          const StateConn &bra_c, const StateConn &ket_c,
                          const std::vector<InfoRef> &ainfos,
                          const std::vector<InfoRef> &binfos,
                          const BlockInfo &cinfo);

// One source of the window-sum kernel:
//   C_window += scale * op(A) (x) op(B)
// A is an am x an matrix (row stride lda) at a_off of the A buffer, B a
// bm x bn matrix (row stride ldb) at b_off of the B buffer; conj bit 0
// transposes A, bit 1 B.  The window is (am' bm') x (an' bn') with op
// dimensions am' x an', bm' x bn'.
struct WinRec {
    uint64_t a_off, b_off;
    uint32_t am, an, lda;
    uint32_t bm, bn, ldb;
    uint32_t conj;
    uint32_t pad;
    double scale;
};

// A destination window (rows x cols, row stride ldc, at c_off of the C
// buffer) with its sources [rec_begin, rec_end), summed in that order.
// Windows of different groups are disjoint.
struct WinGroup {
    uint64_t c_off;
    uint32_t ldc, rows, cols;
    uint32_t rec_begin, rec_end;
    uint32_t pad;
};

// allocator whose resize leaves trivially constructible elements
// uninitialised (record arrays of hundreds of millions of entries are
// written right after sizing; zero-filling them first doubles the traffic)
template <typename T> struct NoInitAlloc : std::allocator<T> {
    template <typename U> struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <typename U> NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <typename U> void construct(U *p) noexcept {
        ::new ((void *)p) U;
    }
    template <typename U, typename... Args>
    void construct(U *p, Args &&...args) {
        ::new ((void *)p) U(std::forward<Args>(args)...);
    }
};

// one record with its destination window, as the plan builders emit them
// (trivially default-constructible: sized arrays of them are not zeroed)
struct WinItem {
    WinGroup first;
    WinRec second;
    WinItem() = default;
    WinItem(const WinGroup &g, const WinRec &r) : first(g), second(r) {}
};
static_assert(std::is_trivially_default_constructible<WinItem>::value,
              "window items are sized without zero-filling");
using WinItems = std::vector<WinItem, NoInitAlloc<WinItem>>;

struct WinPlan {
    std::vector<WinGroup> groups;
    std::vector<WinRec, NoInitAlloc<WinRec>> recs;
    uint64_t c_len = 0;     // elements of the C buffer the plan writes into
    uint64_t a_need = 0;    // minimum element counts of the A / B buffers
    uint64_t b_need = 0;
    uint64_t src_elems = 0; // elements read from A and B (algorithmic)
    uint64_t dst_elems = 0; // elements written (sum of window sizes)
    // group the records by destination window (equal c_off and shape), in
    // first-appearance order of the windows and record order within one;
    // fails if two different windows overlap
    bool finish(WinItems &items,
                std::string &err);
};

} // namespace b2g
