// Host-side plan builder: block-sparse layouts, connection lists and the
// recorded pair list of one effective Hamiltonian.
#pragma once

#include "symmetry.hpp"

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace b2g {

// Block-sparse layout (SparseMatrixInfo, src/core/sparse_matrix.hpp:47-877):
// blocks sorted by composite quantum label; block i is a row-major
// n_bra[i] x n_ket[i] matrix at element offset shift[i].
struct BlockInfo {
    uint64_t dq = 0;
    bool is_fermion = false;
    std::vector<uint64_t> quanta;
    std::vector<uint32_t> nbra, nket;
    std::vector<uint64_t> shift;
    int n() const { return (int)quanta.size(); }
    uint64_t total() const {
        return quanta.empty() ? 0
                              : shift.back() + (uint64_t)nbra.back() *
                                                   (uint64_t)nket.back();
    }
    // binary search (SparseMatrixInfo::find_state, sparse_matrix.hpp:783-789)
    int find_state(uint64_t q) const;
};

// Connection list (SparseMatrixInfo::ConnectionInfo, sparse_matrix.hpp:72-480)
// in structure-of-arrays form.  Class k = entry of the sorted sub-label list
// (conj bits, combined delta quantum); its connections are
// [idx[k], idx[k+1]).  n[c] is the first class with conj == c.
struct Connections {
    int n[5] = {-1, -1, -1, -1, -1};
    std::vector<uint64_t> quanta;
    std::vector<uint32_t> idx;
    std::vector<uint64_t> stride; // output block (wfn) / 0 (diag)
    std::vector<double> factor;
    std::vector<uint32_t> ia, ib, ic;
    int nc() const { return (int)ia.size(); }
    // connection range of the class (conj, abdq); false if absent
    bool find_class(uint8_t conj, uint64_t abdq, int &begin, int &end) const;
};

using SubLabel = std::pair<uint8_t, uint64_t>;
using InfoRef = std::pair<uint64_t, const BlockInfo *>;

struct Operator {
    uint64_t dq;
    double factor;
    int info;          // index into the left/right info list
    uint64_t data_off; // element offset in the operator buffer
};

struct Term {
    int a, b;      // left / right operator index
    uint8_t conj;  // bit0: left transposed, bit1: right transposed
    double factor;
    // delayed three-tensor string (OpSumProd): sub-operators dl[da], dr[db]
    // of the delayed tensor with conj bits dconj
    bool three = false;
    int da = -1, db = -1;
    uint8_t dconj = 0;
};

struct Problem {
    Symm symm = Symm::SU2;
    uint64_t opdq = 0, hop_left_vacuum = 0;
    BlockInfo ket, bra;
    std::vector<BlockInfo> linfos, rinfos; // sorted by dq
    std::vector<Operator> lops, rops;
    std::vector<Term> terms;
    // delayed tensor (0 none, 1 left, 2 right): its sub-operator sets with
    // their infos (data in the delayed side's buffer) and the tensor-product
    // connection list of each info of the delayed side
    int delayed_side = 0;
    std::vector<BlockInfo> dlinfos, drinfos;
    std::vector<Operator> dlops, drops;
    std::vector<Connections> tp; // indexed like linfos / rinfos
};

// One recorded GEMM chain: V[v] += alpha * op(A) * C[c] * op(B)^T.  For a
// three-tensor chain C and V are windows with row strides c_ld / v_ld
// (0: dense) and alpha still lacks the value of the 1 x 1 sub-operator
// element at scal_off of the delayed side's buffer (-1: none).
struct PairRec {
    uint64_t a_off, b_off, c_off, v_off;
    uint32_t c_m, c_n, v_m, v_n;
    uint32_t conj;
    double alpha;
    uint32_t kind = 0;
    uint32_t c_ld = 0, v_ld = 0;
    int64_t scal_off = -1;
};

// D[d](r0 + i, c0 + j) += alpha * A(i,i) * B(j,j) (scal_off as above)
struct DiagRec {
    uint64_t a_off, b_off, d_off;
    uint32_t a_n, b_n;
    double alpha;
    uint32_t r0 = 0, c0 = 0;
    int64_t scal_off = -1;
};

// Unique (conj, delta-quantum) sub-labels of H_eff's strings
// (Partition::get_uniq_sub_labels, src/dmrg/partition.hpp:368-480).
std::vector<SubLabel> uniq_sub_labels(const Problem &p);

// ConnectionInfo::initialize_wfn (sparse_matrix.hpp:163-292)
void build_wfn_connections(Connections &ci, Symm s, const SU2Coupling &cg,
                           uint64_t cdq, uint64_t vdq, uint64_t opdq,
                           const std::vector<SubLabel> &subdq,
                           const std::vector<InfoRef> &ainfos,
                           const std::vector<InfoRef> &binfos,
                           const BlockInfo &cinfo, const BlockInfo &vinfo);

// ConnectionInfo::initialize_diag (sparse_matrix.hpp:81-161)
void build_diag_connections(Connections &ci, Symm s, const SU2Coupling &cg,
                            uint64_t cdq, uint64_t opdq,
                            const std::vector<SubLabel> &subdq,
                            const std::vector<InfoRef> &ainfos,
                            const std::vector<InfoRef> &binfos,
                            const BlockInfo &cinfo);

// Record the pair list of H*psi: for every string, every connection of its
// class (OperatorFunctions::tensor_product_multiply,
// src/core/operator_functions.hpp:474-542); for a delayed string, every
// matching tensor-product connection of the delayed operator's block as
// well (three_tensor_product_multiply, :543-671).
std::vector<PairRec> record_pairs(const Problem &p, const Connections &wfn,
                                  std::string &err);
// Record the diagonal contributions
// (OperatorFunctions::tensor_product_diagonal, operator_functions.hpp:212-247,
// three_tensor_product_diagonal :248-328).
std::vector<DiagRec> record_diag(const Problem &p, const Connections &dg,
                                 std::string &err);

// Reference-order multiply-add count of one H*psi
// (GMatrixFunctions::rotate return value, matrix_functions.hpp:1008).
uint64_t pair_macs(const PairRec &r);

} // namespace b2g
