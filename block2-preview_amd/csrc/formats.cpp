// block2's binary files of block-sparse matrices, read and written on the
// host, so that wavefunctions and operators block2 saved can be handed to
// the device path (and back) without re-deriving them.
//
//   SparseMatrixInfo::save_data / load_data  src/core/sparse_matrix.hpp:504-575
//   SparseMatrix::save_data / load_data      src/core/sparse_matrix.hpp:908-989
//
// Layout (S = SZLongLong / SU2LongLong, 8 bytes; ubond_t of ubond_size
// bytes, 4 with block2's large-bond build):
//   info:   S delta_quantum | int32 n | quanta[n] (8 n bytes) |
//           n_states_bra[n], n_states_ket[n] (ubond_t each, the pair padded
//           to whole 32-bit words) | uint32 block_shifts[n] |
//           uint8 is_fermion | uint8 flags (block_shifts_fl_size << 2 |
//           is_wavefunction; block2 writes 8 << 2 | 1 for an FP64
//           wavefunction)
//   matrix: double factor | uint64 total_memory | double data[total_memory]
// A total_memory of 2^64 - 1 marks block2's compressed storage (an extra
// length and an FPCodec stream follow), which is not read here.
#include "common.hpp"

#include <cstring>
#include <fstream>
#include <memory>
#include <string>

struct b2g_smat_s {
    uint64_t dq = 0;
    int fermion = 0, flags = 0, has_info = 0;
    std::vector<uint64_t> quanta, shift;
    std::vector<uint32_t> nb, nk;
    double factor = 1.0;
    std::vector<double> data;
};

namespace {

// 32-bit words holding the two dimension arrays (_DBL_MEM_SIZE)
size_t dims_words(int n, int ubond_size) {
    return ubond_size == 4 ? (size_t)n << 1
                           : ubond_size == 2 ? (size_t)n
                                             : (size_t)(n - (n >> 1));
}

uint32_t get_ubond(const char *p, int ubond_size, size_t i) {
    if (ubond_size == 4) {
        uint32_t v;
        memcpy(&v, p + 4 * i, 4);
        return v;
    }
    if (ubond_size == 2) {
        uint16_t v;
        memcpy(&v, p + 2 * i, 2);
        return v;
    }
    return (uint8_t)p[i];
}

void put_ubond(char *p, int ubond_size, size_t i, uint32_t v) {
    if (ubond_size == 4)
        memcpy(p + 4 * i, &v, 4);
    else if (ubond_size == 2) {
        const uint16_t w = (uint16_t)v;
        memcpy(p + 2 * i, &w, 2);
    } else
        p[i] = (char)(uint8_t)v;
}

} // namespace

extern "C" int b2g_smat_load(const char *path, int with_info, int ubond_size,
                             b2g_smat *out) {
    B2G_REQUIRE(path && out, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(ubond_size == 1 || ubond_size == 2 || ubond_size == 4,
                B2G_ERR_ARG, "ubond_size must be 1, 2 or 4");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    B2G_REQUIRE(f.good(), B2G_ERR_ARG,
                std::string("cannot open '") + path + "'");
    std::unique_ptr<b2g_smat_s> m(new b2g_smat_s());
    auto rd = [&f](void *p, size_t n) {
        f.read((char *)p, (std::streamsize)n);
        return (size_t)f.gcount() == n;
    };
    const char *trunc = "file ends inside a sparse matrix";
    if (with_info) {
        int32_t n = 0;
        B2G_REQUIRE(rd(&m->dq, 8) && rd(&n, 4), B2G_ERR_ARG, trunc);
        B2G_REQUIRE(n >= 0 && n < (1 << 28), B2G_ERR_ARG,
                    "bad block count in sparse-matrix info");
        const size_t dw = dims_words(n, ubond_size);
        std::vector<char> buf(((size_t)2 * n + dw + n) * 4);
        B2G_REQUIRE(rd(buf.data(), buf.size()), B2G_ERR_ARG, trunc);
        uint8_t fer = 0, iwbs = 0;
        B2G_REQUIRE(rd(&fer, 1) && rd(&iwbs, 1), B2G_ERR_ARG, trunc);
        m->fermion = fer != 0;
        m->flags = iwbs;
        m->quanta.resize(n);
        memcpy(m->quanta.data(), buf.data(), (size_t)8 * n);
        const char *dims = buf.data() + (size_t)8 * n;
        m->nb.resize(n);
        m->nk.resize(n);
        for (int i = 0; i < n; i++) {
            m->nb[i] = get_ubond(dims, ubond_size, (size_t)i);
            m->nk[i] = get_ubond(dims, ubond_size, (size_t)n + i);
        }
        const char *sh = dims + dw * 4;
        m->shift.resize(n);
        for (int i = 0; i < n; i++) {
            uint32_t v;
            memcpy(&v, sh + (size_t)4 * i, 4);
            m->shift[i] = v;
        }
        m->has_info = 1;
    }
    uint64_t total = 0;
    B2G_REQUIRE(rd(&m->factor, 8) && rd(&total, 8), B2G_ERR_ARG, trunc);
    B2G_REQUIRE(total != ~(uint64_t)0, B2G_ERR_ARG,
                "compressed sparse-matrix storage is not supported");
    B2G_REQUIRE(total < ((uint64_t)1 << 40), B2G_ERR_ARG,
                "bad element count in sparse matrix");
    m->data.resize(total);
    B2G_REQUIRE(rd(m->data.data(), total * 8), B2G_ERR_ARG, trunc);
    if (m->has_info) {
        for (size_t i = 0; i < m->quanta.size(); i++)
            B2G_REQUIRE(m->shift[i] + (uint64_t)m->nb[i] * m->nk[i] <= total,
                        B2G_ERR_ARG, "block outside the matrix data");
    }
    *out = m.release();
    return B2G_OK;
}

extern "C" int b2g_smat_get(b2g_smat m, b2g_sparse_info *info,
                            int *flags, double *factor,
                            const double **data, size_t *len) {
    B2G_REQUIRE(m, B2G_ERR_ARG, "NULL arg");
    if (info) {
        B2G_REQUIRE(m->has_info, B2G_ERR_STATE,
                    "sparse matrix was loaded without its info");
        info->delta_quantum = m->dq;
        info->is_fermion = m->fermion;
        info->n = (int32_t)m->quanta.size();
        info->quanta = m->quanta.data();
        info->n_states_bra = m->nb.data();
        info->n_states_ket = m->nk.data();
        info->block_shifts = m->shift.data();
    }
    if (flags)
        *flags = m->flags;
    if (factor)
        *factor = m->factor;
    if (data)
        *data = m->data.data();
    if (len)
        *len = m->data.size();
    return B2G_OK;
}

extern "C" int b2g_smat_save(const char *path, const b2g_sparse_info *info,
                             int flags, int ubond_size,
                             double factor, const double *data, size_t len) {
    B2G_REQUIRE(path && (data || len == 0), B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(ubond_size == 1 || ubond_size == 2 || ubond_size == 4,
                B2G_ERR_ARG, "ubond_size must be 1, 2 or 4");
    std::string bytes;
    auto put = [&bytes](const void *p, size_t n) {
        bytes.append((const char *)p, n);
    };
    if (info != nullptr) {
        const int32_t n = info->n;
        B2G_REQUIRE(n >= 0 && (n == 0 || (info->quanta && info->n_states_bra &&
                                          info->n_states_ket &&
                                          info->block_shifts)),
                    B2G_ERR_ARG, "bad info");
        const uint32_t lim = ubond_size == 4 ? 0xFFFFFFFFu
                                             : ubond_size == 2 ? 0xFFFFu : 0xFFu;
        for (int i = 0; i < n; i++) {
            B2G_REQUIRE(info->n_states_bra[i] <= lim &&
                            info->n_states_ket[i] <= lim,
                        B2G_ERR_ARG, "block dimension above the ubond range");
            B2G_REQUIRE(info->block_shifts[i] <= 0xFFFFFFFFull, B2G_ERR_ARG,
                        "block shift above 2^32 (block2 stores 32 bits)");
            B2G_REQUIRE(info->block_shifts[i] +
                                (uint64_t)info->n_states_bra[i] *
                                    info->n_states_ket[i] <=
                            len,
                        B2G_ERR_ARG, "block outside the matrix data");
        }
        put(&info->delta_quantum, 8);
        put(&n, 4);
        put(info->quanta, (size_t)8 * n);
        std::vector<char> dims(dims_words(n, ubond_size) * 4, 0);
        for (int i = 0; i < n; i++) {
            put_ubond(dims.data(), ubond_size, (size_t)i, info->n_states_bra[i]);
            put_ubond(dims.data(), ubond_size, (size_t)n + i,
                      info->n_states_ket[i]);
        }
        put(dims.data(), dims.size());
        for (int i = 0; i < n; i++) {
            const uint32_t v = (uint32_t)info->block_shifts[i];
            put(&v, 4);
        }
        B2G_REQUIRE(flags >= 0 && flags < 256, B2G_ERR_ARG, "bad info flags");
        const uint8_t fer = info->is_fermion ? 1 : 0;
        const uint8_t iwbs = (uint8_t)flags;
        put(&fer, 1);
        put(&iwbs, 1);
    }
    const uint64_t total = len;
    put(&factor, 8);
    put(&total, 8);
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    B2G_REQUIRE(f.good(), B2G_ERR_ARG,
                std::string("cannot create '") + path + "'");
    f.write(bytes.data(), (std::streamsize)bytes.size());
    f.write((const char *)data, (std::streamsize)(len * 8));
    B2G_REQUIRE(f.good(), B2G_ERR_ARG,
                std::string("write failed on '") + path + "'");
    return B2G_OK;
}

extern "C" int b2g_smat_destroy(b2g_smat m) {
    delete m;
    return B2G_OK;
}
