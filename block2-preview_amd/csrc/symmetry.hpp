// Quantum-number labels and SU(2) coupling coefficients used by the plan
// builder.
//
// Labels are the packed 64-bit words block2 passes around, so a caller can hand
// its SparseMatrixInfo::quanta arrays over unchanged:
//   SZ  (src/core/symmetry.hpp:656-732, SZLongLong):
//       bits 48..63 N (int16), 16..31 2Sz (int16), 0..15 point group.
//   SU2 (src/core/symmetry.hpp:1184-1290, SU2LongLong):
//       bits 48..63 N (int16), 32..47 2S_low, 16..31 2S, 0..15 point group.
// An SU2 word with 2S_low != 2S is either a *range* of total spins produced by
// adding two spins (2S_low..2S in steps of 2) or, for an operator block, the
// composite (bra 2S in 2S_low, ket 2S in 2S).  The operations below implement
// the same algebra with plain field arithmetic.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace b2g {

enum class Symm : int { SZ = 0, SU2 = 1 };

constexpr uint64_t QINVALID = ~0ULL;

inline int q_n(uint64_t q) { return (int)(int16_t)(q >> 48); }
inline int q_pg(uint64_t q) { return (int)(q & 0xFFFFULL); }
// 2S (SU2) or 2Sz (SZ, signed)
inline int q_twos(Symm s, uint64_t q) {
    return s == Symm::SZ ? (int)(int16_t)((q >> 16) & 0xFFFFULL)
                         : (int)((q >> 16) & 0xFFFFULL);
}
inline int q_twos_low(uint64_t q) { return (int)((q >> 32) & 0xFFFFULL); }

inline uint64_t q_make(Symm s, int n, int twos_low, int twos, int pg) {
    uint64_t w = ((uint64_t)(uint16_t)(int16_t)n << 48) |
                 ((uint64_t)(uint16_t)twos << 16) | (uint64_t)(uint16_t)pg;
    if (s == Symm::SU2)
        w |= (uint64_t)(uint16_t)twos_low << 32;
    return w;
}

inline int q_multiplicity(Symm s, uint64_t q) {
    return s == Symm::SZ ? 1 : q_twos(s, q) + 1;
}
inline bool q_is_fermion(uint64_t q) { return (q >> 16) & 1ULL; }

inline uint64_t q_neg(Symm s, uint64_t q) {
    if (s == Symm::SZ)
        return q_make(s, -q_n(q), 0, -q_twos(s, q), q_pg(q));
    return q_make(s, -q_n(q), q_twos_low(q), q_twos(s, q), q_pg(q));
}

// a + b.  For SU2 the result is the range |Sa - Sb| .. Sa + Sb.
inline uint64_t q_add(Symm s, uint64_t a, uint64_t b) {
    int n = q_n(a) + q_n(b), pg = q_pg(a) ^ q_pg(b);
    if (s == Symm::SZ)
        return q_make(s, n, 0, q_twos(s, a) + q_twos(s, b), pg);
    // lowest spin of the coupled range: the smaller non-negative one of
    // (2Sa - 2Sb_low, 2Sb - 2Sa_low), evaluated in the 16-bit field with
    // wrap-around so that a negative candidate never wins
    uint64_t ta = (uint64_t)q_twos(s, a), tb = (uint64_t)q_twos(s, b);
    uint64_t la = (uint64_t)q_twos_low(a), lb = (uint64_t)q_twos_low(b);
    uint64_t lr = (ta << 32) - (lb << 32), rl = (tb << 32) - (la << 32);
    uint64_t lo = lr < rl ? lr : rl;
    uint64_t w = ((uint64_t)(uint16_t)(int16_t)n << 48) |
                 ((uint64_t)(uint16_t)(ta + tb) << 16) | (uint64_t)pg;
    return w | lo;
}
inline uint64_t q_sub(Symm s, uint64_t a, uint64_t b) {
    return q_add(s, a, q_neg(s, b));
}

// number of total spins in a range label
inline int q_count(Symm s, uint64_t q) {
    return s == Symm::SZ ? 1 : ((q_twos(s, q) - q_twos_low(q)) >> 1) + 1;
}
// i-th member of a range label (2S = 2S_low + 2i)
inline uint64_t q_at(Symm s, uint64_t q, int i) {
    if (s == Symm::SZ)
        return q;
    int t = q_twos_low(q) + 2 * i;
    return q_make(s, q_n(q), t, t, q_pg(q));
}
// ket part of a composite operator-block label
inline uint64_t q_get_ket(Symm s, uint64_t q) {
    if (s == Symm::SZ)
        return q;
    return q_make(s, q_n(q), q_twos(s, q), q_twos(s, q), q_pg(q));
}
// bra part of a composite label of an operator with delta quantum dq
inline uint64_t q_get_bra(Symm s, uint64_t q, uint64_t dq) {
    if (s == Symm::SZ)
        return q_add(s, q, dq);
    int t = q_twos_low(q);
    return q_make(s, q_n(q) + q_n(dq), t, t, q_pg(q) ^ q_pg(dq));
}
inline bool su2_triangle(int a, int b, int c) {
    return !((a + b + c) & 1) && c <= a + b && c >= std::abs(a - b);
}
// composite label of the block <bra| op(dq) |ket>, or QINVALID if forbidden
inline uint64_t q_combine(Symm s, uint64_t dq, uint64_t bra, uint64_t ket) {
    if (s == Symm::SZ)
        return q_add(s, ket, dq) == bra ? ket : QINVALID;
    uint64_t k = q_make(s, q_n(ket), q_twos(s, bra), q_twos(s, ket),
                        q_pg(ket));
    if (q_get_bra(s, k, dq) != bra ||
        !su2_triangle(q_twos(s, k), q_twos(s, dq), q_twos(s, bra)))
        return QINVALID;
    return k;
}

// SU(2) recoupling coefficients from the textbook formulas (Racah's closed
// form for the 6j symbol; the 9j symbol as the standard single sum over 6j
// products; A. R. Edmonds, Angular Momentum in Quantum Mechanics, eqs.
// 6.3.7 and 6.4.3).  All arguments are doubled spins.  Evaluated in long
// double with a factorial table.
class SU2Coupling {
  public:
    explicit SU2Coupling(int nfact = 256) : fact_(nfact) {
        fact_[0] = 1.0L;
        for (int i = 1; i < nfact; i++)
            fact_[i] = fact_[i - 1] * (long double)i;
    }
    long double delta(int a, int b, int c) const {
        return sqrtl(fact_[(a + b - c) / 2] * fact_[(a - b + c) / 2] *
                     fact_[(-a + b + c) / 2] / fact_[(a + b + c) / 2 + 1]);
    }
    long double sixj(int a, int b, int c, int d, int e, int f) const {
        if (!su2_triangle(a, b, c) || !su2_triangle(a, e, f) ||
            !su2_triangle(d, b, f) || !su2_triangle(d, e, c))
            return 0.0L;
        int t1 = (a + b + c) / 2, t2 = (a + e + f) / 2, t3 = (d + b + f) / 2,
            t4 = (d + e + c) / 2;
        int u1 = (a + b + d + e) / 2, u2 = (b + c + e + f) / 2,
            u3 = (a + c + d + f) / 2;
        int tmin = std::max(std::max(t1, t2), std::max(t3, t4));
        int tmax = std::min(u1, std::min(u2, u3));
        long double sum = 0.0L;
        for (int t = tmin; t <= tmax; t++) {
            long double den = fact_[t - t1] * fact_[t - t2] * fact_[t - t3] *
                              fact_[t - t4] * fact_[u1 - t] * fact_[u2 - t] *
                              fact_[u3 - t];
            long double term = fact_[t + 1] / den;
            sum += (t & 1) ? -term : term;
        }
        return sum * delta(a, b, c) * delta(a, e, f) * delta(d, b, f) *
               delta(d, e, c);
    }
    // { a b c }
    // { d e f }
    // { g h i }
    long double ninej(int a, int b, int c, int d, int e, int f, int g, int h,
                      int i) const {
        if (!su2_triangle(a, b, c) || !su2_triangle(d, e, f) ||
            !su2_triangle(g, h, i) || !su2_triangle(a, d, g) ||
            !su2_triangle(b, e, h) || !su2_triangle(c, f, i))
            return 0.0L;
        int xmin = std::max(std::abs(a - i),
                            std::max(std::abs(d - h), std::abs(b - f)));
        int xmax = std::min(a + i, std::min(d + h, b + f));
        long double sum = 0.0L;
        for (int x = xmin; x <= xmax; x += 2)
            sum += (long double)(x + 1) * sixj(a, b, c, f, i, x) *
                   sixj(d, e, f, b, x, h) * sixj(g, h, i, x, a, d);
        // (-1)^(2x) is constant over the sum since x steps by 2
        return (xmin & 1) ? -sum : sum;
    }
    // factor relating the reduced matrix of an operator (2S = td) between
    // row spin tl and column spin tr to that of its transpose
    long double transpose(int td, int tl, int tr) const {
        long double r = sqrtl((long double)(tr + 1) / (long double)(tl + 1));
        return ((td + tl - tr) & 2) ? -r : r;
    }

  private:
    std::vector<long double> fact_;
};

} // namespace b2g
