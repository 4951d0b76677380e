/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.  Plain-C CPU restatement of block2's
 * effective-Hamiltonian hot path; see b2g_oracle.h for the reference
 * file:line each part follows.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker.
 */
#include "b2g_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define TINY 1E-20
#define QINV 0xFFFFFFFFFFFFFFFFULL

/* ------------------------------------------------------------- labels */
/* symmetry.hpp:656-732 (SZ), 1184-1290 (SU2) */
static int g_su2; /* set per call site from the problem */

static int qn(uint64_t q) { return (int)(int16_t)(q >> 48); }
static int qpg(uint64_t q) { return (int)(q & 0xFFFF); }
static int qtw(int su2, uint64_t q) {
    return su2 ? (int)((q >> 16) & 0xFFFF) : (int)(int16_t)((q >> 16) & 0xFFFF);
}
static int qtl(uint64_t q) { return (int)((q >> 32) & 0xFFFF); }
static uint64_t qmk(int su2, int n, int tl, int tw, int pg) {
    uint64_t w = ((uint64_t)(uint16_t)(int16_t)n << 48) |
                 ((uint64_t)(uint16_t)tw << 16) | (uint64_t)(uint16_t)pg;
    if (su2)
        w |= (uint64_t)(uint16_t)tl << 32;
    return w;
}
static uint64_t qneg(int su2, uint64_t q) {
    return su2 ? qmk(1, -qn(q), qtl(q), qtw(1, q), qpg(q))
               : qmk(0, -qn(q), 0, -qtw(0, q), qpg(q));
}
static uint64_t qadd(int su2, uint64_t a, uint64_t b) {
    int n = qn(a) + qn(b), pg = qpg(a) ^ qpg(b);
    if (!su2)
        return qmk(0, n, 0, qtw(0, a) + qtw(0, b), pg);
    uint64_t ta = (uint64_t)qtw(1, a), tb = (uint64_t)qtw(1, b);
    uint64_t la = (uint64_t)qtl(a), lb = (uint64_t)qtl(b);
    uint64_t lr = (ta << 32) - (lb << 32), rl = (tb << 32) - (la << 32);
    uint64_t lo = lr < rl ? lr : rl;
    return (((uint64_t)(uint16_t)(int16_t)n << 48) |
            ((uint64_t)(uint16_t)(ta + tb) << 16) | (uint64_t)pg) |
           lo;
}
static uint64_t qsub(int su2, uint64_t a, uint64_t b) {
    return qadd(su2, a, qneg(su2, b));
}
static int qcount(int su2, uint64_t q) {
    return su2 ? ((qtw(1, q) - qtl(q)) >> 1) + 1 : 1;
}
static uint64_t qat(int su2, uint64_t q, int i) {
    if (!su2)
        return q;
    int t = qtl(q) + 2 * i;
    return qmk(1, qn(q), t, t, qpg(q));
}
static uint64_t qket(int su2, uint64_t q) {
    return su2 ? qmk(1, qn(q), qtw(1, q), qtw(1, q), qpg(q)) : q;
}
static uint64_t qbra(int su2, uint64_t q, uint64_t dq) {
    if (!su2)
        return qadd(0, q, dq);
    return qmk(1, qn(q) + qn(dq), qtl(q), qtl(q), qpg(q) ^ qpg(dq));
}
static int tri(int a, int b, int c) {
    return !((a + b + c) & 1) && c <= a + b && c >= abs(a - b);
}
static uint64_t qcombine(int su2, uint64_t dq, uint64_t bra, uint64_t ket) {
    if (!su2)
        return qadd(0, ket, dq) == bra ? ket : QINV;
    uint64_t k = qmk(1, qn(ket), qtw(1, bra), qtw(1, ket), qpg(ket));
    if (qbra(1, k, dq) != bra || !tri(qtw(1, k), qtw(1, dq), qtw(1, bra)))
        return QINV;
    return k;
}
static int qmult(int su2, uint64_t q) { return su2 ? qtw(1, q) + 1 : 1; }
static int qferm(uint64_t q) { return (int)((q >> 16) & 1); }

/* ----------------------------------------------------------- couplings */
/* Racah closed form of the 6j symbol and the 9j symbol as the standard sum
 * over three 6j symbols (clebsch_gordan.hpp:117-176 computes the same
 * quantities). */
static long double g_fact[200];
static int g_fact_ok;
static void init_fact(void) {
    if (g_fact_ok)
        return;
    g_fact[0] = 1.0L;
    for (int i = 1; i < 200; i++)
        g_fact[i] = g_fact[i - 1] * i;
    g_fact_ok = 1;
}
static long double tdelta(int a, int b, int c) {
    return sqrtl(g_fact[(a + b - c) / 2] * g_fact[(a - b + c) / 2] *
                 g_fact[(-a + b + c) / 2] / g_fact[(a + b + c) / 2 + 1]);
}
static long double sixj_l(int a, int b, int c, int d, int e, int f) {
    init_fact();
    if (!tri(a, b, c) || !tri(a, e, f) || !tri(d, b, f) || !tri(d, e, c))
        return 0;
    int t1 = (a + b + c) / 2, t2 = (a + e + f) / 2, t3 = (d + b + f) / 2,
        t4 = (d + e + c) / 2;
    int u1 = (a + b + d + e) / 2, u2 = (b + c + e + f) / 2,
        u3 = (a + c + d + f) / 2;
    int tmin = t1;
    if (t2 > tmin) tmin = t2;
    if (t3 > tmin) tmin = t3;
    if (t4 > tmin) tmin = t4;
    int tmax = u1;
    if (u2 < tmax) tmax = u2;
    if (u3 < tmax) tmax = u3;
    long double s = 0;
    for (int t = tmin; t <= tmax; t++) {
        long double x = g_fact[t + 1] /
                        (g_fact[t - t1] * g_fact[t - t2] * g_fact[t - t3] *
                         g_fact[t - t4] * g_fact[u1 - t] * g_fact[u2 - t] *
                         g_fact[u3 - t]);
        s += (t & 1) ? -x : x;
    }
    return s * tdelta(a, b, c) * tdelta(a, e, f) * tdelta(d, b, f) *
           tdelta(d, e, c);
}
static long double ninej_l(int a, int b, int c, int d, int e, int f, int g,
                           int h, int i) {
    if (!tri(a, b, c) || !tri(d, e, f) || !tri(g, h, i) || !tri(a, d, g) ||
        !tri(b, e, h) || !tri(c, f, i))
        return 0;
    int x0 = abs(a - i), x1 = abs(d - h), x2 = abs(b - f);
    int xmin = x0 > x1 ? x0 : x1;
    if (x2 > xmin) xmin = x2;
    int y0 = a + i, y1 = d + h, y2 = b + f;
    int xmax = y0 < y1 ? y0 : y1;
    if (y2 < xmax) xmax = y2;
    long double s = 0;
    for (int x = xmin; x <= xmax; x += 2)
        s += (x + 1) * sixj_l(a, b, c, f, i, x) * sixj_l(d, e, f, b, x, h) *
             sixj_l(g, h, i, x, a, d);
    return (xmin & 1) ? -s : s;
}
static long double transpose_l(int td, int tl, int tr) {
    long double r = sqrtl((long double)(tr + 1) / (long double)(tl + 1));
    return ((td + tl - tr) & 2) ? -r : r;
}
double orc_sixj(int a, int b, int c, int d, int e, int f) {
    return (double)sixj_l(a, b, c, d, e, f);
}
double orc_ninej(int a, int b, int c, int d, int e, int f, int g, int h,
                 int i) {
    return (double)ninej_l(a, b, c, d, e, f, g, h, i);
}

/* ------------------------------------------------------------ problem */
typedef struct {
    uint64_t dq;
    int ferm, n;
    const uint64_t *q;
    const uint32_t *nb, *nk;
    const uint64_t *sh;
} info_t;

typedef struct {
    int n5[5];
    int64_t nq;
    uint64_t *quanta;
    uint32_t *idx;
    int64_t nc, cap;
    uint32_t *ia, *ib, *ic;
    uint64_t *st;
    double *f;
} conn_t;

typedef struct {
    uint64_t a_off, b_off, c_off, v_off;
    uint32_t cm, cn, vm, vn, conj;
    double alpha;
} pair_t;

typedef struct {
    uint64_t a_off, b_off, d_off;
    uint32_t an, bn;
    double alpha;
} dterm_t;

struct orc_problem {
    int su2;
    uint64_t opdq, hlv;
    info_t ket, bra;
    int nli, nri;
    info_t *li, *ri;
    int nlo, nro;
    const double *lo_f, *ro_f;
    const int32_t *lo_info, *ro_info;
    const uint64_t *lo_off, *ro_off;
    const double *ldata, *rdata;
    int nt;
    const uint8_t *t_conj;
    const int32_t *t_a, *t_b;
    const double *t_f;
    conn_t wfn, dg;
    pair_t *pairs;
    int64_t np;
    dterm_t *dterms;
    int64_t nd;
    char err[256];
};

static int find_state(const info_t *in, uint64_t q) {
    int lo = 0, hi = in->n;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (in->q[mid] < q)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < in->n && in->q[lo] == q) ? lo : -1;
}

static const info_t *find_info(const info_t *infos, int n, uint64_t dq) {
    for (int i = 0; i < n; i++)
        if (infos[i].dq == dq)
            return &infos[i];
    return NULL;
}

static void conn_push(conn_t *c, uint32_t ia, uint32_t ib, uint32_t ic,
                      uint64_t st, double f) {
    if (c->nc == c->cap) {
        c->cap = c->cap ? c->cap * 2 : 256;
        c->ia = realloc(c->ia, c->cap * 4);
        c->ib = realloc(c->ib, c->cap * 4);
        c->ic = realloc(c->ic, c->cap * 4);
        c->st = realloc(c->st, c->cap * 8);
        c->f = realloc(c->f, c->cap * 8);
    }
    c->ia[c->nc] = ia, c->ib[c->nc] = ib, c->ic[c->nc] = ic;
    c->st[c->nc] = st, c->f[c->nc] = f;
    c->nc++;
}

typedef struct {
    uint8_t conj;
    uint64_t q;
} sub_t;
static int sub_cmp(const void *x, const void *y) {
    const sub_t *a = x, *b = y;
    if (a->conj != b->conj)
        return a->conj < b->conj ? -1 : 1;
    return a->q < b->q ? -1 : a->q > b->q ? 1 : 0;
}

/* partition.hpp:368-480 */
static sub_t *sub_labels(orc_problem *p, int *ns) {
    int su2 = p->su2;
    sub_t *r = malloc(sizeof(sub_t) * (p->nt + 1));
    int n = 0;
    for (int t = 0; t < p->nt; t++) {
        uint64_t adq = p->li[p->lo_info[p->t_a[t]]].dq;
        uint64_t bdq = p->ri[p->ro_info[p->t_b[t]]].dq;
        uint8_t cj = p->t_conj[t];
        uint64_t q = qcombine(su2, p->opdq, (cj & 1) ? qneg(su2, adq) : adq,
                              (cj & 2) ? bdq : qneg(su2, bdq));
        if (q != QINV)
            r[n].conj = cj, r[n].q = q, n++;
    }
    uint64_t p0 = qcombine(su2, p->opdq, p->hlv,
                           qat(su2, qsub(su2, p->opdq, p->opdq), 0));
    if (p0 != QINV)
        r[n].conj = 0, r[n].q = p0, n++;
    qsort(r, n, sizeof(sub_t), sub_cmp);
    int m = 0;
    for (int i = 0; i < n; i++)
        if (m == 0 || sub_cmp(&r[m - 1], &r[i]) != 0)
            r[m++] = r[i];
    *ns = m;
    return r;
}

static void conn_classes(conn_t *c, const sub_t *s, int ns) {
    c->nq = ns;
    c->quanta = malloc(8 * (ns + 1));
    for (int i = 0; i < ns; i++)
        c->quanta[i] = s[i].q;
    c->n5[4] = ns;
    for (int i = 3; i >= 0; i--)
        if (c->n5[i] == -1)
            c->n5[i] = c->n5[i + 1];
}

typedef struct {
    double f;
    uint64_t iv;
    uint32_t ia, ib, ic;
} ent_t;

/* sparse_matrix.hpp:163-292 */
static void build_wfn(orc_problem *p, const sub_t *s, int ns) {
    int su2 = p->su2;
    conn_t *c = &p->wfn;
    for (int i = 0; i < 5; i++)
        c->n5[i] = -1;
    c->idx = malloc(4 * (ns + 1));
    uint64_t cdq = p->ket.dq, vdq = p->bra.dq, opdq = p->opdq;
    for (int k = 0; k < ns; k++) {
        uint8_t conj = s[k].conj;
        if (c->n5[conj] == -1)
            c->n5[conj] = k;
        int cja = conj & 1, cjb = (conj & 2) >> 1;
        c->idx[k] = (uint32_t)c->nc;
        uint64_t a0 = qbra(su2, s[k].q, opdq), b0 = qket(su2, s[k].q);
        uint64_t adq = cja ? qneg(su2, a0) : a0;
        uint64_t bdq = cjb ? b0 : qneg(su2, b0);
        const info_t *ai = find_info(p->li, p->nli, adq);
        const info_t *bi = find_info(p->ri, p->nri, bdq);
        if (!ai || !bi)
            continue;
        /* per output block lists, flattened rank-major */
        int nv = p->bra.n;
        ent_t **lists = calloc(nv, sizeof(ent_t *));
        int *cnt = calloc(nv, sizeof(int)), *cap = calloc(nv, sizeof(int));
        int maxr = 0;
        for (int iv = 0; iv < nv; iv++) {
            uint64_t lq = qbra(su2, p->bra.q[iv], vdq);
            uint64_t rq = qneg(su2, qket(su2, p->bra.q[iv]));
            uint64_t rqps = cjb ? qadd(su2, rq, bdq) : qsub(su2, rq, bdq);
            for (int r = 0; r < qcount(su2, rqps); r++) {
                uint64_t rqp = qat(su2, rqps, r);
                int ib = find_state(bi, cjb ? qcombine(su2, bdq, rqp, rq)
                                            : qcombine(su2, bdq, rq, rqp));
                if (ib == -1)
                    continue;
                uint64_t lqps = qsub(su2, cdq, rqp);
                for (int l = 0; l < qcount(su2, lqps); l++) {
                    uint64_t lqp = qat(su2, lqps, l);
                    int ia = find_state(ai, cja ? qcombine(su2, adq, lqp, lq)
                                                : qcombine(su2, adq, lq, lqp));
                    int ic = find_state(&p->ket,
                                        qcombine(su2, cdq, lqp, qneg(su2, rqp)));
                    if (ia == -1 || ic == -1)
                        continue;
                    long double f = 1;
                    if (su2) {
                        f = sqrtl((long double)(qmult(1, cdq) * qmult(1, opdq) *
                                                qmult(1, lq) * qmult(1, rq))) *
                            ninej_l(qtw(1, lqp), qtw(1, rqp), qtw(1, cdq),
                                    qtw(1, adq), qtw(1, bdq), qtw(1, opdq),
                                    qtw(1, lq), qtw(1, rq), qtw(1, vdq));
                    }
                    if (bi->ferm && qferm(lqp))
                        f = -f;
                    if (su2 && cja)
                        f *= transpose_l(qtw(1, adq), qtw(1, lq), qtw(1, lqp));
                    if (su2 && cjb)
                        f *= transpose_l(qtw(1, bdq), qtw(1, rq), qtw(1, rqp));
                    double fd = (double)f;
                    if (fabs(fd) < TINY)
                        continue;
                    if (cnt[iv] == cap[iv]) {
                        cap[iv] = cap[iv] ? 2 * cap[iv] : 8;
                        lists[iv] = realloc(lists[iv], cap[iv] * sizeof(ent_t));
                    }
                    ent_t e = {fd, (uint64_t)iv, (uint32_t)ia, (uint32_t)ib,
                               (uint32_t)ic};
                    lists[iv][cnt[iv]++] = e;
                    if (cnt[iv] > maxr)
                        maxr = cnt[iv];
                }
            }
        }
        for (int r = 0; r < maxr; r++)
            for (int iv = 0; iv < nv; iv++)
                if (r < cnt[iv]) {
                    ent_t *e = &lists[iv][r];
                    conn_push(c, e->ia, e->ib, e->ic, e->iv, e->f);
                }
        for (int iv = 0; iv < nv; iv++)
            free(lists[iv]);
        free(lists), free(cnt), free(cap);
    }
    conn_classes(c, s, ns);
}

/* sparse_matrix.hpp:81-161 */
static void build_diag(orc_problem *p, const sub_t *s, int ns) {
    int su2 = p->su2;
    conn_t *c = &p->dg;
    for (int i = 0; i < 5; i++)
        c->n5[i] = -1;
    c->idx = malloc(4 * (ns + 1));
    uint64_t cdq = p->ket.dq, opdq = p->opdq;
    for (int k = 0; k < ns; k++) {
        uint8_t conj = s[k].conj;
        if (c->n5[conj] == -1)
            c->n5[conj] = k;
        int cja = conj & 1, cjb = (conj & 2) >> 1;
        c->idx[k] = (uint32_t)c->nc;
        uint64_t a0 = qbra(su2, s[k].q, opdq), b0 = qket(su2, s[k].q);
        uint64_t adq = cja ? qneg(su2, a0) : a0;
        uint64_t bdq = cjb ? b0 : qneg(su2, b0);
        if (qat(su2, qadd(su2, adq, bdq), 0) != qat(su2, qsub(su2, adq, adq), 0))
            continue;
        const info_t *ai = find_info(p->li, p->nli, adq);
        const info_t *bi = find_info(p->ri, p->nri, bdq);
        if (!ai || !bi)
            continue;
        for (int ic = 0; ic < p->ket.n; ic++) {
            uint64_t aq = qbra(su2, p->ket.q[ic], cdq);
            uint64_t bq = qneg(su2, qket(su2, p->ket.q[ic]));
            int ia = find_state(ai, aq), ib = find_state(bi, bq);
            if (ia == -1 || ib == -1 || aq != qbra(su2, aq, adq) ||
                bq != qbra(su2, bq, bdq))
                continue;
            long double f = 1;
            if (su2) {
                f = sqrtl((long double)(qmult(1, cdq) * qmult(1, opdq) *
                                        qmult(1, aq) * qmult(1, bq))) *
                    ninej_l(qtw(1, aq), qtw(1, bq), qtw(1, cdq), qtw(1, adq),
                            qtw(1, bdq), qtw(1, opdq), qtw(1, aq), qtw(1, bq),
                            qtw(1, cdq));
                if (cja)
                    f *= transpose_l(qtw(1, adq), qtw(1, aq), qtw(1, aq));
                if (cjb)
                    f *= transpose_l(qtw(1, bdq), qtw(1, bq), qtw(1, bq));
            }
            if (bi->ferm && qferm(aq))
                f = -f;
            double fd = (double)f;
            if (fabs(fd) < TINY)
                continue;
            conn_push(c, (uint32_t)ia, (uint32_t)ib, (uint32_t)ic, 0, fd);
        }
    }
    conn_classes(c, s, ns);
}

static int term_class(const orc_problem *p, const conn_t *c, int t, int *b,
                      int *e) {
    int su2 = p->su2;
    uint64_t adq = p->li[p->lo_info[p->t_a[t]]].dq;
    uint64_t bdq = p->ri[p->ro_info[p->t_b[t]]].dq;
    uint8_t cj = p->t_conj[t];
    uint64_t abdq = qcombine(su2, p->opdq, (cj & 1) ? qneg(su2, adq) : adq,
                             (cj & 2) ? bdq : qneg(su2, bdq));
    for (int k = c->n5[cj]; k < c->n5[cj + 1]; k++)
        if (c->quanta[k] == abdq) {
            *b = (int)c->idx[k];
            *e = k == c->n5[4] - 1 ? (int)c->nc : (int)c->idx[k + 1];
            return 1;
        }
    return 0;
}

/* operator_functions.hpp:474-542 (records each rotate) */
static int record(orc_problem *p) {
    int64_t cap = 1024;
    p->pairs = malloc(cap * sizeof(pair_t));
    p->np = 0;
    for (int t = 0; t < p->nt; t++) {
        int a = p->t_a[t], b = p->t_b[t];
        double scale = p->t_f[t] * p->lo_f[a] * p->ro_f[b];
        if (fabs(scale) < TINY)
            continue;
        int lb, le;
        if (!term_class(p, &p->wfn, t, &lb, &le)) {
            snprintf(p->err, sizeof(p->err), "term %d has no class", t);
            return 1;
        }
        const info_t *ai = &p->li[p->lo_info[a]], *bi = &p->ri[p->ro_info[b]];
        for (int il = lb; il < le; il++) {
            uint32_t ia = p->wfn.ia[il], ib = p->wfn.ib[il], ic = p->wfn.ic[il];
            uint64_t iv = p->wfn.st[il];
            pair_t r;
            r.a_off = p->lo_off[a] + ai->sh[ia];
            r.b_off = p->ro_off[b] + bi->sh[ib];
            r.c_off = p->ket.sh[ic];
            r.v_off = p->bra.sh[iv];
            r.cm = p->ket.nb[ic], r.cn = p->ket.nk[ic];
            r.vm = p->bra.nb[iv], r.vn = p->bra.nk[iv];
            r.conj = p->t_conj[t];
            r.alpha = scale * p->wfn.f[il];
            if (p->np == cap) {
                cap *= 2;
                p->pairs = realloc(p->pairs, cap * sizeof(pair_t));
            }
            p->pairs[p->np++] = r;
        }
    }
    cap = 1024;
    p->dterms = malloc(cap * sizeof(dterm_t));
    p->nd = 0;
    for (int t = 0; t < p->nt; t++) {
        int a = p->t_a[t], b = p->t_b[t];
        double scale = p->t_f[t] * p->lo_f[a] * p->ro_f[b];
        if (fabs(scale) < TINY)
            continue;
        int lb, le;
        if (!term_class(p, &p->dg, t, &lb, &le))
            continue;
        const info_t *ai = &p->li[p->lo_info[a]], *bi = &p->ri[p->ro_info[b]];
        for (int il = lb; il < le; il++) {
            uint32_t ia = p->dg.ia[il], ib = p->dg.ib[il], ic = p->dg.ic[il];
            dterm_t d;
            d.a_off = p->lo_off[a] + ai->sh[ia];
            d.b_off = p->ro_off[b] + bi->sh[ib];
            d.d_off = p->ket.sh[ic];
            d.an = ai->nb[ia], d.bn = bi->nb[ib];
            d.alpha = scale * p->dg.f[il];
            if (p->nd == cap) {
                cap *= 2;
                p->dterms = realloc(p->dterms, cap * sizeof(dterm_t));
            }
            p->dterms[p->nd++] = d;
        }
    }
    return 0;
}

orc_problem *orc_create(
    int sym, uint64_t opdq, uint64_t hlv, uint64_t ket_dq, int ket_n,
    const uint64_t *ket_q, const uint32_t *ket_nb, const uint32_t *ket_nk,
    const uint64_t *ket_sh, uint64_t bra_dq, int bra_n, const uint64_t *bra_q,
    const uint32_t *bra_nb, const uint32_t *bra_nk, const uint64_t *bra_sh,
    int nli, const uint64_t *li_dq, const uint32_t *li_ferm,
    const uint32_t *li_n, const uint32_t *li_off, const uint64_t *li_q,
    const uint32_t *li_nb, const uint32_t *li_nk, const uint64_t *li_sh,
    int nri, const uint64_t *ri_dq, const uint32_t *ri_ferm,
    const uint32_t *ri_n, const uint32_t *ri_off, const uint64_t *ri_q,
    const uint32_t *ri_nb, const uint32_t *ri_nk, const uint64_t *ri_sh,
    int nlo, const double *lo_f, const int32_t *lo_info, const uint64_t *lo_off,
    const double *ldata, int nro, const double *ro_f, const int32_t *ro_info,
    const uint64_t *ro_off, const double *rdata, int nt, const uint8_t *t_conj,
    const int32_t *t_a, const int32_t *t_b, const double *t_f) {
    orc_problem *p = calloc(1, sizeof(orc_problem));
    p->su2 = sym == 1;
    p->opdq = opdq, p->hlv = hlv;
    info_t k = {ket_dq, 0, ket_n, ket_q, ket_nb, ket_nk, ket_sh};
    info_t b = {bra_dq, 0, bra_n, bra_q, bra_nb, bra_nk, bra_sh};
    p->ket = k, p->bra = b;
    p->nli = nli, p->nri = nri;
    p->li = malloc(sizeof(info_t) * (nli + 1));
    p->ri = malloc(sizeof(info_t) * (nri + 1));
    for (int i = 0; i < nli; i++) {
        info_t x = {li_dq[i], (int)li_ferm[i], (int)li_n[i], li_q + li_off[i],
                    li_nb + li_off[i], li_nk + li_off[i], li_sh + li_off[i]};
        p->li[i] = x;
    }
    for (int i = 0; i < nri; i++) {
        info_t x = {ri_dq[i], (int)ri_ferm[i], (int)ri_n[i], ri_q + ri_off[i],
                    ri_nb + ri_off[i], ri_nk + ri_off[i], ri_sh + ri_off[i]};
        p->ri[i] = x;
    }
    p->nlo = nlo, p->lo_f = lo_f, p->lo_info = lo_info, p->lo_off = lo_off;
    p->ldata = ldata;
    p->nro = nro, p->ro_f = ro_f, p->ro_info = ro_info, p->ro_off = ro_off;
    p->rdata = rdata;
    p->nt = nt, p->t_conj = t_conj, p->t_a = t_a, p->t_b = t_b, p->t_f = t_f;
    int ns;
    sub_t *s = sub_labels(p, &ns);
    build_diag(p, s, ns);
    build_wfn(p, s, ns);
    free(s);
    record(p);
    return p;
}

static void conn_free(conn_t *c) {
    free(c->quanta), free(c->idx), free(c->ia), free(c->ib), free(c->ic);
    free(c->st), free(c->f);
}

void orc_destroy(orc_problem *p) {
    if (!p)
        return;
    conn_free(&p->wfn), conn_free(&p->dg);
    free(p->li), free(p->ri), free(p->pairs), free(p->dterms);
    free(p);
}

const char *orc_error(orc_problem *p) { return p->err; }

static uint64_t info_total(const info_t *in) {
    return in->n == 0 ? 0
                      : in->sh[in->n - 1] +
                            (uint64_t)in->nb[in->n - 1] * in->nk[in->n - 1];
}
uint64_t orc_ket_len(const orc_problem *p) { return info_total(&p->ket); }
uint64_t orc_bra_len(const orc_problem *p) { return info_total(&p->bra); }

uint64_t orc_macs(const orc_problem *p) {
    uint64_t s = 0;
    for (int64_t i = 0; i < p->np; i++) {
        const pair_t *r = &p->pairs[i];
        s += (uint64_t)r->cm * r->cn * r->vn + (uint64_t)r->vm * r->cm * r->vn;
    }
    return s;
}

int64_t orc_connections(const orc_problem *p, int which, uint32_t *ia,
                        uint32_t *ib, uint32_t *ic, uint64_t *stride,
                        double *factor, int32_t *n5, int64_t *nq,
                        uint64_t *quanta, uint32_t *idx) {
    const conn_t *c = which == 0 ? &p->wfn : &p->dg;
    if (ia) memcpy(ia, c->ia, c->nc * 4);
    if (ib) memcpy(ib, c->ib, c->nc * 4);
    if (ic) memcpy(ic, c->ic, c->nc * 4);
    if (stride) memcpy(stride, c->st, c->nc * 8);
    if (factor) memcpy(factor, c->f, c->nc * 8);
    if (n5) memcpy(n5, c->n5, 5 * 4);
    if (nq) *nq = c->nq;
    if (quanta) memcpy(quanta, c->quanta, c->nq * 8);
    if (idx) memcpy(idx, c->idx, c->nq * 4);
    return c->nc;
}

/* V += alpha op(A) (C op(B)^T), plain loops (matrix_functions.hpp:998-1009) */
static void rotate_pair(const orc_problem *p, const pair_t *r, const double *x,
                        double *y, double *work) {
    const double *A = p->ldata + r->a_off, *B = p->rdata + r->b_off;
    const double *C = x + r->c_off;
    double *V = y + r->v_off;
    const uint32_t cm = r->cm, cn = r->cn, vm = r->vm, vn = r->vn;
    /* work (cm x vn) = C (cm x cn) * op(B) with op(B)[k][j] =
     * (conj&2) ? B[k*vn + j] : B[j*cn + k] */
    for (uint32_t i = 0; i < cm; i++)
        for (uint32_t j = 0; j < vn; j++) {
            double s = 0;
            if (r->conj & 2)
                for (uint32_t k = 0; k < cn; k++)
                    s += C[i * cn + k] * B[(uint64_t)k * vn + j];
            else
                for (uint32_t k = 0; k < cn; k++)
                    s += C[i * cn + k] * B[(uint64_t)j * cn + k];
            work[(uint64_t)i * vn + j] = s;
        }
    /* V (vm x vn) += alpha op(A) (vm x cm) * work */
    for (uint32_t i = 0; i < vm; i++)
        for (uint32_t j = 0; j < vn; j++) {
            double s = 0;
            if (r->conj & 1)
                for (uint32_t k = 0; k < cm; k++)
                    s += A[(uint64_t)k * vm + i] * work[(uint64_t)k * vn + j];
            else
                for (uint32_t k = 0; k < cm; k++)
                    s += A[(uint64_t)i * cm + k] * work[(uint64_t)k * vn + j];
            V[(uint64_t)i * vn + j] += r->alpha * s;
        }
}

/* pair order for H*psi: by output block, then record order */
static const pair_t *g_sort_pairs;
static int cmp_pair(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    uint64_t vx = g_sort_pairs[x].v_off, vy = g_sort_pairs[y].v_off;
    if (vx != vy)
        return vx < vy ? -1 : 1;
    return x < y ? -1 : x > y ? 1 : 0;
}

/* Pairs grouped by output block so that threads own disjoint outputs; the
 * order inside a block is the record order. */
static uint64_t hpsi_n(const orc_problem *p, const double *x, double *y,
                       int64_t np, int nthreads) {
    uint64_t ylen = orc_bra_len(p);
    memset(y, 0, ylen * sizeof(double));
    int64_t *ord = malloc(sizeof(int64_t) * (np + 1));
    for (int64_t i = 0; i < np; i++)
        ord[i] = i;
    g_sort_pairs = p->pairs;
    qsort(ord, np, sizeof(int64_t), cmp_pair);
    int64_t ngrp = 0;
    int64_t *gs = malloc(sizeof(int64_t) * (np + 1));
    for (int64_t i = 0; i < np; i++)
        if (i == 0 || p->pairs[ord[i]].v_off != p->pairs[ord[i - 1]].v_off)
            gs[ngrp++] = i;
    gs[ngrp] = np;
    uint64_t maxw = 1, macs = 0;
    for (int64_t i = 0; i < np; i++) {
        const pair_t *r = &p->pairs[i];
        if ((uint64_t)r->cm * r->vn > maxw)
            maxw = (uint64_t)r->cm * r->vn;
        macs += (uint64_t)r->cm * r->cn * r->vn + (uint64_t)r->vm * r->cm * r->vn;
    }
    if (nthreads <= 0)
        nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        double *work = malloc(maxw * sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t g = 0; g < ngrp; g++)
            for (int64_t i = gs[g]; i < gs[g + 1]; i++)
                rotate_pair(p, &p->pairs[ord[i]], x, y, work);
        free(work);
    }
    free(ord), free(gs);
    return macs;
}

void orc_hpsi(const orc_problem *p, const double *x, double *y, int nthreads) {
    hpsi_n(p, x, y, p->np, nthreads);
}

uint64_t orc_hpsi_sample(const orc_problem *p, const double *x, double *y,
                         uint64_t max_pairs, int nthreads) {
    int64_t n = (int64_t)max_pairs < p->np ? (int64_t)max_pairs : p->np;
    return hpsi_n(p, x, y, n, nthreads);
}

/* rotate_pair with each GEMM's rows spread over threads (every output
 * element still sums in the same order) */
static void rotate_pair_rows(const orc_problem *p, const pair_t *r,
                             const double *x, double *y, double *work,
                             int nthreads) {
    const double *A = p->ldata + r->a_off, *B = p->rdata + r->b_off;
    const double *C = x + r->c_off;
    double *V = y + r->v_off;
    const int64_t cm = r->cm, cn = r->cn, vm = r->vm, vn = r->vn;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < cm; i++)
        for (int64_t j = 0; j < vn; j++) {
            double s = 0;
            if (r->conj & 2)
                for (int64_t k = 0; k < cn; k++)
                    s += C[i * cn + k] * B[k * vn + j];
            else
                for (int64_t k = 0; k < cn; k++)
                    s += C[i * cn + k] * B[j * cn + k];
            work[i * vn + j] = s;
        }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < vm; i++)
        for (int64_t j = 0; j < vn; j++) {
            double s = 0;
            if (r->conj & 1)
                for (int64_t k = 0; k < cm; k++)
                    s += A[k * vm + i] * work[k * vn + j];
            else
                for (int64_t k = 0; k < cm; k++)
                    s += A[i * cm + k] * work[k * vn + j];
            V[i * vn + j] += r->alpha * s;
        }
}

uint64_t orc_hpsi_blocks(const orc_problem *p, const double *x, double *y,
                         int nsel, const int32_t *blocks, int nthreads) {
    memset(y, 0, orc_bra_len(p) * sizeof(double));
    if (nthreads <= 0)
        nthreads = 1;
    uint64_t macs = 0, maxw = 1;
    for (int64_t i = 0; i < p->np; i++) {
        const pair_t *r = &p->pairs[i];
        if ((uint64_t)r->cm * r->vn > maxw)
            maxw = (uint64_t)r->cm * r->vn;
    }
    double *work = malloc(maxw * sizeof(double));
    for (int b = 0; b < nsel; b++) {
        if (blocks[b] < 0 || blocks[b] >= p->bra.n)
            continue;
        const uint64_t v_off = p->bra.sh[blocks[b]];
        for (int64_t i = 0; i < p->np; i++) {
            const pair_t *r = &p->pairs[i];
            if (r->v_off != v_off)
                continue;
            rotate_pair_rows(p, r, x, y, work, nthreads);
            macs += (uint64_t)r->cm * r->cn * r->vn +
                    (uint64_t)r->vm * r->cm * r->vn;
        }
    }
    free(work);
    return macs;
}

void orc_block_macs(const orc_problem *p, uint64_t *macs) {
    memset(macs, 0, sizeof(uint64_t) * p->bra.n);
    for (int64_t i = 0; i < p->np; i++) {
        const pair_t *r = &p->pairs[i];
        int lo = 0, hi = p->bra.n - 1;
        while (lo < hi) { /* block_shifts ascend */
            int mid = (lo + hi + 1) / 2;
            if (p->bra.sh[mid] <= r->v_off)
                lo = mid;
            else
                hi = mid - 1;
        }
        macs[lo] += (uint64_t)r->cm * r->cn * r->vn +
                    (uint64_t)r->vm * r->cm * r->vn;
    }
}

/* D[ic](i,j) += alpha A(i,i) B(j,j) (matrix_functions.hpp:1204-1212) */
void orc_diag(const orc_problem *p, double *d) {
    memset(d, 0, orc_ket_len(p) * sizeof(double));
    for (int64_t t = 0; t < p->nd; t++) {
        const dterm_t *r = &p->dterms[t];
        const double *A = p->ldata + r->a_off, *B = p->rdata + r->b_off;
        for (uint32_t i = 0; i < r->an; i++)
            for (uint32_t j = 0; j < r->bn; j++)
                d[r->d_off + (uint64_t)i * r->bn + j] +=
                    r->alpha * A[(uint64_t)i * (r->an + 1)] *
                    B[(uint64_t)j * (r->bn + 1)];
    }
}

/* ------------------------------------------------------------ Davidson */
static double vdot(const double *a, const double *b, uint64_t n) {
    double s = 0;
    for (uint64_t i = 0; i < n; i++)
        s += a[i] * b[i];
    return s;
}
static void vaxpy(double a, const double *x, double *y, uint64_t n) {
    for (uint64_t i = 0; i < n; i++)
        y[i] += a * x[i];
}

/* symmetric eigenproblem by cyclic Jacobi; ascending values, row j of v is
 * the eigenvector of w[j] */
static void jacobi(int m, double *a, double *w, double *v) {
    double *z = calloc((size_t)m * m, sizeof(double));
    for (int i = 0; i < m; i++)
        z[i * m + i] = 1;
    for (int sw = 0; sw < 100; sw++) {
        double off = 0, tot = 0;
        for (int i = 0; i < m * m; i++) {
            tot += a[i] * a[i];
            if (i / m != i % m)
                off += a[i] * a[i];
        }
        if (off <= 1E-30 * tot || off == 0)
            break;
        for (int pp = 0; pp < m - 1; pp++)
            for (int q = pp + 1; q < m; q++) {
                double apq = a[pp * m + q];
                if (fabs(apq) < 1E-300)
                    continue;
                double th = (a[q * m + q] - a[pp * m + pp]) / (2 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < m; k++) {
                    double x1 = a[k * m + pp], x2 = a[k * m + q];
                    a[k * m + pp] = c * x1 - s * x2, a[k * m + q] = s * x1 + c * x2;
                }
                for (int k = 0; k < m; k++) {
                    double x1 = a[pp * m + k], x2 = a[q * m + k];
                    a[pp * m + k] = c * x1 - s * x2, a[q * m + k] = s * x1 + c * x2;
                }
                for (int k = 0; k < m; k++) {
                    double x1 = z[k * m + pp], x2 = z[k * m + q];
                    z[k * m + pp] = c * x1 - s * x2, z[k * m + q] = s * x1 + c * x2;
                }
            }
    }
    int *idx = malloc(sizeof(int) * m);
    for (int i = 0; i < m; i++)
        idx[i] = i;
    for (int i = 1; i < m; i++) /* stable insertion sort */
        for (int j = i; j > 0 && a[idx[j] * m + idx[j]] < a[idx[j - 1] * m + idx[j - 1]]; j--) {
            int t = idx[j];
            idx[j] = idx[j - 1], idx[j - 1] = t;
        }
    for (int j = 0; j < m; j++) {
        w[j] = a[idx[j] * m + idx[j]];
        for (int i = 0; i < m; i++)
            v[j * m + i] = z[i * m + idx[j]];
    }
    free(idx), free(z);
}

/* Order of the Ritz values the solver targets (iterative_matrix_functions.hpp
 * :1031-1061): ascending for Normal, else closest to / below / above shift. */
static int ord_before(int type, double shift, double a, double b) {
    if (type & 4) /* CloseTo */
        return fabs(a - shift) < fabs(b - shift);
    if (type & 2) { /* LessThan */
        if ((shift >= a) != (shift >= b))
            return shift >= a;
        if (shift >= a)
            return shift - a < shift - b;
        return a - shift > b - shift;
    }
    if (type & 1) { /* GreaterThan */
        if ((shift > a) != (shift > b))
            return shift > b;
        if (shift > a)
            return shift - a > shift - b;
        return a - shift < b - shift;
    }
    return 0;
}
static void ord_sort(int type, double shift, const double *ld, int *idx,
                     int m) {
    for (int i = 0; i < m; i++)
        idx[i] = i;
    if (!(type & 7))
        return;
    for (int i = 1; i < m; i++) /* stable insertion sort */
        for (int j = i; j > 0 && ord_before(type, shift, ld[idx[j]],
                                            ld[idx[j - 1]]); j--) {
            int t = idx[j];
            idx[j] = idx[j - 1], idx[j - 1] = t;
        }
}

/* IterativeMatrixFunctions::davidson (iterative_matrix_functions.hpp
 * :866-1173): real symmetric, DavidsonTypes Normal / GreaterThan / LessThan /
 * CloseTo with DavidsonPrecond (:66-72) or NoPrecond or the default Olsen
 * preconditioner (:93-108); orthogonal states ors projected out
 * (:912-919, 939-954, 1071-1078, 1134-1146), or with proj_weights added to H
 * as w|v><v| (:891-896, 974-976); convergence on
 * |q|^2 < conv + |lambda|^2 rel^2 once m >= k (:1062-1070, 1098-1103);
 * soft_max_iter (:966, 1150-1154).  Returns 0 on convergence, 1 at max_iter,
 * 2 when soft_max_iter stopped it.  nthreads: threads of the products (their
 * results do not depend on it). */
int orc_davidson(const orc_problem *p, double *x, int k, double conv,
                 double rel_conv, int max_iter, int soft_max_iter,
                 int def_min, int def_max, int type, double shift, int nor,
                 const double *ors_in, const double *proj_weights,
                 double *energies, int *ndav, int nthreads) {
    uint64_t n = orc_ket_len(p);
    int nwg = 0;
    if (nor > 0 && proj_weights != NULL && !(type & 1024))
        nwg = nor, nor = 0;
    if (def_min < k) def_min = k;
    if (def_max < k + k / 2) def_max = k + k / 2;
    const int nv = nwg > nor ? nwg : nor;
    double *ors = malloc(sizeof(double) * n * (nv ? nv : 1));
    if (nv)
        memcpy(ors, ors_in, sizeof(double) * n * nv);
    double *orn = malloc(sizeof(double) * (nv ? nv : 1));
    for (int i = 0; i < nor; i++) {
        for (int j = 0; j < i; j++)
            if (fabs(orn[j]) > 1E-14)
                vaxpy(-vdot(ors + j * n, ors + i * n, n) / orn[j], ors + j * n,
                      ors + i * n, n);
        orn[i] = vdot(ors + i * n, ors + i * n, n);
    }
    double *bs = malloc(sizeof(double) * n * def_max);
    double *ss = malloc(sizeof(double) * n * def_max);
    double *q = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n);
    double *tmp = malloc(sizeof(double) * n * def_max);
    double *aa = malloc(sizeof(double) * n);
    orc_diag(p, aa);
    memcpy(bs, x, sizeof(double) * n * k);
    int m = k;
    for (int i = 0; i < k; i++) {
        for (int j = 0; j < i; j++)
            vaxpy(-vdot(bs + j * n, bs + i * n, n), bs + j * n, bs + i * n, n);
        double nx = sqrt(vdot(bs + i * n, bs + i * n, n));
        if (nx * nx < 1E-14 && i > 0) {
            m = i;
            break;
        }
        for (uint64_t r = 0; r < n; r++)
            bs[i * n + r] /= nx;
    }
    for (int i = 0; i < m; i++) {
        for (int j = 0; j < nor; j++)
            if (fabs(orn[j]) > 1E-14)
                vaxpy(-vdot(ors + j * n, bs + i * n, n) / orn[j], ors + j * n,
                      bs + i * n, n);
        double nx = sqrt(vdot(bs + i * n, bs + i * n, n));
        for (uint64_t r = 0; r < n; r++)
            bs[i * n + r] /= nx;
    }
    double *ld = malloc(sizeof(double) * def_max), *al = malloc(sizeof(double) * def_max * def_max);
    double *ev = malloc(sizeof(double) * def_max * def_max);
    int *idx = malloc(sizeof(int) * def_max);
    int ck = 0, msig = 0, xiter = 0, rc = 1;
    double qq = 0;
    for (int i = 0; i < k; i++) energies[i] = 0;
    for (int i = 0; i < def_max; i++)
        idx[i] = i;
    while (xiter < max_iter && (soft_max_iter == -1 || xiter < soft_max_iter)) {
        xiter++;
        for (int i = msig; i < m; i++, msig++) {
            orc_hpsi(p, bs + i * n, ss + i * n, nthreads);
            for (int j = 0; j < nwg; j++)
                vaxpy(vdot(ors + j * n, bs + i * n, n) * proj_weights[j],
                      ors + j * n, ss + i * n, n);
        }
        for (int i = 0; i < m; i++)
            for (int j = 0; j <= i; j++)
                al[i * m + j] = al[j * m + i] = vdot(bs + i * n, ss + j * n, n);
        jacobi(m, al, ld, ev);
        memcpy(tmp, ss, sizeof(double) * n * m);
        for (int j = 0; j < m; j++) {
            memset(ss + j * n, 0, sizeof(double) * n);
            for (int i = 0; i < m; i++)
                vaxpy(ev[j * m + i], tmp + i * n, ss + j * n, n);
        }
        memcpy(tmp, bs, sizeof(double) * n * m);
        for (int j = 0; j < m; j++) {
            memset(bs + j * n, 0, sizeof(double) * n);
            for (int i = 0; i < m; i++)
                vaxpy(ev[j * m + i], tmp + i * n, bs + j * n, n);
        }
        ord_sort(type, shift, ld, idx, m);
        for (int i = 0; i < ck; i++) {
            int ii = idx[i];
            memcpy(q, ss + ii * n, sizeof(double) * n);
            vaxpy(-ld[ii], bs + ii * n, q, n);
            if (fabs(vdot(q, q, n)) >=
                conv + ld[ii] * ld[ii] * rel_conv * rel_conv) {
                ck = i;
                break;
            }
        }
        int ick = idx[ck];
        memcpy(q, ss + ick * n, sizeof(double) * n);
        vaxpy(-ld[ick], bs + ick * n, q, n);
        for (int j = 0; j < nor; j++)
            if (fabs(orn[j]) > 1E-14)
                vaxpy(-vdot(ors + j * n, q, n) / orn[j], ors + j * n, q, n);
        qq = vdot(q, q, n);
        const double lk = ld[ick];
        if (type & 32) { /* DavidsonPrecond (:66-72) */
            for (uint64_t r = 0; r < n; r++)
                if (fabs(lk - aa[r]) > 1E-12)
                    q[r] /= lk - aa[r];
        } else if (!(type & 64)) { /* Olsen (:93-108) */
            const double *c = bs + ick * n;
            for (uint64_t r = 0; r < n; r++) {
                t[r] = c[r];
                double dd = lk - aa[r];
                if (fabs(dd) > 1E-12)
                    t[r] /= dd, q[r] /= dd;
            }
            vaxpy(-vdot(c, q, n) / vdot(c, t, n), t, q, n);
        }
        for (int i = 0; i <= ck; i++)
            energies[i] = ld[idx[i]];
        for (int i = ck + 1; i < k; i++)
            energies[i] = 0;
        if (fabs(qq) < conv + energies[ck] * energies[ck] * rel_conv * rel_conv &&
            m >= k) {
            ck++;
            if (ck == k) {
                rc = 0;
                break;
            }
        } else {
            int defl = 0;
            if (m >= def_max) {
                m = msig = def_min;
                defl = (type & 7) != 0;
            }
            if (defl) { /* keep the targeted Ritz pairs (:1105-1133) */
                for (int j = 0; j < m; j++)
                    memcpy(tmp + j * n, bs + idx[j] * n, sizeof(double) * n);
                memcpy(bs, tmp, sizeof(double) * n * m);
                for (int j = 0; j < m; j++)
                    memcpy(tmp + j * n, ss + idx[j] * n, sizeof(double) * n);
                memcpy(ss, tmp, sizeof(double) * n * m);
            }
            for (int j = 0; j < m; j++)
                vaxpy(-vdot(bs + j * n, q, n), bs + j * n, q, n);
            for (int j = 0; j < nor; j++)
                if (fabs(orn[j]) > 1E-14)
                    vaxpy(-vdot(ors + j * n, q, n) / orn[j], ors + j * n, q, n);
            double nq = sqrt(vdot(q, q, n));
            for (uint64_t r = 0; r < n; r++)
                bs[m * n + r] = q[r] / nq;
            m++;
        }
        if (xiter == soft_max_iter) {
            rc = 2;
            break;
        }
    }
    for (int i = 0; i < k; i++)
        memcpy(x + i * n, bs + idx[i] * n, sizeof(double) * n);
    *ndav = xiter;
    free(bs), free(ss), free(q), free(t), free(tmp), free(aa), free(ld);
    free(al), free(ev), free(idx), free(ors), free(orn);
    return rc;
}

/* dE/dA and dE/dB of E = sum alpha <V, op(A) C op(B)^T> (the partial
 * contractions of operator_functions.hpp:354-472) */
void orc_grad_ops(const orc_problem *p, const double *v, const double *c,
                  double *gl, double *gr) {
    uint64_t llen = 0, rlen = 0;
    for (int i = 0; i < p->nlo; i++) {
        uint64_t e = p->lo_off[i] + info_total(&p->li[p->lo_info[i]]);
        if (e > llen) llen = e;
    }
    for (int i = 0; i < p->nro; i++) {
        uint64_t e = p->ro_off[i] + info_total(&p->ri[p->ro_info[i]]);
        if (e > rlen) rlen = e;
    }
    memset(gl, 0, llen * sizeof(double));
    memset(gr, 0, rlen * sizeof(double));
    for (int64_t ip = 0; ip < p->np; ip++) {
        const pair_t *r = &p->pairs[ip];
        const double *A = p->ldata + r->a_off, *B = p->rdata + r->b_off;
        const double *C = c + r->c_off, *V = v + r->v_off;
        const uint32_t cm = r->cm, cn = r->cn, vm = r->vm, vn = r->vn;
        /* B'[k][j] (cn x vn) = (conj&2) ? B[k][j] : B[j][k] */
#define BP(k, j) ((r->conj & 2) ? B[(uint64_t)(k)*vn + (j)] : B[(uint64_t)(j)*cn + (k)])
#define OA(i, k) ((r->conj & 1) ? A[(uint64_t)(k)*vm + (i)] : A[(uint64_t)(i)*cm + (k)])
        /* dE/d op(A)[i][k] = alpha sum_j V[i][j] work[k][j],
         * work[k][j] = sum_l C[k][l] B'[l][j] */
        for (uint32_t kk = 0; kk < cm; kk++)
            for (uint32_t j = 0; j < vn; j++) {
                double w = 0;
                for (uint32_t l = 0; l < cn; l++)
                    w += C[(uint64_t)kk * cn + l] * BP(l, j);
                for (uint32_t i = 0; i < vm; i++) {
                    double g = r->alpha * V[(uint64_t)i * vn + j] * w;
                    if (r->conj & 1)
                        gl[r->a_off + (uint64_t)kk * vm + i] += g;
                    else
                        gl[r->a_off + (uint64_t)i * cm + kk] += g;
                }
            }
        /* dE/d B'[l][j] = alpha sum_k (op(A) C)[.]: u[i][l] = sum_k op(A)[i][k] C[k][l];
         * dB'[l][j] = alpha sum_i u[i][l] V[i][j] */
        for (uint32_t i = 0; i < vm; i++)
            for (uint32_t l = 0; l < cn; l++) {
                double u = 0;
                for (uint32_t kk = 0; kk < cm; kk++)
                    u += OA(i, kk) * C[(uint64_t)kk * cn + l];
                for (uint32_t j = 0; j < vn; j++) {
                    double g = r->alpha * u * V[(uint64_t)i * vn + j];
                    if (r->conj & 2)
                        gr[r->b_off + (uint64_t)l * vn + j] += g;
                    else
                        gr[r->b_off + (uint64_t)j * cn + l] += g;
                }
            }
#undef BP
#undef OA
    }
}
