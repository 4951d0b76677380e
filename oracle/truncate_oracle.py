"""ORACLE -- TEST INFRASTRUCTURE ONLY.

numpy restatement of block2's density-matrix truncation of a two-site
wavefunction, the checker for libb2g's b2g_split_density_matrix and
b2g_syev_batched.  Imported only by tests/ (and bench.py's CPU leg).

Follows:
  SZLongLong / SU2LongLong get_bra, get_ket, operator-, multiplicity
      src/core/symmetry.hpp:656-732, 1184-1290
  SparseMatrixInfo::initialize_dm           src/core/sparse_matrix.hpp:677-724
  OperatorFunctions::trans_product (noise 0) src/core/operator_functions.hpp:776-835
  MovingEnvironment::truncate_density_matrix src/dmrg/moving_environment.hpp:3716-3853
  MovingEnvironment::split_density_matrix    src/dmrg/moving_environment.hpp:4250-4345
  trans_product noise, Random::fill          src/core/operator_functions.hpp:776-835,
                                             src/core/utils.hpp:226-252
  rotation_matrix_info_from_density_matrix / wavefunction_info_from_density_matrix
                                             src/dmrg/moving_environment.hpp:4013-4083
  GMatrixFunctions::eigs (xsyev)             src/core/matrix_functions.hpp:1672-1688
The eigenproblems go to numpy.linalg.eigh (LAPACK syevd); eigenvalues come
back ascending like xsyev's.  Eigenvector signs are fixed here by making the
largest-magnitude component of every vector positive (the first one on
ties), the convention libb2g documents; the reference leaves them to LAPACK,
so comparisons with its dumps use sign-free quantities.
"""
import numpy as np

SZ, SU2 = 0, 1
M16 = 0xFFFF


def _n(q):
    return ((q >> 48) & M16) - (1 << 16 if (q >> 63) & 1 else 0)


def q_neg(sym, q):
    """-q: N (and 2Sz for SZ) negated, 2S fields kept."""
    n = (-_n(q)) & M16
    if sym == SZ:
        tz = (q >> 16) & M16
        tz = (-(tz - (1 << 16) if tz & 0x8000 else tz)) & M16
        return (n << 48) | (tz << 16) | (q & M16)
    return (n << 48) | (q & 0xFFFFFFFFFFFF)


def q_get_ket(sym, q):
    """ket part: SU2 copies 2S into the 2S_low field; SZ: q itself."""
    if sym == SZ:
        return q
    return (q & 0xFFFF0000FFFFFFFF) | ((q & 0xFFFF0000) << 16)


def q_get_bra(sym, q, dq):
    """bra part of a composite label for delta quantum dq."""
    if sym == SZ:
        n = (_n(q) + _n(dq)) & M16
        tq, td = (q >> 16) & M16, (dq >> 16) & M16
        tz = (tq + td) & M16
        return (n << 48) | (tz << 16) | ((q ^ dq) & M16)
    n = (_n(q) + _n(dq)) & M16
    lo = (q >> 32) & M16
    return (n << 48) | (lo << 32) | (lo << 16) | ((q ^ dq) & M16)


def multiplicity(sym, q):
    return 1 if sym == SZ else ((q >> 16) & M16) + 1


def signed_eigh(a):
    """eigenvalues ascending, eigenvectors as rows, each signed so that its
    largest-magnitude component (first on ties) is positive"""
    w, v = np.linalg.eigh(a)
    v = v.T.copy()
    for j in range(v.shape[0]):
        i = int(np.argmax(np.abs(v[j])))
        if v[j, i] < 0:
            v[j] = -v[j]
    return w, v


def density_matrix(sym, info, psi, trace_right, scale=1.0):
    """info: dict(quanta, nbra, nket, shift, dq).  Returns (quanta, dims,
    blocks) of the density matrix, blocks dense (dim x dim)."""
    qs = info["quanta"]
    dq = int(info["dq"])
    lab = [q_get_bra(sym, int(q), dq) if trace_right
           else q_neg(sym, q_get_ket(sym, int(q))) for q in qs]
    uq = sorted(set(lab))
    dims = [0] * len(uq)
    for i, l in enumerate(lab):
        dims[uq.index(l)] = int(info["nbra"][i] if trace_right
                                else info["nket"][i])
    blocks = [np.zeros((d, d)) for d in dims]
    for i, l in enumerate(lab):
        m, n, o = int(info["nbra"][i]), int(info["nket"][i]), int(info["shift"][i])
        x = psi[o:o + m * n].reshape(m, n)
        b = uq.index(l)
        blocks[b] += scale * (x @ x.T if trace_right else x.T @ x)
    return uq, dims, blocks


def block2_uniform(seed, n, a, b):
    """n draws of block2's Random::fill (src/core/utils.hpp:226-252) after
    Random::rand_seed(seed): std::mt19937(seed) through libstdc++'s
    uniform_real_distribution<double>(a, b), i.e. generate_canonical with
    two 32-bit outputs x0, x1: u = (x0 + x1 2^32) / 2^64 (below 1), then
    u (b - a) + a.  numpy's RandomState(seed) runs the same MT19937 stream
    (init_genrand seeding); randint over [0, 2^32) returns its raw 32-bit
    outputs."""
    rs = np.random.RandomState(seed)
    x = rs.randint(0, 2 ** 32, size=2 * n, dtype=np.uint64).astype(np.float64)
    u = (x[0::2] + x[1::2] * 2.0 ** 32) / 2.0 ** 64
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return u * (b - a) + a


def add_noise(sym, info, blocks, uq, trace_right, noise, noise_type, seed,
              n_roots=1):
    """OperatorFunctions::trans_product's noise (src/core/operator_functions.
    hpp:776-835) as MovingEnvironment::density_matrix calls it with
    sqrt(noise) (src/dmrg/moving_environment.hpp:3554-3577), once per root
    for density_matrix_with_multi_target (:3582-3608: each root draws its
    own tensor from the continuing stream).  Returns the density-matrix
    blocks as block2 stores them (density-matrix noise makes them
    asymmetric)."""
    qs = info["quanta"]
    dq = int(info["dq"])
    lab = [q_get_bra(sym, int(q), dq) if trace_right
           else q_neg(sym, q_get_ket(sym, int(q))) for q in qs]
    out = [b.copy() for b in blocks]
    if noise_type == 1:  # Wavefunction
        n = sum(int(m) * int(k) for m, k in zip(info["nbra"], info["nket"]))
        ts = block2_uniform(seed, n * n_roots, -0.5, 0.5)
        for r in range(n_roots):
            t = ts[r * n:(r + 1) * n]
            _, _, tb = density_matrix(sym, info, t, trace_right,
                                      noise / float(t @ t))
            out = [b + x for b, x in zip(out, tb)]
        return out
    # DensityMatrix: T over the density matrix's layout, once per
    # wavefunction block of each sector
    n = sum(b.size for b in blocks)
    ts = block2_uniform(seed, n * n_roots, 0.0, 1.0)
    for r in range(n_roots):
        t = ts[r * n:(r + 1) * n]
        ns = noise / float(np.sqrt(t @ t))
        o = 0
        for bi, b in enumerate(blocks):
            cnt = sum(1 for l in lab if uq.index(l) == bi)
            out[bi] = out[bi] + cnt * ns * t[o:o + b.size].reshape(b.shape)
            o += b.size
    return out


def lower_symmetric(b):
    """the symmetric matrix of b's lower triangle (block2's xsyev "U" on
    row-major storage reads the row-major lower triangle)"""
    return np.tril(b) + np.tril(b, -1).T


def truncate(sym, quanta, w, k, cutoff, trunc_type):
    """The reference's selection: w = list of ascending eigenvalue arrays per
    density-matrix block.  Returns (ss, error) with ss the kept (block,
    index) pairs in (block, index) order."""
    wr = []
    for b, wb in enumerate(w):
        mult = float(multiplicity(sym, quanta[b]))
        if trunc_type & 1:
            wr.append(wb * (1.0 / mult))
        elif trunc_type & 2:
            wr.append(wb * mult)
        else:
            wr.append(wb)
    ss = [(b, j) for b in range(len(w)) for j in range(len(w[b]))]
    k_total = len(ss)
    error = 0.0
    if k != -1:
        ss.sort(key=lambda x: -wr[x[0]][x[1]])
        keep = trunc_type // 16
        if keep == 0:
            for i in range(k, k_total):
                x = w[ss[i][0]][ss[i][1]]
                if x > 0:
                    error += x
            k = min(k, k_total)
            while k > 1 and wr[ss[k - 1][0]][ss[k - 1][1]] < cutoff:
                x = w[ss[k - 1][0]][ss[k - 1][1]]
                if x > 0:
                    error += x
                k -= 1
            ss = ss[:k]
        else:
            mask = [0] * len(w)
            smask = [0] * k_total
            for i in range(k_total):
                mask[ss[i][0]] += 1
                smask[i] = mask[ss[i][0]] > keep
            for i in range(k, k_total):
                x = w[ss[i][0]][ss[i][1]]
                if x > 0 and smask[i]:
                    error += x
            k = min(k, k_total)
            while k > 1 and wr[ss[k - 1][0]][ss[k - 1][1]] < cutoff:
                x = w[ss[k - 1][0]][ss[k - 1][1]]
                if x > 0 and smask[k - 1]:
                    error += x
                k -= 1
            ss = ss[:k] + [ss[i] for i in range(k, k_total) if not smask[i]]
        ss.sort()
    return ss, error


def split_density_matrix(sym, info, psi, k, cutoff, trace_right,
                         normalize=True, trunc_type=0, scale=1.0, noise=0.0,
                         noise_type=0, seed=0):
    """Returns dict(error, eigvals (list per block), dm=(quanta, dims,
    blocks), left=(quanta, nbra, nket, data), right=(...), spectra).
    noise > 0 adds block2's wavefunction (noise_type 1) or density-matrix
    (2) noise drawn as after Random::rand_seed(seed); dm then holds the
    blocks as block2 stores them, and the eigenproblems take their lower
    triangles."""
    uq, dims, blocks = density_matrix(sym, info, psi, trace_right, scale)
    if noise:
        blocks = add_noise(sym, info, blocks, uq, trace_right, noise,
                           noise_type, seed)
    eig = [signed_eigh(lower_symmetric(b)) if b.size
           else (np.zeros(0), np.zeros((0, 0))) for b in blocks]
    w = [e[0] for e in eig]
    ss, error = truncate(sym, uq, w, k, cutoff, trunc_type)
    with_multi = bool(trunc_type & 8)
    spectra = []
    for b, wb in enumerate(w):
        reps = multiplicity(sym, uq[b]) if with_multi else 1
        for _ in range(reps):
            spectra += list(np.sqrt(np.maximum(0.0, wb)))
    ilr, im = [], []
    for b, _ in ss:
        if not ilr or ilr[-1] != b:
            ilr.append(b)
            im.append(1)
        else:
            im[-1] += 1
    # rotation matrix
    rq, rnb, rnk, rdata = [], [], [], []
    iss = 0
    for i, b in enumerate(ilr):
        vecs = np.stack([eig[b][1][ss[iss + j][1]] for j in range(im[i])])
        iss += im[i]
        rq.append(uq[b])
        if trace_right:
            rnb.append(dims[b]), rnk.append(im[i]), rdata.append(vecs.T)
        else:
            rnb.append(im[i]), rnk.append(dims[b]), rdata.append(vecs)
    # wavefunction factor
    qs = info["quanta"]
    dq = int(info["dq"])
    lab = [q_get_bra(sym, int(q), dq) if trace_right
           else q_neg(sym, q_get_ket(sym, int(q))) for q in qs]
    wq, wblk = [], []
    for i, b in enumerate(ilr):
        for iw in range(len(qs)):
            if uq.index(lab[iw]) != b:
                continue
            m, n, o = (int(info["nbra"][iw]), int(info["nket"][iw]),
                       int(info["shift"][iw]))
            x = psi[o:o + m * n].reshape(m, n)
            wq.append(int(qs[iw]))
            wblk.append(rdata[i].T @ x if trace_right else x @ rdata[i].T)
    order = sorted(range(len(wq)), key=lambda j: wq[j])
    wq = [wq[j] for j in order]
    wblk = [wblk[j] for j in order]
    if normalize:
        nrm = np.sqrt(sum(float((b * b).sum()) for b in wblk))
        if nrm > 0:
            wblk = [b / nrm for b in wblk]
    rot = (rq, rnb, rnk, rdata)
    wav = (wq, [b.shape[0] for b in wblk], [b.shape[1] for b in wblk], wblk)
    return dict(error=error, eigvals=w, dm=(uq, dims, blocks),
                left=rot if trace_right else wav,
                right=wav if trace_right else rot, spectra=spectra,
                ss=ss)


def multi_split_density_matrix(sym, info, psis, weights, k, cutoff,
                               trace_right, normalize=True, trunc_type=0,
                               scale=1.0, noise=0.0, noise_type=0, seed=0):
    """State-averaged roots of one structure (density_matrix_with_multi_
    target + multi_split_density_matrix, src/dmrg/moving_environment.hpp:
    3582-3608, 4347-4462): density matrix sum_r w_r scale psi_r psi_r^T,
    one rotation matrix, each root's factor normalised alone.  Returns
    dict(error, eigvals, dm, rot=(quanta, nbra, nket, data), wav=[per root
    (quanta, nbra, nket, data)], spectra, ss)."""
    base = None
    for r, (p, w) in enumerate(zip(psis, weights)):
        uq, dims, blocks = density_matrix(sym, info, p, trace_right,
                                          w * scale)
        if base is None:
            base = (uq, dims, [b.copy() for b in blocks])
        else:
            for a, b in zip(base[2], blocks):
                a += b
    uq, dims, blocks = base
    if noise:
        blocks = add_noise(sym, info, blocks, uq, trace_right, noise,
                           noise_type, seed, len(psis))
    eig = [signed_eigh(lower_symmetric(b)) if b.size
           else (np.zeros(0), np.zeros((0, 0))) for b in blocks]
    w = [e[0] for e in eig]
    ss, error = truncate(sym, uq, w, k, cutoff, trunc_type)
    spectra = []
    for b, wb in enumerate(w):
        reps = multiplicity(sym, uq[b]) if trunc_type & 8 else 1
        for _ in range(reps):
            spectra += list(np.sqrt(np.maximum(0.0, wb)))
    ilr, im = [], []
    for b, _ in ss:
        if not ilr or ilr[-1] != b:
            ilr.append(b)
            im.append(1)
        else:
            im[-1] += 1
    rq, rnb, rnk, rdata = [], [], [], []
    iss = 0
    for i, b in enumerate(ilr):
        vecs = np.stack([eig[b][1][ss[iss + j][1]] for j in range(im[i])])
        iss += im[i]
        rq.append(uq[b])
        if trace_right:
            rnb.append(dims[b]), rnk.append(im[i]), rdata.append(vecs.T)
        else:
            rnb.append(im[i]), rnk.append(dims[b]), rdata.append(vecs)
    qs = info["quanta"]
    dq = int(info["dq"])
    lab = [q_get_bra(sym, int(q), dq) if trace_right
           else q_neg(sym, q_get_ket(sym, int(q))) for q in qs]
    wavs = []
    for p in psis:
        wq, wblk = [], []
        for i, b in enumerate(ilr):
            for iw in range(len(qs)):
                if uq.index(lab[iw]) != b:
                    continue
                m, n, o = (int(info["nbra"][iw]), int(info["nket"][iw]),
                           int(info["shift"][iw]))
                x = p[o:o + m * n].reshape(m, n)
                wq.append(int(qs[iw]))
                wblk.append(rdata[i].T @ x if trace_right else x @ rdata[i].T)
        order = sorted(range(len(wq)), key=lambda j: wq[j])
        wq = [wq[j] for j in order]
        wblk = [wblk[j] for j in order]
        if normalize:
            nrm = np.sqrt(sum(float((b * b).sum()) for b in wblk))
            if nrm > 0:
                wblk = [b / nrm for b in wblk]
        wavs.append((wq, [b.shape[0] for b in wblk],
                     [b.shape[1] for b in wblk], wblk))
    return dict(error=error, eigvals=w, dm=(uq, dims, blocks),
                rot=(rq, rnb, rnk, rdata), wav=wavs, spectra=spectra, ss=ss)
