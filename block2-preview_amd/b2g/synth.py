"""Seeded synthetic middle-site effective Hamiltonian (SU(2), no point group).

Builds the block-sparse structure block2 hands to EffectiveHamiltonian at the
middle site of an n-orbital, n-electron singlet DMRG with bond dimension M:

* left and right bases (environment x one site) as SU(2) sectors (N, 2S) with
  a heavy-tailed dimension profile (largest blocks near N = n/2, S = 0,
  dimensions 1 .. max_dim, total about 3 M multiplets per side);
* the singlet two-site wavefunction: one block per (left, right) sector pair
  with N_L + N_R = n and S_L = S_R;
* operator infos for the delta quanta of the quantum-chemistry normal /
  complementary operators: (0,0) (H, 1, B0, Q0), (0,1) (B1, Q1),
  (+-1, 1/2) (a^+, a, R, R^+), (+-2, 0) and (+-2, 1) (A, P);
* strings  f * (L_k x R_k +- L_k^T x R_k^T)  (conj 0 and 3; minus for
  fermionic pairs), so H is symmetric, plus  H_L x 1 + 1 x H_R  with
  sorted diagonal block energies so the low end of the spectrum is gapped
  like a physical site.

Operator data are seeded random numbers; the structure, not the values,
determines the work of one H*psi.
"""
import numpy as np

from .problem import InfoSet, OpSet, Problem

SU2 = 1


def q_su2(n, twos, twos_low=None):
    """Packed SU2LongLong label (src/core/symmetry.hpp:1184-1290 layout)."""
    if twos_low is None:
        twos_low = twos
    return (((n & 0xFFFF) << 48) | ((twos_low & 0xFFFF) << 32)
            | ((twos & 0xFFFF) << 16))


def _fill(n, key):
    """n reproducible values uniform in [-sqrt(3), sqrt(3)) (unit variance)
    from a vectorised splitmix64 hash of (key, index) -- much faster than a
    sequential generator for the hundreds of millions of operator entries."""
    with np.errstate(over="ignore"):
        z = np.arange(n, dtype=np.uint64) + np.uint64(
            (key * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return (u - 0.5) * (2.0 * np.sqrt(3.0))


def _tri(a, b, c):
    return (a + b + c) % 2 == 0 and abs(a - b) <= c <= a + b


def sectors(n_orb, n_elec, m, max_dim, rng, spread_n=None, max_twos=12):
    """Sector list [(N, 2S, dim)] of one side at the middle bond."""
    n0 = n_elec // 2
    spread_n = spread_n or max(3, n_orb // 5)
    out = []
    for dn in range(-2 * spread_n, 2 * spread_n + 1):
        N = n0 + dn
        if N < 0 or N > n_orb * 2:
            continue
        for tw in range(N % 2, max_twos + 1, 2):
            w = np.exp(-(dn / spread_n) ** 2 - 0.55 * tw)
            w *= np.exp(0.35 * rng.standard_normal())
            out.append([N, tw, w])
    w = np.array([x[2] for x in out])
    # scale so that the side holds about 3 M multiplets, capped at max_dim
    target = 3 * m
    s = target / w.sum()
    for _ in range(50):
        d = np.clip(np.round(w * s), 0, max_dim)
        if d.sum() < target:
            s *= 1.05
        else:
            break
    res = [(N, tw, int(di)) for (N, tw, _), di in zip(out, d) if di >= 1]
    return res


def _info(qs, nb, nk):
    order = np.argsort(np.array(qs, dtype=np.uint64), kind="stable")
    q = np.array(qs, dtype=np.uint64)[order]
    nb = np.array(nb, dtype=np.uint32)[order]
    nk = np.array(nk, dtype=np.uint32)[order]
    sh = np.zeros(len(q), dtype=np.uint64)
    if len(q):
        sh[1:] = np.cumsum(nb.astype(np.uint64) * nk.astype(np.uint64))[:-1]
    return q, nb, nk, sh


def op_info(secs, dn, dtw):
    """Blocks of an operator with delta quantum (dn, 2S = dtw) on `secs`."""
    idx = {(N, tw): d for N, tw, d in secs}
    qs, nb, nk = [], [], []
    for Nk, twk, dk in secs:
        for Nb, twb, db in secs:
            if Nb != Nk + dn or not _tri(twk, dtw, twb):
                continue
            # composite label: ket (N, 2S) with 2S_low = bra 2S
            qs.append(q_su2(Nk, twk, twb))
            nb.append(db)
            nk.append(dk)
    return _info(qs, nb, nk)


def middle_site(n_orb=40, m=3000, max_dim=600, n_pairs=None, seed=1234,
                coupling=0.05, pair_scale=1, data="host"):
    """Synthetic middle-site Problem.  n_pairs: number of random operator
    pairs per delta-quantum type (default: 2 per bosonic and 4 per fermionic
    type, times pair_scale).

    data="host" fills every operator on the host; data="device" leaves the
    random coupling operators unfilled (their seeds are kept in
    `problem.fill_keys`) so that fill_device() generates them on the GPU."""
    rng = np.random.default_rng(seed)
    n_elec = n_orb
    left = sectors(n_orb, n_elec, m, max_dim, rng)
    right = sectors(n_orb, n_elec, m, max_dim, rng)
    # two-site wavefunction (singlet): blocks (l, r), N_l + N_r = n, S_l = S_r
    target = q_su2(n_elec, 0)
    rmap = {(N, tw): d for N, tw, d in right}
    kq, knb, knk = [], [], []
    for N, tw, d in left:
        dr = rmap.get((n_elec - N, tw))
        if dr:
            # combine(target, l, -r): ket = -r with 2S_low = 2S_l
            kq.append(q_su2(-(n_elec - N), tw, tw))
            knb.append(d)
            knk.append(dr)
    kq, knb, knk, ksh = _info(kq, knb, knk)
    ket = InfoSet([target], [0], [len(kq)], [0], kq, knb, knk, ksh)
    # operator types: (dn, 2S) on the left; the right partner is (-dn, 2S)
    types = [(0, 0), (0, 2), (1, 1), (-1, 1), (2, 0), (2, 2), (-2, 0), (-2, 2)]
    if n_pairs is None:
        n_pairs = {k: v * pair_scale for k, v in
                   {(0, 0): 2, (0, 2): 2, (1, 1): 4, (-1, 1): 4, (2, 0): 2,
                    (2, 2): 2, (-2, 0): 2, (-2, 2): 2}.items()}
    linf, rinf = {}, {}
    for dn, tw in types:
        linf[(dn, tw)] = op_info(left, dn, tw)
        rinf[(-dn, tw)] = op_info(right, -dn, tw)

    def pack(infos):
        keys = sorted(infos, key=lambda k: q_su2(k[0], k[1]))
        dq, fe, nn, off, qq, nb, nk, sh = [], [], [], [], [], [], [], []
        o = 0
        for k in keys:
            q, b, kk, s = infos[k]
            dq.append(q_su2(k[0], k[1]))
            fe.append(k[0] % 2 != 0)
            nn.append(len(q))
            off.append(o)
            o += len(q)
            qq.append(q), nb.append(b), nk.append(kk), sh.append(s)
        return keys, InfoSet(dq, fe, nn, off, np.concatenate(qq),
                             np.concatenate(nb), np.concatenate(nk),
                             np.concatenate(sh))

    lkeys, linfo = pack(linf)
    rkeys, rinfo = pack(rinf)
    lidx = {k: i for i, k in enumerate(lkeys)}
    ridx = {k: i for i, k in enumerate(rkeys)}

    ldata, rdata = [], []
    lo = dict(dq=[], f=[], info=[], off=[])
    ro = dict(dq=[], f=[], info=[], off=[])
    lpos, rpos = [0], [0]

    fill_keys = {0: [], 1: []}  # (op index, seed, scale) per side

    def add_op(side, key, data, n=None, seed_key=None, scale=1.0):
        dct, pos, buf = ((lo, lpos, ldata) if side == 0 else
                         (ro, rpos, rdata))
        if data is None:
            fill_keys[side].append((len(dct["dq"]), seed_key, scale))
            data = np.zeros(0)
            length = n
        else:
            length = len(data)
        inf = linfo if side == 0 else rinfo
        ii = (lidx if side == 0 else ridx)[key]
        dct["dq"].append(int(inf.dq[ii]))
        dct["f"].append(1.0)
        dct["info"].append(ii)
        dct["off"].append(pos[0])
        buf.append((pos[0], data))
        pos[0] += length
        return len(dct["dq"]) - 1

    def block_energies(secs, info_key, inf_map, scale, n0):
        """Diagonal operator (dq = 0): sorted energies per sector, raised
        away from (N0, S=0); the lowest state of the central sector sits a
        unit below the rest so the site's ground state is gapped."""
        q, nb, nk, sh = inf_map[info_key]
        data = np.zeros(int(sh[-1] + nb[-1].astype(np.uint64) * nk[-1]))
        for i in range(len(q)):
            qi = int(q[i])
            tw, tl = (qi >> 16) & 0xFFFF, (qi >> 32) & 0xFFFF
            if nb[i] == nk[i] and tw == tl:
                d = int(nb[i])
                nn = ((qi >> 48) + (1 << 15)) % (1 << 16) - (1 << 15)
                e = np.sort(rng.uniform(0, scale, d))
                e += 0.3 * abs(nn - n0) + 0.2 * tw
                if nn == n0 and tw == 0:
                    e[0] = -1.0
                data[int(sh[i]):int(sh[i]) + d * d] = np.diag(e).ravel()
        return data

    def identity(info_key, inf_map):
        q, nb, nk, sh = inf_map[info_key]
        data = np.zeros(int(sh[-1] + nb[-1].astype(np.uint64) * nk[-1]))
        for i in range(len(q)):
            if nb[i] == nk[i] and ((int(q[i]) >> 32) & 0xFFFF) == \
                    ((int(q[i]) >> 16) & 0xFFFF):
                d = int(nb[i])
                data[int(sh[i]):int(sh[i]) + d * d] = np.eye(d).ravel()
        return data

    t_conj, t_a, t_b, t_f = [], [], [], []
    # H_L x 1 + 1 x H_R (gapped diagonal part)
    hl = add_op(0, (0, 0), block_energies(left, (0, 0), linf, 4.0,
                                          n_elec // 2))
    il = add_op(0, (0, 0), identity((0, 0), linf))
    hr = add_op(1, (0, 0), block_energies(right, (0, 0), rinf, 4.0,
                                          n_elec - n_elec // 2))
    ir = add_op(1, (0, 0), identity((0, 0), rinf))
    for a, b in ((hl, ir), (il, hr)):
        t_conj.append(0), t_a.append(a), t_b.append(b), t_f.append(1.0)
    # random couplings, each with its transposed partner (conj 0 and 3)
    for dn, tw in types:
        for _ in range(n_pairs.get((dn, tw), 0)):
            lq = linf[(dn, tw)]
            rq = rinf[(-dn, tw)]
            nl = int(lq[3][-1] + lq[1][-1].astype(np.uint64) * lq[2][-1])
            nr = int(rq[3][-1] + rq[1][-1].astype(np.uint64) * rq[2][-1])
            key = int(rng.integers(1 << 62))
            sc = 1.0 / np.sqrt(max_dim)
            if data == "host":
                a = add_op(0, (dn, tw), _fill(nl, key) * sc)
                b = add_op(1, (-dn, tw), _fill(nr, key + 1) * sc)
            else:
                a = add_op(0, (dn, tw), None, nl, key, sc)
                b = add_op(1, (-dn, tw), None, nr, key + 1, sc)
            f = coupling * rng.uniform(0.5, 1.5)
            # the transposed partner of a fermionic pair carries a sign
            # (exchange of the two odd operators)
            for cj, ff in ((0, f), (3, -f if dn % 2 else f)):
                t_conj.append(cj), t_a.append(a), t_b.append(b), t_f.append(ff)

    def assemble(parts, total):
        if data == "host":
            return np.concatenate([d for _, d in parts])
        # device mode: host copy holds only the explicitly built operators;
        # fill_device() produces the full buffer
        return None

    lops = OpSet(lo["dq"], lo["f"], lo["info"], lo["off"],
                 assemble(ldata, lpos[0]))
    rops = OpSet(ro["dq"], ro["f"], ro["info"], ro["off"],
                 assemble(rdata, rpos[0]))
    meta = dict(kind="synthetic-middle-site", n_orb=n_orb, m=m,
                max_dim=max_dim, seed=seed, pair_scale=pair_scale,
                n_terms=len(t_conj), n_left_sectors=len(left),
                n_right_sectors=len(right), n_psi_blocks=len(kq))
    pb = Problem(SU2, q_su2(0, 0), q_su2(0, 0), ket, ket, linfo, rinfo,
                 lops, rops, t_conj, t_a, t_b, t_f, meta)
    pb.fill_keys = fill_keys
    pb.explicit = {0: [(o, d) for o, d in ldata if len(d)],
                   1: [(o, d) for o, d in rdata if len(d)]}
    pb.data_len = {0: lpos[0], 1: rpos[0]}
    return pb


def fill_device(pb, device):
    """Operator buffers of a data="device" problem, generated on the GPU:
    returns (left, right) float64 tensors.  Random couplings are uniform in
    [-sqrt(3), sqrt(3)) * scale from torch's seeded device generator."""
    import torch
    out = []
    for side, ops in ((0, pb.lops), (1, pb.rops)):
        buf = torch.zeros(pb.data_len[side], dtype=torch.float64,
                          device=device)
        for off, d in pb.explicit[side]:
            buf[off:off + len(d)] = torch.from_numpy(d).to(device)
        g = torch.Generator(device=device)
        for k, key, scale in pb.fill_keys[side]:
            o = int(ops.off[k])
            n = _op_len(pb, side, k)
            g.manual_seed(int(key) & 0x7FFFFFFFFFFFFFFF)
            seg = buf[o:o + n]
            seg.uniform_(-np.sqrt(3.0), np.sqrt(3.0), generator=g)
            seg.mul_(scale)
        out.append(buf)
    return out[0], out[1]


def _op_len(pb, side, k):
    ops, infos = (pb.lops, pb.linfo) if side == 0 else (pb.rops, pb.rinfo)
    return infos.total(int(ops.info[k]))
