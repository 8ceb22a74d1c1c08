"""Float64 CPU emulation of the block-plan assembly kernel (test helper).

Decodes the exported plan exactly as the HIP kernel does and checks every
index against its buffer: node slots < halo size, accumulator offsets inside
the block, each CSR value written by exactly one run.  A record may list its
element's vertices in any order (the plan relabels them to spread LDS bank
accesses), so the element matrix is evaluated in the record's order.
"""
import numpy as np


def decode(rec, nb):
    lo = int(rec[0]) | (int(rec[1]) << 32)
    hi = int(rec[2]) | (int(rec[3]) << 32)
    loc = [(lo >> (10 * a)) & 1023 for a in range(nb)]

    def rank(a, j):
        bit = 40 + 6 * (a * (nb - 1) + j)
        return (lo >> bit) & 63 if bit < 64 else (hi >> (bit - 64)) & 63
    return loc, rank


def vertex_order(rec, nb):
    """Element-local index of each record vertex (bits 112-119, 2 bits each)."""
    code = (int(rec[3]) >> 16) & 0xFF
    return [(code >> (2 * a)) & 3 for a in range(nb)]


def det_grid(values):
    """The deterministic mode's grid step for a row whose entries are
    ``values``: 2^k with M = max|v| in [2^e, 2^(e+1)), k = e + 1 - 44."""
    M = float(np.max(np.abs(values))) if len(values) else 0.0
    if not M > 0.0:
        return 1.0
    e = int(np.frexp(M)[1]) - 1
    return float(np.ldexp(1.0, e + 1 - 44))


def emulate_matrix(topo, arrs, element_matrix, det=False, shuffle=None):
    """``element_matrix(e, x[nb, d]) -> [nb, nb]``; returns the CSR values.

    ``det``: the deterministic mode (entries rounded to the block's grid step
    before they are summed); ``shuffle``: a numpy Generator that permutes the
    order in which each block's entries are added (the atomics' order)."""
    nb = topo.nb
    blk = arrs["blk"].reshape(-1, 8)
    recs = arrs["recs"].reshape(-1, 4)
    runs = arrs["runs"].reshape(-1, 4)
    vals = np.zeros(topo.nnz)
    covered = np.zeros(topo.nnz, dtype=np.int64)
    for b in range(blk.shape[0]):
        hoff, nh, nown, roff, nrec, runoff, nrun, nacc = (int(v) for v in blk[b])
        assert 0 < nh <= 1024 and 0 < nown <= nh
        halo = arrs["halo"][hoff:hoff + nh]
        meta = arrs["meta"][hoff:hoff + nh].astype(np.int64)
        assert (halo[:nown] >= 0).all()
        assert (meta[nown:] == 0).all()
        acc = np.zeros(nacc)
        contrib = []                                     # (accumulator index, value) in issue order
        for r in range(nrec):
            loc, rank = decode(recs[roff + r], nb)
            assert max(loc) < nh
            e = int(arrs["rec_elem"][roff + r])
            assert sorted(halo[loc]) == sorted(topo.conn[e])
            order = vertex_order(recs[roff + r], nb)
            assert sorted(order) == list(range(nb))
            assert list(halo[loc]) == [int(topo.conn[e][o]) for o in order]
            K = element_matrix(e, topo.points[halo[loc]])
            for a in range(nb):
                if loc[a] >= nown:
                    continue
                m = int(meta[loc[a]])
                base, dr, ln = m & 0xFFFF, (m >> 16) & 0xFF, m >> 24
                assert dr < ln and base + ln <= nacc
                contrib.append((base + dr, K[a, a], loc[a]))
                j = 0
                for bb in range(nb):
                    if bb == a:
                        continue
                    k = rank(a, j)
                    assert k < ln and k != dr
                    contrib.append((base + k, K[a, bb], loc[a]))
                    j += 1
        if shuffle is not None:
            contrib = [contrib[i] for i in shuffle.permutation(len(contrib))]
        if det:
            rows = {}
            for _, v, rw in contrib:
                rows.setdefault(rw, []).append(v)
            step = {rw: det_grid(vs) for rw, vs in rows.items()}
        for i, v, rw in contrib:
            acc[i] += np.rint(v / step[rw]) * step[rw] if det else v
        for q in range(nrun):
            vs, as_, ln, _ = (int(v) for v in runs[runoff + q])
            assert as_ + ln <= nacc
            vals[vs:vs + ln] = acc[as_:as_ + ln]
            covered[vs:vs + ln] += 1
    assert (covered == 1).all(), "every CSR value must be written by exactly one run"
    return vals


def decode_edge(words, nb):
    """Edge record (2 or 3 words): record-order slots (decoded from the LDS
    node words w(s) = s + (s & ~15)) and ``rank(k)`` of pair k in the order
    (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)."""
    w = [int(x) for x in words]
    W = len(w)
    SB = 10 if (nb == 4 and W == 2) else 11
    s = w[0] | (w[1] << 32)
    nw = [(s >> (SB * a)) & ((1 << SB) - 1) for a in range(nb)]
    assert all(x & 16 == 0 for x in nw), "node words have bit 4 clear"
    loc = [16 * (x >> 5) + (x & 15) for x in nw]

    def rank(k):
        if nb == 4 and W == 2:
            word, bit, R = 1, 8 + 4 * k, 4
        elif k < 3:
            word, bit, R = 1, 12 + 6 * k, 6
        else:
            word, bit, R = 2, 6 * (k - 3), 6
        return (w[word] >> bit) & ((1 << R) - 1)
    return loc, rank


def emulate_edge_matrix(topo, arrs, earrs, words, element_matrix, rowsum=True):
    """The edge-record kernel in float64: each pair (i, j), i < j, with an
    owned first vertex adds K_ij once, to the row of i; the mirror list then
    fills the transposed entries; ``rowsum`` forms the diagonal as minus the
    row's off-diagonal sum (Laplace-type forms), otherwise the owned diagonal
    entries are added directly (mass).  The owned count of each record must
    match the block's group ends; rows are stored from the block's row table
    (16 lanes per row in the kernel), which must cover each value once."""
    nb = topo.nb
    blk = arrs["blk"].reshape(-1, 8)
    erecs = earrs["erecs"].reshape(-1, words)
    eblk = earrs["eblk"].reshape(-1, 4)
    erows = earrs["erows"]
    mir = earrs["emir"]
    rp, ci = topo.row_ptr, topo.col_idx
    vals = np.zeros(topo.nnz)
    covered = np.zeros(topo.nnz, dtype=np.int64)
    pairs = [(i, j) for i in range(nb) for j in range(i + 1, nb)]
    moff = 0
    for b in range(blk.shape[0]):
        hoff, nh, nown, roff, nrec, runoff, nrun, nacc = (int(v) for v in blk[b])
        mo, mne2, e01, rowoff = (int(v) for v in eblk[b])
        mn, e2 = mne2 & 0xFFFF, mne2 >> 16
        assert mo == moff
        moff += mn
        ends = [e01 & 0xFFFF, e01 >> 16, e2]
        halo = arrs["halo"][hoff:hoff + nh]
        meta = arrs["meta"][hoff:hoff + nh].astype(np.int64)
        acc = np.zeros(nacc)
        hit = np.zeros(nacc, dtype=bool)
        for r in range(nrec):
            loc, rank = decode_edge(erecs[roff + r], nb)
            assert max(loc) < nh
            e = int(arrs["rec_elem"][roff + r])
            assert sorted(halo[loc]) == sorted(topo.conn[e])
            nk = sum(1 for s in loc if s < nown)
            assert nk == nb - sum(1 for x in ends if r >= x)
            assert all(s < nown for s in loc[:nk]) and loc[:nk] == sorted(loc[:nk])
            K = element_matrix(e, topo.points[halo[loc]])
            if not rowsum:
                for a in range(nk):
                    m = int(meta[loc[a]])
                    acc[(m & 0xFFFF) + ((m >> 16) & 0xFF)] += K[a, a]
            for k, (i, j) in enumerate(pairs):
                if i >= nk:
                    continue
                m = int(meta[loc[i]])
                base, dr, ln = m & 0xFFFF, (m >> 16) & 0xFF, m >> 24
                rk = rank(k)
                assert rk < ln and rk != dr
                assert ci[rp[halo[loc[i]]] + rk] == halo[loc[j]]
                acc[base + rk] += K[i, j]
                hit[base + rk] = True
        for v in mir[mo:mo + mn]:
            src, dst = int(v) & 0xFFFF, int(v) >> 16
            assert hit[src] and not hit[dst] and dst < nacc
            acc[dst] = acc[src]
        if rowsum:
            for a in range(nown):
                m = int(meta[a])
                base, dr, ln = m & 0xFFFF, (m >> 16) & 0xFF, m >> 24
                acc[base + dr] = -(acc[base:base + ln].sum() - acc[base + dr])
        for i in range(nown):
            vs, ol = int(erows[2 * (rowoff + i)]), int(erows[2 * (rowoff + i) + 1])
            off, ln = (ol & 0xFFFF) // 8, ol >> 16
            assert (ol & 0xFFFF) % 8 == 0 and off + ln <= nacc
            vals[vs:vs + ln] = acc[off:off + ln]
            covered[vs:vs + ln] += 1
    assert (covered == 1).all()
    return vals
