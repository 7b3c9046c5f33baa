"""
Seeded synthetic Hamiltonians and configuration sets for tests and benchmarks.

Term conventions follow the reference's model builders so the tables have the same shape as real
inputs:
  - molecular: spatial integrals expanded to spin orbitals with even = alpha, odd = beta and
    H = E0 + sum h_pq a+_p a_q + 1/2 sum (ij|kl) a+_i a+_k a_l a_j, normal ordered (creation and
    annihilation indices each descending), as qmb/fcidump.py:102-114 builds it through openfermion;
  - Hubbard: on-site U n_up n_down and -t hopping both ways on (i + j*m)*2 + spin
    (qmb/hubbard.py:62-86);
  - free fermion: +1 hopping (qmb/free_fermion.py:49-66);
  - Ising: Pauli terms as hard-core-boson ladder operators (qmb/ising.py:96-194), kind "bose2".
Configs are packed LSB first: site s is bit s % 8 of byte s // 8 (qmb/bitspack.py:8-56).
"""

from __future__ import annotations

import collections
import math
import typing

import numpy as np
import torch

TermDict = dict[tuple[tuple[int, int], ...], complex]
TermTuple = tuple[torch.Tensor, torch.Tensor, torch.Tensor]


# ----------------------------------------------------------------------------------- packing ---
def pack_occupation(occ: np.ndarray) -> np.ndarray:
    """bool/uint8 [N, n_sites] -> uint8 [N, ceil(n_sites / 8)], site s -> bit s % 8 of byte s // 8."""
    return np.packbits(np.asarray(occ, dtype=np.uint8), axis=-1, bitorder="little")


def unpack_occupation(configs: np.ndarray, n_sites: int) -> np.ndarray:
    """Inverse of pack_occupation: uint8 [N, nq] -> uint8 [N, n_sites]."""
    return np.unpackbits(np.asarray(configs, dtype=np.uint8), axis=-1, bitorder="little")[..., :n_sites]


def words_to_configs(words: np.ndarray, n_qubytes: int) -> np.ndarray:
    """uint64 [N, W] little-endian words -> uint8 [N, n_qubytes]."""
    w = np.ascontiguousarray(words.astype("<u8"))
    return w.view(np.uint8).reshape(len(w), -1)[:, :n_qubytes].copy()


def sort_unique_configs(configs: np.ndarray) -> np.ndarray:
    """Rows sorted by bytes (byte 0 most significant) with duplicates removed."""
    configs = np.ascontiguousarray(configs, dtype=np.uint8)
    if len(configs) == 0:
        return configs
    n, nq = configs.shape
    w = (nq + 7) // 8
    padded = np.zeros((n, 8 * w), dtype=np.uint8)
    padded[:, :nq] = configs
    keys = padded.view(">u8").reshape(n, w).astype(np.uint64)  # big-endian: byte 0 most significant
    if w == 1:
        u = np.unique(keys[:, 0]).astype(">u8")
        return u.view(np.uint8).reshape(-1, 8)[:, :nq].copy()
    order = np.lexsort(keys.T[::-1])  # word 0 is the primary key
    sk = keys[order]
    keep = np.ones(n, dtype=bool)
    keep[1:] = (sk[1:] != sk[:-1]).any(axis=1)
    return configs[order[keep]].copy()


def random_psi(n: int, seed: int) -> np.ndarray:
    """complex128 N(0,1) amplitudes, normalised."""
    rng = np.random.default_rng(seed)
    z = rng.normal(size=(n, 2))
    psi = z[:, 0] + 1j * z[:, 1]
    norm = np.linalg.norm(psi)
    return psi / norm if norm > 0 else psi


# --------------------------------------------------------------------------------- term tables -
def terms_from_dict(hamiltonian: TermDict) -> TermTuple:
    """Dict form -> (site int16[T,4], kind uint8[T,4], coef float64[T,2]) in insertion order.

    Same layout as the reference's prepare<4> (qmb/_hamiltonian.cpp:29-64); unused operator columns
    get kind 2 and site 0 (the reference leaves those sites uninitialised).
    """
    n = len(hamiltonian)
    site = np.zeros((n, 4), dtype=np.int16)
    kind = np.full((n, 4), 2, dtype=np.uint8)
    coef = np.zeros((n, 2), dtype=np.float64)
    for index, (key, value) in enumerate(hamiltonian.items()):
        if len(key) > 4:
            raise ValueError(f"a term has {len(key)} operators; at most 4 are supported")
        for op, (s, k) in enumerate(key):
            site[index, op] = s
            kind[index, op] = k
        value = complex(value)
        coef[index, 0] = value.real
        coef[index, 1] = value.imag
    return torch.from_numpy(site), torch.from_numpy(kind), torch.from_numpy(coef)


def molecular_integrals(n_orbit: int, seed: int) -> tuple[np.ndarray, np.ndarray, float]:
    """h_pq symmetric N(0,1) exp(-|p-q|/4); (ij|kl) 8-fold symmetric N(0,0.1) exp(-(|i-j|+|k-l|)/4)."""
    rng = np.random.default_rng(seed)
    n = n_orbit
    idx = np.arange(n)
    dist = np.abs(idx[:, None] - idx[None, :])
    a = rng.normal(0.0, 1.0, size=(n, n))
    h = (np.triu(a) + np.triu(a, 1).T) * np.exp(-dist / 4.0)
    n_pair = n * (n + 1) // 2
    base = rng.normal(0.0, 0.1, size=(n_pair, n_pair))
    hi = np.maximum(idx[:, None], idx[None, :])
    lo = np.minimum(idx[:, None], idx[None, :])
    pair = hi * (hi + 1) // 2 + lo  # [n, n] canonical pair index of (i, j)
    p_ij = pair[:, :, None, None]
    p_kl = pair[None, None, :, :]
    g = base[np.maximum(p_ij, p_kl), np.minimum(p_ij, p_kl)]
    g = g * np.exp(-(dist[:, :, None, None] + dist[None, None, :, :]) / 4.0)
    e0 = float(rng.normal())
    return h, g, e0


def molecular_terms(n_orbit: int, seed: int, threshold: float = 0.0) -> TermTuple:
    """Normal-ordered spin-orbital terms of a synthetic molecular Hamiltonian (constant, a+a, a+a+aa).

    Two-electron integrals with |(ij|kl)| < threshold are dropped before the expansion.
    """
    h, g, e0 = molecular_integrals(n_orbit, seed)
    if threshold > 0:
        g = np.where(np.abs(g) < threshold, 0.0, g)
    n_so = 2 * n_orbit
    spin = np.arange(n_so) % 2
    spat = np.arange(n_so) // 2
    sites: list[np.ndarray] = []
    kinds: list[np.ndarray] = []
    coefs: list[np.ndarray] = []
    # constant
    sites.append(np.zeros((1, 4), dtype=np.int16))
    kinds.append(np.full((1, 4), 2, dtype=np.uint8))
    coefs.append(np.array([[e0, 0.0]]))
    # one body: a+_P a_Q, same spin
    P, Q = np.meshgrid(np.arange(n_so), np.arange(n_so), indexing="ij")
    m = spin[P] == spin[Q]
    P1, Q1 = P[m], Q[m]
    v1 = h[spat[P1], spat[Q1]]
    s1 = np.zeros((len(P1), 4), dtype=np.int16)
    s1[:, 0], s1[:, 1] = P1, Q1
    k1 = np.full((len(P1), 4), 2, dtype=np.uint8)
    k1[:, 0], k1[:, 1] = 1, 0
    sites.append(s1)
    kinds.append(k1)
    coefs.append(np.stack([v1, np.zeros_like(v1)], axis=1))
    # two body: tb[P,Q,R,S] multiplies a+_P a+_Q a_R a_S; tb = (1/2) (p s | q r) with spin(P)=spin(S),
    # spin(Q)=spin(R).  Normal ordering with P > Q and R > S antisymmetrises it.
    ar = np.arange(n_so)

    def tb(p: np.ndarray, q: np.ndarray, r: np.ndarray, s: np.ndarray) -> np.ndarray:
        ok = (spin[p] == spin[s]) & (spin[q] == spin[r])
        return np.where(ok, 0.5 * g[spat[p], spat[s], spat[q], spat[r]], 0.0)

    pq = np.array([(p, q) for p in ar for q in ar if p > q], dtype=np.int64)
    for start in range(0, len(pq), 256):
        block = pq[start:start + 256]
        Pb = np.repeat(block[:, 0], len(pq))
        Qb = np.repeat(block[:, 1], len(pq))
        Rb = np.tile(pq[:, 0], len(block))
        Sb = np.tile(pq[:, 1], len(block))
        V = tb(Pb, Qb, Rb, Sb) - tb(Qb, Pb, Rb, Sb) - tb(Pb, Qb, Sb, Rb) + tb(Qb, Pb, Sb, Rb)
        keep = V != 0.0
        n_keep = int(keep.sum())
        s2 = np.stack([Pb[keep], Qb[keep], Rb[keep], Sb[keep]], axis=1).astype(np.int16)
        k2 = np.tile(np.array([[1, 1, 0, 0]], dtype=np.uint8), (n_keep, 1))
        sites.append(s2)
        kinds.append(k2)
        coefs.append(np.stack([V[keep], np.zeros(n_keep)], axis=1))
    return (
        torch.from_numpy(np.concatenate(sites)),
        torch.from_numpy(np.concatenate(kinds)),
        torch.from_numpy(np.concatenate(coefs)),
    )


def molecular_term_count(n_orbit: int) -> int:
    """Closed-form T = 2 C(n,2)^2 + n^4 + 2 n^2 + 1 for the generator above (no threshold)."""
    n = n_orbit
    return 2 * math.comb(n, 2) ** 2 + n**4 + 2 * n**2 + 1


def hubbard_dict(m: int, n: int, t: float = 1.0, u: float = 4.0) -> TermDict:
    """qmb/hubbard.py:62-86."""

    def index(i: int, j: int, o: int) -> int:
        return (i + j * m) * 2 + o

    d: TermDict = {}
    for i in range(m):
        for j in range(n):
            d[(index(i, j, 0), 1), (index(i, j, 0), 0), (index(i, j, 1), 1), (index(i, j, 1), 0)] = u
            if i != 0:
                d[(index(i, j, 0), 1), (index(i - 1, j, 0), 0)] = -t
                d[(index(i - 1, j, 0), 1), (index(i, j, 0), 0)] = -t
                d[(index(i, j, 1), 1), (index(i - 1, j, 1), 0)] = -t
                d[(index(i - 1, j, 1), 1), (index(i, j, 1), 0)] = -t
            if j != 0:
                d[(index(i, j, 0), 1), (index(i, j - 1, 0), 0)] = -t
                d[(index(i, j - 1, 0), 1), (index(i, j, 0), 0)] = -t
                d[(index(i, j, 1), 1), (index(i, j - 1, 1), 0)] = -t
                d[(index(i, j - 1, 1), 1), (index(i, j, 1), 0)] = -t
    return d


def free_fermion_dict(m: int, n: int) -> TermDict:
    """qmb/free_fermion.py:49-66."""

    def index(i: int, j: int) -> int:
        return i + j * m

    d: TermDict = {}
    for i in range(m):
        for j in range(n):
            if i != 0:
                d[(index(i, j), 1), (index(i - 1, j), 0)] = 1
                d[(index(i - 1, j), 1), (index(i, j), 0)] = 1
            if j != 0:
                d[(index(i, j), 1), (index(i, j - 1), 0)] = 1
                d[(index(i, j - 1), 1), (index(i, j), 0)] = 1
    return d


def ising_dict(m: int, n: int, **coupling: float) -> TermDict:
    """Pauli couplings as hard-core-boson terms, qmb/ising.py:96-194.

    Keyword couplings: x, y, z (on-site), xh/yh/zh (horizontal), xv/yv/zv (vertical),
    xd/yd/zd (diagonal), xa/ya/za (anti-diagonal).
    """

    def index(i: int, j: int) -> int:
        return i + j * m

    def px(i: int, j: int) -> tuple[tuple[tuple[tuple[int, int], ...], complex], ...]:
        return ((((index(i, j), 1),), 1), (((index(i, j), 0),), 1))

    def py(i: int, j: int) -> tuple[tuple[tuple[tuple[int, int], ...], complex], ...]:
        return ((((index(i, j), 1),), -1j), (((index(i, j), 0),), 1j))

    def pz(i: int, j: int) -> tuple[tuple[tuple[tuple[int, int], ...], complex], ...]:
        return ((((index(i, j), 1), (index(i, j), 0)), 1), (((index(i, j), 0), (index(i, j), 1)), -1))

    paulis = {"x": px, "y": py, "z": pz}
    d: collections.defaultdict[tuple[tuple[int, int], ...], complex] = collections.defaultdict(lambda: 0)
    c = {k: float(coupling.get(k, 0.0)) for k in
         ["x", "y", "z", "xh", "yh", "zh", "xv", "yv", "zv", "xd", "yd", "zd", "xa", "ya", "za"]}

    def bond(a: tuple[int, int], b: tuple[int, int], name: str) -> None:
        for axis in "xyz":
            v = c[axis + name]
            if v != 0:
                for k1, v1 in paulis[axis](*a):
                    for k2, v2 in paulis[axis](*b):
                        d[(*k1, *k2)] += v1 * v2 * v

    for i in range(m):
        for j in range(n):
            for axis in "xyz":
                if c[axis] != 0:
                    for k, v in paulis[axis](i, j):
                        d[k] += v * c[axis]
            if i != 0:
                bond((i, j), (i - 1, j), "h")
            if j != 0:
                bond((i, j), (i, j - 1), "v")
            if i != 0 and j != 0:
                bond((i, j), (i - 1, j - 1), "d")
            if i != 0 and j != n - 1:
                bond((i, j), (i - 1, j + 1), "a")
    return dict(d)


def random_terms(n_sites: int, n_terms: int, seed: int, max_ops: int = 4) -> TermTuple:
    """Arbitrary operator strings (any order, repeated sites, identity pads, complex coefficients).

    Not physical; used to exercise the full application semantics of qmb/_hamiltonian_cpu.cpp:48-78.
    """
    rng = np.random.default_rng(seed)
    site = rng.integers(0, n_sites, size=(n_terms, 4)).astype(np.int16)
    kind = rng.integers(0, 2, size=(n_terms, 4)).astype(np.uint8)
    n_ops = rng.integers(0, max_ops + 1, size=n_terms)
    for t in range(n_terms):
        kind[t, n_ops[t]:] = 2
        site[t, n_ops[t]:] = 0
    coef = rng.normal(size=(n_terms, 2))
    return torch.from_numpy(site), torch.from_numpy(kind), torch.from_numpy(coef)


# ------------------------------------------------------------------------------- config sets ---
def _sector_states(n_orbit: int, n_alpha: int, n_beta: int) -> np.ndarray:
    alpha = [c for c in _combinations_bits(n_orbit, n_alpha)]
    beta = [c for c in _combinations_bits(n_orbit, n_beta)]
    a = np.array(alpha, dtype=np.uint64)
    b = np.array(beta, dtype=np.uint64)
    return a, b


def _combinations_bits(n: int, k: int) -> typing.Iterator[int]:
    import itertools

    for comb in itertools.combinations(range(n), k):
        v = 0
        for c in comb:
            v |= 1 << c
        yield v


def _spread(spatial: np.ndarray, n_orbit: int, spin: int) -> np.ndarray:
    """Spatial occupation bit masks (Python ints as uint64 / object) -> spin-orbital bool rows."""
    bits = ((spatial[:, None] >> np.arange(n_orbit, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.uint8)
    occ = np.zeros((len(spatial), 2 * n_orbit), dtype=np.uint8)
    occ[:, spin::2] = bits
    return occ


def sector_configs(n_orbit: int, n_alpha: int, n_beta: int, count: int, seed: int) -> np.ndarray:
    """`count` distinct configs drawn uniformly from the fixed (N_alpha, N_beta) sector, sorted by bytes."""
    rng = np.random.default_rng(seed)
    n_so = 2 * n_orbit
    sector = math.comb(n_orbit, n_alpha) * math.comb(n_orbit, n_beta)
    if count > sector:
        raise ValueError(f"sector has only {sector} states")
    if sector <= 4_000_000 and n_orbit <= 63:
        a, b = _sector_states(n_orbit, n_alpha, n_beta)
        pick = rng.choice(sector, size=count, replace=False)
        ia, ib = pick // len(b), pick % len(b)
        occ = _spread(a[ia], n_orbit, 0) | _spread(b[ib], n_orbit, 1)
        return sort_unique_configs(pack_occupation(occ))
    out = np.zeros((0, (n_so + 7) // 8), dtype=np.uint8)
    while len(out) < count:
        m = int((count - len(out)) * 1.2) + 16
        occ = np.zeros((m, n_so), dtype=np.uint8)
        ra = np.argsort(rng.random((m, n_orbit)), axis=1)[:, :n_alpha]
        rb = np.argsort(rng.random((m, n_orbit)), axis=1)[:, :n_beta]
        np.put_along_axis(occ, 2 * ra, 1, axis=1)
        np.put_along_axis(occ, 2 * rb + 1, 1, axis=1)
        out = sort_unique_configs(np.concatenate([out, pack_occupation(occ)]))
    keep = np.sort(rng.choice(len(out), size=count, replace=False))
    return out[keep]


def excitation_configs(n_orbit: int, n_alpha: int, n_beta: int, count: int, seed: int) -> np.ndarray:
    """HF + all singles and doubles + seeded random triples/quadruples, `count` distinct, sorted.

    Occupied orbitals of the reference determinant are the lowest n_alpha / n_beta spatial orbitals.
    """
    rng = np.random.default_rng(seed)
    n_so = 2 * n_orbit
    hf = np.zeros(n_so, dtype=np.uint8)
    hf[0:2 * n_alpha:2] = 1
    hf[1:2 * n_beta:2] = 1
    occ_a = np.arange(n_alpha)
    vir_a = np.arange(n_alpha, n_orbit)
    occ_b = np.arange(n_beta)
    vir_b = np.arange(n_beta, n_orbit)
    rows = [hf[None, :]]

    def excite(holes: np.ndarray, parts: np.ndarray) -> np.ndarray:
        """hf with, per row, the sites holes[r] emptied and parts[r] filled."""
        c = np.tile(hf, (len(holes), 1))
        r = np.arange(len(holes))[:, None]
        c[r, holes] = 0
        c[r, parts] = 1
        return c

    def pairs(x: np.ndarray) -> np.ndarray:
        i, j = np.triu_indices(len(x), 1)
        return np.stack([x[i], x[j]], axis=1)

    def product(a: np.ndarray, b: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Every (row of a, row of b) combination, a-major."""
        return np.repeat(a, len(b), axis=0), np.tile(b, (len(a), 1))

    for spin, occ, vir in ((0, occ_a, vir_a), (1, occ_b, vir_b)):
        h1, p1 = product((2 * occ + spin)[:, None], (2 * vir + spin)[:, None])
        rows.append(excite(h1, p1))
        h2, p2 = product(2 * pairs(occ) + spin, 2 * pairs(vir) + spin)
        rows.append(excite(h2, p2))
    ia, ja = product((2 * occ_a)[:, None], (2 * vir_a)[:, None])
    ib, jb = product((2 * occ_b + 1)[:, None], (2 * vir_b + 1)[:, None])
    holes_ab, parts_ab = product(np.concatenate([ia, ja], axis=1), np.concatenate([ib, jb], axis=1))
    rows.append(excite(np.stack([holes_ab[:, 0], parts_ab[:, 0]], axis=1), np.stack([holes_ab[:, 1], parts_ab[:, 1]], axis=1)))
    base = sort_unique_configs(pack_occupation(np.concatenate(rows)))
    if len(base) >= count:
        keep = np.sort(rng.choice(len(base), size=count, replace=False))
        return base[keep]
    extra = np.zeros((0, base.shape[1]), dtype=np.uint8)
    need = count - len(base)
    stalled = 0
    while len(extra) < need:
        before = len(extra)
        m = int((need - len(extra)) * 1.3) + 64
        level = rng.integers(3, 5, size=m)
        cap_a = min(n_alpha, n_orbit - n_alpha)
        cap_b = min(n_beta, n_orbit - n_beta)
        lo = np.maximum(0, level - cap_b)
        hi = np.minimum(level, cap_a)
        ea = lo + np.floor(rng.random(m) * (hi - lo + 1)).astype(np.int64)
        eb = level - ea
        occ = np.tile(hf, (m, 1))
        ra_h = np.argsort(np.argsort(rng.random((m, n_alpha)), axis=1), axis=1) < ea[:, None]
        ra_p = np.argsort(np.argsort(rng.random((m, len(vir_a))), axis=1), axis=1) < ea[:, None]
        rb_h = np.argsort(np.argsort(rng.random((m, n_beta)), axis=1), axis=1) < eb[:, None]
        rb_p = np.argsort(np.argsort(rng.random((m, len(vir_b))), axis=1), axis=1) < eb[:, None]
        occ[:, 2 * occ_a] = np.where(ra_h, 0, occ[:, 2 * occ_a])
        occ[:, 2 * vir_a] = np.where(ra_p, 1, occ[:, 2 * vir_a])
        occ[:, 2 * occ_b + 1] = np.where(rb_h, 0, occ[:, 2 * occ_b + 1])
        occ[:, 2 * vir_b + 1] = np.where(rb_p, 1, occ[:, 2 * vir_b + 1])
        extra = sort_unique_configs(np.concatenate([extra, pack_occupation(occ)]))
        stalled = stalled + 1 if len(extra) == before else 0
        if stalled >= 8:
            raise ValueError(f"only {len(base) + len(extra)} configs up to quadruple excitations exist; asked for {count}")
    keep = np.sort(rng.choice(len(extra), size=need, replace=False))
    return sort_unique_configs(np.concatenate([base, extra[keep]]))


def hubbard_chain_configs(n_site: int, count: int, seed: int, n_seed_configs: int = 65536) -> np.ndarray:
    """Half-filled configs of an open 1-D Hubbard chain: random seeds plus all their single hops.

    Sites follow qmb/hubbard.py's (i*2 + spin) numbering.  Returns `count` distinct configs sorted
    by bytes (or all of them if fewer exist).
    """
    rng = np.random.default_rng(seed)
    n_so = 2 * n_site
    n_up = n_site // 2
    n_dn = n_site - n_up
    occ = np.zeros((n_seed_configs, n_so), dtype=np.uint8)
    ra = np.argsort(rng.random((n_seed_configs, n_site)), axis=1)[:, :n_up]
    rb = np.argsort(rng.random((n_seed_configs, n_site)), axis=1)[:, :n_dn]
    np.put_along_axis(occ, 2 * ra, 1, axis=1)
    np.put_along_axis(occ, 2 * rb + 1, 1, axis=1)
    seeds = pack_occupation(occ)
    parts = [seeds]
    for spin in (0, 1):
        for i in range(n_site - 1):
            a, b = 2 * i + spin, 2 * (i + 1) + spin
            for src, dst in ((a, b), (b, a)):
                ok = (occ[:, src] == 1) & (occ[:, dst] == 0)
                moved = occ[ok].copy()
                moved[:, src] = 0
                moved[:, dst] = 1
                parts.append(pack_occupation(moved))
    allc = sort_unique_configs(np.concatenate(parts))
    if len(allc) <= count:
        return allc
    keep = np.sort(rng.choice(len(allc), size=count, replace=False))
    return allc[keep]


def batch_instances(seed0: int = 4000, count: int = 64, threads: int = 8,
                    progress: typing.Callable[[int, int], None] | None = None) -> list[dict[str, object]]:
    """Seeded molecular instances of uneven sizes: spin orbitals uniform over the even numbers in
    [24, 80], ~30% filling split alpha/beta, |(pq|rs)| < 1e-10 dropped, config-set sizes
    log-uniform in [1e3, 5e5] (capped by the sector) built as HF + singles + doubles +
    triples/quadruples.  Instance k uses seeds seed0 + k (terms), + 1000 + k (configs), + 2000 + k
    (psi) and + 3000 + k (a second vector); its terms are ordered by |coef| descending (stable).
    The shapes are drawn first from one generator, then the instances are built on `threads`
    host threads (the result does not depend on the thread count)."""
    import concurrent.futures  # pylint: disable=import-outside-toplevel

    rng = np.random.default_rng(seed0)
    specs = []
    for k in range(count):
        n_so = 2 * int(rng.integers(12, 41))
        n_orb = n_so // 2
        n_e = int(round(0.3 * n_so))
        n_a, n_b = (n_e + 1) // 2, n_e // 2
        n = int(round(math.exp(rng.uniform(math.log(1e3), math.log(5e5)))))
        specs.append((k, n_orb, n_a, n_b, min(n, math.comb(n_orb, n_a) * math.comb(n_orb, n_b))))
    done = [0]

    def make(spec: tuple[int, int, int, int, int]) -> dict[str, object]:
        k, n_orb, n_a, n_b, n = spec
        site, kind, coef = molecular_terms(n_orb, seed0 + k, threshold=1e-10)
        order = np.argsort(-np.linalg.norm(coef.numpy(), axis=1), kind="stable")
        try:
            configs = excitation_configs(n_orb, n_a, n_b, n, seed0 + 1000 + k)
        except ValueError:
            configs = sector_configs(n_orb, n_a, n_b, n, seed0 + 1000 + k)
        item = {"site": site[order], "kind": kind[order], "coef": coef[order], "configs": configs, "n_so": 2 * n_orb,
                "electrons": (n_a, n_b), "psi": random_psi(len(configs), seed0 + 2000 + k),
                "grad": random_psi(len(configs), seed0 + 3000 + k)}
        done[0] += 1
        if progress is not None:
            progress(done[0], count)
        return item

    with concurrent.futures.ThreadPoolExecutor(max(1, threads)) as pool:
        return list(pool.map(make, specs))
