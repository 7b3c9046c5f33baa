"""
The hop path (qmb_strings_build / qmb_apply_hops / qmb_apply_diagonal_terms) against the oracle and
the enumeration path, for tables whose off-diagonal terms are single excitations between sites of
one parity (qmb/hubbard.py's spinful hops) and diagonal terms.

  - Hubbard chains and lattices (one- and two-word configs): every row of H psi and H^H g equals the
    enumeration path (which the oracle pins in test_gpu_apply), and sampled rows equal the oracle
    to 1e-12 of each row's sum of |contributions|.
  - A random spinful table with conditioned-site hops (c+_p n_k c_q), cross-parity hops (left to the
    enumeration as residual terms), complex coefficients and diagonal terms: every row against the
    enumeration path.
  - The strings index: every row once, groups are runs of equal keys holding equal other-parity
    strings, and its pair-test count equals the sum of group size^2 over groups of two or more.
  - A set whose rows all share one string falls back to the enumeration (the pair tests would be
    n^2).
  - The 64-site Hubbard chain at 4M configs: the hop path equals the enumeration on every row.
  - The C ABI: qmb_apply_within / qmb_apply_within_adjoint on one config buffer with a workspace of
    qmb_apply_within_hops_workspace_size take the hop path and equal the enumeration they run with a
    workspace of qmb_apply_within_workspace_size.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from qmb_hip import prepared as pp
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu


def _prepare(site: np.ndarray, kind: np.ndarray, coef: np.ndarray, cfg: np.ndarray, nq: int, dev: torch.device) -> dict[str, object]:
    st, kt, ct = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (site, kind, coef))
    cfg_t = torch.from_numpy(np.ascontiguousarray(cfg)).to(dev)
    terms = pp.PreparedTerms(st, kt, ct, nq, 1)
    targets = pp.PreparedTargets(cfg_t, nq, distinct=True)
    return {"terms": terms, "targets": targets, "cfg_t": cfg_t, "dev": dev, "np": (site, kind, coef), "cfg": cfg}


def _apply(w: dict[str, object], vec: np.ndarray, mode: int, hops: bool | None) -> np.ndarray:
    v = torch.view_as_real(torch.from_numpy(vec).to(w["dev"])).contiguous()
    out = pp.apply(w["terms"], w["targets"], w["cfg_t"], v, mode, hops=hops)
    return torch.view_as_complex(out).cpu().numpy()


def _check_against_enumeration(w: dict[str, object], seed: int) -> None:
    n = len(w["cfg"])
    assert pp.can_use_hops(w["terms"], w["targets"], w["cfg_t"])
    for mode, s in ((pp.MODE_FORWARD, seed), (pp.MODE_ADJOINT, seed + 1)):
        vec = sy.random_psi(n, s)
        hop = _apply(w, vec, mode, None)
        enum = _apply(w, vec, mode, False)
        scale = float(np.abs(enum).max())
        assert scale > 0.0
        assert float(np.abs(hop - enum).max()) <= 1e-12 * scale, mode


def _check_oracle_rows(w: dict[str, object], seed: int, rows: int = 6) -> None:
    cfg = w["cfg"]
    vec = sy.random_psi(len(cfg), seed)
    fwd = _apply(w, vec, pp.MODE_FORWARD, None)
    for j in np.random.default_rng(seed).choice(len(cfg), size=rows, replace=False):
        # the tables are Hermitian: row j of H psi is <column j of H, psi>
        column, asum, _ = orc.apply_within(1, cfg[j:j + 1], np.ones(1, dtype=np.complex128), cfg, *w["np"])
        want = np.vdot(column, vec)
        assert abs(fwd[j] - want) <= 1e-12 * float(asum @ np.abs(vec)) + 1e-300, (int(j), fwd[j], want)


@pytest.mark.parametrize("m,n_lat,count", [(8, 1, 5000), (4, 4, 6000), (40, 1, 20000), (8, 8, 30000)],
                         ids=["chain-16so", "lattice-4x4", "chain-80so", "lattice-8x8"])
def test_hubbard_hops_match_enumeration_and_oracle(cuda_device: torch.device, m: int, n_lat: int, count: int) -> None:
    site, kind, coef = (x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(m, n_lat, 1.0, 4.0)))
    sites = m * n_lat
    nq = (2 * sites + 7) // 8
    cfg = _hubbard_configs(sites, count, 7000 + sites)
    w = _prepare(site, kind, coef, cfg, nq, cuda_device)
    assert w["terms"].hops and w["terms"].n_residual_terms == 0
    _check_against_enumeration(w, 7100 + sites)
    _check_oracle_rows(w, 7200 + sites)


def _hubbard_configs(sites: int, count: int, seed: int, n_up: int | None = None, n_dn: int | None = None) -> np.ndarray:
    """Configs of n_up + n_dn particles (default: half filling): random seeds plus all their
    same-spin moves between any two sites (so lattice hops meet their targets), distinct, sorted."""
    rng = np.random.default_rng(seed)
    n_up = sites // 2 if n_up is None else n_up
    n_dn = sites - sites // 2 if n_dn is None else n_dn
    n_seed = max(4, count // (2 * n_up * (sites - n_up)) + 1)
    occ = np.zeros((n_seed, 2 * sites), dtype=np.uint8)
    for r in range(n_seed):
        occ[r, 2 * rng.choice(sites, n_up, replace=False)] = 1
        occ[r, 2 * rng.choice(sites, n_dn, replace=False) + 1] = 1
    parts = [occ]
    for spin in (0, 1):
        for a in range(sites):
            for b in range(sites):
                src, dst = 2 * a + spin, 2 * b + spin
                ok = (occ[:, src] == 1) & (occ[:, dst] == 0)
                moved = occ[ok].copy()
                moved[:, src], moved[:, dst] = 0, 1
                parts.append(moved)
    allc = sy.sort_unique_configs(sy.pack_occupation(np.concatenate(parts)))
    if len(allc) > count:
        allc = allc[np.sort(rng.choice(len(allc), size=count, replace=False))]
    return allc


def test_hubbard_hops_mixed_particle_counts(cuda_device: torch.device) -> None:
    """Configs of three (up, down) particle counts: rows of one group then hold hop-parity strings of
    different particle counts, where the hop kernel checks that one differing site is occupied in
    each string (with equal counts the number of differing sites decides)."""
    m = 12
    site, kind, coef = (x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(m, 1, 1.0, 4.0)))
    nq = (2 * m + 7) // 8
    parts = [_hubbard_configs(m, 4000, 7900 + k, n_up=m // 2 + du, n_dn=m // 2 + dd)
             for k, (du, dd) in enumerate(((0, 0), (1, 0), (0, 1)))]
    cfg = sy.sort_unique_configs(np.concatenate(parts))
    occ = np.unpackbits(cfg, axis=1, bitorder="little")[:, :2 * m]
    counts = {(int(u), int(d)) for u, d in zip(occ[:, 0::2].sum(axis=1), occ[:, 1::2].sum(axis=1))}
    assert counts == {(6, 6), (7, 6), (6, 7)}
    w = _prepare(site, kind, coef, cfg, nq, cuda_device)
    assert w["terms"].hops and w["terms"].n_residual_terms == 0
    _check_against_enumeration(w, 7910)
    _check_oracle_rows(w, 7920)


def _spinful_random_terms(n_sites: int, seed: int, n_diag_pairs: int | None = None) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Hermitian spinful table: same-parity hops (some conditioned on another site's occupation),
    cross-parity hops, and diagonal n_p and n_p n_q terms (n_diag_pairs draws, n_sites by
    default), complex coefficients."""
    rng = np.random.default_rng(seed)
    d: dict[tuple[tuple[int, int], ...], complex] = {}
    n_so = 2 * n_sites

    def add(ops: tuple[tuple[int, int], ...], c: complex) -> None:
        d[ops] = d.get(ops, 0.0) + c
        adj = tuple((s, 1 - k) for s, k in reversed(ops))
        d[adj] = d.get(adj, 0.0) + np.conj(c)

    for _ in range(3 * n_sites):
        p, q = rng.choice(n_sites, 2, replace=False)
        spin = int(rng.integers(2))
        add(((2 * p + spin, 1), (2 * q + spin, 0)), complex(rng.normal(), rng.normal()))
    for _ in range(n_sites):
        p, q, k = rng.choice(n_so, 3, replace=False)
        q = 2 * (q // 2) + (p & 1)  # same parity as p
        if q == p or k in (p, q):
            continue
        add(((p, 1), (k, 1), (k, 0), (q, 0)), complex(rng.normal(), rng.normal()))
    for _ in range(2):
        p = 2 * int(rng.integers(n_sites))
        q = 2 * int(rng.integers(n_sites)) + 1
        add(((p, 1), (q, 0)), complex(rng.normal(), 0.0))
    for p in range(n_so):
        d[((p, 1), (p, 0))] = d.get(((p, 1), (p, 0)), 0.0) + rng.normal()
    for _ in range(n_sites if n_diag_pairs is None else n_diag_pairs):
        p, q = rng.choice(n_so, 2, replace=False)
        d[((p, 1), (p, 0), (q, 1), (q, 0))] = d.get(((p, 1), (p, 0), (q, 1), (q, 0)), 0.0) + rng.normal()
    s, k, c = sy.terms_from_dict(d)
    return s.numpy(), k.numpy(), c.numpy()


@pytest.mark.parametrize("n_sites,n_diag_pairs", [(6, None), (40, None), (40, 700)], ids=["one-word", "two-word", "many-diagonal"])
def test_spinful_table_with_conditioned_and_cross_hops(cuda_device: torch.device, n_sites: int, n_diag_pairs: int | None) -> None:
    # many-diagonal: over 256 diagonal terms, walked a tile at a time by the diagonal kernel
    site, kind, coef = _spinful_random_terms(n_sites, 7300 + n_sites, n_diag_pairs)
    nq = (2 * n_sites + 7) // 8
    cfg = _hubbard_configs(n_sites, 8000, 7400 + n_sites)
    w = _prepare(site, kind, coef, cfg, nq, cuda_device)
    assert w["terms"].hops and w["terms"].n_residual_terms > 0
    if n_diag_pairs is not None:
        assert w["terms"].counts()[1] > 256
    _check_against_enumeration(w, 7500 + n_sites)
    _check_oracle_rows(w, 7600 + n_sites, rows=4)


def test_diagonal_values_match_oracle_diagonal_term(cuda_device: torch.device) -> None:
    """qmb_diagonal_values on a set equals the oracle's diagonal_term of its rows (the same terms
    summed in original order, so bit-exact), for a table with n_p and n_p n_q terms over 256 of
    them (two LDS tiles), and qmb_apply_diagonal_values adds d * in (conj(d) for the adjoint)."""
    site, kind, coef = _spinful_random_terms(40, 7350, 700)
    cfg = _hubbard_configs(40, 3000, 7351)
    w = _prepare(site, kind, coef, cfg, 10, cuda_device)
    assert w["terms"].counts()[1] > 256
    d = torch.view_as_complex(w["targets"].diagonal_values(w["terms"])).cpu().numpy()
    want = orc.diagonal_term(1, cfg, site, kind, coef)
    assert np.array_equal(d, want)
    from qmb_hip import _lib  # pylint: disable=import-outside-toplevel

    lib = _lib.library()
    vec = sy.random_psi(len(cfg), 7352)
    v = torch.view_as_real(torch.from_numpy(vec).to(cuda_device)).contiguous()
    for mode, dd in ((0, d), (1, np.conj(d))):
        out = torch.zeros_like(v)
        _lib.check(lib.qmb_apply_diagonal_values(len(cfg), ctypes.c_void_p(w["targets"].diagonal_values(w["terms"]).data_ptr()),
                                                 ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(out.data_ptr()), mode, None),
                   "qmb_apply_diagonal_values")
        torch.cuda.synchronize()
        got = torch.view_as_complex(out).cpu().numpy()
        assert float(np.abs(got - dd * vec).max()) <= 1e-14 * float(np.abs(dd * vec).max())


def test_strings_index_structure(cuda_device: torch.device) -> None:
    cfg = _hubbard_configs(40, 20000, 7700)
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    targets = pp.PreparedTargets(cfg_t, 10, distinct=True)
    occ = sy.unpack_occupation(cfg, 80)
    for parity in (0, 1):
        desc = targets.strings(parity)
        buf = targets._strings[parity][0]  # pylint: disable=protected-access
        base = buf.data_ptr()
        n = len(cfg)

        def arr(ptr: int, count: int) -> np.ndarray:
            off = int(ptr) - base
            return (buf[off:off + 4 * count].view(torch.int32).cpu().numpy().astype(np.int64)) & 0xFFFFFFFF

        stats = targets.strings_stats(parity)
        keys, rows, group = arr(desc.keys, n), arr(desc.rows, n), arr(desc.group, n)
        starts = arr(desc.starts, stats["groups"] + 1)
        assert stats["entries"] == n and sorted(rows.tolist()) == list(range(n))
        assert (np.diff(keys) >= 0).all()
        sizes = np.diff(starts)
        assert starts[0] == 0 and starts[-1] == n and (sizes > 0).all()
        assert (group == np.repeat(np.arange(stats["groups"]), sizes)).all()
        # a group holds one key, and (unless two strings' keys collide: a 32-bit key over these
        # ~20,000 strings merges a pair about once in twenty sets) one other-parity string
        other = occ[rows][:, (1 - parity)::2]
        merged = 0
        for g in np.flatnonzero(sizes > 1):
            lo, hi = starts[g], starts[g + 1]
            assert (keys[lo:hi] == keys[lo]).all()
            merged += int(not (other[lo:hi] == other[lo]).all())
        assert merged <= 1
        # and no string is split across groups
        first = {}
        for g in range(stats["groups"]):
            for r in range(starts[g], starts[g + 1]):
                assert first.setdefault(other[r].tobytes(), g) == g
        assert stats["pair_tests"] == int((sizes[sizes > 1] ** 2).sum())
        assert stats["largest"] == int(sizes.max())
        # the compacted entries: the rows of groups of two or more, in index order, each with its
        # group as a range of that order and its config words
        multi = np.repeat(sizes > 1, sizes)
        m = int(multi.sum())
        assert stats["multi_entries"] == m > 0
        mrows = arr(desc.mrows, m)
        assert (mrows == rows[multi]).all()
        off = int(desc.mrange) - base
        mrange = (buf[off:off + 8 * m].view(torch.int32).cpu().numpy().astype(np.int64) & 0xFFFFFFFF).reshape(m, 2)
        msizes = sizes[sizes > 1]
        mstart = np.repeat(np.concatenate(([0], np.cumsum(msizes)[:-1])), msizes)
        end, merged_bit = mrange[:, 1] & 0x7FFFFFFF, mrange[:, 1] >> 31
        assert (mrange[:, 0] == mstart).all() and (end == mstart + np.repeat(msizes, msizes)).all()
        off = int(desc.mcfg) - base
        mcfg = buf[off:off + 16 * m].view(torch.int64).cpu().numpy().reshape(m, 2)
        toff = int(targets.desc.keys) - targets.buffer.data_ptr()
        words = targets.buffer[toff:toff + 16 * n].view(torch.int64).cpu().numpy().reshape(n, 2)
        assert (mcfg == words[mrows]).all()
        # the hop-parity string: bit j = site 2 j + parity; merged groups flagged on every entry
        off = int(desc.mstr) - base
        mstr = buf[off:off + 8 * m].view(torch.int64).cpu().numpy().view(np.uint64)
        mocc = occ[mrows]
        want = np.zeros(m, dtype=np.uint64)
        for j in range(40):
            want |= mocc[:, 2 * j + parity].astype(np.uint64) << np.uint64(j)
        assert (mstr == want).all()
        mother = mocc[:, (1 - parity)::2]
        for lo in np.unique(mrange[:, 0]):
            hi = int(end[lo])
            pure = bool((mother[lo:hi] == mother[lo]).all())
            assert (merged_bit[lo:hi] == (0 if pure else 1)).all()


def test_shared_string_set_falls_back_to_enumeration(cuda_device: torch.device) -> None:
    # every row holds the same down string: one group of n rows for the up hops
    site, kind, coef = (x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(12, 1, 1.0, 4.0)))
    rng = np.random.default_rng(7800)
    ups = set()
    while len(ups) < 900:
        ups.add(tuple(sorted(rng.choice(12, 6, replace=False).tolist())))
    occ = np.zeros((len(ups), 24), dtype=np.uint8)
    for r, u in enumerate(sorted(ups)):
        occ[r, [2 * s for s in u]] = 1
        occ[r, [2 * s + 1 for s in range(6)]] = 1
    cfg = sy.sort_unique_configs(sy.pack_occupation(occ))
    w = _prepare(site, kind, coef, cfg, 3, cuda_device)
    assert w["terms"].hops
    assert not pp.can_use_hops(w["terms"], w["targets"], w["cfg_t"])
    vec = sy.random_psi(len(cfg), 7801)
    auto = _apply(w, vec, pp.MODE_FORWARD, None)
    enum = _apply(w, vec, pp.MODE_FORWARD, False)
    assert float(np.abs(auto - enum).max()) <= 1e-13 * float(np.abs(enum).max())
    with pytest.raises(RuntimeError, match="hop path"):
        _apply(w, vec, pp.MODE_FORWARD, True)


def test_hubbard128_hops_match_enumeration(cuda_device: torch.device) -> None:
    site, kind, coef = (x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(64, 1, 1.0, 4.0)))
    cfg = sy.hubbard_chain_configs(64, 4_000_000, 3000)
    w = _prepare(site, kind, coef, cfg, 16, cuda_device)
    stats = [w["targets"].strings_stats(p) for p in (0, 1)]
    print("hubbard128 strings", stats, flush=True)
    _check_against_enumeration(w, 7900)


def test_hop_entry_points_check_arguments(cuda_device: torch.device) -> None:
    from qmb_hip import _lib  # pylint: disable=import-outside-toplevel

    lib = _lib.library()
    cfg = _hubbard_configs(8, 500, 7950)
    w = _prepare(*(x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(8, 1, 1.0, 4.0))), cfg, 2, cuda_device)
    out = ctypes.c_void_p(0)
    rc = lib.qmb_strings_build(ctypes.byref(w["targets"].desc), 2, ctypes.c_void_p(1), 1 << 20, ctypes.byref(_lib.QmbStrings()), out)
    assert rc == _lib.QMB_ERR_ARGUMENT
    size = ctypes.c_size_t(0)
    assert lib.qmb_strings_workspace_size(17, 100, ctypes.byref(size)) != 0


def _c_abi_call(lib: object, nq: int, cfg_t: torch.Tensor, terms: tuple[torch.Tensor, ...], vec: np.ndarray, adjoint: bool,
                ws_bytes: int, dev: torch.device) -> tuple[np.ndarray, float]:
    import time  # pylint: disable=import-outside-toplevel

    from qmb_hip import _lib  # pylint: disable=import-outside-toplevel

    n = cfg_t.size(0)
    v = torch.view_as_real(torch.from_numpy(vec).to(dev)).contiguous()
    out = torch.empty((n, 2), dtype=torch.float64, device=dev)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    cp = ctypes.c_void_p(cfg_t.data_ptr())
    tp = [ctypes.c_void_p(t.data_ptr()) for t in terms]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if adjoint:
        rc = lib.qmb_apply_within_adjoint(nq, 1, cp, n, cp, ctypes.c_void_p(v.data_ptr()), n, *tp, terms[0].size(0),
                                          ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws_bytes, stream)
    else:
        rc = lib.qmb_apply_within(nq, 1, cp, ctypes.c_void_p(v.data_ptr()), n, cp, n, *tp, terms[0].size(0),
                                  ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws_bytes, stream)
    _lib.check(rc, "qmb_apply_within" + ("_adjoint" if adjoint else ""))
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    return torch.view_as_complex(out).cpu().numpy(), seconds


@pytest.mark.parametrize("sites,count", [(40, 20000), (64, 4_000_000)], ids=["chain-80so", "hubbard128"])
def test_c_abi_takes_hop_path(cuda_device: torch.device, sites: int, count: int) -> None:
    from qmb_hip import _lib  # pylint: disable=import-outside-toplevel

    lib = _lib.library()
    site, kind, coef = (x.contiguous().to(cuda_device) for x in sy.terms_from_dict(sy.hubbard_dict(sites, 1, 1.0, 4.0)))
    nq = (2 * sites + 7) // 8
    cfg = sy.hubbard_chain_configs(sites, count, 7990 + sites)
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    n = len(cfg)
    hops = ctypes.c_size_t(0)
    _lib.check(lib.qmb_apply_within_hops_workspace_size(nq, n, site.size(0), ctypes.byref(hops)), "hops size")
    small = ctypes.c_size_t(0)
    _lib.check(lib.qmb_apply_within_workspace_size(nq, n, n, site.size(0), ctypes.byref(small)), "enumeration size")
    assert hops.value > small.value
    modes = (False, True) if count < 100_000 else (False,)
    for adjoint in modes:
        vec = sy.random_psi(n, 7995 + int(adjoint))
        _c_abi_call(lib, nq, cfg_t, (site, kind, coef), vec, adjoint, hops.value, cuda_device)  # first call: page in
        got, t_hop = _c_abi_call(lib, nq, cfg_t, (site, kind, coef), vec, adjoint, hops.value, cuda_device)
        want, t_enum = _c_abi_call(lib, nq, cfg_t, (site, kind, coef), vec, adjoint, small.value, cuda_device)
        print(f"{sites} sites, {n} configs, adjoint={adjoint}: hop path {t_hop * 1e3:.2f} ms, enumeration {t_enum * 1e3:.2f} ms",
              flush=True)
        scale = float(np.abs(want).max())
        assert float(np.abs(got - want).max()) <= 1e-12 * scale
