"""
Host-side checks of libqmb_hip.so that need no GPU: the library loads, it exports every entry point
include/qmb_hip.h declares, and its term compiler + record application (the code every kernel runs)
agree bit for bit with the oracle's restatement of qmb/_hamiltonian_cpu.cpp:48-78.
"""

from __future__ import annotations

import ctypes
import pathlib
import re

import numpy as np
import pytest

from oracle import oracle as orc
from qmb_hip import _lib

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "qmb_hip.h"


def declared_functions() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(qmb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function() -> None:
    lib = _lib.library()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.exported_symbols())


def test_version_and_error_channel() -> None:
    lib = _lib.library()
    assert lib.qmb_version() >= 10000
    out = ctypes.c_size_t(0)
    assert lib.qmb_terms_workspace_size(0, 10, ctypes.byref(out)) != 0
    assert b"n_qubytes" in lib.qmb_last_error()


def _compile(site4: np.ndarray, kind4: np.ndarray, pc: int, nq: int) -> tuple[int, int, int, int]:
    lib = _lib.library()
    s = np.ascontiguousarray(site4, dtype=np.int16)
    k = np.ascontiguousarray(kind4, dtype=np.uint8)
    rec = ctypes.c_uint64(0)
    status = ctypes.c_int32(0)
    bucket = ctypes.c_int64(0)
    diag = ctypes.c_int32(0)
    _lib.check(lib.qmb_compile_term_host(s.ctypes.data, k.ctypes.data, pc, nq, ctypes.byref(rec), ctypes.byref(status),
                                         ctypes.byref(bucket), ctypes.byref(diag)), "compile")
    return rec.value, status.value, bucket.value, diag.value


def _apply(rec: int, config: np.ndarray) -> tuple[bool, np.ndarray, int]:
    lib = _lib.library()
    cfg = np.ascontiguousarray(config, dtype=np.uint8)
    out = np.zeros_like(cfg)
    par = ctypes.c_int32(0)
    ok = lib.qmb_apply_record_host(rec, len(cfg), cfg.ctypes.data, out.ctypes.data, ctypes.byref(par))
    return bool(ok), out, par.value


@pytest.mark.parametrize("pc", [1, 2])
@pytest.mark.parametrize("nq", [1, 2, 9, 17, 32])
def test_compiled_records_match_oracle(pc: int, nq: int) -> None:
    rng = np.random.default_rng(1000 * pc + nq)
    n_bits = 8 * nq
    for _ in range(400):
        n_ops = int(rng.integers(0, 5))
        pool = rng.integers(0, n_bits, size=3)  # few distinct sites -> repeated sites are common
        site = np.zeros(4, dtype=np.int16)
        kind = np.full(4, 2, dtype=np.uint8)
        site[:n_ops] = rng.choice(pool, size=n_ops)
        kind[:n_ops] = rng.integers(0, 2, size=n_ops)
        rec, status, bucket, diag = _compile(site, kind, pc, nq)
        assert status in (0, 1)
        for _ in range(8):
            cfg = rng.integers(0, 256, size=nq, dtype=np.uint8)
            if rng.random() < 0.5:  # bias toward configs on which the term can succeed
                for op in range(n_ops):
                    s = int(site[op])
                    want_bit = 1 - int(kind[op])
                    cfg[s // 8] = (int(cfg[s // 8]) & (255 ^ (1 << (s % 8)))) | (want_bit << (s % 8))
            ok_o, out_o, par_o = orc.apply_term(cfg, site, kind, pc)
            ok_l, out_l, par_l = _apply(rec, cfg) if status == 0 else (False, cfg, 0)
            assert ok_o == ok_l, (site, kind, cfg)
            if ok_o:
                assert np.array_equal(out_o, out_l) and par_o == par_l, (site, kind, cfg)
                if diag:
                    assert np.array_equal(out_o, cfg)
        assert (bucket >= 0) == (status == 0)


def test_compile_rejects_bad_terms() -> None:
    assert _compile(np.array([9, 0, 0, 0]), np.array([1, 2, 2, 2]), 1, 1)[1] == 2  # site outside one byte
    assert _compile(np.array([1, 0, 0, 0]), np.array([3, 2, 2, 2]), 1, 1)[1] == 3  # kind 3
    assert _compile(np.array([1, 1, 0, 0]), np.array([0, 0, 2, 2]), 1, 1)[1] == 1  # a_1 a_1 is zero


def test_bucket_is_required_occupied_set() -> None:
    nq = 2
    n_bits = 16
    # a+_5 a+_3 a_9 a_2: requires 2 and 9 occupied -> pair bucket of (2, 9)
    _, status, bucket, _ = _compile(np.array([5, 3, 9, 2]), np.array([1, 1, 0, 0]), 1, nq)
    assert status == 0 and bucket == 1 + n_bits + 9 * 8 // 2 + 2
    # n_4 = a+_4 a_4: requires 4 -> single bucket
    assert _compile(np.array([4, 4, 0, 0]), np.array([1, 0, 2, 2]), 1, nq)[2] == 1 + 4
    # identity -> bucket 0
    assert _compile(np.array([0, 0, 0, 0]), np.array([2, 2, 2, 2]), 1, nq)[2] == 0


def test_philox_known_answers() -> None:
    """Philox4x32-10 known-answer vectors of the Random123 distribution (Salmon et al., SC'11)."""
    lib = _lib.library()
    cases = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0], [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in cases:
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        out = (ctypes.c_uint32 * 4)()
        lib.qmb_philox4x32_10_host(c, k, out)
        assert list(out) == want


def test_philox_uniform_range_and_moments() -> None:
    lib = _lib.library()
    u = np.array([lib.qmb_philox_uniform_host(12345, i) for i in range(20000)])
    assert np.all(u > 0) and np.all(u <= 1)
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005


def test_validate_terms_host() -> None:
    lib = _lib.library()
    site = np.array([[1, 2, 0, 0], [9, 0, 0, 0], [3, 0, 0, 0]], dtype=np.int16)
    kind = np.array([[1, 0, 2, 2], [1, 2, 2, 2], [7, 2, 2, 2]], dtype=np.uint8)
    bad = ctypes.c_int64(0)
    _lib.check(lib.qmb_validate_terms_host(1, site.ctypes.data, kind.ctypes.data, 3, ctypes.byref(bad)), "validate")
    assert bad.value == 2


def test_hermiticity_check() -> None:
    """The exact adjoint test accepts Hermitian model tables and rejects perturbed ones."""
    import torch  # pylint: disable=import-outside-toplevel

    from qmb_hip import prepared as pp  # pylint: disable=import-outside-toplevel
    from qmb_hip import synthetic as sy  # pylint: disable=import-outside-toplevel

    s, k, c = sy.molecular_terms(4, 3)
    assert pp.is_hermitian(s, k, c, 1, 1)
    c2 = c.clone()
    c2[5, 1] = 0.3  # a complex coefficient without its conjugate partner
    assert not pp.is_hermitian(s, k, c2, 1, 1)
    # a+_3 a_1 alone is not Hermitian; adding its adjoint a+_1 a_3 (written in any order) makes it so
    hop = {((3, 1), (1, 0)): 0.5 + 0.25j}
    assert not pp.is_hermitian(*sy.terms_from_dict(hop), 1, 1)
    hop[((1, 1), (3, 0))] = 0.5 - 0.25j
    assert pp.is_hermitian(*sy.terms_from_dict(hop), 1, 1)
    # (a+_0 a+_2 a_3 a_1)^dagger = a+_1 a+_3 a_2 a_0 = a+_3 a+_1 a_0 a_2 (two swaps): same coefficient
    pair = {((0, 1), (2, 1), (3, 0), (1, 0)): 0.7, ((3, 1), (1, 1), (0, 0), (2, 0)): 0.7}
    assert pp.is_hermitian(*sy.terms_from_dict(pair), 1, 1)
    pair_wrong = {((0, 1), (2, 1), (3, 0), (1, 0)): 0.7, ((3, 1), (1, 1), (0, 0), (2, 0)): -0.7}
    assert not pp.is_hermitian(*sy.terms_from_dict(pair_wrong), 1, 1)
    # one swap: (a+_0 a_1)^dagger = a+_1 a_0 = -a_0 a+_1
    hop2 = {((0, 1), (1, 0)): 0.3, ((0, 0), (1, 1)): -0.3}
    assert pp.is_hermitian(*sy.terms_from_dict(hop2), 1, 1)
    assert pp.is_hermitian(*sy.terms_from_dict(sy.ising_dict(3, 2, x=0.5, y=0.2, zh=1.0, yv=0.4)), 1, 2)
    assert not pp.is_hermitian(*sy.random_terms(10, 50, 1), 2, 1)
    assert isinstance(torch.zeros(1), torch.Tensor)


def _pair_parity(config: np.ndarray, sites: tuple[int, int, int, int]) -> int:
    """popcount(c & (L(r0) ^ L(r1) ^ L(c0) ^ L(c1))) mod 2 with L(s) = the sites below s."""
    bits = np.unpackbits(config, bitorder="little").astype(np.int64)
    return int(sum(bits[:s].sum() for s in sites) & 1)


@pytest.mark.parametrize("pc", [1, 2])
@pytest.mark.parametrize("nq", [1, 2, 7, 16])
def test_pair_excitation_classifier_matches_oracle(pc: int, nq: int) -> None:
    """Every operator order of two creations and two annihilations on distinct sites is a pair
    excitation whose oracle sign is (-1)^(pair_parity(c) ^ delta) on every config it applies to;
    terms with repeated sites, other kind mixes or fewer ops are not."""
    lib = _lib.library()
    rng = np.random.default_rng(7000 + 10 * nq + pc)
    n_bits = 8 * nq
    out = (ctypes.c_int32 * 5)()
    seen_pair = seen_other = 0
    for _ in range(300):
        if rng.random() < 0.7 and n_bits >= 4:
            site = rng.choice(n_bits, size=4, replace=False).astype(np.int16)
            kind = rng.permutation(np.array([0, 0, 1, 1], dtype=np.uint8))
        else:
            n_ops = int(rng.integers(1, 5))
            site = np.zeros(4, dtype=np.int16)
            kind = np.full(4, 2, dtype=np.uint8)
            site[:n_ops] = rng.integers(0, min(n_bits, 6), size=n_ops)
            kind[:n_ops] = rng.integers(0, 2, size=n_ops)
        rec, status, _, _ = _compile(site, kind, pc, nq)
        is_pair = bool(lib.qmb_pair_excitation_host(rec, pc, out))
        distinct = len(set(site.tolist())) == 4 and np.all(kind < 2)
        expect = status == 0 and distinct and sorted(kind.tolist()) == [0, 0, 1, 1]
        assert is_pair == expect, (site, kind)
        if not is_pair:
            seen_other += 1
            continue
        seen_pair += 1
        r0, r1, c0, c1, delta = list(out)
        ann = sorted(int(s) for s, k in zip(site, kind) if k == 0)
        cre = sorted(int(s) for s, k in zip(site, kind) if k == 1)
        assert [r0, r1] == ann and [c0, c1] == cre
        for _ in range(6):
            cfg = rng.integers(0, 256, size=nq, dtype=np.uint8)
            bits = np.unpackbits(cfg, bitorder="little")
            bits[[r0, r1]] = 1
            bits[[c0, c1]] = 0
            cfg = np.packbits(bits, bitorder="little")
            ok, res, par = orc.apply_term(cfg, site, kind, pc)
            assert ok
            want = bits.copy()
            want[[r0, r1]] = 0
            want[[c0, c1]] = 1
            np.testing.assert_array_equal(res, np.packbits(want, bitorder="little"))
            sign = (_pair_parity(cfg, (r0, r1, c0, c1)) ^ delta) if pc == 1 else 0
            assert par == sign, (site, kind, cfg)
    assert seen_pair > 50 and seen_other > 20


def _single_parity(config: np.ndarray, r: int, p: int) -> int:
    bits = np.unpackbits(config, bitorder="little").astype(np.int64)
    return int((bits[:r].sum() + bits[:p].sum()) & 1)


@pytest.mark.parametrize("pc", [1, 2])
@pytest.mark.parametrize("nq", [1, 2, 7, 16])
def test_single_excitation_classifier_matches_oracle(pc: int, nq: int) -> None:
    """Hops, hops dressed by an occupied or empty site in every operator order, and repeated-op forms
    are single excitations whose oracle sign is (-1)^(single_parity(c) ^ delta) on every config they
    apply to; pair excitations, diagonal terms and particle-number changes are not."""
    lib = _lib.library()
    rng = np.random.default_rng(9000 + 10 * nq + pc)
    n_bits = 8 * nq
    out = (ctypes.c_int32 * 5)()
    seen = {"single": 0, "other": 0}
    for _ in range(400):
        site = np.zeros(4, dtype=np.int16)
        kind = np.full(4, 2, dtype=np.uint8)
        u = rng.random()
        if u < 0.3 and n_bits >= 2:
            site[:2] = rng.choice(n_bits, size=2, replace=False)
            kind[:2] = rng.permutation([0, 1])
        elif u < 0.7 and n_bits >= 3:
            p, k, r = rng.choice(n_bits, size=3, replace=False)
            order = rng.permutation(4)
            site[order] = [p, k, k, r]
            kind[order] = [1, 1, 0, 0] if rng.random() < 0.5 else [1, 0, 1, 0]
        else:
            n_ops = int(rng.integers(1, 5))
            site[:n_ops] = rng.integers(0, min(n_bits, 5), size=n_ops)
            kind[:n_ops] = rng.integers(0, 2, size=n_ops)
        rec, status, _, _ = _compile(site, kind, pc, nq)
        if status != 0:
            continue
        is_single = bool(lib.qmb_single_excitation_host(rec, pc, out))
        # the host classifier must agree with the action of the term on configs: exactly one site
        # emptied, one filled, at most one other site looked at
        used = sorted({int(s) for s, kk in zip(site, kind) if kk < 2})
        if not is_single:
            seen["other"] += 1
            continue
        seen["single"] += 1
        r, p, k, kocc, delta = list(out)
        assert r in used and p in used and r != p and (k == -1 or (k in used and k not in (r, p)))
        hits = 0
        for _ in range(24):
            cfg = rng.integers(0, 256, size=nq, dtype=np.uint8)
            ok, res, par = orc.apply_term(cfg, site, kind, pc)
            bits = np.unpackbits(cfg, bitorder="little")
            expect_ok = bits[r] == 1 and bits[p] == 0 and (k == -1 or bits[k] == kocc)
            assert ok == expect_ok, (site, kind, cfg)
            if not ok:
                continue
            hits += 1
            want = bits.copy()
            want[r], want[p] = 0, 1
            np.testing.assert_array_equal(res, np.packbits(want, bitorder="little"))
            sign = (_single_parity(cfg, r, p) ^ delta) if pc == 1 else 0
            assert par == sign, (site, kind, cfg)
    assert seen["single"] > 100 and seen["other"] > 20


def test_batch_plan_rejects_bad_items_without_touching_the_device() -> None:
    """qmb_batch_plan validates every item on the host before any copy; these calls never reach HIP."""
    lib = _lib.library()
    size = ctypes.c_size_t(0)
    assert lib.qmb_batch_workspace_size(0, ctypes.byref(size)) != 0
    assert b"n_items" in lib.qmb_last_error()
    assert lib.qmb_batch_workspace_size(3, ctypes.byref(size)) == 0 and size.value > 0
    items = (_lib.QmbBatchItem * 2)()
    plan = _lib.QmbBatch()
    buf = ctypes.create_string_buffer(size.value)
    # item 0 without its pair table
    assert lib.qmb_batch_plan(items, 2, ctypes.cast(buf, ctypes.c_void_p), size.value, ctypes.byref(plan), None) != 0
    assert b"item 0" in lib.qmb_last_error()
    # a pair table whose core index is missing
    pairs = _lib.QmbPairs(n_qubytes=1, particle_cut=1, n_bits=8)
    targets = _lib.QmbTargets(n_qubytes=1, n_words=1, n=0)
    items[0].pairs = ctypes.pointer(pairs)
    items[0].targets = ctypes.pointer(targets)
    assert lib.qmb_batch_plan(items, 1, ctypes.cast(buf, ctypes.c_void_p), size.value, ctypes.byref(plan), None) != 0
    assert b"core" in lib.qmb_last_error()
    # a buffer that is too small
    assert lib.qmb_batch_plan(items, 2, ctypes.cast(buf, ctypes.c_void_p), 16, ctypes.byref(plan), None) != 0
    assert b"too small" in lib.qmb_last_error()
    # mode checks of the apply entry points happen before any launch
    assert lib.qmb_apply_batch(None, 0, 0, None) != 0


def test_structured_builders_reject_bad_arguments() -> None:
    lib = _lib.library()
    size = ctypes.c_size_t(0)
    assert lib.qmb_pairs_workspace_size(17, ctypes.byref(size)) != 0  # the dense pair table stops at 16 bytes
    assert lib.qmb_pairs_workspace_size(7, ctypes.byref(size)) == 0
    n_pairs = 56 * 55 // 2
    assert size.value >= 16 * n_pairs * n_pairs
    assert lib.qmb_singles_workspace_size(7, ctypes.byref(size)) == 0 and size.value >= 16 * 56 ** 3
    assert lib.qmb_cores_workspace_size(7, 1000, 14, 3, ctypes.byref(size)) != 0  # order must be 1 or 2
    assert lib.qmb_cores_workspace_size(7, 1000, 57, 2, ctypes.byref(size)) != 0  # more occupied sites than sites
    assert b"max_occupied" in lib.qmb_last_error()
    buf = ctypes.create_string_buffer(64)
    terms = _lib.QmbTerms()
    assert lib.qmb_terms_build_residual(7, 1, None, None, None, 0, 8, ctypes.cast(buf, ctypes.c_void_p), 64,
                                        ctypes.byref(terms), None) != 0  # exclude mask out of range
