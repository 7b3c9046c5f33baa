"""
GPU parity at the plain-forward workload's full size: 52 spin orbitals, 7a+7b, |(pq|rs)| < 1e-10
dropped (669,579 terms), 1,000,000 configs (HF + singles + doubles + seeded triples/quadruples).

  - 16 random rows of H psi and of H^H g from the structured path equal the oracle's, to 1e-12 of
    each row's sum of |contributions|.  The oracle gives column j of H (configs_i = row j, psi = 1,
    configs_j = the whole set, with the per-target |contribution| sums); the table is Hermitian, so
    (H psi)_j = <H[:, j], psi>.
  - The structured path (pair / single core groups + folded diagonal) equals the enumeration path
    (hash lookups of every generated target) on every row, forward and adjoint.
  - <phi, H psi> = conj(<psi, H phi>) over the whole set.
  - find_relative(K = 100,000, exclude = the set): the output properties of the reference
    (qmb/_hamiltonian_cpu.cpp:442-546) and nesting of the K = 50,000 selection.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import prepared as pp
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def molecular52(cuda_device: torch.device) -> dict[str, object]:
    s, k, c = sorted_terms(*sy.molecular_terms(26, 2001, threshold=1e-10))
    assert len(s) == 669_579
    cfg = sy.excitation_configs(26, 7, 7, 1_000_000, 2002)
    st, kt, ct = (torch.from_numpy(np.ascontiguousarray(x)).to(cuda_device) for x in (s, k, c))
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, 7, 1)
    targets = pp.PreparedTargets(cfg_t, 7, distinct=True, max_occupied=14)
    psi = sy.random_psi(len(cfg), 2003)
    g = sy.random_psi(len(cfg), 2004)
    assert pp.can_use_pairs(terms, targets, cfg_t)
    return {"terms_np": (s, k, c), "cfg": cfg, "cfg_t": cfg_t, "terms": terms, "targets": targets, "psi": psi, "g": g,
            "device": cuda_device}


def _apply(w: dict[str, object], vec: np.ndarray, mode: int, pairs: bool | None = None) -> np.ndarray:
    v = torch.view_as_real(torch.from_numpy(vec).to(w["device"])).contiguous()
    out = pp.apply(w["terms"], w["targets"], w["cfg_t"], v, mode, pairs=pairs)
    return torch.view_as_complex(out).cpu().numpy()


def test_molecular52_rows_match_oracle(molecular52: dict[str, object]) -> None:
    w = molecular52
    cfg, psi, g = w["cfg"], w["psi"], w["g"]
    fwd = _apply(w, psi, pp.MODE_FORWARD)
    adj = _apply(w, g, pp.MODE_ADJOINT)
    for j in np.random.default_rng(2005).choice(len(cfg), size=16, replace=False):
        column, asum, (_, m_hit) = orc.apply_within(1, cfg[j:j + 1], np.ones(1, dtype=np.complex128), cfg, *w["terms_np"])
        assert m_hit > 1  # the row reaches others, not just itself
        for got, vec in ((fwd[j], psi), (adj[j], g)):
            want = np.vdot(column, vec)
            bound = 1e-12 * float(asum @ np.abs(vec))
            assert abs(got - want) <= bound, (int(j), got, want, bound)


def test_molecular52_structured_matches_enumeration(molecular52: dict[str, object]) -> None:
    w = molecular52
    for vec, mode in ((w["psi"], pp.MODE_FORWARD), (w["g"], pp.MODE_ADJOINT)):
        structured = _apply(w, vec, mode)
        enumerated = _apply(w, vec, mode, pairs=False)
        scale = float(np.abs(enumerated).max())
        assert float(np.abs(structured - enumerated).max()) <= 1e-12 * scale


def test_molecular52_hermitian_identity(molecular52: dict[str, object]) -> None:
    w = molecular52
    phi, psi = w["g"], w["psi"]
    lhs = np.vdot(phi, _apply(w, psi, pp.MODE_FORWARD))
    rhs = np.conj(np.vdot(psi, _apply(w, phi, pp.MODE_FORWARD)))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_molecular52_find_relative_properties(molecular52: dict[str, object]) -> None:
    """find_relative(K = 100,000, exclude = the set) at full size: at most K rows, distinct, sorted by
    bytes, none in the exclude set, all in the 7a+7b sector; the rows kept for K = 50,000 are among
    those kept for K = 100,000 (top-50k entries are top-100k entries; a boundary tie class may swap a
    few rows)."""
    from qmb_hip import ops  # pylint: disable=import-outside-toplevel

    w = molecular52
    dev = w["device"]
    s, k, c = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in w["terms_np"])
    psi = torch.view_as_real(torch.from_numpy(w["psi"]).to(dev)).contiguous()
    ns = ops.load(7, 1)
    big = ns.find_relative(w["cfg_t"], psi, 100_000, s, k, c, w["cfg_t"]).cpu().numpy()
    small = ns.find_relative(w["cfg_t"], psi, 50_000, s, k, c, w["cfg_t"]).cpu().numpy()
    for rows, kk in ((big, 100_000), (small, 50_000)):
        assert 0.9 * kk <= len(rows) <= kk
        keys = [r.tobytes() for r in rows]
        assert keys == sorted(set(keys))
        assert not set(keys) & {r.tobytes() for r in w["cfg"]}
        occ = sy.unpack_occupation(rows, 52)
        assert (occ[:, 0::2].sum(axis=1) == 7).all() and (occ[:, 1::2].sum(axis=1) == 7).all()
    assert len({r.tobytes() for r in small} - {r.tobytes() for r in big}) <= 64


def test_two_word_structured_matches_enumeration(cuda_device: torch.device) -> None:
    """Configs of two words (76 spin orbitals, 13a+13b, 3.06M terms, 200,000 configs): the
    structured path equals the enumeration path on every row, forward and adjoint, and the oracle
    on six random rows.  The set's 65M order-2 entries hold thousands of groups where cores with
    equal 34-bit hashes meet (7,586 of 28.2M groups), many with an entry of the other core first
    and cores with sites in the second word: the exact core comparison must hold there for every
    word."""
    s, k, c = sorted_terms(*sy.molecular_terms(38, 5001, threshold=1e-8))
    cfg = sy.excitation_configs(38, 13, 13, 200_000, 5002)
    st, kt, ct = (torch.from_numpy(np.ascontiguousarray(x)).to(cuda_device) for x in (s, k, c))
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, 10, 1)
    targets = pp.PreparedTargets(cfg_t, 10, distinct=True, max_occupied=26)
    assert pp.can_use_pairs(terms, targets, cfg_t)
    w = {"terms": terms, "targets": targets, "cfg_t": cfg_t, "device": cuda_device}
    results = []
    for seed, mode in ((5003, pp.MODE_FORWARD), (5004, pp.MODE_ADJOINT)):
        vec = sy.random_psi(len(cfg), seed)
        structured = _apply(w, vec, mode)
        enumerated = _apply(w, vec, mode, pairs=False)
        scale = float(np.abs(enumerated).max())
        assert float(np.abs(structured - enumerated).max()) <= 1e-12 * scale
        results.append((structured, vec))
    # six rows against the oracle's column of H (the table is Hermitian: row j = conj(column j))
    for j in np.random.default_rng(5005).choice(len(cfg), size=6, replace=False):
        column, asum, (_, m_hit) = orc.apply_within(1, cfg[j:j + 1], np.ones(1, dtype=np.complex128), cfg, s, k, c)
        assert m_hit > 1
        for got, vec in results:
            want = np.vdot(column, vec)
            bound = 1e-12 * float(asum @ np.abs(vec))
            assert abs(got[j] - want) <= bound, (int(j), got[j], want, bound)


def test_molecular52_c_abi_takes_structured_path(molecular52: dict[str, object]) -> None:
    """qmb_apply_within / qmb_apply_within_adjoint through the C ABI with one config buffer for both
    sets and a workspace of qmb_apply_within_structured_workspace_size: the library builds the pair,
    single and residual tables and the core indexes in that workspace and applies the structured
    path.  The 16 oracle rows of test_molecular52_rows_match_oracle hold to 1e-12 of each row's
    sum of |contributions|, every row equals the prepared structured path, and the call is far
    faster than the enumeration it takes with a workspace of qmb_apply_within_workspace_size."""
    import ctypes  # pylint: disable=import-outside-toplevel
    import time  # pylint: disable=import-outside-toplevel

    from qmb_hip import _lib  # pylint: disable=import-outside-toplevel

    w = molecular52
    dev = w["device"]
    lib = _lib.library()
    cfg, n = w["cfg"], len(w["cfg"])
    s, k, c = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in w["terms_np"])
    n_terms = s.size(0)
    size = ctypes.c_size_t(0)
    _lib.check(lib.qmb_apply_within_structured_workspace_size(7, n, n_terms, 14, ctypes.byref(size)), "structured size")
    small = ctypes.c_size_t(0)
    _lib.check(lib.qmb_apply_within_workspace_size(7, n, n, n_terms, ctypes.byref(small)), "enumeration size")
    assert size.value > small.value
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    cfg_p = ctypes.c_void_p(w["cfg_t"].data_ptr())
    tp = [ctypes.c_void_p(t.data_ptr()) for t in (s, k, c)]

    def call(vec: np.ndarray, adjoint: bool, ws_bytes: int) -> tuple[np.ndarray, float]:
        v = torch.view_as_real(torch.from_numpy(vec).to(dev)).contiguous()
        out = torch.empty((n, 2), dtype=torch.float64, device=dev)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if adjoint:
            rc = lib.qmb_apply_within_adjoint(7, 1, cfg_p, n, cfg_p, ctypes.c_void_p(v.data_ptr()), n, *tp, n_terms,
                                              ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws_bytes, stream)
        else:
            rc = lib.qmb_apply_within(7, 1, cfg_p, ctypes.c_void_p(v.data_ptr()), n, cfg_p, n, *tp, n_terms,
                                      ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws_bytes, stream)
        _lib.check(rc, "qmb_apply_within" + ("_adjoint" if adjoint else ""))
        torch.cuda.synchronize()
        return torch.view_as_complex(out).cpu().numpy(), time.perf_counter() - t0

    call(w["psi"], False, size.value)  # first call: page in the workspace
    fwd, t_struct = call(w["psi"], False, size.value)
    adj, _ = call(w["g"], True, size.value)
    enum, t_enum = call(w["psi"], False, small.value)
    print(f"C ABI apply_within at 1M configs: structured {t_struct * 1e3:.1f} ms, enumeration {t_enum * 1e3:.1f} ms", flush=True)
    for j in np.random.default_rng(2005).choice(n, size=16, replace=False):
        column, asum, _ = orc.apply_within(1, cfg[j:j + 1], np.ones(1, dtype=np.complex128), cfg, *w["terms_np"])
        for got, vec in ((fwd[j], w["psi"]), (adj[j], w["g"])):
            want = np.vdot(column, vec)
            assert abs(got - want) <= 1e-12 * float(asum @ np.abs(vec)), (int(j), got, want)
    prepared = _apply(w, w["psi"], pp.MODE_FORWARD)
    scale = float(np.abs(prepared).max())
    assert float(np.abs(fwd - prepared).max()) <= 1e-12 * scale
    assert float(np.abs(fwd - enum).max()) <= 1e-12 * scale
    assert t_struct < 0.5 * t_enum
