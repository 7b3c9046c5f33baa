"""
The CPU dispatch key's operators (the library's host implementation, qmb_cpu_*) against the golden
vectors the reference's own CPU op wrote (tests/golden/), through the reference's op names:

  - apply_within: within 1e-12 of the reference relative to each entry's sum of |contributions|
    (the condition-aware bound; the reference sums in a different order under -ffast-math), on all
    golden cases including the 10,000-config molecular set;
  - find_relative: the same rows as the reference, up to configs whose heaviest contribution ties
    the K-th weight (the reference's heap order among equal weights is not reproduced);
  - the adjoint: <g, A psi> = <A^H g, psi> and, for Hermitian tables on one set, A^H = A;
  - autograd through apply_within on CPU tensors reaches psi through the adjoint;
  - results do not depend on the thread count (bit for bit for find_relative and the adjoint);
  - argument errors carry the reference's messages; diagonal_term / single_relative stay
    GPU-only, as in the reference.
"""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest
import torch

from brute import contribution_weights
from oracle import oracle as orc
from qmb_hip import Hamiltonian
from qmb_hip import ops
from qmb_hip import synthetic as sy

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
MANIFEST = json.loads((GOLDEN / "manifest.json").read_text())
APPLY = [m["name"] for m in MANIFEST if m["op"] == "apply_within"]
FIND = [m["name"] for m in MANIFEST if m["op"] == "find_relative"]


def load(prefix: str, name: str) -> dict[str, np.ndarray]:
    with np.load(GOLDEN / f"{prefix}_{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _t(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a))


def _real(psi: np.ndarray) -> torch.Tensor:
    return _t(np.ascontiguousarray(psi, dtype=np.complex128).view(np.float64).reshape(-1, 2))


def _complex(out: torch.Tensor) -> np.ndarray:
    a = out.numpy()
    return a[:, 0] + 1j * a[:, 1]


@pytest.mark.parametrize("name", APPLY)
def test_cpu_apply_within_matches_reference(name: str) -> None:
    d = load("aw", name)
    pc = int(d["particle_cut"])
    nq = d["configs_i"].shape[1]
    ns = ops.load(nq, pc)
    out = ns.apply_within(_t(d["configs_i"]), _real(d["psi_i"]), _t(d["configs_j"]), _t(d["site"]), _t(d["kind"]), _t(d["coef"]))
    _, asum, _ = orc.apply_within(pc, d["configs_i"], d["psi_i"], d["configs_j"], d["site"], d["kind"], d["coef"])
    err = np.abs(_complex(out) - d["result"])
    assert np.all(err <= 1e-12 * asum), (name, float(err.max()))
    assert np.all(_complex(out)[asum == 0] == 0)


@pytest.mark.parametrize("name", FIND)
def test_cpu_find_relative_matches_reference(name: str) -> None:
    d = load("fr", name)
    pc = int(d["particle_cut"])
    k_sel = int(d["count_selected"])
    nq = d["configs_i"].shape[1]
    ns = ops.load(nq, pc)
    got = ns.find_relative(_t(d["configs_i"]), _real(d["psi_i"]), k_sel, _t(d["site"]), _t(d["kind"]), _t(d["coef"]),
                           _t(d["configs_exclude"])).numpy()
    rows = [bytes(r) for r in got]
    assert rows == sorted(set(rows)), "rows must be distinct and sorted by bytes"
    diff = set(rows) ^ {bytes(r) for r in d["result"]}
    if diff:
        _, kth = orc.find_relative(pc, d["configs_i"], d["psi_i"], k_sel, d["site"], d["kind"], d["coef"], d["configs_exclude"])
        best: dict[bytes, float] = {}
        for w, key in contribution_weights(pc, d["configs_i"], d["psi_i"], d["configs_exclude"], d["site"], d["kind"], d["coef"]):
            best[key] = max(best.get(key, -1.0), w)
        assert all(best[key] == kth for key in diff), name


def test_cpu_adjoint_identity_and_hermitian() -> None:
    d = load("aw", "molecular_20so_distinct_sets")
    pc = int(d["particle_cut"])
    ci, cj = d["configs_i"], d["configs_j"]
    ns = ops.load(ci.shape[1], pc)
    rng = np.random.default_rng(7)
    g = rng.normal(size=len(cj)) + 1j * rng.normal(size=len(cj))
    args = (_t(d["site"]), _t(d["kind"]), _t(d["coef"]))
    a_psi = _complex(ns.apply_within(_t(ci), _real(d["psi_i"]), _t(cj), *args))
    ah_g = _complex(ns.apply_within_adjoint(_t(ci), _real(g), _t(cj), *args))
    lhs = np.vdot(g, a_psi)
    rhs = np.vdot(ah_g, d["psi_i"])
    assert abs(lhs - rhs) <= 1e-12 * (np.abs(g).sum() * np.abs(a_psi).max() + 1.0)
    # a Hermitian table on one set: the adjoint is the forward
    s, k, c = (x.numpy() for x in sy.molecular_terms(5, 11))
    cfg = sy.sector_configs(5, 2, 2, 80, 12)
    psi = sy.random_psi(len(cfg), 13)
    ns1 = ops.load(cfg.shape[1], 1)
    fwd = ns1.apply_within(_t(cfg), _real(psi), _t(cfg), _t(s), _t(k), _t(c))
    adj = ns1.apply_within_adjoint(_t(cfg), _real(psi), _t(cfg), _t(s), _t(k), _t(c))
    assert torch.allclose(fwd, adj, rtol=0, atol=1e-12 * float(fwd.abs().max()))


def test_cpu_autograd_through_hamiltonian() -> None:
    ham = Hamiltonian(sy.hubbard_dict(3, 2, 1.0, 4.0), kind="fermi")
    cfg = _t(sy.sector_configs(6, 3, 3, 200, 1))
    psi = _t(sy.random_psi(len(cfg), 2)).requires_grad_(True)
    g = _t(sy.random_psi(len(cfg), 3))
    y = ham.apply_within(cfg, psi, cfg)
    y.backward(g)
    # Hermitian table: grad = A^H g = A g
    want = ham.apply_within(cfg, g, cfg)
    assert torch.allclose(psi.grad, want, rtol=0, atol=1e-12 * float(want.abs().max()))
    s, k, c = ham.site.numpy(), ham.kind.numpy(), ham.coef.numpy()
    ref, asum, _ = orc.apply_within(1, cfg.numpy(), psi.detach().numpy(), cfg.numpy(), s, k, c)
    assert np.all(np.abs(y.detach().numpy() - ref) <= 1e-12 * asum)


def test_cpu_results_independent_of_threads() -> None:
    d = load("aw", "c1_molecular_20so_10k")
    ns = ops.load(3, 1)
    args = (_t(d["site"]), _t(d["kind"]), _t(d["coef"]))
    saved = torch.get_num_threads()
    try:
        outs, adjs, finds = [], [], []
        for threads in (1, 3, 8):
            torch.set_num_threads(threads)
            outs.append(ns.apply_within(_t(d["configs_i"]), _real(d["psi_i"]), _t(d["configs_j"]), *args))
            adjs.append(ns.apply_within_adjoint(_t(d["configs_i"]), _real(d["psi_i"]), _t(d["configs_j"]), *args))
            finds.append(ns.find_relative(_t(d["configs_i"][:2000]), _real(d["psi_i"][:2000]), 3000, *args, _t(d["configs_i"])))
    finally:
        torch.set_num_threads(saved)
    scale = float(outs[0].abs().max())
    for o in outs[1:]:
        assert torch.allclose(o, outs[0], rtol=0, atol=1e-13 * scale)
    for a in adjs[1:]:
        assert torch.equal(a, adjs[0])
    for f in finds[1:]:
        assert torch.equal(f, finds[0])


def test_cpu_argument_checks_and_gpu_only_ops() -> None:
    ns = ops.load(1, 1)
    configs = torch.tensor([[0b011], [0b101]], dtype=torch.uint8)
    psi = torch.zeros((2, 2), dtype=torch.float64)
    site = torch.tensor([[0, 1, 0, 0]], dtype=torch.int16)
    kind = torch.tensor([[1, 0, 2, 2]], dtype=torch.uint8)
    coef = torch.tensor([[1.0, 0.0]], dtype=torch.float64)
    with pytest.raises(RuntimeError, match="psi must be float64."):
        ns.apply_within(configs, psi.float(), configs, site, kind, coef)
    with pytest.raises(RuntimeError, match="configs must have 1 columns."):
        ns.apply_within(torch.zeros((2, 2), dtype=torch.uint8), psi, configs, site, kind, coef)
    with pytest.raises(RuntimeError, match="count_selected must be non-negative."):
        ns.find_relative(configs, psi, -1, site, kind, coef, configs)
    assert ns.find_relative(configs, psi, 0, site, kind, coef, configs).shape == (0, 1)
    # a_0^+ a_1 (right to left): 0b010 -> 0b001 with sign +1 (no occupied site below either
    # operator); 0b001 has site 1 empty, so the term fails there
    hop = torch.tensor([[0b010], [0b001]], dtype=torch.uint8)
    out = ns.apply_within(hop, torch.tensor([[1.5, -0.5], [2.0, 0.0]], dtype=torch.float64), hop, site, kind, coef)
    assert out.tolist() == [[0.0, 0.0], [1.5, -0.5]]
    with pytest.raises(NotImplementedError):
        ns.diagonal_term(configs, site, kind, coef)
    with pytest.raises(NotImplementedError):
        ns.single_relative(configs, site, kind, coef)
