"""
GPU checks of the VMC energy losses (qmb/vmc.py:80-125) on the electron up/down network:

  - energy_backward (Rayleigh quotient, adjoint kernel, chunked pull-back) equals PyTorch autograd
    through a dense H built by the oracle, energy to 1e-12 and every parameter gradient to 1e-9
    relative, with the network in float64;
  - closure_energy (the reference's closure) equals its dense formula on one set and on a set
    extended by other configs; on one set its gradient is half the Rayleigh gradient;
  - at the 76-spin-orbital, 13a+13b shape (3.06M terms) the energy on a set near the reference
    determinant equals the oracle's, and H psi per entry to 1e-12 of its sum of |contributions|;
  - a VMC cycle (both losses) lowers the energy and never goes below the lowest eigenvalue of H on
    the sampled set.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import Hamiltonian
from qmb_hip import mlp
from qmb_hip import prepared as pp
from qmb_hip import synthetic as sy
from qmb_hip import vmc

pytestmark = pytest.mark.gpu


def _net(sites: int, up: int, down: int, hidden: tuple[int, ...], dtype: torch.dtype, device: torch.device,
         seed: int = 0, scale: float = 0.3) -> mlp.WaveFunctionElectronUpDown:
    torch.manual_seed(seed)
    net = mlp.WaveFunctionElectronUpDown(double_sites=2 * sites, physical_dim=2, is_complex=True, spin_up=up, spin_down=down,
                                         hidden_size=hidden, ordering=+1)
    with torch.no_grad():
        for p in net.parameters():
            p.normal_(0.0, scale)
    return net.to(device=device, dtype=dtype)


def _dense(cfg_i: np.ndarray, cfg_j: np.ndarray, terms: tuple[np.ndarray, np.ndarray, np.ndarray]) -> np.ndarray:
    """H[j, i] between configs_j rows and configs_i rows, one oracle application per column."""
    out = np.zeros((len(cfg_j), len(cfg_i)), dtype=np.complex128)
    for i in range(len(cfg_i)):
        e = np.zeros(len(cfg_i), dtype=np.complex128)
        e[i] = 1.0
        out[:, i], _, _ = orc.apply_within(1, cfg_i, e, cfg_j, *terms)
    return out


def _grads(net: torch.nn.Module) -> list[torch.Tensor]:
    return [p.grad.detach().clone() if p.grad is not None else torch.zeros_like(p) for p in net.parameters()]


def _small_case(device: torch.device) -> tuple[Hamiltonian, tuple[np.ndarray, ...], np.ndarray, torch.Tensor]:
    terms = sorted_terms(*sy.molecular_terms(5, 7))
    cfg = sy.sector_configs(5, 2, 2, 100, 0)  # the whole 100-state sector
    ham = Hamiltonian(tuple(torch.from_numpy(np.ascontiguousarray(x)) for x in terms), kind="fermi")  # type: ignore[arg-type]
    return ham, terms, cfg, torch.from_numpy(cfg).to(device)


def test_energy_backward_matches_dense_autograd(cuda_device: torch.device) -> None:
    ham, terms, cfg, cfg_t = _small_case(cuda_device)
    dense = torch.from_numpy(_dense(cfg, cfg, terms)).to(cuda_device)
    net = _net(5, 2, 2, (8, 8, 8), torch.float64, cuda_device)
    psi = net(cfg_t)
    e_ref = (torch.vdot(psi, dense @ psi) / torch.vdot(psi, psi)).real
    e_ref.backward()
    want = _grads(net)
    net.zero_grad()
    e = vmc.energy_backward(net, vmc.EnergySet(ham, cfg_t), chunk=17)
    got = _grads(net)
    assert abs(e - e_ref.item()) < 1e-12 * max(1.0, abs(e_ref.item()))
    scale = max(float(g.abs().max()) for g in want if g.numel())
    for g, w in zip(got, want):
        torch.testing.assert_close(g, w, atol=1e-9 * scale, rtol=1e-9)


def test_closure_energy_matches_dense_formula(cuda_device: torch.device) -> None:
    ham, terms, cfg, cfg_t = _small_case(cuda_device)
    rows = np.sort(np.random.default_rng(1).choice(len(cfg), 40, replace=False))
    src, src_t = cfg[rows], cfg_t[torch.from_numpy(rows).to(cuda_device)]
    net = _net(5, 2, 2, (8,), torch.float64, cuda_device, seed=2)
    for dst, dst_t in ((src, src_t), (cfg, cfg_t)):
        h_sd = torch.from_numpy(_dense(dst, src, terms)).to(cuda_device)  # rows: src, columns: dst
        psi_src = net(src_t)
        with torch.no_grad():
            psi_dst = net(dst_t)
        e_ref = ((psi_src.conj() @ (h_sd @ psi_dst)) / (psi_src.conj() @ psi_src.detach())).real
        e_ref.backward()
        want = _grads(net)
        net.zero_grad()
        e = vmc.closure_energy(ham, net, src_t, dst_t)
        e.backward()
        got = _grads(net)
        net.zero_grad()
        assert abs(e.item() - e_ref.item()) < 1e-12 * max(1.0, abs(e_ref.item()))
        scale = max(float(g.abs().max()) for g in want if g.numel())
        for g, w in zip(got, want):
            torch.testing.assert_close(g, w, atol=1e-9 * scale, rtol=1e-9)
    # one set: the closure's gradient is half the Rayleigh quotient's (H Hermitian)
    e = vmc.closure_energy(ham, net, src_t, src_t)
    e.backward()
    half = _grads(net)
    net.zero_grad()
    vmc.energy_backward(net, vmc.EnergySet(ham, src_t), chunk=1000)
    full = _grads(net)
    scale = max(float(g.abs().max()) for g in full if g.numel())
    for h, f in zip(half, full):
        torch.testing.assert_close(2 * h, f, atol=1e-9 * scale, rtol=1e-9)


def test_energy_at_76_spin_orbitals_matches_oracle(cuda_device: torch.device) -> None:
    """The workload's Hamiltonian (38 orbitals, 13a+13b, |(pq|rs)| < 1e-8 dropped) on 64 configs
    within two excitations of the reference determinant, psi from a float32 network."""
    terms = sorted_terms(*sy.molecular_terms(38, 5001, threshold=1e-8))
    assert len(terms[0]) == 3_063_651
    pool = sy.excitation_configs(38, 13, 13, 5_000, 5002)  # a sample of the singles and doubles
    hf = np.zeros((1, 76), dtype=np.uint8)
    hf[0, 0:26:2] = 1
    hf[0, 1:26:2] = 1
    hf = sy.pack_occupation(hf)
    others = pool[~np.all(pool == hf, axis=1)]
    sub = sy.sort_unique_configs(np.concatenate([hf, others[np.random.default_rng(3).choice(len(others), 63, replace=False)]]))
    assert len(sub) == 64
    ham = Hamiltonian(tuple(torch.from_numpy(np.ascontiguousarray(x)) for x in terms), kind="fermi")  # type: ignore[arg-type]
    sub_t = torch.from_numpy(sub).to(cuda_device)
    net = _net(38, 13, 13, (64,), torch.float32, cuda_device, seed=4, scale=0.1)
    with torch.no_grad():
        psi = net(sub_t)
    energy_set = vmc.EnergySet(ham, sub_t, max_occupied=26)
    e_gpu = energy_set.energy(psi).item()
    hpsi = torch.view_as_complex(pp.apply(energy_set.terms, energy_set.targets, sub_t, torch.view_as_real(psi.contiguous()),
                                          pp.MODE_FORWARD))
    psi_np = psi.cpu().numpy()
    want, asum, _ = orc.apply_within(1, sub, psi_np, sub, *terms)
    # off-diagonal elements take part: the reference determinant reaches most other rows
    column, _, _ = orc.apply_within(1, hf, np.ones(1, dtype=np.complex128), sub, *terms)
    assert np.count_nonzero(column) > 32
    np.testing.assert_array_less(np.abs(hpsi.cpu().numpy() - want), 1e-12 * np.maximum(asum, 1e-300) + 1e-300)
    norm2 = float(np.vdot(psi_np, psi_np).real)
    e_ref = float(np.vdot(psi_np, want).real) / norm2
    bound = 1e-12 * float(np.abs(psi_np) @ asum) / norm2
    assert abs(e_gpu - e_ref) <= bound + 1e-14, (e_gpu, e_ref, bound)


@pytest.mark.parametrize("rayleigh", [True, False])
def test_vmc_cycle_lowers_energy(cuda_device: torch.device, rayleigh: bool) -> None:
    terms = sorted_terms(*sy.molecular_terms(6, 11))
    ham = Hamiltonian(tuple(torch.from_numpy(np.ascontiguousarray(x)) for x in terms), kind="fermi")  # type: ignore[arg-type]
    net = _net(6, 3, 3, (32,), torch.float32, cuda_device, seed=5, scale=0.2)
    opt = torch.optim.Adam(net.parameters(), lr=1e-2)
    res = vmc.vmc_cycle(ham, net, opt, sampling_count=150, local_step=40, relative_count=0 if rayleigh else 250,
                        rayleigh=rayleigh, seed=9)
    cfg = res.configs.cpu().numpy()
    assert len(np.unique(cfg, axis=0)) == len(cfg) <= 150
    assert res.energies[-1] < res.energies[0] - 1e-3
    if rayleigh:
        # a Rayleigh quotient on the sampled set never goes below the lowest eigenvalue of H there
        lowest = float(np.linalg.eigvalsh(_dense(cfg, cfg, terms)).min())
        assert min(res.energies) >= lowest - 1e-9
