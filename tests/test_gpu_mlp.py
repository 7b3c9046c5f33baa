"""
GPU checks of the config producer (qmb_hip.mlp, qmb/mlp.py:61-314) against a plain PyTorch
statement of the reference algorithm run on the same parameters (the per-site loop of networks,
masks, normalisation and Gumbel-top-k of qmb/mlp.py:169-314):

  - forward psi in float64 within 1e-12 (float32 within 1e-5), on every config of a sector plus
    configs outside it (psi = 0 for both), for natural, reversed and permuted site orderings and
    with no hidden layer;
  - the gradient of a real loss of psi with respect to every parameter within 1e-10 of autograd
    through the PyTorch statement (the hand-written backward of the log-amplitude);
  - generate_unique with injected uniforms returns the very configs, in the very order, of the
    PyTorch statement with the same uniforms, and the same psi;
  - with Philox uniforms: distinct configs with the right electron counts, at most batch_size of
    them, the whole sector once batch_size reaches its size, the same configs for the same seed;
  - a 52-spin-orbital network (26 sites, 7 + 7 electrons) samples 20,000 distinct valid configs;
  - the fused float32 kernels (one hidden layer of 64 to 512, no gradients): forward within 2e-5
    of the statement and of the GEMM path, sampler with injected uniforms returning the
    statement's configs, and fused against GEMM forward at the 26-site, 512-wide shape.
"""

from __future__ import annotations

import itertools
import math

import numpy as np
import pytest
import torch

from qmb_hip import bitspack
from qmb_hip import mlp

pytestmark = pytest.mark.gpu


def _net(sites: int, up: int, down: int, hidden: tuple[int, ...], ordering: object, dtype: torch.dtype, device: torch.device,
         seed: int = 0) -> mlp.WaveFunctionElectronUpDown:
    torch.manual_seed(seed)
    net = mlp.WaveFunctionElectronUpDown(double_sites=2 * sites, physical_dim=2, is_complex=True, spin_up=up, spin_down=down,
                                         hidden_size=hidden, ordering=ordering)
    with torch.no_grad():
        for p in net.parameters():
            p.normal_(0.0, 0.5)  # larger than the default init, so the masks and normalisation matter
    return net.to(device=device, dtype=dtype)


def _ref_mask(net: mlp.WaveFunctionElectronUpDown, x: torch.Tensor) -> torch.Tensor:
    """qmb/mlp.py:113-154 as written there."""
    current = x.shape[1]
    number = x.sum(dim=1)
    up_e, down_e = number[:, 0], number[:, 1]
    add_up = torch.stack([current - up_e < net.sites - net.spin_up, up_e < net.spin_up], dim=-1).unsqueeze(-1)
    add_down = torch.stack([current - down_e < net.sites - net.spin_down, down_e < net.spin_down], dim=-1).unsqueeze(-2)
    return torch.logical_and(add_up, add_down)


def _ref_normalize(x: torch.Tensor) -> torch.Tensor:
    return x - ((2 * x).exp().sum(dim=[-2, -1]).log() / 2).unsqueeze(-1).unsqueeze(-1)


def _ref_forward(net: mlp.WaveFunctionElectronUpDown, x: torch.Tensor) -> torch.Tensor:
    """qmb/mlp.py:169-205 with torch modules, site by site."""
    b = x.shape[0]
    xs = bitspack.unpack_int(x, size=1, last_dim=net.double_sites).view([b, net.sites, 2])
    xs = torch.index_select(xs, 1, net.ordering_reversed)
    xf = xs.to(net.dummy_param.dtype)
    arange = torch.arange(b, device=x.device)
    total = torch.zeros(b, dtype=torch.float64, device=x.device)
    for i, m in enumerate(net.amplitude):
        d = m(xf[:, :i].reshape([b, 2 * i])).view([b, 2, 2])
        d = d + torch.where(_ref_mask(net, xs[:, :i]), 0, -torch.inf)
        d = _ref_normalize(d)
        xi = xs[:, i, :].to(torch.int64)
        total = total + d[arange, xi[:, 0], xi[:, 1]].double()
    phase = net.phase(xf.view([b, net.double_sites])).view([b])
    return torch.view_as_complex(torch.stack([total, phase.double()], dim=-1)).exp()


def _ref_generate(net: mlp.WaveFunctionElectronUpDown, batch: int, uniforms: list[torch.Tensor]) -> torch.Tensor:
    """qmb/mlp.py:207-295 with torch.rand_like replaced by the given [n, 4] uniforms."""
    device, dtype = net.dummy_param.device, net.dummy_param.dtype
    x = torch.empty([1, 0, 2], device=device, dtype=torch.uint8)
    unp = torch.tensor([0], dtype=dtype, device=device)
    per = torch.tensor([0], dtype=dtype, device=device)
    for i, m in enumerate(net.amplitude):
        n = x.shape[0]
        d = m(x.to(dtype).view([n, 2 * i])).view([n, 2, 2])
        d = _ref_normalize(d + torch.where(_ref_mask(net, x), 0, -torch.inf))
        l = unp.view([n, 1]) + (2 * d).view([n, 4])
        big_l = l - (-uniforms[i].log()).log()
        z = big_l.max(dim=-1).values.view([n, 1])
        tilde = -torch.log(torch.exp(-per.view([n, 1])) - torch.exp(-z) + torch.exp(-big_l))
        kids = [torch.cat([x, torch.tensor([[a, c]], device=device, dtype=torch.uint8).expand(n, -1, -1)], dim=1)
                for a, c in ((0, 0), (0, 1), (1, 0), (1, 1))]
        x = torch.cat(kids)
        unp = l.permute(1, 0).reshape([4 * n])
        per = tilde.permute(1, 0).reshape([4 * n])
        sel = per.argsort(descending=True)[:batch]
        x, unp, per = x[sel], unp[sel], per[sel]
        keep = per.isfinite()
        x, unp, per = x[keep], unp[keep], per[keep]
    x = torch.index_select(x, 1, net.ordering).reshape([x.size(0), net.double_sites])
    return bitspack.pack_int(x, size=1)


def _sector(sites: int, up: int, down: int) -> np.ndarray:
    rows = []
    for ua in itertools.combinations(range(sites), up):
        for da in itertools.combinations(range(sites), down):
            bits = np.zeros(2 * sites, dtype=np.uint8)
            bits[[2 * s for s in ua]] = 1
            bits[[2 * s + 1 for s in da]] = 1
            rows.append(np.packbits(bits, bitorder="little"))
    return np.array(rows, dtype=np.uint8)


ORDERINGS = [+1, -1, [2, 0, 5, 1, 4, 3]]


@pytest.mark.parametrize("ordering", ORDERINGS, ids=["natural", "reversed", "permuted"])
@pytest.mark.parametrize("hidden", [(16, 8), ()], ids=["two-hidden", "no-hidden"])
def test_forward_matches_reference_algorithm(cuda_device: torch.device, ordering: object, hidden: tuple[int, ...]) -> None:
    net = _net(6, 3, 2, hidden, ordering, torch.float64, cuda_device)
    sector = _sector(6, 3, 2)
    rng = np.random.default_rng(5)
    outside = rng.integers(0, 256, size=(40, 2), dtype=np.uint8)
    outside[:, 1] &= 0x0F  # 12 bits
    x = torch.from_numpy(np.concatenate([sector, outside])).to(cuda_device)
    got = net(x)
    want = _ref_forward(net, x)
    torch.testing.assert_close(got, want, rtol=1e-12, atol=1e-300)
    assert torch.all(got[: len(sector)].abs() > 0)
    # float32 network: same within single-precision rounding
    net32 = _net(6, 3, 2, hidden, ordering, torch.float32, cuda_device)
    torch.testing.assert_close(net32(x), _ref_forward(net32, x), rtol=1e-5, atol=1e-7)


def test_backward_matches_autograd(cuda_device: torch.device) -> None:
    net = _net(6, 3, 2, (16, 8), [2, 0, 5, 1, 4, 3], torch.float64, cuda_device, seed=3)
    x = torch.from_numpy(_sector(6, 3, 2)).to(cuda_device)
    w = torch.linspace(0.5, 1.5, x.size(0), device=cuda_device, dtype=torch.float64)
    grads = []
    for fn in (net, lambda t: _ref_forward(net, t)):
        net.zero_grad()
        psi = fn(x)
        loss = (w * psi.abs() ** 2).sum() + psi.real.sum() - 0.3 * psi.imag.sum()
        loss.backward()
        grads.append({k: p.grad.detach().clone() for k, p in net.named_parameters() if p.grad is not None})
    assert grads[0].keys() == grads[1].keys() and len(grads[0]) > 10
    for k in grads[0]:
        torch.testing.assert_close(grads[0][k], grads[1][k], rtol=1e-10, atol=1e-12, msg=k)


@pytest.mark.parametrize("ordering", ORDERINGS, ids=["natural", "reversed", "permuted"])
@pytest.mark.parametrize("batch", [7, 64, 1000])
def test_generate_unique_matches_reference_algorithm(cuda_device: torch.device, ordering: object, batch: int) -> None:
    net = _net(6, 3, 2, (16,), ordering, torch.float64, cuda_device, seed=11)
    gen = torch.Generator(device=cuda_device)
    gen.manual_seed(batch)
    uniforms: list[torch.Tensor] = []

    def uniform(site: int, n: int) -> torch.Tensor:
        u = torch.rand((n, 4), dtype=torch.float64, device=cuda_device, generator=gen).clamp_min(1e-300)
        uniforms.append(u)
        return u.t().contiguous()

    configs, psi, _, _ = net.generate_unique(batch, uniform=uniform)
    want = _ref_generate(net, batch, uniforms)
    assert torch.equal(configs, want)
    torch.testing.assert_close(psi, _ref_forward(net, want), rtol=1e-12, atol=1e-300)


def test_generate_unique_properties(cuda_device: torch.device) -> None:
    net = _net(6, 3, 2, (16,), [2, 0, 5, 1, 4, 3], torch.float64, cuda_device, seed=13)
    sector = {bytes(r) for r in _sector(6, 3, 2)}
    for batch in (1, 50, 299, 300, 5000):
        configs, psi, _, _ = net.generate_unique(batch, seed=1234)
        rows = [bytes(r) for r in configs.cpu().numpy()]
        assert len(rows) == len(set(rows)) == min(batch, len(sector))
        assert set(rows) <= sector
        assert torch.all(psi.abs() > 0)
    again, _, _, _ = net.generate_unique(50, seed=1234)
    first, _, _, _ = net.generate_unique(50, seed=1234)
    assert torch.equal(again, first)
    other, _, _, _ = net.generate_unique(50, seed=99)
    assert not torch.equal(other, first)
    # block_num only splits the work
    blocked, _, _, _ = net.generate_unique(50, block_num=3, seed=1234)
    assert torch.equal(blocked, first)


def test_molecular_scale_sampling(cuda_device: torch.device) -> None:
    net = _net(26, 7, 7, (64,), +1, torch.float32, cuda_device, seed=17)
    configs, psi, _, _ = net.generate_unique(20_000, seed=7)
    assert configs.shape == (20_000, 7)
    bits = bitspack.unpack_int(configs, size=1, last_dim=52).view(-1, 26, 2).to(torch.int64)
    assert torch.all(bits[:, :, 0].sum(dim=1) == 7) and torch.all(bits[:, :, 1].sum(dim=1) == 7)
    assert torch.unique(configs, dim=0).size(0) == 20_000
    assert torch.all(torch.isfinite(psi.real))
    assert math.comb(26, 7) ** 2 > 20_000


@pytest.mark.parametrize("hidden", [64, 128, 256, 512])
def test_fused_forward_matches_reference_algorithm(cuda_device: torch.device, hidden: int) -> None:
    """One hidden layer in float32 without gradients runs the fused kernel."""
    net = _net(6, 3, 2, (hidden,), [2, 0, 5, 1, 4, 3], torch.float32, cuda_device, seed=hidden)
    assert net._fused_ok()  # pylint: disable=protected-access
    sector = _sector(6, 3, 2)
    rng = np.random.default_rng(hidden)
    outside = rng.integers(0, 256, size=(40, 2), dtype=np.uint8)
    outside[:, 1] &= 0x0F
    x = torch.from_numpy(np.concatenate([sector, outside])).to(cuda_device)
    with torch.no_grad():
        got = net(x)
        net.fused = False
        plain = net(x)
        net.fused = None
        want = _ref_forward(net, x)
    torch.testing.assert_close(got, want, rtol=2e-5, atol=1e-30)
    torch.testing.assert_close(plain, want, rtol=2e-5, atol=1e-30)
    assert torch.all(got[len(sector):][want[len(sector):] == 0] == 0)


def test_fused_forward_molecular_shape(cuda_device: torch.device) -> None:
    net = _net(26, 7, 7, (512,), -1, torch.float32, cuda_device, seed=21)
    with torch.no_grad():
        for p in net.parameters():
            p.mul_(0.2)  # keep 26-site sums of logits in a moderate range
        configs, _, _, _ = net.generate_unique(3000, seed=3)
        got = net(configs)
        net.fused = False
        plain = net(configs)
        net.fused = None
    torch.testing.assert_close(got, plain, rtol=1e-4, atol=1e-30)


@pytest.mark.parametrize("batch", [7, 64, 1000])
def test_fused_sampler_matches_reference_algorithm(cuda_device: torch.device, batch: int) -> None:
    net = _net(6, 3, 2, (64,), [2, 0, 5, 1, 4, 3], torch.float32, cuda_device, seed=31)
    gen = torch.Generator(device=cuda_device)
    gen.manual_seed(7 + batch)
    uniforms: list[torch.Tensor] = []

    def uniform(site: int, n: int) -> torch.Tensor:
        u = torch.rand((n, 4), dtype=torch.float32, device=cuda_device, generator=gen).clamp_min(1e-30)
        uniforms.append(u)
        return u.t().contiguous()

    configs, psi, _, _ = net.generate_unique(batch, uniform=uniform)
    with torch.no_grad():
        want = _ref_generate(net, batch, uniforms)
    assert torch.equal(configs, want)
