"""
The autoregressive electron up/down network and its unique sampler (qmb/mlp.py:10-314): the
config producer that feeds configs_i / psi_i to the Hamiltonian operators.

Parameters, buffers and their names are the reference's (amplitude.{i}.model.{2l}.weight/bias,
phase.model.*, ordering, ordering_reversed, dummy_param), so a state_dict moves between the two.
What runs differs:
  - forward: the first layers of all S site networks are one GEMM.  Site i reads the first 2 i
    inputs, so its weight rows are padded with zeros to 2 S columns and the S blocks are stacked
    (an exact rewrite: the padding multiplies inputs site i never sees by zero).  The further layers
    of all sites are one batched GEMM each.  Everything is computed transposed ([features, batch]),
    so the HIP kernels that build the inputs from packed configs (qmb_network_inputs) and reduce the
    [S][4][B] logits to the log-amplitude (qmb_network_amplitude: particle-number masks,
    normalisation, selection, double-precision sum over sites) read and write coalesced rows.  The
    log-amplitude has a hand-written backward (the same kernel in gradient mode).
  - one hidden layer in float32 (width 64 to 512), without gradients: qmb_network_forward_fused
    runs the whole forward in one kernel.  With 0/1 inputs the first layer is a sum of weight
    columns over the set bits, so it is a few vector adds per row; the hidden activations stay in
    registers, the site weights in LDS, and nothing but the configs is read from HBM.  The sampler
    uses the same fusion per step (qmb_sampler_step_fused: site network plus children).
  - generate_unique (Gumbel-top-k without replacement, arXiv:2408.07625): prefixes live as two bits
    per site in uint64 words; per site one GEMM chain evaluates that site's network on the kept
    prefixes, one kernel (qmb_sampler_children) masks, normalises and perturbs the four children of
    every prefix with Philox4x32-10 uniforms, a device sort keeps the best `batch_size`, and one
    kernel (qmb_sampler_select) gathers the kept prefixes and log-probabilities.  The last kernel
    (qmb_sampler_prefix_configs) applies the site ordering and packs the configs.
The GEMMs run in PyTorch's BLAS (hipBLASLt).  The uniforms come from Philox under a seed drawn
from torch's host generator (the reference draws them with torch.rand_like), so the same seed
reproduces the same samples; tests inject uniforms to compare against the reference algorithm.
"""

from __future__ import annotations

import ctypes
import itertools
import typing

import torch

from . import _lib
from .ops import _ptr, _stream


class FakeLinear(torch.nn.Module):
    """A zero-input layer: its output is its bias (qmb/mlp.py:10-25)."""

    def __init__(self, dim_in: int, dim_out: int) -> None:
        super().__init__()
        assert dim_in == 0
        self.bias = torch.nn.Parameter(torch.zeros([dim_out]))

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        return self.bias.view([1, -1]).expand([x.shape[0], -1])


def select_linear_layer(dim_in: int, dim_out: int) -> torch.nn.Module:
    """qmb/mlp.py:28-35."""
    return FakeLinear(dim_in, dim_out) if dim_in == 0 else torch.nn.Linear(dim_in, dim_out)


class MLP(torch.nn.Module):
    """Linear layers with SiLU between them (qmb/mlp.py:38-58); `model` is the Sequential."""

    def __init__(self, dim_input: int, dim_output: int, hidden_size: tuple[int, ...]) -> None:
        super().__init__()
        self.dim_input = dim_input
        self.dim_output = dim_output
        self.hidden_size = tuple(hidden_size)
        dims = [dim_input, *hidden_size, dim_output]
        modules: list[torch.nn.Module] = []
        for a, b in itertools.pairwise(dims):
            modules += [select_linear_layer(a, b), torch.nn.SiLU()]
        self.model = torch.nn.Sequential(*modules[:-1])

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        return self.model(x)

    def layers(self) -> list[tuple[torch.Tensor | None, torch.Tensor]]:
        """(weight or None for a zero-input layer, bias) of every linear layer, in order."""
        return [(getattr(m, "weight", None), m.bias) for m in self.model if not isinstance(m, torch.nn.SiLU)]


def _dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return 0
    if dtype == torch.float64:
        return 1
    raise RuntimeError(f"the network kernels run in float32 or float64, not {dtype}")


def _transposed_chain(layers: list[tuple[torch.Tensor | None, torch.Tensor]], xt: torch.Tensor, n: int) -> torch.Tensor:
    """An MLP applied to column inputs xt [in, n]: returns [out, n]."""
    h = xt
    for idx, (w, b) in enumerate(layers):
        h = b[:, None].expand(-1, n) if w is None else torch.addmm(b[:, None], w, h)
        if idx < len(layers) - 1:
            h = torch.nn.functional.silu(h)
    return h


class _LogAmplitude(torch.autograd.Function):
    """sum over sites of the selected normalised logit (qmb_network_amplitude) with its backward."""

    @staticmethod
    def forward(ctx: torch.autograd.function.FunctionCtx, logits: torch.Tensor, configs: torch.Tensor,  # type: ignore[override]
                net: "WaveFunctionElectronUpDown") -> torch.Tensor:
        logits = logits.contiguous()
        n = configs.size(0)
        out = torch.empty(n, dtype=torch.float64, device=configs.device)
        with torch.cuda.device(configs.device):
            _lib.check(_lib.library().qmb_network_amplitude(
                _dtype_code(logits.dtype), n, net.sites, net.spin_up, net.spin_down, _ptr(configs), configs.size(1),
                _ptr(net.ordering_reversed), _ptr(logits), _ptr(out), None, None, _stream(configs.device)), "qmb_network_amplitude")
        ctx.save_for_backward(logits, configs)
        ctx.net = net  # type: ignore[attr-defined]
        return out

    @staticmethod
    def backward(ctx: torch.autograd.function.FunctionCtx, grad: torch.Tensor) -> tuple[object, ...]:  # type: ignore[override]
        logits, configs = ctx.saved_tensors  # type: ignore[attr-defined]
        net = ctx.net  # type: ignore[attr-defined]
        grad = grad.contiguous().to(torch.float64)
        grad_logits = torch.empty_like(logits)
        with torch.cuda.device(configs.device):
            _lib.check(_lib.library().qmb_network_amplitude(
                _dtype_code(logits.dtype), configs.size(0), net.sites, net.spin_up, net.spin_down, _ptr(configs), configs.size(1),
                _ptr(net.ordering_reversed), _ptr(logits), None, _ptr(grad), _ptr(grad_logits), _stream(configs.device)),
                "qmb_network_amplitude")
        return grad_logits, None, None


class WaveFunctionElectronUpDown(torch.nn.Module):
    """Autoregressive wave function conserving the up and down electron counts
    (qmb/mlp.py:61-314): psi(x) = exp(sum_i log-amplitude_i + i phase(x))."""

    # pylint: disable=too-many-instance-attributes

    def __init__(self, *, double_sites: int, physical_dim: int, is_complex: bool, spin_up: int, spin_down: int,
                 hidden_size: tuple[int, ...], ordering: int | list[int]) -> None:
        super().__init__()
        assert double_sites % 2 == 0
        assert physical_dim == 2
        assert is_complex
        self.double_sites = double_sites
        self.sites = double_sites // 2
        self.spin_up = spin_up
        self.spin_down = spin_down
        self.hidden_size = tuple(hidden_size)
        self.amplitude = torch.nn.ModuleList([MLP(i * 2, 4, self.hidden_size) for i in range(self.sites)])
        self.phase = MLP(self.double_sites, 1, self.hidden_size)
        if isinstance(ordering, int) and ordering == +1:
            ordering = list(range(self.sites))
        if isinstance(ordering, int) and ordering == -1:
            ordering = list(reversed(range(self.sites)))
        order = torch.tensor(ordering, dtype=torch.int64)
        if sorted(order.tolist()) != list(range(self.sites)):
            raise ValueError("ordering must be a permutation of the sites")
        self.ordering: torch.Tensor
        self.register_buffer("ordering", order)
        self.ordering_reversed: torch.Tensor
        self.register_buffer("ordering_reversed", torch.scatter(torch.zeros(self.sites, dtype=torch.int64), 0, order,
                                                                torch.arange(self.sites, dtype=torch.int64)))
        self.dummy_param = torch.nn.Parameter(torch.empty(0))
        # None: use the fused kernels whenever they apply; False: never (the GEMM path)
        self.fused: bool | None = None
        self._fused_cache: tuple[tuple[object, ...], tuple[torch.Tensor, ...]] | None = None

    # ------------------------------------------------------------------------------ fused path
    def _fused_ok(self) -> bool:
        """The fused float32 kernels cover one hidden layer of width 64, 128, 256 or 512 whose
        first-layer columns fit LDS."""
        if self.fused is False or self.dummy_param.dtype != torch.float32 or len(self.hidden_size) != 1:
            return False
        return bool(_lib.library().qmb_network_fused_supported(self.sites, self.hidden_size[0]))

    def _fused_weights(self) -> tuple[torch.Tensor, ...]:
        """(w1, b1, w2, b2, wp1, bp1, wp2, bp2) in the fused kernels' layout, rebuilt when a
        parameter changes."""
        params = list(self.parameters())
        key = tuple((p.data_ptr(), p._version) for p in params)  # pylint: disable=protected-access
        if self._fused_cache is not None and self._fused_cache[0] == key:
            return self._fused_cache[1]
        with torch.no_grad():
            per = [m.layers() for m in self.amplitude]
            dev = self.dummy_param.device
            h = self.hidden_size[0]
            w1 = torch.cat([per[i][0][0].t() for i in range(1, self.sites)]) if self.sites > 1 else torch.zeros((1, h), device=dev)
            b1 = torch.stack([layers[0][1] for layers in per])
            w2 = torch.stack([layers[1][0] for layers in per])
            b2 = torch.stack([layers[1][1] for layers in per])
            (pw1, pb1), (pw2, pb2) = self.phase.layers()
            assert pw1 is not None and pw2 is not None
            packed = tuple(t.detach().float().contiguous() for t in (w1, b1, w2, b2, pw1.t(), pb1, pw2.reshape(-1), pb2))
        self._fused_cache = (key, packed)
        return packed

    def _forward_fused(self, x: torch.Tensor) -> torch.Tensor:
        n = x.size(0)
        amplitude = torch.empty(n, dtype=torch.float64, device=x.device)
        phase = torch.empty(n, dtype=torch.float32, device=x.device)
        w = self._fused_weights()
        with torch.cuda.device(x.device):
            _lib.check(_lib.library().qmb_network_forward_fused(
                n, self.sites, self.hidden_size[0], self.spin_up, self.spin_down, _ptr(x), x.size(1), _ptr(self.ordering_reversed),
                *[_ptr(t) for t in w], _ptr(amplitude), _ptr(phase), _stream(x.device)), "qmb_network_forward_fused")
        return torch.view_as_complex(torch.stack([amplitude, phase.double()], dim=-1)).exp()

    # ------------------------------------------------------------------------------ network
    def _stacked(self) -> tuple[torch.Tensor, torch.Tensor, list[tuple[torch.Tensor, torch.Tensor]]]:
        """First layers of all sites as one [S h1, 2 S] matrix (site i's columns >= 2 i are zero)
        and every further layer as [S, out, in] (differentiable views of the parameters)."""
        per_site = [m.layers() for m in self.amplitude]
        w1 = []
        for i, layers in enumerate(per_site):
            w, b = layers[0]
            if w is None:
                w = torch.zeros((b.numel(), self.double_sites), dtype=b.dtype, device=b.device)
            else:
                w = torch.nn.functional.pad(w, (0, self.double_sites - 2 * i))
            w1.append(w)
        first_w = torch.cat(w1, dim=0)
        first_b = torch.cat([layers[0][1] for layers in per_site])
        rest = [(torch.stack([layers[l][0] for layers in per_site]), torch.stack([layers[l][1] for layers in per_site]))
                for l in range(1, len(per_site[0]))]
        return first_w, first_b, rest

    def _amplitude_logits(self, xt: torch.Tensor) -> torch.Tensor:
        """[S, 4, B] logits of every site network on the inputs xt [2 S, B]."""
        n = xt.size(1)
        first_w, first_b, rest = self._stacked()
        h = torch.addmm(first_b[:, None], first_w, xt)
        if not rest:
            return h.view(self.sites, 4, n)
        h = torch.nn.functional.silu(h).view(self.sites, -1, n)
        for idx, (w, b) in enumerate(rest):
            h = torch.baddbmm(b[:, :, None], w, h)
            if idx < len(rest) - 1:
                h = torch.nn.functional.silu(h)
        return h

    def _inputs(self, x: torch.Tensor) -> torch.Tensor:
        """[2 S, B] network inputs of packed configs (sampling order)."""
        dtype = self.dummy_param.dtype
        n = x.size(0)
        xt = torch.empty((self.double_sites, n), dtype=dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.library().qmb_network_inputs(_dtype_code(dtype), n, self.sites, _ptr(x), x.size(1), _ptr(self.ordering_reversed),
                                                         _ptr(xt), _stream(x.device)), "qmb_network_inputs")
        return xt

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """psi of packed configs x (uint8 [B, >= ceil(2 S / 8)]), complex128 (qmb/mlp.py:169-205)."""
        if x.dtype != torch.uint8 or x.dim() != 2 or x.size(1) * 8 < self.double_sites or not x.is_cuda:
            raise RuntimeError("x must be a uint8 [batch, n_qubytes] tensor on a HIP device")
        x = x.contiguous()
        needs_grad = torch.is_grad_enabled() and any(p.requires_grad for p in self.parameters())
        if not needs_grad and self._fused_ok():
            return self._forward_fused(x)
        n = x.size(0)
        xt = self._inputs(x)
        amplitude = _LogAmplitude.apply(self._amplitude_logits(xt), x, self)
        phase = _transposed_chain(self.phase.layers(), xt, n).view(n)
        return torch.view_as_complex(torch.stack([amplitude, phase.double()], dim=-1)).exp()

    # ------------------------------------------------------------------------------ sampler
    def _site_logits(self, site: int, xt: torch.Tensor, n: int, block_num: int) -> torch.Tensor:
        layers = self.amplitude[site].layers()
        if block_num <= 1 or n <= 1:
            return _transposed_chain(layers, xt, n).contiguous()
        parts = []
        for cols in torch.tensor_split(torch.arange(n), block_num):
            if cols.numel():
                lo, hi = int(cols[0]), int(cols[-1]) + 1
                parts.append(_transposed_chain(layers, xt[:, lo:hi], hi - lo))
        return torch.cat(parts, dim=1).contiguous()

    def generate_unique(self, batch_size: int, block_num: int = 1, *, seed: int | None = None,
                        uniform: typing.Callable[[int, int], torch.Tensor] | None = None
                        ) -> tuple[torch.Tensor, torch.Tensor, None, None]:
        """Up to batch_size distinct configs drawn without replacement, and their psi
        (qmb/mlp.py:207-314).  `uniform(site, n)` may supply the [4, n] uniforms of a step."""
        device = self.dummy_param.device
        dtype = self.dummy_param.dtype
        code = _dtype_code(dtype)
        if device.type != "cuda":
            raise RuntimeError("generate_unique runs on a HIP device")
        if seed is None:
            seed = int(torch.randint(0, 2**62, (1,)).item())
        lib = _lib.library()
        stream = _stream(device)
        words = (self.double_sites + 63) // 64
        prefix = torch.zeros((1, words), dtype=torch.int64, device=device)
        unperturbed = torch.zeros(1, dtype=dtype, device=device)
        perturbed = torch.zeros(1, dtype=dtype, device=device)
        fused = self._fused_ok()
        with torch.cuda.device(device), torch.no_grad():
            fw = self._fused_weights() if fused else None
            for site in range(self.sites):
                n = prefix.size(0)
                u = None
                if uniform is not None:
                    u = uniform(site, n).to(device=device, dtype=dtype).contiguous()
                    if u.shape != (4, n):
                        raise RuntimeError("uniform(site, n) must return a [4, n] tensor")
                child_unperturbed = torch.empty(4 * n, dtype=dtype, device=device)
                child_perturbed = torch.empty(4 * n, dtype=dtype, device=device)
                if fw is not None:
                    # the site network and the children in one kernel
                    _lib.check(lib.qmb_sampler_step_fused(n, site, self.sites, self.hidden_size[0], self.spin_up, self.spin_down, _ptr(prefix),
                                                          words, *[_ptr(t) for t in fw[:4]], _ptr(unperturbed), _ptr(perturbed), _ptr(u),
                                                          ctypes.c_uint64(seed), _ptr(child_unperturbed), _ptr(child_perturbed), stream),
                               "qmb_sampler_step_fused")
                else:
                    xt = torch.empty((2 * site, n), dtype=dtype, device=device)
                    _lib.check(lib.qmb_sampler_prefix_inputs(code, n, site, _ptr(prefix), words, _ptr(xt), stream),
                               "qmb_sampler_prefix_inputs")
                    logits = self._site_logits(site, xt, n, block_num)
                    _lib.check(lib.qmb_sampler_children(code, n, site, self.sites, self.spin_up, self.spin_down, _ptr(prefix), words,
                                                        _ptr(logits), _ptr(unperturbed), _ptr(perturbed), _ptr(u), ctypes.c_uint64(seed),
                                                        _ptr(child_unperturbed), _ptr(child_perturbed), stream), "qmb_sampler_children")
                values, index = torch.sort(child_perturbed, descending=True)
                values, index = values[:batch_size], index[:batch_size]
                index = index[torch.isfinite(values)].contiguous()
                n_sel = index.numel()
                new_prefix = torch.empty((n_sel, words), dtype=torch.int64, device=device)
                unperturbed = torch.empty(n_sel, dtype=dtype, device=device)
                perturbed = torch.empty(n_sel, dtype=dtype, device=device)
                _lib.check(lib.qmb_sampler_select(code, n_sel, _ptr(index), n, site, _ptr(prefix), words, _ptr(child_unperturbed),
                                                  _ptr(child_perturbed), _ptr(new_prefix), _ptr(unperturbed), _ptr(perturbed), stream),
                           "qmb_sampler_select")
                prefix = new_prefix
            n = prefix.size(0)
            n_qubytes = (self.double_sites + 7) // 8
            configs = torch.empty((n, n_qubytes), dtype=torch.uint8, device=device)
            _lib.check(lib.qmb_sampler_prefix_configs(n, self.sites, _ptr(prefix), words, _ptr(self.ordering), n_qubytes, _ptr(configs),
                                                      stream), "qmb_sampler_prefix_configs")
        parts = [self(configs[rows]) for rows in torch.tensor_split(torch.arange(n, device=device), max(block_num, 1)) if rows.numel()]
        amplitude = torch.cat(parts) if parts else torch.empty(0, dtype=torch.complex128, device=device)
        return configs, amplitude, None, None
