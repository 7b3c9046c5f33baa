"""
PyTorch-ROCm operator layer over libqmb_hip.so.

Registers, per (n_qubytes, particle_cut), the reference's op namespace and schemas
(qmb/_hamiltonian.cpp:80-92):

    torch.ops.qmb_hamiltonian_{nq}_{pc}.apply_within(configs_i, psi_i, configs_j, site, kind, coef)
    torch.ops.qmb_hamiltonian_{nq}_{pc}.find_relative(configs_i, psi_i, count_selected, site, kind, coef, configs_exclude)
    torch.ops.qmb_hamiltonian_{nq}_{pc}.diagonal_term(configs, site, kind, coef)
    torch.ops.qmb_hamiltonian_{nq}_{pc}.single_relative(configs, site, kind, coef)

with implementations on the GPU dispatch key (HIP tensors report device type "cuda" in
PyTorch-ROCm) and, for apply_within, its adjoint and find_relative, on the CPU key (the library's
host operators, as the reference registers those two for CPU).  Extra ops with no reference counterpart live in the same namespace:
apply_within_adjoint, apply_within_plain, single_relative_keyed.  apply_within is differentiable
with respect to psi_i through apply_within_adjoint.

Each impl validates its arguments with the reference's messages (qmb/_hamiltonian_cuda.cu:186-232),
takes scratch memory from PyTorch's caching allocator and enqueues the C ABI call on the current
stream.  Like the reference, find_relative synchronises to learn its output size; diagonal_term and
single_relative do not.

apply_within (and its adjoint) takes the structured path when the table has pair excitations (or
hops between sites of one parity, without pair excitations):
the term table is compiled once per (site, kind, coef) tensors and kept (a small cache keyed by
the tensors and their versions), and the pair, single and diagonal kernels apply it within one
config set -- configs_i itself when configs_j is the very same tensor, else configs_j joined with
the rows of configs_i it lacks (prepared.apply_across).  That path synchronises to size the set's
core index.  Calls it does not cover (repeated target rows, no pair excitations, nq > 16) run the
one-shot enumeration kernel (qmb_apply_within), as stateless as the reference's op.
"""

from __future__ import annotations

import collections
import ctypes
import functools
import threading

import torch

from . import _lib

_REGISTRY: dict[str, torch.library.Library] = {}
_REGISTRY_LOCK = threading.Lock()
_VALIDATED: dict[tuple[int, int, int, int, int, int], bool] = {}
_PREPARED: "collections.OrderedDict[tuple[object, ...], object]" = collections.OrderedDict()
_PREPARED_MAX = 4
_PREPARED_LOCK = threading.Lock()


def _ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    if t is None or t.numel() == 0:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())


def _stream(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def _size_query(fn: object, *args: object) -> int:
    out = ctypes.c_size_t(0)
    _lib.check(fn(*args, ctypes.byref(out)), getattr(fn, "__name__", "workspace size"))
    return int(out.value)


def _check(cond: bool, message: str) -> None:
    if not cond:
        raise RuntimeError(message)


def _check_tensor(t: torch.Tensor, name: str, device: torch.device, dtype: torch.dtype, cols: int | None,
                  rows: int | None = None, rows_name: str = "batch size") -> None:
    """Checks of qmb/_hamiltonian_cuda.cu:186-232 (CPU: qmb/_hamiltonian_cpu.cpp:178-224): device,
    contiguity, dtype, rank, widths."""
    if device.type == "cpu":
        _check(t.device.type == "cpu", f"{name} must be on CPU.")
    else:
        _check(t.device.type == "cuda", f"{name} must be on CUDA.")
    _check(t.device == device, f"{name} must be on the same device as others.")
    _check(t.is_contiguous(), f"{name} must be contiguous.")
    _check(t.dtype == dtype, f"{name} must be {str(dtype).replace('torch.', '')}.")
    _check(t.dim() == 2, f"{name} must be 2D.")
    if rows is not None:
        _check(t.size(0) == rows, f"{name} {rows_name} must match the provided {rows_name.replace(' ', '_')}.")
    if cols is not None:
        _check(t.size(1) == cols, f"{name} must have {cols} columns.")


def _check_terms(site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor, device: torch.device, n_qubytes: int) -> int:
    n_terms = site.size(0) if site.dim() == 2 else -1
    _check_tensor(site, "site", device, torch.int16, 4, n_terms, "size")
    _check_tensor(kind, "kind", device, torch.uint8, 4, n_terms, "size")
    _check_tensor(coef, "coef", device, torch.float64, 2, n_terms, "size")
    key = (site.data_ptr(), site._version, kind.data_ptr(), kind._version, n_terms, n_qubytes)
    if key not in _VALIDATED:
        active = kind != 2
        bad = (kind > 2) | (active & ((site < 0) | (site >= 8 * n_qubytes)))
        n_bad = int(bad.any(dim=1).sum().item()) if n_terms > 0 else 0
        _check(n_bad == 0, f"{n_bad} terms have a kind outside {{0,1,2}} or a site outside [0, {8 * n_qubytes}).")
        if len(_VALIDATED) > 4096:
            _VALIDATED.clear()
        _VALIDATED[key] = True
    return n_terms


# ----------------------------------------------------------------------------------- impls -----
def _prepared_terms(nq: int, pc: int, site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor) -> object:
    """The compiled table of these term tensors, kept while they are (the cache holds them, so their
    storage cannot be reused by other tensors meanwhile); an in-place change bumps a version and
    rebuilds."""
    from . import prepared as pp  # pylint: disable=import-outside-toplevel

    key = (str(site.device), nq, pc, site.data_ptr(), kind.data_ptr(), coef.data_ptr(), site.size(0), site._version, kind._version,
           coef._version)  # pylint: disable=protected-access
    with _PREPARED_LOCK:
        terms = _PREPARED.get(key)
        if terms is not None:
            _PREPARED.move_to_end(key)
            return terms
    terms = pp.PreparedTerms(site, kind, coef, nq, pc)
    with _PREPARED_LOCK:
        _PREPARED[key] = terms
        while len(_PREPARED) > _PREPARED_MAX:
            _PREPARED.popitem(last=False)
    return terms


def _within_structured(nq: int, pc: int, configs: torch.Tensor, vec: torch.Tensor, site: torch.Tensor, kind: torch.Tensor,
                       coef: torch.Tensor, mode: int) -> torch.Tensor | None:
    """apply_within (mode 0) or its adjoint (mode 1) within one config set through the structured
    path (pair-excitation core groups, or the hop path for tables of hops and diagonal terms), or
    None when neither applies (no such terms, repeated rows, nq > 16)."""
    from . import prepared as pp  # pylint: disable=import-outside-toplevel

    if configs.size(0) == 0 or nq > pp.PAIR_MAX_QUBYTES:
        return None
    terms = _prepared_terms(nq, pc, site, kind, coef)
    if terms.pairs_desc is None and not terms.hops:  # type: ignore[attr-defined]
        return None
    targets = pp.PreparedTargets(configs, nq)
    if not (pp.can_use_pairs(terms, targets, configs) or pp.can_use_hops(terms, targets, configs)):  # type: ignore[arg-type]
        return None
    return pp.apply(terms, targets, configs, vec, mode)  # type: ignore[arg-type]


def _same_set(configs_i: torch.Tensor, configs_j: torch.Tensor) -> bool:
    return configs_i.data_ptr() == configs_j.data_ptr() and configs_i.shape == configs_j.shape and configs_i.is_contiguous()


def _structured(nq: int, pc: int, configs_i: torch.Tensor, vec: torch.Tensor, configs_j: torch.Tensor, site: torch.Tensor,
                kind: torch.Tensor, coef: torch.Tensor, mode: int) -> torch.Tensor | None:
    """apply_within (mode 0) or its adjoint (mode 1) through the structured path: within one set
    when configs_j is configs_i, else across the two sets; None when it does not apply."""
    from . import prepared as pp  # pylint: disable=import-outside-toplevel

    if _same_set(configs_i, configs_j):
        return _within_structured(nq, pc, configs_i, vec.contiguous(), site, kind, coef, mode)
    if configs_i.size(0) == 0 or configs_j.size(0) == 0 or nq > pp.PAIR_MAX_QUBYTES:
        return None
    terms = _prepared_terms(nq, pc, site, kind, coef)
    return pp.apply_across(terms, configs_i, vec, configs_j, mode)  # type: ignore[arg-type]


def _apply_within(nq: int, pc: int, configs_i: torch.Tensor, psi_i: torch.Tensor, configs_j: torch.Tensor,
                  site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor) -> torch.Tensor:
    device = configs_i.device
    _check(device.type == "cuda", "configs must be on CUDA.")
    n_i = configs_i.size(0)
    n_j = configs_j.size(0)
    _check_tensor(configs_i, "configs", device, torch.uint8, nq, n_i)
    _check_tensor(psi_i, "psi", device, torch.float64, 2, n_i)
    _check_tensor(configs_j, "result_configs", device, torch.uint8, nq, n_j)
    n_terms = _check_terms(site, kind, coef, device, nq)
    out_s = _structured(nq, pc, configs_i, psi_i, configs_j, site, kind, coef, 0)
    if out_s is not None:
        return out_s
    lib = _lib.library()
    with torch.cuda.device(device):
        ws_bytes = _size_query(lib.qmb_apply_within_workspace_size, nq, n_i, n_j, n_terms)
        ws = _workspace(ws_bytes, device)
        out = torch.empty((n_j, 2), dtype=torch.float64, device=device)
        _lib.check(
            lib.qmb_apply_within(nq, pc, _ptr(configs_i), _ptr(psi_i), n_i, _ptr(configs_j), n_j, _ptr(site), _ptr(kind),
                                 _ptr(coef), n_terms, _ptr(out), _ptr(ws), ws.numel(), _stream(device)),
            "qmb_apply_within",
        )
    return out


def _apply_within_adjoint(nq: int, pc: int, configs_i: torch.Tensor, grad_j: torch.Tensor, configs_j: torch.Tensor,
                          site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor) -> torch.Tensor:
    device = configs_i.device
    _check(device.type == "cuda", "configs must be on CUDA.")
    n_i = configs_i.size(0)
    n_j = configs_j.size(0)
    _check_tensor(configs_i, "configs", device, torch.uint8, nq, n_i)
    _check_tensor(configs_j, "result_configs", device, torch.uint8, nq, n_j)
    _check_tensor(grad_j, "grad", device, torch.float64, 2, n_j)
    n_terms = _check_terms(site, kind, coef, device, nq)
    out_s = _structured(nq, pc, configs_i, grad_j, configs_j, site, kind, coef, 1)
    if out_s is not None:
        return out_s
    lib = _lib.library()
    with torch.cuda.device(device):
        ws_bytes = _size_query(lib.qmb_apply_within_workspace_size, nq, n_i, n_j, n_terms)
        ws = _workspace(ws_bytes, device)
        out = torch.empty((n_i, 2), dtype=torch.float64, device=device)
        _lib.check(
            lib.qmb_apply_within_adjoint(nq, pc, _ptr(configs_i), n_i, _ptr(configs_j), _ptr(grad_j), n_j, _ptr(site), _ptr(kind),
                                         _ptr(coef), n_terms, _ptr(out), _ptr(ws), ws.numel(), _stream(device)),
            "qmb_apply_within_adjoint",
        )
    return out


def _apply_within_plain(nq: int, pc: int, configs_i: torch.Tensor, psi_i: torch.Tensor, configs_j: torch.Tensor,
                        site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor) -> torch.Tensor:
    device = configs_i.device
    _check(device.type == "cuda", "configs must be on CUDA.")
    n_i = configs_i.size(0)
    n_j = configs_j.size(0)
    _check_tensor(configs_i, "configs", device, torch.uint8, nq, n_i)
    _check_tensor(psi_i, "psi", device, torch.float64, 2, n_i)
    _check_tensor(configs_j, "result_configs", device, torch.uint8, nq, n_j)
    n_terms = _check_terms(site, kind, coef, device, nq)
    lib = _lib.library()
    with torch.cuda.device(device):
        ws = _workspace(_size_query(lib.qmb_apply_within_plain_workspace_size, nq, n_j), device)
        out = torch.empty((n_j, 2), dtype=torch.float64, device=device)
        _lib.check(
            lib.qmb_apply_within_plain(nq, pc, _ptr(configs_i), _ptr(psi_i), n_i, _ptr(configs_j), n_j, _ptr(site), _ptr(kind),
                                       _ptr(coef), n_terms, _ptr(out), _ptr(ws), ws.numel(), _stream(device)),
            "qmb_apply_within_plain",
        )
    return out


def _find_relative(nq: int, pc: int, configs_i: torch.Tensor, psi_i: torch.Tensor, count_selected: int, site: torch.Tensor,
                   kind: torch.Tensor, coef: torch.Tensor, configs_exclude: torch.Tensor) -> torch.Tensor:
    device = configs_i.device
    _check(device.type == "cuda", "configs must be on CUDA.")
    n_i = configs_i.size(0)
    n_x = configs_exclude.size(0)
    _check_tensor(configs_i, "configs", device, torch.uint8, nq, n_i)
    _check_tensor(psi_i, "psi", device, torch.float64, 2, n_i)
    _check_tensor(configs_exclude, "configs", device, torch.uint8, nq, n_x)
    n_terms = _check_terms(site, kind, coef, device, nq)
    _check(count_selected >= 0, "count_selected must be non-negative.")
    if count_selected == 0:
        return torch.empty((0, nq), dtype=torch.uint8, device=device)
    lib = _lib.library()
    with torch.cuda.device(device):
        ws = _workspace(_size_query(lib.qmb_find_relative_workspace_size, nq, n_i, n_x, n_terms, count_selected), device)
        out = torch.empty((count_selected, nq), dtype=torch.uint8, device=device)
        rows = torch.zeros((1,), dtype=torch.int64, device=device)
        _lib.check(
            lib.qmb_find_relative(nq, pc, _ptr(configs_i), _ptr(psi_i), n_i, count_selected, _ptr(site), _ptr(kind), _ptr(coef),
                                  n_terms, _ptr(configs_exclude), n_x, _ptr(out), _ptr(rows), _ptr(ws), ws.numel(),
                                  _stream(device)),
            "qmb_find_relative",
        )
        n_rows = int(rows.item())
    return out[:n_rows]


def _diagonal_term(nq: int, pc: int, configs: torch.Tensor, site: torch.Tensor, kind: torch.Tensor,
                   coef: torch.Tensor) -> torch.Tensor:
    device = configs.device
    _check(device.type == "cuda", "configs must be on CUDA.")
    n = configs.size(0)
    _check_tensor(configs, "configs", device, torch.uint8, nq, n)
    n_terms = _check_terms(site, kind, coef, device, nq)
    lib = _lib.library()
    with torch.cuda.device(device):
        ws = _workspace(_size_query(lib.qmb_diagonal_term_workspace_size, nq, n, n_terms), device)
        out = torch.empty((n, 2), dtype=torch.float64, device=device)
        _lib.check(
            lib.qmb_diagonal_term(nq, pc, _ptr(configs), n, _ptr(site), _ptr(kind), _ptr(coef), n_terms, _ptr(out), _ptr(ws),
                                  ws.numel(), _stream(device)),
            "qmb_diagonal_term",
        )
    return out


def _single_relative_impl(nq: int, pc: int, configs: torch.Tensor, site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor,
                          seed: int, keys: torch.Tensor | None) -> torch.Tensor:
    device = configs.device
    _check(device.type == "cuda", "configs must be on CUDA.")
    n = configs.size(0)
    _check_tensor(configs, "configs", device, torch.uint8, nq, n)
    n_terms = _check_terms(site, kind, coef, device, nq)
    if keys is not None:
        _check(keys.device == device and keys.dtype == torch.float64 and keys.is_contiguous() and keys.numel() == n_terms,
               "keys must be a contiguous float64 tensor with one entry per term on the configs' device.")
    lib = _lib.library()
    with torch.cuda.device(device):
        ws = _workspace(_size_query(lib.qmb_single_relative_workspace_size, nq, n, n_terms), device)
        out = torch.empty((n, nq), dtype=torch.uint8, device=device)
        _lib.check(
            lib.qmb_single_relative(nq, pc, _ptr(configs), n, _ptr(site), _ptr(kind), _ptr(coef), n_terms, seed & (2**64 - 1),
                                    _ptr(keys), _ptr(out), _ptr(ws), ws.numel(), _stream(device)),
            "qmb_single_relative",
        )
    return out


def _single_relative(nq: int, pc: int, configs: torch.Tensor, site: torch.Tensor, kind: torch.Tensor,
                     coef: torch.Tensor) -> torch.Tensor:
    # One draw from torch's host generator, as the reference does (qmb/_hamiltonian_cuda.cu:963-964).
    seed = int(torch.randint(0, 2**63 - 1, (), dtype=torch.int64).item())
    return _single_relative_impl(nq, pc, configs, site, kind, coef, seed, None)


def _single_relative_keyed(nq: int, pc: int, configs: torch.Tensor, site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor,
                           keys: torch.Tensor) -> torch.Tensor:
    return _single_relative_impl(nq, pc, configs, site, kind, coef, 0, keys)


def _single_relative_seeded(nq: int, pc: int, configs: torch.Tensor, site: torch.Tensor, kind: torch.Tensor,
                            coef: torch.Tensor, seed: int) -> torch.Tensor:
    return _single_relative_impl(nq, pc, configs, site, kind, coef, seed, None)


# -------------------------------------------------------------------------------- CPU impls -----
# The CPU dispatch key's implementations (qmb/_hamiltonian_cpu.cpp:558-561 registers apply_within
# and find_relative for CPU): the library's host operators on the tensors' host memory, on torch's
# intra-op thread count.  They return when the result is written, like the reference's.
def _host_threads() -> int:
    return max(1, torch.get_num_threads())


def _apply_within_cpu(nq: int, pc: int, configs_i: torch.Tensor, psi_i: torch.Tensor, configs_j: torch.Tensor, site: torch.Tensor,
                      kind: torch.Tensor, coef: torch.Tensor) -> torch.Tensor:
    device = configs_i.device
    _check(device.type == "cpu", "configs must be on CPU.")
    n_i = configs_i.size(0)
    n_j = configs_j.size(0)
    _check_tensor(configs_i, "configs", device, torch.uint8, nq, n_i)
    _check_tensor(psi_i, "psi", device, torch.float64, 2, n_i)
    _check_tensor(configs_j, "result_configs", device, torch.uint8, nq, n_j)
    n_terms = _check_terms(site, kind, coef, device, nq)
    out = torch.empty((n_j, 2), dtype=torch.float64)
    _lib.check(_lib.library().qmb_cpu_apply_within(nq, pc, _ptr(configs_i), _ptr(psi_i), n_i, _ptr(configs_j), n_j, _ptr(site),
                                                   _ptr(kind), _ptr(coef), n_terms, _ptr(out), _host_threads()),
               "qmb_cpu_apply_within")
    return out


def _apply_within_adjoint_cpu(nq: int, pc: int, configs_i: torch.Tensor, grad_j: torch.Tensor, configs_j: torch.Tensor,
                              site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor) -> torch.Tensor:
    device = configs_i.device
    _check(device.type == "cpu", "configs must be on CPU.")
    n_i = configs_i.size(0)
    n_j = configs_j.size(0)
    _check_tensor(configs_i, "configs", device, torch.uint8, nq, n_i)
    _check_tensor(configs_j, "result_configs", device, torch.uint8, nq, n_j)
    _check_tensor(grad_j, "grad", device, torch.float64, 2, n_j)
    n_terms = _check_terms(site, kind, coef, device, nq)
    out = torch.empty((n_i, 2), dtype=torch.float64)
    _lib.check(_lib.library().qmb_cpu_apply_within_adjoint(nq, pc, _ptr(configs_i), n_i, _ptr(configs_j), _ptr(grad_j), n_j,
                                                           _ptr(site), _ptr(kind), _ptr(coef), n_terms, _ptr(out), _host_threads()),
               "qmb_cpu_apply_within_adjoint")
    return out


def _find_relative_cpu(nq: int, pc: int, configs_i: torch.Tensor, psi_i: torch.Tensor, count_selected: int, site: torch.Tensor,
                       kind: torch.Tensor, coef: torch.Tensor, configs_exclude: torch.Tensor) -> torch.Tensor:
    device = configs_i.device
    _check(device.type == "cpu", "configs must be on CPU.")
    n_i = configs_i.size(0)
    n_x = configs_exclude.size(0)
    _check_tensor(configs_i, "configs", device, torch.uint8, nq, n_i)
    _check_tensor(psi_i, "psi", device, torch.float64, 2, n_i)
    _check_tensor(configs_exclude, "configs", device, torch.uint8, nq, n_x)
    n_terms = _check_terms(site, kind, coef, device, nq)
    _check(count_selected >= 0, "count_selected must be non-negative.")
    out = torch.empty((count_selected, nq), dtype=torch.uint8)
    rows = ctypes.c_int64(0)
    _lib.check(_lib.library().qmb_cpu_find_relative(nq, pc, _ptr(configs_i), _ptr(psi_i), n_i, count_selected, _ptr(site), _ptr(kind),
                                                    _ptr(coef), n_terms, _ptr(configs_exclude), n_x, _ptr(out), ctypes.byref(rows),
                                                    _host_threads()),
               "qmb_cpu_find_relative")
    return out[:rows.value].clone()


# ---------------------------------------------------------------------------- registration -----
SCHEMAS = {
    "apply_within": "apply_within(Tensor configs_i, Tensor psi_i, Tensor configs_j, Tensor site, Tensor kind, Tensor coef) -> Tensor",
    "find_relative": ("find_relative(Tensor configs_i, Tensor psi_i, int count_selected, Tensor site, Tensor kind, Tensor coef, "
                      "Tensor configs_exclude) -> Tensor"),
    "diagonal_term": "diagonal_term(Tensor configs, Tensor site, Tensor kind, Tensor coef) -> Tensor",
    "single_relative": "single_relative(Tensor configs, Tensor site, Tensor kind, Tensor coef) -> Tensor",
    "apply_within_adjoint": ("apply_within_adjoint(Tensor configs_i, Tensor grad_j, Tensor configs_j, Tensor site, Tensor kind, "
                             "Tensor coef) -> Tensor"),
    "apply_within_plain": ("apply_within_plain(Tensor configs_i, Tensor psi_i, Tensor configs_j, Tensor site, Tensor kind, "
                           "Tensor coef) -> Tensor"),
    "single_relative_keyed": "single_relative_keyed(Tensor configs, Tensor site, Tensor kind, Tensor coef, Tensor keys) -> Tensor",
    "single_relative_seeded": "single_relative_seeded(Tensor configs, Tensor site, Tensor kind, Tensor coef, int seed) -> Tensor",
}

# impls on the CPU dispatch key (the others raise there, as the reference's do)
CPU_IMPLS = {
    "apply_within": _apply_within_cpu,
    "find_relative": _find_relative_cpu,
    "apply_within_adjoint": _apply_within_adjoint_cpu,
}

IMPLS = {
    "apply_within": _apply_within,
    "find_relative": _find_relative,
    "diagonal_term": _diagonal_term,
    "single_relative": _single_relative,
    "apply_within_adjoint": _apply_within_adjoint,
    "apply_within_plain": _apply_within_plain,
    "single_relative_keyed": _single_relative_keyed,
    "single_relative_seeded": _single_relative_seeded,
}


def namespace(n_qubytes: int, particle_cut: int) -> str:
    """Op namespace of the reference (qmb/_hamiltonian.cpp:81-83)."""
    return f"qmb_hamiltonian_{n_qubytes}_{particle_cut}"


def load(n_qubytes: int, particle_cut: int) -> object:
    """Registers (once) and returns torch.ops.qmb_hamiltonian_{n_qubytes}_{particle_cut}."""
    if particle_cut not in (1, 2):
        raise ValueError(f"particle_cut must be 1 or 2, got {particle_cut}")
    if not 1 <= n_qubytes <= 32:
        raise ValueError(f"n_qubytes must be in [1, 32], got {n_qubytes}")
    _lib.library()  # fail loudly here if the HIP library is missing
    name = namespace(n_qubytes, particle_cut)
    with _REGISTRY_LOCK:
        if name not in _REGISTRY:
            lib = torch.library.Library(name, "DEF")
            for op, schema in SCHEMAS.items():
                lib.define(schema)
                lib.impl(op, functools.partial(IMPLS[op], n_qubytes, particle_cut), "CUDA")
                if op in CPU_IMPLS:
                    lib.impl(op, functools.partial(CPU_IMPLS[op], n_qubytes, particle_cut), "CPU")
            _REGISTRY[name] = lib
            ns = getattr(torch.ops, name)

            def _setup(ctx: object, inputs: tuple[object, ...], output: object) -> None:
                configs_i, _, configs_j, site, kind, coef = inputs
                ctx.save_for_backward(configs_i, configs_j, site, kind, coef)  # type: ignore[attr-defined]

            def _backward(ctx: object, grad: torch.Tensor) -> tuple[object, ...]:
                configs_i, configs_j, site, kind, coef = ctx.saved_tensors  # type: ignore[attr-defined]
                grad_i = ns.apply_within_adjoint(configs_i, grad.contiguous(), configs_j, site, kind, coef)
                return None, grad_i, None, None, None, None

            torch.library.register_autograd(f"{name}::apply_within", _backward, setup_context=_setup, lib=lib)
    return getattr(torch.ops, name)
