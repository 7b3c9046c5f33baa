"""
Prepared objects: a compiled term table and a hash index over a config set, built once on the GPU
and reused across many applications (Krylov loops, forward + adjoint of one autograd step, the
batched launch).  Thin holders around ``qmb_terms`` / ``qmb_targets`` of include/qmb_hip.h whose
device buffers come from PyTorch's caching allocator.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .ops import _check_terms, _ptr, _stream, _workspace

MODE_FORWARD = 0
MODE_ADJOINT = 1
MODE_COUNT = 2


STRATEGY_AUTO = 0
STRATEGY_TILES = 1
STRATEGY_DIRECT = 2
# auto picks the direct kernel up to this many terms (kDirectMaxTerms in the library)
DIRECT_MAX_TERMS = 4096
# the dense pair-excitation table covers configs of up to this many bytes
PAIR_MAX_QUBYTES = 16
# qmb_terms_build_residual exclusion mask
EXCLUDE_PAIRS = 1
EXCLUDE_SINGLES = 2
EXCLUDE_DIAGONAL = 4
EXCLUDE_HOPS = 8
# The hop path tests the row pairs of each strings group; it is taken while those tests stay within
# this many per (row, hop term) -- the enumeration's work per row is a few lookups per hop term.
HOP_PAIRS_PER_TERM_ROW = 4


def _device_count(buffer: torch.Tensor, counts_ptr: int) -> int:
    """First int64 of a device counter array living inside `buffer` (synchronises)."""
    offset = int(counts_ptr) - buffer.data_ptr()
    return int(buffer[offset:offset + 8].view(torch.int64).item())


# A Hermitian table gives the same map in both modes; its structured kernels run in the forward
# orientation, where a coefficient load reads one table row with the group's columns ascending
# across the lanes (the adjoint orientation reads one column of many rows).
HERMITIAN_MODE = MODE_FORWARD

# qmb_launch.symmetric values
SYMMETRIC_OFF = 0
SYMMETRIC_GATHER = 1   # forward computed as the adjoint gather (A = A^H): no atomics per contribution
SYMMETRIC_HALF = 2     # each connected pair looked up once, both directions scattered with atomics


def _launch(tile: int = 0, groups: int = 0, strategy: int = 0, symmetric: bool | int = False) -> _lib.QmbLaunch | None:
    sym = SYMMETRIC_GATHER if symmetric is True else int(symmetric)
    if tile == 0 and groups == 0 and strategy == 0 and sym == 0:
        return None
    return _lib.QmbLaunch(tile, groups, strategy, sym)


def can_use_symmetric(terms: "PreparedTerms", targets: "PreparedTargets", configs_i: torch.Tensor) -> bool:
    """Hermitian operator applied within the indexed set itself, rows distinct."""
    return terms.hermitian and targets.indexes(configs_i) and targets.duplicates() == 0


def is_hermitian(site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor, n_qubytes: int, particle_cut: int,
                 tol: float = 1e-12) -> bool:
    """Whether sum_t coef_t O_t equals its adjoint (exact test on the compiled terms, host side)."""
    s = site.detach().cpu().contiguous()
    k = kind.detach().cpu().contiguous()
    c = coef.detach().cpu().contiguous()
    out = ctypes.c_int32(0)
    lib = _lib.library()
    _lib.check(lib.qmb_terms_hermitian_host(n_qubytes, particle_cut, _ptr(s), _ptr(k), _ptr(c), s.shape[0], tol, ctypes.byref(out)),
               "qmb_terms_hermitian_host")
    return bool(out.value)


class PreparedTerms:
    """Compiled, bucketed term table on one device (qmb_terms_build)."""

    def __init__(self, site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor, n_qubytes: int, particle_cut: int,
                 check_hermitian: bool = True, pairs: bool = True) -> None:
        device = site.device
        site, kind, coef = site.contiguous(), kind.contiguous(), coef.contiguous()
        self.n_terms = _check_terms(site, kind, coef, device, n_qubytes)
        self.n_qubytes = n_qubytes
        self.particle_cut = particle_cut
        self.device = device
        lib = _lib.library()
        size = ctypes.c_size_t(0)
        _lib.check(lib.qmb_terms_workspace_size(n_qubytes, self.n_terms, ctypes.byref(size)), "qmb_terms_workspace_size")
        self.buffer = _workspace(size.value, device)
        self.desc = _lib.QmbTerms()
        # the raw table must outlive the asynchronous build
        self._inputs = (site, kind, coef)
        with torch.cuda.device(device):
            _lib.check(lib.qmb_terms_build(n_qubytes, particle_cut, _ptr(site), _ptr(kind), _ptr(coef), self.n_terms, _ptr(self.buffer),
                                           self.buffer.numel(), ctypes.byref(self.desc), _stream(device)), "qmb_terms_build")
        self.hermitian = is_hermitian(site, kind, coef, n_qubytes, particle_cut) if check_hermitian else False
        # Structured path (used when the table has pair excitations): merged pair and single
        # excitation tables, the diagonal records of the full table, and the rest for enumeration.
        self.pairs_desc: _lib.QmbPairs | None = None
        self.singles_desc: _lib.QmbSingles | None = None
        self.residual_desc: _lib.QmbTerms | None = None
        self.n_pair_terms = 0
        self.n_single_terms = 0
        self.n_residual_terms = self.n_terms
        if pairs and n_qubytes <= PAIR_MAX_QUBYTES and self.n_terms > 0:
            _lib.check(lib.qmb_pairs_workspace_size(n_qubytes, ctypes.byref(size)), "qmb_pairs_workspace_size")
            pbuf = _workspace(size.value, device)
            pdesc = _lib.QmbPairs()
            with torch.cuda.device(device):
                _lib.check(lib.qmb_pairs_build(n_qubytes, particle_cut, _ptr(site), _ptr(kind), _ptr(coef), self.n_terms, _ptr(pbuf),
                                               pbuf.numel(), ctypes.byref(pdesc), _stream(device)), "qmb_pairs_build")
            self.n_pair_terms = _device_count(pbuf, pdesc.counts)
            if self.n_pair_terms > 0:
                self.pairs_buffer, self.pairs_desc = pbuf, pdesc
                if bool((coef[:, 1] == 0).all().item()):
                    # real table: the pair kernels read 8-byte coefficients
                    with torch.cuda.device(device):
                        _lib.check(lib.qmb_pairs_set_real(ctypes.byref(pdesc), 1, _stream(device)), "qmb_pairs_set_real")
                _lib.check(lib.qmb_singles_workspace_size(n_qubytes, ctypes.byref(size)), "qmb_singles_workspace_size")
                sbuf = _workspace(size.value, device)
                sdesc = _lib.QmbSingles()
                with torch.cuda.device(device):
                    _lib.check(lib.qmb_singles_build(n_qubytes, particle_cut, _ptr(site), _ptr(kind), _ptr(coef), self.n_terms,
                                                     _ptr(sbuf), sbuf.numel(), ctypes.byref(sdesc), _stream(device)), "qmb_singles_build")
                self.n_single_terms = _device_count(sbuf, sdesc.counts)
                if self.n_single_terms > 0 and pdesc.real:
                    # real table: the single-excitation kernels read 8-byte coefficients too
                    with torch.cuda.device(device):
                        _lib.check(lib.qmb_singles_set_real(ctypes.byref(sdesc), 1, _stream(device)), "qmb_singles_set_real")
                exclude = EXCLUDE_PAIRS | EXCLUDE_DIAGONAL
                if self.n_single_terms > 0:
                    self.singles_buffer, self.singles_desc = sbuf, sdesc
                    exclude |= EXCLUDE_SINGLES
                _lib.check(lib.qmb_terms_workspace_size(n_qubytes, self.n_terms, ctypes.byref(size)), "qmb_terms_workspace_size")
                self.residual_buffer = _workspace(size.value, device)
                self.residual_desc = _lib.QmbTerms()
                with torch.cuda.device(device):
                    _lib.check(lib.qmb_terms_build_residual(n_qubytes, particle_cut, _ptr(site), _ptr(kind), _ptr(coef), self.n_terms,
                                                            exclude, _ptr(self.residual_buffer), self.residual_buffer.numel(),
                                                            ctypes.byref(self.residual_desc), _stream(device)),
                               "qmb_terms_build_residual")
                self.n_residual_terms = _device_count(self.residual_buffer, self.residual_desc.counts)
        # Hop path (a table without pair excitations whose off-diagonal terms include single
        # excitations between two sites of one parity): the single-excitation table, the diagonal
        # records of the full table, and the rest for enumeration.
        self.hops = False
        self.n_hop_terms = 0
        if pairs and n_qubytes <= PAIR_MAX_QUBYTES and self.n_terms > 0 and self.pairs_desc is None:
            self._prepare_hops(site, kind, coef)

    def _prepare_hops(self, site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor) -> None:
        lib = _lib.library()
        device, nq, pc = self.device, self.n_qubytes, self.particle_cut
        size = ctypes.c_size_t(0)
        _lib.check(lib.qmb_terms_workspace_size(nq, self.n_terms, ctypes.byref(size)), "qmb_terms_workspace_size")
        rbuf = _workspace(size.value, device)
        rdesc = _lib.QmbTerms()
        with torch.cuda.device(device):
            _lib.check(lib.qmb_terms_build_residual(nq, pc, _ptr(site), _ptr(kind), _ptr(coef), self.n_terms, EXCLUDE_DIAGONAL | EXCLUDE_HOPS,
                                                    _ptr(rbuf), rbuf.numel(), ctypes.byref(rdesc), _stream(device)),
                       "qmb_terms_build_residual")
        n_residual = _device_count(rbuf, rdesc.counts)
        valid, diagonal, _ = self.counts()
        n_hops = valid - diagonal - n_residual
        if n_hops <= 0:
            return
        _lib.check(lib.qmb_singles_workspace_size(nq, ctypes.byref(size)), "qmb_singles_workspace_size")
        sbuf = _workspace(size.value, device)
        sdesc = _lib.QmbSingles()
        with torch.cuda.device(device):
            _lib.check(lib.qmb_singles_build(nq, pc, _ptr(site), _ptr(kind), _ptr(coef), self.n_terms, _ptr(sbuf), sbuf.numel(),
                                             ctypes.byref(sdesc), _stream(device)), "qmb_singles_build")
            if bool((coef[:, 1] == 0).all().item()):
                # real table: the hop kernel reads 8-byte coefficients
                _lib.check(lib.qmb_singles_set_real(ctypes.byref(sdesc), 1, _stream(device)), "qmb_singles_set_real")
        self.singles_buffer, self.singles_desc = sbuf, sdesc
        self.n_single_terms = _device_count(sbuf, sdesc.counts)
        self.residual_buffer, self.residual_desc = rbuf, rdesc
        self.n_residual_terms = n_residual
        self.n_hop_terms = n_hops
        self.hops = True

    def counts(self) -> tuple[int, int, int]:
        """(valid terms, diagonal terms, invalid terms) -- synchronises."""
        offset = int(self.desc.counts) - self.buffer.data_ptr()
        v = self.buffer[offset:offset + 32].view(torch.int64).tolist()
        return int(v[0]), int(v[1]), int(v[2])


class PreparedTargets:
    """Hash index over a config set on one device (qmb_targets_build)."""

    def __init__(self, configs: torch.Tensor, n_qubytes: int, distinct: bool | None = None, max_occupied: int | None = None) -> None:
        """distinct=True asserts the rows are pairwise different and max_occupied bounds the number of
        occupied sites per row; either left as None is measured on first use (a device sync)."""
        configs = configs.contiguous()
        if configs.dtype != torch.uint8 or configs.dim() != 2 or configs.size(1) != n_qubytes:
            raise RuntimeError("configs must be a uint8 [n, n_qubytes] tensor.")
        device = configs.device
        self.configs = configs
        self.n = configs.size(0)
        self.n_qubytes = n_qubytes
        self.device = device
        lib = _lib.library()
        size = ctypes.c_size_t(0)
        _lib.check(lib.qmb_targets_workspace_size(n_qubytes, self.n, ctypes.byref(size)), "qmb_targets_workspace_size")
        self.buffer = _workspace(size.value, device)
        # the key words and hash columns now; the hash table, filter and duplicate count on first use
        # of `desc` (the core, strings and hop paths read only keys_desc)
        self._desc = _lib.QmbTargets()
        self._indexed = False
        with torch.cuda.device(device):
            _lib.check(lib.qmb_targets_build_keys(n_qubytes, _ptr(configs), self.n, _ptr(self.buffer), self.buffer.numel(),
                                                  ctypes.byref(self._desc), _stream(device)), "qmb_targets_build_keys")
        self._duplicates: int | None = 0 if distinct else None
        self._max_occupied = max_occupied
        # a bound from the caller is checked against the rows once per core index (a measured one
        # holds by construction)
        self._occupied_hint = max_occupied is not None
        self._cores: dict[int, tuple[torch.Tensor, _lib.QmbCores]] = {}
        self._unchecked: set[int] = set()
        self._strings: dict[int, tuple[torch.Tensor, _lib.QmbStrings]] = {}
        self._strings_stats: dict[int, dict[str, int]] = {}
        self._diagonal: dict[int, tuple[PreparedTerms, torch.Tensor]] = {}

    @property
    def desc(self) -> _lib.QmbTargets:
        """The descriptor with its hash index (built on first use)."""
        if not self._indexed:
            with torch.cuda.device(self.device):
                _lib.check(_lib.library().qmb_targets_build_index(ctypes.byref(self._desc), _stream(self.device)), "qmb_targets_build_index")
            self._indexed = True
        return self._desc

    @property
    def keys_desc(self) -> _lib.QmbTargets:
        """The descriptor for kernels that read only the key words and hash columns."""
        return self._desc

    def duplicates(self) -> int:
        """Rows equal to an earlier row (synchronises once)."""
        if self._duplicates is None:
            offset = int(self.desc.duplicates) - self.buffer.data_ptr()
            self._duplicates = int(self.buffer[offset:offset + 8].view(torch.int64).item())
        return self._duplicates

    def indexes(self, configs: torch.Tensor) -> bool:
        """True when `configs` is the very tensor this index was built over."""
        return configs.data_ptr() == self.configs.data_ptr() and configs.shape == self.configs.shape

    def max_occupied(self) -> int:
        """Largest number of occupied sites in a row (synchronises unless given at construction)."""
        if self._max_occupied is None:
            if self.n == 0:
                self._max_occupied = 0
            else:
                table = torch.tensor([bin(v).count("1") for v in range(256)], dtype=torch.int32, device=self.device)
                self._max_occupied = int(table[self.configs.long()].sum(dim=1).max().item())
        return self._max_occupied

    def has_cores(self, order: int) -> bool:
        """True once the core index of this order is built (or being built on a stream)."""
        return order in self._cores

    def cores(self, order: int = 2, check: bool = True) -> _lib.QmbCores:
        """Core index of the set (order 2: pair-excitation cores, 1: single-excitation cores), built
        on first use on the current stream.  With a max_occupied bound from the caller, the build's
        overflow count is read (a sync) on the first call with check=True: an index built on a
        side stream is checked after the caller has joined that stream."""
        if order not in self._cores:
            lib = _lib.library()
            occ = self.max_occupied()
            size = ctypes.c_size_t(0)
            _lib.check(lib.qmb_cores_workspace_size(self.n_qubytes, self.n, occ, order, ctypes.byref(size)), "qmb_cores_workspace_size")
            buf = _workspace(size.value, self.device)
            desc = _lib.QmbCores()
            with torch.cuda.device(self.device):
                _lib.check(lib.qmb_cores_build(ctypes.byref(self.keys_desc), occ, order, _ptr(buf), buf.numel(), ctypes.byref(desc),
                                               _stream(self.device)), "qmb_cores_build")
            self._cores[order] = (buf, desc)
            if self._occupied_hint:
                self._unchecked.add(order)
        if check and order in self._unchecked:
            buf, desc = self._cores[order]
            self._unchecked.discard(order)
            check_core_overflow(_device_count(buf, desc.counts + 16), self.max_occupied())
        return self._cores[order][1]


    def has_strings(self, parity: int) -> bool:
        """True once the strings index of this parity is built (or being built on a stream)."""
        return parity in self._strings

    def has_diagonal_values(self, terms: "PreparedTerms") -> bool:
        """True once the diagonal values under `terms` are computed (or being computed)."""
        return id(terms) in self._diagonal

    def strings(self, parity: int) -> _lib.QmbStrings:
        """Strings index of the set for the hops between sites of `parity` (0 even, 1 odd): the rows
        grouped by their string on the other parity's sites, built on first use."""
        if parity not in self._strings:
            lib = _lib.library()
            size = ctypes.c_size_t(0)
            _lib.check(lib.qmb_strings_workspace_size(self.n_qubytes, self.n, ctypes.byref(size)), "qmb_strings_workspace_size")
            buf = _workspace(size.value, self.device)
            desc = _lib.QmbStrings()
            with torch.cuda.device(self.device):
                _lib.check(lib.qmb_strings_build(ctypes.byref(self.keys_desc), parity, _ptr(buf), buf.numel(), ctypes.byref(desc),
                                                 _stream(self.device)), "qmb_strings_build")
            self._strings[parity] = (buf, desc)
        return self._strings[parity][1]

    def diagonal_values(self, terms: "PreparedTerms") -> torch.Tensor:
        """The diagonal value of every row under `terms` (float64 [n, 2]), computed on first use and
        kept with this set, so that each further application adds it with one pass."""
        key = id(terms)
        if key not in self._diagonal:
            d = torch.empty((max(self.n, 1), 2), dtype=torch.float64, device=self.device)
            with torch.cuda.device(self.device):
                _lib.check(_lib.library().qmb_diagonal_values(ctypes.byref(terms.desc), ctypes.byref(self.keys_desc), _ptr(d),
                                                              _stream(self.device)), "qmb_diagonal_values")
            self._diagonal[key] = (terms, d)
        return self._diagonal[key][1]

    def strings_stats(self, parity: int) -> dict[str, int]:
        """Sizes of a strings index (synchronises): entries, groups, the row pairs the hop kernel
        tests (sum of size^2 over the groups of two or more rows), the largest group and the entries
        of the groups of two or more rows."""
        if parity not in self._strings_stats:
            desc = self.strings(parity)
            buf = self._strings[parity][0]
            off = int(desc.counts) - buf.data_ptr()
            v = buf[off:off + 40].view(torch.int64).tolist()
            self._strings_stats[parity] = {"entries": int(v[0]), "groups": int(v[1]), "pair_tests": int(v[2]), "largest": int(v[3]),
                                           "multi_entries": int(v[4])}
        return self._strings_stats[parity]


def core_stats(targets: PreparedTargets, order: int) -> dict[str, int]:
    """Sizes of a set's core index (synchronises): entries, groups, and the entries and groups of the
    groups holding two or more entries (the only ones the group kernels walk)."""
    cores = targets.cores(order)
    buf = targets._cores[order][0]  # pylint: disable=protected-access
    base = buf.data_ptr()
    counts = buf[int(cores.counts) - base:int(cores.counts) - base + 16].view(torch.int64).tolist()
    entries, groups = int(counts[0]), int(counts[1])
    off = int(cores.starts) - base
    starts = buf[off:off + 4 * (groups + 1)].view(torch.int32).to(torch.int64) & 0x7FFFFFFF  # bit 31: mixed group
    sizes = starts[1:] - starts[:-1]
    multi = sizes > 1
    return {"entries": entries, "groups": groups, "multi_entries": int(sizes[multi].sum().item()),
            "multi_groups": int(multi.sum().item()), "largest": int(sizes.max().item()) if groups else 0}


def check_core_overflow(overflow: int, max_occupied: int) -> None:
    """Raises when a core index found rows with more occupied sites than the bound it was built
    for (the kernels would leave those rows' pair and single terms out)."""
    if overflow != 0:
        raise RuntimeError(f"{overflow} configs have more than max_occupied={max_occupied} occupied sites; "
                           "pass a larger max_occupied or None to measure it")


def can_use_pairs(terms: PreparedTerms, targets: PreparedTargets, configs_i: torch.Tensor) -> bool:
    """The pair-excitation path applies within one distinct-row set indexed by `targets`."""
    return terms.pairs_desc is not None and targets.indexes(configs_i) and targets.duplicates() == 0


def can_use_hops(terms: PreparedTerms, targets: PreparedTargets, configs_i: torch.Tensor) -> bool:
    """The hop path applies within one distinct-row set indexed by `targets`, for a table with hops
    and no pair excitations, while its pairwise tests (the strings indexes' sum of group size^2,
    built here: a synchronisation) stay within HOP_PAIRS_PER_TERM_ROW per row and hop term."""
    if not terms.hops or terms.pairs_desc is not None or not targets.indexes(configs_i) or targets.duplicates() != 0:
        return False
    # the two parities' indexes build side by side (one on a second stream), then one sync; the
    # second stream first computes the rows' diagonal values the hop path adds next, a single kernel
    # that fills the GPU while the host enqueues the indexes' many small launches
    if not targets.has_strings(1):
        device = targets.device
        side = _side_stream(device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            if not targets.has_diagonal_values(terms):
                targets.diagonal_values(terms).record_stream(torch.cuda.current_stream(device))
            targets.strings(1)
        targets._strings[1][0].record_stream(torch.cuda.current_stream(device))  # pylint: disable=protected-access
        targets.strings(0)
        torch.cuda.current_stream(device).wait_stream(side)
    tests = sum(targets.strings_stats(parity)["pair_tests"] for parity in (0, 1))
    return tests <= HOP_PAIRS_PER_TERM_ROW * max(targets.n, 1) * terms.n_hop_terms


def apply_hops(terms: PreparedTerms, targets: PreparedTargets, vec: torch.Tensor, mode: int, out: torch.Tensor) -> None:
    """The hop path's kernels, accumulated into out: the diagonal terms (their per-row values kept
    with the set), the hops of both parities and the residual enumeration (terms outside those
    classes)."""
    lib = _lib.library()
    device = targets.device
    stream = _stream(device)
    with torch.cuda.device(device):
        # a first application computes the rows' diagonal values on a second stream beside the hop
        # kernels, and adds them after the join
        side = None
        if not targets.has_diagonal_values(terms):
            side = _side_stream(device)
            side.wait_stream(torch.cuda.current_stream(device))
            with torch.cuda.stream(side):
                targets.diagonal_values(terms).record_stream(torch.cuda.current_stream(device))
        if terms.n_residual_terms > 0:
            _lib.check(lib.qmb_apply_prepared(ctypes.byref(terms.residual_desc), ctypes.byref(targets.desc), _ptr(targets.configs),
                                              targets.n, _ptr(vec), _ptr(out), mode, None, stream), "qmb_apply_prepared")
        for parity in (0, 1):
            _lib.check(lib.qmb_apply_hops(ctypes.byref(terms.singles_desc), ctypes.byref(targets.strings(parity)), ctypes.byref(targets.keys_desc),
                                          _ptr(vec), _ptr(out), mode, stream), "qmb_apply_hops")
        if side is not None:
            torch.cuda.current_stream(device).wait_stream(side)
        _lib.check(lib.qmb_apply_diagonal_values(targets.n, _ptr(targets.diagonal_values(terms)), _ptr(vec), _ptr(out), mode, stream),
                   "qmb_apply_diagonal_values")


_SIDE_STREAMS: dict[int, torch.cuda.Stream] = {}


def _side_stream(device: torch.device) -> torch.cuda.Stream:
    """A second stream per device (default priority) for index builds beside the apply kernels."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx not in _SIDE_STREAMS:
        _SIDE_STREAMS[idx] = torch.cuda.Stream(device=torch.device("cuda", idx))
    return _SIDE_STREAMS[idx]


def apply(terms: PreparedTerms, targets: PreparedTargets, configs_i: torch.Tensor, vec: torch.Tensor, mode: int,
          out: torch.Tensor | None = None, tile: int = 0, groups: int = 0, zero: bool = True, strategy: int = 0,
          symmetric: bool | int | None = None, pairs: bool | None = None, hops: bool | None = None) -> torch.Tensor:
    """mode 0: psi_j = A psi_i (vec = psi_i as float64 [n_i, 2]); mode 1: grad_i = A^H grad_j.

    `out` (float64 [n, 2]) is accumulated into when given, after zeroing unless zero=False.
    symmetric=None uses the symmetric kernel whenever can_use_symmetric() holds.
    pairs=None takes the pair-excitation path whenever can_use_pairs() holds: the pair-excitation
    terms are applied by the core-group kernel and the rest by the enumeration kernel.
    hops=None takes the hop path whenever can_use_hops() holds (and the pair path does not).
    """
    configs_i = configs_i.contiguous()
    vec = vec.contiguous()
    n_i = configs_i.size(0)
    n_out = targets.n if mode == MODE_FORWARD else n_i
    device = configs_i.device
    if out is None:
        out = torch.zeros((n_out, 2), dtype=torch.float64, device=device)
    elif zero:
        out.zero_()
    if symmetric is None:
        symmetric = mode != MODE_COUNT and can_use_symmetric(terms, targets, configs_i)
    use_pairs = mode != MODE_COUNT and (pairs is None or pairs) and can_use_pairs(terms, targets, configs_i)
    if pairs and not use_pairs:
        raise RuntimeError("the pair-excitation path needs pair terms and configs_i equal to the indexed, distinct-row set.")
    use_hops = not use_pairs and mode != MODE_COUNT and (hops is None or hops) and can_use_hops(terms, targets, configs_i)
    if hops and not use_hops:
        raise RuntimeError("the hop path needs hop terms, no pair excitations and configs_i equal to the indexed, distinct-row set.")
    if use_hops:
        apply_hops(terms, targets, vec, mode, out)
        return out
    launch = _launch(tile, groups, strategy, symmetric)
    lib = _lib.library()
    stream = _stream(device)
    with torch.cuda.device(device):
        if not use_pairs:
            _lib.check(lib.qmb_apply_prepared(ctypes.byref(terms.desc), ctypes.byref(targets.desc), _ptr(configs_i), n_i, _ptr(vec),
                                              _ptr(out), mode, ctypes.byref(launch) if launch else None, stream), "qmb_apply_prepared")
            return out
        # a Hermitian table: forward and adjoint are the same map, applied in the faster orientation
        pmode = HERMITIAN_MODE if symmetric else mode
        if terms.n_residual_terms > 0:
            _lib.check(lib.qmb_apply_prepared(ctypes.byref(terms.residual_desc), ctypes.byref(targets.desc), _ptr(configs_i), n_i,
                                              _ptr(vec), _ptr(out), mode, ctypes.byref(launch) if launch else None, stream),
                       "qmb_apply_prepared")
        # A first application builds the order-1 index on a side stream beside the pair kernel (which
        # leaves memory bandwidth and room on the CUs unused): forked once the order-2 index is
        # built, joined before the single kernel.
        cores2 = targets.cores(2)
        side = None
        if terms.singles_desc is not None and not targets.has_cores(1):
            side = _side_stream(device)
            side.wait_stream(torch.cuda.current_stream(device))
            with torch.cuda.stream(side):
                targets.cores(1, check=False)
            # the index's memory is allocated on the side stream and read on this one
            targets._cores[1][0].record_stream(torch.cuda.current_stream(device))  # pylint: disable=protected-access
        _lib.check(lib.qmb_apply_diagonal(ctypes.byref(terms.pairs_desc), ctypes.byref(targets.keys_desc), _ptr(vec), _ptr(out), pmode,
                                          stream), "qmb_apply_diagonal")
        _lib.check(lib.qmb_apply_pairs(ctypes.byref(terms.pairs_desc), ctypes.byref(cores2), ctypes.byref(targets.keys_desc),
                                       _ptr(vec), _ptr(out), pmode, stream), "qmb_apply_pairs")
        if terms.singles_desc is not None:
            if side is not None:
                torch.cuda.current_stream(device).wait_stream(side)
            _lib.check(lib.qmb_apply_singles(ctypes.byref(terms.singles_desc), ctypes.byref(targets.cores(1)), ctypes.byref(targets.keys_desc),
                                             _ptr(vec), _ptr(out), pmode, stream), "qmb_apply_singles")
    return out


def apply_dual(terms: PreparedTerms, targets: PreparedTargets, vec: torch.Tensor, vec2: torch.Tensor,
               mode: int = MODE_ADJOINT) -> tuple[torch.Tensor, torch.Tensor]:
    """A^H vec and A^H vec2 (mode 1; A with mode 0) on the indexed set in one pass per kernel
    (qmb_apply_structured_dual): for a Hermitian table the forward and the backward of a training
    step, then run in HERMITIAN_MODE.  Needs the structured path (pair excitations, distinct rows)
    and no residual terms."""
    if terms.hermitian:
        mode = HERMITIAN_MODE
    if not can_use_pairs(terms, targets, targets.configs) or terms.n_residual_terms > 0:
        raise RuntimeError("apply_dual needs the structured path: pair excitations, distinct rows and no residual terms.")
    vec, vec2 = vec.contiguous(), vec2.contiguous()
    out = torch.zeros((targets.n, 2), dtype=torch.float64, device=targets.device)
    out2 = torch.zeros_like(out)
    singles = ctypes.byref(terms.singles_desc) if terms.singles_desc is not None else None
    cores1 = ctypes.byref(targets.cores(1)) if terms.singles_desc is not None else None
    with torch.cuda.device(targets.device):
        _lib.check(_lib.library().qmb_apply_structured_dual(ctypes.byref(terms.pairs_desc), singles, ctypes.byref(targets.cores(2)), cores1,
                                                            ctypes.byref(targets.keys_desc), _ptr(vec), _ptr(out), _ptr(vec2), _ptr(out2),
                                                            mode, _stream(targets.device)), "qmb_apply_structured_dual")
    return out, out2


def apply_across(terms: PreparedTerms, configs_i: torch.Tensor, vec: torch.Tensor, configs_j: torch.Tensor, mode: int,
                 max_occupied: int | None = None) -> torch.Tensor | None:
    """The structured path between two config sets: mode 0 returns A psi_i on the rows of configs_j
    (vec = psi_i, float64 [n_i, 2]); mode 1 returns A^H g on the rows of configs_i (vec = g on
    configs_j).  Both are computed within one set that holds both -- configs_j followed by the rows
    of configs_i that configs_j lacks -- with the input placed on its rows (repeated rows of
    configs_i add up) and the output read from them.  That is exact: the element between two rows
    of the joint set is the element between the two configs.  Returns None when the path does not
    apply (no pair excitations, configs_j repeats a row, or a missing row of configs_i repeats),
    leaving the caller to the enumeration kernel.  Synchronises to size the joint set."""
    if terms.pairs_desc is None or mode not in (MODE_FORWARD, MODE_ADJOINT):
        return None
    n_i, n_j = configs_i.size(0), configs_j.size(0)
    if n_i == 0 or n_j == 0:
        return None
    nq = terms.n_qubytes
    device = configs_i.device
    configs_i = configs_i.contiguous()
    configs_j = configs_j.contiguous()
    tj = PreparedTargets(configs_j, nq, max_occupied=max_occupied)
    if tj.duplicates() != 0:
        return None
    rows = torch.empty(n_i, dtype=torch.int64, device=device)
    with torch.cuda.device(device):
        _lib.check(_lib.library().qmb_targets_lookup(ctypes.byref(tj.desc), _ptr(configs_i), n_i, _ptr(rows), _stream(device)),
                   "qmb_targets_lookup")
    missing = rows < 0
    n_miss = int(missing.sum().item())
    if n_miss:
        union = torch.cat([configs_j, configs_i[missing]])
        pos = rows.clone()
        pos[missing] = torch.arange(n_j, n_j + n_miss, device=device)
        targets = PreparedTargets(union, nq, max_occupied=max_occupied)
        if targets.duplicates() != 0:
            return None
    else:
        union, pos, targets = configs_j, rows, tj
    uvec = torch.zeros((union.size(0), 2), dtype=torch.float64, device=device)
    if mode == MODE_FORWARD:
        uvec.index_add_(0, pos, vec.contiguous())
        return apply(terms, targets, union, uvec, MODE_FORWARD)[:n_j]
    uvec[:n_j] = vec
    return apply(terms, targets, union, uvec, MODE_ADJOINT)[pos]


def count(terms: PreparedTerms, targets: PreparedTargets, configs_i: torch.Tensor, tile: int = 0, groups: int = 0,
          strategy: int = 0) -> tuple[int, int]:
    """(M_gen, M_hit): successful term applications and those whose target is indexed -- synchronises."""
    configs_i = configs_i.contiguous()
    device = configs_i.device
    counters = torch.zeros(2, dtype=torch.int64, device=device)
    launch = _launch(tile, groups, strategy)
    lib = _lib.library()
    with torch.cuda.device(device):
        _lib.check(lib.qmb_apply_prepared(ctypes.byref(terms.desc), ctypes.byref(targets.desc), _ptr(configs_i), configs_i.size(0), None,
                                          _ptr(counters), MODE_COUNT, ctypes.byref(launch) if launch else None, _stream(device)),
                   "qmb_apply_prepared(count)")
    gen, hit = counters.tolist()
    return int(gen), int(hit)
