"""
Batched application of many Hamiltonians to many config sets in two persistent launches.

Each instance is a PreparedTerms (with the structured tables: pair and single excitations,
diagonal records) and a PreparedTargets over its distinct config set.  The inputs and outputs of all
instances are packed into two device buffers, so the plan (qmb_batch_plan) is built once and every
launch (qmb_apply_batch) is two persistent kernels over all instances' work (the pair units, then the
single and diagonal units); the whole step
can be captured in a hipGraph (torch.cuda.CUDAGraph) and replayed.  Instances whose tables keep
terms outside the structured classes get those applied by one enumeration launch each after the
batch.

The autotuner picks two launch parameters: the number of persistent waves, and a split -- the
`split` instances with the most core-index entries run their own structured launches (the
single-set kernels, which fill the GPU alone and keep more registers than the batch kernel), the
rest share the batched launches (one plan per split, built on first use).
"""

from __future__ import annotations

import ctypes
import typing

import torch

from . import _lib
from . import prepared as pp
from .autotune import TUNER
from .ops import _ptr, _stream, _workspace

WAVE_CANDIDATES = (1024, 2048, 4096, 8192, 16384, 32768)
SPLIT_CANDIDATES = (0, 1, 2, 4, 8, 16)


class Batch:
    """Packed inputs/outputs + the device plan for a list of (terms, targets) instances."""

    def __init__(self, instances: typing.Sequence[tuple[pp.PreparedTerms, pp.PreparedTargets]], two_vectors: bool = False) -> None:
        """two_vectors=True (Hermitian instances only) adds a second packed input/output, so that
        launch_both() applies every instance to both in one pass: the forward and the backward of
        a training step."""
        if not instances:
            raise ValueError("a batch needs at least one instance")
        device = instances[0][1].device
        self.device = device
        self.instances = list(instances)
        self.rows = []
        off = 0
        for terms, targets in self.instances:
            if targets.device != device or terms.device != device:
                raise ValueError("all instances must live on one device")
            if terms.pairs_desc is None:
                raise ValueError("batched instances need the structured tables (a table with pair excitations, n_qubytes <= 16)")
            if targets.duplicates() != 0:
                raise ValueError("batched config sets must have distinct rows")
            targets.cores(2)
            if terms.singles_desc is not None:
                targets.cores(1)
            self.rows.append((off, off + targets.n))
            off += targets.n
        self.n_rows = off
        self.inp = torch.zeros((max(off, 1), 2), dtype=torch.float64, device=device)
        self.out = torch.zeros((max(off, 1), 2), dtype=torch.float64, device=device)
        self.hermitian = all(t.hermitian for t, _ in self.instances)
        self.two_vectors = two_vectors
        if two_vectors:
            if not self.hermitian:
                raise ValueError("a two-vector batch needs Hermitian instances (both vectors take the same map)")
            self.inp2 = torch.zeros_like(self.inp)
            self.out2 = torch.zeros_like(self.out)
        # instances by decreasing core-index entries (the work of their group walks)
        work = [targets.cores(2).capacity + (targets.cores(1).capacity if terms.singles_desc is not None else 0)
                for terms, targets in self.instances]
        self.order = sorted(range(len(self.instances)), key=lambda k: (-work[k], k))
        self._plans: dict[int, tuple[object, torch.Tensor, _lib.QmbBatch]] = {}
        self.split = 0
        self.plan = self._plan(0)
        self.waves = 0
        self.residual = [k for k, (t, _) in enumerate(self.instances) if t.n_residual_terms > 0]
        for k in self.residual:
            # the residual enumeration looks targets up: build their hash index now, outside any
            # graph capture of the launches
            _ = self.instances[k][1].desc

    def _plan(self, split: int) -> _lib.QmbBatch:
        """The persistent launch's plan over every instance but the `split` largest (built once)."""
        if split not in self._plans:
            members = sorted(self.order[split:])
            n = len(members)
            items = (_lib.QmbBatchItem * max(n, 1))()
            for j, k in enumerate(members):
                terms, targets = self.instances[k]
                lo = self.rows[k][0]
                it = items[j]
                it.terms = ctypes.pointer(terms.desc)
                it.pairs = ctypes.pointer(terms.pairs_desc)
                it.singles = ctypes.pointer(terms.singles_desc) if terms.singles_desc is not None else None
                it.targets = ctypes.pointer(targets.keys_desc)
                it.cores2 = ctypes.pointer(targets.cores(2))
                it.cores1 = ctypes.pointer(targets.cores(1)) if terms.singles_desc is not None else None
                it.in_ = self.inp.data_ptr() + 16 * lo
                it.out = self.out.data_ptr() + 16 * lo
                if self.two_vectors:
                    it.in2 = self.inp2.data_ptr() + 16 * lo
                    it.out2 = self.out2.data_ptr() + 16 * lo
            plan = _lib.QmbBatch()
            buffer = torch.empty(0, dtype=torch.uint8, device=self.device)
            if n > 0:
                lib = _lib.library()
                size = ctypes.c_size_t(0)
                _lib.check(lib.qmb_batch_workspace_size(n, ctypes.byref(size)), "qmb_batch_workspace_size")
                buffer = _workspace(size.value, self.device)
                with torch.cuda.device(self.device):
                    _lib.check(lib.qmb_batch_plan(items, n, _ptr(buffer), buffer.numel(), ctypes.byref(plan), _stream(self.device)),
                               "qmb_batch_plan")
            self._plans[split] = (items, buffer, plan)
        return self._plans[split][2]

    def _solo(self, mode: int, stream: ctypes.c_void_p) -> None:
        """The `split` largest instances through the single-set structured kernels."""
        lib = _lib.library()
        for k in self.order[:self.split]:
            terms, targets = self.instances[k]
            lo = self.rows[k][0]
            singles = ctypes.byref(terms.singles_desc) if terms.singles_desc is not None else None
            cores1 = ctypes.byref(targets.cores(1)) if terms.singles_desc is not None else None
            vin, vout = ctypes.c_void_p(self.inp.data_ptr() + 16 * lo), ctypes.c_void_p(self.out.data_ptr() + 16 * lo)
            if self.two_vectors:
                _lib.check(lib.qmb_apply_structured_dual(ctypes.byref(terms.pairs_desc), singles, ctypes.byref(targets.cores(2)), cores1,
                                                         ctypes.byref(targets.keys_desc), vin, vout,
                                                         ctypes.c_void_p(self.inp2.data_ptr() + 16 * lo),
                                                         ctypes.c_void_p(self.out2.data_ptr() + 16 * lo), mode, stream),
                           "qmb_apply_structured_dual")
            else:
                _lib.check(lib.qmb_apply_diagonal(ctypes.byref(terms.pairs_desc), ctypes.byref(targets.keys_desc), vin, vout, mode, stream),
                           "qmb_apply_diagonal")
                if terms.singles_desc is not None:
                    _lib.check(lib.qmb_apply_singles(ctypes.byref(terms.singles_desc), cores1, ctypes.byref(targets.keys_desc), vin, vout,
                                                     mode, stream), "qmb_apply_singles")
                _lib.check(lib.qmb_apply_pairs(ctypes.byref(terms.pairs_desc), ctypes.byref(targets.cores(2)),
                                               ctypes.byref(targets.keys_desc), vin, vout, mode, stream), "qmb_apply_pairs")

    def configure(self, split: int) -> None:
        """Runs the `split` largest instances by their own launches and the rest batched."""
        if not 0 <= split <= len(self.instances):
            raise ValueError(f"split must be in [0, {len(self.instances)}], got {split}")
        self.plan = self._plan(split)
        self.split = split

    # ------------------------------------------------------------------ inputs and outputs
    def set_inputs(self, psis: typing.Sequence[torch.Tensor], second: typing.Sequence[torch.Tensor] | None = None) -> None:
        """Copies each instance's complex128 vector into the packed input buffer (and, for a
        two-vector batch, the second vectors into the second buffer)."""
        for buf, vecs in ((self.inp, psis), (getattr(self, "inp2", None), second)):
            if vecs is None:
                continue
            if buf is None:
                raise ValueError("second vectors need a two-vector batch")
            if len(vecs) != len(self.instances):
                raise ValueError("one vector per instance")
            for (lo, hi), psi in zip(self.rows, vecs):
                if psi.dtype != torch.complex128 or psi.shape != (hi - lo,):
                    raise RuntimeError("each vector must be complex128 with one entry per config of its instance")
                buf[lo:hi].copy_(torch.view_as_real(psi.resolve_conj()))

    def output(self, k: int, second: bool = False) -> torch.Tensor:
        """Instance k's result as a complex128 view of the packed output buffer (or the second)."""
        lo, hi = self.rows[k]
        return torch.view_as_complex((self.out2 if second else self.out)[lo:hi])

    def input(self, k: int) -> torch.Tensor:
        lo, hi = self.rows[k]
        return torch.view_as_complex(self.inp[lo:hi])

    # ------------------------------------------------------------------------------ launch
    def launch_both(self, waves: int | None = None) -> None:
        """out = H in and out2 = H in2 for every (Hermitian) instance in one batched pass, plus
        the residual enumeration per instance that needs it.  Enqueue-only (graph capturable)."""
        if not self.two_vectors:
            raise RuntimeError("launch_both needs a batch built with two_vectors=True")
        self.out.zero_()
        self.out2.zero_()
        lib = _lib.library()
        stream = _stream(self.device)
        w = self.waves if waves is None else waves
        with torch.cuda.device(self.device):
            self._solo(pp.HERMITIAN_MODE, stream)
            _lib.check(lib.qmb_apply_batch(ctypes.byref(self.plan), pp.HERMITIAN_MODE, w, stream), "qmb_apply_batch")
            for k in self.residual:
                for inp, out in ((self.inp, self.out), (self.inp2, self.out2)):
                    self._residual(k, pp.HERMITIAN_MODE, inp, out, stream)

    def _residual(self, k: int, mode: int, inp: torch.Tensor, out: torch.Tensor, stream: ctypes.c_void_p) -> None:
        terms, targets = self.instances[k]
        lo, _ = self.rows[k]
        launch = pp._launch(0, 0, 0, self.hermitian)  # pylint: disable=protected-access
        _lib.check(_lib.library().qmb_apply_prepared(ctypes.byref(terms.residual_desc), ctypes.byref(targets.desc), _ptr(targets.configs),
                                                     targets.n, ctypes.c_void_p(inp.data_ptr() + 16 * lo),
                                                     ctypes.c_void_p(out.data_ptr() + 16 * lo), mode,
                                                     ctypes.byref(launch) if launch else None, stream), "qmb_apply_prepared")

    def launch(self, mode: int, waves: int | None = None) -> None:
        """out = A in (mode 0) or A^H in (mode 1) for every instance: one batched pass plus one
        enumeration launch per instance with residual terms.  Enqueue-only (graph capturable)."""
        if self.two_vectors:
            raise RuntimeError("a two-vector batch runs with launch_both()")
        self.out.zero_()
        lib = _lib.library()
        stream = _stream(self.device)
        pmode = pp.HERMITIAN_MODE if self.hermitian else mode
        w = self.waves if waves is None else waves
        with torch.cuda.device(self.device):
            self._solo(pmode, stream)
            _lib.check(lib.qmb_apply_batch(ctypes.byref(self.plan), pmode, w, stream), "qmb_apply_batch")
            for k in self.residual:
                self._residual(k, mode, self.inp, self.out, stream)

    def tune(self, retune: bool = False) -> int:
        """Picks the number of persistent waves and the split with the autotuner (runs the step on
        the current inputs; overwrites the outputs).  Returns the waves."""
        key = (f"batch:n={len(self.instances)}:rows={self.n_rows}:units={self._plan(0).total_units}:two={int(self.two_vectors)}"
               ":split")
        splits = [k for k in SPLIT_CANDIDATES if k < len(self.instances)]

        def run(c: dict[str, int]) -> None:
            self.configure(int(c["split"]))
            if self.two_vectors:
                self.launch_both(int(c["waves"]))
            else:
                self.launch(pp.MODE_FORWARD, int(c["waves"]))

        best = TUNER.best(key, [{"waves": w, "split": k} for k in splits for w in WAVE_CANDIDATES], run, retune=retune)
        self.configure(int(best["split"]))
        self.waves = int(best["waves"])
        return self.waves
