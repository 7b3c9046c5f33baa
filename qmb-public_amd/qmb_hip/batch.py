"""
Batched application of many Hamiltonians to many config sets in one persistent launch.

Each instance is a PreparedTerms (with the structured tables: pair and single excitations,
diagonal records) and a PreparedTargets over its distinct config set.  The inputs and outputs of all
instances are packed into two device buffers, so the plan (qmb_batch_plan) is built once and every
launch (qmb_apply_batch) is a single persistent kernel over all instances' work; the whole step
can be captured in a hipGraph (torch.cuda.CUDAGraph) and replayed.  The number of persistent
waves is picked by the autotuner.  Instances whose tables keep terms outside the structured
classes get those applied by one enumeration launch each after the batch.
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
        n = len(self.instances)
        self._items = (_lib.QmbBatchItem * n)()
        for k, (terms, targets) in enumerate(self.instances):
            lo = self.rows[k][0]
            it = self._items[k]
            it.terms = ctypes.pointer(terms.desc)
            it.pairs = ctypes.pointer(terms.pairs_desc)
            it.singles = ctypes.pointer(terms.singles_desc) if terms.singles_desc is not None else None
            it.targets = ctypes.pointer(targets.keys_desc)
            it.cores2 = ctypes.pointer(targets.cores(2))
            it.cores1 = ctypes.pointer(targets.cores(1)) if terms.singles_desc is not None else None
            it.in_ = self.inp.data_ptr() + 16 * lo
            it.out = self.out.data_ptr() + 16 * lo
            if two_vectors:
                it.in2 = self.inp2.data_ptr() + 16 * lo
                it.out2 = self.out2.data_ptr() + 16 * lo
        lib = _lib.library()
        size = ctypes.c_size_t(0)
        _lib.check(lib.qmb_batch_workspace_size(n, ctypes.byref(size)), "qmb_batch_workspace_size")
        self.buffer = _workspace(size.value, device)
        self.plan = _lib.QmbBatch()
        with torch.cuda.device(device):
            _lib.check(lib.qmb_batch_plan(self._items, n, _ptr(self.buffer), self.buffer.numel(), ctypes.byref(self.plan),
                                          _stream(device)), "qmb_batch_plan")
        self.waves = 0
        self.residual = [k for k, (t, _) in enumerate(self.instances) if t.n_residual_terms > 0]
        for k in self.residual:
            # the residual enumeration looks targets up: build their hash index now, outside any
            # graph capture of the launches
            _ = self.instances[k][1].desc

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
        """out = H in and out2 = H in2 for every (Hermitian) instance in ONE persistent launch, plus
        the residual enumeration per instance that needs it.  Enqueue-only (graph capturable)."""
        if not self.two_vectors:
            raise RuntimeError("launch_both needs a batch built with two_vectors=True")
        self.out.zero_()
        self.out2.zero_()
        lib = _lib.library()
        stream = _stream(self.device)
        w = self.waves if waves is None else waves
        with torch.cuda.device(self.device):
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
        """out = A in (mode 0) or A^H in (mode 1) for every instance: one persistent launch plus one
        enumeration launch per instance with residual terms.  Enqueue-only (graph capturable)."""
        if self.two_vectors:
            raise RuntimeError("a two-vector batch runs with launch_both()")
        self.out.zero_()
        lib = _lib.library()
        stream = _stream(self.device)
        pmode = pp.HERMITIAN_MODE if self.hermitian else mode
        w = self.waves if waves is None else waves
        with torch.cuda.device(self.device):
            _lib.check(lib.qmb_apply_batch(ctypes.byref(self.plan), pmode, w, stream), "qmb_apply_batch")
            for k in self.residual:
                self._residual(k, mode, self.inp, self.out, stream)

    def tune(self, retune: bool = False) -> int:
        """Picks the number of persistent waves with the autotuner (runs the batch on the current
        inputs; overwrites the outputs)."""
        key = f"batch:n={len(self.instances)}:rows={self.n_rows}:units={self.plan.total_units}:two={int(self.two_vectors)}"
        run = (lambda c: self.launch_both(c["waves"])) if self.two_vectors else (lambda c: self.launch(pp.MODE_FORWARD, c["waves"]))
        best = TUNER.best(key, [{"waves": w} for w in WAVE_CANDIDATES], run, retune=retune)
        self.waves = int(best["waves"])
        return self.waves
