"""
The op layer on a host without a GPU: `ops.load` registers the reference's namespace with its four
schemas word for word (qmb/_hamiltonian.cpp:83-88), the GPU-only ops raise on CPU tensors, and
without the HIP library loading the ops raises.
"""

from __future__ import annotations

import os
import pathlib
import subprocess
import sys

import pytest
import torch

from qmb_hip import ops

# qmb/_hamiltonian.cpp:84-88, the reference's op definitions.
REFERENCE_SCHEMAS = {
    "apply_within": "apply_within(Tensor configs_i, Tensor psi_i, Tensor configs_j, Tensor site, Tensor kind, Tensor coef) -> Tensor",
    "find_relative": ("find_relative(Tensor configs_i, Tensor psi_i, int count_selected, Tensor site, Tensor kind, Tensor coef, "
                      "Tensor configs_exclude) -> Tensor"),
    "diagonal_term": "diagonal_term(Tensor configs, Tensor site, Tensor kind, Tensor coef) -> Tensor",
    "single_relative": "single_relative(Tensor configs, Tensor site, Tensor kind, Tensor coef) -> Tensor",
}


@pytest.mark.parametrize("nq,pc", [(7, 1), (3, 2)])
def test_namespace_and_reference_schemas(nq: int, pc: int) -> None:
    ns = ops.load(nq, pc)
    assert ns is getattr(torch.ops, f"qmb_hamiltonian_{nq}_{pc}")
    assert ops.load(nq, pc) is ns  # registered once
    for name, schema in REFERENCE_SCHEMAS.items():
        assert str(getattr(ns, name).default._schema) == f"qmb_hamiltonian_{nq}_{pc}::{schema}"  # pylint: disable=protected-access


def test_gpu_only_ops_have_no_cpu_kernel() -> None:
    """diagonal_term and single_relative exist for the GPU only, as in the reference
    (qmb/_hamiltonian_cuda.cu:1008-1013); apply_within and find_relative run on CPU tensors
    (tests/test_cpu_ops.py)."""
    ns = ops.load(1, 1)
    configs = torch.tensor([[0b011], [0b101]], dtype=torch.uint8)
    site = torch.tensor([[0, 1, 0, 0]], dtype=torch.int16)
    kind = torch.tensor([[1, 0, 2, 2]], dtype=torch.uint8)
    coef = torch.tensor([[1.0, 0.0]], dtype=torch.float64)
    with pytest.raises(NotImplementedError):
        ns.diagonal_term(configs, site, kind, coef)
    with pytest.raises(NotImplementedError):
        ns.single_relative(configs, site, kind, coef)


def test_load_rejects_unsupported_shapes() -> None:
    with pytest.raises(ValueError, match="particle_cut"):
        ops.load(7, 3)
    with pytest.raises(ValueError, match="n_qubytes"):
        ops.load(0, 1)
    with pytest.raises(ValueError, match="n_qubytes"):
        ops.load(33, 1)


def test_missing_library_fails_loudly(tmp_path: pathlib.Path) -> None:
    """With the HIP library absent, loading the ops raises instead of falling back to any other
    path (run in a fresh process: the library is loaded once per process)."""
    root = pathlib.Path(__file__).resolve().parent.parent
    env = dict(os.environ, QMB_HIP_LIBRARY=str(tmp_path / "absent.so"),
               PYTHONPATH=os.pathsep.join([str(root / "qmb-public_amd"), os.environ.get("PYTHONPATH", "")]))
    code = "from qmb_hip import ops\nops.load(7, 1)\n"
    proc = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120, check=False)
    assert proc.returncode != 0
    assert "ImportError" in proc.stderr and "no non-HIP fallback" in proc.stderr, proc.stderr[-2000:]
