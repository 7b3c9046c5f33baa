"""
Condition-aware error bounds for comparing two FP64 applications of one term table.

For an entry y_j = sum over contributions s_t c_t x_i, two summation orders (or merged
coefficients) differ by a few ulps of sum |c_t| |x_i| over the same contributions, not of |y_j|,
which cancellation can make arbitrarily small.  That sum is itself an application: the table with
coefficients |c_t| applied to |x| without signs, i.e. with particle_cut 2 (hard-core bosons: the
same success conditions and targets, no Jordan-Wigner sign).  It is computed here by the library's
enumeration kernel (every connection looked up, no merged coefficients).  Test infrastructure.
"""

from __future__ import annotations

import numpy as np
import torch

from qmb_hip import prepared as pp


def abs_terms(site: np.ndarray, kind: np.ndarray, coef: np.ndarray, n_qubytes: int, device: torch.device) -> pp.PreparedTerms:
    """The table with coefficients (|c_t|, 0) and particle_cut 2, for the enumeration kernel."""
    magnitude = np.zeros_like(np.asarray(coef, dtype=np.float64))
    magnitude[:, 0] = np.hypot(coef[:, 0], coef[:, 1])
    dev = [torch.from_numpy(np.ascontiguousarray(x)).to(device) for x in (site, kind, magnitude)]
    return pp.PreparedTerms(*dev, n_qubytes, 2, check_hermitian=False, pairs=False)


def abs_bound(terms_abs: pp.PreparedTerms, targets: pp.PreparedTargets, configs: torch.Tensor, vec: np.ndarray, mode: int) -> np.ndarray:
    """sum |c_t| |vec_i| per output entry of mode 0 (forward) or mode 1 (adjoint)."""
    mag = np.zeros((len(vec), 2), dtype=np.float64)
    mag[:, 0] = np.abs(vec)
    out = pp.apply(terms_abs, targets, configs, torch.from_numpy(mag).to(configs.device), mode, symmetric=False)
    return out[:, 0].cpu().numpy()
