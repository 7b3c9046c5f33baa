"""
Vectorised brute force over (term, config) pairs for test-side checks of selection rules.

Applies each term to all configs at once on 64-bit words, following qmb/_hamiltonian_cpu.cpp:48-78
(operators right to left, bit test, set, parity from the occupied sites below the operator's site).
Test infrastructure only.
"""

from __future__ import annotations

import numpy as np


def to_words(configs: np.ndarray) -> np.ndarray:
    """uint8 [N, nq] -> uint64 [N, W] little-endian words."""
    nq = configs.shape[1]
    w = (nq + 7) // 8
    padded = np.zeros((len(configs), 8 * w), dtype=np.uint8)
    padded[:, :nq] = configs
    return padded.view("<u8").reshape(len(configs), w).astype(np.uint64)


def from_words(words: np.ndarray, nq: int) -> np.ndarray:
    return np.ascontiguousarray(words.astype("<u8")).view(np.uint8).reshape(len(words), -1)[:, :nq]


def apply_all(words: np.ndarray, site4: np.ndarray, kind4: np.ndarray, particle_cut: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """One term on every config: (success mask, resulting words, parity bits)."""
    cur = words.copy()
    ok = np.ones(len(words), dtype=bool)
    par = np.zeros(len(words), dtype=np.int64)
    for op in range(3, -1, -1):
        k = int(kind4[op])
        if k == 2:
            continue
        s = int(site4[op])
        wi, b = s // 64, np.uint64(s % 64)
        bit = (cur[:, wi] >> b) & np.uint64(1)
        ok &= bit != k
        if k:
            cur[:, wi] |= np.uint64(1) << b
        else:
            cur[:, wi] &= ~(np.uint64(1) << b)
        if particle_cut == 1:
            for w in range(cur.shape[1]):
                lo = s - 64 * w
                if lo <= 0:
                    continue
                mask = np.uint64(0xFFFFFFFFFFFFFFFF) if lo >= 64 else (np.uint64(1) << np.uint64(lo)) - np.uint64(1)
                par ^= np.bitwise_count(cur[:, w] & mask).astype(np.int64) & 1
    return ok, cur, par


def contribution_weights(pc: int, configs: np.ndarray, psi: np.ndarray, exclude: np.ndarray, site: np.ndarray, kind: np.ndarray,
                         coef: np.ndarray) -> list[tuple[float, bytes]]:
    """(|sign coef psi|^2, target bytes) of every contribution whose target is outside `exclude`."""
    nq = configs.shape[1]
    words = to_words(configs)
    ex = {bytes(r) for r in exclude}
    out: list[tuple[float, bytes]] = []
    for t in range(len(site)):
        ok, res, par = apply_all(words, site[t], kind[t], pc)
        idx = np.nonzero(ok)[0]
        if len(idx) == 0:
            continue
        rows = from_words(res[idx], nq)
        sign = np.where(par[idx] == 1, -1.0, 1.0)
        re = sign * (coef[t, 0] * psi[idx].real - coef[t, 1] * psi[idx].imag)
        im = sign * (coef[t, 0] * psi[idx].imag + coef[t, 1] * psi[idx].real)
        w = re * re + im * im
        for j in range(len(idx)):
            key = rows[j].tobytes()
            if key not in ex:
                out.append((float(w[j]), key))
    return out
