"""
ctypes/numpy wrapper of the CPU oracle (oracle/qmb_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, never by the product package.  Each function mirrors one reference function; see the C file's
header for the citations and for what pins each function.
"""

from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

ORACLE_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "_build" / "libqmb_oracle.so"
REF_DIR = ORACLE_DIR / "_ref"

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_LIB: ctypes.CDLL | None = None


def build() -> None:
    """Compiles the oracle (and, where the reference sources exist, the reference CPU op)."""
    subprocess.run(["make", "-C", str(ORACLE_DIR), "-s"], check=True)


def lib() -> ctypes.CDLL:
    global _LIB  # pylint: disable=global-statement
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.oracle_apply_term.argtypes = [_vp, _vp, _vp, _int, ctypes.POINTER(_int)]
        L.oracle_apply_term.restype = _int
        L.oracle_apply_within.argtypes = [_int, _int, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp]
        L.oracle_apply_within.restype = _int
        L.oracle_find_relative.argtypes = [_int, _int, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _i64, _vp,
                                           ctypes.POINTER(ctypes.c_double)]
        L.oracle_find_relative.restype = _i64
        L.oracle_diagonal_term.argtypes = [_int, _int, _vp, _i64, _vp, _vp, _vp, _i64, _vp]
        L.oracle_diagonal_term.restype = _int
        L.oracle_single_relative.argtypes = [_int, _int, _vp, _i64, _vp, _vp, _vp, _i64, _vp]
        L.oracle_single_relative.restype = _int
        L.oracle_count_connections.argtypes = [_int, _int, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp]
        L.oracle_count_connections.restype = _int
        _LIB = L
    return _LIB


def _c(a: np.ndarray, dtype: object) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def _terms(site: object, kind: object, coef: object) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    return (_c(np.asarray(site), np.int16).reshape(-1, 4), _c(np.asarray(kind), np.uint8).reshape(-1, 4),
            _c(np.asarray(coef), np.float64).reshape(-1, 2))


def _psi(psi: object) -> np.ndarray:
    a = np.asarray(psi)
    if np.iscomplexobj(a):
        a = np.stack([a.real, a.imag], axis=-1)
    return _c(a, np.float64).reshape(-1, 2)


def apply_term(config: np.ndarray, site4: object, kind4: object, particle_cut: int) -> tuple[bool, np.ndarray, int]:
    """qmb/_hamiltonian_cpu.cpp:48-78 on one config: (success, resulting config, parity)."""
    cur = _c(config, np.uint8).copy()
    s = _c(site4, np.int16)
    k = _c(kind4, np.uint8)
    par = ctypes.c_int(0)
    ok = lib().oracle_apply_term(_p(cur), _p(s), _p(k), particle_cut, ctypes.byref(par))
    return bool(ok), cur, int(par.value)


def apply_within(particle_cut: int, configs_i: np.ndarray, psi_i: object, configs_j: np.ndarray, site: object, kind: object,
                 coef: object) -> tuple[np.ndarray, np.ndarray, tuple[int, int]]:
    """Returns (psi_j complex128 [n_j], sum |contribution| per target, (M_gen, M_hit))."""
    ci = _c(configs_i, np.uint8)
    cj = _c(configs_j, np.uint8)
    nq = ci.shape[1] if ci.ndim == 2 and ci.shape[0] else cj.shape[1]
    pi = _psi(psi_i)
    s, k, c = _terms(site, kind, coef)
    out = np.zeros((len(cj), 2), dtype=np.float64)
    asum = np.zeros(len(cj), dtype=np.float64)
    counts = np.zeros(2, dtype=np.int64)
    rc = lib().oracle_apply_within(nq, particle_cut, _p(ci), _p(pi), len(ci), _p(cj), len(cj), _p(s), _p(k), _p(c), len(s),
                                   _p(out), _p(asum), _p(counts))
    if rc != 0:
        raise ValueError("oracle_apply_within rejected its arguments")
    return out[:, 0] + 1j * out[:, 1], asum, (int(counts[0]), int(counts[1]))


def apply_within_threads(particle_cut: int, configs_i: np.ndarray, psi_i: object, configs_j: np.ndarray, site: object,
                         kind: object, coef: object, threads: int) -> tuple[np.ndarray, tuple[int, int]]:
    """apply_within with configs_i split over `threads` host threads (timing only; no reference
    counterpart).  Returns (psi_j complex128 [n_j], (M_gen, M_hit))."""
    L = lib()
    fn = L.oracle_apply_within_threads
    fn.argtypes = [_int, _int, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _i64, _int, _vp, _vp]
    fn.restype = _int
    ci = _c(configs_i, np.uint8)
    cj = _c(configs_j, np.uint8)
    nq = ci.shape[1] if ci.ndim == 2 and ci.shape[0] else cj.shape[1]
    pi = _psi(psi_i)
    s, k, c = _terms(site, kind, coef)
    out = np.zeros((len(cj), 2), dtype=np.float64)
    counts = np.zeros(2, dtype=np.int64)
    rc = fn(nq, particle_cut, _p(ci), _p(pi), len(ci), _p(cj), len(cj), _p(s), _p(k), _p(c), len(s), int(threads), _p(out), _p(counts))
    if rc != 0:
        raise ValueError(f"oracle_apply_within_threads failed ({rc})")
    return out[:, 0] + 1j * out[:, 1], (int(counts[0]), int(counts[1]))


def find_relative(particle_cut: int, configs_i: np.ndarray, psi_i: object, count_selected: int, site: object, kind: object,
                  coef: object, configs_exclude: np.ndarray | None = None) -> tuple[np.ndarray, float]:
    """Returns (unique selected configs sorted by bytes, K-th entry weight of the final heap)."""
    ci = _c(configs_i, np.uint8)
    cx = ci if configs_exclude is None else _c(configs_exclude, np.uint8)
    nq = ci.shape[1]
    pi = _psi(psi_i)
    s, k, c = _terms(site, kind, coef)
    out = np.zeros((count_selected, nq), dtype=np.uint8)
    kth = ctypes.c_double(0.0)
    n = lib().oracle_find_relative(nq, particle_cut, _p(ci), _p(pi), len(ci), count_selected, _p(s), _p(k), _p(c), len(s), _p(cx),
                                   len(cx), _p(out), ctypes.byref(kth))
    if n < 0:
        raise ValueError("oracle_find_relative rejected its arguments")
    return out[:n].copy(), float(kth.value)


def diagonal_term(particle_cut: int, configs: np.ndarray, site: object, kind: object, coef: object) -> np.ndarray:
    cf = _c(configs, np.uint8)
    s, k, c = _terms(site, kind, coef)
    out = np.zeros((len(cf), 2), dtype=np.float64)
    if lib().oracle_diagonal_term(cf.shape[1], particle_cut, _p(cf), len(cf), _p(s), _p(k), _p(c), len(s), _p(out)) != 0:
        raise ValueError("oracle_diagonal_term rejected its arguments")
    return out[:, 0] + 1j * out[:, 1]


def single_relative(particle_cut: int, configs: np.ndarray, site: object, kind: object, keys: np.ndarray) -> np.ndarray:
    cf = _c(configs, np.uint8)
    s, k, _ = _terms(site, kind, np.zeros((len(np.asarray(site).reshape(-1, 4)), 2)))
    kk = _c(keys, np.float64)
    out = np.zeros_like(cf)
    if lib().oracle_single_relative(cf.shape[1], particle_cut, _p(cf), len(cf), _p(s), _p(k), _p(kk), len(s), _p(out)) != 0:
        raise ValueError("oracle_single_relative rejected its arguments")
    return out


def best_weights(particle_cut: int, configs_i: np.ndarray, psi_i: object, site: object, kind: object, coef: object,
                 configs_t: np.ndarray) -> np.ndarray:
    """Per row of configs_t, the largest single-contribution weight |sign coef psi|^2 landing on it
    (-1 where none does)."""
    L = lib()
    fn = L.oracle_best_weights
    fn.argtypes = [_int, _int, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _i64, _vp]
    fn.restype = _int
    ci = _c(configs_i, np.uint8)
    ct = _c(configs_t, np.uint8)
    pi = _psi(psi_i)
    s, k, c = _terms(site, kind, coef)
    best = np.zeros(len(ct), dtype=np.float64)
    if fn(ct.shape[1], particle_cut, _p(ci), _p(pi), len(ci), _p(s), _p(k), _p(c), len(s), _p(ct), len(ct), _p(best)) != 0:
        raise ValueError("oracle_best_weights rejected its arguments")
    return best


def count_connections(particle_cut: int, configs_i: np.ndarray, configs_j: np.ndarray, site: object, kind: object) -> tuple[int, int]:
    ci = _c(configs_i, np.uint8)
    cj = _c(configs_j, np.uint8)
    s, k, _ = _terms(site, kind, np.zeros((len(np.asarray(site).reshape(-1, 4)), 2)))
    counts = np.zeros(2, dtype=np.int64)
    lib().oracle_count_connections(ci.shape[1], particle_cut, _p(ci), len(ci), _p(cj), len(cj), _p(s), _p(k), len(s), _p(counts))
    return int(counts[0]), int(counts[1])


def ref_library(n_qubytes: int, particle_cut: int) -> pathlib.Path:
    """Path of the compiled reference CPU op for (nq, pc) (built by `make -C oracle ref`)."""
    return REF_DIR / f"qmb_ref_{n_qubytes}_{particle_cut}.so"
