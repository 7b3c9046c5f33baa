"""
ctypes binding of libqmb_hip.so, the C ABI declared in include/qmb_hip.h.

The library is built in-tree (``make -C qmb-public_amd/csrc``, or ``__graft_entry__.build()``) and
loaded from this directory.  There is no fallback: if the library is missing or fails to load, every
entry point raises, so a GPU run can never silently compute on another path.
"""

from __future__ import annotations

import ctypes
import os
import pathlib

_HERE = pathlib.Path(__file__).resolve().parent
LIBRARY_PATH = pathlib.Path(os.environ.get("QMB_HIP_LIBRARY") or str(_HERE / "libqmb_hip.so"))

QMB_OK = 0
QMB_ERR_ARGUMENT = 1
QMB_ERR_HIP = 2
QMB_ERR_WORKSPACE = 3
QMB_ERR_UNSUPPORTED = 4

_c_int = ctypes.c_int
_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_u64 = ctypes.c_uint64
_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p


class QmbTerms(ctypes.Structure):
    """Mirror of ``struct qmb_terms`` (include/qmb_hip.h)."""

    _fields_ = [
        ("n_qubytes", _c_i32),
        ("particle_cut", _c_i32),
        ("n_words", _c_i32),
        ("n_bits", _c_i32),
        ("n_terms", _c_i64),
        ("n_buckets", _c_i64),
        ("buffer", _vp),
        ("bytes", _c_size),
        ("recs", _vp),
        ("coefs", _vp),
        ("tidx", _vp),
        ("bucket_start", _vp),
        ("drecs", _vp),
        ("dcoefs", _vp),
        ("counts", _vp),
    ]


class QmbTargets(ctypes.Structure):
    """Mirror of ``struct qmb_targets`` (include/qmb_hip.h)."""

    _fields_ = [
        ("n_qubytes", _c_i32),
        ("n_words", _c_i32),
        ("n", _c_i64),
        ("capacity", _c_u64),
        ("filter_words", _c_u64),
        ("buffer", _vp),
        ("bytes", _c_size),
        ("keys", _vp),
        ("table", _vp),
        ("filter", _vp),
        ("cols", _vp),
        ("duplicates", _vp),
    ]


class QmbPairs(ctypes.Structure):
    """Mirror of ``struct qmb_pairs`` (include/qmb_hip.h)."""

    _fields_ = [
        ("n_qubytes", _c_i32),
        ("particle_cut", _c_i32),
        ("n_bits", _c_i32),
        ("real", _c_i32),
        ("n_pairs", _c_i64),
        ("buffer", _vp),
        ("bytes", _c_size),
        ("coef", _vp),
        ("coef_re", _vp),
        ("diag", _vp),
        ("counts", _vp),
    ]


class QmbSingles(ctypes.Structure):
    """Mirror of ``struct qmb_singles`` (include/qmb_hip.h)."""

    _fields_ = [
        ("n_qubytes", _c_i32),
        ("particle_cut", _c_i32),
        ("n_bits", _c_i32),
        ("real", _c_i32),
        ("buffer", _vp),
        ("bytes", _c_size),
        ("a", _vp),
        ("kd", _vp),
        ("counts", _vp),
        ("a_re", _vp),
        ("kd_re", _vp),
    ]


class QmbCores(ctypes.Structure):
    """Mirror of ``struct qmb_cores`` (include/qmb_hip.h)."""

    _fields_ = [
        ("n_qubytes", _c_i32),
        ("n_words", _c_i32),
        ("max_occupied", _c_i32),
        ("per_config", _c_i32),
        ("order", _c_i32),
        ("key_bits", _c_i32),
        ("n", _c_i64),
        ("capacity", _c_i64),
        ("buffer", _vp),
        ("bytes", _c_size),
        ("keys", _vp),
        ("vals", _vp),
        ("starts", _vp),
        ("counts", _vp),
    ]


class QmbStrings(ctypes.Structure):
    """Mirror of ``struct qmb_strings`` (include/qmb_hip.h)."""

    _fields_ = [
        ("n_qubytes", _c_i32),
        ("n_words", _c_i32),
        ("parity", _c_i32),
        ("pad", _c_i32),
        ("n", _c_i64),
        ("buffer", _vp),
        ("bytes", _c_size),
        ("keys", _vp),
        ("rows", _vp),
        ("group", _vp),
        ("starts", _vp),
        ("mrows", _vp),
        ("mrange", _vp),
        ("mcfg", _vp),
        ("mstr", _vp),
        ("counts", _vp),
    ]


class QmbBatchItem(ctypes.Structure):
    """Mirror of ``struct qmb_batch_item`` (include/qmb_hip.h)."""

    _fields_ = [
        ("terms", ctypes.POINTER(QmbTerms)),
        ("pairs", ctypes.POINTER(QmbPairs)),
        ("singles", ctypes.POINTER(QmbSingles)),
        ("targets", ctypes.POINTER(QmbTargets)),
        ("cores2", ctypes.POINTER(QmbCores)),
        ("cores1", ctypes.POINTER(QmbCores)),
        ("in_", _vp),
        ("out", _vp),
        ("in2", _vp),
        ("out2", _vp),
    ]


class QmbBatch(ctypes.Structure):
    """Mirror of ``struct qmb_batch`` (include/qmb_hip.h)."""

    _fields_ = [
        ("n_items", _c_i32),
        ("dual", _c_i32),
        ("total_units", _c_i64),
        ("buffer", _vp),
        ("bytes", _c_size),
        ("items", _vp),
        ("offsets", _vp),
    ]


class QmbLaunch(ctypes.Structure):
    """Mirror of ``struct qmb_launch`` (include/qmb_hip.h)."""

    _fields_ = [("tile", _c_i32), ("groups", _c_i32), ("strategy", _c_i32), ("symmetric", _c_i32)]


# name -> (restype, argtypes); kept in the order of include/qmb_hip.h
SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "qmb_last_error": (ctypes.c_char_p, []),
    "qmb_version": (_c_int, []),
    "qmb_validate_terms_host": (_c_int, [_c_int, _vp, _vp, _c_i64, ctypes.POINTER(_c_i64)]),
    "qmb_compile_term_host": (
        _c_int,
        [_vp, _vp, _c_int, _c_int, ctypes.POINTER(_c_u64), ctypes.POINTER(_c_i32), ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i32)],
    ),
    "qmb_apply_record_host": (_c_int, [_c_u64, _c_int, _vp, _vp, ctypes.POINTER(_c_i32)]),
    "qmb_terms_hermitian_host": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _c_i64, ctypes.c_double, ctypes.POINTER(_c_i32)]),
    "qmb_philox4x32_10_host": (None, [_vp, _vp, _vp]),
    "qmb_philox_uniform_host": (ctypes.c_double, [_c_u64, _c_u64]),
    "qmb_apply_within_workspace_size": (_c_int, [_c_int, _c_i64, _c_i64, _c_i64, ctypes.POINTER(_c_size)]),
    "qmb_apply_within_structured_workspace_size": (_c_int, [_c_int, _c_i64, _c_i64, _c_i32, ctypes.POINTER(_c_size)]),
    "qmb_apply_within_hops_workspace_size": (_c_int, [_c_int, _c_i64, _c_i64, ctypes.POINTER(_c_size)]),
    "qmb_apply_within": (_c_int, [_c_int, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _vp, _c_size, _vp]),
    "qmb_find_relative_workspace_size": (_c_int, [_c_int, _c_i64, _c_i64, _c_i64, _c_i64, ctypes.POINTER(_c_size)]),
    "qmb_find_relative": (
        _c_int,
        [_c_int, _c_int, _vp, _vp, _c_i64, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _c_i64, _vp, _vp, _vp, _c_size, _vp],
    ),
    "qmb_diagonal_term_workspace_size": (_c_int, [_c_int, _c_i64, _c_i64, ctypes.POINTER(_c_size)]),
    "qmb_diagonal_term": (_c_int, [_c_int, _c_int, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _vp, _c_size, _vp]),
    "qmb_single_relative_workspace_size": (_c_int, [_c_int, _c_i64, _c_i64, ctypes.POINTER(_c_size)]),
    "qmb_single_relative": (_c_int, [_c_int, _c_int, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _c_u64, _vp, _vp, _vp, _c_size, _vp]),
    "qmb_apply_within_adjoint": (_c_int, [_c_int, _c_int, _vp, _c_i64, _vp, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _vp, _c_size, _vp]),
    "qmb_static_scratch_bytes": (_c_size, []),
    "qmb_apply_within_noalloc": (_c_int, [_c_int, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _vp]),
    "qmb_apply_within_plain_workspace_size": (_c_int, [_c_int, _c_i64, ctypes.POINTER(_c_size)]),
    "qmb_apply_within_plain": (_c_int, [_c_int, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _vp, _c_size, _vp]),
    "qmb_terms_workspace_size": (_c_int, [_c_int, _c_i64, ctypes.POINTER(_c_size)]),
    "qmb_terms_build": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _c_i64, _vp, _c_size, ctypes.POINTER(QmbTerms), _vp]),
    "qmb_targets_workspace_size": (_c_int, [_c_int, _c_i64, ctypes.POINTER(_c_size)]),
    "qmb_targets_build": (_c_int, [_c_int, _vp, _c_i64, _vp, _c_size, ctypes.POINTER(QmbTargets), _vp]),
    "qmb_targets_build_keys": (_c_int, [_c_int, _vp, _c_i64, _vp, _c_size, ctypes.POINTER(QmbTargets), _vp]),
    "qmb_targets_build_index": (_c_int, [ctypes.POINTER(QmbTargets), _vp]),
    "qmb_targets_lookup": (_c_int, [ctypes.POINTER(QmbTargets), _vp, _c_i64, _vp, _vp]),
    "qmb_apply_prepared": (
        _c_int,
        [ctypes.POINTER(QmbTerms), ctypes.POINTER(QmbTargets), _vp, _c_i64, _vp, _vp, _c_int, ctypes.POINTER(QmbLaunch), _vp],
    ),
    "qmb_pairs_workspace_size": (_c_int, [_c_int, ctypes.POINTER(_c_size)]),
    "qmb_pairs_build": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _c_i64, _vp, _c_size, ctypes.POINTER(QmbPairs), _vp]),
    "qmb_pairs_set_real": (_c_int, [ctypes.POINTER(QmbPairs), _c_i32, _vp]),
    "qmb_singles_set_real": (_c_int, [ctypes.POINTER(QmbSingles), _c_i32, _vp]),
    "qmb_singles_workspace_size": (_c_int, [_c_int, ctypes.POINTER(_c_size)]),
    "qmb_singles_build": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _c_i64, _vp, _c_size, ctypes.POINTER(QmbSingles), _vp]),
    "qmb_terms_build_residual": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _c_i64, _c_i32, _vp, _c_size, ctypes.POINTER(QmbTerms), _vp]),
    "qmb_cores_workspace_size": (_c_int, [_c_int, _c_i64, _c_i32, _c_i32, ctypes.POINTER(_c_size)]),
    "qmb_cores_build": (_c_int, [ctypes.POINTER(QmbTargets), _c_i32, _c_i32, _vp, _c_size, ctypes.POINTER(QmbCores), _vp]),
    "qmb_apply_pairs": (
        _c_int,
        [ctypes.POINTER(QmbPairs), ctypes.POINTER(QmbCores), ctypes.POINTER(QmbTargets), _vp, _vp, _c_int, _vp],
    ),
    "qmb_apply_singles": (
        _c_int,
        [ctypes.POINTER(QmbSingles), ctypes.POINTER(QmbCores), ctypes.POINTER(QmbTargets), _vp, _vp, _c_int, _vp],
    ),
    "qmb_apply_diagonal": (_c_int, [ctypes.POINTER(QmbPairs), ctypes.POINTER(QmbTargets), _vp, _vp, _c_int, _vp]),
    "qmb_strings_workspace_size": (_c_int, [_c_int, _c_i64, ctypes.POINTER(_c_size)]),
    "qmb_strings_build": (_c_int, [ctypes.POINTER(QmbTargets), _c_i32, _vp, _c_size, ctypes.POINTER(QmbStrings), _vp]),
    "qmb_apply_hops": (
        _c_int,
        [ctypes.POINTER(QmbSingles), ctypes.POINTER(QmbStrings), ctypes.POINTER(QmbTargets), _vp, _vp, _c_int, _vp],
    ),
    "qmb_apply_diagonal_terms": (_c_int, [ctypes.POINTER(QmbTerms), ctypes.POINTER(QmbTargets), _vp, _vp, _c_int, _vp]),
    "qmb_diagonal_values": (_c_int, [ctypes.POINTER(QmbTerms), ctypes.POINTER(QmbTargets), _vp, _vp]),
    "qmb_apply_diagonal_values": (_c_int, [_c_i64, _vp, _vp, _vp, _c_int, _vp]),
    "qmb_apply_structured_dual": (
        _c_int,
        [ctypes.POINTER(QmbPairs), ctypes.POINTER(QmbSingles), ctypes.POINTER(QmbCores), ctypes.POINTER(QmbCores),
         ctypes.POINTER(QmbTargets), _vp, _vp, _vp, _vp, _c_int, _vp],
    ),
    "qmb_batch_workspace_size": (_c_int, [_c_int, ctypes.POINTER(_c_size)]),
    "qmb_batch_plan": (_c_int, [ctypes.POINTER(QmbBatchItem), _c_int, _vp, _c_size, ctypes.POINTER(QmbBatch), _vp]),
    "qmb_apply_batch": (_c_int, [ctypes.POINTER(QmbBatch), _c_int, _c_int, _vp]),
    "qmb_sampler_children": (_c_int, [_c_int, _c_i64, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _c_u64, _vp, _vp, _vp]),
    "qmb_sampler_select": (_c_int, [_c_int, _c_i64, _vp, _c_i64, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "qmb_sampler_prefix_inputs": (_c_int, [_c_int, _c_i64, _c_int, _vp, _c_int, _vp, _vp]),
    "qmb_sampler_prefix_configs": (_c_int, [_c_i64, _c_int, _vp, _c_int, _vp, _c_int, _vp, _vp]),
    "qmb_network_inputs": (_c_int, [_c_int, _c_i64, _c_int, _vp, _c_int, _vp, _vp, _vp]),
    "qmb_network_amplitude": (_c_int, [_c_int, _c_i64, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "qmb_network_fused_supported": (_c_int, [_c_int, _c_int]),
    "qmb_network_forward_fused": (_c_int, [_c_i64, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "qmb_sampler_step_fused": (_c_int, [_c_i64, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_u64, _vp, _vp, _vp]),
    "qmb_lanczos_workspace_size": (_c_int, [ctypes.POINTER(_c_size)]),
    "qmb_lanczos_step": (_c_int, [_c_i64, _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "qmb_cpu_apply_within": (_c_int, [_c_int, _c_int, _vp, _vp, _c_i64, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _c_int]),
    "qmb_cpu_apply_within_adjoint": (_c_int, [_c_int, _c_int, _vp, _c_i64, _vp, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _c_int]),
    "qmb_cpu_find_relative": (
        _c_int,
        [_c_int, _c_int, _vp, _vp, _c_i64, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _c_i64, _vp, ctypes.POINTER(_c_i64), _c_int],
    ),
    "qmb_pair_excitation_host": (_c_int, [_c_u64, _c_int, ctypes.POINTER(_c_i32)]),
    "qmb_single_excitation_host": (_c_int, [_c_u64, _c_int, ctypes.POINTER(_c_i32)]),
}


class QmbError(RuntimeError):
    """A nonzero status from libqmb_hip.so, carrying qmb_last_error()."""


_LIB: ctypes.CDLL | None = None


def library() -> ctypes.CDLL:
    """Loads libqmb_hip.so once and binds every declared entry point; raises if it is absent."""
    global _LIB  # pylint: disable=global-statement
    if _LIB is None:
        if not LIBRARY_PATH.exists():
            raise ImportError(
                f"libqmb_hip.so not found at {LIBRARY_PATH}; build it with `make -C qmb-public_amd/csrc` "
                "(or __graft_entry__.build()). There is no non-HIP fallback."
            )
        lib = ctypes.CDLL(str(LIBRARY_PATH))
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _LIB = lib
    return _LIB


def check(status: int, what: str) -> None:
    """Raises QmbError with the library's message when status is nonzero."""
    if status != QMB_OK:
        message = library().qmb_last_error().decode("utf-8", "replace")
        raise QmbError(f"{what}: {message} (status {status})")


def exported_symbols() -> list[str]:
    """Names of the entry points this binding expects the library to export."""
    return list(SIGNATURES)
