"""
Source rules of the HIP kernels and their host code that no GPU run can be relied on to show.

A cross-lane operation (shuffle, ballot, readlane, DPP move) reads the registers of other lanes, so
every lane it reads from must execute it.  On the right of `&&` or `||` it runs only on the lanes
where the left side did not decide the result: the lanes reading from a lane that skipped it get a
stale value.  This once dropped contributions in the exact core comparison of two-word configs
(a lane whose first word already differed skipped the second word's shuffle).  Every such
operation is therefore evaluated into a value of its own before it is combined.
"""

from __future__ import annotations

import pathlib
import re

CSRC = pathlib.Path(__file__).resolve().parent.parent / "qmb-public_amd" / "csrc"
CROSS_LANE = re.compile(r"\b(__shfl\w*|shfl_u64|shfl_d2|__ballot\w*|__builtin_amdgcn_readlane|__builtin_amdgcn_update_dpp|"
                        r"__builtin_amdgcn_mov_dpp|__builtin_amdgcn_ds_bpermute)\s*\(")


def statements(text: str) -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return [s for s in re.split(r"[;{}]", text) if s.strip()]


def short_circuited(stmt: str) -> bool:
    """A cross-lane call to the right of the first `&&` / `||` of a statement."""
    m = re.search(r"&&|\|\|", stmt)
    return bool(m and CROSS_LANE.search(stmt, m.end()))


def test_rule_catches_the_fault_pattern() -> None:
    assert short_circuited("same = same && shfl_u64(core.w[w], src) == core.w[w]")
    assert short_circuited("mixed = mixed || __ballot(diff) != 0ull")
    assert not short_circuited("const bool eq = shfl_u64(core.w[w], src) == core.w[w]")
    assert not short_circuited("same = same & eq")


def test_no_cross_lane_operation_behind_a_short_circuit() -> None:
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.hpp"))
    assert len(sources) >= 8
    bad = []
    for path in sources:
        for stmt in statements(path.read_text()):
            if short_circuited(stmt):
                bad.append(f"{path.name}: {' '.join(stmt.split())[:160]}")
    assert not bad, "\n".join(bad)


PACKAGE = CSRC.parent


def test_hip_code_is_built_without_packed_fp32() -> None:
    """Every HIP compile line turns the packed-FP32 instructions off (some of their forms have
    given wrong results without an error); the kernels write no such instruction by hand."""
    flags = (CSRC / "Makefile").read_text()
    assert "-Xclang -target-feature -Xclang -packed-fp32-ops" in flags
    for path in sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.hpp")):
        assert not re.search(r"v_pk_(add|mul|fma)_f32", path.read_text()), path.name


def test_no_priority_streams_memset_nodes_or_host_registration() -> None:
    """Streams keep the default priority; device memory is cleared by fill kernels, never by a
    memset node a graph capture would record; host memory is allocated pinned, never registered
    after the fact."""
    banned = re.compile(r"hipStreamCreateWithPriority|hipMemsetAsync|hipMemset\b|hipHostRegister|"
                        r"is_high_priority_stream|Stream\(priority")
    files = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.hpp")) + sorted((PACKAGE / "qmb_hip").glob("*.py"))
    assert len(files) >= 15
    bad = [f"{p.name}: {m.group(0)}" for p in files for m in banned.finditer(p.read_text())]
    assert not bad, bad
