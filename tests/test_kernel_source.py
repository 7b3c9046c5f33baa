"""
Source rules of the HIP kernels that no GPU run can be relied on to show.

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
