"""
The oracle against the golden vectors written by the reference's own CPU op (tests/golden/).

apply_within: within 1e-12 of the reference relative to each entry's sum of |contributions|.
find_relative: the same rows, up to configs whose heaviest contribution ties the K-th weight.
"""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from brute import contribution_weights
from oracle import oracle as orc

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
MANIFEST = json.loads((GOLDEN / "manifest.json").read_text())
APPLY = [m["name"] for m in MANIFEST if m["op"] == "apply_within"]
FIND = [m["name"] for m in MANIFEST if m["op"] == "find_relative"]


def load(prefix: str, name: str) -> dict[str, np.ndarray]:
    with np.load(GOLDEN / f"{prefix}_{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("name", APPLY)
def test_oracle_apply_within_matches_reference(name: str) -> None:
    d = load("aw", name)
    got, asum, _ = orc.apply_within(int(d["particle_cut"]), d["configs_i"], d["psi_i"], d["configs_j"], d["site"], d["kind"],
                                    d["coef"])
    err = np.abs(got - d["result"])
    assert np.all(err <= 1e-12 * asum), (name, err.max())
    assert np.all(d["result"][asum == 0] == 0)


def _entries(pc: int, cfg: np.ndarray, psi: np.ndarray, exclude: np.ndarray, s: np.ndarray, k: np.ndarray,
             c: np.ndarray) -> list[tuple[float, bytes]]:
    return contribution_weights(pc, cfg, psi, exclude, s, k, c)


@pytest.mark.parametrize("name", FIND)
def test_oracle_find_relative_matches_reference(name: str) -> None:
    d = load("fr", name)
    pc = int(d["particle_cut"])
    k_sel = int(d["count_selected"])
    got, kth = orc.find_relative(pc, d["configs_i"], d["psi_i"], k_sel, d["site"], d["kind"], d["coef"], d["configs_exclude"])
    want = d["result"]
    assert [bytes(r) for r in got] == sorted(bytes(r) for r in got)
    diff = {bytes(r) for r in got} ^ {bytes(r) for r in want}
    if diff:
        best: dict[bytes, float] = {}
        for w, key in _entries(pc, d["configs_i"], d["psi_i"], d["configs_exclude"], d["site"], d["kind"], d["coef"]):
            best[key] = max(best.get(key, -1.0), w)
        assert all(best[key] == kth for key in diff), name


@pytest.mark.parametrize("name", FIND)
def test_find_relative_brute_force(name: str) -> None:
    """Selection rule restated by brute force: top-K contributions by weight, zeros dropped, distinct."""
    d = load("fr", name)
    pc = int(d["particle_cut"])
    k_sel = int(d["count_selected"])
    entries = _entries(pc, d["configs_i"], d["psi_i"], d["configs_exclude"], d["site"], d["kind"], d["coef"])
    entries.sort(key=lambda e: -e[0])
    top = entries[:k_sel]
    kth = top[-1][0] if len(top) == k_sel else 0.0
    strict = {key for w, key in top if w > kth}
    allowed = strict | {key for w, key in entries if w == kth and w > 0}
    want = {bytes(r) for r in d["result"]}
    assert strict <= want <= allowed, name


@pytest.mark.parametrize("name", FIND)
def test_oracle_best_weights_brute_force(name: str) -> None:
    """best_weights equals the per-target maximum over a brute-force enumeration of every
    contribution (exclusion off: the targets are the selected configs and a few excluded ones)."""
    d = load("fr", name)
    pc = int(d["particle_cut"])
    targets = np.concatenate([d["result"], d["configs_exclude"][:5]])
    got = orc.best_weights(pc, d["configs_i"], d["psi_i"], d["site"], d["kind"], d["coef"], targets)
    no_exclude = np.zeros((0, targets.shape[1]), dtype=np.uint8)
    best: dict[bytes, float] = {}
    for w, key in _entries(pc, d["configs_i"], d["psi_i"], no_exclude, d["site"], d["kind"], d["coef"]):
        best[key] = max(best.get(key, -1.0), w)
    want = np.array([best.get(bytes(r), -1.0) for r in targets])
    assert np.array_equal(got, want), name
