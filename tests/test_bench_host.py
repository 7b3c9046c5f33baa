"""
CPU checks of bench.py's measurement bookkeeping: the structured and hop bytes models add up the
terms their docstrings state, and a profile summary's traffic figures are used only when the
summary was made on the current library sources (its `source_sha:` line).
"""

from __future__ import annotations

import importlib.util
import pathlib
import types

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def bench() -> types.ModuleType:
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    module = importlib.util.module_from_spec(spec)
    assert spec.loader is not None
    spec.loader.exec_module(module)
    return module


def _stats(entries: int, groups: int, multi_entries: int, multi_groups: int) -> dict[str, int]:
    return {"entries": entries, "groups": groups, "multi_entries": multi_entries, "multi_groups": multi_groups, "largest": 9}


@pytest.mark.parametrize("nq,real", [(7, True), (9, False)])
def test_structured_model_terms(bench: types.ModuleType, nq: int, real: bool) -> None:
    st = {2: _stats(910, 355, 680, 125), 1: _stats(140, 93, 66, 20)}
    pp = types.SimpleNamespace(core_stats=lambda targets, order: st[order])
    n_pairs = 8 * nq * (8 * nq - 1) // 2
    terms = types.SimpleNamespace(pairs_desc=types.SimpleNamespace(real=int(real), n_pairs=n_pairs),
                                  singles_desc=types.SimpleNamespace(real=int(real)))
    n = 100
    words = (nq + 7) // 8
    coef = 8 if real else 16
    n_bits = 8 * nq
    out = bench.structured_model(pp, terms, None, nq, n)
    # per launch: starts, one config per group of >= 2 entries, entry + amplitude per walked entry,
    # the table once, the output once
    assert out["pair_kernel"]["bytes"] == 4 * 355 + 8 * words * 125 + 24 * 680 + coef * n_pairs ** 2 + 16 * n
    assert out["single_kernel"]["bytes"] == 4 * 93 + 8 * words * 20 + 24 * 66 + coef * (n_bits ** 2 + n_bits ** 3) + 16 * n
    assert out["diag_apply_kernel"]["bytes"] == n * (8 * words + 32) + 16 * (1 + n_bits + n_pairs)
    # per build: 52 per entry, a config per entry of a group of >= 2, a start per group
    assert out["core_index_build"]["bytes"] == sum(52 * s["entries"] + 8 * words * s["multi_entries"] + 4 * s["groups"]
                                                   for s in st.values())
    assert out["pair_kernel"]["multi_groups"] == 125


def test_hop_model_terms(bench: types.ModuleType) -> None:
    stats = {0: {"entries": 400, "groups": 210, "pair_tests": 9000, "largest": 20, "multi_entries": 190},
             1: {"entries": 400, "groups": 205, "pair_tests": 8000, "largest": 18, "multi_entries": 195}}
    targets = types.SimpleNamespace(strings_stats=lambda parity: stats[parity])
    terms = types.SimpleNamespace(singles_desc=types.SimpleNamespace(real=1))
    nq, n = 16, 400
    words, n_bits = 2, 128
    out = bench.hop_model(terms, targets, nq, n)
    for parity in (0, 1):
        assert out[f"hop_kernel_{parity}"]["bytes"] == stats[parity]["multi_entries"] * (8 + 4 + 8 * words + 16 + 32) + 8 * n_bits ** 2
    assert out["diag_scale_kernel"]["bytes"] == 64 * n
    build = sum(n * (8 * words + 8 + 16 + 8) + stats[p]["multi_entries"] * (20 + 8 * words) for p in (0, 1))
    assert out["strings_build"]["bytes"] == build + n * (8 * words + 16)


def test_profile_traffic_only_on_matching_sources(bench: types.ModuleType, tmp_path: pathlib.Path,
                                                  monkeypatch: pytest.MonkeyPatch) -> None:
    sha = bench.source_sha()
    assert len(sha) == 16 and int(sha, 16) >= 0
    (tmp_path / "profiles").mkdir()
    summary = tmp_path / "profiles" / "some_kernel_work.md"
    monkeypatch.setattr(bench, "ROOT", tmp_path)
    monkeypatch.setattr(bench, "source_sha", lambda: sha)  # the sources stay where they are
    assert bench.read_traffic("work", "some_kernel")[:2] == (None, None)  # no summary
    summary.write_text(f"# title\n\nsource_sha: {sha}\n\nhbm_bytes_per_launch: 1.5e9\nl2_hit_rate: 0.25\n")
    traffic, hit, source = bench.read_traffic("work", "some_kernel")
    assert (traffic, hit) == (1.5e9, 0.25) and source == "profiles/some_kernel_work.md"
    summary.write_text("source_sha: 0123456789abcdef\nhbm_bytes_per_launch: 1.5e9\nl2_hit_rate: 0.25\n")
    traffic, hit, source = bench.read_traffic("work", "some_kernel")
    assert (traffic, hit) == (None, None) and "other sources" in source


def test_committed_profiles_name_their_sources(bench: types.ModuleType) -> None:
    """Every committed kernel summary that bench.py reads carries a source hash and the two figures
    it takes from the summary."""
    for name in ("pair_kernel_molecular52.md", "hop_kernel_hubbard128.md"):
        text = (ROOT / "profiles" / name).read_text()
        keys = {line.split(":", 1)[0] for line in text.splitlines() if ":" in line}
        assert {"source_sha", "hbm_bytes_per_launch", "l2_hit_rate"} <= keys, name
