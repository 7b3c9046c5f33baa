"""
Host-side behaviour of the launch-shape autotuner (qmb_hip/autotune.py) that needs no GPU: a
cached key answers without running a candidate, the cache file is read back by a new tuner and a
corrupt one is ignored, and an empty candidate list is an error.  Timing itself runs on the GPU
(tests/test_gpu_batch.py).
"""

from __future__ import annotations

import json
import pathlib

import pytest

from qmb_hip.autotune import Autotuner


def _never(params: dict[str, object]) -> None:
    raise AssertionError(f"a cached key ran candidate {params}")


def test_cached_key_answers_without_running(tmp_path: pathlib.Path) -> None:
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"apply/n=1000": {"params": {"tile": 256, "groups": 8}, "ms": 1.5, "tried": []}}))
    tuner = Autotuner(str(path))
    got = tuner.best("apply/n=1000", [{"tile": 128}], _never)
    assert got == {"tile": 256, "groups": 8}
    got["tile"] = 0  # the caller gets a copy
    assert tuner.best("apply/n=1000", [], _never) == {"tile": 256, "groups": 8}


def test_corrupt_cache_file_starts_empty(tmp_path: pathlib.Path) -> None:
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    tuner = Autotuner(str(path))
    assert tuner.cache == {}
    with pytest.raises(ValueError, match="no candidates"):
        tuner.best("apply/n=1", [], _never)


def test_save_round_trips_through_a_new_tuner(tmp_path: pathlib.Path) -> None:
    path = tmp_path / "tune.json"
    first = Autotuner(str(path))
    first.cache["k"] = {"params": {"waves": 4096}, "ms": 2.0, "tried": [[{"waves": 4096}, 2.0]]}
    first.save()
    second = Autotuner(str(path))
    assert second.best("k", [{"waves": 1}], _never) == {"waves": 4096}
    assert Autotuner(None).path is None  # no file: memory only, save() writes nothing
    Autotuner(None).save()
