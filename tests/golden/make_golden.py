"""
Writes the golden fixtures in tests/golden/ from the reference's own CPU op.

The reference op (qmb/_hamiltonian.cpp + qmb/_hamiltonian_cpu.cpp) is compiled from its sources by
`make -C oracle ref` into oracle/_ref/ and run through oracle/ref_runner.py in a subprocess.  Each
fixture stores the inputs and the reference output, so the tests need neither the reference nor the
generators to reproduce them:

    python tests/golden/make_golden.py          (needs /root/reference and a built oracle/_ref)

Fixtures:  aw_*.npz  apply_within cases,  fr_*.npz  find_relative cases.
"""

from __future__ import annotations

import json
import pathlib
import subprocess
import sys
import tempfile

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path[:0] = [str(ROOT / "qmb-public_amd"), str(ROOT)]

from qmb_hip import synthetic as sy  # noqa: E402  pylint: disable=wrong-import-position


def _sorted(site: object, kind: object, coef: object) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    s, k, c = (np.asarray(x) for x in (site, kind, coef))
    order = np.argsort(-np.linalg.norm(c, axis=1), kind="stable")
    return s[order], k[order], c[order]


def _run(op: str, inputs: dict[str, object]) -> np.ndarray:
    with tempfile.TemporaryDirectory() as d:
        inp = pathlib.Path(d) / "in.npz"
        out = pathlib.Path(d) / "out.npz"
        np.savez(inp, **inputs)
        subprocess.run([sys.executable, str(ROOT / "oracle" / "ref_runner.py"), op, str(inp), str(out)], check=True)
        return np.load(out)["result"]


def apply_case(name: str, pc: int, terms: tuple[np.ndarray, np.ndarray, np.ndarray], cfg_i: np.ndarray, cfg_j: np.ndarray,
               psi_seed: int) -> dict[str, object]:
    s, k, c = terms
    psi = sy.random_psi(len(cfg_i), psi_seed)
    inputs = dict(configs_i=cfg_i, psi_i=psi, configs_j=cfg_j, site=s, kind=k, coef=c, particle_cut=pc)
    result = _run("apply_within", inputs)
    np.savez_compressed(HERE / f"aw_{name}.npz", **inputs, result=result)
    return {"op": "apply_within", "name": name, "n_i": len(cfg_i), "n_j": len(cfg_j), "n_terms": len(s), "n_qubytes": cfg_i.shape[1],
            "particle_cut": pc}


def find_case(name: str, pc: int, terms: tuple[np.ndarray, np.ndarray, np.ndarray], cfg_i: np.ndarray, exclude: np.ndarray,
              k_sel: int, psi_seed: int) -> dict[str, object]:
    s, k, c = terms
    psi = sy.random_psi(len(cfg_i), psi_seed)
    inputs = dict(configs_i=cfg_i, psi_i=psi, configs_exclude=exclude, site=s, kind=k, coef=c, particle_cut=pc,
                  count_selected=k_sel)
    result = _run("find_relative", inputs)
    np.savez_compressed(HERE / f"fr_{name}.npz", **inputs, result=result)
    return {"op": "find_relative", "name": name, "n_i": len(cfg_i), "n_exclude": len(exclude), "count_selected": k_sel,
            "rows": int(len(result)), "n_qubytes": cfg_i.shape[1], "particle_cut": pc}


def main() -> None:
    manifest = []
    # C1: the reference's CPU-runnable configuration (20 spin orbitals, 5a+5b, 10k configs).
    c1 = _sorted(*sy.molecular_terms(10, 1001))
    cfg_c1 = sy.sector_configs(10, 5, 5, 10000, 1002)
    manifest.append(apply_case("c1_molecular_20so_10k", 1, c1, cfg_c1, cfg_c1, 1003))
    mol8 = _sorted(*sy.molecular_terms(4, 11))
    full8 = sy.sector_configs(4, 2, 2, 36, 12)
    manifest.append(apply_case("molecular_8so_full_sector", 1, mol8, full8, full8, 13))
    rnd = tuple(x.numpy() for x in sy.random_terms(12, 400, 21))
    cfg12 = sy.sort_unique_configs(np.random.default_rng(22).integers(0, 256, size=(600, 2), dtype=np.uint8) & np.array([255, 15], np.uint8))
    manifest.append(apply_case("random_ops_12_sites", 1, rnd, cfg12, cfg12, 23))
    manifest.append(apply_case("random_ops_12_sites_bose2", 2, rnd, cfg12, cfg12, 24))
    ising = tuple(x.numpy() for x in sy.terms_from_dict(sy.ising_dict(4, 3, x=0.7, zh=1.0, zv=0.5, yh=0.3)))
    manifest.append(apply_case("ising_4x3_bose2", 2, ising, cfg12, cfg12, 25))
    hub = _sorted(*sy.terms_from_dict(sy.hubbard_dict(40, 1, 1.0, 4.0)))
    cfg80 = sy.hubbard_chain_configs(40, 4000, 41, n_seed_configs=512)
    manifest.append(apply_case("hubbard_chain_80so", 1, hub, cfg80, cfg80, 26))
    rnd200 = tuple(x.numpy() for x in sy.random_terms(200, 300, 51))
    rng = np.random.default_rng(52)
    base = rng.integers(0, 256, size=(64, 25), dtype=np.uint8)
    flips = (rng.integers(0, 2, size=(64, 25)) << rng.integers(0, 8, size=(64, 25))).astype(np.uint8)
    cfg200 = sy.sort_unique_configs(np.concatenate([base, base ^ flips]))
    manifest.append(apply_case("random_ops_200_sites", 1, rnd200, cfg200, cfg200, 27))
    cfg_i = sy.sector_configs(10, 5, 5, 1500, 7)
    cfg_j = sy.sector_configs(10, 5, 5, 2500, 8)
    manifest.append(apply_case("molecular_20so_distinct_sets", 1, c1, cfg_i, cfg_j, 28))

    cfg300 = sy.sector_configs(10, 5, 5, 300, 4)
    manifest.append(find_case("molecular_20so_k2000", 1, c1, cfg300, cfg300, 2000, 9))
    manifest.append(find_case("molecular_20so_k7", 1, c1, cfg300, cfg300, 7, 9))
    hub42 = _sorted(*sy.terms_from_dict(sy.hubbard_dict(4, 2, 1.0, 4.0)))
    cfg16 = sy.sector_configs(8, 4, 4, 200, 5)
    manifest.append(find_case("hubbard_4x2_k150", 1, hub42, cfg16, cfg16, 150, 10))
    exclude = sy.sector_configs(10, 5, 5, 5000, 15)
    cfg200b = sy.sector_configs(10, 5, 5, 200, 14)
    manifest.append(find_case("molecular_20so_distinct_exclude_k3000", 1, c1, cfg200b, exclude, 3000, 16))
    manifest.append(find_case("ising_bose2_k40", 2, ising, cfg12[:60], cfg12[:60], 40, 17))
    (HERE / "manifest.json").write_text(json.dumps(manifest, indent=1) + "\n")


if __name__ == "__main__":
    main()
