"""
FCIDUMP ingestion and the cached term-table format (qmb/fcidump.py:50-199), without openfermion.

The reference reads the integrals, fills the 8-fold symmetric two-electron tensor, spreads both
tensors over spin orbitals (even = alpha, odd = beta), and hands them to openfermion
(InteractionOperator -> get_fermion_operator -> normal_ordered) to obtain the normal-ordered
spin-orbital term dict (qmb/fcidump.py:102-114).  openfermion ~= 1.7 is not part of this build; its
published algorithm for that chain is restated here, vectorised with numpy:

  1. get_fermion_operator iterates the non-zero tensor entries -- the constant, then the one-body
     entries, then the two-body entries, each in itertools.product order of the spin-orbital
     indices -- and adds each as one term to an empty FermionOperator.  Every addition to a
     SymbolicOperator drops a term whose coefficient is below EQ_TOLERANCE = 1e-8 in magnitude.
  2. normal_ordered walks those terms in insertion order and adds each term's normal form to an
     empty FermionOperator (same small-coefficient rule at every addition).  For a+_P a_Q and the
     constant the normal form is the term itself; for a+_P a+_Q a_R a_S it is zero when P == Q or
     R == S, else the term with the creators and the annihilators each in descending index order,
     with one sign flip per swap.  Terms that meet on the same normal form are summed in that
     order; a partial sum that falls below 1e-8 is deleted and the next contribution re-inserts the
     term at the end of the dict.
The result is the same dict -- keys, FP64 values and insertion order -- the reference's
_read_fcidump returns.  It is computed straight into (site, kind, coef) arrays; read_fcidump builds
the dict form from them for callers that want the reference's return value.

The cache (qmb/fcidump.py:143-169) is the reference's own: torch.save((site, kind, coef)) under the
user cache directory (platformdirs "qmb"/"kclab"), named by the SHA-256 of the FCIDUMP file plus
"v5", loaded with torch.load(weights_only=True).  A cache written by either implementation is read
by the other.
"""

from __future__ import annotations

import dataclasses
import gzip
import hashlib
import logging
import os
import pathlib
import re
import typing

import numpy as np
import torch

from .hamiltonian import Hamiltonian

QMB_MODEL_PATH = "QMB_MODEL_PATH"
EQ_TOLERANCE = 1e-8
CACHE_VERSION = "v5"

TermDict = dict[tuple[tuple[int, int], ...], complex]
TermArrays = tuple[np.ndarray, np.ndarray, np.ndarray]


def _open(file_name: pathlib.Path) -> typing.TextIO:
    if file_name.name.endswith(".gz"):
        return gzip.open(file_name, "rt", encoding="utf-8")
    return open(file_name, "rt", encoding="utf-8")  # pylint: disable=consider-using-with


def _wrap(index: np.ndarray, n: int) -> np.ndarray:
    """Tensor indexing semantics: -n <= i < n, negative from the end; else IndexError."""
    if index.size and (index.min() < -n or index.max() >= n):
        raise IndexError(f"FCIDUMP orbital index out of range for NORB={n}")
    return index % n


def _last_write(flat: np.ndarray, values: np.ndarray, size: int) -> np.ndarray:
    """Dense float64 [size] after assigning values[t] at flat[t] for t = 0, 1, ... in order."""
    out = np.zeros(size, dtype=np.float64)
    if flat.size:
        # the last occurrence of each position wins
        rev_unique, rev_first = np.unique(flat[::-1], return_index=True)
        out[rev_unique] = values[::-1][rev_first]
    return out


def read_integrals(file_name: pathlib.Path | str, *, cached: bool = False
                   ) -> tuple[tuple[int, int, int], float, np.ndarray | None, np.ndarray | None]:
    """Header (NORB, NELEC, MS2) and, unless cached, E0, h [n, n] and the 8-fold symmetric (ij|kl)
    [n, n, n, n] exactly as qmb/fcidump.py:50-100 fills them (later lines overwrite earlier ones)."""
    file_name = pathlib.Path(file_name)
    with _open(file_name) as file:
        n_orbit: int | None = None
        n_electron: int | None = None
        n_spin: int | None = None
        for line in file:
            data = line.lower()
            if (match := re.search(r"norb\s*=\s*(\d+)", data)) is not None:
                n_orbit = int(match.group(1))
            if (match := re.search(r"nelec\s*=\s*(\d+)", data)) is not None:
                n_electron = int(match.group(1))
            if (match := re.search(r"ms2\s*=\s*(\d+)", data)) is not None:
                n_spin = int(match.group(1))
            if "&end" in data:
                break
        if n_orbit is None or n_electron is None or n_spin is None:
            raise AssertionError("the FCIDUMP header must give NORB, NELEC and MS2 before &END")
        if cached:
            return (n_orbit, n_electron, n_spin), 0.0, None, None
        energy_0 = 0.0
        one: list[tuple[int, int, float]] = []
        two: list[tuple[int, int, int, int, float]] = []
        for line in file:
            pieces = line.split()
            if not pieces:
                continue  # blank lines carry nothing (the reference would fail on pieces[0])
            coefficient = float(pieces[0])
            sites = tuple(int(i) - 1 for i in pieces[1:])
            match sites:
                case (-1, -1, -1, -1):
                    energy_0 = coefficient
                case (_, -1, -1, -1):
                    pass  # Psi4's extra one-electron lines, omitted as the reference does
                case (i, j, -1, -1):
                    one.append((i, j, coefficient))
                case (_, _, _, -1):
                    raise ValueError(f"Invalid FCIDUMP format: {sites}")
                case (i, j, k, l):
                    two.append((i, j, k, l, coefficient))
                case _:
                    raise ValueError(f"Invalid FCIDUMP format: {sites}")
    n = n_orbit
    # one-body: h[i, j] = h[j, i] = c, line by line
    if one:
        a = np.array(one, dtype=np.float64)
        i, j = _wrap(a[:, 0].astype(np.int64), n), _wrap(a[:, 1].astype(np.int64), n)
        flat = np.stack([i * n + j, j * n + i], axis=1).reshape(-1)
        h = _last_write(flat, np.repeat(a[:, 2], 2), n * n).reshape(n, n)
    else:
        h = np.zeros((n, n))
    # two-body: the eight index orders of qmb/fcidump.py:89-96, in that order, line by line
    if two:
        a = np.array(two, dtype=np.float64)
        i, j, k, l = (_wrap(a[:, c].astype(np.int64), n) for c in range(4))
        orders = [(i, j, k, l), (i, j, l, k), (j, i, k, l), (j, i, l, k), (l, k, j, i), (k, l, j, i), (l, k, i, j), (k, l, i, j)]
        flat = np.stack([((p * n + q) * n + r) * n + s for p, q, r, s in orders], axis=1).reshape(-1)
        g = _last_write(flat, np.repeat(a[:, 4], 8), n**4).reshape(n, n, n, n)
    else:
        g = np.zeros((n, n, n, n))
    return (n_orbit, n_electron, n_spin), energy_0, h, g


def normal_ordered_terms(energy_0: float, h: np.ndarray, g: np.ndarray) -> tuple[TermArrays, np.ndarray]:
    """Normal-ordered spin-orbital terms of E0 + sum h a+a + sum E2b a+a+aa with
    E2b[p s1, q s2, r s2, s s1] = (ps|qr)/2 (qmb/fcidump.py:100-114), in the dict order openfermion
    leaves them.  Returns ((site int16 [T, 4], kind uint8 [T, 4], coef float64 [T, 2]), positions)."""
    n = h.shape[0]
    big_n = 2 * n
    # stage 1: the tensor entries as get_fermion_operator adds them (product order, exact-zero
    # entries skipped, |value| < EQ_TOLERANCE dropped by the addition)
    # one-body: E1b[2p+s, 2q+s] = h[p, q]
    p, q = np.nonzero(h)
    sig = np.repeat(np.arange(2), len(p))
    big_p = 2 * np.tile(p, 2) + sig
    big_q = 2 * np.tile(q, 2) + sig
    v1 = np.tile(h[p, q], 2)
    order = np.argsort(big_p * big_n + big_q, kind="stable")
    big_p, big_q, v1 = big_p[order], big_q[order], v1[order]
    keep = np.abs(v1) >= EQ_TOLERANCE
    big_p, big_q, v1 = big_p[keep], big_q[keep], v1[keep]
    # two-body: E2[p, q, r, s] = g[p, s, q, r] / 2 (the reference's permute(0, 2, 3, 1) / 2)
    e2 = np.ascontiguousarray(g.transpose(0, 2, 3, 1)) / 2
    p, q, r, s = np.nonzero(e2)
    v = e2[p, q, r, s]
    blocks = [(0, 0), (0, 1), (1, 0), (1, 1)]  # (spin of P and S, spin of Q and R)
    a_p = np.concatenate([2 * p + s1 for s1, _ in blocks])
    a_q = np.concatenate([2 * q + s2 for _, s2 in blocks])
    a_r = np.concatenate([2 * r + s2 for _, s2 in blocks])
    a_s = np.concatenate([2 * s + s1 for s1, _ in blocks])
    v2 = np.tile(v, len(blocks))
    order = np.argsort(((a_p * big_n + a_q) * big_n + a_r) * big_n + a_s, kind="stable")
    a_p, a_q, a_r, a_s, v2 = a_p[order], a_q[order], a_r[order], a_s[order], v2[order]
    keep = np.abs(v2) >= EQ_TOLERANCE
    a_p, a_q, a_r, a_s, v2 = a_p[keep], a_q[keep], a_r[keep], a_s[keep], v2[keep]
    has_constant = abs(energy_0) >= EQ_TOLERANCE
    base_one = int(has_constant)
    base_two = base_one + len(v1)
    # stage 2: normal forms, summed per form in stage-1 order
    valid = (a_p != a_q) & (a_r != a_s)
    pos2 = base_two + np.nonzero(valid)[0]
    a_p, a_q, a_r, a_s, v2 = a_p[valid], a_q[valid], a_r[valid], a_s[valid], v2[valid]
    sign = np.where(a_q > a_p, -1.0, 1.0) * np.where(a_s > a_r, -1.0, 1.0)
    c1, c2 = np.maximum(a_p, a_q), np.minimum(a_p, a_q)
    d1, d2 = np.maximum(a_r, a_s), np.minimum(a_r, a_s)
    key = ((c1 * big_n + c2) * big_n + d1) * big_n + d2
    x = sign * v2  # a sign flip is exact
    order = np.argsort(key, kind="stable")  # groups in stage-1 order
    key, x, pos2 = key[order], x[order], pos2[order]
    starts = np.nonzero(np.r_[True, key[1:] != key[:-1]])[0]
    sizes = np.diff(np.r_[starts, len(key)])
    n_groups = len(starts)
    acc = np.zeros(n_groups)
    alive = np.zeros(n_groups, dtype=bool)
    inserted = np.zeros(n_groups, dtype=np.int64)
    for rnd in range(int(sizes.max()) if n_groups else 0):
        m = sizes > rnd
        idx = starts[m] + rnd
        prev_alive = alive[m]
        new = np.where(prev_alive, acc[m], 0.0) + x[idx]
        inserted[m] = np.where(prev_alive, inserted[m], pos2[idx])
        acc[m] = new
        alive[m] = np.abs(new) >= EQ_TOLERANCE
    gk = key[starts][alive]
    g_val = acc[alive]
    g_pos = inserted[alive]
    d2g = gk % big_n
    d1g = (gk // big_n) % big_n
    c2g = (gk // big_n**2) % big_n
    c1g = gk // big_n**3
    # assemble in dict order: the constant, the one-body terms, the two-body terms by insertion
    n_one, n_two = len(v1), len(gk)
    total = base_one + n_one + n_two
    site = np.zeros((total, 4), dtype=np.int16)
    kind = np.full((total, 4), 2, dtype=np.uint8)
    coef = np.zeros((total, 2), dtype=np.float64)
    positions = np.zeros(total, dtype=np.int64)
    if has_constant:
        coef[0, 0] = energy_0
    site[base_one:base_two, 0], site[base_one:base_two, 1] = big_p, big_q
    kind[base_one:base_two, 0], kind[base_one:base_two, 1] = 1, 0
    coef[base_one:base_two, 0] = v1
    positions[base_one:base_two] = np.arange(base_one, base_two)
    o = np.argsort(g_pos, kind="stable")
    sl = slice(base_two, total)
    site[sl] = np.stack([c1g[o], c2g[o], d1g[o], d2g[o]], axis=1)
    kind[sl] = np.array([1, 1, 0, 0], dtype=np.uint8)
    coef[sl, 0] = g_val[o]
    positions[sl] = g_pos[o]
    return (site, kind, coef), positions


def terms_to_dict(site: np.ndarray, kind: np.ndarray, coef: np.ndarray) -> TermDict:
    """(site, kind, coef) arrays -> the reference's term dict (identity columns dropped)."""
    out: TermDict = {}
    for s_row, k_row, c_row in zip(site.tolist(), kind.tolist(), coef.tolist()):
        out[tuple((s, k) for s, k in zip(s_row, k_row) if k != 2)] = complex(c_row[0], c_row[1])
    return out


def fcidump_terms(file_name: pathlib.Path | str) -> tuple[tuple[int, int, int], TermArrays]:
    """(NORB, NELEC, MS2) and the normal-ordered terms as (site, kind, coef) arrays in dict order."""
    meta, energy_0, h, g = read_integrals(file_name)
    if h is None or g is None:
        raise ValueError(f"{file_name} holds no one- and two-electron integrals")
    terms, _ = normal_ordered_terms(energy_0, h, g)
    return meta, terms


def read_fcidump(file_name: pathlib.Path | str, *, cached: bool = False) -> tuple[tuple[int, int, int], TermDict]:
    """qmb/fcidump.py:50-114: (NORB, NELEC, MS2) and the normal-ordered term dict ({} when cached)."""
    if cached:
        meta, _, _, _ = read_integrals(file_name, cached=True)
        return meta, {}
    meta, terms = fcidump_terms(file_name)
    return meta, terms_to_dict(*terms)


def write_fcidump(file_name: pathlib.Path | str, energy_0: float, h: np.ndarray, g: np.ndarray, n_electron: int, ms2: int = 0) -> None:
    """Writes a standard FCIDUMP (unique (ij|kl) with i >= j, k >= l, ij >= kl; h with i >= j;
    E0 last) whose floats read back exactly."""
    n = h.shape[0]
    lines = [f" &FCI NORB={n},NELEC={n_electron},MS2={ms2},", "  ORBSYM=" + ",".join("1" * n) + ",", "  ISYM=1,", " &END"]
    for i in range(n):
        for j in range(i + 1):
            for k in range(n):
                for l in range(k + 1):
                    if i * (i + 1) // 2 + j < k * (k + 1) // 2 + l:
                        continue
                    if g[i, j, k, l] != 0.0:
                        lines.append(f"{float(g[i, j, k, l])!r} {i + 1} {j + 1} {k + 1} {l + 1}")
    for i in range(n):
        for j in range(i + 1):
            if h[i, j] != 0.0:
                lines.append(f"{float(h[i, j])!r} {i + 1} {j + 1} 0 0")
    lines.append(f"{float(energy_0)!r} 0 0 0 0")
    text = "\n".join(lines) + "\n"
    path = pathlib.Path(file_name)
    if path.name.endswith(".gz"):
        with gzip.open(path, "wt", encoding="utf-8") as file:
            file.write(text)
    else:
        path.write_text(text, encoding="utf-8")


def cache_directory() -> pathlib.Path:
    """The reference's cache folder (platformdirs user cache of app "qmb", author "kclab"); the
    QMB_HIP_CACHE_DIR environment variable points it elsewhere."""
    if "QMB_HIP_CACHE_DIR" in os.environ:
        return pathlib.Path(os.environ["QMB_HIP_CACHE_DIR"])
    import platformdirs  # pylint: disable=import-outside-toplevel

    return platformdirs.user_cache_path("qmb", "kclab")


def cache_file_for(model_file_name: pathlib.Path) -> pathlib.Path:
    """qmb/fcidump.py:146-147: SHA-256 of the FCIDUMP bytes + the format version."""
    return cache_directory() / (hashlib.sha256(model_file_name.read_bytes()).hexdigest() + CACHE_VERSION)


def save_terms(path: pathlib.Path | str, site: torch.Tensor, kind: torch.Tensor, coef: torch.Tensor) -> None:
    """The cached tuple format of qmb/fcidump.py:168: torch.save((site, kind, coef))."""
    path = pathlib.Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    torch.save((site.cpu(), kind.cpu(), coef.cpu()), path)


def load_terms(path: pathlib.Path | str) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Reads the cached tuple with torch.load(weights_only=True) and checks its layout."""
    data = torch.load(path, map_location="cpu", weights_only=True)
    if not (isinstance(data, tuple) and len(data) == 3):
        raise ValueError(f"{path} does not hold a (site, kind, coef) tuple")
    site, kind, coef = data
    if site.dtype != torch.int16 or kind.dtype != torch.uint8 or coef.dtype != torch.float64 or site.dim() != 2 or site.size(1) != 4 \
            or kind.shape != site.shape or coef.shape != (site.size(0), 2):
        raise ValueError(f"{path} does not hold site int16 [T, 4], kind uint8 [T, 4], coef float64 [T, 2]")
    return site, kind, coef


@dataclasses.dataclass
class ModelConfig:
    """qmb/fcidump.py:27-47: model name, models folder (default $QMB_MODEL_PATH or "models"), reference energy."""

    model_name: str
    model_path: pathlib.Path | None = None
    ref_energy: float | None = None

    def __post_init__(self) -> None:
        if self.model_path is not None:
            self.model_path = pathlib.Path(self.model_path)
        elif QMB_MODEL_PATH in os.environ:
            self.model_path = pathlib.Path(os.environ[QMB_MODEL_PATH])
        else:
            self.model_path = pathlib.Path("models")


class Model:
    """An FCIDUMP model (qmb/fcidump.py:117-214): its Hamiltonian (from the cache when present),
    qubit/electron/spin counts, reference energy, and the ModelProto methods."""

    network_dict: dict[str, type] = {}
    config_t = ModelConfig

    @classmethod
    def default_group_name(cls, config: ModelConfig) -> str:
        return config.model_name

    def __init__(self, args: ModelConfig) -> None:
        model_name = args.model_name
        model_path = args.model_path
        assert model_path is not None
        model_file_name = model_path / f"{model_name}.FCIDUMP.gz"
        model_file_name = model_file_name if model_file_name.exists() else model_path / model_name
        cache_file = cache_file_for(model_file_name)
        if cache_file.exists():
            logging.info("Loading FCIDUMP Hamiltonian '%s' from cache %s", model_name, cache_file)
            (n_orbit, n_electron, n_spin), _ = read_fcidump(model_file_name, cached=True)
            self.hamiltonian = Hamiltonian(load_terms(cache_file), kind="fermi")
        else:
            logging.info("Loading FCIDUMP Hamiltonian '%s' from file: %s", model_name, model_file_name)
            (n_orbit, n_electron, n_spin), (site, kind, coef) = fcidump_terms(model_file_name)
            # the dict constructor's path: terms in dict order, then sorted by |coef| descending
            self.hamiltonian = Hamiltonian((torch.from_numpy(site), torch.from_numpy(kind), torch.from_numpy(coef)), kind="fermi")
            self.hamiltonian._sort_site_kind_coef()  # pylint: disable=protected-access
            save_terms(cache_file, self.hamiltonian.site, self.hamiltonian.kind, self.hamiltonian.coef)
        self.n_qubit: int = n_orbit * 2
        self.n_electron: int = n_electron
        self.n_spin: int = n_spin
        self.ref_energy: float
        if args.ref_energy is not None:
            self.ref_energy = args.ref_energy
        else:
            yaml_file = model_file_name.parent / "FCIDUMP.yaml"
            if yaml_file.exists():
                import yaml  # pylint: disable=import-outside-toplevel

                with open(yaml_file, "rt", encoding="utf-8") as file:
                    data = yaml.safe_load(file)
                self.ref_energy = data.get(pathlib.Path(model_name).name, 0)
            else:
                self.ref_energy = 0
        logging.info("Model '%s': %d qubits, %d electrons, spin %d, reference energy %.10f", model_name, self.n_qubit, self.n_electron,
                     self.n_spin, self.ref_energy)

    def apply_within(self, configs_i: torch.Tensor, psi_i: torch.Tensor, configs_j: torch.Tensor) -> torch.Tensor:
        return self.hamiltonian.apply_within(configs_i, psi_i, configs_j)

    def find_relative(self, configs_i: torch.Tensor, psi_i: torch.Tensor, count_selected: int,
                      configs_exclude: torch.Tensor | None = None) -> torch.Tensor:
        return self.hamiltonian.find_relative(configs_i, psi_i, count_selected, configs_exclude)

    def diagonal_term(self, configs: torch.Tensor) -> torch.Tensor:
        return self.hamiltonian.diagonal_term(configs)

    def single_relative(self, configs: torch.Tensor) -> torch.Tensor:
        return self.hamiltonian.single_relative(configs)

    def show_config(self, config: torch.Tensor) -> str:
        """qmb/fcidump.py:201-204: one glyph per spatial orbital (alpha bit first)."""
        string = "".join(f"{i:08b}"[::-1] for i in config.cpu().numpy())
        return "[" + "".join(self._show_config_site(string[index:index + 2]) for index in range(0, self.n_qubit, 2)) + "]"

    def _show_config_site(self, string: str) -> str:
        match string:
            case "00":
                return " "
            case "10":
                return "↑"
            case "01":
                return "↓"
            case "11":
                return "↕"
            case _:
                raise ValueError(f"Invalid string: {string}")
