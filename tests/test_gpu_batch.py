"""
GPU checks of the batched persistent launch (qmb_batch_plan / qmb_apply_batch through
qmb_hip.batch.Batch):

  - every instance's output of one batched launch equals its own structured application (and so
    the oracle, which that path is checked against) in both modes, for a mix of Hermitian
    molecular instances of different sizes and word counts, to 1e-12 of each entry's scale;
  - a non-Hermitian batch (a table with residual terms on two config words) takes mode 0 and 1
    apart and adds the residual enumeration per instance;
  - a hipGraph of the batched forward + adjoint, replayed several times with eager work between
    replays, reproduces the eager result on every replay;
  - a two-vector batch applies every instance to two vectors in one launch (forward and backward of
    a training step for Hermitian instances), equal to the per-instance results, graph replays too;
  - the autotuner picks a wave count among its candidates without changing the result.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import batch as qb
from qmb_hip import prepared as pp
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu


def _dev(x: np.ndarray, device: torch.device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _molecular(n_orb: int, n_a: int, n_b: int, count: int, seed: int, device: torch.device) -> tuple[pp.PreparedTerms, pp.PreparedTargets]:
    s, k, c = sorted_terms(*sy.molecular_terms(n_orb, seed))
    cfg = sy.sector_configs(n_orb, n_a, n_b, count, seed + 1)
    nq = cfg.shape[1]
    terms = pp.PreparedTerms(_dev(s, device), _dev(k, device), _dev(c, device), nq, 1)
    targets = pp.PreparedTargets(_dev(cfg, device), nq)
    return terms, targets


def _reference(terms: pp.PreparedTerms, targets: pp.PreparedTargets, psi: torch.Tensor, mode: int) -> np.ndarray:
    vec = torch.view_as_real(psi).contiguous()
    out = pp.apply(terms, targets, targets.configs, vec, mode)
    return torch.view_as_complex(out).cpu().numpy()


def _scale(ref: np.ndarray, terms_coef_abs: float, psi: torch.Tensor) -> np.ndarray:
    return np.abs(ref) + terms_coef_abs * float(psi.abs().sum())


def test_batch_matches_per_instance_hermitian(cuda_device: torch.device) -> None:
    specs = [(4, 2, 2, 36, 11), (6, 3, 2, 120, 12), (8, 3, 3, 800, 13), (10, 5, 5, 3000, 14), (5, 2, 1, 40, 15)]
    instances = [_molecular(*spec, cuda_device) for spec in specs]
    b = qb.Batch(instances)
    assert b.hermitian
    psis = [torch.from_numpy(sy.random_psi(t.n, 100 + k)).to(cuda_device) for k, (_, t) in enumerate(instances)]
    b.set_inputs(psis)
    for mode in (pp.MODE_FORWARD, pp.MODE_ADJOINT):
        b.launch(mode)
        for k, (terms, targets) in enumerate(instances):
            ref = _reference(terms, targets, psis[k], mode)
            got = b.output(k).cpu().numpy()
            np.testing.assert_array_less(np.abs(got - ref), 1e-12 * _scale(ref, 10.0, psis[k]) + 1e-300)


def test_batch_oracle_spot_check(cuda_device: torch.device) -> None:
    """The first instance of a batch against the oracle directly."""
    s, k, c = sorted_terms(*sy.molecular_terms(6, 21))
    cfg = sy.sector_configs(6, 3, 3, 300, 22)
    terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), 2, 1)
    targets = pp.PreparedTargets(_dev(cfg, cuda_device), 2)
    other = _molecular(4, 2, 2, 36, 23, cuda_device)
    b = qb.Batch([(terms, targets), other])
    psi = sy.random_psi(len(cfg), 24)
    b.set_inputs([torch.from_numpy(psi).to(cuda_device), torch.from_numpy(sy.random_psi(36, 25)).to(cuda_device)])
    b.launch(pp.MODE_FORWARD)
    want, asum, _ = orc.apply_within(1, cfg, psi, cfg, s, k, c)
    got = b.output(0).cpu().numpy()
    assert np.all(np.abs(got - want) <= 1e-12 * asum + 1e-300)


def _two_word_instance(device: torch.device, seed: int) -> tuple[pp.PreparedTerms, pp.PreparedTargets, tuple]:
    from test_gpu_pairs import _fixed_weight_configs, _pair_heavy_table  # pylint: disable=import-outside-toplevel

    sites = np.array([0, 2, 4, 9, 31, 40, 63, 64, 65, 67, 70], dtype=np.int64)
    s, k, c = _pair_heavy_table(sites, 600, seed)
    cfg = _fixed_weight_configs(sites, 4, 9)
    terms = pp.PreparedTerms(_dev(s, device), _dev(k, device), _dev(c, device), 9, 1)
    targets = pp.PreparedTargets(_dev(cfg, device), 9)
    return terms, targets, (s, k, c, cfg)


def test_batch_non_hermitian_with_residual(cuda_device: torch.device) -> None:
    t1, g1, raw1 = _two_word_instance(cuda_device, 51)
    t2, g2, _ = _two_word_instance(cuda_device, 52)
    assert not t1.hermitian and t1.n_residual_terms > 0
    b = qb.Batch([(t1, g1), (t2, g2)])
    assert b.residual == [0, 1]
    psis = [torch.from_numpy(sy.random_psi(g.n, 60 + k)).to(cuda_device) for k, g in enumerate((g1, g2))]
    b.set_inputs(psis)
    b.launch(pp.MODE_FORWARD)
    s, k, c, cfg = raw1
    want, asum, _ = orc.apply_within(1, cfg, psis[0].cpu().numpy(), cfg, s, k, c)
    assert np.all(np.abs(b.output(0).cpu().numpy() - want) <= 1e-12 * asum + 1e-300)
    b.launch(pp.MODE_ADJOINT)
    for idx, (terms, targets) in enumerate(((t1, g1), (t2, g2))):
        ref = _reference(terms, targets, psis[idx], pp.MODE_ADJOINT)
        np.testing.assert_array_less(np.abs(b.output(idx).cpu().numpy() - ref), 1e-12 * _scale(ref, 10.0, psis[idx]) + 1e-300)


def test_batch_mixed_word_counts(cuda_device: torch.device) -> None:
    """Instances of one and two config words in one launch (waves of one block then run both), with
    groups large enough for the staged path: a regression test for per-wave staging slots."""
    s1, k1, c1 = sorted_terms(*sy.molecular_terms(10, 71))
    cfg1 = sy.excitation_configs(10, 4, 4, 3000, 72)
    s2, k2, c2 = sorted_terms(*sy.molecular_terms(34, 73, threshold=1e-6))
    cfg2 = sy.excitation_configs(34, 3, 3, 4000, 74)
    assert cfg1.shape[1] <= 8 < cfg2.shape[1]
    instances = []
    for s, k, c, cfg in ((s1, k1, c1, cfg1), (s2, k2, c2, cfg2), (s1, k1, c1, cfg1)):
        nq = cfg.shape[1]
        terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), nq, 1)
        instances.append((terms, pp.PreparedTargets(_dev(cfg, cuda_device), nq)))
    b = qb.Batch(instances)
    psis = [torch.from_numpy(sy.random_psi(t.n, 400 + k)).to(cuda_device) for k, (_, t) in enumerate(instances)]
    b.set_inputs(psis)
    for waves in (64, 4096):
        b.launch(pp.MODE_FORWARD, waves=waves)
        for k, (terms, targets) in enumerate(instances):
            ref = _reference(terms, targets, psis[k], pp.MODE_FORWARD)
            got = b.output(k).cpu().numpy()
            np.testing.assert_array_less(np.abs(got - ref), 1e-12 * _scale(ref, 10.0, psis[k]) + 1e-300)


def test_batch_two_vectors_in_one_launch(cuda_device: torch.device) -> None:
    """A two-vector batch gives, for both vectors, what each instance's own application gives;
    one and two config words mixed in the batch, and its graph replays match."""
    s1, k1, c1 = sorted_terms(*sy.molecular_terms(10, 81))
    cfg1 = sy.excitation_configs(10, 4, 4, 2500, 82)
    s2, k2, c2 = sorted_terms(*sy.molecular_terms(34, 83, threshold=1e-6))
    cfg2 = sy.excitation_configs(34, 3, 3, 3000, 84)
    instances = []
    for s, k, c, cfg in ((s1, k1, c1, cfg1), (s2, k2, c2, cfg2)):
        nq = cfg.shape[1]
        terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), nq, 1)
        instances.append((terms, pp.PreparedTargets(_dev(cfg, cuda_device), nq)))
    b = qb.Batch(instances, two_vectors=True)
    psis = [torch.from_numpy(sy.random_psi(t.n, 500 + k)).to(cuda_device) for k, (_, t) in enumerate(instances)]
    grads = [torch.from_numpy(sy.random_psi(t.n, 600 + k)).to(cuda_device) for k, (_, t) in enumerate(instances)]
    b.set_inputs(psis, grads)
    b.launch_both()
    eager = (b.out.clone(), b.out2.clone())
    for k, (terms, targets) in enumerate(instances):
        for second, vec in ((False, psis[k]), (True, grads[k])):
            ref = _reference(terms, targets, vec, pp.MODE_FORWARD)
            got = b.output(k, second).cpu().numpy()
            np.testing.assert_array_less(np.abs(got - ref), 1e-12 * _scale(ref, 10.0, vec) + 1e-300)
    with pytest.raises(RuntimeError):
        b.launch(pp.MODE_FORWARD)
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream(device=cuda_device)
    side.wait_stream(torch.cuda.current_stream(cuda_device))
    with torch.cuda.stream(side):
        b.launch_both()
    torch.cuda.current_stream(cuda_device).wait_stream(side)
    with torch.cuda.graph(graph):
        b.launch_both()
    for _ in range(3):
        b.out.fill_(float("nan"))
        b.out2.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for got, want in zip((b.out, b.out2), eager):
            assert (got - want).abs().max().item() <= 1e-12 * want.abs().max().item()


def test_batch_graph_replays_match_eager(cuda_device: torch.device) -> None:
    specs = [(6, 3, 3, 300, 31), (8, 4, 3, 1000, 32), (4, 2, 2, 36, 33)]
    instances = [_molecular(*spec, cuda_device) for spec in specs]
    b = qb.Batch(instances)
    psis = [torch.from_numpy(sy.random_psi(t.n, 200 + k)).to(cuda_device) for k, (_, t) in enumerate(instances)]
    b.set_inputs(psis)
    b.launch(pp.MODE_FORWARD)
    eager = b.out.clone()
    side = torch.cuda.Stream(device=cuda_device)
    side.wait_stream(torch.cuda.current_stream(cuda_device))
    with torch.cuda.stream(side):
        b.launch(pp.MODE_FORWARD)  # warm up on the capture stream
    torch.cuda.current_stream(cuda_device).wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        b.launch(pp.MODE_FORWARD)
    for replay in range(4):
        b.out.fill_(float("nan"))  # eager work between replays; the graph must clear it itself
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        diff = (b.out - eager).abs().max().item()
        scale = eager.abs().max().item()
        assert diff <= 1e-12 * scale, (replay, diff)


def test_batch_autotune_keeps_result(cuda_device: torch.device) -> None:
    instances = [_molecular(8, 3, 3, 600, 41, cuda_device), _molecular(6, 2, 2, 200, 42, cuda_device)]
    b = qb.Batch(instances)
    b.set_inputs([torch.from_numpy(sy.random_psi(t.n, 300 + k)).to(cuda_device) for k, (_, t) in enumerate(instances)])
    b.launch(pp.MODE_FORWARD, waves=qb.WAVE_CANDIDATES[0])
    first = b.out.clone()
    waves = b.tune(retune=True)
    assert waves in qb.WAVE_CANDIDATES and b.split in qb.SPLIT_CANDIDATES
    b.launch(pp.MODE_FORWARD)
    assert (b.out - first).abs().max().item() <= 1e-12 * first.abs().max().item()


@pytest.mark.parametrize("two", [False, True], ids=["one-vector", "two-vector"])
def test_batch_split_matches_full_batch(cuda_device: torch.device, two: bool) -> None:
    """Every split of a batch (the largest instances by their own structured launches, the rest
    batched) gives each instance's own application, both vectors of a two-vector batch and both
    modes of a one-vector one; the instances are ordered by core-index entries, largest first."""
    specs = [(6, 3, 3, 300, 51), (8, 3, 3, 1500, 52), (4, 2, 2, 36, 53), (8, 4, 3, 900, 54)]
    instances = [_molecular(*spec, cuda_device) for spec in specs]
    b = qb.Batch(instances, two_vectors=two)
    work = [t.cores(2).capacity + t.cores(1).capacity for _, t in instances]
    assert [work[k] for k in b.order] == sorted(work, reverse=True)
    psis = [torch.from_numpy(sy.random_psi(t.n, 700 + k)).to(cuda_device) for k, (_, t) in enumerate(instances)]
    grads = [torch.from_numpy(sy.random_psi(t.n, 800 + k)).to(cuda_device) for k, (_, t) in enumerate(instances)]
    b.set_inputs(psis, grads if two else None)
    for split in (0, 1, 3, 4):
        b.configure(split)
        modes = (pp.MODE_FORWARD,) if two else (pp.MODE_FORWARD, pp.MODE_ADJOINT)
        for mode in modes:
            if two:
                b.launch_both(4096)
            else:
                b.launch(mode, 4096)
            for k, (terms, targets) in enumerate(instances):
                for second, vec in (((False, psis[k]), (True, grads[k])) if two else ((False, psis[k]),)):
                    ref = _reference(terms, targets, vec, mode)
                    got = b.output(k, second).cpu().numpy()
                    np.testing.assert_array_less(np.abs(got - ref), 1e-12 * _scale(ref, 10.0, vec) + 1e-300)


def test_per_instance_two_vector_matches_two_applications(cuda_device: torch.device) -> None:
    """pp.apply_dual (qmb_apply_structured_dual: diagonal, single and pair kernels each applying A^H
    to two vectors in one pass) equals two separate adjoint applications, one- and two-word sets."""
    for n_orb, n_e, count, seed in ((8, 3, 800, 31), (34, 5, 3000, 33)):
        s, k, c = sorted_terms(*sy.molecular_terms(n_orb, seed))
        cfg = sy.excitation_configs(n_orb, n_e, n_e, count, seed + 3)  # HF + singles + doubles: many connections
        nq = cfg.shape[1]
        assert nq == (2 * n_orb + 7) // 8
        terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), nq, 1)
        targets = pp.PreparedTargets(_dev(cfg, cuda_device), nq, distinct=True)
        x = torch.from_numpy(sy.random_psi(targets.n, seed + 1)).to(cuda_device)
        y = torch.from_numpy(sy.random_psi(targets.n, seed + 2)).to(cuda_device)
        d1, d2 = pp.apply_dual(terms, targets, torch.view_as_real(x), torch.view_as_real(y))
        for got, vec in ((d1, x), (d2, y)):
            ref = _reference(terms, targets, vec, pp.MODE_ADJOINT)
            diff = np.abs(torch.view_as_complex(got).cpu().numpy() - ref)
            np.testing.assert_array_less(diff, 1e-12 * _scale(ref, 10.0, vec) + 1e-300)
