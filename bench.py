#!/usr/bin/env python3
"""
apply_within forward + adjoint throughput on one MI355X, in matrix elements per second.

A step is one differentiable application of a seeded synthetic Hamiltonian to a seeded config set:
build the index over the set (hash index, plus the core indexes of the structured path when the
table has pair excitations), forward psi_j = H psi_i, then the adjoint grad_i = H^H g for a random
upstream gradient g through torch.autograd.  Molecular tables run the structured path (pair / single
core groups, folded diagonal); tables without pair excitations run the enumeration with hash
lookups, as a gather when the operator is Hermitian on its own set.  The term table is compiled once
outside the timed region (it belongs to the Hamiltonian, like the reference's one-off _prepare_data).  A matrix element is a contribution: a (config, term) pair whose application
succeeds and whose target is in configs_j; a step computes M_hit of them forward and M_hit backward.

Workloads (--workload):
  molecular52 (default): 52 spin orbitals, 7a+7b, |(pq|rs)| < 1e-10 dropped (669,579 terms),
                         1,000,000 configs = HF + all singles/doubles + seeded triples/quadruples,
                         configs_j = configs_i.
  hubbard128:            open 1-D Hubbard chain, 64 sites (128 spin orbitals), t=1, U=4, half
                         filling, 4,000,000 configs (random seeds + single hops), configs_j = configs_i.
  molecular20:           20 spin orbitals, 5a+5b, 10,000 configs uniform in the sector.
  batch64:               64 seeded molecular instances, forward + backward in one batched pass.
  lanczos52:             Lanczos iterations on the molecular52 basis (application + vector step).
  pt2_52:                the perturbation step (E0, find_relative, application into the external set).
  sampler52:             the config producer: unique sampling of 1M configs from a 26-site network.
  vmc76:                 one VMC training step: 76 spin orbitals, 13a+13b, 3.06M terms, 2M configs,
                         psi from the electron up/down network (float32), energy fwd+bwd through the
                         adjoint, network backward, Adam step.

Prints one JSON line with the metric, a roofline entry for the forward kernel (algorithmic bytes per
launch / its HIP-event time) and the reference CPU op timed on this host on a bounded sample.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import pathlib
import platform
import subprocess
import sys
import tempfile
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "qmb-public_amd"), str(ROOT)]

HBM_PEAK_GBPS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str) -> dict[str, object]:
    import numpy as np  # pylint: disable=import-outside-toplevel

    from qmb_hip import synthetic as sy  # pylint: disable=import-outside-toplevel

    t0 = time.time()
    if name == "molecular52":
        site, kind, coef = sy.molecular_terms(26, 2001, threshold=1e-10)
        configs = sy.excitation_configs(26, 7, 7, 1_000_000, 2002)
        pc, seeds, sample, occ = 1, (2003, 2004), 300, 14
        desc = {"workload": "molecular52", "spin_orbitals": 52, "electrons": "7a+7b", "n_configs": 1_000_000,
                "configs": "HF + singles + doubles + seeded triples/quadruples", "threshold": 1e-10}
    elif name == "hubbard128":
        site, kind, coef = sy.terms_from_dict(sy.hubbard_dict(64, 1, 1.0, 4.0))
        configs = sy.hubbard_chain_configs(64, 4_000_000, 3000)
        pc, seeds, sample, occ = 1, (3001, 3002), 200_000, 64
        desc = {"workload": "hubbard128", "sites": 64, "spin_orbitals": 128, "t": 1.0, "U": 4.0, "filling": "half",
                "n_configs": int(len(configs))}
    elif name == "molecular20":
        site, kind, coef = sy.molecular_terms(10, 1001)
        configs = sy.sector_configs(10, 5, 5, 10_000, 1002)
        pc, seeds, sample, occ = 1, (1003, 1004), 10_000, 10
        desc = {"workload": "molecular20", "spin_orbitals": 20, "electrons": "5a+5b", "n_configs": 10_000}
    else:
        raise SystemExit(f"unknown workload {name}")
    order = np.argsort(-np.linalg.norm(coef.numpy(), axis=1), kind="stable")
    site, kind, coef = site[order], kind[order], coef[order]
    psi = sy.random_psi(len(configs), seeds[0])
    grad = sy.random_psi(len(configs), seeds[1])
    log(f"workload {name}: {len(site)} terms, {len(configs)} configs x {configs.shape[1]} bytes, built in {time.time() - t0:.1f}s")
    return {"site": site, "kind": kind, "coef": coef, "configs": configs, "psi": psi, "grad": grad, "pc": pc,
            "sample": min(sample, len(configs)), "desc": desc, "max_occupied": occ}


def source_sha() -> str:
    """SHA-256 (first 16 hex digits) of the library's sources: a profile's counters describe the
    build they were taken on only."""
    import hashlib  # pylint: disable=import-outside-toplevel

    h = hashlib.sha256()
    csrc = ROOT / "qmb-public_amd" / "csrc"
    for path in sorted(list(csrc.glob("*.hip")) + list(csrc.glob("*.hpp")) + [ROOT / "include" / "qmb_hip.h"]):
        h.update(path.name.encode())
        h.update(path.read_bytes())
    return h.hexdigest()[:16]


def read_traffic(workload: str, kernel: str) -> tuple[float | None, float | None, str]:
    """(HBM bytes per launch, L2 hit rate) of `kernel` from profiles/<kernel>_<workload>.md, taken
    only when that summary was made on the current sources (its `source_sha:` line)."""
    path = ROOT / "profiles" / f"{kernel}_{workload}.md"
    if not path.exists():
        return None, None, "no profile"
    traffic = hit = sha = None
    for line in path.read_text().splitlines():
        if line.startswith("source_sha:"):
            sha = line.split(":", 1)[1].strip()
        elif line.startswith("hbm_bytes_per_launch:"):
            traffic = float(line.split(":", 1)[1])
        elif line.startswith("l2_hit_rate:"):
            hit = float(line.split(":", 1)[1])
    if sha != source_sha():
        return None, None, f"profile made on other sources ({sha})"
    return traffic, hit, str(path.relative_to(ROOT))


def structured_model(pp: object, terms: object, targets: object, nq: int, n: int) -> dict[str, object]:
    """Bytes the structured application must move at least once (synchronises to size the core
    indexes).  Per kernel launch:
      pair_kernel:   4 per group (its start) + per group of >= 2 entries its first entry's config
                     words (8 W: every row of a group of one core has that core) + per entry of such a
                     group: the entry (8, row and removed sites) and its row's amplitude (16); the
                     coefficient table once (8 or 16 per entry of n_pairs^2); the output once (16 per
                     row);
      single_kernel: the same over the order-1 index, with the single-excitation tables (n_bits^2 +
                     n_bits^3 entries) once;
      diag_apply:    per row its config words, amplitude and output (8 W + 32); the folded table once.
    Per core index build (packed entries): every entry emitted (8 written), sorted once (8 read + 8
    written), split and flagged (8 read; 4 key + 8 value + 4 flag written; the config words of each
    entry of a group of >= 2 entries read, 8 W), its flag read again for the starts (4), and 4 per
    group start -- the sort's extra passes, the per-chunk head counts and the hash index are not
    counted."""
    words = (nq + 7) // 8
    out: dict[str, object] = {}
    coef_b = 8 if terms.pairs_desc.real else 16
    n_bits = 8 * nq

    def walk(st: dict[str, int]) -> int:
        return 4 * st["groups"] + 8 * words * st["multi_groups"] + 24 * st["multi_entries"]

    def build(st: dict[str, int]) -> int:
        return 52 * st["entries"] + 8 * words * st["multi_entries"] + 4 * st["groups"]

    st2 = pp.core_stats(targets, 2)
    out["pair_kernel"] = dict(st2, bytes=walk(st2) + coef_b * terms.pairs_desc.n_pairs ** 2 + 16 * n)
    index_bytes = build(st2)
    if terms.singles_desc is not None:
        st1 = pp.core_stats(targets, 1)
        sb = 8 if terms.singles_desc.real else 16
        out["single_kernel"] = dict(st1, bytes=walk(st1) + sb * (n_bits ** 2 + n_bits ** 3) + 16 * n)
        index_bytes += build(st1)
    out["diag_apply_kernel"] = {"bytes": n * (8 * words + 32) + 16 * (1 + n_bits + terms.pairs_desc.n_pairs)}
    out["core_index_build"] = {"bytes": index_bytes}
    return out


def hop_model(terms: object, targets: object, nq: int, n: int) -> dict[str, object]:
    """Bytes the hop path must move at least once (reads the strings indexes' sizes).  Per launch:
      hop_kernel (one parity): per entry of a group of >= 2 rows its group range (8), row (4) and
                     config words (8 W), its row's amplitude (16) and output, read and written (32);
                     the single-excitation table (n_bits^2 entries) once;
      diag_scale_kernel: per row its diagonal value, amplitude and output (16 + 16 + 32).
    Per set (once for the step's applications): the diagonal values, per row its config words read
    and its value written (8 W + 16); per strings index build (one parity): per row its config words
    read (8 W), key and row emitted (8), sorted (8 read + 8 written), group and start found (8); per
    entry of a group of >= 2 rows its compacted row, range, config words and string written (20 +
    8 W)."""
    words = (nq + 7) // 8
    n_bits = 8 * nq
    sb = 8 if terms.singles_desc.real else 16
    out: dict[str, object] = {}
    build = 0
    for parity in (0, 1):
        st = targets.strings_stats(parity)
        out[f"hop_kernel_{parity}"] = dict(st, bytes=st["multi_entries"] * (8 + 4 + 8 * words + 16 + 32) + sb * n_bits ** 2)
        build += n * (8 * words + 8 + 16 + 8) + st["multi_entries"] * (20 + 8 * words)
    out["diag_scale_kernel"] = {"bytes": n * 64}
    out["strings_build"] = {"bytes": build + n * (8 * words + 16)}
    return out


def host_cpu_model() -> str:
    """The host CPU's model name from /proc/cpuinfo (the machine type where that is unreadable)."""
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.machine()


def sample_rows(n_total: int, n_sample: int, seed: int = 7001) -> "object":
    """A seeded random sample of configs_i rows (sorted indices), the CPU baseline's workload."""
    import numpy as np  # pylint: disable=import-outside-toplevel

    if n_sample >= n_total:
        return np.arange(n_total)
    return np.sort(np.random.default_rng(seed).choice(n_total, size=n_sample, replace=False))


def cpu_baseline(w: dict[str, object], rows: object, m_hit_sample: int, m_hit_per_row_set: float) -> dict[str, object]:
    """The reference CPU op (oracle/_ref) applying a seeded random sample of configs_i rows against
    all of configs_j: one warm-up run, then the median of five (std::chrono-like wall time of each
    op call).  Beside it, the same sample through the C port split over the host threads."""
    import numpy as np  # pylint: disable=import-outside-toplevel

    configs = w["configs"]
    sub, sub_psi = configs[rows], w["psi"][rows]
    n = len(sub)
    nq = configs.shape[1]
    pc = int(w["pc"])
    ref = ROOT / "oracle" / "_ref" / f"qmb_ref_{nq}_{pc}.so"
    sample_desc = (f"apply_within of {n} seeded random configs_i rows against all {len(configs)} configs_j "
                   f"({m_hit_sample / n:.0f} contributions per sampled row, {m_hit_per_row_set:.0f} per row over the set); "
                   "1 warm-up + median of 5")
    with tempfile.TemporaryDirectory() as d:
        inp = pathlib.Path(d) / "in.npz"
        out = pathlib.Path(d) / "out.npz"
        np.savez(inp, configs_i=sub, psi_i=sub_psi, configs_j=configs, site=w["site"].numpy(), kind=w["kind"].numpy(),
                 coef=w["coef"].numpy(), particle_cut=pc)
        from oracle import oracle as orc  # pylint: disable=import-outside-toplevel

        args = (pc, sub, sub_psi, configs, w["site"].numpy(), w["kind"].numpy(), w["coef"].numpy())
        result = None
        if ref.exists():
            env = dict(os.environ, OMP_NUM_THREADS="1")
            proc = subprocess.run([sys.executable, str(ROOT / "oracle" / "ref_runner.py"), "apply_within", str(inp), str(out),
                                   "--repeat", "6", "--threads", "1"], env=env, capture_output=True, text=True, check=False)
            if proc.returncode == 0:
                times = np.load(out)["times"]
                t = float(np.median(times[1:]))
                result = {"value": m_hit_sample / t, "unit": "matrix elements/s", "cores": 1, "kind": "reference",
                          "sample": sample_desc + " (qmb/_hamiltonian_cpu.cpp built from source, -O3 -ffast-math, 1 thread)",
                          "seconds": t, "runs_s": [float(x) for x in times]}
            else:
                log(f"reference op failed: {proc.stderr[-500:]}")
        if result is None:
            times = []
            for _ in range(6):
                t0 = time.perf_counter()
                orc.apply_within(*args)
                times.append(time.perf_counter() - t0)
            t = float(np.median(times[1:]))
            result = {"value": m_hit_sample / t, "unit": "matrix elements/s", "cores": 1, "kind": "port",
                      "sample": sample_desc + " (oracle/qmb_oracle.c, 1 thread)", "seconds": t, "runs_s": times}
        # The same sample over the host cores this process may use (at most 16, one GPU's share of
        # the host): the port's slices of configs_i on threads, reported beside the serial figure.
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        times, hit = [], 0
        for _ in range(6):
            t0 = time.perf_counter()
            _, (_, hit) = orc.apply_within_threads(*args, threads)
            times.append(time.perf_counter() - t0)
        t = float(np.median(times[1:]))
        if hit != m_hit_sample:
            log(f"WARNING: the threaded CPU port found {hit} contributions in the sample, the GPU count {m_hit_sample}")
        result["host_cpu"] = host_cpu_model()
        result["host_cores_allowed"] = len(os.sched_getaffinity(0))
        result["all_cores"] = {"value": hit / t, "cores": threads, "kind": "port", "seconds": t, "m_hit": hit,
                               "sample": "the same sample, oracle/qmb_oracle.c apply_within_threads (configs_i split over threads)"}
        log(f"CPU baseline: {result['kind']} 1 thread {result['value']:.3e} elements/s; "
            f"port on {threads} threads {hit / t:.3e} elements/s")
        return result


def build_batch_instances(seed0: int = 4000, count: int = 64) -> list[dict[str, object]]:
    """The batch64 workload (qmb_hip.synthetic.batch_instances), built on up to 16 host threads."""
    from qmb_hip import synthetic as sy  # pylint: disable=import-outside-toplevel

    threads = max(1, min(16, len(os.sched_getaffinity(0))))

    def progress(done: int, total: int) -> None:
        if done % 8 == 0:
            log(f"batch64: generated {done}/{total} instances")

    return sy.batch_instances(seed0, count, threads, progress)


def run_batch(args: argparse.Namespace, dev: object) -> int:
    """64 molecular instances, forward + adjoint of all of them: one persistent batched launch per
    direction (graph-captured) against per-instance structured applications."""
    import numpy as np  # pylint: disable=import-outside-toplevel
    import torch  # pylint: disable=import-outside-toplevel

    from qmb_hip import batch as qb  # pylint: disable=import-outside-toplevel
    from qmb_hip import prepared as pp  # pylint: disable=import-outside-toplevel

    t0 = time.time()
    raw = build_batch_instances()
    log(f"batch64: generated {len(raw)} instances in {time.time() - t0:.1f}s")
    instances, psis, grads = [], [], []
    m_gen = m_hit = 0
    algo_bytes = 0
    for w in raw:
        configs = torch.from_numpy(w["configs"]).to(dev)
        nq = configs.shape[1]
        site, kind, coef = (t.contiguous().to(dev) for t in (w["site"], w["kind"], w["coef"]))
        terms = pp.PreparedTerms(site, kind, coef, nq, 1)
        targets = pp.PreparedTargets(configs, nq, distinct=True, max_occupied=sum(w["electrons"]))
        g, h = pp.count(terms, targets, configs)
        m_gen += g
        m_hit += h
        n, t = configs.shape[0], site.shape[0]
        algo_bytes += 2 * n * (nq + 16) + t * 28 + g * nq + h * 16
        instances.append((terms, targets))
        if len(instances) % 8 == 0:
            log(f"batch64: prepared {len(instances)}/{len(raw)} instances")
        psis.append(torch.from_numpy(w["psi"]).to(dev))
        grads.append(torch.from_numpy(w["grad"]).to(dev))
    torch.cuda.synchronize()
    log(f"batch64: prepared in {time.time() - t0:.1f}s; M_gen={m_gen} M_hit={m_hit}")
    # forward and backward of every instance in one launch: the instances are Hermitian, so the
    # backward A^H g is the same map applied to the upstream gradient
    b = qb.Batch(instances, two_vectors=True)
    b.set_inputs(psis, grads)
    waves = b.tune()
    # every (waves, split) the autotuner timed on this batch, for the record
    tuning = [entry for entry in qb.TUNER.cache.values() if "split" in entry["params"]]
    tried = sorted(tuning[-1]["tried"], key=lambda x: x[1])[:8] if tuning else []
    b1 = qb.Batch(instances)
    b1.set_inputs(psis)
    b1.waves = waves
    stream = torch.cuda.current_stream(dev)

    def time_it(fn: object, reps: int) -> float:
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        e.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(e) / reps

    # per-instance: the structured application of each instance, forward then adjoint
    outs = [torch.zeros((t.n, 2), dtype=torch.float64, device=dev) for _, t in instances]
    vins = [torch.view_as_real(p).contiguous() for p in psis]
    gins = [torch.view_as_real(g).contiguous() for g in grads]

    def per_instance() -> None:
        for (terms, targets), o, v, g in zip(instances, outs, vins, gins):
            pp.apply(terms, targets, targets.configs, v, pp.MODE_FORWARD, out=o)
            pp.apply(terms, targets, targets.configs, g, pp.MODE_ADJOINT, out=o)

    def per_instance_dual() -> None:
        # the same work as the batched launch, instance by instance: each kernel applies A^H to psi
        # and to g in one pass (forward and backward of a Hermitian instance)
        for (terms, targets), v, g in zip(instances, vins, gins):
            pp.apply_dual(terms, targets, v, g)

    def batched() -> None:
        b.launch_both()

    def batched_two_launches() -> None:
        b1.launch(pp.MODE_FORWARD)
        b1.launch(pp.MODE_ADJOINT)

    steps = max(args.steps, 1)
    eager_inst = time_it(per_instance, steps)
    eager_dual = time_it(per_instance_dual, steps)
    eager_batch = time_it(batched, steps)
    two_launches = time_it(batched_two_launches, steps)
    graphs = {}
    for name, fn in (("per_instance", per_instance), ("per_instance_dual", per_instance_dual), ("batched", batched)):
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(stream)
        with torch.cuda.stream(side):
            fn()
        stream.wait_stream(side)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            fn()
        graphs[name] = g
    graph_inst = time_it(graphs["per_instance"].replay, steps)
    graph_dual = time_it(graphs["per_instance_dual"].replay, steps)
    graph_batch = time_it(graphs["batched"].replay, steps)
    fwd_only = time_it(lambda: b1.launch(pp.MODE_FORWARD), steps)
    # check: both batched outputs equal the per-instance forward and adjoint
    b.launch_both()
    worst = 0.0
    for k, ((terms, targets), v, gv) in enumerate(zip(instances, vins, gins)):
        lo, hi = b.rows[k]
        for got, vec, mode in ((b.out, v, pp.MODE_FORWARD), (b.out2, gv, pp.MODE_ADJOINT)):
            ref = pp.apply(terms, targets, targets.configs, vec, mode)
            worst = max(worst, float((got[lo:hi] - ref).abs().max()) / max(float(ref.abs().max()), 1e-300))
    for k, ((terms, targets), v, gv) in enumerate(zip(instances, vins, gins)):
        lo, hi = b.rows[k]
        d1, d2 = pp.apply_dual(terms, targets, v, gv)
        for got, ref in ((d1, b.out[lo:hi]), (d2, b.out2[lo:hi])):
            worst = max(worst, float((got - ref).abs().max()) / max(float(ref.abs().max()), 1e-300))
    log(f"batch64: per-instance eager {eager_inst:.2f} ms, graph {graph_inst:.2f} ms; per-instance two-vector eager "
        f"{eager_dual:.2f} ms, graph {graph_dual:.2f} ms; batched (one launch for forward and backward) eager {eager_batch:.2f} ms, "
        f"graph {graph_batch:.2f} ms; batched as two launches {two_launches:.2f} ms (waves {waves}); max rel diff {worst:.2e}")
    # bytes the structured kernels must move at least once, summed over the instances (one direction)
    model_bytes = 0
    for terms, targets in instances:
        parts_b = structured_model(pp, terms, targets, targets.n_qubytes, targets.n)
        model_bytes += sum(v["bytes"] for key, v in parts_b.items() if key != "core_index_build")
    achieved = model_bytes / (fwd_only * 1e-3) / 1e9
    if achieved > HBM_PEAK_GBPS:
        print(f"bench.py: refusing to report {achieved:.0f} GB/s for the batch kernel, above the {HBM_PEAK_GBPS:.0f} GB/s peak",
              file=sys.stderr)
        return 3
    result = {
        "metric": "apply_within fwd+bwd matrix elements/s",
        "value": 2 * m_hit / (graph_batch * 1e-3),
        "unit": "matrix elements/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 2,
        "ms_per_step": graph_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": "batch64", "instances": len(raw), "spin_orbitals": [w["n_so"] for w in raw],
                   "n_configs": [int(len(w["configs"])) for w in raw], "n_terms_total": int(sum(len(w["site"]) for w in raw)),
                   "M_gen": m_gen, "M_hit": m_hit, "waves": waves, "split": b.split,
                   "tuning_fastest": [{"waves": p["waves"], "split": p["split"], "ms": ms} for p, ms in tried]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": None, "kernel": "batch_kernel<adjoint, one vector> (one direction, all instances)", "kernel_ms": fwd_only,
                     "algorithmic_bytes": model_bytes,
                     "model": "structured path, bytes moved at least once: group starts, one config per group of >= 2 entries, "
                              "the entries of those groups with their rows' amplitudes, the coefficient tables, the outputs "
                              "(summed over the instances)",
                     "lookup_model": {"what": "one key read per generated connection (the enumeration path's bound, not this path's)",
                                      "bytes": algo_bytes, "gbps_equivalent": algo_bytes / (fwd_only * 1e-3) / 1e9}},
        "batch": {"batched_graph_ms": graph_batch, "batched_eager_ms": eager_batch, "batched_two_launches_ms": two_launches,
                  "per_instance_eager_ms": eager_inst,
                  "per_instance_graph_ms": graph_inst, "speedup_vs_per_instance_eager": eager_inst / graph_batch,
                  "speedup_vs_per_instance_graph": graph_inst / graph_batch, "max_rel_diff_vs_per_instance": worst,
                  "per_instance_two_vector_eager_ms": eager_dual, "per_instance_two_vector_graph_ms": graph_dual,
                  "batching_only_speedup": graph_dual / graph_batch},
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)
    return 0


def run_lanczos(args: argparse.Namespace, dev: object) -> int:
    """Lanczos on the molecular52 basis (qmb/haar.py:107-142): every iteration is one application
    through the structured path plus the fused vector step (alpha, three-term update, norm,
    normalisation) and one 16-byte read-back; timed per iteration, with its parts timed alone and
    the reference's unfused tensor arithmetic for the same vector work beside it."""
    import numpy as np  # pylint: disable=import-outside-toplevel
    import torch  # pylint: disable=import-outside-toplevel

    from qmb_hip import Hamiltonian  # pylint: disable=import-outside-toplevel
    from qmb_hip import krylov as kr  # pylint: disable=import-outside-toplevel
    from qmb_hip import prepared as pp  # pylint: disable=import-outside-toplevel

    w = build_workload("molecular52")
    configs = torch.from_numpy(w["configs"]).to(dev)
    ham = Hamiltonian(tuple(t.contiguous().to(dev) for t in (w["site"], w["kind"], w["coef"])), kind="fermi")
    psi = torch.from_numpy(w["psi"]).to(dev)
    t0 = time.perf_counter()
    op = kr.WithinOperator(ham, configs)
    op(psi)  # builds the core indexes
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    m_gen, m_hit = pp.count(op.terms, op.targets, configs)
    stream = torch.cuda.current_stream(dev)
    steps = max(args.steps, 2)
    gen = kr.lanczos(op, psi, threshold=1e-12)
    for _ in range(max(args.warmup, 1)):
        next(gen)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        alpha, beta, v = next(gen)
    torch.cuda.synchronize()
    iter_ms = (time.perf_counter() - t0) * 1e3 / steps
    energy, _ = kr.lowest_ritz_pair(alpha, beta, v, dev)

    def timed(fn: object, reps: int = 10) -> float:
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    vk, vprev = v[-1], v[-2]
    wvec = op(vk)
    fused = kr.LanczosStep(dev)
    scratch = wvec.clone()

    def fused_step() -> None:
        scratch.copy_(wvec)
        fused(vk, vprev, 0.5, scratch)

    def unfused_step() -> None:
        # the reference's arithmetic between two applications (qmb/haar.py:130-138)
        a = (wvec.conj() @ vk).real
        x = wvec - a * vk - 0.5 * vprev
        nrm = torch.linalg.norm(x)  # pylint: disable=not-callable
        _ = bool(nrm < 1e-12)
        _ = x / nrm

    copy_ms = timed(lambda: scratch.copy_(wvec))
    apply_ms = timed(lambda: op(vk))
    fused_ms = timed(fused_step) - copy_ms
    unfused_ms = timed(unfused_step)
    log(f"lanczos52: {iter_ms:.3f} ms/iteration (application {apply_ms:.3f} ms, fused vector step {fused_ms:.3f} ms, "
        f"reference tensor arithmetic {unfused_ms:.3f} ms); lowest Ritz value after {len(alpha)} iterations {float(energy):.10f}")
    result = {
        "metric": "Lanczos iterations/s",
        "value": 1e3 / iter_ms,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": iter_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": "lanczos52", "basis": w["desc"], "M_gen": m_gen, "M_hit": m_hit},
        "lanczos": {"application_ms": apply_ms, "fused_vector_step_ms": fused_ms, "reference_tensor_arithmetic_ms": unfused_ms,
                    "setup_s": setup_s, "iterations": len(alpha), "lowest_ritz_value": float(energy),
                    "matrix_elements_per_s": m_hit / (iter_ms * 1e-3)},
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)
    return 0


def run_pt2(args: argparse.Namespace, dev: object) -> int:
    """The perturbation step (qmb/pert.py:37-63) on the molecular52 basis: E0 on the set,
    find_relative with count = len(configs), the application into the selected external set and its
    diagonal, each timed with events."""
    import torch  # pylint: disable=import-outside-toplevel

    from qmb_hip import Hamiltonian  # pylint: disable=import-outside-toplevel
    from qmb_hip import krylov as kr  # pylint: disable=import-outside-toplevel
    from qmb_hip import pt2  # pylint: disable=import-outside-toplevel

    w = build_workload("molecular52")
    configs = torch.from_numpy(w["configs"]).to(dev)
    ham = Hamiltonian(tuple(t.contiguous().to(dev) for t in (w["site"], w["kind"], w["coef"])), kind="fermi")
    psi = torch.from_numpy(w["psi"]).to(dev)
    stream = torch.cuda.current_stream(dev)
    for _ in range(max(args.warmup, 1)):
        res = pt2.perturbation(ham, configs, psi)
    torch.cuda.synchronize()
    steps = max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        res = pt2.perturbation(ham, configs, psi)
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) * 1e3 / steps

    def timed(fn: object) -> float:
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b)

    op = kr.WithinOperator(ham, configs)
    parts = {"energy0_ms": timed(lambda: op(psi)),
             "find_relative_ms": timed(lambda: ham.find_relative(configs, psi, configs.size(0), configs)),
             "apply_into_external_ms": timed(lambda: pt2._apply_foreign(ham, configs, psi, res.other)),  # pylint: disable=protected-access
             "diagonal_term_ms": timed(lambda: ham.diagonal_term(res.other))}
    log(f"pt2_52: {step_ms:.1f} ms per step, parts {parts}; E0 {res.energy0:.10f} E2 {res.energy2:.10f} over {res.other.size(0)} configs")
    result = {
        "metric": "perturbation steps/s",
        "value": 1e3 / step_ms,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": "pt2_52", "basis": w["desc"], "count": int(configs.size(0)), "external_configs": int(res.other.size(0))},
        "pt2": dict(parts, energy0=res.energy0, energy2=res.energy2),
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)
    return 0


def vmc_energy_reference(terms: tuple[object, object, object], configs: object, psi: object) -> dict[str, object]:
    """The reference CPU op (oracle/_ref, else the oracle) applied on a small config set:
    <psi|H|psi> / <psi|psi> there and the time it took (1 thread)."""
    import numpy as np  # pylint: disable=import-outside-toplevel

    nq = configs.shape[1]
    ref = ROOT / "oracle" / "_ref" / f"qmb_ref_{nq}_1.so"
    site, kind, coef = (np.ascontiguousarray(t.numpy()) for t in terms)
    with tempfile.TemporaryDirectory() as d:
        inp = pathlib.Path(d) / "in.npz"
        out = pathlib.Path(d) / "out.npz"
        np.savez(inp, configs_i=configs, psi_i=psi, configs_j=configs, site=site, kind=kind, coef=coef, particle_cut=1)
        hpsi, seconds, kind_name = None, 0.0, "port"
        if ref.exists():
            env = dict(os.environ, OMP_NUM_THREADS="1")
            proc = subprocess.run([sys.executable, str(ROOT / "oracle" / "ref_runner.py"), "apply_within", str(inp), str(out),
                                   "--repeat", "1", "--threads", "1"], env=env, capture_output=True, text=True, check=False)
            if proc.returncode == 0:
                data = np.load(out)
                hpsi, seconds, kind_name = data["result"], float(data["times"][0]), "reference"
            else:
                log(f"reference op failed: {proc.stderr[-500:]}")
        if hpsi is None:
            from oracle import oracle as orc  # pylint: disable=import-outside-toplevel

            t0 = time.perf_counter()
            hpsi, _, _ = orc.apply_within(1, configs, psi, configs, site, kind, coef)
            seconds = time.perf_counter() - t0
    energy = float(np.vdot(psi, hpsi).real / np.vdot(psi, psi).real)
    return {"energy": energy, "seconds": seconds, "kind": kind_name}


def run_vmc(args: argparse.Namespace, dev: object) -> int:
    """One VMC training step (qmb/vmc.py:104-118) at 76 spin orbitals, 13a+13b, |(pq|rs)| < 1e-8
    dropped (3.06M terms), over 2M configs: the electron up/down network in float32 (each site network
    and the phase network 4 layers of width 512 by default), psi cast to complex128, the energy
    <psi|H|psi>/<psi|psi> forward and backward through the adjoint kernel, the network's backward
    (in row chunks) and an Adam step.  The config set and its index are the cycle's, built once per
    cycle of local steps, outside the timed steps (reported as setup)."""
    import numpy as np  # pylint: disable=import-outside-toplevel
    import torch  # pylint: disable=import-outside-toplevel

    from qmb_hip import Hamiltonian  # pylint: disable=import-outside-toplevel
    from qmb_hip import mlp  # pylint: disable=import-outside-toplevel
    from qmb_hip import prepared as pp  # pylint: disable=import-outside-toplevel
    from qmb_hip import synthetic as sy  # pylint: disable=import-outside-toplevel
    from qmb_hip import vmc  # pylint: disable=import-outside-toplevel

    n_orbit, n_e, n_configs = 38, 13, 2_000_000
    hidden = tuple(int(h) for h in args.hidden.split(",")) if args.hidden else (512, 512, 512)
    t0 = time.time()
    site, kind, coef = sy.molecular_terms(n_orbit, 5001, threshold=1e-8)
    order = np.argsort(-np.linalg.norm(coef.numpy(), axis=1), kind="stable")
    site, kind, coef = site[order], kind[order], coef[order]
    configs_np = sy.excitation_configs(n_orbit, n_e, n_e, n_configs, 5002)
    log(f"vmc76: {len(site)} terms, {len(configs_np)} configs built in {time.time() - t0:.1f}s")
    ham = Hamiltonian(tuple(t.contiguous().to(dev) for t in (site, kind, coef)), kind="fermi")
    configs = torch.from_numpy(configs_np).to(dev)
    torch.manual_seed(5003)
    net = mlp.WaveFunctionElectronUpDown(double_sites=2 * n_orbit, physical_dim=2, is_complex=True, spin_up=n_e, spin_down=n_e,
                                         hidden_size=hidden, ordering=+1).to(dev)
    opt = torch.optim.Adam(net.parameters(), lr=1e-3)
    chunk = args.chunk
    stream = torch.cuda.current_stream(dev)

    def timed(fn: object) -> float:
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b)

    holder: dict[str, object] = {}

    def setup() -> None:
        es = vmc.EnergySet(ham, configs, max_occupied=2 * n_e)
        es.targets.cores(2)
        es.targets.cores(1)
        holder["set"] = es

    setup_ms = timed(setup)
    energy_set = holder["set"]
    energies: list[float] = []

    def step() -> None:
        opt.zero_grad()
        energies.append(vmc.energy_backward(net, energy_set, chunk))
        opt.step()

    for i in range(max(args.warmup, 1)):
        step()
        log(f"warmup {i + 1}: energy {energies[-1]:.10f}")
    torch.cuda.synchronize()
    steps = max(args.steps, 1)
    t1 = time.perf_counter()
    for i in range(steps):
        step()
        log(f"step {i + 1}/{steps}: energy {energies[-1]:.10f}")
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t1) * 1e3 / steps

    # the parts of a step, each timed alone with events
    psi_leaf = vmc.network_forward(net, configs, chunk).requires_grad_(True)

    def energy_fwd_bwd() -> None:
        psi_leaf.grad = None
        energy_set.energy(psi_leaf).backward()

    parts = {"network_forward_ms": timed(lambda: vmc.network_forward(net, configs, chunk)), "energy_forward_backward_ms": timed(energy_fwd_bwd)}
    grad_psi = psi_leaf.grad
    net.zero_grad()
    parts["network_backward_ms"] = timed(lambda: vmc.network_backward(net, configs, grad_psi, chunk))
    terms, targets = energy_set.terms, energy_set.targets
    vec = torch.view_as_real(psi_leaf.detach()).contiguous()
    out = torch.zeros_like(vec)
    fwd = []
    for _ in range(5):
        fwd.append(timed(lambda: pp.apply(terms, targets, configs, vec, pp.MODE_FORWARD, out=out)))
    forward_ms = float(np.median(fwd))
    m_gen, m_hit = pp.count(terms, targets, configs)
    nq = configs.size(1)
    # the structured path's bytes model (pair + single + diagonal kernels of one application); the
    # lookup model is kept as a labelled comparison only
    model = structured_model(pp, terms, targets, nq, n_configs)
    algo_bytes = sum(v["bytes"] for k, v in model.items() if k != "core_index_build")
    lookup_bytes = 2 * n_configs * (nq + 16) + len(site) * 28 + m_gen * nq + m_hit * 16
    achieved = algo_bytes / (forward_ms * 1e-3) / 1e9
    if achieved > HBM_PEAK_GBPS:
        print(f"bench.py: refusing to report {achieved:.0f} GB/s for the structured application, above the {HBM_PEAK_GBPS:.0f} GB/s "
              "peak; the model does not bound this path", file=sys.stderr)
        return 3
    log(f"vmc76: {step_ms:.1f} ms/step, setup {setup_ms:.1f} ms, parts {parts}, H forward {forward_ms:.2f} ms, M_hit {m_hit}")

    sampler_ms = None
    if args.sample:
        with torch.no_grad():  # psi of the sample feeds find_relative only
            sampler_ms = timed(lambda: net.generate_unique(n_configs, max(1, n_configs // chunk), seed=5004))
        log(f"generate_unique({n_configs}): {sampler_ms:.1f} ms")

    cpu = None
    if not args.no_cpu:
        # energy error on a set near the reference determinant: HF + 255 seeded singles/doubles of the set
        occ = sy.unpack_occupation(configs_np, 2 * n_orbit)
        hf = np.zeros(2 * n_orbit, dtype=np.uint8)
        hf[0:2 * n_e:2] = 1
        hf[1:2 * n_e:2] = 1
        level = (occ != hf).sum(axis=1) // 2
        rows_hf = np.flatnonzero(level == 0)
        rows_sd = np.flatnonzero((level >= 1) & (level <= 2))
        rows = np.sort(np.concatenate([rows_hf, np.random.default_rng(5005).choice(rows_sd, 255, replace=False)]))
        sub = configs_np[rows]
        sub_t = torch.from_numpy(sub).to(dev)
        psi_sub = vmc.network_forward(net, sub_t, chunk)
        e_gpu = float(vmc.EnergySet(ham, sub_t, max_occupied=2 * n_e).energy(psi_sub))
        ref = vmc_energy_reference((site, kind, coef), sub, psi_sub.cpu().numpy())
        m_sub = pp.count(terms, pp.PreparedTargets(sub_t, nq, distinct=True), sub_t)[1]
        cpu = {"value": m_sub / ref["seconds"], "unit": "matrix elements/s", "cores": 1, "kind": ref["kind"],
               "sample": f"apply_within on {len(sub)} configs (HF + 255 seeded singles/doubles of the set) into themselves, "
                         + ("qmb/_hamiltonian_cpu.cpp built from source, -O3 -ffast-math, 1 thread" if ref["kind"] == "reference"
                            else "oracle/qmb_oracle.c, 1 thread"),
               "seconds": ref["seconds"], "energy_cpu": ref["energy"], "energy_gpu": e_gpu,
               "energy_error": abs(e_gpu - ref["energy"])}
        log(f"energy on the sample: GPU {e_gpu:.15f} CPU {ref['energy']:.15f} ({ref['seconds']:.1f} s)")

    result = {
        "metric": "VMC training step time",
        "value": step_ms,
        "unit": "ms",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": step_ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 network, f64 Hamiltonian",
        "data": "synthetic (seeded Hamiltonian and configs, seeded network init)",
        "config": {"workload": "vmc76", "spin_orbitals": 2 * n_orbit, "electrons": f"{n_e}a+{n_e}b", "threshold": 1e-8,
                   "n_terms": len(site), "n_configs": n_configs, "configs": "HF + singles + doubles + seeded triples/quadruples",
                   "network": "WaveFunctionElectronUpDown", "hidden": list(hidden), "chunk": chunk, "optimizer": "Adam(lr=1e-3)",
                   "M_gen": m_gen, "M_hit": m_hit},
        "vmc": dict(parts, setup_ms=setup_ms, energies=energies, h_forward_ms=forward_ms, generate_unique_ms=sampler_ms,
                    h_elements_per_s=2 * m_hit / (parts["energy_forward_backward_ms"] * 1e-3)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": None, "kernel": "H forward (pair_kernel<W=2> + single_kernel<W=2> + diagonal)", "kernel_ms": forward_ms,
                     "algorithmic_bytes": algo_bytes,
                     "model": "structured path, bytes moved at least once per application (pair, single and diagonal kernels)",
                     "model_parts": model,
                     "lookup_model": {"what": "one key read per generated connection (the enumeration path's bound; the structured "
                                              "path performs no such lookups, so this is not its roofline)",
                                      "bytes": lookup_bytes, "gbps_equivalent": lookup_bytes / (forward_ms * 1e-3) / 1e9}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    return 0


def run_sampler(args: argparse.Namespace, dev: object) -> int:
    """The config producer on molecular52's shape (qmb/mlp.py:61-314): 26 sites, 7 + 7 electrons,
    one hidden layer of 512, float32; generate_unique of 1,000,000 distinct configs with their psi,
    and the forward alone on them."""
    import torch  # pylint: disable=import-outside-toplevel

    from qmb_hip import mlp  # pylint: disable=import-outside-toplevel

    torch.manual_seed(5001)
    net = mlp.WaveFunctionElectronUpDown(double_sites=52, physical_dim=2, is_complex=True, spin_up=7, spin_down=7, hidden_size=(512,),
                                         ordering=+1).to(dev)
    batch = 1_000_000
    steps = max(args.steps, 1)
    with torch.no_grad():
        for _ in range(max(args.warmup, 1)):
            configs, psi, _, _ = net.generate_unique(batch, seed=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            configs, psi, _, _ = net.generate_unique(batch, seed=2 + k)
        torch.cuda.synchronize()
        gen_ms = (time.perf_counter() - t0) * 1e3 / steps
    stream = torch.cuda.current_stream(dev)

    def timed(fn: object) -> float:
        with torch.no_grad():
            fn()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(steps):
                fn()
            b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    def per_site_forward(x: torch.Tensor) -> torch.Tensor:
        """The reference's forward arithmetic (qmb/mlp.py:169-205): one small network per site."""
        from qmb_hip import bitspack  # pylint: disable=import-outside-toplevel

        n = x.shape[0]
        xs = bitspack.unpack_int(x, size=1, last_dim=net.double_sites).view([n, net.sites, 2])
        xs = torch.index_select(xs, 1, net.ordering_reversed)
        xf = xs.to(torch.float32)
        number = torch.zeros((n, 2), dtype=torch.int64, device=x.device)
        total = torch.zeros(n, dtype=torch.float64, device=x.device)
        rows = torch.arange(n, device=x.device)
        for i, m in enumerate(net.amplitude):
            d = m(xf[:, :i].reshape([n, 2 * i])).view([n, 2, 2])
            ok_up = torch.stack([i - number[:, 0] < net.sites - net.spin_up, number[:, 0] < net.spin_up], dim=-1).unsqueeze(-1)
            ok_dn = torch.stack([i - number[:, 1] < net.sites - net.spin_down, number[:, 1] < net.spin_down], dim=-1).unsqueeze(-2)
            d = d + torch.where(torch.logical_and(ok_up, ok_dn), 0, -torch.inf)
            d = d - ((2 * d).exp().sum(dim=[-2, -1]).log() / 2)[:, None, None]
            xi = xs[:, i, :].to(torch.int64)
            total = total + d[rows, xi[:, 0], xi[:, 1]].double()
            number += xi
        phase = net.phase(xf.view([n, net.double_sites])).view([n])
        return torch.view_as_complex(torch.stack([total, phase.double()], dim=-1)).exp()

    fwd_ms = timed(lambda: net(configs))
    net.fused = False
    gemm_ms = timed(lambda: net(configs))
    net.fused = None
    per_site_ms = timed(lambda: per_site_forward(configs))
    log(f"sampler52: generate_unique {configs.size(0)} configs in {gen_ms:.1f} ms; forward on them: fused {fwd_ms:.2f} ms, "
        f"stacked GEMMs {gemm_ms:.2f} ms, per-site networks (the reference's arithmetic) {per_site_ms:.2f} ms")
    result = {
        "metric": "unique configs sampled/s",
        "value": configs.size(0) / (gen_ms * 1e-3),
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": gen_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-init network)",
        "config": {"workload": "sampler52", "sites": 26, "electrons": "7a+7b", "hidden": [512], "batch": batch,
                   "sampled": int(configs.size(0))},
        "sampler": {"generate_unique_ms": gen_ms, "forward_fused_ms": fwd_ms, "forward_stacked_gemm_ms": gemm_ms,
                    "forward_per_site_ms": per_site_ms, "forward_configs_per_s": configs.size(0) / (fwd_ms * 1e-3),
                    "forward_speedup_vs_per_site": per_site_ms / fwd_ms},
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)
    return 0


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--workload", default="molecular52", choices=["molecular52", "hubbard128", "molecular20", "batch64", "lanczos52", "pt2_52", "sampler52", "vmc76"])
    parser.add_argument("--hidden", default="", help="vmc76: network hidden widths, comma separated (default 512,512,512)")
    parser.add_argument("--chunk", type=int, default=65536, help="vmc76: rows per network backward chunk")
    parser.add_argument("--sample", action="store_true", help="vmc76: also time generate_unique of the 2M configs")
    parser.add_argument("--tile", type=int, default=0)
    parser.add_argument("--groups", type=int, default=0)
    parser.add_argument("--no-plain", action="store_true", help="skip the reference-scheme comparator")
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    parser.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                        help="also run a plain device-to-device copy of BYTES bytes three times (HBM counter calibration)")
    parser.add_argument("--gather-calibration", type=int, default=0, metavar="ROWS",
                        help="also gather ROWS random 64-byte rows of a 4 GiB table three times (random-request counter calibration)")
    parser.add_argument("--source-sha", action="store_true", help="print the library sources' hash (profiles record it) and exit")
    args = parser.parse_args()
    if args.source_sha:
        print(source_sha())
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for {args.gpus}); the kernels run on one GPU per process and are not "
              "sharded across GPUs.", file=sys.stderr)
        return 2

    import numpy as np  # pylint: disable=import-outside-toplevel
    import torch  # pylint: disable=import-outside-toplevel

    from qmb_hip import _lib  # pylint: disable=import-outside-toplevel
    from qmb_hip import autograd as ag  # pylint: disable=import-outside-toplevel
    from qmb_hip import ops  # pylint: disable=import-outside-toplevel
    from qmb_hip import prepared as pp  # pylint: disable=import-outside-toplevel

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP GPU")
    dev = torch.device("cuda:0")
    if args.workload == "batch64":
        return run_batch(args, dev)
    if args.workload == "lanczos52":
        return run_lanczos(args, dev)
    if args.workload == "pt2_52":
        return run_pt2(args, dev)
    if args.workload == "sampler52":
        return run_sampler(args, dev)
    if args.workload == "vmc76":
        return run_vmc(args, dev)
    w = build_workload(args.workload)
    configs = torch.from_numpy(w["configs"]).to(dev)
    nq = configs.shape[1]
    pc = int(w["pc"])
    site, kind, coef = (t.contiguous().to(dev) for t in (w["site"], w["kind"], w["coef"]))
    psi = torch.from_numpy(w["psi"]).to(dev)
    grad = torch.from_numpy(w["grad"]).to(dev)
    n = configs.shape[0]
    n_terms = site.shape[0]
    terms = pp.PreparedTerms(site, kind, coef, nq, pc)
    torch.cuda.synchronize()
    valid_terms, diag_terms, bad_terms = terms.counts()
    assert bad_terms == 0

    occ = int(w["max_occupied"])

    def step() -> torch.Tensor:
        # configs are distinct by construction; occ bounds the occupied sites of every row
        targets = pp.PreparedTargets(configs, nq, distinct=True, max_occupied=occ)
        x = psi.detach().requires_grad_(True)
        y = ag.apply_within(x, configs, terms, targets, args.tile, args.groups)
        y.backward(grad)
        return x.grad

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup}")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = wall * 1e3 / args.steps
    log(f"timed {args.steps} steps: {step_ms:.3f} ms/step (events {ev0.elapsed_time(ev1) / args.steps:.3f} ms)")

    # The forward application (the dominant work), timed alone with events on the stream it runs on,
    # then its parts: the pair-excitation kernel and the enumeration kernel over the other terms.
    targets = pp.PreparedTargets(configs, nq, distinct=True, max_occupied=occ)
    m_gen, m_hit = pp.count(terms, targets, configs, args.tile, args.groups)
    symmetric = pp.can_use_symmetric(terms, targets, configs)
    use_pairs = pp.can_use_pairs(terms, targets, configs)
    direct = n_terms <= pp.DIRECT_MAX_TERMS
    vec = torch.view_as_real(psi).contiguous()
    gvec = torch.view_as_real(grad).contiguous()
    out = torch.zeros((n, 2), dtype=torch.float64, device=dev)
    reps = max(3, min(20, args.steps))
    lib = _lib.library()

    def timed(fn: object) -> float:
        ts = []
        for _ in range(reps):
            out.zero_()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ts.append(a.elapsed_time(b))
        return float(np.median(ts))

    kernel_ms = timed(lambda: pp.apply(terms, targets, configs, vec, pp.MODE_FORWARD, out=out, tile=args.tile, groups=args.groups,
                                       zero=False))
    adjoint_ms = timed(lambda: pp.apply(terms, targets, configs, gvec, pp.MODE_ADJOINT, out=out, tile=args.tile, groups=args.groups,
                                        zero=False))
    parts = None
    use_hops = not use_pairs and pp.can_use_hops(terms, targets, configs)
    if use_hops:
        # the hop path's kernels, each timed alone on the same stream
        s_ptr = ctypes.c_void_p(stream.cuda_stream)
        vin, vout = ctypes.c_void_p(vec.data_ptr()), ctypes.c_void_p(out.data_ptr())
        parts = {"hop_terms": terms.n_hop_terms, "residual_terms": terms.n_residual_terms,
                 "strings": [targets.strings_stats(parity) for parity in (0, 1)]}
        parts["hop_kernel_ms"] = [timed(lambda parity=parity: lib.qmb_apply_hops(
            ctypes.byref(terms.singles_desc), ctypes.byref(targets.strings(parity)), ctypes.byref(targets.desc), vin, vout,
            pp.MODE_FORWARD, s_ptr)) for parity in (0, 1)]
        dvals = targets.diagonal_values(terms)
        parts["diagonal_values_ms"] = timed(lambda: lib.qmb_diagonal_values(ctypes.byref(terms.desc), ctypes.byref(targets.keys_desc),
                                                                            ctypes.c_void_p(dvals.data_ptr()), s_ptr))
        parts["diagonal_apply_ms"] = timed(lambda: lib.qmb_apply_diagonal_values(n, ctypes.c_void_p(dvals.data_ptr()), vin, vout,
                                                                                 pp.MODE_FORWARD, s_ptr))

        def rebuild_strings() -> None:
            targets._strings.clear()  # pylint: disable=protected-access
            targets.strings(0)
            targets.strings(1)

        parts["strings_build_ms"] = timed(rebuild_strings)
        log(f"hop path parts: {parts}")
    if use_pairs:
        # the structured path's kernels, each timed alone on the same stream
        s_ptr = ctypes.c_void_p(stream.cuda_stream)
        pmode = pp.HERMITIAN_MODE if symmetric else pp.MODE_FORWARD
        vin, vout = ctypes.c_void_p(vec.data_ptr()), ctypes.c_void_p(out.data_ptr())
        cores2, cores1 = targets.cores(2), targets.cores(1)
        parts = {"pair_excitation_terms": terms.n_pair_terms, "single_excitation_terms": terms.n_single_terms,
                 "residual_terms": terms.n_residual_terms}
        parts["pair_kernel_ms"] = timed(lambda: lib.qmb_apply_pairs(ctypes.byref(terms.pairs_desc), ctypes.byref(cores2),
                                                                    ctypes.byref(targets.desc), vin, vout, pmode, s_ptr))
        if terms.singles_desc is not None:
            parts["single_kernel_ms"] = timed(lambda: lib.qmb_apply_singles(ctypes.byref(terms.singles_desc), ctypes.byref(cores1),
                                                                            ctypes.byref(targets.desc), vin, vout, pmode, s_ptr))
        parts["diagonal_kernel_ms"] = timed(lambda: lib.qmb_apply_diagonal(ctypes.byref(terms.pairs_desc), ctypes.byref(targets.desc),
                                                                           vin, vout, pmode, s_ptr))
        if terms.n_residual_terms > 0:
            launch = pp._launch(args.tile, args.groups, 0, symmetric)  # pylint: disable=protected-access
            parts["enumeration_kernel_ms"] = timed(lambda: lib.qmb_apply_prepared(
                ctypes.byref(terms.residual_desc), ctypes.byref(targets.desc), ctypes.c_void_p(configs.data_ptr()), n, vin, vout,
                pp.MODE_FORWARD, ctypes.byref(launch) if launch else None, s_ptr))

        def rebuild(order: int) -> None:
            targets._cores.pop(order, None)  # pylint: disable=protected-access
            targets.cores(order)

        parts["core_index_build_ms"] = {"order2": timed(lambda: rebuild(2)), "order1": timed(lambda: rebuild(1))}
        log(f"structured path parts: {parts}")

        # The C-ABI drop-in (qmb_apply_within with one config buffer and the structured workspace):
        # one call builds the term tables and core indexes in the workspace, reads the plan
        # counters back and applies; timed whole, tables and indexes included.
        ws_size = ctypes.c_size_t(0)
        _lib.check(lib.qmb_apply_within_structured_workspace_size(nq, n, n_terms, occ, ctypes.byref(ws_size)), "structured size")
        ws = torch.empty(ws_size.value, dtype=torch.uint8, device=dev)
        cfg_p = ctypes.c_void_p(configs.data_ptr())
        term_p = [ctypes.c_void_p(t.data_ptr()) for t in (site, kind, coef)]

        def c_abi_call() -> None:
            _lib.check(lib.qmb_apply_within(nq, pc, cfg_p, vin, n, cfg_p, n, *term_p, n_terms, vout, ctypes.c_void_p(ws.data_ptr()),
                                            ws_size.value, s_ptr), "qmb_apply_within")

        parts["c_abi_structured_call_ms"] = timed(c_abi_call)
        parts["c_abi_workspace_bytes"] = ws_size.value
        del ws

    # achievable HBM bandwidth: a plain 1 GiB device copy
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    cps = []
    for _ in range(5):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        dst.copy_(src)
        b.record(stream)
        b.synchronize()
        cps.append(a.elapsed_time(b))
    copy_gbps = 2 * src.numel() * 8 / (float(np.median(cps)) * 1e-3) / 1e9
    del src, dst
    # The lookup model (one key read per generated connection): the bound of the
    # enumeration path, reported for the structured path only as a labelled comparison.
    lookup_bytes = n * (nq + 16) + n * (nq + 16) + n_terms * 28 + m_gen * nq + m_hit * 16
    log(f"M_gen={m_gen} M_hit={m_hit}; forward {kernel_ms:.3f} ms, adjoint {adjoint_ms:.3f} ms, copy {copy_gbps:.0f} GB/s")
    if use_pairs:
        model = structured_model(pp, terms, targets, nq, n)
        kname = "pair_kernel"
        k_bytes, k_ms = model["pair_kernel"]["bytes"], parts["pair_kernel_ms"]
        label = f"pair_kernel<{(nq + 7) // 8}, {'true' if pmode == pp.MODE_ADJOINT else 'false'}, 1>"
        app_bytes = sum(v["bytes"] for k, v in model.items() if k != "core_index_build")
        step_bytes = model["core_index_build"]["bytes"] + 2 * app_bytes
        model_desc = ("structured path, bytes moved at least once per launch: group starts, one config per group of >= 2 "
                      "entries, the entries of those groups with their rows' amplitudes, the coefficient table, the output")
    elif use_hops:
        model = hop_model(terms, targets, nq, n)
        kname = "hop_kernel"
        k_bytes, k_ms = model["hop_kernel_0"]["bytes"], parts["hop_kernel_ms"][0]
        label = f"hop_kernel<{(nq + 7) // 8}, false>"
        app_bytes = sum(v["bytes"] for k, v in model.items() if k != "strings_build")
        step_bytes = model["strings_build"]["bytes"] + 2 * app_bytes
        model_desc = ("hop path, bytes moved at least once per launch: the entries of groups of >= 2 rows with their rows' "
                      "config words, amplitudes and outputs, the single-excitation table")
    else:
        model = None
        kname = "apply_kernel"
        k_bytes, k_ms = lookup_bytes, kernel_ms
        label = f"apply_kernel<{(nq + 7) // 8}, {1 if symmetric else 0}, {1 if direct else 0}>"
        app_bytes = lookup_bytes
        step_bytes = 2 * lookup_bytes
        model_desc = "enumeration path: configs and amplitudes in and out, the term table, one key read per connection, one 16-byte accumulate per contribution"
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    step_gbps = step_bytes / (step_ms * 1e-3) / 1e9
    if achieved > HBM_PEAK_GBPS or step_gbps > HBM_PEAK_GBPS:
        print(f"bench.py: refusing to report: the bytes model gives {achieved:.0f} GB/s for {label} and {step_gbps:.0f} GB/s per "
              f"step, above the {HBM_PEAK_GBPS:.0f} GB/s peak; the model does not bound this path", file=sys.stderr)
        return 3
    traffic, l2_hit, traffic_src = read_traffic(args.workload, kname)
    log(f"roofline: {label} {k_bytes:.4g} B in {k_ms:.3f} ms = {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBPS:.3f} of peak); "
        f"step {step_bytes:.4g} B = {step_gbps:.0f} GB/s; traffic {traffic} ({traffic_src})")

    if args.copy_calibration > 0:
        src = torch.empty(args.copy_calibration // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {src.numel() * 8} bytes")
        del src, dst

    if args.gather_calibration > 0:
        # random 64-byte rows of a table far above the 256 MB last-level cache: each row is one
        # 64-byte HBM read, so FETCH_SIZE over ROWS x (64 + 8 index bytes) calibrates random requests
        table = torch.empty(((4 << 30) // 64, 8), dtype=torch.float64, device=dev).fill_(1.0)
        idx = torch.randint(0, table.size(0), (args.gather_calibration,), device=dev, generator=torch.Generator(device=dev).manual_seed(9))
        for _ in range(3):
            table.index_select(0, idx)
        torch.cuda.synchronize()
        log(f"gather calibration: 3 x {args.gather_calibration} random 64-byte rows of a {table.numel() * 8} byte table")
        del table, idx

    plain = None
    if not args.no_plain:
        # The reference scheme (full term x config grid, per-call sort, binary search, atomics);
        # its backward is the same kernel on swapped arguments with conjugated coefficients, which
        # equals the adjoint because the synthetic tables are Hermitian.
        ns = ops.load(nq, pc)
        psi_r = torch.view_as_real(psi).contiguous()
        coef_conj = coef.clone()
        coef_conj[:, 1].neg_()
        small = configs[:64]
        ns.apply_within_plain(small, psi_r[:64], small, site, kind, coef)
        torch.cuda.synchronize()
        p0 = time.perf_counter()
        ns.apply_within_plain(configs, psi_r, configs, site, kind, coef)
        ns.apply_within_plain(configs, gvec, configs, site, kind, coef_conj)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - p0
        plain = {"ms_per_step": plain_s * 1e3, "value": 2 * m_hit / plain_s, "speedup": plain_s * 1e3 / step_ms}
        log(f"plain comparator: {plain_s * 1e3:.1f} ms per fwd+bwd")

    cpu = None
    if not args.no_cpu:
        rows = sample_rows(n, int(w["sample"]))
        _, h_s = pp.count(terms, targets, configs[torch.from_numpy(rows).to(dev)])
        cpu = cpu_baseline(w, rows, h_s, m_hit / n)
        cpu["m_hit_sample"] = h_s
        log(f"cpu baseline: {cpu['value']:.3e} elements/s over {cpu['seconds']:.2f}s ({cpu['kind']})")

    value = 2 * m_hit / (step_ms * 1e-3)
    result = {
        "metric": "apply_within fwd+bwd matrix elements/s",
        "value": value,
        "unit": "matrix elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded)",
        "config": dict(w["desc"], n_terms=n_terms, n_qubytes=nq, M_gen=m_gen, M_hit=m_hit, valid_terms=valid_terms,
                       diagonal_terms=diag_terms, tile=args.tile or "auto", groups=args.groups or 8,
                       kernel_path=("pair-excitation core groups + " if use_pairs else "hop strings groups + " if use_hops else "")
                       + ("symmetric " if symmetric else "scatter/gather ") + ("direct" if direct else "bucketed tiles")),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic, "traffic_source": traffic_src, "l2_hit_rate": l2_hit,
                     "kernel": label, "kernel_ms": k_ms, "algorithmic_bytes": k_bytes, "model": model_desc,
                     "model_parts": model, "parts": parts, "copy_gbps": copy_gbps,
                     "application": {"forward_ms": kernel_ms, "adjoint_ms": adjoint_ms, "bytes": app_bytes,
                                     "gbps": app_bytes / (kernel_ms * 1e-3) / 1e9},
                     "step": {"bytes": step_bytes, "gbps": step_gbps},
                     "lookup_model": {"what": "one key read per generated connection (the enumeration path's bound; the "
                                              "structured path performs no such lookups, so this is not its roofline)",
                                      "bytes": lookup_bytes, "gbps_equivalent": lookup_bytes / (kernel_ms * 1e-3) / 1e9}},
        "cpu_baseline": cpu,
        "plain_kernel": plain,
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
