"""
Runs the compiled reference CPU op (oracle/_ref/qmb_ref_<nq>_<pc>.so) in its own process.

Test / baseline infrastructure.  The reference op registers torch.ops.qmb_hamiltonian_<nq>_<pc>,
the same namespace the product registers, so it must never share a process with qmb_hip; callers
run this file as a subprocess:

    python oracle/ref_runner.py apply_within  <in.npz> <out.npz> [--repeat R]
    python oracle/ref_runner.py find_relative <in.npz> <out.npz> [--repeat R]

in.npz holds configs_i, psi_i (complex128), configs_j or configs_exclude, site, kind, coef,
particle_cut and (find_relative) count_selected.  out.npz receives the result and, with --repeat,
the wall time of every run in seconds (std::chrono-like timing via time.perf_counter around the op).
"""

from __future__ import annotations

import argparse
import os
import pathlib
import sys
import time

import numpy as np


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("op", choices=["apply_within", "find_relative"])
    parser.add_argument("inp")
    parser.add_argument("out")
    parser.add_argument("--repeat", type=int, default=1)
    parser.add_argument("--threads", type=int, default=1)
    args = parser.parse_args()

    import torch  # pylint: disable=import-outside-toplevel

    torch.set_num_threads(args.threads)
    data = np.load(args.inp, allow_pickle=False)
    pc = int(data["particle_cut"])
    configs_i = torch.from_numpy(np.ascontiguousarray(data["configs_i"]))
    nq = configs_i.shape[1]
    so = pathlib.Path(__file__).resolve().parent / "_ref" / f"qmb_ref_{nq}_{pc}.so"
    if not so.exists():
        print(f"reference op not built: {so}", file=sys.stderr)
        return 2
    torch.ops.load_library(str(so))
    ns = getattr(torch.ops, f"qmb_hamiltonian_{nq}_{pc}")
    psi = torch.from_numpy(np.ascontiguousarray(data["psi_i"]))
    psi_r = torch.view_as_real(psi).contiguous()
    site = torch.from_numpy(np.ascontiguousarray(data["site"]))
    kind = torch.from_numpy(np.ascontiguousarray(data["kind"]))
    coef = torch.from_numpy(np.ascontiguousarray(data["coef"]))
    times = []
    result = None
    for _ in range(max(1, args.repeat)):
        t0 = time.perf_counter()
        if args.op == "apply_within":
            configs_j = torch.from_numpy(np.ascontiguousarray(data["configs_j"]))
            result = ns.apply_within(configs_i, psi_r, configs_j, site, kind, coef)
        else:
            configs_x = torch.from_numpy(np.ascontiguousarray(data["configs_exclude"]))
            result = ns.find_relative(configs_i, psi_r, int(data["count_selected"]), site, kind, coef, configs_x)
        times.append(time.perf_counter() - t0)
    assert result is not None
    res = result.numpy()
    if args.op == "apply_within":
        res = res[:, 0] + 1j * res[:, 1]
    np.savez(args.out, result=res, times=np.array(times), threads=args.threads, pid=os.getpid())
    return 0


if __name__ == "__main__":
    sys.exit(main())
