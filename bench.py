"""Benchmark of the MI355X LULESH Lagrange leapfrog step.

One step = one cycle of the reference's main loop (TimeIncrement +
LagrangeLeapFrog, lulesh.cc:2745-2748) over the Sedov problem.  Default
workload: s=192 (7,077,888 elements), 21 regions, balance 2, cost 2, FP64 --
the load-imbalanced s=192 FP64 configuration the headline FOM is quoted on.
Prints one JSON line.

  value      FOM = numElem * steps / elapsed_us (zones/us), state resident in
             HBM, per-cycle dt read back to the host as the reference's loop
             needs (--device-loop keeps TimeIncrement on the device)
  roofline   dominant kernel: algorithmic bytes per launch / mean launch time
             (HIP events on the context stream, measured in the timed region)
  cpu_baseline  the reference's own OpenMP build (oracle/_ref/lulesh2.0) on
             this host's cores, on a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def kernel_bytes(name: str, s: int, precision: str) -> float:
    """Algorithmic (compulsory) HBM bytes of one launch of each kernel: every
    array the kernel must read or write counted once, node gathers assumed
    perfectly cached (r = nodes per element)."""
    E = s ** 3
    r = ((s + 1) / s) ** 3
    cb = 4 if precision == "mixed" else 8   # corner force storage
    gb = 4 if precision == "mixed" else 8   # gradient storage
    per_elem = {
        # nodelist 32, x..zd 48r, p q v volo ss elemMass 48 -> 48 corner values
        "k_force": 32 + 48 * r + 48 + 48 * cb,
        # start 4r, corner list 32, 48 corner values, mass 8r, bc 1r, x..zd r/w 96r
        "k_node": 4 * r + 32 + 48 * cb + 8 * r + 1 * r + 96 * r,
        # fused brick nodal pass: x..zd read 48r, p q v volo ss elemMass 48,
        # nodalMass 8r, node BC mask 1r, x..zd written 48r
        "k_nodal_brick": 48 * r + 48 + 8 * r + 1 * r + 48 * r,
        # nodelist 32, x..zd 48r, volo v 16 -> vnew delv arealg vdov 32, grads 6
        "k_kin": 32 + 48 * r + 16 + 32 + 6 * gb,
        # perm 4, elemBC 4, 6 nbr 24, grads 6, 10 fields 80 -> 5 fields 40
        "k_eos": 4 + 4 + 24 + 6 * gb + 80 + 40,
    }
    return per_elem.get(name, 0.0) * E


def cpu_baseline(s: int, r: int, b: int, c: int, seconds: float) -> dict | None:
    """Times the reference's OpenMP build on a bounded sample (a few cycles of
    the same problem) with all usable cores."""
    exe = ROOT / "oracle" / "_ref" / "lulesh2.0"
    if not exe.exists():
        return None
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    # ~0.25 us per zone-cycle on 8 cores; size the sample to the budget.
    its = max(2, int(seconds / (s ** 3 * 0.25e-6 * 8 / cores)))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    try:
        out = subprocess.run([str(exe), "-s", str(s), "-r", str(r), "-b", str(b), "-c", str(c),
                              "-i", str(its)], env=env, capture_output=True, text=True,
                             timeout=600, check=True).stdout
    except (subprocess.SubprocessError, OSError):
        return None
    m = re.search(r"Grind time \(us/z/c\)\s*=\s*([0-9.eE+-]+)", out)
    if not m:
        return None
    grind = float(m.group(1))
    return {"value": round(1.0 / grind, 4), "unit": "zones/us", "cores": cores,
            "kind": "reference",
            "sample": f"oracle/_ref/lulesh2.0 -s {s} -r {r} -b {b} -c {c} -i {its} "
                      f"(OpenMP {cores} threads, cycle loop only)"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=192)
    ap.add_argument("--regions", type=int, default=21)
    ap.add_argument("--balance", type=int, default=2)
    ap.add_argument("--cost", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--device-loop", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU halo exchange is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch

    import lulesh_amd as la

    if not torch.cuda.is_available():
        sys.exit("bench.py: no GPU visible")
    prec = la.MIXED if a.precision == "mixed" else la.FP64
    print(f"[bench] building s={a.size} r={a.regions} b={a.balance} c={a.cost} {a.precision}",
          file=sys.stderr, flush=True)
    dom = la.HostDomain(a.size, a.regions, a.balance, a.cost)
    dom.attach_device(prec)
    ctx = dom.device()
    lib = la.gpu_lib()

    st = dom.run(a.warmup, a.device_loop)
    if st != 0:
        sys.exit(f"warm-up failed with status {st}")
    if not a.no_kernel_timing:
        la.check(lib.lgpu_set_option(ctx, 2, 1))
        la.check(lib.lgpu_reset_kernel_stats(ctx))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = dom.run(a.warmup + a.steps, a.device_loop)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if st != 0:
        sys.exit(f"timed run failed with status {st}")
    steps = int(dom.get("cycle")) - a.warmup
    elapsed = t1 - t0
    E = a.size ** 3
    fom = E * steps / (elapsed * 1e6)

    roof = None
    kstats = {}
    if not a.no_kernel_timing:
        import ctypes
        for k in range(lib.lgpu_num_kernels()):
            n, ms = ctypes.c_int64(), ctypes.c_double()
            la.check(lib.lgpu_kernel_stats(ctx, k, ctypes.byref(n), ctypes.byref(ms)))
            if n.value:
                kstats[lib.lgpu_kernel_name(k).decode()] = (n.value, ms.value)
        dom_k = max(kstats, key=lambda k: kstats[k][1])
        n, ms = kstats[dom_k]
        avg_s = ms / n / 1e3
        byts = kernel_bytes(dom_k, a.size, a.precision)
        ach = byts / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": dom_k, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_us": round(avg_s * 1e6, 2), "alg_bytes": int(byts)}

    cpu = None
    if not a.no_cpu_baseline:
        print("[bench] timing the reference OpenMP build on a bounded sample", file=sys.stderr,
              flush=True)
        cpu = cpu_baseline(a.size, a.regions, a.balance, a.cost, a.cpu_seconds)

    line = {
        "metric": "FOM zones/us @s=192 FP64 (LULESH Sedov, LagrangeLeapFrog cycle)",
        "value": round(fom, 2),
        "unit": "zones/us",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed * 1e3 / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f64+f32",
        "data": "synthetic (Sedov blast initial conditions, deterministic)",
        "config": {"workload": f"sedov s={a.size} ({E} elements) r={a.regions} b={a.balance} "
                               f"c={a.cost} {a.precision}",
                   "size": a.size, "regions": a.regions, "balance": a.balance, "cost": a.cost,
                   "precision": a.precision, "device_loop": a.device_loop,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels_ms_per_step": {k: round(v[1] / v[0], 4) for k, v in kstats.items()},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
