"""Benchmark of the MI355X LULESH Lagrange leapfrog step.

One step = one cycle of the reference's main loop (TimeIncrement +
LagrangeLeapFrog, lulesh.cc:2745-2748) over the Sedov problem.  Default
workload: s=192 (7,077,888 elements), 21 regions, balance 2, cost 2, FP64 --
the load-imbalanced s=192 FP64 configuration the headline FOM is quoted on.
Prints one JSON line.

  value      FOM = numElem * steps / elapsed_us (zones/us), state resident in
             HBM, TimeIncrement evaluated on the device (lgpu_run); the
             reference-style loop with a per-cycle dt read-back to the host is
             reported beside it as host_loop_zones_per_us (--host-loop makes it
             the headline)
             The timed region holds the cycles only: no events, no profiling.
  roofline   dominant kernel: algorithmic bytes per launch / mean launch time,
             from HIP events on the context stream around that kernel alone,
             in cycles run after the timed region
  kernels_ms_per_step  every kernel timed with HIP events, per cycle, over
             steady-state cycles after the timed region
  cpu_baseline  the reference's own OpenMP build (oracle/_ref/lulesh2.0) on
             every host core this process may use (affinity mask and cgroup
             quota), on a bounded sample of the same workload: median of 3
             runs as built, and median of 3 with huge pages for malloc
  fp64fma    the FP64 kernels with fused multiply-adds on the same workload,
             with the origin-energy error against the reference's number
  mixed      mixed-precision form vs FP64 at s=192, 11 regions, 200 cycles,
             with the origin-energy error against the reference's number
  ensemble   64 Sedov problems (std::mt19937_64(1234) draw), 100 cycles,
             batched vs run one after another, FP64 and mixed
  setup      problem set-up outside the timed loop: host Domain + device
             context against the device-side set-up (lgpu_create_sedov)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector
# Algorithmic FP64 operations per element of the fused nodal pass (one
# element's stress integration 376, hourglass control 1379, node update 66;
# counted from the reference formulas, lulesh.cc:381-474, 591-706, 1139-1219).
NODAL_FLOPS_PER_ELEM = 1821


def kernel_bytes(name: str, s: int, precision: str, brick: bool = True) -> float:
    """Algorithmic (compulsory) HBM bytes of one launch of each kernel: every
    array the kernel must read or write counted once, node gathers assumed
    perfectly cached (r = nodes per element).  On the brick the element-to-
    node map, face neighbours and BC masks are computed, not read."""
    E = s ** 3
    r = ((s + 1) / s) ** 3
    # FP32 in the mixed form: corner forces, Q gradients, p q ss
    n4 = 4 if precision == "mixed" else 8
    nl = 0 if brick else 32                 # element-to-node map
    faces = 0 if brick else 28              # elemBC + 6 face neighbours
    per_elem = {
        # nodelist 32, x..zd 48r, p q ss, elemMass v volo 24 -> 48 corner values
        "k_force": 32 + 48 * r + 3 * n4 + 24 + 48 * n4,
        # start 4r, corner list 32, 48 corner values, mass 8r, bc 1r, x..zd r/w 96r
        "k_node": 4 * r + 32 + 48 * n4 + 8 * r + 1 * r + 96 * r,
        # fused brick nodal pass: x..zd read 48r, p q ss, elemMass v volo 24,
        # nodalMass 8r, node BC mask 1r, x..zd written 48r
        "k_nodal_brick": 48 * r + 3 * n4 + 24 + 8 * r + 1 * r + 48 * r,
        # nodelist, x..zd 48r, volo 8 -> vnew arealg vdov 24, 6 gradients
        # (delv = vnew - v is formed by k_eos, not stored)
        "k_kin": nl + 48 * r + 8 + 24 + 6 * n4,
        "k_kin_brick": 48 * r + 8 + 24 + 6 * n4,
        # perm 4, faces, 6 gradients, p q, vdov elemMass arealg volo vnew v e
        # 56 -> p q ss, e v 16
        "k_eos": 4 + faces + 6 * n4 + 2 * n4 + 56 + 3 * n4 + 16,
    }
    return per_elem.get(name, 0.0) * E


def host_cores() -> dict:
    """The host cores this process may use: the CPUs of its affinity mask,
    limited by its cgroup's CPU quota (cpu.max) where one is set."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    usable = affinity if quota is None else max(1, min(affinity, int(quota)))
    return {"usable": usable, "affinity_cpus": affinity, "cgroup_cpu_quota": quota,
            "logical_cpus": os.cpu_count()}


def cpu_baseline(s: int, r: int, b: int, c: int, seconds: float) -> dict | None:
    """Times the reference's OpenMP build on a bounded sample (a few cycles of
    the same problem) with every core this process may use (one OpenMP thread
    per usable core, at least 2: with one thread the reference takes its
    serial scatter path)."""
    exe = ROOT / "oracle" / "_ref" / "lulesh2.0"
    if not exe.exists():
        return None
    hc = host_cores()
    cores = max(2, hc["usable"])
    # ~0.25 us per zone-cycle on 8 cores; size the sample to the budget.
    its = max(2, int(seconds / (s ** 3 * 0.25e-6 * 8 / cores)))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")

    def fom(extra_env):
        try:
            out = subprocess.run([str(exe), "-s", str(s), "-r", str(r), "-b", str(b), "-c", str(c),
                                  "-i", str(its)], env=dict(env, **extra_env),
                                 capture_output=True, text=True, timeout=600, check=True).stdout
        except (subprocess.SubprocessError, OSError):
            return None
        m = re.search(r"Grind time \(us/z/c\)\s*=\s*([0-9.eE+-]+)", out)
        return round(1.0 / float(m.group(1)), 4) if m else None

    runs = [fom({}) for _ in range(3)]
    if any(v is None for v in runs):
        return None
    # The reference mallocs and frees ~9 corner arrays of 8E doubles per
    # cycle; with huge pages for malloc the page faults mostly go away.
    huge = [fom({"GLIBC_TUNABLES": "glibc.malloc.hugetlb=1"}) for _ in range(3)]
    huge = sorted(huge)[1] if all(v is not None for v in huge) else None
    return {"value": sorted(runs)[1], "unit": "zones/us", "cores": cores, "kind": "reference",
            "sample": f"oracle/_ref/lulesh2.0 -s {s} -r {r} -b {b} -c {c} -i {its} "
                      f"(OpenMP {cores} threads, close/cores binding, cycle loop only; "
                      f"median of 3 runs)",
            "runs": runs, "cpu_model": cpu_model(), "host_cores": hc,
            "hugetlb_value": huge}


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or "unknown"."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measured_traffic(kernel: str, precision: str, s: int):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs recorded in profiles/traffic.json (s=192 only), or None.
    Counters cannot be read inside this run (rocprofv3 would have to wrap
    it), so the figure is the recorded one."""
    try:
        d = json.loads((ROOT / "profiles" / "traffic.json").read_text())
        k = d[precision][kernel]
    except (OSError, KeyError, ValueError):
        return None
    if s != 192:
        return None
    return int(1024 * (d["fetch_scale"] * k["fetch_kb"] + d["write_scale"] * k["write_kb"]))


def golden_e0(s: int, r: int, b: int, c: int, its: int) -> float | None:
    """Final origin energy of the reference build for this configuration, if
    recorded in tests/golden/ref_summaries.json."""
    p = ROOT / "tests" / "golden" / "ref_summaries.json"
    try:
        d = json.loads(p.read_text())
        return float(d[f"s{s}_r{r}_b{b}_c{c}_i{its}_m1"]["e0"])
    except (OSError, KeyError, ValueError):
        return None


def timed_run(la, torch, s, r, b, c, prec, warmup, steps, device_loop):
    """Builds the problem, runs `warmup` untimed cycles and times `steps`
    cycles with nothing else on the stream (no events, no host work between
    the cycles but the loop's own).  Returns (zones_per_us, ms_per_step,
    domain)."""
    dom = la.HostDomain(s, r, b, c)
    dom.attach_device(prec)
    if dom.run(warmup, device_loop) != 0:
        raise RuntimeError("warm-up failed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = dom.run(warmup + steps, device_loop)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if st != 0:
        raise RuntimeError(f"timed run failed with status {st}")
    done = int(dom.get("cycle")) - warmup
    return s ** 3 * done / ((t1 - t0) * 1e6), (t1 - t0) * 1e3 / done, dom


def profile_run(la, torch, dom, device_loop, cycles):
    """After a timed run: `cycles` more cycles with every kernel timed by HIP
    events (the per-kernel breakdown of steady-state cycles, ms per cycle),
    then as many again with events around the dominant kernel only (its mean
    launch duration, the roofline's figure)."""
    ctx = dom.device()
    lib = la.gpu_lib()
    names = [lib.lgpu_kernel_name(k).decode() for k in range(lib.lgpu_num_kernels())]

    def stats():
        out = {}
        for k, name in enumerate(names):
            n, ms = ctypes.c_int64(), ctypes.c_double()
            la.check(lib.lgpu_kernel_stats(ctx, k, ctypes.byref(n), ctypes.byref(ms)))
            if n.value:
                out[name] = (n.value, ms.value)
        return out

    la.check(lib.lgpu_set_option(ctx, 2, 1))           # kernel timing on
    la.check(lib.lgpu_set_option(ctx, 6, 0xffffffff))  # every kernel
    la.check(lib.lgpu_reset_kernel_stats(ctx))
    c0 = int(dom.get("cycle"))
    torch.cuda.synchronize()
    p0 = time.perf_counter()
    if dom.run(c0 + cycles, device_loop) != 0:
        raise RuntimeError("profile run failed")
    torch.cuda.synchronize()
    p1 = time.perf_counter()
    n_prof = int(dom.get("cycle")) - c0
    per_cycle = {k: v[1] / n_prof for k, v in stats().items()}
    dom_k = max(per_cycle, key=per_cycle.get)
    la.check(lib.lgpu_set_option(ctx, 6, 1 << names.index(dom_k)))
    la.check(lib.lgpu_reset_kernel_stats(ctx))
    c1 = int(dom.get("cycle"))
    if dom.run(c1 + cycles, device_loop) != 0:
        raise RuntimeError("profile run failed")
    n, ms = stats()[dom_k]
    la.check(lib.lgpu_set_option(ctx, 2, 0))
    return {"kernels_ms_per_step": {k: round(v, 4) for k, v in per_cycle.items()},
            "kernels_sum_ms": round(sum(per_cycle.values()), 4),
            "profile_ms_per_step": round((p1 - p0) * 1e3 / n_prof, 4),
            "profile_cycles": n_prof, "dominant": dom_k, "dominant_launches": n,
            "dominant_avg_ms": ms / n}


def mixed_leg(la, torch, s: int, steps: int, profile_cycles: int) -> dict:
    """Mixed vs FP64 on the s, 11-region, b=1, c=1 problem, both run from cycle
    0 for the same number of cycles; energy error against the reference's
    FP64 number where it is recorded, L_inf/scale per field against the FP64
    device run.  Per-kernel breakdowns come from cycles after that."""
    warm = 5
    total = warm + steps
    fm, msm, dm = timed_run(la, torch, s, 11, 1, 1, la.MIXED, warm, steps, True)
    fd, msd, dd = timed_run(la, torch, s, 11, 1, 1, la.FP64, warm, steps, True)
    dm.sync()
    dd.sync()
    e_m, e_d = float(dm.real("e")[0]), float(dd.real("e")[0])
    ref = golden_e0(s, 11, 1, 1, total)
    out = {"workload": f"sedov s={s} r=11 b=1 c=1, cycles {total}",
           "mixed_zones_per_us": round(fm, 2), "fp64_zones_per_us": round(fd, 2),
           "speedup": round(fm / fd, 3), "cycles": total,
           "e0_mixed": e_m, "e0_fp64": e_d,
           "e0_rel_err_vs_fp64_gpu": abs(e_m - e_d) / abs(e_d)}
    if ref is not None:
        out["e0_reference"] = ref
        out["e0_rel_err_vs_reference"] = abs(e_m - ref) / abs(ref)
        out["fp64_e0_rel_err_vs_reference"] = abs(e_d - ref) / abs(ref)
    for f in ("e", "p", "q", "v", "xd"):
        a, b_ = dm.real(f), dd.real(f)
        out[f"linf_{f}"] = float(abs(a - b_).max() / max(abs(b_).max(), 1e-300))
    if profile_cycles:
        out["mixed_kernels_ms_per_step"] = profile_run(la, torch, dm, True,
                                                       profile_cycles)["kernels_ms_per_step"]
        out["fp64_kernels_ms_per_step"] = profile_run(la, torch, dd, True,
                                                      profile_cycles)["kernels_ms_per_step"]
    dm.close()
    dd.close()
    return out


def fp64fma_leg(la, torch, s: int, r: int, b: int, c: int, fp64_fom: float,
                profile_cycles: int) -> dict:
    """The FP64_FMA form (FP64 kernels with fused multiply-adds) on the
    headline workload: 5 + 195 cycles from cycle 0, so that the final origin
    energy compares with the reference's 200-cycle number."""
    warm, steps = 5, 195
    f, ms, d = timed_run(la, torch, s, r, b, c, la.FP64_FMA, warm, steps, True)
    d.sync()
    e0 = float(d.real("e")[0])
    out = {"workload": f"sedov s={s} r={r} b={b} c={c}, cycles {warm + steps}",
           "zones_per_us": round(f, 2), "ms_per_step": round(ms, 4),
           "speedup_vs_fp64": round(f / fp64_fom, 3), "e0": e0,
           "symmetry": list(d.symmetry())}
    ref = golden_e0(s, r, b, c, warm + steps)
    if ref is not None:
        out["e0_reference"] = ref
        out["e0_rel_err_vs_reference"] = abs(e0 - ref) / abs(ref)
    if profile_cycles:
        out["kernels_ms_per_step"] = profile_run(la, torch, d, True,
                                                 profile_cycles)["kernels_ms_per_step"]
    d.close()
    return out


def ensemble_leg(la, torch, members: int, cycles: int, seed: int = 1234) -> dict:
    """The batched ensemble (one cycle graph per cycle for all members) against
    running the same members one after another (each on its own context,
    device-side time stepping), in both precisions."""
    sz, em, b, c = la.ensemble_draw(members, seed)
    zones = int((sz.astype(np.int64) ** 3).sum()) * cycles
    out = {"members": members, "seed": seed, "cycles": cycles,
           "total_elements": int((sz.astype(np.int64) ** 3).sum())}
    for name, prec in (("fp64", la.FP64), ("mixed", la.MIXED)):
        ens = la.Ensemble(sz, em, b, c, precision=prec)
        tune = ens.autotune(trial_cycles=3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, status = ens.run(cycles)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if st != 0:
            raise RuntimeError(f"ensemble failed: {status}")
        out[f"{name}_batched_zones_per_us"] = round(zones / ((t1 - t0) * 1e6), 2)
        out[f"{name}_autotune"] = tune
        ens.close()
        singles = []
        for i in range(members):
            d = la.HostDomain(int(sz[i]), 11, int(b[i]), int(c[i]), float(em[i]))
            d.attach_device(prec)
            singles.append(d)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for d in singles:
            if d.run(cycles, True) != 0:
                raise RuntimeError("sequential member failed")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        out[f"{name}_sequential_zones_per_us"] = round(zones / ((t1 - t0) * 1e6), 2)
        out[f"{name}_speedup_vs_sequential"] = round(out[f"{name}_batched_zones_per_us"] /
                                                     out[f"{name}_sequential_zones_per_us"], 3)
        for d in singles:
            d.close()
    return out


def setup_leg(la, torch, s: int, r: int, b: int, c: int) -> dict:
    """Problem set-up (Domain::Domain, lulesh-init.cc:16-194) outside the timed
    loop: the host Domain plus its device context (lgpu_create + uploads)
    against the device-side set-up (lgpu_create_sedov), same problem."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    h = la.HostDomain(s, r, b, c)
    t1 = time.perf_counter()
    h.attach_device(la.FP64)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    h.close()
    t3 = time.perf_counter()
    d = la.SedovDevice(s, r, b, c)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    d.close()
    return {"workload": f"sedov s={s} r={r} b={b} c={c}", "host_domain_s": round(t1 - t0, 3),
            "host_context_s": round(t2 - t1, 3), "device_setup_s": round(t4 - t3, 3),
            "speedup": round((t2 - t0) / (t4 - t3), 1)}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=192)
    ap.add_argument("--regions", type=int, default=21)
    ap.add_argument("--balance", type=int, default=2)
    ap.add_argument("--cost", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "fp64fma", "mixed"], default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--host-loop", action="store_true",
                    help="TimeIncrement on the host with a per-cycle dt read-back "
                         "(default: on the device)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-fma", action="store_true", help="skip the FP64_FMA leg")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-setup", action="store_true", help="skip the set-up timing leg")
    ap.add_argument("--ensemble-members", type=int, default=64)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one device copy of 512 MiB (float64), the known-size "
                         "kernel that scales the FETCH_SIZE/WRITE_SIZE counters")
    a = ap.parse_args()
    a.device_loop = not a.host_loop
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU halo exchange is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch

    import lulesh_amd as la

    if not torch.cuda.is_available():
        sys.exit("bench.py: no GPU visible")
    if a.calibrate_copy:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    prec = {"fp64": la.FP64, "fp64fma": la.FP64_FMA, "mixed": la.MIXED}[a.precision]
    print(f"[bench] s={a.size} r={a.regions} b={a.balance} c={a.cost} {a.precision}",
          file=sys.stderr, flush=True)
    fom, ms_step, dom = timed_run(la, torch, a.size, a.regions, a.balance, a.cost, prec,
                                  a.warmup, a.steps, a.device_loop)
    steps = int(dom.get("cycle")) - a.warmup
    E = a.size ** 3
    prof_cycles = 0 if a.no_kernel_timing else min(max(a.steps, 10), 50)

    roof, prof = None, {}
    if prof_cycles:
        print("[bench] per-kernel profile (HIP events, after the timed region)",
              file=sys.stderr, flush=True)
        prof = profile_run(la, torch, dom, a.device_loop, prof_cycles)
        dom_k = prof["dominant"]
        avg_s = prof["dominant_avg_ms"] / 1e3
        byts = kernel_bytes(dom_k, a.size, a.precision)
        ach = byts / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": dom_k, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(dom_k, a.precision, a.size),
                "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE, WRITE_SIZE "
                                  "of this kernel at this size, copy-calibrated)",
                "avg_us": round(avg_s * 1e6, 2), "alg_bytes": int(byts),
                "launches_timed": prof["dominant_launches"]}
        if dom_k == "k_nodal_brick":
            tflops = NODAL_FLOPS_PER_ELEM * E / avg_s / 1e12
            roof["fp64_tflops"] = round(tflops, 2)
            roof["fp64_peak_tflops"] = FP64_PEAK_TFLOPS
            roof["fp64_frac"] = round(tflops / FP64_PEAK_TFLOPS, 4)
    dom.close()
    del dom

    extra = {}
    print("[bench] host-loop variant", file=sys.stderr, flush=True)
    fom_h, ms_h, dom_h = timed_run(la, torch, a.size, a.regions, a.balance, a.cost, prec,
                                   a.warmup, min(a.steps, 100), False)
    dom_h.close()
    del dom_h
    extra["host_loop_zones_per_us"] = round(fom_h, 2)
    if not a.no_fma and a.precision == "fp64":
        print("[bench] FP64_FMA leg", file=sys.stderr, flush=True)
        extra["fp64fma"] = fp64fma_leg(la, torch, a.size, a.regions, a.balance, a.cost, fom,
                                       prof_cycles)
    if not a.no_mixed:
        print("[bench] mixed-precision leg", file=sys.stderr, flush=True)
        extra["mixed"] = mixed_leg(la, torch, a.size, 195, prof_cycles)
    if not a.no_ensemble:
        print("[bench] ensemble leg", file=sys.stderr, flush=True)
        extra["ensemble"] = ensemble_leg(la, torch, a.ensemble_members, 100)
    if not a.no_setup:
        print("[bench] set-up leg", file=sys.stderr, flush=True)
        extra["setup"] = setup_leg(la, torch, a.size, a.regions, a.balance, a.cost)

    cpu = None
    if not a.no_cpu_baseline:
        print("[bench] timing the reference OpenMP build on a bounded sample", file=sys.stderr,
              flush=True)
        cpu = cpu_baseline(a.size, a.regions, a.balance, a.cost, a.cpu_seconds)
        if cpu:
            cpu["gpu_over_cpu"] = round(fom / cpu["value"], 1)

    line = {
        # the headline metric; value is its first part (FP64 FOM), the mixed
        # form, the HBM fraction (roofline) and the ensemble are reported in
        # their own fields of this line
        "metric": "FOM zones/\u00b5s @s=192 FP64 vs mixed; HBM GB/s % of peak; ensemble zones/\u00b5s",
        "value": round(fom, 2),
        "unit": "zones/us",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if a.precision == "mixed" else "f64",
        "data": "synthetic (Sedov blast initial conditions, deterministic)",
        "config": {"workload": f"sedov s={a.size} ({E} elements) r={a.regions} b={a.balance} "
                               f"c={a.cost} {a.precision}",
                   "size": a.size, "regions": a.regions, "balance": a.balance, "cost": a.cost,
                   "precision": a.precision, "device_loop": a.device_loop,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        # steady-state cycles after the timed region, every kernel timed
        "kernels_ms_per_step": prof.get("kernels_ms_per_step"),
        "kernels_profile": {k: prof[k] for k in ("kernels_sum_ms", "profile_ms_per_step",
                                                 "profile_cycles") if k in prof},
        **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
