#!/usr/bin/env python3
"""Split-step throughput of GPEtask.solve_split on one MI355X.

Workload (default): grid3d(12,256,12,256,12,256), V = (X^2+Y^2+16Z^2)/2,
Ntotal = 1e6, g = 0.02, gamma = 0.01, n_recalc = 10, dt = 0.002, seeded
Thomas-Fermi * exp(i theta) initial state, complex128 throughout.

A step is one Strang split step over the whole 256^3 grid; the timed region
runs --steps steps as solve_split outer iterations of 50 inner steps (block
boundaries with renormalisation and chemical potential included), with the
state resident in HBM.  Prints ONE JSON line.

Legs:
  parity       the GPU and the FP64 oracle run the same first 10 steps from the
               same state; rel-L2 of the GPU state against the oracle
  timed        warmup, then --steps steps bracketed by device synchronisation
  profile      one outer iteration with HIP events around every sweep launch
               (on the library's stream) -> per-kernel average durations
  cpu_baseline the oracle's steps/s on this host (the same 10 steps)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "gpe-matlab_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

PEAK_HBM_GBPS = 8000.0
# algorithmic HBM bytes per grid point per launch of each sweep sequence
SEQ_BYTES = {"fwd": 32, "inv": 32, "kin": 32, "open": 40, "mid": 40, "close": 40, "kex": 24, "ke": 16,
             "itpa": 40, "itpb": 48}
STEP_BYTES_3D = 136  # 4 sweeps x 32 B + 8 B of V per step


def committed_traffic(n):
    """HBM bytes per launch per sweep from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", f"sweep{n}_c2_hbm_traffic.md")
    if not os.path.exists(path):
        return {}
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return {k: int(v) for k, v in (kv.split("=") for kv in line.split(":", 1)[1].strip().split(", "))}
    return {}


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256, help="grid points per axis")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain 256 MiB device copy (known-size kernel for HBM counter scaling)")
    args = ap.parse_args()

    if args.copy_calib:
        import torch
        x = torch.empty(16 * 1024 * 1024, dtype=torch.complex128, device="cuda").fill_(1.0)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_read": x.numel() * 16, "copy_calib_bytes_written": x.numel() * 16}))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU slab decomposition is designed but not built", file=sys.stderr)
        sys.exit(2)

    import gpe_hip
    from gpe_hip import workloads
    from gpe_hip._lib import AXIS_NAMES, SEQ_NAMES

    ni = 50 if (args.steps % 50 == 0 and args.warmup % 50 == 0) else 10
    if args.steps % ni or args.warmup % ni:
        sys.exit("--steps and --warmup must be multiples of 10")
    ddt = 0.002
    t_setup = time.perf_counter()
    grid, V, psi0, props = workloads.c2_state(n=args.n)
    log(f"workload built in {time.perf_counter() - t_setup:.1f}s; device {gpe_hip.device_info()}")
    task = gpe_hip.GPEtask(grid, V)
    task.Ntotal, task.g, task.gamma = props["Ntotal"], props["g"], props["gamma"]
    dev = task._device()
    task._set_params(task.Ntotal)
    dev.set_potential(0, V)
    npts = int(np.prod(grid.size))

    # ---------------------------------------------------------------- parity leg
    parity = None
    gpu10 = None
    cs = args.cpu_steps
    if not args.no_parity or not args.no_cpu_baseline:
        dev.set_state(0, psi0)
        dev.split_begin(ddt, cs)
        dev.split_outer(1)
        gpu10 = dev.get_state(0)
        log(f"parity leg: {cs} GPU steps done")

    cpu_baseline = None
    if not args.no_cpu_baseline or not args.no_parity:
        import oracle
        from oracle import grids as ogrids
        O = oracle.grid3d(grid.xmax, grid.nx, grid.ymax, grid.ny, grid.zmax, grid.nz)
        to = oracle.GPEtask(O, V)
        to.Ntotal, to.g, to.gamma = props["Ntotal"], props["g"], props["gamma"]
        to.init_state = psi0
        t0 = time.perf_counter()
        po = to.solve_split(ddt, cs, 1)
        el = time.perf_counter() - t0
        cpu_baseline = {"value": cs / el, "unit": "split-steps/s", "cores": ogrids.WORKERS, "kind": "port",
                        "sample": f"FP64 numpy/scipy restatement of solve_split.m, {cs} steps (one n_rec block) "
                                  f"of the same 256^3 workload in {el:.1f}s; scipy.fft workers={ogrids.WORKERS}, "
                                  "elementwise passes single-threaded"}
        parity = float(np.linalg.norm(gpu10 - po) / np.linalg.norm(po))
        log(f"cpu leg: {cs} oracle steps in {el:.1f}s; GPU rel-L2 vs oracle {parity:.3e}")

    # ---------------------------------------------------------------- timed leg
    dev.set_state(0, psi0)
    dev.split_begin(ddt, ni)
    j = 0
    for _ in range(args.warmup // ni):
        j += 1
        dev.split_outer(j)
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps // ni):
        j += 1
        _, muc, ncur = dev.split_outer(j)
    sync()
    el = time.perf_counter() - t0
    steps_per_s = args.steps / el
    log(f"timed: {args.steps} steps in {el:.3f}s -> {steps_per_s:.1f} steps/s (mu {muc:.6f}, n {ncur:.6f})")

    # ---------------------------------------------------------------- profile leg
    dev.profile_enable(True)
    j += 1
    dev.split_outer(j)
    prof = dev.profile_read()
    dev.profile_enable(False)
    kernels = []
    for (seq, ax), (ms, cnt) in prof.items():
        name = f"{SEQ_NAMES[seq]}_{AXIS_NAMES[ax]}"
        avg = ms / cnt
        b = SEQ_BYTES[SEQ_NAMES[seq]] * npts
        kernels.append({"kernel": name, "launches": cnt, "avg_ms": avg, "total_ms": ms,
                        "alg_bytes": b, "alg_GBps": b / (avg * 1e-3) / 1e9})
    kernels.sort(key=lambda k: -k["total_ms"])
    for k in kernels:
        log(f"  {k['kernel']:8s} x{k['launches']:3d} avg {k['avg_ms'] * 1e3:8.1f} us  {k['alg_GBps']:7.1f} GB/s alg")
    dom = kernels[0]
    traffic = committed_traffic(args.n).get(dom["kernel"])
    roofline = {"bound": "hbm", "achieved": round(dom["alg_GBps"], 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                "frac": round(dom["alg_GBps"] / PEAK_HBM_GBPS, 4), "traffic": traffic, "kernel": dom["kernel"],
                "alg_bytes_per_launch": dom["alg_bytes"],
                "traffic_source": "profiles/sweep256_c2_hbm_traffic.md (FETCH_SIZE/WRITE_SIZE, copy-calibrated)"
                if traffic else None}
    step_gbps = STEP_BYTES_3D * npts * steps_per_s / 1e9

    out = {
        "metric": "split-steps/s (GPEtask.solve_split, 256^3 complex128)",
        "value": round(steps_per_s, 2),
        "unit": "split-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Thomas-Fermi x exp(i U[0,0.1]) state)",
        "config": {"workload": f"GPEtask 3D drop-in grid3d(12,{args.n},12,{args.n},12,{args.n}) "
                               "V=0.5(X^2+Y^2+16Z^2) Ntotal=1e6 g=0.02 gamma=0.01 n_recalc=10 "
                               f"solve_split dt=0.002, {ni} inner steps per outer",
                   "grid": [args.n] * 3, "parallelism": "single GPU"},
        "roofline": roofline,
        "step_alg_GBps": round(step_gbps, 1),
        "step_roofline_frac": round(step_gbps / PEAK_HBM_GBPS, 4),
        "parity_rel_l2_vs_oracle": parity,
        "cpu_baseline": cpu_baseline,
        "kernels": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in kk.items()} for kk in kernels],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
