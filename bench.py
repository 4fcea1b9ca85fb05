"""Benchmark: agent-steps/s of the per-step mechanics path on one MI355X.

Workload: 3-D two-species cell sorting, Langevin, N = 1,000,000 agents (volume
fraction 0.3, species-Morse R 1 / cutoff 2.5 / strength 0.5 x [[1, .2], [.2, 1]],
gamma 2, kT 0.01, dt 0.005, AB3).  One step = forces -> integrator -> boundary
-> counting-sort re-binning of all agents, inputs resident in HBM.

Prints one JSON line.  `value` is the mixed-precision path (float pair forces
with compensated sums, double state); the FP64 path, the force error of the
mixed path against FP64, the force kernel's roofline figures and the CPU
baseline (the oracle's chili restatement, threaded) are reported beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def candidates_per_agent(wl, start):
    """Mean number of candidate agents over the clipped 3x3x3 stencil (own voxel included)."""
    dom = wl.instances[0].domain
    n = dom.get_n_voxels()
    cnt = np.diff(start.astype(np.int64)).reshape(n)
    pad = np.pad(cnt, 1)
    stencil = np.zeros_like(cnt)
    for a in (0, 1, 2):
        for b in (0, 1, 2):
            for c in (0, 1, 2):
                stencil += pad[a:a + n[0], b:b + n[1], c:c + n[2]]
    return float((stencil * cnt).sum() / cnt.sum())


PROFILE = os.path.join(ROOT, "profiles", "c2_cell_sorting_1M.md")
FORCE_KERNEL = "k_force_tiled<3, 3, true>"


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (counters scaled by a copy kernel of known size), or None."""
    try:
        text = open(PROFILE).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)["hbm_bytes_per_dispatch"].get(kernel)
    except (OSError, IndexError, KeyError, ValueError):
        return None


def force_bytes_per_agent(cbar, record_bytes, dim, stencil=27):
    # own record + (start, end) per stencil voxel + candidate records + force write (double)
    return record_bytes + 8 * stencil + cbar * record_bytes + 8 * dim


def run_gpu(wl, precision, steps, warmup, seed=1, timing_steps=20):
    from cellular_raza_amd import MechanicsBackend

    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, noise="philox",
                         seed=seed, precision=precision, graph=True)
    g.set_agents(wl.pos, wl.vel, wl.species)
    f0 = g.compute_forces()
    g.step(warmup)
    t0 = time.perf_counter()
    g.step(steps)  # synchronous: returns after the stream drained
    t1 = time.perf_counter()
    sec = t1 - t0
    _, start = g.get_voxels()
    g.close()
    # per-kernel HIP-event timing over a second timed region on the same stream
    gt = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, noise="philox",
                          seed=seed, precision=precision, graph=False, timing=True)
    gt.set_agents(wl.pos, wl.vel, wl.species)
    gt.step(warmup)
    gt.reset_timing()
    gt.step(timing_steps)
    tim = gt.timing()
    gt.close()
    return sec, f0, start, tim


def cpu_baseline(wl, threads, steps):
    from oracle.oracle import OracleSim

    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, noise="philox", seed=1,
                  n_subdomains=threads)
    o.set_agents(wl.pos, wl.vel, wl.species)
    o.step_timed(1)  # warm-up step
    sec = o.step_timed(steps)
    f = o.compute_forces() if threads == 1 else None
    o.close()
    return wl.n_agents * steps / sec, sec, f


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (reference for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; --gpus must be 1 (slab sharding across "
                                   "GPUs is designed, not built)"}))
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return

    from cellular_raza_amd import workloads

    wl = workloads.c2_sorting_3d(n=args.n, seed=2, steps=args.steps)
    N = wl.n_agents
    print(f"bench: workload {wl.name}, N={N}, steps={args.steps}, warmup={args.warmup}", file=sys.stderr,
          flush=True)

    sec_m, f_mixed, start, tim_m = run_gpu(wl, "mixed", args.steps, args.warmup)
    print(f"bench: mixed {N * args.steps / sec_m:.4e} agent-steps/s", file=sys.stderr, flush=True)
    fp64_steps = max(args.steps // 4, 10)
    sec_d, f_fp64, _, tim_d = run_gpu(wl, "fp64", fp64_steps, args.warmup)
    print(f"bench: fp64 {N * fp64_steps / sec_d:.4e} agent-steps/s", file=sys.stderr, flush=True)

    def err(a, b):
        return float(np.linalg.norm(a - b, axis=1).max() / np.linalg.norm(b, axis=1).max())

    force_err_mixed_vs_fp64 = err(f_mixed, f_fp64)

    value = N * args.steps / sec_m
    value_fp64 = N * fp64_steps / sec_d
    cbar = candidates_per_agent(wl, start)
    bf = force_bytes_per_agent(cbar, 16, 3)
    t_force_ms = tim_m["force"]["total_ms"] / max(tim_m["force"]["launches"], 1)
    achieved = bf * N / (t_force_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        cpu_val, cpu_sec, f_orc = cpu_baseline(wl, threads, args.cpu_steps)
        cpu = {
            "value": cpu_val,
            "unit": "agent-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"same {N}-agent workload, {args.cpu_steps} steps after 1 warm-up step, "
                      f"chili restatement with {threads} subdomain threads ({cpu_sec:.1f} s)",
        }
        print(f"bench: cpu {cpu_val:.4e} agent-steps/s on {threads} threads", file=sys.stderr, flush=True)

    line = {
        "metric": "agent-steps/s on 1 MI355X, mixed-FP32 vs FP64 at <=1e-5 rel force err",
        "value": value,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_m / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair forces / f64 state",
        "data": "synthetic (seeded uniform positions, Bernoulli(0.5) species)",
        "config": {
            "workload": "3-D two-species cell sorting, Langevin, N=1,000,000, volume fraction 0.3",
            "n_agents": N,
            "precision": "mixed",
            "mechanics_solver_order": wl.solver_order,
            "dt": wl.dt,
        },
        "fp64_value": value_fp64,
        "mixed_over_fp64": value / value_fp64,
        "force_rel_err_mixed_vs_fp64": force_err_mixed_vs_fp64,
        "kernel_ms": {k: v["total_ms"] / max(v["launches"], 1) for k, v in tim_m.items()},
        "kernel_ms_fp64": {k: v["total_ms"] / max(v["launches"], 1) for k, v in tim_d.items()},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(FORCE_KERNEL),
            "traffic_source": "profiles/c2_cell_sorting_1M.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "kernel": FORCE_KERNEL,
            "bytes_per_agent": bf,
            "candidates_per_agent": cbar,
        },
        "cpu_baseline": cpu,
    }
    if cpu is not None:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
