"""Benchmark: agent-steps/s of the per-step mechanics path on one MI355X.

Default workload (the headline line): 3-D two-species cell sorting, Langevin,
N = 1,000,000 agents (volume fraction 0.3, species-Morse R 1 / cutoff 2.5 /
strength 0.5 x [[1, .2], [.2, 1]], gamma 2, kT 0.01, dt 0.005, AB3).  One step =
forces -> integrator -> boundary -> counting-sort re-binning of all agents,
inputs resident in HBM.

Prints one JSON line.  `value` is the mixed-precision path (float pair forces
and sums, double state) where the workload has one, else FP64; the
FP64 path, the force error of the mixed path against FP64, the force kernel's
roofline figures and the CPU baseline (the oracle's chili restatement, threaded)
are reported beside it.  `--workload` selects the other configurations:
  c1  2-D Morse NewtonDamped, 1000 agents, Euler
  c3  2-D bacterial rods, 200,000 rods x 8 vertices, AB3
  c4  512-instance 2-D Morse ensemble in one context, plus the same instances
      run one after another (one context each)
  c5  3-D Mie lattice, 4,000,000 agents, AB3
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
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32
# Useful FP32 work of the force kernel (FMA = 2 flops; sqrt, reciprocal, exp = 1):
# per candidate the displacement and squared distance (3 sub, 1 mul, 2 FMA), per
# pair inside the cutoff about 30 more (root and reciprocal, the potential's
# exponential or powers, direction and weight, the sum of 3 components).
FLOPS_PER_CANDIDATE = 8
FLOPS_PER_PAIR = 30
# committed rocprofv3 summaries of the workloads that have one
PROFILES = {"c2": "profiles/c2_cell_sorting_1M.md", "c3": "profiles/c3_rods_2d_200k.md",
            "c4": "profiles/c4_ensemble_512.md", "c5": "profiles/c5_mie_3d_4M.md"}


def candidates_per_agent(wl, start):
    """Mean number of candidate agents over the clipped 3^D stencil (own voxel included),
    averaged over all agents of all instances."""
    tot, num = 0.0, 0.0
    off = 0
    for spec in wl.instances:
        n = spec.domain.get_n_voxels()
        V = int(np.prod(n))
        cnt = np.diff(start[off:off + V + 1].astype(np.int64)).reshape(n)
        off += V
        pad = np.pad(cnt, 1)
        stencil = np.zeros_like(cnt)
        D = len(n)
        shifts = [(a, b, c) for a in range(3) for b in range(3) for c in range(3)] if D == 3 else \
            [(a, b) for a in range(3) for b in range(3)]
        for sh in shifts:
            sl = tuple(slice(s, s + n[k]) for k, s in enumerate(sh))
            stencil += pad[sl]
        tot += float((stencil * cnt).sum())
        num += float(cnt.sum())
    return tot / num


def profile_figure(table, kernel, workload):
    """A per-kernel figure from the workload's committed rocprofv3 summary:
    `hbm_bytes_per_dispatch` (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of
    known size) or `valu_issue_utilisation`; None when absent."""
    if workload not in PROFILES:
        return None
    try:
        text = open(os.path.join(ROOT, PROFILES[workload])).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)[table].get(kernel)
    except (OSError, IndexError, KeyError, ValueError):
        return None


def force_bytes_per_agent(cbar, record_bytes, dim, stencil, force_bytes):
    """Algorithmic bytes of the force kernel per agent: own record + (start, end) per
    stencil voxel + candidate records + force write."""
    return record_bytes + 8 * stencil + cbar * record_bytes + force_bytes


def time_rods_colony(steps=100):
    """The cycle pipeline at scale: 200,000 2-D rods of 8 vertices (the rods workload)
    with per-rod spring lengths U[0.5, 0.6], growth 0.5 per unit time, division at
    0.62 with radius 0.5 and sensing within 1.0: rod-steps/s over `steps` steps while
    the colony divides, beside the plain step of the same rods (FP64 both)."""
    from cellular_raza_amd import MechanicsBackend, RodGrowthCycle, workloads

    wl = workloads.c3_rods_2d(n=200_000, seed=3)
    n = wl.n_agents
    rng = np.random.default_rng(7)
    spring = rng.uniform(0.5, 0.6, n)
    cyc = RodGrowthCycle(growth_base_rate=0.5, division_threshold=0.62, division_radius=0.5, sensing_radius=1.0,
                         max_agents=3 * n)
    out = {}
    for mode in ("plain", "cycle"):
        g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=2)
        if mode == "cycle":
            g.set_rod_cycle(cyc)
        g.set_agents(wl.pos.reshape(n, -1), wl.vel.reshape(n, -1))
        g.set_rod_spring_lengths(spring)
        g.step(5)
        n0 = g.n_agents
        t0 = time.perf_counter()
        g.step(steps)
        dt = time.perf_counter() - t0
        out[mode] = {"rods_start": n0, "rods_end": g.n_agents, "s": dt,
                     "rod_steps_per_s": steps * 0.5 * (n0 + g.n_agents) / dt}
        g.close()
    return {"steps": steps, **out}


def time_removal(wl, precision, frac=0.1, repeats=3):
    """CycleEvent::Remove at scale: `frac` of the agents of the workload removed
    between steps (crm_remove_agents: flags to the device, compaction, set-up sort),
    `repeats` times in a row; wall clock per removal and removed agents per second."""
    from cellular_raza_amd import MechanicsBackend

    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, noise=wl.noise, seed=1,
                         precision=precision)
    g.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    g.step(4)
    rng = np.random.default_rng(11)
    secs, removed = [], []
    for _ in range(repeats):
        rm = rng.random(g.n_agents) < frac
        t0 = time.perf_counter()
        g.remove_agents(rm)
        secs.append(time.perf_counter() - t0)
        removed.append(int(rm.sum()))
        g.step(2)
    g.close()
    return {"agents_before": wl.n_agents, "removed": removed, "s": secs,
            "removed_per_s": sum(removed) / sum(secs), "ms_per_removal": 1e3 * sum(secs) / len(secs)}


def time_rods_example():
    """The bacterial rods example end to end (5 rods growing and dividing for 25 h of
    simulated time, dt 0.1 min): GPU wall clock for all 15000 steps, and for the
    first 5000 beside the oracle's single-thread run of the same 5000 steps."""
    from cellular_raza_amd import MechanicsBackend, workloads
    from oracle.oracle import OracleSim

    ex = workloads.bacterial_rods_example()
    wl = ex.workload

    def gpu():
        g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=2)
        g.set_rod_cycle(ex.cycle)
        g.set_agents(wl.pos.reshape(len(wl.pos), -1), wl.vel.reshape(len(wl.pos), -1))
        g.set_rod_spring_lengths(ex.spring_length)
        return g

    g = gpu()
    t0 = time.perf_counter()
    g.step(5000)
    t_5000 = time.perf_counter() - t0
    n_5000 = g.n_agents
    g.step(10000)
    t_all = time.perf_counter() - t0
    n_all = g.n_agents
    g.close()
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=2)
    o.set_agents(wl.pos.reshape(len(wl.pos), -1), wl.vel.reshape(len(wl.pos), -1))
    o.set_rod_cycle(ex.cycle)
    o.set_spring_lengths(ex.spring_length)
    t_cpu = o.step_timed(5000)
    return {"steps": 15000, "gpu_s": t_all, "rods_at_end": n_all, "gpu_s_first_5000": t_5000,
            "rods_at_5000": n_5000, "oracle_s_first_5000": t_cpu, "oracle_rods_at_5000": o.n_agents,
            "oracle_threads": 1}


def time_snapshot(wl, precision):
    """One save point after the history ring has filled: crm_get_snapshot (device gather
    + copy of the agent-order records) and crm_save_cells_json (gather + threaded JSON
    formatting + write), wall clock; the JSON goes to a temporary folder."""
    import shutil
    import tempfile

    from cellular_raza_amd import storage as st

    g = make_backend(wl, precision, graph=True, timing=False, seed=1)
    g.step(4)
    g.snapshot()  # first call allocates the record buffer
    t0 = time.perf_counter()
    snap = g.snapshot()
    t_gather = time.perf_counter() - t0
    rec_bytes = sum(v.nbytes for v in snap.values())
    n_sp = int(wl.species.max()) + 1 if wl.species is not None else 1
    templates = st.cell_templates(wl.instances[0].mechanics, wl.instances[0].interaction, n_species=n_sp)
    tmp = tempfile.mkdtemp(prefix="crm_snapshot_")
    try:
        t0 = time.perf_counter()
        g.save_cells_json(tmp, 4, templates)
        t_json = time.perf_counter() - t0
        size = sum(os.path.getsize(os.path.join(r, f)) for r, _, fs in os.walk(tmp) for f in fs)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
        g.close()
    return {"gather_and_copy_s": t_gather, "record_bytes": rec_bytes, "json_write_s": t_json, "json_bytes": size,
            "json_agents_per_s": wl.n_agents / t_json}


def make_backend(wl, precision, graph, timing, seed):
    """Context with the tiling chosen by the library's autotuner (timed on these agents)."""
    from cellular_raza_amd import MechanicsBackend

    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, noise=wl.noise,
                         seed=seed, precision=precision, graph=graph, timing=timing, autotune=True)
    g.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    return g


def run_gpu(wl, precision, steps, warmup, seed=1, timing_steps=20):
    t_set0 = time.perf_counter()
    g = make_backend(wl, precision, True, False, seed)
    t_set = time.perf_counter() - t_set0  # context + upload + initial placement (host buffers)
    f0 = g.compute_forces()
    g.step(warmup)
    t0 = time.perf_counter()
    g.step(steps)  # synchronous: returns after the stream drained
    t1 = time.perf_counter()
    sec = t1 - t0
    _, start = g.get_voxels()
    tiling = None
    if getattr(wl.instances[0].mechanics, "n_vertices", 1) == 1:
        tiling = dict(g.tiling(0), policy=g.tiling_policy)
    t_get0 = time.perf_counter()
    pos_end = g.get_state()[0]
    t_get = time.perf_counter() - t_get0  # download of positions and velocities to host buffers
    in_range = None
    cut = pair_cutoff(wl.instances[0])
    if tiling is not None and cut:  # pairs inside the cutoff per agent, on the end state
        g.set_neighbor_sensing(cut)
        g.compute_forces()
        in_range = float(g.neighbor_counts().mean())
        g.set_neighbor_sensing(None)
    g.close()
    # per-kernel HIP-event timing over a second timed region on the same stream
    gt = make_backend(wl, precision, False, True, seed)
    gt.step(warmup)
    gt.reset_timing()
    gt.step(timing_steps)
    tim = gt.timing()
    gt.close()
    return sec, f0, start, tim, pos_end, t_set + t_get, tiling, in_range


def pair_cutoff(spec):
    """The interaction's cutoff radius (None when it has none)."""
    inter = spec.interaction
    c = getattr(inter, "cutoff", None)
    return float(c) if c else None


def run_sequential(wl, precision, steps, warmup, seed=1):
    """Every instance of an ensemble in a context of its own, run one after another."""
    from cellular_raza_amd import MechanicsBackend

    total = 0.0
    off = 0
    for spec, n in zip(wl.instances, wl.counts):
        g = MechanicsBackend([spec], dt=wl.dt, mechanics_solver_order=wl.solver_order, noise=wl.noise, seed=seed,
                             precision=precision, graph=True)
        g.set_agents(wl.pos[off:off + n], wl.vel[off:off + n])
        off += n
        g.step(warmup)
        t0 = time.perf_counter()
        g.step(steps)
        total += time.perf_counter() - t0
        g.close()
    return total


def cpu_baseline(wl, threads, steps):
    from oracle.oracle import OracleSim

    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, noise=wl.noise, seed=1,
                  n_subdomains=threads)
    o.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    o.step_timed(1)  # warm-up step
    sec = o.step_timed(steps)
    o.close()
    return wl.n_agents * steps / sec, sec


def oracle_forces(wl):
    from oracle.oracle import OracleSim

    # one subdomain: pairs across subdomain borders are weighted differently
    # (full instead of half), so only the single-subdomain oracle matches
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, noise=wl.noise, seed=1)
    o.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    f = o.compute_forces()
    o.close()
    return f


def rel_err(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return float(np.linalg.norm(a - b, axis=1).max() / np.linalg.norm(b, axis=1).max())


def rms_rel_err(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, axis=1))) / np.sqrt(np.mean(np.sum(b ** 2, axis=1))))


def forces_at(wl, pos, precision):
    """Forces of the workload's instances at positions `pos` (one context)."""
    from cellular_raza_amd import MechanicsBackend

    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, precision=precision)
    g.set_agents(pos, None, wl.species, wl.counts)
    f = g.compute_forces()
    g.close()
    return f


def trajectory_error(wl, steps, seed=1):
    """Mixed against FP64 from the same initial state after `steps` steps (same Philox
    stream): max and RMS of |x_mixed - x_fp64| per agent (absolute, length units), and the
    force error of the mixed path on the state the mixed trajectory reaches."""
    from cellular_raza_amd import MechanicsBackend

    end = {}
    for precision in ("fp64", "mixed"):
        g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, noise=wl.noise,
                             seed=seed, precision=precision, graph=True)
        g.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
        g.step(steps)
        end[precision] = g.get_state()[0]
        g.close()
    n = wl.n_agents
    d = np.linalg.norm((end["mixed"] - end["fp64"]).reshape(n, -1), axis=1)
    pm = end["mixed"]
    f_m, f_d = forces_at(wl, pm, "mixed"), forces_at(wl, pm, "fp64")
    return {"steps": steps, "position_max_abs": float(d.max()), "position_rms_abs": float(np.sqrt((d ** 2).mean())),
            "force_rel_err_on_mixed_state": rel_err(f_m, f_d), "force_rms_rel_err_on_mixed_state": rms_rel_err(f_m, f_d)}


def host_cpu():
    """Model name and core counts of this host (lscpu)."""
    import subprocess

    info = {"logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "CPU family", "Model", "Thread(s) per core", "Core(s) per socket", "Socket(s)",
                     "CPU max MHz", "L3 cache"):
                info[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_calibration(threads):
    """The reference's own benchmark workload (cell_sorting.rs sim-size, k = 6 and 8:
    81,920 and 1,310,720 cells) on the oracle, 1 thread and `threads` threads, beside
    the published 1-thread numbers (per executed step, 2.0 GHz desktop cores), and the
    GPU (FP64 and mixed) on the same 1,310,720-cell workload."""
    from cellular_raza_amd import workloads as W

    published = {"81920": {"3700X-at-2000MHz": 2.11e4, "3960X-at-2000MHz": 2.30e4, "12700H-at-2000MHz": 2.06e4},
                 "1310720": {"3700X-at-2000MHz": 2.02e4, "3960X-at-2000MHz": 2.22e4, "12700H-at-2000MHz": 2.04e4}}
    out = {"host": host_cpu(), "published_1_thread_per_executed_step": published, "oracle": {}}
    for k, steps1 in ((6, 10), (8, 2)):
        wl = W.cell_sorting_benchmark(k)
        row = {}
        for t, st in ((1, steps1), (threads, 5 * steps1)):
            v, sec = cpu_baseline(wl, t, st)
            row[f"{t}_threads"] = {"agent_steps_per_s": v, "steps": st, "s": sec}
            print(f"bench: calibration N={wl.n_agents} {t} threads {v:.3e} agent-steps/s", file=sys.stderr,
                  flush=True)
        out["oracle"][str(wl.n_agents)] = row
    wl = W.cell_sorting_benchmark(8)
    gpu = {}
    for precision in ("fp64", "mixed"):
        sec = run_gpu(wl, precision, 200, 10, timing_steps=5)[0]
        gpu[precision] = wl.n_agents * 200 / sec
    out["gpu_same_workload_agent_steps_per_s"] = gpu
    out["gpu_mixed_over_published_3700X"] = gpu["mixed"] / published["1310720"]["3700X-at-2000MHz"]
    return out


# name -> (factory(args) -> Workload, mixed available, description, force kernel name for the profile lookup)
def workload_table():
    from cellular_raza_amd import workloads as W

    return {
        "c2": (lambda a: W.c2_sorting_3d(n=a.n or 1_000_000, seed=2, steps=a.steps), True,
               "3-D two-species cell sorting, Langevin, N=1,000,000, volume fraction 0.3",
               "k_force_tiled<3, 3, true, false, false>"),
        "c1": (lambda a: W.c1_morse_2d(n=a.n or 1000, seed=1, steps=a.steps), False,
               "2-D Morse NewtonDamped, N=1000 in 250x250, Euler", "k_force_tiled<2, 0, false, false, true>"),
        "c3": (lambda a: W.c3_rods_2d(n=a.n or 200_000, seed=3, steps=a.steps), True,
               "2-D bacterial rods, 200,000 rods x 8 vertices in 2000x2000, AB3", "k_force_rod<2, 0, true>"),
        "c4": (lambda a: W.c4_ensemble_2d(n_instances=a.n or 512, seed=4, steps=a.steps), True,
               "512-instance 2-D Morse ensemble (500..20000 agents each), Euler, one context",
               "k_force_tiled<2, 0, true, false, false>"),
        "c5": (lambda a: W.c5_mie_3d(n=a.n or 4_000_000, seed=5, steps=a.steps), True,
               "3-D Mie lattice, N=4,000,000, AB3", "k_force_tiled<3, 1, true, false, false>"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--n", type=int, default=0, help="agents (c4: instances); 0 = the workload's own size")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--seq-steps", type=int, default=100, help="c4: steps of the one-after-another run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rods-example", action="store_true",
                    help="also time the bacterial rods example (Cycle + NeighborSensing, 25 h = 15000 steps) on the "
                         "GPU, and its first 5000 steps on the GPU and on the oracle")
    ap.add_argument("--rods-colony", action="store_true",
                    help="also time the rods cycle at scale: the 200,000-rod 2-D workload growing and dividing "
                         "(the rods example's cycle model), against the same rods without the cycle")
    ap.add_argument("--snapshot", action="store_true",
                    help="also time one save point: the device gather into agent order with its copy to the host, "
                         "and the cells written in the reference's JSON layout (to a temporary folder)")
    ap.add_argument("--removal", action="store_true",
                    help="also time agent removal (CycleEvent::Remove): 10 %% of the workload's agents removed "
                         "between steps, three times")
    ap.add_argument("--cpu-calibration", action="store_true",
                    help="also time the oracle (1 thread and --cpu-threads) and the GPU on the reference's own "
                         "cell-sorting benchmark workload (81,920 and 1,310,720 cells) beside its published numbers")
    ap.add_argument("--traj-steps", type=int, default=0,
                    help="steps of the mixed-vs-FP64 trajectory comparison (0: 500 for c5, 100 otherwise)")
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

    factory, has_mixed, desc, force_kernel = workload_table()[args.workload]
    clamped = None
    if args.workload == "c1" and args.warmup + args.steps > 35:
        # 1000 discs of radius 6 in 250 x 250 overlap 1.8-fold: the overlapping pairs
        # throw an agent out of the domain (BoundaryError) within 40 to 100 steps --
        # the system is chaotic, so the step depends on the summation order of the
        # forces (FP64 GPU step 40, oracle step 100; every GPU step agrees with the
        # oracle restarted from the GPU state) -- so the run stays in the first 35
        # steps, in FP64
        clamped = f"warmup {args.warmup} + steps {args.steps} clamped to 5 + 30"
        args.warmup, args.steps = 5, 30
    wl = factory(args)
    N = wl.n_agents
    D = wl.dim
    nv = getattr(wl.instances[0].mechanics, "n_vertices", 1)
    print(f"bench: workload {wl.name}, N={N}, steps={args.steps}, warmup={args.warmup}", file=sys.stderr,
          flush=True)

    main_prec = "mixed" if has_mixed else "fp64"
    sec_m, f_main, start, tim_m, pos_main, host_io, tiling, in_range = run_gpu(wl, main_prec, args.steps, args.warmup)
    value = N * args.steps / sec_m
    print(f"bench: {main_prec} {value:.4e} agent-steps/s", file=sys.stderr, flush=True)
    extra = {
        # the same run with the host hand-over (context set-up from host arrays and the
        # final download) inside the clock
        "host_io_s": host_io,
        "value_with_host_io": N * args.steps / (sec_m + host_io),
    }
    if has_mixed:
        fp64_steps = max(args.steps // 4, 10)
        sec_d, f_fp64, _, tim_d = run_gpu(wl, "fp64", fp64_steps, args.warmup)[:4]
        value_fp64 = N * fp64_steps / sec_d
        print(f"bench: fp64 {value_fp64:.4e} agent-steps/s", file=sys.stderr, flush=True)
        extra["fp64_value"] = value_fp64
        extra["mixed_over_fp64"] = value / value_fp64
        extra["force_rel_err_mixed_vs_fp64"] = rel_err(f_main, f_fp64)
        extra["force_rms_rel_err_mixed_vs_fp64"] = rms_rel_err(f_main, f_fp64)
        extra["kernel_ms_fp64"] = {k: v["total_ms"] / max(v["launches"], 1) for k, v in tim_d.items()}
        # the mixed path's force error on the state its own timed run ended in, and the
        # trajectory error against FP64 from the same start
        f_m, f_d = forces_at(wl, pos_main, "mixed"), forces_at(wl, pos_main, "fp64")
        extra["force_rel_err_mixed_vs_fp64_at_end"] = rel_err(f_m, f_d)
        extra["force_rms_rel_err_mixed_vs_fp64_at_end"] = rms_rel_err(f_m, f_d)
        extra["mixed_end_state_step"] = args.warmup + args.steps
        traj_steps = args.traj_steps or (500 if args.workload == "c5" else 100)
        extra["trajectory_mixed_vs_fp64"] = trajectory_error(wl, traj_steps)
        print(f"bench: trajectory {extra['trajectory_mixed_vs_fp64']}", file=sys.stderr, flush=True)
    if args.workload == "c4" and args.seq_steps > 0:
        sec_s = run_sequential(wl, main_prec, args.seq_steps, 5)
        seq_value = N * args.seq_steps / sec_s
        extra["sequential_value"] = seq_value
        extra["batched_over_sequential"] = value / seq_value
        extra["sequential_sample"] = f"{len(wl.instances)} contexts one after another, {args.seq_steps} steps each"
        print(f"bench: sequential {seq_value:.4e} agent-steps/s", file=sys.stderr, flush=True)
    if args.rods_colony:
        extra["rods_colony"] = time_rods_colony()
        print(f"bench: rods colony {extra['rods_colony']}", file=sys.stderr, flush=True)
    if args.rods_example:
        extra["rods_example"] = time_rods_example()
        print(f"bench: rods example {extra['rods_example']}", file=sys.stderr, flush=True)
    if args.removal:
        extra["removal"] = time_removal(wl, main_prec)
        print(f"bench: removal {extra['removal']}", file=sys.stderr, flush=True)
    if args.snapshot:
        extra["snapshot"] = time_snapshot(wl, main_prec)
        print(f"bench: snapshot {extra['snapshot']}", file=sys.stderr, flush=True)
    if args.workload in ("c1", "c3"):
        # force error of the benchmarked path against the oracle on the benchmarked input
        extra["force_rel_err_vs_oracle"] = rel_err(f_main, oracle_forces(wl))

    cbar = candidates_per_agent(wl, start)
    if nv > 1:
        rec_bytes = 8 * nv * D + 4  # vertex positions + species
        bf = force_bytes_per_agent(cbar, rec_bytes, D, 3 ** D, 8 * nv * D)
    else:
        rec_bytes = (4 * D + 4) if has_mixed else 8 * D + 8
        # mixed point forces are stored in float, FP64 ones in double
        bf = force_bytes_per_agent(cbar, rec_bytes, D, 3 ** D, (4 if has_mixed else 8) * D)
    t_force_ms = tim_m["force"]["total_ms"] / max(tim_m["force"]["launches"], 1)
    achieved = bf * N / (t_force_ms * 1e-3) / 1e9
    flops_agent = None
    achieved_tflops = None
    if in_range is not None:
        flops_agent = FLOPS_PER_CANDIDATE * cbar + FLOPS_PER_PAIR * in_range
    elif nv > 1:  # rods: a vertex-segment distance (~20 flops) per vertex pair of a candidate rod
        flops_agent = 20.0 * cbar * nv * (nv - 1)
    if flops_agent is not None:
        achieved_tflops = flops_agent * N / (t_force_ms * 1e-3) / 1e12
    traffic = profile_figure("hbm_bytes_per_dispatch", force_kernel, args.workload)
    step_bytes = None
    if nv == 1:  # the step's byte model: force + integrator + counting sort (points)
        langevin = wl.noise == "philox"
        force_w = 4 * D if has_mixed else 8 * D
        b_i = (8 * D + 8 * D + force_w + 32 * D + (8 * D if langevin else 0)) + (8 * D + 8 * D + 16 * D + 4)
        n_vox = len(start) - 1
        b_s = 4 + 4 + 8 * n_vox / N + 8 + 2 * (48 * D + 8)
        step_bytes = bf + b_i + b_s

    cpu = None
    if not args.no_cpu:
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        cpu_val, cpu_sec = cpu_baseline(wl, threads, args.cpu_steps)
        one_steps = max(2, args.cpu_steps // 10)
        cpu1_val, cpu1_sec = cpu_baseline(wl, 1, one_steps)
        cpu = {
            "value": cpu_val,
            "unit": "agent-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"same {N}-agent workload, {args.cpu_steps} steps after 1 warm-up step, "
                      f"chili restatement with {threads} subdomain threads ({cpu_sec:.1f} s); "
                      f"1 thread: {one_steps} steps ({cpu1_sec:.1f} s)",
            "value_1_thread": cpu1_val,
            "host": host_cpu(),
        }
        print(f"bench: cpu {cpu_val:.4e} agent-steps/s on {threads} threads", file=sys.stderr, flush=True)

    if args.cpu_calibration:
        extra["cpu_calibration"] = cpu_calibration(min(args.cpu_threads, os.cpu_count() or 1))

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
        "dtype": "f32 pair forces / f64 state" if has_mixed else "f64",
        "data": "synthetic (seeded positions" + (", Bernoulli(0.5) species)" if wl.species is not None else ")"),
        "config": {
            "workload": desc,
            "n_agents": N,
            "n_instances": len(wl.instances),
            "vertices_per_agent": nv,
            "precision": main_prec,
            "mechanics_solver_order": wl.solver_order,
            "dt": wl.dt,
            "force_tiling": tiling,
        },
        **{k: v for k, v in extra.items() if k in ("fp64_value", "mixed_over_fp64", "force_rel_err_mixed_vs_fp64",
                                                  "force_rms_rel_err_mixed_vs_fp64")},
        "kernel_ms": {k: v["total_ms"] / max(v["launches"], 1) for k, v in tim_m.items()},
        # The force kernel serves its candidates from LDS and is bound by VALU
        # instruction issue (the profile's SQ_INSTS_VALU): `achieved` is its useful
        # FP32 work over the live HIP-event time; the HBM byte model and the
        # measured HBM traffic stand beside it.
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_PEAK_TFLOPS if achieved_tflops is not None else None,
            "traffic": traffic,
            "valu_issue_frac": profile_figure("valu_issue_utilisation", force_kernel, args.workload),
            "kernel": force_kernel,
            "kernel_us": t_force_ms * 1e3,
            "flops_per_agent": flops_agent,
            "candidates_per_agent": cbar,
            "pairs_in_cutoff_per_agent": in_range,
            "hbm_model": {"bytes_per_agent": bf, "achieved_GBs": achieved, "frac": achieved / HBM_PEAK_GBS,
                          "note": "every candidate record charged to HBM (the kernel reads them from LDS)"},
            "hbm_traffic_frac": (traffic / (t_force_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
            "source": PROFILES[args.workload] + " (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated; "
                      "SQ_INSTS_VALU for the VALU issue fraction)" if args.workload in PROFILES else None,
            "step_model_bytes_per_agent": step_bytes,
            "step_model_frac": (step_bytes * value / 1e9 / HBM_PEAK_GBS) if step_bytes else None,
        },
        "cpu_baseline": cpu,
    }
    for k, v in extra.items():
        if k not in line:
            line[k] = v
    if clamped:
        line["clamped"] = clamped
    if cpu is not None:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
