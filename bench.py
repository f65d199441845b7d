"""Benchmark: 256^3 polycrystal C_eff solves/s on one MI355X.

Workload (the cell the metric is quoted on): 256^3 periodic Voronoi
polycrystal, 200 grains, cubic C11/C12/C44 = 168.4/121.4/75.4 GPa, one
uniform random rotation per grain (grain label + per-grain rotated tensor),
six unit engineering-strain loads solved together.  A step is one full C_eff
solve with the int32 labels already resident in HBM, timed like the
reference's workload (set-up included, geometry generation excluded): device
label -> row front end, reference-medium selection, operator and FFT
preconditioner set-up, RHS, block PCG (FP32 operator / preconditioner / CG
with FP64 reliable updates), FP64 true-residual certificate <= 1e-8, stress
recovery and teardown.  The same solve on a resident operator is reported
beside it, and the plain FP64 solve of the same cell (same step definition)
is timed with the C_eff difference between the two.

roofline: the FP32 six-RHS stencil apply (the dominant kernel), timed with HIP
events on the stream it runs on; algorithmic bytes per launch =
N (12 R read + 12 R write + 1 label byte), R = 6.  `traffic` is read from the
committed FETCH_SIZE/WRITE_SIZE summary of the same launch (profiles/).
cpu_baseline: the FP64 NumPy oracle (operator apply + Green preconditioner)
timed at 64^3 on one host thread (child process, BLAS/OpenMP threads = 1,
started before the GPU is initialised) and extrapolated to a 256^3 six-load
solve with the GPU's iteration count (labelled as extrapolated).
cpu_baseline_all_cores: the same sample in one single-thread process per host
core at once, rates summed (the oracle's element loop is serial, so this is
how the host's cores are all used).  cpu_end_to_end: whole oracle solves on one
host thread, not extrapolated: the 32^3 sphere cell by FP64 FFT-PCG and a 16^3
sphere cell by sparse direct (the reference's CPU-backend size limit), run in
a child process while the GPU legs run; gpu_small_cells: the same two cells
solved on the GPU (set-up included), FP64 and mixed.
mfma_stencil: the same six-RHS apply on the matrix cores (vxl_cell_set_mfma),
same buffers and events.  block_cg: the resident 256^3 solve (and the batch,
batch.block_cg) with block CG over the six loads instead of six CG recurrences.  batch.graph_replay: the 64-cell batch again with the
check blocks replayed from HIP graphs.
woven_reference_workload: the reference's own published headline workload, its
analytic woven cell at 192^3 and 256^3, as complete `homogenize` calls (mixed
and FP64) beside the reference's published complete-call times on its GPU, with
C_eff against the reference's certified record of the same cell.
adjoint: forward C_eff and backward gradient of the 96^3 SIMP density cell and
of the tow angles of the 64^3 woven copy (FP64), beside the central-difference
cost the adjoint replaces.  woven_like_128: the 128^3 woven-like cell with
per-voxel tow tangents, mixed against FP64.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "voxel-lmfem_amd"))

HBM_PEAK_GBS = 8000.0


def cpu_sample(n=64, budget_s=12.0):
    """Seconds per RHS-iteration of the FP64 NumPy oracle at n^3 (operator apply + Green apply),
    repeated until about budget_s seconds of CPU work have been timed."""
    sys.path.insert(0, str(ROOT))
    from oracle import fem, spectral
    from voxel_lmfem.cells import voronoi_polycrystal
    from voxel_lmfem import materials as pm
    (phase, mats, dom, _), _ = voronoi_polycrystal(n, grains=200, seed=3)
    f = pm.build_field(phase, mats)
    full = pm.unpack21(f.rows[f.index.astype(np.int64)])
    h = dom / n
    op = fem.MatrixFree(full, phase.shape, h)
    u = np.random.default_rng(0).normal(size=(n ** 3, 3))
    t0 = time.perf_counter()
    op.apply(u)
    spectral.green_apply(u, phase.shape, h, 50.0, 40.0)
    reps = max(3, int(budget_s / max(time.perf_counter() - t0, 1e-3)))
    t0 = time.perf_counter()
    for _ in range(reps):
        op.apply(u)
    t_apply = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        spectral.green_apply(u, phase.shape, h, 50.0, 40.0)
    t_green = (time.perf_counter() - t0) / reps
    return dict(n=n, reps=reps, t_apply=t_apply, t_green=t_green)


def host_info():
    """Host cores visible to this process and the CPU model."""
    model = None
    try:
        for line in open("/proc/cpuinfo", encoding="utf-8"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return dict(cpu_count=os.cpu_count(), affinity=affinity, model=model)


def cpu_end_to_end():
    """Whole oracle solves on one host thread: the FP64 matrix-free FFT-PCG
    C_eff of the 32^3 sphere cell (r = 3, vf >= 0.30, E 3/70 GPa, tol 1e-8)
    and the sparse-direct C_eff of a 16^3 sphere cell (4096 voxels, the
    reference's CPU-backend limit)."""
    sys.path.insert(0, str(ROOT))
    from oracle import homogenize as oracle
    from voxel_lmfem.cells import periodic_spheres
    (phase, mats, dom, _), _ = periodic_spheres(32, seed=1)
    t0 = time.perf_counter()
    r = oracle.homogenize_matfree(phase, mats, dom)
    t_c1 = time.perf_counter() - t0
    its = [int(c["iterations"]) for c in r["calls"]] if r.get("calls") and "iterations" in r["calls"][0] else None
    (phase, mats, dom, _), _ = periodic_spheres(16, radius=1.5, seed=1)
    t0 = time.perf_counter()
    d = oracle.homogenize_direct(phase, mats, dom)
    t_d16 = time.perf_counter() - t0
    return dict(c1_seconds=t_c1, c1_iterations=its, direct16_seconds=t_d16,
                c1_CH=np.asarray(r["CH"]).tolist(), direct16_CH=np.asarray(d["CH"]).tolist())


def _child_env(threads):
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
        env[k] = str(threads)
    return env


def cpu_baseline_child(threads=1):
    """Run cpu_sample in a child process with `threads` BLAS/OpenMP threads
    (NumPy's FFT and element loops run on one thread either way).

    Called before the GPU is initialised; the child never touches the GPU."""
    import subprocess
    out = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-sample"], env=_child_env(threads),
                         capture_output=True, text=True, timeout=600, check=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline_concurrent(copies):
    """`copies` single-thread cpu_sample children at once: the host's aggregate
    oracle throughput with every core busy (NumPy's element loop is serial,
    so more BLAS threads in one process do not use the other cores)."""
    import subprocess
    procs = [subprocess.Popen([sys.executable, str(Path(__file__).resolve()), "--cpu-sample"], env=_child_env(1),
                              stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True) for _ in range(copies)]
    outs = []
    for p in procs:
        text, _ = p.communicate(timeout=900)
        if p.returncode != 0:
            raise RuntimeError("cpu sample child failed")
        outs.append(json.loads(text.strip().splitlines()[-1]))
    return outs


def cpu_baseline_aggregate(samples, iterations_total, n_target):
    """Sum of the per-process rates of concurrent single-thread samples."""
    legs = [cpu_baseline(sm, iterations_total, n_target) for sm in samples]
    worst = max(sm["t_apply"] + sm["t_green"] for sm in samples)
    return dict(value=float(sum(leg["value"] for leg in legs)), unit="solves/s", cores=len(samples), kind="port",
                sample=(f"{len(samples)} concurrent single-thread FP64 NumPy oracle processes, each timing "
                        f"operator apply + FFT Green apply per RHS-iteration at {samples[0]['n']}^3 (slowest "
                        f"{worst:.3f} s), each extrapolated to {n_target}^3 x {iterations_total} RHS-iterations; "
                        f"aggregate throughput, extrapolated"))


def cpu_baseline(sample, iterations_total, n_target, threads=1):
    """Extrapolate the per-iteration oracle sample to a full n_target^3 solve."""
    n = sample["n"]
    scale = (n_target / n) ** 3
    per_iter = (sample["t_apply"] + sample["t_green"]) * scale
    seconds = per_iter * iterations_total
    return dict(value=1.0 / seconds, unit="solves/s", cores=threads, kind="port",
                sample=(f"FP64 NumPy oracle in a child process with OMP/OpenBLAS/MKL threads = {threads}: Gauss-B "
                        f"operator apply {sample['t_apply']:.3f} s + FFT Green apply {sample['t_green']:.3f} s per "
                        f"RHS-iteration at {n}^3 (200-grain polycrystal, {sample['reps']} reps each), x{scale:.0f} to "
                        f"{n_target}^3, x{iterations_total} RHS-iterations of the GPU run; extrapolated"))


def frames_leg(n=256, R=6, reps=20):
    """FP32 six-RHS stencil on an n^3 single-phase cell whose every voxel has
    its own random frame (uniform rotations, seed 1), for a cubic crystal
    (C11 168.4, C12 121.4, C44 75.4 GPa: the cubic closed form), an
    orthotropic tow (E1 230, E2 = E3 15 GPa: the orthotropic closed form, on
    the 32 x 14 and the 32 x 8 tile) and a general anisotropic phase (the
    Bond rotation): per-voxel coefficient rows (the device front end's i32
    index and one FP64 + FP32 row per voxel) against phase rows rotated in
    the stencil by a per-voxel FP32 quaternion (vxl_cell_set_frames).  HIP
    events on the stream the kernels run on."""
    import torch
    from scipy.spatial.transform import Rotation
    from voxel_lmfem import materials as pm
    from voxel_lmfem.gpu import Cell
    cubic = [[168.4, 121.4, 121.4, 0, 0, 0], [121.4, 168.4, 121.4, 0, 0, 0], [121.4, 121.4, 168.4, 0, 0, 0],
             [0, 0, 0, 75.4, 0, 0], [0, 0, 0, 0, 75.4, 0], [0, 0, 0, 0, 0, 75.4]]
    tow = dict(kind="orthotropic", frame="local", E1=230., E2=15., E3=15., nu12=.2, nu13=.2, nu23=.3,
               G12=27., G13=27., G23=15. / 2.6)
    X = np.random.default_rng(3).normal(size=(6, 6))
    general = (X @ X.T + 6 * np.eye(6)) * 10.0
    shape = (n, n, n)
    N = n ** 3
    rot = Rotation.random(N, random_state=1).as_matrix().reshape(shape + (3, 3))
    phase = np.ones(shape, np.int32)
    labels = torch.from_numpy(phase).cuda()
    stream = torch.cuda.current_stream()

    def timed(cell):
        u = torch.randn(R, 3, N, dtype=torch.float32, device="cuda")
        o = torch.empty_like(u)
        for _ in range(3):
            cell.apply_planes(u, o)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            cell.apply_planes(u, o)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = dict(workload=f"single-phase {n}^3 cell, random frame per voxel", n=n, rhs=R,
               bytes_per_voxel_as_scheduled=dict(rows=R * (24 + 84) + 4, frames=R * (24 + 17)),
               bytes_per_voxel_minimal=dict(rows=24 * R + 4 + 84, frames=24 * R + 17))
    for name, spec in (("cubic", dict(kind="anisotropic", frame="local", C=cubic)), ("orthotropic", tow),
                       ("anisotropic", dict(kind="anisotropic", frame="local", C=general.tolist()))):
        mats = {1: spec}
        field = pm.build_field_device(labels, mats, rot)
        cell = Cell(field, shape, (1.0 / n,) * 3)
        try:
            ms_rows = timed(cell)
            cell.set_frames(*pm.phase_frames(phase, mats, rot))
            ms_frames = timed(cell)
            res = dict(rows_ms=ms_rows, frames_ms=ms_frames, frames_vs_rows=ms_frames / ms_rows,
                       frames_tile=[32, cell.stencil_config()[0]])
            if name == "orthotropic":  # the closed form on the 32 x 8 tile too
                ty, sched, zmin = cell.stencil_config()
                cell.set_stencil_config("float32", 8, sched, zmin)
                res["frames_ms_tile_32x8"] = timed(cell)
                cell.set_stencil_config("float32", ty, sched, zmin)
        finally:
            cell.close()
        del field
        out[name] = res
        print(f"# frames {name} {n}^3 R={R}: per-voxel rows {ms_rows:.3f} ms, frames in the stencil "
              f"{ms_frames:.3f} ms", file=sys.stderr, flush=True)
    return out


# The reference's published complete-call times of its analytic woven cell
# (RTX 5090; FP64 operator and PCG vectors, FP32 FFT preconditioner, FP64
# true residual <= 1e-8; material construction, operator preparation, RHS,
# solves, residual checks, stress recovery and cleanup included; median of
# three after a same-grid warm-up): repro/generated/figure_9_optimization_chain.csv:7
# (192^3) and repro/generated/figure_10_grid_cost.csv:11 (256^3).
REFERENCE_WOVEN_SECONDS = {192: (1.8897, "repro/generated/figure_9_optimization_chain.csv:7"),
                           256: (4.3867, "repro/generated/figure_10_grid_cost.csv:11")}


def woven_leg(grids=(192, 256), repeats=3):
    """The reference's own headline workload on this GPU: the analytic woven
    cell (four orthotropic yarns in local frames, matrix), six unit loads,
    through `homogenize` (the drop-in call) in mixed and FP64 precision.
    Complete-call wall time (host validation and material front end
    included, geometry generation excluded), median of `repeats` calls after
    one warm-up call, beside the reference's published time; C_eff against
    the reference's certified record of the same cell (tests/golden)."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import reference_weave
    records = json.loads((Path(__file__).resolve().parent / "tests" / "golden" / "woven_records.json").read_text())
    out = {}
    for g in grids:
        (phase, mats, dom, ori), meta = reference_weave(g)
        rec = records[f"certify_{g}"]
        ref_s, ref_src = REFERENCE_WOVEN_SECONDS[g]
        row = dict(input_hash_matches_reference=meta["input_hash"] == rec["input_hash"],
                   reference_complete_call_s=ref_s, reference_source=ref_src,
                   reference_hardware="RTX 5090 (FP64 operator, FP32 FFT preconditioner)")
        for precision in ("mixed", "fp64"):
            homogenize(phase, mats, dom, orientation=ori, precision=precision)
            walls, inner, res = [], [], None
            for _ in range(repeats):
                t0 = time.perf_counter()
                res = homogenize(phase, mats, dom, orientation=ori, precision=precision)
                walls.append(time.perf_counter() - t0)
                inner.append(res["seconds"])
            wall = float(np.median(walls))
            ch = np.asarray(res["CH"])
            row[precision] = dict(
                complete_call_s=wall, complete_call_s_all=[round(w, 4) for w in walls],
                workload_s=float(np.median(inner)), speedup_vs_reference=ref_s / wall,
                stages_s={st["name"]: round(st["seconds"], 4) for st in res["stages"]},
                iterations=[c["iterations"] for c in res["calls"]], max_true_residual=res["max_true_residual"],
                ceff_rel_diff_vs_reference_record=float(np.linalg.norm(ch - np.asarray(rec["CH"])) /
                                                        np.linalg.norm(np.asarray(rec["CH"]))))
            print(f"# woven {g}^3 {precision}: {wall:.3f} s per complete call (reference {ref_s:.3f} s on its "
                  f"GPU), iterations {row[precision]['iterations']}, C_eff vs the reference's certified record "
                  f"{row[precision]['ceff_rel_diff_vs_reference_record']:.1e}", file=sys.stderr, flush=True)
        out[f"{g}"] = row
    return out


def c2_leg(n=128, repeats=3):
    """The 128^3 woven-like cell (three labels: matrix and warp / weft tows,
    transversely isotropic tows with one path tangent per voxel) as complete
    `homogenize` calls, mixed and FP64, medians of `repeats` after one untimed
    call, with the mixed C_eff's error against the FP64 one."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import woven_like
    (phase, mats, dom, ori), meta = woven_like(n)
    out = dict(workload=f"woven_like_{n}^3_ti_tows_per_voxel_tangents_6loads", tow_fraction=meta.get("tow_fraction"))
    CH = {}
    for precision in ("mixed", "fp64"):
        homogenize(phase, mats, dom, orientation=ori, precision=precision)
        walls, res = [], None
        for _ in range(repeats):
            t0 = time.perf_counter()
            res = homogenize(phase, mats, dom, orientation=ori, precision=precision)
            walls.append(time.perf_counter() - t0)
        CH[precision] = np.asarray(res["CH"])
        out[precision] = dict(complete_call_s=float(np.median(walls)), workload_s=float(res["seconds"]),
                              iterations=[c["iterations"] for c in res["calls"]],
                              max_true_residual=res["max_true_residual"])
    out["ceff_rel_err_mixed_vs_fp64"] = float(np.linalg.norm(CH["mixed"] - CH["fp64"]) / np.linalg.norm(CH["fp64"]))
    print(f"# woven-like {n}^3: mixed {out['mixed']['complete_call_s'] * 1e3:.1f} ms, fp64 "
          f"{out['fp64']['complete_call_s'] * 1e3:.1f} ms per complete call, mixed vs fp64 "
          f"{out['ceff_rel_err_mixed_vs_fp64']:.1e}", file=sys.stderr, flush=True)
    return out


def adjoint_leg(repeats=3):
    """dC_eff/d(per-voxel parameter) through the adjoint (torch.autograd):
    the SIMP density cell at 96^3 (Gaussian random field, sigma 4 voxels,
    E = 1e-3 + rho^3 (70 - 1e-3), nu 0.3) and the fiber angles of the tow
    voxels of the 64^3 woven copy, J = C_1111 + K_V(C_eff), FP64 solves.
    Forward (C_eff) and backward (the whole gradient field from the fields
    the forward solve kept) timed separately, medians of `repeats` after one
    untimed pass; beside them the central-difference cost the adjoint
    replaces, two forward solves per voxel (not run)."""
    import torch
    from voxel_lmfem import adjoint, materials as pm
    from voxel_lmfem.cells import simp_density, woven_like

    def timed(make):
        fw, bw = [], []
        for k in range(repeats + 1):
            x = make()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            C, param = x()
            t1 = time.perf_counter()
            J = C[0, 0] + C[:3, :3].sum() / 9.0
            J.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if k:
                fw.append(t1 - t0)
                bw.append(t2 - t1)
        return float(np.median(fw)), float(np.median(bw)), param.grad

    out = {}
    rho, C1, _, _ = simp_density(96, sigma=4.0, seed=5)

    def density():
        r = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
        return lambda: (adjoint.density_ceff(r, np.ones(3), C1=C1), r)
    fw, bw, g = timed(density)
    n = rho.size
    out["density_96"] = dict(voxels=n, forward_s=fw, backward_s=bw, gradient_entries_per_s=n / bw,
                             central_difference_s=2 * n * fw, adjoint_vs_fd=2 * n * fw / (fw + bw),
                             gradient_finite=bool(torch.isfinite(g).all()))
    (phase, mats, dom, ori), _ = woven_like(64, section=(2, 6), amplitude=3.0)
    labels, tensors, local = pm.phase_table(mats)
    slot = np.searchsorted(labels, phase.reshape(-1))
    mask = local[slot]
    idx = np.flatnonzero(mask)
    frames = pm.frames(ori, phase.shape, idx)
    iso = pm.pack21(tensors[slot])
    C_loc = tensors[slot[idx[0]]]

    def angle():
        t = torch.zeros(phase.shape, dtype=torch.float64, requires_grad=True)
        return lambda: (adjoint.angle_ceff(t, dom, C_loc, frames, iso, mask), t)
    fw, bw, g = timed(angle)
    out["angle_woven_64"] = dict(voxels=int(phase.size), tow_voxels=int(len(idx)), forward_s=fw, backward_s=bw,
                                 central_difference_s=2 * len(idx) * fw, adjoint_vs_fd=2 * len(idx) * fw / (fw + bw),
                                 gradient_finite=bool(torch.isfinite(g).all()))
    for name, leg in out.items():
        print(f"# adjoint {name}: forward {leg['forward_s'] * 1e3:.1f} ms, backward {leg['backward_s'] * 1e3:.1f} ms "
              f"(central differences would take {leg['central_difference_s']:.0f} s)", file=sys.stderr, flush=True)
    return out


def batch_leg(steps=2, warmup=1, count=64, streams=4, workers=8, method="cg", stack=1, stack_voxels=64 ** 3,
              sizes=(32, 48, 64, 80, 96)):
    """Uneven batch: `count` two-phase random-inclusion cells (edges drawn from
    {32, 48, 64, 80, 96}, volume fraction U[0.1, 0.4], radius 2-5 voxels,
    contrast log-uniform in [10, 1000], nu = 0.3), six loads each, mixed
    precision, certified <= 1e-8.  Labels are resident in HBM before timing; a
    step is every cell's device front end, reference medium, operator and
    preconditioner set-up, the concurrent solves (vxl_solver_run_batch,
    largest first, `streams` HIP streams, `workers` host threads) and
    teardown.  stack > 1: cells of the same edge with at most `stack_voxels`
    voxels run up to `stack` at a time as one stacked cell (every launch
    covers all of them)."""
    import torch
    from voxel_lmfem import autotune as tuner
    from voxel_lmfem.cells import random_inclusion_batch
    from voxel_lmfem import materials as pm
    from voxel_lmfem.gpu import Cell, Solver, run_batch
    from voxel_lmfem.reference_medium import select
    t0 = time.perf_counter()
    cells = random_inclusion_batch(count=count, seed=7, sizes=tuple(sizes))
    print(f"# generated {count} inclusion cells in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    resident = [(torch.from_numpy(ph.astype(np.int32)).to("cuda"), ph, mats, dom) for (ph, mats, dom, _), _ in cells]
    pool = [torch.cuda.Stream() for _ in range(streams)]
    entries = tuner.load()  # autotuned stencil configurations, per grid shape
    torch.cuda.synchronize()
    configs = {}

    split = {"setup": 0.0, "solve": 0.0, "teardown": 0.0}

    by_edge = {}
    for k, (_, ph, _, _) in enumerate(resident):
        by_edge.setdefault(ph.shape, []).append(k)
    # A stack runs until its slowest cell certifies: stack cells of similar
    # stiffness contrast (iteration counts follow it).
    contrast = [max(m["E"] for m in mats.values()) / min(m["E"] for m in mats.values())
                for _, _, mats, _ in resident]
    for ks in by_edge.values():
        ks.sort(key=lambda k: contrast[k])
    groups = []
    for shape, ks in by_edge.items():
        width = max(1, stack) if int(np.prod(shape)) <= stack_voxels else 1
        groups += [ks[i:i + width] for i in range(0, len(ks), width)]

    def step():
        built = []
        t_a = time.perf_counter()
        for j, ks in enumerate(groups):
            fields, lams, mus = [], [], []
            for k in ks:
                lab, ph, mats, dom = resident[k]
                field = pm.build_field_device(lab, mats)
                lam, mu, _ = select(field, ph, mats)
                fields.append(field)
                lams.append(lam)
                mus.append(mu)
            h = dom / np.array(ph.shape[::-1], dtype=float)
            if len(ks) == 1:
                cell = Cell(fields[0], ph.shape, h, stream=pool[j % streams])
                lam, mu = lams[0], mus[0]
            else:
                cell = Cell.stack(fields, ph.shape, h, stream=pool[j % streams])
                lam, mu = lams, mus
            got = tuner.tune(cell, "float32", 6 * len(ks), entries=entries)  # tunes only on a cache miss
            configs[(ph.shape[0], len(ks))] = (got["tile_y"], got["schedule"], got["zmin"])
            built.append((cell, Solver(cell, lam, mu, mode="mixed", method=method)))
        t_b = time.perf_counter()
        outs = run_batch([b[1] for b in built], tol=1e-8, workers=workers)
        t_c = time.perf_counter()
        for cell, solver in built:
            solver.close()
            cell.close()
        t_d = time.perf_counter()
        split["setup"] += t_b - t_a
        split["solve"] += t_c - t_b
        split["teardown"] += t_d - t_c
        return outs

    for _ in range(warmup):
        step()
    for key in split:
        split[key] = 0.0
    tuner.save(entries, strict=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        outs = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    its = [int(i) for o in outs for i in o["iterations"]]
    worst = max(float(r) for o in outs for r in o["true_residuals"])
    voxels = int(sum(ph.size for _, ph, _, _ in resident))
    edges = "-".join(str(v) for v in (min(sizes), max(sizes))) if len(sizes) > 1 else str(sizes[0])
    out = dict(workload=f"random_inclusion_batch_{count}_cells_edges_{edges}_6loads", method=method, cells=count,
               voxels=voxels, stack=stack, stack_voxels=stack_voxels, solvers=len(groups),
               cells_per_s=count / dt, ms_per_step=1e3 * dt, voxel_loads_per_s=6 * voxels / dt, steps=steps,
               warmup=warmup, streams=streams, workers=workers, iterations_mean=float(np.mean(its)),
               stencil_config_by_edge={f"{k[0]}x{k[1]}" if k[1] > 1 else str(k[0]):
                                       dict(tile=f"32x{t}", schedule=v, zmin=z)
                                       for k, (t, v, z) in sorted(configs.items())},
               graph_blocks=int(sum(o["graph_blocks"] for o in outs)),
               host_split_ms={k: 1e3 * v / steps for k, v in split.items()},
               iterations_max=max(its), max_true_residual=worst)
    print(f"# batch ({len(groups)} solvers, stack {stack}): {1e3 * dt:.1f} ms per {count}-cell batch "
          f"({count / dt:.1f} cells/s), iterations mean "
          f"{np.mean(its):.1f} max {max(its)}, worst residual {worst:.2e}; set-up "
          f"{1e3 * split['setup'] / steps:.0f} ms, solves {1e3 * split['solve'] / steps:.0f} ms, teardown "
          f"{1e3 * split['teardown'] / steps:.0f} ms", file=sys.stderr, flush=True)
    return out


TRAFFIC_SUMMARY = ROOT / "profiles" / "apply_kernel_v2_fp32_r6_polycrystal256_traffic.md"


def measured_traffic():
    """HBM bytes per launch of the FP32 six-RHS stencil from the committed
    FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or (None, None)."""
    import re
    if not TRAFFIC_SUMMARY.is_file():
        return None, None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", TRAFFIC_SUMMARY.read_text(), flags=re.M)
    return (float(m.group(1)), str(TRAFFIC_SUMMARY.relative_to(ROOT))) if m else (None, None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--preconditioner", default="fft", choices=["fft", "fft_bf16", "jacobi"],
                    help="mixed-mode preconditioner of the timed solve (fft: FP32 spectra; fft_bf16: BF16-stored "
                         "spectra; jacobi: 1/diag K control)")
    ap.add_argument("--split", type=int, default=1, choices=[1, 2, 3, 6],
                    help="solve the six loads as this many concurrent block solves on separate streams")
    ap.add_argument("--breakdown", action="store_true", help="time the parts of one step (diagnostic)")
    ap.add_argument("--skip-batch", action="store_true",
                    help="skip the 64-cell uneven batch reported beside the headline number")
    ap.add_argument("--batch-only", action="store_true", help="run only the 64-cell batch leg")
    ap.add_argument("--cpu-sample", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-end-to-end", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--skip-mfma", action="store_true", help="skip the matrix-core stencil A/B")
    ap.add_argument("--skip-block", action="store_true", help="skip the block-CG A/B")
    ap.add_argument("--skip-small", action="store_true", help="skip the 32^3 / 16^3 sphere cells on the GPU")
    ap.add_argument("--skip-woven", action="store_true",
                    help="skip the reference's woven 192^3 / 256^3 cells (its published headline workload)")
    ap.add_argument("--woven-only", action="store_true", help="run only the woven reference-workload leg")
    ap.add_argument("--skip-adjoint", action="store_true", help="skip the adjoint-gradient timing leg")
    ap.add_argument("--skip-c2", action="store_true", help="skip the 128^3 woven-like cell")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint-gradient leg and the 128^3 woven-like cell")
    ap.add_argument("--woven-grids", type=int, nargs="+", default=[192, 256], choices=[192, 256],
                    help="woven cells of the reference-workload leg")
    ap.add_argument("--woven-repeats", type=int, default=3, help="timed calls per woven cell and precision")
    ap.add_argument("--stencil", default="modal", choices=["modal", "mfma"],
                    help="FP32 operator of the timed mixed solve: modal VALU stencil or matrix-core form")
    ap.add_argument("--frames-only", action="store_true",
                    help="only time the FP32 six-RHS stencil at 256^3 on random per-voxel frames: "
                         "per-voxel coefficient rows against frames rebuilt in the stencil")
    ap.add_argument("--stencil-only", action="store_true",
                    help="only run 10 six-RHS FP32 stencil applies and a 1 GiB device copy "
                         "(the calibration kernel for FETCH_SIZE/WRITE_SIZE counter runs)")
    args = ap.parse_args()
    if args.cpu_sample:
        print(json.dumps(cpu_sample()), flush=True)
        return
    if args.cpu_end_to_end:
        print(json.dumps(cpu_end_to_end()), flush=True)
        return
    if args.frames_only:
        print(json.dumps(dict(frames_stencil=frames_leg(args.n))), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps(dict(error="this benchmark runs on one GPU; sharding across GPUs is not built",
                              n_gpus=args.gpus)))
        sys.exit(2)
    if args.batch_only:
        import torch
        torch.cuda.set_device(0)
        print(json.dumps(dict(batch=batch_leg(steps=args.steps, warmup=args.warmup))), flush=True)
        return
    if args.woven_only:
        import torch
        torch.cuda.set_device(0)
        print(json.dumps(dict(woven_reference_workload=woven_leg(args.woven_grids, args.woven_repeats))), flush=True)
        return
    if args.adjoint_only:
        import torch
        torch.cuda.set_device(0)
        print(json.dumps(dict(adjoint=adjoint_leg(), woven_like_128=c2_leg())), flush=True)
        return
    sample = samples_all = None
    e2e_proc = None
    host = host_info()
    all_threads = max(1, min(16, host["affinity"] or host["cpu_count"] or 1))
    if not (args.skip_cpu or args.stencil_only):
        t0 = time.perf_counter()
        sample = cpu_baseline_child(1)
        print(f"# cpu sample {sample} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        if all_threads > 1:
            t0 = time.perf_counter()
            samples_all = cpu_baseline_concurrent(all_threads)
            print(f"# cpu sample, {all_threads} concurrent processes ({time.perf_counter() - t0:.1f} s)",
                  file=sys.stderr, flush=True)
        # Whole oracle solves (C1 FFT-PCG, 16^3 sparse direct) on one host core
        # while the GPU legs run; the child never touches the GPU.
        import subprocess
        e2e_proc = subprocess.Popen([sys.executable, str(Path(__file__).resolve()), "--cpu-end-to-end"],
                                    env=_child_env(1), stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
    import torch
    from voxel_lmfem.cells import voronoi_polycrystal
    from voxel_lmfem import materials as pm
    from voxel_lmfem.gpu import Cell, Solver
    from voxel_lmfem.reference_medium import select

    n = args.n
    t0 = time.perf_counter()
    (phase, mats, dom, _), meta = voronoi_polycrystal(n, grains=200, seed=3)
    h = dom / np.array(phase.shape[::-1], dtype=float)
    print(f"# generated {n}^3 polycrystal in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.set_device(0)
    # Inputs resident in HBM before any timing: the int32 grain labels (the
    # per-grain tensors are a 200-entry host table).
    labels = torch.from_numpy(phase.astype(np.int32)).to("cuda")
    torch.cuda.synchronize()

    def setup(mode):
        field = pm.build_field_device(labels, mats)
        lam, mu, rmeta = select(field, phase, mats)
        cell = Cell(field, phase.shape, h)
        if mode == "mixed" and args.stencil == "mfma":
            cell.set_mfma(True)
        kind = args.preconditioner if mode == "mixed" else "fft"
        return cell, Solver(cell, lam, mu, mode=mode, nrhs=6, preconditioner=kind), rmeta

    split_streams = [torch.cuda.Stream() for _ in range(args.split)] if args.split > 1 else []

    def step(mode):
        """One full C_eff: device label front end, reference medium, operator and
        preconditioner set-up, RHS, block PCG, FP64 certificate, stress, teardown."""
        if args.split > 1 and mode == "mixed":
            return split_step(mode)
        cell, solver, rmeta = setup(mode)
        out = solver.run()
        solver.close()
        cell.close()
        return out, rmeta

    def split_step(mode):
        """The six loads as `split` independent block solves on their own
        streams, driven concurrently (vxl_solver_run_batch)."""
        from voxel_lmfem.gpu import run_batch
        field = pm.build_field_device(labels, mats)
        lam, mu, rmeta = select(field, phase, mats)
        torch.cuda.current_stream().synchronize()
        per = 6 // args.split
        cells = [Cell(field, phase.shape, h, stream=st) for st in split_streams]
        solvers = [Solver(c, lam, mu, mode=mode, nrhs=per) for c in cells]
        outs = run_batch(solvers, macro=np.eye(6), workers=args.split)
        for sv, c in zip(solvers, cells):
            sv.close()
            c.close()
        out = dict(CH=np.concatenate([o["CH"] for o in outs], axis=1),
                   iterations=sum((o["iterations"] for o in outs), []),
                   restarts=sum((o["restarts"] for o in outs), []),
                   true_residuals=sum((o["true_residuals"] for o in outs), []),
                   device_ms={}, graph_blocks=sum(o["graph_blocks"] for o in outs))
        return out, rmeta

    if args.breakdown:
        # Host-side split of one step (synchronising between the parts).
        for _ in range(3):
            marks = [("start", time.perf_counter())]

            def mark(name):
                torch.cuda.synchronize()
                marks.append((name, time.perf_counter()))
            field = pm.build_field_device(labels, mats)
            mark("label front end")
            lam, mu, _ = select(field, phase, mats)
            mark("reference medium")
            cell = Cell(field, phase.shape, h)
            mark("cell")
            solver = Solver(cell, lam, mu, mode="mixed", nrhs=6)
            mark("solver + preconditioner set-up")
            solver.run()
            mark("run")
            solver.close()
            cell.close()
            mark("teardown")
            print("# breakdown ms: " + ", ".join(f"{marks[i][0]} {1e3 * (marks[i][1] - marks[i - 1][1]):.2f}"
                                                 for i in range(1, len(marks))), file=sys.stderr, flush=True)
        return

    if args.stencil_only:
        cell, solver, _ = setup("mixed")
        u = torch.randn(6, 3, cell.n, dtype=torch.float32, device="cuda")
        o = torch.empty_like(u)
        for _ in range(10):
            cell.apply_planes(u, o)
        a = torch.empty(2 ** 28, dtype=torch.float32, device="cuda").fill_(1.0)  # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps(dict(stencil_only=True, n=n, copy_bytes=a.numel() * 4)), flush=True)
        return

    for _ in range(args.warmup):
        step("mixed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = rmeta = None
    step_ms = []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        out, rmeta = step("mixed")  # ends with teardown, which synchronises the stream
        step_ms.append(1e3 * (time.perf_counter() - t1))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    its = out["iterations"]
    print(f"# mixed: {ms_per_step:.1f} ms/solve (set-up included), iterations {its}, restarts {out['restarts']}, "
          f"residuals {max(out['true_residuals']):.2e}, device ms {out['device_ms']}, reference medium "
          f"{rmeta['status']} lam={rmeta['lam']:.3f} mu={rmeta['mu']:.3f}", file=sys.stderr, flush=True)
    CH_mixed = out["CH"]

    # The same solve on a resident operator (set-up excluded), for the breakdown,
    # launched kernel by kernel and with the check blocks replayed from graphs.
    cell, solver, _ = setup("mixed")
    resident = {}
    saved = os.environ.get("VXL_GRAPH")
    for mode in ("0", "1"):
        os.environ["VXL_GRAPH"] = mode
        solver.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            got = solver.run()
        torch.cuda.synchronize()
        resident[mode] = (1e3 * (time.perf_counter() - t0) / 3, int(got["graph_blocks"]))
    if saved is None:
        del os.environ["VXL_GRAPH"]
    else:
        os.environ["VXL_GRAPH"] = saved
    resident_ms = resident["0"][0]
    solver.close()
    # Block CG over the six loads on the same resident operator.
    block = None
    if not args.skip_block:
        bsolver = Solver(cell, rmeta["lam"], rmeta["mu"], mode="mixed", nrhs=6, method="block")
        bsolver.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        got = bsolver.run()
        torch.cuda.synchronize()
        block = dict(method="block", resident_ms_per_solve=1e3 * (time.perf_counter() - t0),
                     iterations=[int(i) for i in got["iterations"]],
                     ceff_rel_diff_vs_cg=float(np.linalg.norm(got["CH"] - CH_mixed) / np.linalg.norm(CH_mixed)))
        bsolver.close()
        print(f"# block CG, resident operator: {block['resident_ms_per_solve']:.1f} ms/solve, iterations "
              f"{block['iterations']}", file=sys.stderr, flush=True)
    print(f"# mixed, resident operator: {resident_ms:.1f} ms/solve; graph replay {resident['1'][0]:.1f} ms "
          f"({resident['1'][1]} blocks)", file=sys.stderr, flush=True)

    # Plain FP64 GPU solve of the same cell, same step definition.
    fp64 = None
    if not args.skip_fp64:
        step("fp64")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        o64, _ = step("fp64")
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        err = float(np.linalg.norm(CH_mixed - o64["CH"]) / np.linalg.norm(o64["CH"]))
        fp64 = dict(ms_per_step=1e3 * t64, iterations=o64["iterations"], speedup=t64 / (elapsed / args.steps),
                    ceff_rel_err_mixed_vs_fp64=err)
        print(f"# fp64: {1e3 * t64:.1f} ms/solve, iterations {o64['iterations']}, mixed-vs-fp64 rel err {err:.2e}",
              file=sys.stderr, flush=True)

    # Stencil roofline: FP32 six-RHS apply timed with HIP events recorded on
    # the stream the kernel is launched on (the cell's stream).
    cell.set_mfma(False)
    N = cell.n
    R = 6
    stream = torch.cuda.current_stream()
    u = torch.randn(R, 3, N, dtype=torch.float32, device="cuda")
    o = torch.empty_like(u)
    for _ in range(3):
        cell.apply_planes(u, o)
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        cell.apply_planes(u, o)
    e1.record(stream)
    torch.cuda.synchronize()
    t_apply = e0.elapsed_time(e1) / reps * 1e-3
    alg_bytes = N * (12 * R + 12 * R + 1)
    achieved = alg_bytes / t_apply / 1e9
    u64 = torch.randn(1, 3, N, dtype=torch.float64, device="cuda")
    o64b = torch.empty_like(u64)
    cell.apply_planes(u64, o64b)
    e0.record(stream)
    for _ in range(reps):
        cell.apply_planes(u64, o64b)
    e1.record(stream)
    torch.cuda.synchronize()
    t_apply64 = e0.elapsed_time(e1) / reps * 1e-3
    print(f"# stencil fp32 R=6: {t_apply * 1e3:.3f} ms/launch, {achieved:.0f} GB/s algorithmic; "
          f"fp64 R=1: {t_apply64 * 1e3:.3f} ms/launch, {N * 49 / t_apply64 / 1e9:.0f} GB/s",
          file=sys.stderr, flush=True)

    # A/B: the same FP32 six-RHS apply on the matrix cores (label-grouped dense
    # element matrices, FP16 hi/lo MFMA products), same buffers and events.
    mfma = None
    if not args.skip_mfma:
        cell.set_mfma(True)
        cell.apply_planes(u, o)
        e0.record(stream)
        for _ in range(reps):
            cell.apply_planes(u, o)
        e1.record(stream)
        torch.cuda.synchronize()
        t_mfma = e0.elapsed_time(e1) / reps * 1e-3
        cell.set_mfma(False)
        # Issued matrix-core work: every element of a layer (33 x 8 per 32 x 7
        # output tile, one recomputed layer per z chunk) runs 2 row tiles x 3
        # products of 16x16x32 (16384 flop each) per 16 element columns.
        tiles = -(-n // 32) * -(-n // 7)
        chunks = max(1, min(-(-4096 // (tiles * R)), n // 8))
        zc = -(-n // chunks)
        chunks = -(-n // zc)
        elems = tiles * 33 * 8 * (n + chunks) * R
        issued = elems * 6 * 16384 / 16
        mfma = dict(kernel="mfma_apply_kernel<uint8> R=6, 32x7 tile, v_mfma_f32_16x16x32_f16 (3 products)",
                    ms_per_launch=t_mfma * 1e3, achieved=alg_bytes / t_mfma / 1e9,
                    frac=alg_bytes / t_mfma / 1e9 / HBM_PEAK_GBS, vs_modal=t_apply / t_mfma,
                    elements_per_launch=elems, mfma_tflops_issued=issued / t_mfma / 1e12,
                    mfma_frac_of_fp16_peak=issued / t_mfma / 2.5e15)
        print(f"# stencil fp32 R=6 on the matrix cores: {t_mfma * 1e3:.3f} ms/launch "
              f"({t_apply / t_mfma:.2f}x the modal stencil)", file=sys.stderr, flush=True)
    cell.close()

    cpu = cpu_all = None
    if sample is not None:
        cpu = cpu_baseline(sample, int(sum(its)), n)
        cpu["host"] = host
        print(f"# cpu baseline: {cpu}", file=sys.stderr, flush=True)
    if samples_all is not None:
        cpu_all = cpu_baseline_aggregate(samples_all, int(sum(its)), n)

    batch = None
    if not args.skip_batch:
        # Check blocks launched kernel by kernel (VXL_GRAPH=0) and replayed
        # from captured HIP graphs (VXL_GRAPH=1), same cells.
        saved = os.environ.get("VXL_GRAPH")
        legs = {}
        for mode in ("0", "1"):
            os.environ["VXL_GRAPH"] = mode
            legs[mode] = batch_leg(steps=2, warmup=1)
        if saved is None:
            del os.environ["VXL_GRAPH"]
        else:
            os.environ["VXL_GRAPH"] = saved
        batch = legs["0"]
        batch["graph_replay"] = legs["1"]
        if not args.skip_block:
            batch["block_cg"] = batch_leg(steps=2, warmup=1, method="block")
        # Many small cells: 64 inclusion cells of 32^3, one solver per cell
        # against stacks of ten cells per launch.
        batch["small_cells"] = dict(per_cell=batch_leg(steps=3, warmup=1, sizes=(32,)),
                                    stacked=batch_leg(steps=3, warmup=1, sizes=(32,), stack=10))

    # The same small cells on the GPU (set-up included, as a step above): the
    # 32^3 sphere cell by FP64 and mixed PCG, and the 16^3 sphere cell.
    small = None
    if not args.skip_small:
        from voxel_lmfem.cells import periodic_spheres
        small = {}
        for name, (nn, rad) in (("c1_32", (32, 3.0)), ("spheres_16", (16, 1.5))):
            (sph, smats, sdom, _), _ = periodic_spheres(nn, radius=rad, seed=1)
            slab = torch.from_numpy(sph.astype(np.int32)).to("cuda")
            for mode in ("fp64", "mixed"):
                def small_step():
                    field = pm.build_field_device(slab, smats)
                    lam, mu, _ = select(field, sph, smats)
                    c = Cell(field, sph.shape, sdom / np.array(sph.shape[::-1], dtype=float))
                    sv = Solver(c, lam, mu, mode=mode, nrhs=6)
                    o = sv.run()
                    sv.close()
                    c.close()
                    return o
                small_step()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(10):
                    o = small_step()
                torch.cuda.synchronize()
                dt = (time.perf_counter() - t0) / 10
                small[f"{name}_{mode}"] = dict(ms_per_solve=1e3 * dt, solves_per_s=1.0 / dt,
                                               iterations=[int(i) for i in o["iterations"]],
                                               CH=np.asarray(o["CH"]).tolist())
        print("# small cells on the GPU: " + ", ".join(f"{k} {v['ms_per_solve']:.2f} ms" for k, v in small.items()),
              file=sys.stderr, flush=True)

    cpu_e2e = None
    if e2e_proc is not None:
        text, _ = e2e_proc.communicate(timeout=900)
        if e2e_proc.returncode == 0:
            got = json.loads(text.strip().splitlines()[-1])
            cpu_e2e = dict(
                c1=dict(value=1.0 / got["c1_seconds"], unit="solves/s", cores=1, kind="port",
                        sample=("whole FP64 matrix-free FFT-PCG C_eff of the 32^3 sphere cell (six loads, "
                                f"tol 1e-8) on one host thread: {got['c1_seconds']:.1f} s, not extrapolated")),
                direct16=dict(value=1.0 / got["direct16_seconds"], unit="solves/s", cores=1, kind="port",
                              sample=("sparse-direct (SuperLU) C_eff of a 16^3 sphere cell, the reference's "
                                      f"CPU-backend path, on one host thread: {got['direct16_seconds']:.1f} s")))
            # C_eff of the GPU solves against the FP64 CPU oracle on the same cells.
            if small is not None:
                for name, key in (("c1_32", "c1_CH"), ("spheres_16", "direct16_CH")):
                    ref = np.asarray(got[key])
                    for mode in ("fp64", "mixed"):
                        leg = small[f"{name}_{mode}"]
                        leg["ceff_rel_err_vs_cpu_oracle"] = float(np.linalg.norm(np.asarray(leg["CH"]) - ref) /
                                                                  np.linalg.norm(ref))
            print(f"# cpu end-to-end: c1 {got['c1_seconds']:.1f} s, direct16 {got['direct16_seconds']:.1f} s",
                  file=sys.stderr, flush=True)

    woven = None if args.skip_woven else woven_leg(args.woven_grids, args.woven_repeats)
    adj = None if args.skip_adjoint else adjoint_leg()
    c2 = None if args.skip_c2 else c2_leg()

    traffic, traffic_src = measured_traffic()
    line = dict(
        metric="256^3 polycrystal C_eff solves/s (mixed FP32/FP64, 6 loads, FP64 residual <= 1e-8)",
        value=args.steps / elapsed, unit="solves/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms_per_step, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32 inner + f64 refinement", data="synthetic",
        config=dict(workload=f"voronoi_polycrystal_{n}^3_200grains_cubic_6loads", grid=[n, n, n], grains=200,
                    loads=6, precision="mixed", preconditioner=args.preconditioner, stencil=args.stencil,
                    tol=1e-8, iterations=its,
                    step=("device label front end + reference medium + operator/preconditioner set-up + "
                          "RHS + block PCG + FP64 certificate + stress + teardown; labels resident in HBM"),
                    resident_operator_ms_per_solve=resident_ms,
                    resident_graph_replay_ms_per_solve=resident["1"][0], graph_blocks_per_solve=resident["1"][1],
                    reference_medium=rmeta,
                    step_ms=[round(t, 2) for t in step_ms], median_ms_per_step=float(np.median(step_ms))),
        roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=achieved / HBM_PEAK_GBS, traffic=traffic, traffic_source=traffic_src,
                      kernel="apply_kernel_v2<float,uint8> R=6, 32x14 tile, cubic-voxel form (schedule 3)",
                      ms_per_launch=t_apply * 1e3, algorithmic_bytes=alg_bytes),
        fp64=fp64,
        mfma_stencil=mfma,
        block_cg=block,
        batch=batch,
        cpu_baseline=cpu,
        cpu_baseline_all_cores=cpu_all,
        cpu_end_to_end=cpu_e2e,
        gpu_small_cells=small,
        woven_reference_workload=woven,
        adjoint=adj,
        woven_like_128=c2,
    )
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
