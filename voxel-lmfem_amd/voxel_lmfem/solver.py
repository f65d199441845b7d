"""Public solve API: ``homogenize``, a drop-in for the reference's GPU backend.

Same signature, input conventions, errors and result keys as
src/voxel_lmfem/solver.py:9-85 (GPU branch) with the result schema of
src/voxel_lmfem/_core/runtime/benchmark.py:110-111.  The computation runs in
libvxl (hand-written gfx950 HIP kernels): matrix-free FI HEX8 operator, FFT
Green preconditioner and block multi-RHS PCG, certified in FP64.

backend: "cuda" (the reference's name of its GPU backend and the default;
no CUDA code is involved), or "gpu" / "hip" for the same backend.  The result's
``backend`` key repeats the caller's spelling, so a drop-in caller checking
``result["backend"] == "cuda"`` sees what the reference returns.  "cpu" names
the reference's sparse-direct oracle, which this GPU build does not ship.
precision: "mixed" (FP32 operator/preconditioner/CG with FP64 reliable
updates, the default) or "fp64" (everything FP64).
element: "fi" (fully integrated HEX8, the reference's element; default) or
"c3d8r_hg" (the reference's reduced-integration control, femlib/c3d8r.py:62-78,
with `hourglass_coefficient`, default 0.1; its FFT preconditioner uses that
element's symbol).
preconditioner: "fft" (Green operator of the isotropic reference medium, the
default), "fft_bf16" (same, spectra stored in BF16 between FFT passes; mixed
precision only) or "jacobi" (1 / diag K, the reference's scalar control in
repro/bench/chapter_aligned/study.py:84-90).
solver: "cg" (one preconditioned CG recurrence per load case, all six sharing
every launch; default) or "block" (block CG over the six loads: one Krylov
space, R x R step and direction matrices; FFT preconditioner).
stencil: "modal" (the modal VALU stencil, default) or "mfma" (the FP32
operator of the mixed solve as label-grouped dense element matrices on the
matrix cores, the reference's grouped-kernel form,
_core/runtime/fi/grouped_kernels.cu:12-74; FI element, at most 4096
coefficient rows) or "frames" (the modal stencil rebuilding each element's
stiffness from its phase row and a per-voxel FP32 quaternion of its frame
instead of reading a pre-rotated row; FI element, at most 256 labels).  The
FP64 operator and the certificate are the same.
"""
import time

import numpy as np

from . import materials as mat
from .inputs import validate
from .reference_medium import select

GPU_BACKENDS = ("gpu", "hip", "cuda")
TOL = 1e-8


def _prepare_device(phase, table, domain, orientation=None, compact=None):
    """Device bookkeeping: labels uploaded once; label -> row, and the frames and
    Bond-rotated rows of local-frame phases, computed on the GPU (from the
    local voxels' frames alone when `compact` holds them)."""
    import torch
    if phase.dtype in (np.uint8, np.int8, np.int16, np.int32):   # uploaded as is, widened on the device
        labels = torch.from_numpy(np.ascontiguousarray(phase)).to("cuda").to(torch.int32)
    elif phase.dtype == np.uint16:
        labels = torch.from_numpy(np.ascontiguousarray(phase).view(np.int16)).to("cuda").to(torch.int32) & 0xFFFF
    else:
        small = phase.min() >= np.iinfo(np.int32).min and phase.max() <= np.iinfo(np.int32).max
        labels = torch.from_numpy(np.ascontiguousarray(phase, dtype=np.int32 if small else np.int64)).to("cuda")
    field = mat.build_field_device(labels, table, orientation, compact=compact)
    h = domain / np.array(phase.shape[::-1], dtype=float)
    lam, mu, meta = select(field, phase, table)
    return field, h, lam, mu, meta


def _field_statistics(cell, u, phase, loads=(0, 5)):
    """Element-centre energy/stress diagnostics of the reference workload
    (benchmark.py:99-107), with per-phase mean energies in place of its
    woven-specific yarn/matrix split."""
    flat = np.asarray(phase).reshape(-1)
    used = np.unique(flat)
    stats, slices = {}, {}
    nz = cell.shape[0]
    for j in loads:
        _, sig, w = cell.center_fields_planes(u[j:j + 1].contiguous(), np.eye(6)[j][None], strain=False)
        w = w[0].cpu().numpy()
        sig = sig[0].cpu().numpy()
        stats[str(j)] = dict(mean_energy=float(w.mean()), stress_l2=float(np.linalg.norm(sig)),
                             phase_mean_energy={str(int(k)): float(w[flat == k].mean()) for k in used})
        slices[f"energy_load{j}"] = w.reshape(cell.shape)[nz // 2]
        slices[f"stress11_load{j}"] = sig[0].reshape(cell.shape)[nz // 2]
    return stats, slices


def _workload(phase, table, domain, orientation, precision, tol, local, fields=False, preconditioner="fft",
              element="fi", hourglass_coefficient=0.1, max_iter=10000, stencil="modal", solver_method="cg",
              compact=None):
    from .gpu import Cell, Solver
    import torch
    torch.cuda.synchronize()
    start = time.perf_counter()
    stages = []

    def stage(name):
        torch.cuda.synchronize()
        now = time.perf_counter()
        stages.append(dict(name=name, seconds=now - stage.t, peak=torch.cuda.max_memory_allocated()))
        stage.t = now
    stage.t = start
    torch.cuda.reset_peak_memory_stats()
    field, h, lam, mu, meta = _prepare_device(phase, table, domain, orientation, compact)
    stage("material")
    cell = solver = None
    try:
        cell = Cell(field, phase.shape, h, element=element, hourglass_coefficient=hourglass_coefficient)
        if stencil == "mfma":
            cell.set_mfma(True)
        elif stencil == "frames":
            rows_p, local_p, slot_p, quat_p = mat.phase_frames(phase, table, orientation)
            cell.set_frames(rows_p, local_p, slot_p, quat_p)
        solver = Solver(cell, lam, mu, mode=precision, nrhs=6, preconditioner=preconditioner, method=solver_method)
        stage("operator")
        out = solver.run(np.eye(6), tol=tol, return_fields=fields, max_iter=max_iter)
        stage("solve")
        field_stats, slices = {}, {}
        if fields:
            field_stats, slices = _field_statistics(cell, out["fields"], phase)
            stage("fields")
        index_bytes = int(cell.index.numel() * cell.index.element_size())
    finally:  # a failed certificate (ConvergenceError) must not hold the device buffers
        if solver is not None:
            solver.close()
        if cell is not None:
            cell.close()
    seconds = time.perf_counter() - start
    calls = [dict(iterations=int(out["iterations"][j]), restarts=int(out["restarts"][j]),
                  true_residuals=[float(out["true_residuals"][j])],
                  trace=[dict(iteration=abs(i), relative_residual=r, true_residual=i < 0)
                         for i, r in out["trace"][j]]) for j in range(6)]
    result = dict(seconds=seconds, CH=np.asarray(out["CH"]).tolist(), calls=calls, stages=stages,
                  field_statistics=field_stats, peak_allocated_bytes=int(torch.cuda.max_memory_allocated()),
                  peak_reserved_bytes=int(torch.cuda.max_memory_reserved()),
                  device_ms=out["device_ms"], graph_blocks=out["graph_blocks"], stencil=stencil,
                  solver=solver_method)
    rows = int(getattr(field, "row_count", len(field.rows)))
    storage = dict(mode="label_rows", n_elements=int(phase.size), pool_rows=rows, index_bytes=index_bytes,
                   coefficient_bytes=rows * 21 * 8,
                   front_end="device" if getattr(field, "index_dev", None) is not None else "host")
    if stencil == "frames":
        storage["fp32_operator"] = dict(mode="phase_frames", frame_bytes=int(phase.size) * 17)
    operator = dict(variant=("mfma_label_grouped_dense_element" if stencil == "mfma" else
                             "modal_fi_lds_stencil_frames" if stencil == "frames" else
                             "modal_fi_lds_stencil" if element == "fi" else "modal_c3d8r_hg_lds_stencil"),
                    element=element, precision=precision, preconditioner=preconditioner, reference_medium=meta)
    if element != "fi":
        operator["hourglass_coefficient"] = float(hourglass_coefficient)
    if fields:
        result["slices"] = slices
    return result, storage, operator


def homogenize(phase_id, materials, domain_size, *, orientation=None, backend="cuda", warmup=False,
               precision="mixed", tol=TOL, fields=False, preconditioner="fft", element="fi",
               hourglass_coefficient=0.1, max_iter=10000, stencil="modal", solver="cg"):
    """Return effective stiffness in engineering-Voigt order 11,22,33,23,13,12.

    max_iter caps the PCG iterations per load (the reference's 10,000,
    benchmark.py:81-90); a load that has not certified by then raises
    RuntimeError, as a residual above the tolerance does (solver.py:78-79).

    fields=True adds the reference workload's field diagnostics for loads 0
    and 5 (``field_statistics`` and mid-plane ``slices`` of the element-centre
    energy and sigma_11)."""
    if backend not in GPU_BACKENDS + ("cpu",):
        raise ValueError("backend must be 'gpu', 'hip', 'cuda' or 'cpu'")
    if precision not in ("mixed", "fp64"):
        raise ValueError("precision must be 'mixed' or 'fp64'")
    if preconditioner not in ("fft", "fft_bf16", "jacobi"):
        raise ValueError("preconditioner must be 'fft', 'fft_bf16' or 'jacobi'")
    if preconditioner == "fft_bf16" and precision != "mixed":
        raise ValueError("preconditioner='fft_bf16' needs precision='mixed'")
    if element not in ("fi", "c3d8r_hg"):
        raise ValueError("element must be 'fi' or 'c3d8r_hg'")
    if element == "c3d8r_hg" and not (np.isfinite(hourglass_coefficient) and hourglass_coefficient > 0):
        raise ValueError("hourglass_coefficient must be finite and positive")
    if isinstance(max_iter, bool) or not isinstance(max_iter, (int, np.integer)) or max_iter < 1:
        raise ValueError("max_iter must be a positive integer")
    if stencil not in ("modal", "mfma", "frames"):
        raise ValueError("stencil must be 'modal', 'mfma' or 'frames'")
    if stencil in ("mfma", "frames") and element != "fi":
        raise ValueError(f"stencil='{stencil}' implements the FI element only")
    if solver not in ("cg", "block"):
        raise ValueError("solver must be 'cg' or 'block'")
    if solver == "block" and preconditioner == "jacobi":
        raise ValueError("solver='block' runs with the FFT preconditioner")
    phase, table, domain, orientation, labels, tensors, local = validate(
        phase_id, materials, domain_size, orientation)
    if backend == "cpu":
        raise ValueError("backend='cpu' is the reference's sparse-direct oracle and is not part of this "
                         "GPU build; use backend='gpu'")
    compact = None
    if local.any():  # every local frame is validated before any GPU work; the values read are kept
        flat = phase.reshape(-1)
        sels = [mat.voxels_with_label(flat, labels[k]) for k in np.flatnonzero(local)]
        compact = (sels, mat.check_frames(orientation, phase.shape, sels, keep=True))
    try:
        import torch
    except ImportError as exc:
        raise RuntimeError("Install a ROCm-enabled PyTorch build for device memory") from exc
    if not torch.cuda.is_available():
        raise RuntimeError("No GPU is available. This backend runs only on an MI355X (gfx950).")
    if warmup:
        _workload(phase, table, domain, orientation, precision, tol, local, fields, preconditioner, element,
                  hourglass_coefficient, int(max_iter), stencil, solver, compact)
    start = time.perf_counter()
    result, storage, operator = _workload(phase, table, domain, orientation, precision, tol, local, fields,
                                          preconditioner, element, hourglass_coefficient, int(max_iter),
                                          stencil, solver, compact)
    result.update(backend=backend, storage=storage, operator=operator,
                  scope=(("FI HEX8" if element == "fi" else "C3D8R-HG") + " matrix-free stiffness; "
                         + {"fft": "FFT Green preconditioner; ", "fft_bf16": "FFT Green preconditioner (BF16 spectra); ",
                            "jacobi": "Jacobi preconditioner; "}[preconditioner]
                         + ("FP32 PCG with FP64 reliable updates" if precision == "mixed" else "FP64 PCG")
                         + "; FP64 true-residual certificate; six unit engineering strains"))
    residuals = [r for call in result["calls"] for r in call["true_residuals"]]
    if not np.isfinite(residuals).all() or max(residuals) > tol:
        raise RuntimeError(f"True-residual check failed: {residuals}")
    if not np.isfinite(result["CH"]).all():
        raise RuntimeError("Effective stiffness contains nonfinite values")
    result["max_true_residual"] = max(residuals)
    result["elapsed_seconds"] = time.perf_counter() - start
    result["voigt_order"] = ["11", "22", "33", "23", "13", "12"]
    return result


def homogenize_batch(cells, *, precision="mixed", tol=TOL, streams=4, workers=4, autotune=True, tune_cache=None,
                     stack=10, stack_voxels=64 ** 3):
    """C_eff of many global-frame cells of uneven size, solved concurrently.

    cells: iterable of (phase_id, materials, domain_size).  Each cell gets the
    device label front end and its own reference medium.  Cells of the same
    grid and voxel size with at most `stack_voxels` voxels run `stack` at a
    time as one stacked cell (``Cell.stack``: every kernel launch covers all
    of them, six loads each, each cell with its own medium; measured 2x the
    throughput of one solver per cell on 64 cells of 32^3, and no gain where
    80^3 and 96^3 cells dominate, whose launches already fill the GPU);
    stack=1 gives every cell its own solver.
    The solvers sit on `streams` HIP streams and ``vxl_solver_run_batch``
    drives them from `workers` host threads, largest first, so small cells
    overlap on the GPU.
    autotune: pick each solver's stencil tile, schedule and chunk depth with
    the autotuner (tuned once per grid shape, load count and machine, then
    read from the JSON cache `tune_cache`, default $VXL_TUNE_CACHE or
    ~/.cache/voxel_lmfem/stencil_tuning.json).
    Returns one dict per cell, in input order (CH, calls, max_true_residual,
    grid_shape, stacked_with).
    """
    from . import autotune as tuner
    import torch
    from .gpu import Cell, Solver, run_batch
    stack = max(1, min(int(stack), 10))
    pool = [torch.cuda.Stream() for _ in range(max(1, int(streams)))]
    prepared = []
    for phase_id, materials, domain_size in cells:
        phase, table, domain, orientation, labels, tensors, local = validate(phase_id, materials, domain_size)
        if local.any():
            raise ValueError("homogenize_batch takes global-frame cells; use homogenize for local frames")
        field, h, lam, mu, meta = _prepare_device(phase, table, domain)
        prepared.append((phase.shape, field, h, lam, mu, meta))
    by_grid = {}
    for k, item in enumerate(prepared):
        by_grid.setdefault((item[0], tuple(float(v) for v in item[2])), []).append(k)
    groups = []
    for (shape, _), ks in by_grid.items():
        width = stack if int(np.prod(shape)) <= stack_voxels else 1
        groups += [ks[i:i + width] for i in range(0, len(ks), width)]
    built = []
    entries = tuner.load(tune_cache) if autotune else None
    tuned_any = False
    try:
        for j, ks in enumerate(groups):
            shape, h = prepared[ks[0]][0], prepared[ks[0]][2]
            stream = pool[j % len(pool)]
            if len(ks) == 1:
                cell = Cell(prepared[ks[0]][1], shape, h, stream=stream)
                lam, mu = prepared[ks[0]][3], prepared[ks[0]][4]
            else:
                cell = Cell.stack([prepared[k][1] for k in ks], shape, h, stream=stream)
                lam, mu = [prepared[k][3] for k in ks], [prepared[k][4] for k in ks]
            try:
                if autotune:
                    prec = "float32" if precision == "mixed" else "float64"
                    tuned_any |= tuner.tune(cell, prec, 6 * len(ks), entries=entries)["tuned"]
                solver = Solver(cell, lam, mu, mode=precision)
            except BaseException:
                cell.close()  # not yet in `built`, which the finally below closes
                raise
            built.append((cell, solver, ks))
        if tuned_any:
            tuner.save(entries, tune_cache, strict=False)
        outs = run_batch([b[1] for b in built], tol=tol, workers=workers)
    finally:
        for cell, solver, _ in built:
            solver.close()
            cell.close()
    results = [None] * len(prepared)
    for (cell, solver, ks), out in zip(built, outs):
        for g, k in enumerate(ks):
            cols = slice(6 * g, 6 * g + 6)
            residuals = [float(r) for r in out["true_residuals"][cols]]
            if not np.isfinite(residuals).all() or max(residuals) > tol:
                raise RuntimeError(f"True-residual check failed: {residuals}")
            results[k] = dict(CH=np.asarray(out["CH"])[:, cols].tolist(), grid_shape=list(prepared[k][0]),
                              max_true_residual=max(residuals), graph_blocks=out["graph_blocks"],
                              stacked_with=len(ks),
                              calls=[dict(iterations=int(i), restarts=int(r), true_residuals=[float(t)])
                                     for i, r, t in zip(out["iterations"][cols], out["restarts"][cols],
                                                        out["true_residuals"][cols])],
                              reference_medium=prepared[k][5], voigt_order=["11", "22", "33", "23", "13", "12"])
    return results
