"""End-to-end GPU C_eff against the oracle, the reference's KATs and its
archived records."""
import json

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import homogenize as oracle
from test_oracle import gpu_fixture_cell, iso, laminate_expected

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reference_gpu_fixture(torch_gpu, precision):
    """tests/test_gpu.py:10-29 of the reference (4^3 rotated anisotropic)."""
    from voxel_lmfem import homogenize
    print(f"case gpu-fixture {precision}", flush=True)
    phase, table, domain, rot = gpu_fixture_cell()
    ref = oracle.homogenize_direct(phase, table, domain, orientation=rot)
    out = homogenize(phase, table, domain, orientation=rot, backend="cuda", warmup=True, precision=precision)
    assert rel(out["CH"], ref["CH"]) < 1e-8
    assert out["max_true_residual"] <= 1e-8
    assert set(out) >= {"seconds", "CH", "calls", "stages", "field_statistics", "peak_allocated_bytes",
                        "peak_reserved_bytes", "backend", "storage", "operator", "scope", "max_true_residual",
                        "elapsed_seconds", "voigt_order"}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_kats(torch_gpu, precision):
    from voxel_lmfem import homogenize
    print(f"case kats {precision}", flush=True)
    r = homogenize(np.full((2, 3, 4), 17, dtype=np.int32), {17: dict(kind="isotropic", E=3500., nu=.35)},
                   [2., 3., 5.], precision=precision)
    np.testing.assert_allclose(r["CH"], iso(3500., .35), rtol=2e-13, atol=2e-10)
    phase = np.zeros((4, 2, 3), dtype=np.uint8)
    phase[:2] = 1
    specs = {0: dict(kind="isotropic", E=3500., nu=.35), 1: dict(kind="isotropic", E=70000., nu=.2)}
    r = homogenize(phase, specs, [1., 1., 1.], precision=precision)
    np.testing.assert_allclose(r["CH"], laminate_expected([iso(3500., .35), iso(70000., .2)]), rtol=1e-8, atol=1e-5)
    R = np.array([[0., -1., 0.], [1., 0., 0.], [0., 0., 1.]])
    r = homogenize(np.zeros((2, 2, 2), dtype=np.uint8),
                   {0: dict(kind="anisotropic", C=np.diag([10., 20., 30., 4., 5., 6.]), frame="local")},
                   [1., 2., 3.], orientation=R, precision=precision)
    np.testing.assert_allclose(r["CH"], np.diag([20., 10., 30., 5., 4., 6.]), rtol=2e-13, atol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_archived_woven_8(torch_gpu, precision):
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import reference_weave
    print(f"case woven8 {precision}", flush=True)
    rec = json.loads((GOLDEN / "woven_records.json").read_text())["8_fp64_r0"]
    (phase, mats, dom, ori), meta = reference_weave(8)
    assert meta["input_hash"] == rec["input_hash"]
    out = homogenize(phase, mats, dom, orientation=ori, precision=precision)
    assert rel(out["CH"], rec["CH"]) < 1e-8
    ref = out["operator"]["reference_medium"]
    assert (ref["lam"], ref["mu"]) == (rec["reference_lam"], rec["reference_mu"])
    if precision == "fp64":
        its = [c["iterations"] for c in out["calls"]]
        print("  iterations", its, "reference", rec["iterations"], flush=True)
        assert all(abs(a - b) <= 1 for a, b in zip(its, rec["iterations"]))


def test_spheres_32_mixed_vs_oracle(torch_gpu):
    """Config 1 (32^3 random spheres, vf 0.30): mixed GPU vs FP64 CPU oracle."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import periodic_spheres
    print("case spheres32", flush=True)
    (phase, mats, dom, ori), meta = periodic_spheres(32, seed=1)
    ref = oracle.homogenize_matfree(phase, mats, dom)
    for precision in ("fp64", "mixed"):
        out = homogenize(phase, mats, dom, precision=precision)
        err = rel(out["CH"], ref["CH"])
        print(f"  {precision} rel err {err:.2e} its {[c['iterations'] for c in out['calls']]}", flush=True)
        assert err < 1e-7
        assert out["max_true_residual"] <= 1e-8


@pytest.mark.parametrize("grid", [64, 128, 192, 256])
def test_archived_woven_certify(torch_gpu, grid):
    """The reference's certified woven C_eff (certify_woven_{grid}.json, FP64
    operator, residual <= 1e-8) from the same analytic weave."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import reference_weave
    print(f"case woven{grid} certify", flush=True)
    rec = json.loads((GOLDEN / "woven_records.json").read_text())[f"certify_{grid}"]
    (phase, mats, dom, ori), meta = reference_weave(grid)
    assert meta["input_hash"] == rec["input_hash"]
    out = homogenize(phase, mats, dom, orientation=ori, precision="mixed")
    err = rel(out["CH"], rec["CH"])
    print(f"  rel err {err:.2e} its {[c['iterations'] for c in out['calls']]} "
          f"seconds {out['seconds']:.3f}", flush=True)
    assert err < 1e-7
    assert out["max_true_residual"] <= 1e-8


@pytest.mark.parametrize("ring", ["1", "0"])
@pytest.mark.parametrize("precision,preconditioner", [("mixed", "fft"), ("fp64", "fft"), ("mixed", "jacobi")])
def test_graph_blocks_match_eager_launches(torch_gpu, monkeypatch, precision, preconditioner, ring):
    """Blocks between host checks replayed from one captured graph (VXL_GRAPH=1)
    give the same bits, iterations and residual history as the same blocks
    launched one by one under the same rules (VXL_GRAPH=2), with the direction
    ring and the fused r update on (ring "1") or the per-iteration x update;
    against the default launches (ring flushed only when full) the C_eff agrees
    to rounding with the same iterations."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=24, grains=10, seed=8)
    print(f"case graph vs eager {precision} {preconditioner} ring={ring} 24^3", flush=True)
    out = {}
    monkeypatch.setenv("VXL_DIRECTION_RING", ring)
    for flag in ("0", "2", "1"):
        monkeypatch.setenv("VXL_GRAPH", flag)
        out[flag] = homogenize(phase, mats, dom, precision=precision, preconditioner=preconditioner)
    a, b = out["2"], out["1"]
    assert np.array_equal(np.asarray(a["CH"]), np.asarray(b["CH"]))
    assert [c["iterations"] for c in a["calls"]] == [c["iterations"] for c in b["calls"]]
    assert [c["trace"] for c in a["calls"]] == [c["trace"] for c in b["calls"]]
    assert max(c["iterations"] for c in b["calls"]) > 8
    assert a["graph_blocks"] == 0 and b["graph_blocks"] > 0  # the blocks really were replayed
    d = out["0"]  # other flush points: x64 sums in another grouping, certified either way
    assert all(abs(c["iterations"] - e["iterations"]) <= 1 for c, e in zip(d["calls"], b["calls"]))
    assert rel(d["CH"], b["CH"]) < 1e-8 and d["graph_blocks"] == 0


def test_solve_under_other_stencil_configs(torch_gpu, monkeypatch):
    """The CG form of the stencil (p.Kp from element energies) under other tiles,
    schedules and chunk depths: same certified C_eff; the block graph is
    re-captured for the new configuration."""
    from voxel_lmfem import materials as pm
    from voxel_lmfem.cells import voronoi_polycrystal
    from voxel_lmfem.gpu import Cell, Solver
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=24, grains=10, seed=8)
    monkeypatch.setenv("VXL_GRAPH", "1")
    field = pm.build_field(phase, mats)
    cell = Cell(field, phase.shape, dom / 24)
    solver = Solver(cell, 100.0, 40.0, mode="mixed")
    base = solver.run()
    print("case stencil configs in the solve", flush=True)
    for ty, v, z in ((8, 3, 2), (6, 2, 4), (14, 3, 8)):
        cell.set_stencil_config("float32", ty, v, z)
        cell.set_stencil_config("float64", (8, 6, 4)[(14, 8, 6).index(ty)], v, z)
        out = solver.run()
        rel = np.linalg.norm(out["CH"] - base["CH"]) / np.linalg.norm(base["CH"])
        print(f"  tile {ty} schedule {v} zmin {z}: rel {rel:.2e}, its {out['iterations']}", flush=True)
        assert rel < 1e-8 and max(out["true_residuals"]) <= 1e-8
        assert out["graph_blocks"] > 0
    solver.close()
    cell.close()


@pytest.mark.parametrize("precision,preconditioner", [("mixed", "fft"), ("fp64", "fft"), ("mixed", "jacobi")])
def test_direction_ring_matches_per_iteration_update(torch_gpu, monkeypatch, precision, preconditioner):
    """x advanced from the stored directions in one FP64 pass per certification
    against x += alpha p every iteration: same iterations, C_eff to rounding."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=24, grains=10, seed=9)
    print(f"case direction ring {precision} {preconditioner} 24^3", flush=True)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("VXL_DIRECTION_RING", flag)
        out[flag] = homogenize(phase, mats, dom, precision=precision, preconditioner=preconditioner)
    a, b = out["0"], out["1"]
    rel = np.linalg.norm(np.asarray(a["CH"]) - np.asarray(b["CH"])) / np.linalg.norm(np.asarray(a["CH"]))
    its_a = [c["iterations"] for c in a["calls"]]
    its_b = [c["iterations"] for c in b["calls"]]
    print(f"  rel {rel:.2e} iterations {its_a} vs {its_b}", flush=True)
    assert rel < 1e-9
    assert max(abs(x - y) for x, y in zip(its_a, its_b)) <= 1
    assert b["max_true_residual"] <= 1e-8


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_zero_load_among_others(torch_gpu, mode):
    """A zero macro strain among five others (benchmark.py:79: zero RHS -> 0):
    its displacement stays exactly zero through the direction ring and its
    stress column is zero; the other columns equal the six-load solve's.
    Solved twice on one solver, so the second run starts from direction
    slots the first one filled."""
    import torch
    from voxel_lmfem import materials as pm
    from voxel_lmfem.cells import voronoi_polycrystal
    from voxel_lmfem.gpu import Cell, Solver
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=24, grains=10, seed=4)
    print(f"case zero load {mode} 24^3", flush=True)
    field = pm.build_field(phase, mats)
    h = dom / np.array(phase.shape[::-1], dtype=float)
    cell = Cell(field, phase.shape, h)
    solver = Solver(cell, 100.0, 40.0, mode=mode, nrhs=6)
    full = solver.run(return_fields=True)
    macro = np.eye(6)
    macro[3] = 0.0
    for _ in range(2):
        out = solver.run(macro, return_fields=True)
        u = out["fields"]
        assert torch.count_nonzero(u[3]).item() == 0
        assert np.all(out["CH"][:, 3] == 0.0) and out["true_residuals"][3] == 0.0
        keep = [0, 1, 2, 4, 5]
        err = rel(out["CH"][:, keep], full["CH"][:, keep])
        du = float((u[keep] - full["fields"][keep]).abs().max() / full["fields"][keep].abs().max())
        print(f"  CH rel {err:.2e}, u rel {du:.2e}", flush=True)
        assert err < 1e-8 and du < 1e-5
        assert max(out["true_residuals"]) <= 1e-8
    solver.close()
    cell.close()


@pytest.mark.parametrize("ring", ["1", "0"])
def test_fused_residual_update_matches_separate_pass(torch_gpu, monkeypatch, ring):
    """r -= alpha q folded into the preconditioner's x pass (with the direction
    ring, or with x += alpha p there too) against the separate update pass:
    same iterations to one, C_eff to rounding."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=32, grains=12, seed=5)
    print(f"case fused update ring={ring} 32^3", flush=True)
    monkeypatch.setenv("VXL_DIRECTION_RING", ring)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("VXL_FUSED_UPDATE", flag)
        out[flag] = homogenize(phase, mats, dom, precision="mixed")
    a, b = out["0"], out["1"]
    err = rel(b["CH"], a["CH"])
    its_a = [c["iterations"] for c in a["calls"]]
    its_b = [c["iterations"] for c in b["calls"]]
    print(f"  rel {err:.2e} iterations {its_a} vs {its_b}", flush=True)
    assert err < 1e-9
    assert max(abs(x - y) for x, y in zip(its_a, its_b)) <= 1
    assert b["max_true_residual"] <= 1e-8


@pytest.mark.parametrize("shape", [(32, 32, 32), (40, 24, 48)])
def test_spectral_residual_matches_real_space(torch_gpu, monkeypatch, shape):
    """The residual kept as its x spectrum inside the preconditioner (r -= alpha
    q applied to R^ in the x pass, ||r||^2 by Parseval; the default mixed
    form with the direction ring and the fused pass) against the real-space
    residual (VXL_SPECTRAL_RESIDUAL=0): same iterations to one, C_eff to
    rounding, both certified; on a cubic and on a mixed-radix grid."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=shape[0], grains=12, seed=5)
    phase = np.ascontiguousarray(np.resize(phase, shape))
    print(f"case spectral residual {shape}", flush=True)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("VXL_SPECTRAL_RESIDUAL", flag)
        out[flag] = homogenize(phase, mats, dom, precision="mixed")
    a, b = out["0"], out["1"]
    err = rel(b["CH"], a["CH"])
    its_a = [c["iterations"] for c in a["calls"]]
    its_b = [c["iterations"] for c in b["calls"]]
    print(f"  rel {err:.2e} iterations {its_a} vs {its_b}", flush=True)
    assert err < 1e-9
    assert max(abs(x - y) for x, y in zip(its_a, its_b)) <= 1
    assert a["max_true_residual"] <= 1e-8 and b["max_true_residual"] <= 1e-8


@pytest.mark.parametrize("form", ["pair_vectors", "pair_mixed_shapes", "matrix_grid"])
def test_one_frame_for_every_voxel_through_homogenize(torch_gpu, form):
    """One frame for every voxel, given as a broadcast axis pair, as a
    per-voxel axis1 with one axis2, or as a per-voxel rotation array: the
    front end (host check keeping the frames it reads, device rows of the
    selected voxels) gives a one-phase local-frame cell the rotated tensor,
    C_eff = diag(20, 10, 30, 5, 4, 6) for diag(10, 20, 30, 4, 5, 6) turned
    90 degrees about z (the reference's tests/test_solver.py:49-56)."""
    from voxel_lmfem import homogenize
    print(f"case one frame for every voxel, {form}", flush=True)
    shape = (2, 3, 4)
    a1, a2 = np.array([0., 1., 0.]), np.array([-1., 0., 0.])   # R = [a1, c x a1, c], c = e_z
    if form == "pair_vectors":
        ori = dict(axis1=a1, axis2=a2)
    elif form == "pair_mixed_shapes":
        ori = dict(axis1=np.broadcast_to(a1, shape + (3,)).copy(), axis2=a2)
    else:
        R = np.array([[0., -1., 0.], [1., 0., 0.], [0., 0., 1.]])
        ori = np.broadcast_to(R, shape + (3, 3)).copy()
    mats = {0: dict(kind="anisotropic", C=np.diag([10., 20., 30., 4., 5., 6.]), frame="local")}
    for precision in ("fp64", "mixed"):
        r = homogenize(np.zeros(shape, dtype=np.uint8), mats, [1., 2., 3.], orientation=ori, precision=precision)
        np.testing.assert_allclose(r["CH"], np.diag([20., 10., 30., 5., 4., 6.]), rtol=2e-13, atol=1e-12)
