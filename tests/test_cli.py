"""The `solve` command line on the CPU: argument handling, refusal of an
existing output, exit status 2 with "error: ..." (the reference's
src/voxel_lmfem/cli.py:141-142,166-171,213-235 and tests/test_cli.py:13-26).
The GPU solve itself is tests/test_gpu_boundary.py."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT


def run_cli(*args, cwd, optimize=False):
    env = dict(os.environ, PYTHONPATH=os.pathsep.join([str(ROOT / "voxel-lmfem_amd"),
                                                       os.environ.get("PYTHONPATH", "")]))
    cmd = [sys.executable] + (["-O"] if optimize else []) + ["-m", "voxel_lmfem", *map(str, args)]
    return subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, env=env, timeout=300)


@pytest.fixture
def model(tmp_path):
    phase = np.zeros((4, 4, 4), dtype=np.uint8)
    phase[:2] = 1
    np.savez(tmp_path / "model.npz", phase_id=phase, domain_size=np.array([1.0, 1.0, 1.0]))
    (tmp_path / "materials.json").write_text(json.dumps({"0": dict(kind="isotropic", E=3500.0, nu=0.35),
                                                         "1": dict(kind="isotropic", E=70000.0, nu=0.2)}))
    return tmp_path / "model.npz", tmp_path / "materials.json"


def test_help_lists_the_commands(tmp_path):
    out = run_cli("--help", cwd=tmp_path)
    assert out.returncode == 0
    assert "solve" in out.stdout and "import-model" in out.stdout
    sub = run_cli("solve", "--help", cwd=tmp_path)
    assert sub.returncode == 0
    for flag in ("--input", "--materials", "--output", "--backend", "--warmup"):
        assert flag in sub.stdout
    assert "{cuda,gpu,hip,cpu}" in sub.stdout


def test_existing_output_is_refused_and_kept(tmp_path, model):
    npz, mats = model
    output = tmp_path / "result.json"
    output.write_text('{"CH": "earlier result"}')
    before = output.read_bytes()
    out = run_cli("solve", "--input", npz, "--materials", mats, "--output", output, cwd=tmp_path)
    assert out.returncode == 2
    assert out.stderr.startswith("error: ") and "result.json" in out.stderr
    assert output.read_bytes() == before


@pytest.mark.parametrize("args,fragment", [
    (["--backend", "cpu"], "backend='cpu' is the reference's sparse-direct oracle"),
    (["--input", "missing.npz"], "missing.npz"),
])
def test_input_errors_exit_2(tmp_path, model, args, fragment):
    npz, mats = model
    base = {"--input": str(npz), "--materials": str(mats), "--output": str(tmp_path / "r.json")}
    for key, value in zip(args[::2], args[1::2]):
        base[key] = value
    flat = [x for kv in base.items() for x in kv]
    out = run_cli("solve", *flat, cwd=tmp_path)
    assert out.returncode == 2, out.stderr
    assert out.stderr.startswith("error: ") and fragment in out.stderr
    assert not (tmp_path / "r.json").exists()


def test_npz_without_domain_is_an_input_error(tmp_path, model):
    _, mats = model
    np.savez(tmp_path / "bad.npz", phase_id=np.zeros((2, 2, 2), dtype=np.uint8))
    out = run_cli("solve", "--input", tmp_path / "bad.npz", "--materials", mats, "--output", tmp_path / "r.json",
                  cwd=tmp_path)
    assert out.returncode == 2
    assert "error: Input NPZ must contain phase_id and domain_size" in out.stderr


def test_invalid_material_is_an_input_error(tmp_path, model):
    npz, _ = model
    (tmp_path / "m.json").write_text(json.dumps({"0": dict(kind="isotropic", E=-1.0, nu=0.3),
                                                  "1": dict(kind="isotropic", E=1.0, nu=0.3)}))
    out = run_cli("solve", "--input", npz, "--materials", tmp_path / "m.json", "--output", tmp_path / "r.json",
                  cwd=tmp_path)
    assert out.returncode == 2
    assert "error: Isotropic elasticity requires E>0 and -1<nu<0.5" in out.stderr


def test_unknown_backend_is_an_argument_error(tmp_path, model):
    npz, mats = model
    out = run_cli("solve", "--input", npz, "--materials", mats, "--output", tmp_path / "r.json",
                  "--backend", "opencl", cwd=tmp_path)
    assert out.returncode == 2 and "invalid choice" in out.stderr


def test_optimised_interpreter_is_refused(tmp_path, model):
    npz, mats = model
    out = run_cli("solve", "--input", npz, "--materials", mats, "--output", tmp_path / "r.json", cwd=tmp_path,
                  optimize=True)
    assert out.returncode != 0 and "Run without -O" in out.stderr
    assert not (tmp_path / "r.json").exists()


def test_no_gpu_is_a_runtime_error_without_output(tmp_path, model):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible: the solve would run (tests/test_gpu_boundary.py covers it)")
    npz, mats = model
    out = run_cli("solve", "--input", npz, "--materials", mats, "--output", tmp_path / "r.json", cwd=tmp_path)
    assert out.returncode == 2 and out.stderr.startswith("error: ")
    assert not (tmp_path / "r.json").exists()


def test_import_then_solve_input_round_trip(tmp_path):
    """import-model output is what `solve --input/--materials` reads."""
    np.save(tmp_path / "vol.npy", np.arange(8, dtype=np.int32).reshape(2, 2, 2) % 2)
    (tmp_path / "m.json").write_text(json.dumps({"0": dict(kind="isotropic", E=1.0, nu=0.3),
                                                  "1": dict(kind="isotropic", E=2.0, nu=0.3)}))
    out = run_cli("import-model", "--input", tmp_path / "vol.npy", "--materials", tmp_path / "m.json",
                  "--domain-size", 1, 2, 3, "--out", tmp_path / "imp", cwd=tmp_path)
    assert out.returncode == 0, out.stderr
    printed = json.loads(out.stdout)
    assert printed["status"] == "imported" and printed["shape"] == [2, 2, 2]
    assert printed["domain_size"] == [1.0, 2.0, 3.0]
    again = run_cli("import-model", "--input", tmp_path / "vol.npy", "--materials", tmp_path / "m.json",
                    "--domain-size", 1, 2, 3, "--out", tmp_path / "imp", cwd=tmp_path)
    assert again.returncode == 2 and "error: Choose a new output directory" in again.stderr
    from voxel_lmfem.inputs import load_input
    phase, table, domain, orientation = load_input(tmp_path / "imp" / "model.npz", tmp_path / "imp" / "materials.json")
    assert phase.shape == (2, 2, 2) and orientation is None and set(table) == {"0", "1"}
