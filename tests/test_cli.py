"""The `emc` driver binary (emc_amd/bin/emc): configuration errors on CPU, a short synthetic run
on the GPU."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "emc_amd", "bin", "emc")

CONF = """model_side = 24;
rotations_n = 2;
N_images = 64;
max_iterations = 2;
image_prefix = "synthetic";
synth_atoms = 20; synth_radius = 4.0; synth_photons = 800.0; synth_seed = 5;
beamstop = 2.0; outer_radius = 12.0;
model_blur = 0.5;
isOutput = 1; output_loop = 1; output_period = 1;
"""


def test_cli_rejects_bad_config(tmp_path):
    p = tmp_path / "emc.conf"
    p.write_text("model_side = 24;\nno_such_key = 1;\n")
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0 and "no_such_key" in r.stderr


def test_cli_rejects_file_images(tmp_path):
    p = tmp_path / "emc.conf"
    p.write_text('image_prefix = "/data/img_";\n')
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 2 and "synthetic" in r.stderr


@pytest.mark.gpu
def test_cli_synthetic_run(tmp_path):
    print("[cli] synthetic 24^3 run", flush=True)
    p = tmp_path / "emc.conf"
    p.write_text(CONF)
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    assert [l["iteration"] for l in lines] == [0, 1]
    assert all(np.isfinite(l["total_logl"]) for l in lines)
    times = (tmp_path / "ExeTime.data").read_text().split("\n")
    assert times[0].split()[:2] == ["0", "0"]
    m = np.fromfile(tmp_path / "output" / "model_0001.bin", np.float32)
    w = np.fromfile(tmp_path / "output" / "weight_0001.bin", np.float32)
    assert m.size == 24 ** 3 and w.size == 24 ** 3
    assert abs(m[m > 0].mean() - 1.0) < 1e-4   # normalised to mean 1 over positive voxels
    assert np.all((w > 0) | (m == -1) | (m < 0))
    assert all(l["search"] == "full" for l in lines)


@pytest.mark.gpu
def test_cli_with_proposal(tmp_path):
    """proposal_coarse_n = 1 (60 cells): warm-up full searches, then proposal iterations."""
    print("[cli] synthetic 24^3 run with the learned proposal", flush=True)
    p = tmp_path / "emc.conf"
    p.write_text(CONF.replace("max_iterations = 2;", "max_iterations = 5;") +
                 "proposal_coarse_n = 1; proposal_top_k = 8;\n")
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    assert [l["search"] for l in lines][:3] == ["full"] * 3
    assert all(np.isfinite(l["total_logl"]) for l in lines)
    assert lines[3]["search"] in ("proposal", "full")  # full when too many patterns fall back


@pytest.mark.gpu
def test_cli_reference_default_scaling_run(tmp_path):
    """The reference's own emc.conf choices: diff_type = "poisson" (chi^2) with known_intensity =
    false, so every (rotation, pattern) pair gets its intensity scaling (main.cpp:283-295)."""
    print("[cli] synthetic 24^3 run, chi^2 with intensity scaling", flush=True)
    p = tmp_path / "emc.conf"
    p.write_text(CONF + 'diff_type = "poisson";\nknown_intensity = false;\n'
                 "sigma_start = 0.4; sigma_final = 0.12; sigma_half_life = 25;\n")
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    assert [l["iteration"] for l in lines] == [0, 1]
    assert all(np.isfinite(l["total_logl"]) for l in lines)
    m = np.fromfile(tmp_path / "output" / "model_0001.bin", np.float32)
    assert abs(m[m > 0].mean() - 1.0) < 1e-4
