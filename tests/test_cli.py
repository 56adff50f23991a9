"""The `emc` driver binary (emc_amd/bin/emc): configuration errors on CPU, a short synthetic run
on the GPU."""
import json
import os
import subprocess

import numpy as np
import pytest

from emc_amd import h5

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


def test_cli_missing_image_files(tmp_path):
    p = tmp_path / "emc.conf"
    p.write_text('image_prefix = "/nonexistent/img_";\nmask_file = "/nonexistent/mask.h5";\n')
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path)
    if h5.available():
        assert r.returncode == 1 and "mask_file" in r.stderr
    else:
        assert r.returncode == 2 and "libemc_h5" in r.stderr


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
    # normalised to mean 1 over the voxels with weight, then blurred (which moves that mean a little)
    print(f"   mean over the voxels with weight {m[w > 0].mean():.4f}", flush=True)
    assert abs(m[w > 0].mean() - 1.0) < 0.1
    assert np.all(m >= 0) and np.all((m == 0) | (w > 0))  # write_model's mask (file_helper.cpp:50-65)
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
    w = np.fromfile(tmp_path / "output" / "weight_0001.bin", np.float32)
    assert abs(m[w > 0].mean() - 1.0) < 0.1  # normalised, then blurred


@pytest.mark.gpu
def test_cli_reference_conf_choices(tmp_path):
    """The rest of the reference's emc.conf choices: normalize_images = true (non-integer patterns,
    on the matrix cores), the radial average initial model (model_input = 3) and the responsibility
    dump of isDebug."""
    print("[cli] normalize_images, model_input = 3, isDebug", flush=True)
    p = tmp_path / "emc.conf"
    p.write_text(CONF + 'diff_type = "poisson";\nknown_intensity = false;\nnormalize_images = true;\n'
                 "model_input = 3;\nisDebug = 4;\n")
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    assert [l["iteration"] for l in lines] == [0, 1] and all(np.isfinite(l["total_logl"]) for l in lines)
    S = 420  # rotations_n = 2: 10 (n + 5 n^3)
    for it in (0, 1):
        R = np.fromfile(tmp_path / "debug" / f"respons_{it:04d}.data", np.float32)
        assert R.size == S * 64
        assert np.allclose(R.reshape(S, 64).sum(0), 1.0, atol=1e-4)  # normalised over rotations


@pytest.mark.gpu
def test_cli_model_from_file(tmp_path):
    """model_input = 2 reads the initial model from model_file (raw float32, the driver's own output
    format); model_input = 1 inserts the patterns at random orientations."""
    print("[cli] model_input = 2 and 1", flush=True)
    L = 24
    rng = np.random.default_rng(0)
    rng.uniform(0.5, 1.5, L ** 3).astype(np.float32).tofile(tmp_path / "init.bin")
    for extra in (f'model_input = 2;\nmodel_file = "{tmp_path / "init.bin"}";\n', "model_input = 1;\n"):
        p = tmp_path / "emc.conf"
        p.write_text(CONF.replace("max_iterations = 2;", "max_iterations = 1;") + extra)
        r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
        assert r.returncode == 0, r.stderr
        lines = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
        assert len(lines) == 1 and np.isfinite(lines[0]["total_logl"])


def test_cli_rejects_bad_model_input(tmp_path):
    p = tmp_path / "emc.conf"
    p.write_text(CONF + "model_input = 7;\n")
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 2 and "model_input" in r.stderr
    p.write_text(CONF + 'model_input = 2;\nmodel_file = "/nonexistent/model.bin";\n')
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 1 and "model_read_raw" in r.stderr


@pytest.mark.gpu
@pytest.mark.skipif(not h5.available(), reason="libemc_h5.so not built (no HDF5 library)")
def test_cli_reads_hdf5_images(tmp_path):
    """image_prefix%04d.h5 patterns (48 x 48 detector, read_stride 2 onto 24 x 24) and an HDF5
    mask_file; the model is also written as output/model_%04d.h5."""
    print("[cli] HDF5 images", flush=True)
    rng = np.random.default_rng(3)
    for i in range(32):
        h5.write_image(tmp_path / f"img_{i:04d}.h5", rng.poisson(2.0, size=(48, 48)).astype(np.float32))
    h5.write_image(tmp_path / "mask.h5", np.ones((48, 48), np.float32))
    conf = CONF.replace('image_prefix = "synthetic";', f'image_prefix = "{tmp_path}/img_";') \
               .replace("N_images = 64;", "N_images = 32;").replace("max_iterations = 2;", "max_iterations = 1;")
    p = tmp_path / "emc.conf"
    p.write_text(conf + f'mask_file = "{tmp_path}/mask.h5";\nread_stride = 2;\nmodel_input = 3;\n')
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1 and np.isfinite(lines[0]["total_logl"])
    m = h5.read_model(tmp_path / "output" / "model_0000.h5", 24)
    raw = np.fromfile(tmp_path / "output" / "model_0000.bin", np.float32)
    assert np.array_equal(m, raw)  # the same write_model mask on both files
