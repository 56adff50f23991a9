"""Optional HDF5 formats (libemc_h5.so): the read_images binning, read_mask, write_model and the model
reader against closed forms, and the files against the HDF5 tools' own reading.  CPU only; skipped
where the library was not built."""
import shutil
import subprocess

import numpy as np
import pytest

import emc_amd
from emc_amd import h5

pytestmark = pytest.mark.skipif(not h5.available(), reason="libemc_h5.so not built (no HDF5 library)")


def test_binned_pattern_closed_form(tmp_path):
    """nx = ny = 8, L = 4, read_stride = 2: model pixel (x, y) takes detector pixels 2x, 2x+1 and
    2y, 2y+1 (emc_helper.cpp:229), value = sum / mask sum, mask = mask sum > 1."""
    n, L = 8, 4
    rng = np.random.default_rng(0)
    img = rng.uniform(-3, 3, size=(n, n)).astype(np.float32)  # [y][x]
    msk = np.ones((n, n), np.float32)
    msk[0, 0] = 0.0   # block (0, 0): three of four pixels
    msk[6:8, 6:8] = 0  # block (3, 3): none
    p = tmp_path / "img_0000.h5"
    h5.write_image(p, img, msk)
    assert h5.image_dims(p) == (n, n)
    pat, mask = h5.read_pattern(p, L, 2)
    want = np.zeros((L, L), np.float32)
    wmask = np.zeros((L, L), np.int32)
    for y in range(L):
        for x in range(L):
            blk = np.abs(img[2 * y:2 * y + 2, 2 * x:2 * x + 2]).astype(np.float64)
            ms = msk[2 * y:2 * y + 2, 2 * x:2 * x + 2].sum()
            want[y, x] = blk.sum() / ms if ms else 0.0
            wmask[y, x] = int(ms > 1)
    assert np.allclose(pat.reshape(L, L), want, rtol=1e-6)
    assert np.array_equal(mask.reshape(L, L), wmask)
    assert wmask[0, 0] == 1 and wmask[3, 3] == 0
    _, m1 = h5.read_pattern(p, 8, 1)
    assert not m1.any()  # stride 1: every mask sum is 1, not > 1 (the reference's test)
    with pytest.raises(emc_amd.EmcError):
        h5.read_pattern(p, 8, 2)  # 8 x 2 pixels do not fit an 8-pixel image


def test_mask_reader(tmp_path):
    n, L = 16, 8
    img = np.ones((n, n), np.float32)
    img[4:6, :] = 0.0  # detector rows 4, 5 are zero: model row y = 2 samples detector row 4
    p = tmp_path / "mask.h5"
    h5.write_image(p, img)
    m = h5.read_mask(p, L, 2).reshape(L, L)
    c = np.arange(L) - L / 2 + 0.5
    inside = np.sqrt(c[None, :] ** 2 + c[:, None] ** 2) <= L / 2
    want = inside.copy()
    want[2, :] = False
    assert np.array_equal(m.astype(bool), want)


def test_model_writer_and_reader(tmp_path):
    L = 6
    rng = np.random.default_rng(1)
    M = rng.normal(size=L ** 3).astype(np.float32)
    W = rng.uniform(-1, 1, size=L ** 3).astype(np.float32)
    p = tmp_path / "model_0000.h5"
    h5.write_model(p, L, M, W)
    on = (W > 0) & (M > 0)
    back = h5.read_model(p, L)
    assert np.array_equal(back, np.where(on, M, 0).astype(np.float32))
    dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if shutil.which(dump):  # the HDF5 tools read the file the same way
        head = subprocess.run([dump, "-H", str(p)], capture_output=True, text=True).stdout
        assert 'DATASET "real"' in head and 'DATASET "mask"' in head
        assert f"( {L}, {L}, {L} )" in head and "H5T_STD_I32LE" in head and "H5T_IEEE_F32LE" in head
    with pytest.raises(emc_amd.EmcError):
        h5.read_model(p, L + 1)
    with pytest.raises(emc_amd.EmcError):
        h5.read_model(tmp_path / "missing.h5", L)


def test_library_exports_every_declared_symbol():
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(root, "include", "emc_h5.h")).read(), flags=re.S)
    syms = sorted(set(re.findall(r"\b(emc_h5_[a-z0-9_]+)\s*\(", txt)))
    assert len(syms) == 6
    L = h5.lib()
    for s in syms:
        assert hasattr(L, s), s
