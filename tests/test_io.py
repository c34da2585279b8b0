"""Frame reading, alpha schedule and .mat writing of the example scripts
(examples/PyHSchunck_Fs3_4.py:35-51,63-130).  CPU only."""
import os

import numpy as np
import pytest
from PIL import Image

from ofri import io


def test_imread_8bit_reference_pair(golden, tmp_path):
    g = golden("piv_pair")
    for k in ("im0", "im1"):
        p = str(tmp_path / (k + ".tif"))
        Image.fromarray(g[k]).save(p)
        a = io.imread(p)
        assert a.dtype == np.float32 and a.shape == (512, 512)
        assert np.array_equal(a, g[k].astype(np.float32))
    assert 5.0 < io.imread(str(tmp_path / "im0.tif")).mean() < 10.0   # the pair's mean intensity is about 7.7


def test_imread_16bit_and_colour(tmp_path):
    rng = np.random.default_rng(2)
    a = rng.integers(0, 4096, (40, 56), dtype=np.uint16)      # 12-bit data in a 16-bit TIFF
    p = str(tmp_path / "f12.tif")
    Image.fromarray(a).save(p)
    assert np.array_equal(io.imread(p), a.astype(np.float32))
    Image.fromarray(np.zeros((8, 8, 3), np.uint8)).save(str(tmp_path / "rgb.tif"))
    with pytest.raises(ValueError):
        io.imread(str(tmp_path / "rgb.tif"))


def test_read_sequence_pairs(tmp_path):
    paths = []
    for k in range(4):
        p = str(tmp_path / ("s_%d.tif" % k))
        Image.fromarray(np.full((6, 7), k * 10, np.uint8)).save(p)
        paths.append(p)
    pairs = list(io.read_sequence(paths))
    assert len(pairs) == 3
    assert [(a[0, 0], b[0, 0]) for a, b in pairs] == [(0, 10), (10, 20), (20, 30)]


def test_alpha_schedule_table():
    base = os.path.join("testImages", "Bits08", "Ni06")
    assert io.alpha_schedule(base, 1, 1) == [21]
    assert io.alpha_schedule(base, 2, 1) == [21, 45]
    assert io.alpha_schedule(base, 2, 2) == [21, 21, 45, 45]
    assert io.alpha_schedule(os.path.join("x", "Bits12", "Ni16"), 3, 1) == [550, 1460, 1460]
    assert io.alpha_schedule("Bits10/Ni01", 1, 3) == [20, 20, 20]
    with pytest.raises(KeyError):
        io.alpha_schedule(os.path.join("testImages", "Bits16", "Ni06"), 1, 1)


def test_save_flow_schema_round_trip(tmp_path):
    import scipy.io
    rng = np.random.default_rng(4)
    U = rng.standard_normal((24, 31)).astype(np.float32)
    V = rng.standard_normal((24, 31)).astype(np.float32)
    fn = str(tmp_path / "flow.mat")
    io.save_flow(U, V, fn)
    u, v = io.load_flow(fn)
    assert u.dtype == np.float32 and np.array_equal(u, U) and np.array_equal(v, V)
    m = scipy.io.loadmat(fn)
    vel, par = m["velocities"][0, 0], m["parameters"][0, 0]
    assert set(vel.dtype.names) == {"u", "v", "iaWidth", "iaHeight", "margins"}
    assert set(vel["margins"][0, 0].dtype.names) == {"top", "left", "bottom", "right"}
    assert int(vel["iaWidth"][0, 0]) == 1 and int(vel["iaHeight"][0, 0]) == 1
    assert set(par.dtype.names) == {"overlapFactor", "imageHeight", "imageWidth"}
    assert float(par["overlapFactor"][0, 0]) == 1.0
    assert int(par["imageHeight"][0, 0]) == 24 and int(par["imageWidth"][0, 0]) == 31
