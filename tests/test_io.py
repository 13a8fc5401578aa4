"""Writers and post-processing of the reconstruction driver (test.py:82-93):
TIFF pages read back by an independent reader (Pillow), NPY round trip, the
background subtraction against its closed form."""
import os

import numpy as np
import pytest

from ptychotomo_amd import io


def test_write_tiff_multipage(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(0)
    a = rng.standard_normal((3, 5, 7)).astype(np.float32)
    fn = io.write_tiff(a, str(tmp_path / "sub" / "vol"))
    assert fn.endswith("vol.tiff") and os.path.exists(fn)
    im = Image.open(fn)
    assert im.n_frames == 3
    for k in range(3):
        im.seek(k)
        assert im.mode == "F"
        np.testing.assert_array_equal(np.array(im), a[k])


def test_write_tiff_2d_and_rejects_complex(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    a = np.arange(12, dtype=np.float64).reshape(3, 4)
    fn = io.write_tiff(a, str(tmp_path / "img.tiff"))
    np.testing.assert_array_equal(np.array(Image.open(fn)), a.astype(np.float32))
    with pytest.raises(TypeError):
        io.write_tiff(a + 1j, str(tmp_path / "c"))


def test_write_npy(tmp_path):
    a = np.arange(6, dtype=np.float32).reshape(2, 3)
    fn = io.write_npy(a, str(tmp_path / "d" / "x"))
    np.testing.assert_array_equal(np.load(fn), a)


def test_subtract_background_numpy_and_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    u = (rng.standard_normal((10, 4, 4)) + 1j * rng.standard_normal((10, 4, 4))).astype(np.complex64)
    expect = u.copy()
    expect.real -= u[4:8].real.mean()
    got = io.subtract_background(u.copy())
    np.testing.assert_allclose(got, expect, rtol=1e-6, atol=1e-6)
    t = io.subtract_background(torch.from_numpy(u.copy()))
    np.testing.assert_allclose(t.numpy(), expect, rtol=1e-6, atol=1e-6)
    assert np.array_equal(got.imag, u.imag)


def test_save_reconstruction_layout(tmp_path):
    u = np.ones((2, 3, 3), np.complex64) * (0.5 - 0.25j)
    psi = np.exp(1j * np.ones((4, 2, 3), np.float32)).astype(np.complex64)
    files = io.save_reconstruction(u, psi, "run", str(tmp_path), fmt="npy")
    names = [os.path.relpath(f, tmp_path) for f in files]
    assert names == ["beta/betarun.npy", "delta/deltarun.npy", "psi/psianglerun.npy",
                     "psi/psiamprun.npy"]
    np.testing.assert_allclose(np.load(files[0]), -0.25)
    np.testing.assert_allclose(np.load(files[1]), -0.5)   # sign change of delta
    np.testing.assert_allclose(np.load(files[2]), 1.0, rtol=1e-6)
    np.testing.assert_allclose(np.load(files[3]), 1.0, rtol=1e-6)
