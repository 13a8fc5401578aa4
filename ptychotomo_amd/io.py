"""Output post-processing and writers of the reconstruction driver
(test.py:82-93).

The reference subtracts a background from delta and writes beta, -delta and
the amplitude/phase of psi with dxchange.write_tiff.  dxchange is not a
dependency here: ``write_tiff`` writes the same kind of file (one
uncompressed little-endian float32 page per leading-axis slice, baseline
TIFF 6.0 with SampleFormat = IEEE float), and ``write_npy`` is the NumPy
alternative.
"""
import os
import struct

import numpy as np


def _host(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def subtract_background(u, rows=slice(4, 8)):
    """test.py:83: u.real -= mean(u[4:8].real), in place; returns u."""
    if hasattr(u, "real") and hasattr(u, "detach"):
        u.real.sub_(u[rows].real.mean())
        return u
    u.real -= np.mean(u[rows].real)
    return u


def _fname(path, ext):
    if not path.endswith(ext):
        path = path + ext
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    return path


def write_tiff(data, fname):
    """Write a 2D or 3D real array as a float32 (multi-page) TIFF, creating
    the directory and appending ".tiff" like dxchange.write_tiff.  Returns
    the file name."""
    a = _host(data)
    if np.iscomplexobj(a):
        raise TypeError("write_tiff: complex data; write real and imaginary parts separately")
    a = np.ascontiguousarray(a, dtype="<f4")
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise ValueError("write_tiff: expected a 2D or 3D array")
    npage, h, w = a.shape
    fname = _fname(fname, ".tiff")
    ntag = 10
    ifd_size = 2 + ntag * 12 + 4
    page_bytes = h * w * 4
    with open(fname, "wb") as fh:
        fh.write(b"II" + struct.pack("<HI", 42, 8))
        off = 8
        for k in range(npage):
            data_off = off + ifd_size
            nxt = data_off + page_bytes if k + 1 < npage else 0
            tags = [
                (256, 4, 1, w),            # ImageWidth
                (257, 4, 1, h),            # ImageLength
                (258, 3, 1, 32),           # BitsPerSample
                (259, 3, 1, 1),            # Compression: none
                (262, 3, 1, 1),            # Photometric: min-is-black
                (273, 4, 1, data_off),     # StripOffsets
                (277, 3, 1, 1),            # SamplesPerPixel
                (278, 4, 1, h),            # RowsPerStrip
                (279, 4, 1, page_bytes),   # StripByteCounts
                (339, 3, 1, 3),            # SampleFormat: IEEE float
            ]
            fh.write(struct.pack("<H", ntag))
            for tag, typ, cnt, val in tags:
                if typ == 3:
                    fh.write(struct.pack("<HHIHH", tag, typ, cnt, val, 0))
                else:
                    fh.write(struct.pack("<HHII", tag, typ, cnt, val))
            fh.write(struct.pack("<I", nxt))
            fh.write(a[k].tobytes())
            off = data_off + page_bytes
    return fname


def write_npy(data, fname):
    """np.save with the directory created; returns the file name."""
    fname = _fname(fname, ".npy")
    np.save(fname, _host(data))
    return fname


def save_reconstruction(u, psi, name, outdir=".", fmt="tiff"):
    """test.py:86-93: beta = u.imag, -delta = -u.real, angle and amplitude of
    psi, under outdir/{beta,delta,psi}/.  Returns the written file names."""
    import torch

    write = write_tiff if fmt == "tiff" else write_npy
    u = _host(u)
    if hasattr(psi, "detach"):
        psi_ang = torch.angle(psi).cpu().numpy()
        psi_amp = torch.abs(psi).cpu().numpy()
    else:
        psi_ang, psi_amp = np.angle(psi), np.abs(psi)
    return [
        write(u.imag, os.path.join(outdir, "beta", "beta" + name)),
        write(-u.real, os.path.join(outdir, "delta", "delta" + name)),
        write(psi_ang, os.path.join(outdir, "psi", "psiangle" + name)),
        write(psi_amp, os.path.join(outdir, "psi", "psiamp" + name)),
    ]
