"""Oracle for the ptychography operator (TEST INFRASTRUCTURE ONLY).

``fwd`` / ``adj``
    FP32 restatement of ``ptychofft::fwd``/``adj`` (``src/cuda/ptychofft.cu:58-93``)
    with the kernels ``mul``, ``mula``, ``takeshifts``, ``shifts``,
    ``shiftsa`` (``src/include/kernels.cuh:243-355``).  Unnormalised, like the
    native class; ``Solver`` multiplies by ``coefptycho``.

``fwd_exact``
    FP64 per-patch closed form: ``c * DFT2(embed_D(prb * patch^T)) * e^{2 pi i
    (frac sx + frac sy)}`` evaluated with explicit DFT matrices (no FFT).

Index conventions (kernels.cuh:251-263): probe row ``ix`` multiplies object
column ``stx+ix``; the detector frame is ``[ix*dety + iy]`` with the probe
centred at offset ((detx-P)/2, (dety-P)/2); scan positions are rounded half
away from zero (``roundf``) and patches with a negative rounded coordinate
are skipped.
"""
import numpy as np
import scipy.fft as sfft


def roundf(x):
    """C ``roundf`` (half away from zero), exact for float32 input."""
    x = np.asarray(x, dtype=np.float32).astype(np.float64)
    return (np.sign(x) * np.floor(np.abs(x) + 0.5)).astype(np.int64)


def shifts_factor(scan, cdt=np.complex64):
    """takeshifts (kernels.cuh:306-317): e^{2 pi i (s - round(s))} for x and
    y, FP64 then complex64.  scan: [2, pt, S] float32."""
    scan = np.asarray(scan, dtype=np.float32)
    fr = scan.astype(np.float64) - roundf(scan)
    ph = np.exp(2j * np.pi * fr)
    return ph[0].astype(cdt), ph[1].astype(cdt)


def _patch_index(scan, P, nz, n):
    stx = roundf(scan[0])          # [pt, S]
    sty = roundf(scan[1])
    valid = (stx >= 0) & (sty >= 0)
    ix = np.arange(P)
    # object row = sty + iy, object column = stx + ix; frame[ix][iy]
    rows = np.clip(sty[:, :, None, None] + ix[None, None, None, :], 0, nz - 1)
    cols = np.clip(stx[:, :, None, None] + ix[None, None, :, None], 0, n - 1)
    return rows, cols, valid


def fwd(psi, prb, scan, detx, dety, dtype=np.complex64):
    """ptychofft::fwd: psi [pt, nz, n] -> g [pt, S, detx, dety], complex64.
    ``dtype=complex128`` gives the FP64 twin of the same steps."""
    cdt = np.dtype(dtype)
    rdt = np.float32 if cdt == np.complex64 else np.float64
    psi = np.asarray(psi, dtype=cdt)
    prb = np.asarray(prb, dtype=cdt)
    pt, nz, n = psi.shape
    P = prb.shape[0]
    S = scan.shape[2]
    rows, cols, valid = _patch_index(scan, P, nz, n)
    t = np.arange(pt)[:, None, None, None]
    patch = psi[t, rows, cols]                          # [pt, S, ix, iy]
    c = rdt(1) / np.sqrt(rdt(detx * dety))
    frame = np.zeros((pt, S, detx, dety), cdt)
    ox, oy = (detx - P) // 2, (dety - P) // 2
    frame[:, :, ox:ox + P, oy:oy + P] = (c * prb)[None, None] * patch
    frame[~valid] = 0
    g = sfft.fft2(frame, axes=(-2, -1)).astype(cdt)
    sx, sy = shifts_factor(scan, cdt)
    g = g * sx[:, :, None, None]
    g = g * sy[:, :, None, None]
    return g


def adj(g, prb, scan, nz, n, dtype=np.complex64):
    """ptychofft::adj: g [pt, S, detx, dety] -> psi [pt, nz, n], complex64.
    Overlap-add is accumulated in FP64 (the reference uses float atomics)."""
    cdt = np.dtype(dtype)
    rdt = np.float32 if cdt == np.complex64 else np.float64
    g = np.asarray(g, dtype=cdt)
    prb = np.asarray(prb, dtype=cdt)
    pt, S, detx, dety = g.shape
    P = prb.shape[0]
    sx, sy = shifts_factor(scan, cdt)
    g = g * np.conj(sx)[:, :, None, None]
    g = g * np.conj(sy)[:, :, None, None]
    g = (sfft.ifft2(g, axes=(-2, -1)) * (detx * dety)).astype(cdt)
    ox, oy = (detx - P) // 2, (dety - P) // 2
    c = rdt(1) / np.sqrt(rdt(detx * dety))
    contrib = (c * np.conj(prb))[None, None] * g[:, :, ox:ox + P, oy:oy + P]
    rows, cols, valid = _patch_index(scan, P, nz, n)
    contrib = contrib * valid[:, :, None, None]
    out = np.zeros((pt, nz * n), np.complex128)
    lin = (rows * n + cols).reshape(pt, -1)
    cf = contrib.reshape(pt, -1).astype(np.complex128)
    for t in range(pt):
        re = np.bincount(lin[t], weights=cf[t].real, minlength=nz * n)
        im = np.bincount(lin[t], weights=cf[t].imag, minlength=nz * n)
        out[t] = re + 1j * im
    return out.reshape(pt, nz, n).astype(cdt)


def _dft_matrix(D, sign):
    k = np.arange(D)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / D)


def fwd_exact(psi, prb, scan, detx, dety):
    """FP64 closed form with explicit DFT matrices (no FFT library)."""
    psi = np.asarray(psi, dtype=np.complex128)
    prb = np.asarray(prb, dtype=np.complex128)
    pt, nz, n = psi.shape
    P = prb.shape[0]
    S = scan.shape[2]
    rows, cols, valid = _patch_index(scan, P, nz, n)
    t = np.arange(pt)[:, None, None, None]
    patch = psi[t, rows, cols]
    frame = np.zeros((pt, S, detx, dety), np.complex128)
    ox, oy = (detx - P) // 2, (dety - P) // 2
    frame[:, :, ox:ox + P, oy:oy + P] = prb[None, None] * patch
    frame[~valid] = 0
    Fx = _dft_matrix(detx, -1)
    Fy = _dft_matrix(dety, -1)
    out = np.einsum('ab,tsbc,dc->tsad', Fx, frame, Fy) / np.sqrt(detx * dety)
    fr = np.asarray(scan, np.float32).astype(np.float64) - roundf(scan)
    ph = np.exp(2j * np.pi * (fr[0] + fr[1]))
    return out * ph[:, :, None, None]


def adj_exact(g, prb, scan, nz, n):
    """FP64 transpose of ``fwd_exact``."""
    g = np.asarray(g, dtype=np.complex128)
    prb = np.asarray(prb, dtype=np.complex128)
    pt, S, detx, dety = g.shape
    P = prb.shape[0]
    fr = np.asarray(scan, np.float32).astype(np.float64) - roundf(scan)
    ph = np.exp(-2j * np.pi * (fr[0] + fr[1]))
    g = g * ph[:, :, None, None]
    Fx = _dft_matrix(detx, 1)
    Fy = _dft_matrix(dety, 1)
    frame = np.einsum('ab,tsbc,dc->tsad', Fx, g, Fy) / np.sqrt(detx * dety)
    ox, oy = (detx - P) // 2, (dety - P) // 2
    contrib = np.conj(prb)[None, None] * frame[:, :, ox:ox + P, oy:oy + P]
    rows, cols, valid = _patch_index(scan, P, nz, n)
    out = np.zeros((pt, nz, n), np.complex128)
    for t in range(pt):
        for s in range(S):
            if valid[t, s]:
                np.add.at(out[t], (rows[t, s], cols[t, s]), contrib[t, s])
    return out
