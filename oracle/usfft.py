"""Oracle for the USFFT Radon transform (TEST INFRASTRUCTURE ONLY).

Two independent statements of the reference operator ``radonusfft``:

``fwd`` / ``adj``
    FP32 line-by-line restatement of the reference kernel chain
    (``src/cuda/radonusfft.cu:54-139`` driving ``src/include/kernels.cuh``).
    Every index map, pad, checkerboard and scale follows the reference; the
    FFTs are ``scipy.fft`` in complex64, the 15x15 Gaussian interpolation is a
    sparse matrix built once per geometry (float32 weights, float32 operands).

``fwd_exact`` / ``adj_exact``
    FP64 direct non-uniform DFT closed form of the same operator, evaluated at
    the operator's own float32 sample positions (``takexy`` incl. its clamp).
    The difference ``fwd - fwd_exact`` is the USFFT's own gridding error
    (eps = 1e-6 design accuracy, ``radonusfft.cu:10``).

Both are unnormalised, like the reference native class: the Python solver
multiplies by ``coeftomo`` (``solver.py:71,79``).
"""
import functools
import math

import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp

PI_REF = 3.1415926535  # kernels.cuh:1


def params(n):
    """Oversampled grid size N, Gaussian width mu and half-width M.

    radonusfft.cu:7-12: N = 3n/2, eps = 1e-6 (float), mu = -log(eps)/(2N^2)
    (stored as float), M = ceil(2N/pi*sqrt(-mu*log(eps) + (mu*N)^2/4)).
    """
    if n % 4:
        raise ValueError("n must be a multiple of 4 (N/6 embedding offset)")
    N = n * 3 // 2
    eps = float(np.float32(1e-6))
    mu = np.float32(-math.log(eps) / (2 * N * N))
    M = int(math.ceil(2 * N / PI_REF * math.sqrt(-float(mu) * math.log(eps)
                                                + (float(mu) * N) ** 2 / 4)))
    return N, mu, M


def sample_positions(theta, N):
    """Polar sample positions, kernels.cuh:84-95 (``takexy``).

    x = k*cos(theta)/N, y = -k*sin(theta)/N with k = i - N/2, clamped to
    0.5 - 1e-5 where >= 0.5.  Evaluated in FP64 and rounded to float32 (the
    reference uses fast-math ``__cosf``/``__sinf``, whose exact rounding is
    not reproduced).  Returns float32 arrays of shape [ntheta, N].
    """
    th = np.asarray(theta, dtype=np.float32).astype(np.float64)
    k = (np.arange(N) - N // 2).astype(np.float64)
    x = (k[None, :] / N * np.cos(th)[:, None]).astype(np.float32)
    y = (-k[None, :] / N * np.sin(th)[:, None]).astype(np.float32)
    lim = np.float32(0.5 - 1e-5)
    x[x >= np.float32(0.5)] = lim
    y[y >= np.float32(0.5)] = lim
    return x, y


def circ_mask(N):
    """kernels.cuh:160-172 with r = 1/N (radonusfft.cu:75): keep where
    4x^2 + 4y^2 < 1 - 1/N, x = (i - N/2)/N, evaluated in float32."""
    i = np.arange(N)
    x = (i - N // 2).astype(np.float32) / np.float32(N)
    four = np.float32(4)
    val = four * x[None, :] * x[None, :] + four * x[:, None] * x[:, None]
    return val < (np.float32(1) - np.float32(1) / np.float32(N))


def _checker(L):
    """Sign (1 - 2*((i+1) % 2)) of kernels.cuh:34 and :45."""
    i = np.arange(L)
    return (1 - 2 * ((i + 1) % 2)).astype(np.float32)


def _divphi_factor(N, mu, rdt=np.float32):
    """1/phi/(4N^2) of kernels.cuh:12-14 and :23-25, FP64 then float32."""
    a = (np.arange(N) - N // 2).astype(np.float64)
    phi = np.exp(-float(mu) * (a[None, :] ** 2 + a[:, None] ** 2))
    return (1.0 / phi / (4.0 * N * N)).astype(rdt)


@functools.lru_cache(maxsize=8)
def _interp_matrix_cached(theta_bytes, ntheta, n, rdt):
    theta = np.frombuffer(theta_bytes, dtype=np.float32)
    return _build_interp(theta, n, rdt)


def interp_matrix(theta, n, rdt=np.float32):
    """Sparse [ntheta*N, G*G] matrix of the gather (kernels.cuh:97-126).

    Row (ty*N + tx) holds the 15x15 weights
    w = pi/mu * exp(-pi^2/mu*(w0^2 + w1^2)) / N at grid column
    (N+M+ell1)*G + (N+M+ell0), ell = floor(2N*pos) - M + i, w = ell/(2N) - pos.
    Weights are evaluated in FP64 from float32 operands and rounded to float32.
    """
    theta = np.ascontiguousarray(theta, dtype=np.float32)
    return _interp_matrix_cached(theta.tobytes(), theta.size, n, np.dtype(rdt).name)


def _build_interp(theta, n, rdt="float32"):
    N, mu, M = params(n)
    G = 2 * N + 2 * M
    x, y = sample_positions(theta, N)
    x = x.ravel()
    y = y.ravel()
    ns = x.size
    twoN = np.float32(2 * N)
    fx = np.floor(twoN * x).astype(np.int64)
    fy = np.floor(twoN * y).astype(np.int64)
    i = np.arange(2 * M + 1)
    ell0 = fx[:, None] - M + i[None, :]          # [ns, 15]
    ell1 = fy[:, None] - M + i[None, :]
    w0 = (ell0.astype(np.float32) / twoN - x[:, None]).astype(np.float64)
    w1 = (ell1.astype(np.float32) / twoN - y[:, None]).astype(np.float64)
    c = PI_REF * PI_REF / float(mu)
    ex = np.exp(-c * w0 * w0)
    ey = np.exp(-c * w1 * w1)
    w = (PI_REF / abs(float(mu))) * ey[:, :, None] * ex[:, None, :] / N
    cols = (N + M + ell1)[:, :, None] * G + (N + M + ell0)[:, None, :]
    rows = np.broadcast_to(np.arange(ns)[:, None, None], cols.shape)
    mat = sp.csr_matrix((w.astype(rdt).ravel(),
                         (rows.ravel(), cols.ravel())), shape=(ns, G * G))
    return mat


def _pad2_index(N, n):
    o = N // 6
    return np.clip(np.arange(N) - o, 0, n - 1)


def fwd(u, theta, n=None, workers=None, dtype=np.complex64):
    """radonusfft::fwd (radonusfft.cu:54-92), unnormalised.

    u: [pz, n, n] complex64 -> [ntheta, pz, n] complex64.  ``dtype=complex128``
    gives the FP64 twin of the same chain.
    """
    cdt = np.dtype(dtype)
    rdt = np.float32 if cdt == np.complex64 else np.float64
    u = np.asarray(u, dtype=cdt)
    pz, n0, n1 = u.shape
    n = n0 if n is None else n
    assert n0 == n1 == n
    N, mu, M = params(n)
    G = 2 * N + 2 * M
    ntheta = len(theta)
    idx = _pad2_index(N, n)
    f = u[:, idx[:, None], idx[None, :]]                      # embed + pad2
    f = f * circ_mask(N)                                      # circ
    fde = np.zeros((pz, 2 * N, 2 * N), cdt)
    fde[:, N // 2:N // 2 + N, N // 2:N // 2 + N] = f * _divphi_factor(N, mu, rdt)
    s2 = _checker(2 * N)
    fde *= s2[None, :, None] * s2[None, None, :]              # fftshiftc
    F = sfft.fft2(fde, axes=(-2, -1), workers=workers)        # CUFFT_FORWARD
    fdee = np.zeros((pz, G, G), cdt)
    fdee[:, M:M + 2 * N, M:M + 2 * N] = F
    sG = _checker(G)
    fdee *= sG[None, :, None] * sG[None, None, :]             # fftshiftc
    j = (np.arange(G) - M + 2 * N) % (2 * N) + M              # wrap
    fdee = fdee[:, j[:, None], j[None, :]]
    W = interp_matrix(theta, n, rdt)
    g = (W @ fdee.reshape(pz, G * G).T).T                     # gather, [pz, ns]
    g = np.asarray(g, dtype=cdt).reshape(pz, ntheta, N).transpose(1, 0, 2)
    s1 = _checker(N)
    g = g * s1                                                # fftshift1c
    g = (sfft.ifft(g, axis=-1, workers=workers) * N).astype(cdt)  # CUFFT_INVERSE
    g = g * s1                                                # fftshift1c
    o = N // 6
    return np.ascontiguousarray(g[:, :, o:o + n])


def ramp_filter(N):
    """applyfilter (kernels.cuh:174-187) weights per detector frequency index
    i: rho = (i - N/2)/N, w = |rho| N 4 sin(rho)/rho, w = 0 at rho = 0, in
    float32 as there."""
    rho = ((np.arange(N) - N // 2) / np.float32(N)).astype(np.float32)
    w = np.zeros(N, np.float32)
    nz = rho != 0
    w[nz] = (np.abs(rho[nz]) * np.float32(N) * np.float32(4) * np.sin(rho[nz]) / rho[nz]).astype(np.float32)
    return w


def adj(g, theta, n=None, workers=None, dtype=np.complex64, filt=None):
    """radonusfft::adj (radonusfft.cu:94-139), unnormalised.

    g: [ntheta, pz, n] complex64 -> [pz, n, n] complex64.  Note the detector
    pad is edge-replicating (kernels.cuh:189-206), so this is not the exact
    transpose of ``fwd`` at the boundary.
    """
    cdt = np.dtype(dtype)
    rdt = np.float32 if cdt == np.complex64 else np.float64
    g = np.asarray(g, dtype=cdt)
    ntheta, pz, n0 = g.shape
    n = n0 if n is None else n
    N, mu, M = params(n)
    G = 2 * N + 2 * M
    o = N // 6
    gg = np.zeros((ntheta, pz, N), cdt)
    gg[:, :, o:o + n] = g
    gg[:, :, :o] = gg[:, :, o:o + 1]                          # pad
    gg[:, :, o + n:] = gg[:, :, o + n - 1:o + n]
    s1 = _checker(N)
    gg = gg * s1
    gg = sfft.fft(gg, axis=-1, workers=workers).astype(cdt)  # CUFFT_FORWARD
    gg = gg * s1
    if filt == "ramp":                                       # applyfilter
        gg = gg * ramp_filter(N)
    W = interp_matrix(theta, n, rdt)
    samples = gg.transpose(1, 0, 2).reshape(pz, ntheta * N)
    fdee = np.asarray((W.T @ samples.T).T, dtype=cdt).reshape(pz, G, G)  # scatter
    F = np.zeros((pz, 2 * N, 2 * N), cdt)                    # wrapadj
    t = (np.arange(G) - M + 2 * N) % (2 * N)
    rows = np.broadcast_to(t[:, None], (G, G)).ravel()
    cols = np.broadcast_to(t[None, :], (G, G)).ravel()
    lin = rows * (2 * N) + cols
    for z in range(pz):
        re = np.bincount(lin, weights=fdee[z].real.ravel().astype(np.float64), minlength=4 * N * N)
        im = np.bincount(lin, weights=fdee[z].imag.ravel().astype(np.float64), minlength=4 * N * N)
        F[z] = (re + 1j * im).reshape(2 * N, 2 * N)
    # fftshiftc on the G grid restricted to the interior [M, 2N+M)
    sG = _checker(G)[M:M + 2 * N]
    F *= sG[None, :, None] * sG[None, None, :]
    fde = (sfft.ifft2(F, axes=(-2, -1), workers=workers) * (4 * N * N)).astype(cdt)
    s2 = _checker(2 * N)
    fde *= s2[None, :, None] * s2[None, None, :]
    f = fde[:, N // 2:N // 2 + N, N // 2:N // 2 + N] * _divphi_factor(N, mu, rdt)  # unpaddivphi
    f = f * circ_mask(N)
    return np.ascontiguousarray(f[:, o:o + n, o:o + n])


# --------------------------------------------------------------------------
# FP64 closed forms
# --------------------------------------------------------------------------

def _ftilde(u, n):
    """circ(pad2(embed(u))) on the N x N grid, FP64."""
    N = n * 3 // 2
    idx = _pad2_index(N, n)
    f = np.asarray(u, dtype=np.complex128)[:, idx[:, None], idx[None, :]]
    return f * circ_mask(N)


def fwd_exact(u, theta, n=None):
    """FP64 direct NUDFT of the forward operator.

    R[th,z,j] = (-1)^(N/2)/N * sum_k e^{2 pi i k s/N}
                * sum_{p,q} ft[z,q,p] e^{-2 pi i (p X + q Y)}
    with s = j - n/2, k in [-N/2, N/2), (X, Y) the float32 sample positions,
    ft = circ(pad2(embed(u))), p = col - N/2, q = row - N/2.
    """
    u = np.asarray(u)
    pz = u.shape[0]
    n = u.shape[-1] if n is None else n
    N = n * 3 // 2
    ft = _ftilde(u, n)
    x, y = sample_positions(theta, N)
    X = x.astype(np.float64).ravel()
    Y = y.astype(np.float64).ravel()
    p = np.arange(N) - N // 2
    Ex = np.exp(-2j * np.pi * np.outer(p, X))                  # [p, s]
    Ey = np.exp(-2j * np.pi * np.outer(p, Y))                  # [q, s]
    T = (ft.reshape(pz * N, N) @ Ex).reshape(pz, N, -1)       # [z, q, s]
    Gs = np.einsum('zqs,qs->zs', T, Ey)                        # [z, s]
    ntheta = len(theta)
    Gs = Gs.reshape(pz, ntheta, N)
    k = np.arange(N) - N // 2
    s = np.arange(n) - n // 2
    E1 = np.exp(2j * np.pi * np.outer(k, s) / N)              # [k, j]
    sign = (-1.0) ** (N // 2)
    R = sign / N * np.einsum('ztk,kj->tzj', Gs, E1)
    return R


def adj_exact(g, theta, n=None):
    """FP64 direct adjoint NUDFT with the reference's boundary handling:
    edge-replicated detector pad (kernels.cuh:189-206), circ, crop."""
    g = np.asarray(g, dtype=np.complex128)
    ntheta, pz, n0 = g.shape
    n = n0 if n is None else n
    N = n * 3 // 2
    o = N // 6
    gg = np.zeros((ntheta, pz, N), np.complex128)
    gg[:, :, o:o + n] = g
    gg[:, :, :o] = gg[:, :, o:o + 1]
    gg[:, :, o + n:] = gg[:, :, o + n - 1:o + n]
    k = np.arange(N) - N // 2
    sN = np.arange(N) - N // 2
    E1 = np.exp(-2j * np.pi * np.outer(sN, k) / N)            # [s, k]
    sign = (-1.0) ** (N // 2)
    H = sign / N * np.einsum('tzs,sk->ztk', gg, E1).reshape(pz, ntheta * N)
    x, y = sample_positions(theta, N)
    X = x.astype(np.float64).ravel()
    Y = y.astype(np.float64).ravel()
    p = np.arange(N) - N // 2
    Ex = np.exp(2j * np.pi * np.outer(X, p))                   # [s, p]
    Ey = np.exp(2j * np.pi * np.outer(Y, p))                   # [s, q]
    # f[z,q,p] = sum_s H[z,s] Ey[s,q] Ex[s,p]
    f = np.einsum('zs,sq,sp->zqp', H, Ey, Ex, optimize=True)
    f = f * circ_mask(N)
    return f[:, o:o + n, o:o + n]


def core_fwd(f, theta, n):
    """Core pair, forward: N x N masked grid -> full polar lines [ntheta,pz,N]
    (no embed/pad2/crop).  Built from the same FP32 steps as ``fwd``."""
    f = np.asarray(f, dtype=np.complex64) * circ_mask(n * 3 // 2)
    N, mu, M = params(n)
    G = 2 * N + 2 * M
    pz = f.shape[0]
    fde = np.zeros((pz, 2 * N, 2 * N), np.complex64)
    fde[:, N // 2:N // 2 + N, N // 2:N // 2 + N] = f * _divphi_factor(N, mu)
    s2 = _checker(2 * N)
    fde *= s2[None, :, None] * s2[None, None, :]
    F = sfft.fft2(fde, axes=(-2, -1))
    fdee = np.zeros((pz, G, G), np.complex64)
    fdee[:, M:M + 2 * N, M:M + 2 * N] = F
    sG = _checker(G)
    fdee *= sG[None, :, None] * sG[None, None, :]
    j = (np.arange(G) - M + 2 * N) % (2 * N) + M
    fdee = fdee[:, j[:, None], j[None, :]]
    W = interp_matrix(theta, n)
    g = np.asarray((W @ fdee.reshape(pz, G * G).T).T, np.complex64)
    g = g.reshape(pz, len(theta), N).transpose(1, 0, 2)
    s1 = _checker(N)
    return ((sfft.ifft(g * s1, axis=-1) * N) * s1).astype(np.complex64)


def core_adj(g, theta, n):
    """Core pair, adjoint of ``core_fwd``: full polar lines -> N x N grid."""
    g = np.asarray(g, dtype=np.complex64)
    N, mu, M = params(n)
    G = 2 * N + 2 * M
    ntheta, pz, _ = g.shape
    s1 = _checker(N)
    gg = (sfft.fft(g * s1, axis=-1) * s1).astype(np.complex64)
    W = interp_matrix(theta, n)
    samples = gg.transpose(1, 0, 2).reshape(pz, ntheta * N)
    fdee = np.asarray((W.T @ samples.T).T, dtype=np.complex64).reshape(pz, G, G)
    t = (np.arange(G) - M + 2 * N) % (2 * N)
    F = np.zeros((pz, 2 * N, 2 * N), np.complex128)
    np.add.at(F, (slice(None), t[:, None], t[None, :]), fdee)
    sG = _checker(G)[M:M + 2 * N]
    F *= sG[None, :, None] * sG[None, None, :]
    fde = sfft.ifft2(F, axes=(-2, -1)) * (4 * N * N)
    s2 = _checker(2 * N)
    fde *= s2[None, :, None] * s2[None, None, :]
    f = fde[:, N // 2:N // 2 + N, N // 2:N // 2 + N] * _divphi_factor(N, mu)
    return (f * circ_mask(N)).astype(np.complex64)
