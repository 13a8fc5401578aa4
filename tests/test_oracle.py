"""CPU pinning of the oracle restatement against independent checks: FP64
closed forms, special cases with hand-derivable answers, exact adjointness
and textbook identities.  Each check is chosen so that a plausible mistake
(transposed frame, wrong sign, dropped term, wrong rounding) fails it."""
import numpy as np
import pytest

from oracle import ptycho as optycho
from oracle import solver as osolver
from oracle import usfft as ousfft


def crandn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ------------------------------------------------------------------ ptychography
def small_ptycho(rng, pt=2, nz=14, n=17, P=4, D=8, S=6):
    psi = crandn(rng, (pt, nz, n))
    prb = crandn(rng, (P, P))
    scan = np.stack([rng.uniform(0, n - P, (pt, S)), rng.uniform(0, nz - P, (pt, S))]).astype(np.float32)
    return psi, prb, scan


def test_ptycho_fwd_matches_dft_brute_force():
    rng = np.random.default_rng(1)
    psi, prb, scan = small_ptycho(rng)
    a = optycho.fwd(psi, prb, scan, 8, 8)
    b = optycho.fwd_exact(psi, prb, scan, 8, 8)
    assert rel(a, b) < 1e-6


def test_ptycho_adj_matches_transpose_and_is_adjoint():
    rng = np.random.default_rng(2)
    psi, prb, scan = small_ptycho(rng)
    g = crandn(rng, (2, 6, 8, 8))
    assert rel(optycho.adj(g, prb, scan, 14, 17), optycho.adj_exact(g, prb, scan, 14, 17)) < 1e-6
    lhs = np.vdot(optycho.fwd(psi, prb, scan, 8, 8).astype(np.complex128), g)
    rhs = np.vdot(psi.astype(np.complex128), optycho.adj(g, prb, scan, 14, 17))
    assert abs(lhs - rhs) / abs(lhs) < 1e-6


def test_ptycho_single_pixel_pins_transposed_frame():
    """psi = delta at (row r0, col c0), probe = ones, integer scan position:
    the frame is nonzero at [ix = c0 - stx][iy = r0 - sty] only
    (kernels.cuh:251-263), so the far field is c*exp(-2 pi i (kx ix + ky iy)/D)."""
    P = D = 8
    nz, n = 20, 24
    psi = np.zeros((1, nz, n), np.complex64)
    r0, c0 = 9, 13
    psi[0, r0, c0] = 1.0
    prb = np.ones((P, P), np.complex64)
    stx, sty = 10, 5
    scan = np.array([[[stx]], [[sty]]], np.float32)
    g = optycho.fwd(psi, prb, scan, D, D)[0, 0]
    ix, iy = c0 - stx, r0 - sty
    k = np.arange(D)
    expect = np.exp(-2j * np.pi * (k[:, None] * ix + k[None, :] * iy) / D) / D
    assert np.abs(g - expect).max() < 1e-6
    # the transposed convention would put the phase ramps on the other axes
    wrong = np.exp(-2j * np.pi * (k[:, None] * iy + k[None, :] * ix) / D) / D
    assert np.abs(g - wrong).max() > 0.1


def test_ptycho_constant_patch_is_dc_spike_and_shift_phase():
    P = D = 8
    psi = np.full((1, 16, 16), 2.0 + 1.0j, np.complex64)
    prb = np.ones((P, P), np.complex64)
    scan = np.array([[[3.25]], [[4.0]]], np.float32)  # frac x = 0.25 -> phase i
    g = optycho.fwd(psi, prb, scan, D, D)[0, 0]
    expect = np.zeros((D, D), np.complex128)
    expect[0, 0] = (2.0 + 1.0j) * D * np.exp(2j * np.pi * 0.25)  # c*D^2*value*phase
    assert np.abs(g - expect).max() < 1e-5


def test_roundf_half_away_from_zero():
    x = np.array([0.5, 1.5, 2.5, -0.5, -1.5, 0.49999997], np.float32)
    assert list(optycho.roundf(x)) == [1, 2, 3, -1, -2, 0]


def test_ptycho_negative_position_skipped():
    rng = np.random.default_rng(3)
    psi, prb, scan = small_ptycho(rng)
    scan[0, 0, 0] = -0.6  # rounds to -1: the reference kernel returns early
    g = optycho.fwd(psi, prb, scan, 8, 8)
    assert np.all(g[0, 0] == 0)


# ------------------------------------------------------------------ USFFT
@pytest.mark.parametrize("n", [16, 20, 32])
def test_usfft_fwd_matches_direct_nudft(n):
    rng = np.random.default_rng(n)
    theta = np.linspace(0, np.pi, 10, endpoint=False).astype(np.float32)
    u = crandn(rng, (2, n, n))
    e = rel(ousfft.fwd(u, theta), ousfft.fwd_exact(u, theta))
    assert e < 2e-6, e


@pytest.mark.parametrize("n", [16, 20])
def test_usfft_theta0_is_signed_column_sum(n):
    """theta = 0: R = (-1)^(N/2) * sum over rows of circ(pad2(embed(u)));
    n = 20 has N/2 odd and pins the global sign."""
    rng = np.random.default_rng(7)
    u = crandn(rng, (2, n, n))
    N = 3 * n // 2
    ft = ousfft._ftilde(u, n)
    o = N // 6
    col = ft.sum(axis=1)[:, o:o + n]
    sign = (-1.0) ** (N // 2)
    R = ousfft.fwd(u, np.zeros(1, np.float32))[0]
    assert rel(R, sign * col) < 2e-6
    assert rel(R, -sign * col) > 1.0


def test_usfft_theta_half_pi_is_reversed_row_sum():
    """theta = pi/2: R[sigma] = sum over columns of row -sigma (orientation)."""
    n = 32
    rng = np.random.default_rng(8)
    u = crandn(rng, (1, n, n))
    N = 3 * n // 2
    ft = ousfft._ftilde(u, n)
    sig = np.arange(n) - n // 2
    rows = ft.sum(axis=2)[:, -sig + N // 2]
    R = ousfft.fwd(u, np.array([np.pi / 2], np.float32))[0]
    # float32(pi/2) leaves cos ~ -4e-8, a sub-1e-4 phase; the wrong orientation is O(1) off
    assert rel(R, rows) < 2e-4
    assert rel(R, ft.sum(axis=2)[:, sig + N // 2]) > 0.5


def test_usfft_core_pair_adjoint():
    n = 24
    rng = np.random.default_rng(9)
    theta = np.linspace(0, np.pi, 12, endpoint=False).astype(np.float32)
    N = 3 * n // 2
    f = crandn(rng, (2, N, N))
    g = crandn(rng, (12, 2, N))
    lhs = np.vdot(ousfft.core_fwd(f, theta, n).astype(np.complex128), g)
    rhs = np.vdot(f.astype(np.complex128), ousfft.core_adj(g, theta, n))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_usfft_adj_matches_direct_adjoint_with_pad():
    n = 20
    rng = np.random.default_rng(10)
    theta = np.linspace(0, np.pi, 8, endpoint=False).astype(np.float32)
    g = crandn(rng, (8, 2, n))
    assert rel(ousfft.adj(g, theta), ousfft.adj_exact(g, theta)) < 2e-6


def test_usfft_boundary_pair_not_adjoint():
    """The reference pads are not transposes of each other (pad2 vs pad):
    the full pair is adjoint only for zero borders."""
    n = 16
    rng = np.random.default_rng(11)
    theta = np.linspace(0, np.pi, 8, endpoint=False).astype(np.float32)
    u = crandn(rng, (1, n, n))
    g = crandn(rng, (8, 1, n))
    lhs = np.vdot(ousfft.fwd(u, theta).astype(np.complex128), g)
    rhs = np.vdot(u.astype(np.complex128), ousfft.adj(g, theta))
    assert abs(lhs - rhs) / abs(lhs) > 1e-3
    u[:, 0, :] = u[:, -1, :] = u[:, :, 0] = u[:, :, -1] = 0
    g[:, :, 0] = g[:, :, -1] = 0
    lhs = np.vdot(ousfft.fwd(u, theta).astype(np.complex128), g)
    rhs = np.vdot(u.astype(np.complex128), ousfft.adj(g, theta))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_usfft_params_reference_values():
    # radonusfft.cu:7-12 gives M = 7 for every configuration size
    for n in [64, 96, 128, 192, 256, 512]:
        N, mu, M = ousfft.params(n)
        assert N == 3 * n // 2 and M == 7
        assert abs(float(mu) - (-np.log(1e-6) / (2 * N * N))) < 1e-7 * float(mu) + 1e-12


# ------------------------------------------------------------------ solver pieces
def tiny_solver(dtype=np.complex64):
    rng = np.random.default_rng(0)
    n, nz, nt, P = 16, 16, 8, 8
    theta = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    prb = crandn(rng, (P, P))
    scan = np.stack([rng.uniform(0, n - P, (nt, 9)), rng.uniform(0, nz - P, (nt, 9))]).astype(np.float32)
    return osolver.Solver(prb, scan, theta, [8, 8], 1e-6, 8.8, nt, nz, n, 4, 8, dtype=dtype), rng


def test_reg_pair_adjoint_on_ADMM_subspace():
    s, rng = tiny_solver(np.complex128)
    u = crandn(rng, (16, 16, 16)).astype(np.complex128)
    g = crandn(rng, (3, 16, 16, 16)).astype(np.complex128)
    g[0, :, :, -1] = 0
    g[1, :, -1, :] = 0
    g[2, -1, :, :] = 0
    lhs = np.vdot(s.fwd_reg(u), g)
    rhs = np.vdot(u, s.adj_reg(g))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_solve_reg_is_21_prox():
    """z* = argmin_z a||z||_{2,1} + 1/2||z - v||^2: per voxel, z* = 0 if
    |v| <= a else v (1 - a/|v|); check optimality by perturbation."""
    s, rng = tiny_solver(np.complex128)
    u = crandn(rng, (16, 16, 16)).astype(np.complex128)
    mu = crandn(rng, (3, 16, 16, 16)).astype(np.complex128)
    tau, alpha = 0.7, 1.3
    z = s.solve_reg(u, mu, tau, alpha)
    v = s.fwd_reg(u) + mu / tau
    a = alpha / tau

    def obj(zz):
        return a * np.sum(np.sqrt(np.sum(np.abs(zz) ** 2, 0))) + 0.5 * np.sum(np.abs(zz - v) ** 2)
    f0 = obj(z)
    for _ in range(5):
        dz = 1e-3 * crandn(rng, z.shape)
        assert obj(z + dz) >= f0 - 1e-9
    za = np.sqrt(np.sum(np.abs(v) ** 2, 0))
    assert np.all(z[:, za <= a] == 0)


def test_cg_tomo_functional_is_quadratic_in_gamma():
    s, rng = tiny_solver(np.complex128)
    xi0 = crandn(rng, (8, 16, 16)).astype(np.complex128)
    xi1 = crandn(rng, (3, 16, 16, 16)).astype(np.complex128)
    K = crandn(rng, (8, 16, 16)).astype(np.complex128)
    u = crandn(rng, (16, 16, 16)).astype(np.complex128)
    d = crandn(rng, (16, 16, 16)).astype(np.complex128)
    rho, tau = 0.5, 0.25

    def minf(g):
        return rho * np.linalg.norm(K * s.fwd_tomo_batch(u + g * d) - xi0) ** 2 + \
            tau * np.linalg.norm(s.fwd_reg(u + g * d) - xi1) ** 2
    gs = np.array([0.0, 0.5, 1.0, 2.0, 3.0])
    fs = np.array([minf(g) for g in gs])
    c = np.polyfit(gs[:3], fs[:3], 2)
    assert np.allclose(np.polyval(c, gs[3:]), fs[3:], rtol=1e-8)


def test_line_search_steps_are_halvings():
    s, rng = tiny_solver()
    calls = []

    def minf(u, fu):
        calls.append(u)
        return (u - 0.3) ** 2
    g = s.line_search(minf, 2.0, 0.0, 0.0, 1.0, 0.0)
    assert g in [2.0 * 0.5 ** j for j in range(45)]
    assert (0.0 - 0.3) ** 2 - (g - 0.3) ** 2 >= 0
    # no descent direction: the rule ends at gamma = 0
    assert s.line_search(lambda u, fu: u, 2.0, 0.0, 0.0, 1.0, 0.0) == 0


def test_lagrangian_sum_omits_tau_term():
    """take_lagr's Sum is sum(lagr[0:5]) (solver.py:304): lagr[5] excluded."""
    s, rng = tiny_solver()
    sh = (8, 16, 16)
    psi = np.exp(1j * 0.1 * rng.standard_normal(sh)).astype(np.complex64)
    data = np.abs(rng.standard_normal((8, 9, 8, 8))).astype(np.float32)
    phi = crandn(rng, (3, 16, 16, 16))
    e = crandn(rng, (3, 16, 16, 16))
    lagr = s.take_lagr(psi, phi, data, psi, e, crandn(rng, sh), crandn(rng, phi.shape), 1e-3,
                       0.5, 0.5, 'poisson')
    assert abs(lagr[6] - np.sum(lagr[0:5], dtype=np.float64)) <= 1e-5 * abs(lagr[6])
    assert lagr[5] > 0


def test_usfft_ramp_filter_inverts_radon():
    """applyfilter's weight ~ 4N|k|/N makes the filtered adjoint a filtered
    backprojection: FBP(R u) = (4 N ntheta / pi) u up to the filter's own
    error (sin(rho)/rho taper, removed DC) for a smooth object.  The
    unfiltered adjoint is a blurred u (error > 0.5), so a filter applied in
    the wrong domain or with a wrong weight fails one of the two bounds."""
    n, ntheta = 64, 96
    N = 3 * n // 2
    theta = np.linspace(0, np.pi, ntheta, endpoint=False).astype(np.float32)
    y, x = np.mgrid[:n, :n] - n / 2 + 0.5
    u = np.exp(-(x ** 2 + y ** 2) / (2 * (n / 10) ** 2)).astype(np.complex64)[None]
    g = ousfft.fwd(u, theta)
    for filt, lo, hi in (("ramp", 0.0, 0.12), (None, 0.5, 10.0)):
        f = ousfft.adj(g, theta, filt=filt)
        c = np.vdot(u, f).real / np.vdot(u, u).real
        err = np.linalg.norm(f / c - u) / np.linalg.norm(u)
        assert lo <= err < hi, (filt, err)
        if filt:
            assert abs(c / (4 * N * ntheta / np.pi) - 1) < 0.05, c


def test_usfft_ramp_filter_weights():
    w = ousfft.ramp_filter(12)
    assert w[6] == 0                                   # rho = 0
    assert np.allclose(w, w[::-1][np.r_[11, 0:11]])    # even in rho about i = N/2
    rho = (np.arange(12) - 6) / 12
    m = rho != 0
    assert np.allclose(w[m], 4 * 12 * np.abs(np.sin(rho[m])), rtol=1e-6)
