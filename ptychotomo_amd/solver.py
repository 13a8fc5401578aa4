"""Drop-in ``Solver`` for joint ptycho-tomography by ADMM on MI355X.

Mirrors the reference class ``Solver`` (src/ptychotomo/solver.py:19-336):
same constructor, method names, argument meaning, shapes, dtypes and
normalisation.  Arrays are PyTorch tensors on the GPU (complex64 / float32);
NumPy inputs are copied to the device.  Every computation runs in the HIP
library through its C ABI; PyTorch only holds the buffers.

Implementation notes (mathematically identical to the reference, different
in rounding only):
* ``fwd_tomo_batch`` / ``adj_tomo_batch`` and ``cg_ptycho_batch`` process the
  whole volume in one native call; the angle chunking of ``ptheta`` is kept
  where it changes the mathematics (per-chunk CG step and Dai-Yuan beta,
  solver.py:245-254), as segmented reductions.
* ``cg_tomo`` updates R u by linearity (R(u + g d) = R u + g R d) instead of
  a fresh transform per iteration.
* The line searches evaluate up to 7 halving candidates per sweep in one
  fused pass (``cg_ptycho``: far field and rho term; ``cg_tomo``: K R and J
  terms).  Each candidate is formed as the reference forms it -- x + g*dx
  element-wise in FP32 -- so a step below FP32 resolution reproduces the
  gamma = 0 value exactly and the halving rule stops where the reference's
  does instead of collapsing to gamma = 0.
* Scalars (norms, dots, steps) are accumulated in FP64.
"""
import ctypes
import warnings

import numpy as np
import torch

from . import _lib
from ._dev import c64, device, empty, f32, f64, ptr, stream, to_dev, zeros
from .objects import wavenumber as _wavenumber
from .ptychofft import ptychofft
from .radonusfft import radonusfft

PLANCK_CONSTANT = 6.58211928e-19  # [keV*s]   solver.py:15
SPEED_OF_LIGHT = 299792458e+2     # [cm/s]    solver.py:16

_MAXGAM = 8


def _call(name, *args):
    return _lib.call(name, *args)


def _dai_yuan(gnorm2, den):
    """Dai-Yuan beta = ||g||^2 / sum(conj(d)(g - g0)) (solver.py:195-196,
    230-231, complex).  After a zero step (the halving rule found no descent)
    g == g0 and the reference divides 0 by 0; here beta = 0 instead, i.e. a
    steepest-descent restart, so the iterate stays finite."""
    den = np.asarray(den, np.complex128)
    ok = (den != 0) & np.isfinite(den)
    beta = np.zeros(den.shape, np.complex128)
    beta[ok] = np.asarray(gnorm2, np.float64)[ok] / den[ok]
    beta[~np.isfinite(beta)] = 0
    return beta


class Solver(object):
    """solver.py:19-46."""

    def __init__(self, prb, scan, theta, det, voxelsize, energy, ntheta, nz, n, ptheta, pnz):
        self.voxelsize = voxelsize
        self.energy = energy
        self.scan = to_dev(scan, f32)
        self.prb = to_dev(prb, c64)
        self.theta = to_dev(theta, f32)
        self.ntheta = int(ntheta)
        self.nz = int(nz)
        self.n = int(n)
        self.nscan = int(self.scan.shape[2])
        self.ndety = int(det[0])
        self.ndetx = int(det[1])
        self.nprb = int(self.prb.shape[0])
        self.ptheta = int(ptheta)
        self.pnz = int(pnz)
        if self.ntheta % self.ptheta:
            raise ValueError("ntheta must be a multiple of ptheta")
        if self.nz % self.pnz:
            raise ValueError("nz must be a multiple of pnz")
        self.nchunk = self.ntheta // self.ptheta
        # chunk-sized ptychography operator (solver.py:35-36; det[0] -> detx)
        self.cl_ptycho = ptychofft(self.ptheta, self.nz, self.n, self.ptheta, self.nscan,
                                   self.ndety, self.ndetx, self.nprb)
        # slab-sized tomography operator (solver.py:39-40)
        self.cl_tomo = radonusfft(self.ntheta, self.pnz, self.n)
        self.cl_tomo.setobj(self.theta.data_ptr())
        # whole-volume operators used by the batch methods and the solvers
        if self.pnz == self.nz:
            self._tomo = self.cl_tomo
        else:
            self._tomo = radonusfft(self.ntheta, self.nz, self.n)
            self._tomo.setobj(self.theta.data_ptr())
        self._ptycho = ptychofft(self.ntheta, self.nz, self.n, self.ntheta, self.nscan,
                                 self.ndety, self.ndetx, self.nprb)
        self._ptycho.setobj(self.scan.data_ptr(), self.prb.data_ptr())
        # normalisation coefficients (solver.py:43-46)
        prb_h = self.prb.cpu().numpy()
        self.coeftomo = np.float32(1 / np.sqrt(self.ntheta * self.n / 2))
        self.coefptycho = np.float32(1) / np.abs(prb_h).max().astype(np.float32)
        self.coefdata = np.float32(1) / (np.float32(self.ndetx * self.ndety)
                                         * (np.abs(prb_h) ** 2).max().astype(np.float32))
        self._sc = torch.zeros(64, dtype=f64, device=device())  # scalar results
        self._data_cache = None
        self.history = []
        self.state = None

    # ------------------------------------------------------------------ helpers
    def _scalars(self, k):
        """first k FP64 results (synchronises)."""
        return self._sc[:k].cpu().numpy()

    def _dev_data(self, data):
        """data [ntheta, nscan, ndety, ndetx] float32 on the device (cached)."""
        if isinstance(data, torch.Tensor) and data.is_cuda and data.dtype == f32 \
                and data.is_contiguous():
            return data
        key = id(data)
        if self._data_cache is not None and self._data_cache[0] == key:
            return self._data_cache[1]
        d = to_dev(data, f32)
        self._data_cache = (key, d, data)
        return d

    # ------------------------------------------------------------------ element-wise
    def mlog(self, psi):
        """solver.py:48-52."""
        psi = to_dev(psi, c64)
        res = empty(psi.shape)
        _call("pt_k_mlog", ptr(res), ptr(psi), psi.numel(), stream())
        return res

    def wavenumber(self):
        """solver.py:55-56."""
        return 2 * np.pi / (2 * np.pi * PLANCK_CONSTANT * SPEED_OF_LIGHT / self.energy)

    def exptomo(self, psi):
        """solver.py:59-60: exp(i*psi*voxelsize*k/coeftomo)."""
        psi = to_dev(psi, c64)
        res = empty(psi.shape)
        c = self.voxelsize * self.wavenumber() / float(self.coeftomo)
        _call("pt_k_exptomo", ptr(res), ptr(psi), c, 0.0, psi.numel(), stream())
        return res

    def logtomo(self, psi):
        """solver.py:63-64: -i/k * mlog(psi)/voxelsize*coeftomo."""
        m = self.mlog(psi)
        a = -float(self.coeftomo) / (self.wavenumber() * self.voxelsize)
        _call("pt_k_lincomb3", ptr(m), 0.0, a, ptr(m), 0.0, 0.0, None, 0.0, 0.0, None,
              m.numel(), stream())
        return m

    # ------------------------------------------------------------------ tomography
    def fwd_tomo(self, u):
        """solver.py:67-72 (one pnz slab)."""
        u = to_dev(u, c64)
        res = empty((self.ntheta, self.pnz, self.n))
        self.cl_tomo.fwd_scaled(res.data_ptr(), u.data_ptr(), float(self.coeftomo))
        return res

    def adj_tomo(self, data):
        """solver.py:75-80 (one pnz slab)."""
        data = to_dev(data, c64)
        res = empty((self.pnz, self.n, self.n))
        self.cl_tomo.adj_scaled(res.data_ptr(), data.data_ptr(), float(self.coeftomo))
        return res

    def fwd_tomo_batch(self, u):
        """solver.py:83-89 (whole volume in one call; pnz is memory-only)."""
        u = to_dev(u, c64)
        res = empty((self.ntheta, self.nz, self.n))
        self._tomo.fwd_scaled(res.data_ptr(), u.data_ptr(), float(self.coeftomo))
        return res

    def adj_tomo_batch(self, data, scale=1.0):
        """solver.py:92-97; ``scale`` multiplies the result (fused)."""
        data = to_dev(data, c64)
        res = empty((self.nz, self.n, self.n))
        self._tomo.adj_scaled(res.data_ptr(), data.data_ptr(), float(self.coeftomo) * scale)
        return res

    def fbp_tomo_batch(self, data):
        """Filtered backprojection of projections scaled like fwd_tomo_batch:
        the adjoint with the reference's ramp filter (applyfilter,
        kernels.cuh:174-187; commented out at radonusfft.cu:121), normalised so
        that fbp_tomo_batch(fwd_tomo_batch(u)) approximates u.  The filter
        weight is about 4N|k|/N per detector frequency k, N = 3n/2, so the
        filtered backprojection over ntheta angles in [0, pi) is
        4 N ntheta / pi times the inverse Radon transform; coeftomo enters
        once from the data and once from the adjoint."""
        data = to_dev(data, c64)
        res = empty((self.nz, self.n, self.n))
        N = 3 * self.n // 2
        scale = float(np.pi / (4.0 * N * self.ntheta * float(self.coeftomo)))
        self._tomo.setfilter("ramp")
        try:
            self._tomo.adj_scaled(res.data_ptr(), data.data_ptr(), scale)
        finally:
            self._tomo.setfilter(None)
        return res

    # ------------------------------------------------------------------ ptychography
    def fwd_ptycho(self, psi):
        """solver.py:100-105 (chunk set by cl_ptycho.setobj)."""
        psi = to_dev(psi, c64)
        res = empty((self.ptheta, self.nscan, self.ndety, self.ndetx))
        self.cl_ptycho.fwd_scaled(res.data_ptr(), psi.data_ptr(), float(self.coefptycho))
        return res

    def fwd_ptycho_batch(self, psi):
        """solver.py:108-117: |F psi|^2/coefdata for all angles, returned as a
        host float32 array like the reference."""
        return self.fwd_ptycho_batch_dev(psi).cpu().numpy()

    def fwd_ptycho_batch_dev(self, psi):
        psi = to_dev(psi, c64)
        far = empty((self.ntheta, self.nscan, self.ndety, self.ndetx))
        self._ptycho.fwd_scaled(far.data_ptr(), psi.data_ptr(), float(self.coefptycho))
        data = empty(far.shape, f32)
        _call("pt_k_intensity", ptr(data), ptr(far), float(1.0 / self.coefdata), far.numel(),
              stream())
        return data

    def adj_ptycho(self, data):
        """solver.py:120-125 (chunk set by cl_ptycho.setobj)."""
        data = to_dev(data, c64)
        res = empty((self.ptheta, self.nz, self.n))
        self.cl_ptycho.adj_scaled(res.data_ptr(), data.data_ptr(), float(self.coefptycho))
        return res

    # ------------------------------------------------------------------ regulariser
    def fwd_reg(self, u):
        """solver.py:128-134."""
        u = to_dev(u, c64)
        res = empty((3, *u.shape))
        _call("pt_k_reg_fwd", ptr(res), ptr(u), *u.shape, stream())
        return res

    def adj_reg(self, gr):
        """solver.py:137-146."""
        gr = to_dev(gr, c64)
        res = empty(gr.shape[1:])
        _call("pt_k_reg_adj", ptr(res), ptr(gr), *gr.shape[1:], stream())
        return res

    def solve_reg(self, u, mu, tau, alpha):
        """solver.py:257-264."""
        u = to_dev(u, c64)
        mu = to_dev(mu, c64)
        res = empty(mu.shape)
        _call("pt_k_reg_solve", ptr(res), ptr(u), ptr(mu), float(tau), float(alpha), *u.shape,
              stream())
        return res

    # ------------------------------------------------------------------ takexi
    def takexi(self, psi, phi, lamd, mu, rho, tau):
        """solver.py:149-168 (r = nprb // 2)."""
        psi = to_dev(psi, c64)
        lamd = to_dev(lamd, c64)
        phi = to_dev(phi, c64)
        mu = to_dev(mu, c64)
        r = self.nprb // 2
        _call("pt_k_takexi_stats", ptr(self._sc), ptr(psi), ptr(lamd), float(rho), self.ntheta,
              self.nz, self.n, r, stream())
        s1, s2, tmax = self._scalars(3)
        cnt = self.ntheta * self.nz * r
        pshift = float(np.float32((np.float32(s1 / cnt) + np.float32(s2 / cnt)) / 2))
        kc = self.voxelsize * self.wavenumber() / float(self.coeftomo)
        kscale = 1.0 / (kc * tmax) if tmax > 0 else 0.0
        xi0 = empty(psi.shape)
        K = empty(psi.shape)
        _call("pt_k_takexi", ptr(xi0), ptr(K), ptr(psi), ptr(lamd), float(rho), pshift, kc,
              kscale, self.voxelsize * self.wavenumber(), float(self.coeftomo), psi.numel(),
              stream())
        xi1 = empty(phi.shape)
        _call("pt_k_lincomb3", ptr(xi1), 1.0, 0.0, ptr(phi), -1.0 / tau, 0.0, ptr(mu), 0.0, 0.0,
              None, phi.numel(), stream())
        return xi0, xi1, K, pshift

    # ------------------------------------------------------------------ line search
    def line_search(self, minf, gamma, u, fu, d, fd):
        """solver.py:171-177 (generic form, callable minf)."""
        while (minf(u, fu) - minf(u + gamma * d, fu + gamma * fd) < 0 and gamma > 1e-12):
            gamma *= 0.5
        if gamma <= 1e-12:
            gamma = 0
        return gamma

    @staticmethod
    def _halving_sweeps(gamma, evalf):
        """Halving rule of line_search (solver.py:171-177) for a vector of
        independent segments.  ``evalf(cand)`` returns the functional at the
        candidate steps cand[s, j] (column 0 is gamma = 0); each sweep tries
        the current step and its next 6 halvings for every segment."""
        nseg = len(gamma)
        cur = np.array(gamma, dtype=np.float64)
        out = np.zeros(nseg)
        done = np.zeros(nseg, bool)
        while not done.all():
            cand = np.zeros((nseg, _MAXGAM))
            cand[:, 1:] = cur[:, None] * 0.5 ** np.arange(_MAXGAM - 1)[None, :]
            f = evalf(cand)
            for s in range(nseg):
                if done[s]:
                    continue
                for j in range(1, _MAXGAM):
                    g = cand[s, j]
                    if (f[s, 0] - f[s, j] < 0) and g > 1e-12:
                        continue
                    out[s] = 0.0 if g <= 1e-12 else g
                    done[s] = True
                    break
                else:
                    cur[s] = cand[s, -1] * 0.5
        return out

    # ------------------------------------------------------------------ cg_tomo
    def cg_tomo(self, xi0, xi1, K, init, rho, tau, titer):
        """solver.py:180-203."""
        u, _ = self._cg_tomo(xi0, xi1, K, init, rho, tau, titer)
        return u

    def _cg_tomo(self, xi0, xi1, K, init, rho, tau, titer, Ru=None):
        xi0, xi1, K = to_dev(xi0, c64), to_dev(xi1, c64), to_dev(K, c64)
        u = to_dev(init, c64).clone()
        nz, n = self.nz, self.n
        if Ru is None:
            Ru = self.fwd_tomo_batch(u)
        else:
            Ru = Ru.clone()
        lenp = Ru.numel()
        lenu = u.numel()
        sc = self._sc
        st = stream()
        r_t = empty(Ru.shape)
        d = empty(u.shape)
        grad0 = None
        gamma = 2.0
        gammas = []
        beta_dev = torch.zeros(2, dtype=f64, device=device())
        gam_dev = torch.zeros(_MAXGAM, dtype=f64, device=device())
        for i in range(titer):
            # grad = rho*R*(conj(K)(K Ru - xi0)) + tau*J*(J u - xi1)
            _call("pt_k_tomo_resid", ptr(r_t), ptr(sc[0:]), ptr(K), ptr(Ru), ptr(xi0), lenp, st)
            grad = self.adj_tomo_batch(r_t, scale=rho)
            _call("pt_k_reg_grad", ptr(grad), ptr(sc[1:]), ptr(u), ptr(xi1), float(tau), nz, n,
                  n, st)
            if i == 0:
                _call("pt_k_dy_update", ptr(d), ptr(grad), None, 1, lenu, 1, st)
            else:
                _call("pt_k_dy_reduce", ptr(sc[2:]), ptr(grad), ptr(grad0), ptr(d), lenu, 1, st)
                gn, dr, di = self._scalars(5)[2:5]
                beta = complex(_dai_yuan(np.array([gn]), np.array([complex(dr, di)]))[0])
                beta_dev.copy_(torch.tensor([beta.real, beta.imag], dtype=f64))
                _call("pt_k_dy_update", ptr(d), ptr(grad), ptr(beta_dev), 0, lenu, 1, st)
            grad0 = grad
            Rd = self.fwd_tomo_batch(d)

            def evalf(cand):
                gam_dev.copy_(torch.from_numpy(cand[0]))
                _call("pt_k_tomo_minf", ptr(sc[16:]), ptr(K), ptr(Ru), ptr(Rd), ptr(xi0),
                      ptr(gam_dev), _MAXGAM, lenp, st)
                _call("pt_k_reg_minf", ptr(sc[24:]), ptr(u), ptr(d), ptr(xi1), ptr(gam_dev),
                      _MAXGAM, nz, n, n, st)
                v = self._scalars(32)
                return (rho * v[16:24] + tau * v[24:32])[None, :]

            gamma = float(self._halving_sweeps([gamma], evalf)[0])
            gammas.append(gamma)
            _call("pt_k_lincomb3", ptr(u), 1.0, 0.0, ptr(u), gamma, 0.0, ptr(d), 0.0, 0.0, None,
                  lenu, st)
            _call("pt_k_lincomb3", ptr(Ru), 1.0, 0.0, ptr(Ru), gamma, 0.0, ptr(Rd), 0.0, 0.0,
                  None, lenp, st)
        self._last_tomo_gammas = gammas
        return u, Ru

    # ------------------------------------------------------------------ cg_ptycho
    def cg_ptycho(self, data, init, h, lamd, rho, piter, model):
        """solver.py:206-242 on the chunk set by ``cl_ptycho.setobj`` (data
        already multiplied by coefdata, as the reference passes it)."""
        data = to_dev(data, f32)
        psi = self._cg_ptycho(self.cl_ptycho, 1, data, init, h, lamd, rho, piter, model,
                              coefdata=1.0)
        return psi

    def cg_ptycho_batch(self, data, init, h, lamd, rho, piter, model):
        """solver.py:245-254: all angle chunks in one fused pass with
        per-chunk step lengths and Dai-Yuan coefficients."""
        d = self._dev_data(data)
        return self._cg_ptycho(self._ptycho, self.nchunk, d, init, h, lamd, rho, piter, model,
                               coefdata=float(self.coefdata))

    def _cg_ptycho(self, op, nseg, data, init, h, lamd, rho, piter, model, coefdata):
        mdl = _lib.PT_MODEL[model]
        psi = to_dev(init, c64).clone()
        h = to_dev(h, c64)
        lamd = to_dev(lamd, c64)
        st = stream()
        L = psi.numel()
        sc = torch.zeros(8 * nseg + 8, dtype=f64, device=device())
        sums = torch.zeros(nseg * _MAXGAM, dtype=f64, device=device())
        rsums = torch.zeros(nseg * _MAXGAM, dtype=f64, device=device())
        gam_dev = torch.zeros(nseg * _MAXGAM, dtype=f64, device=device())
        beta_dev = torch.zeros(2 * nseg, dtype=f64, device=device())
        step_dev = torch.zeros(2 * nseg, dtype=f64, device=device())
        grad = empty(psi.shape)
        grad0 = empty(psi.shape)
        d = empty(psi.shape)
        gamma = np.full(nseg, 2.0)
        chunk = op.Ntheta // nseg
        cp, cd = float(self.coefptycho), float(coefdata)
        gam_hist = [[] for _ in range(nseg)]
        for i in range(piter):
            _call("pt_ptycho_grad", op._h, ptr(grad), ptr(psi), ptr(data), ptr(h), ptr(lamd),
                  float(rho), cp, cd, mdl, st)
            if i == 0:
                _call("pt_k_dy_update", ptr(d), ptr(grad), None, 1, L, nseg, st)
            else:
                _call("pt_k_dy_reduce", ptr(sc), ptr(grad), ptr(grad0), ptr(d), L, nseg, st)
                v = sc[:3 * nseg].cpu().numpy().reshape(nseg, 3)
                beta = _dai_yuan(v[:, 0], v[:, 1] + 1j * v[:, 2])
                beta_dev.copy_(torch.from_numpy(np.stack([beta.real, beta.imag], 1).ravel()))
                _call("pt_k_dy_update", ptr(d), ptr(grad), ptr(beta_dev), 0, L, nseg, st)
            grad, grad0 = grad0, grad

            def evalf(cand):
                gam_dev.copy_(torch.from_numpy(cand.ravel()))
                _call("pt_ptycho_minf", op._h, ptr(sums), ptr(psi), ptr(d), ptr(data),
                      ptr(gam_dev), _MAXGAM, chunk, cp, cd, mdl, 0, st)
                _call("pt_k_rho_minf", ptr(rsums), ptr(h), ptr(psi), ptr(lamd), ptr(d),
                      float(rho), ptr(gam_dev), _MAXGAM, L, nseg, st)
                dat = sums.cpu().numpy().reshape(nseg, _MAXGAM)
                rr = rsums.cpu().numpy().reshape(nseg, _MAXGAM)
                return dat + rho * rr

            gamma = self._halving_sweeps(gamma, evalf)
            for s in range(nseg):
                gam_hist[s].append(float(gamma[s]))
            step_dev.copy_(torch.from_numpy(np.stack([gamma, np.zeros(nseg)], 1).ravel()))
            _call("pt_k_axpy_seg", ptr(psi), ptr(step_dev), ptr(d), L, nseg, st)
        _call("pt_k_max_angle", ptr(sc), ptr(psi), L, st)
        mx = float(sc[0].item())
        if mx > 3.14:
            print('possible phase wrap, max computed angle', mx)
        self._chunk_gammas = gam_hist
        return psi

    # ------------------------------------------------------------------ penalties / Lagrangian
    def update_penalty(self, psi, h, h0, phi, e, e0, rho, tau):
        """solver.py:267-282."""
        st = stream()
        _call("pt_k_norm2_lin", ptr(self._sc[0:]), ptr(psi), ptr(h), 1.0, 0.0, -1.0, 0.0,
              psi.numel(), 1, st)
        _call("pt_k_norm2_lin", ptr(self._sc[1:]), ptr(h), ptr(h0), rho, 0.0, -rho, 0.0,
              h.numel(), 1, st)
        _call("pt_k_norm2_lin", ptr(self._sc[2:]), ptr(phi), ptr(e), 1.0, 0.0, -1.0, 0.0,
              phi.numel(), 1, st)
        _call("pt_k_norm2_lin", ptr(self._sc[3:]), ptr(e), ptr(e0), tau, 0.0, -tau, 0.0,
              e.numel(), 1, st)
        r, s, r2, s2 = self._scalars(4)
        if r > 10 * s:
            rho *= 2
        elif s > 10 * r:
            rho *= 0.5
        if r2 > 10 * s2:
            tau *= 2
        elif s2 > 10 * r2:
            tau *= 0.5
        return rho, tau

    def take_lagr(self, psi, phi, data, h, e, lamd, mu, alpha, rho, tau, model):
        """solver.py:285-305 (lagr[6] = sum(lagr[0:5]) as in the reference)."""
        st = stream()
        dd = self._dev_data(data)
        lagr = np.zeros(7, dtype=np.float32)
        sums = torch.zeros(self.nchunk, dtype=f64, device=device())
        gz = torch.zeros(self.nchunk, dtype=f64, device=device())
        _call("pt_ptycho_minf", self._ptycho._h, ptr(sums), ptr(psi), None, ptr(dd), ptr(gz), 1,
              self.ptheta, float(self.coefptycho), float(self.coefdata), _lib.PT_MODEL[model],
              1, st)
        plane = self.nz * self.n * self.n
        _call("pt_k_tv_norm", ptr(self._sc[0:]), ptr(phi), plane, st)
        _call("pt_k_dot_diff", ptr(self._sc[1:]), ptr(lamd), ptr(h), ptr(psi), h.numel(), st)
        _call("pt_k_norm2_lin", ptr(self._sc[3:]), ptr(h), ptr(psi), 1.0, 0.0, -1.0, 0.0,
              h.numel(), 1, st)
        _call("pt_k_dot_diff", ptr(self._sc[4:]), ptr(mu), ptr(e), ptr(phi), e.numel(), st)
        _call("pt_k_norm2_lin", ptr(self._sc[6:]), ptr(e), ptr(phi), 1.0, 0.0, -1.0, 0.0,
              e.numel(), 1, st)
        v = self._scalars(7)
        lagr[0] = sums.sum().item()
        lagr[1] = alpha * v[0]
        lagr[2] = 2 * v[1]
        lagr[3] = rho * v[3]
        lagr[4] = 2 * v[4]
        lagr[5] = tau * v[6]
        lagr[6] = np.sum(lagr[0:5])
        return lagr

    # ------------------------------------------------------------------ ADMM
    def admm(self, data, h, e, psi, phi, lamd, mu, u, alpha, piter, titer, niter, model,
             lagr_every=10, verbose=True, rho=0.5, tau=0.5, m0=0, Ru=None, callback=None):
        """solver.py:308-336.  Returns (u, psi) like the reference.

        Beyond the reference: the full final state (h, e, psi, phi, lamd, mu,
        u, rho, tau, Ru) is kept in ``self.state``; ``rho``, ``tau``, ``m0``
        (index of the first iteration, which sets the piter+32 schedule of
        the first two iterations and the Lagrangian cadence) and ``Ru``
        (R u of the initial u) resume a run exactly where ``self.state`` left
        it.  ``callback(m, solver_locals)`` is called after every iteration.
        """
        data = self._dev_data(data)
        h, e, psi = to_dev(h, c64), to_dev(e, c64), to_dev(psi, c64)
        phi, lamd, mu, u = to_dev(phi, c64), to_dev(lamd, c64), to_dev(mu, c64), to_dev(u, c64)
        st = stream()
        self.history = []
        for m in range(m0, m0 + niter):
            h0, e0 = h, e
            psi = self.cg_ptycho_batch(data, psi, h, lamd, rho, piter + (m < 2) * 32, model)
            xi0, xi1, K, pshift = self.takexi(psi, phi, lamd, mu, rho, tau)
            u, Ru = self._cg_tomo(xi0, xi1, K, u, rho, tau, titer, Ru=Ru)
            phi = self.solve_reg(u, mu, tau, alpha)
            # h = exptomo(fwd_tomo_batch(u)) * exp(i pshift), R u from the CG recursion
            h = empty(psi.shape)
            c = self.voxelsize * self.wavenumber() / float(self.coeftomo)
            _call("pt_k_exptomo", ptr(h), ptr(Ru), c, float(pshift), h.numel(), st)
            e = self.fwd_reg(u)
            lamd_n = empty(lamd.shape)
            _call("pt_k_lincomb3", ptr(lamd_n), 1.0, 0.0, ptr(lamd), rho, 0.0, ptr(h), -rho, 0.0,
                  ptr(psi), lamd.numel(), st)
            mu_n = empty(mu.shape)
            _call("pt_k_lincomb3", ptr(mu_n), 1.0, 0.0, ptr(mu), tau, 0.0, ptr(e), -tau, 0.0,
                  ptr(phi), mu.numel(), st)
            lamd, mu = lamd_n, mu_n
            rec = {"m": m, "rho_used": rho, "tau_used": tau, "pshift": float(pshift),
                   "ptycho_gammas": self._chunk_gammas, "tomo_gammas": self._last_tomo_gammas}
            rho, tau = self.update_penalty(psi, h, h0, phi, e, e0, rho, tau)
            rec["rho"], rec["tau"] = rho, tau
            if lagr_every and np.mod(m, lagr_every) == 0:
                lagr = self.take_lagr(psi, phi, data, h, e, lamd, mu, alpha, rho, tau, model)
                rec["lagr"] = [float(v) for v in lagr]
                if verbose:
                    print("%d/%d) rho=%.2e, tau=%.2e, Lagrangian terms:  %.2e %.2e %.2e %.2e %.2e %.2e, Sum: %.2e" %
                          (m, niter, rho, tau, *lagr))
            self.history.append(rec)
            if callback is not None:
                callback(m, dict(h=h, e=e, psi=psi, phi=phi, lamd=lamd, mu=mu, u=u, rho=rho,
                                 tau=tau, Ru=Ru))
        self.state = dict(h=h, e=e, psi=psi, phi=phi, lamd=lamd, mu=mu, u=u, rho=rho, tau=tau,
                          Ru=Ru, m=m0 + niter)
        return u, psi

    def admm_resume(self, data, alpha, piter, titer, niter, model, **kw):
        """Continue from ``self.state`` (checkpoint/resume of the full state)."""
        s = self.state
        return self.admm(data, s["h"], s["e"], s["psi"], s["phi"], s["lamd"], s["mu"], s["u"],
                         alpha, piter, titer, niter, model, rho=s["rho"], tau=s["tau"],
                         m0=s["m"], Ru=s["Ru"], **kw)
