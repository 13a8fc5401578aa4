"""Writes the golden fixtures of the operator parity tests.

The reference ships no test vectors and cannot run here (CUDA + CuPy), so the
fixtures hold seeded inputs and the FP64 closed-form outputs of the oracle
(direct NUDFT for the USFFT Radon transform, explicit DFT matrices for the
ptychography operator), which the FP32 restatement and the device operators
must reproduce.  Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ptycho as optycho  # noqa: E402
from oracle import usfft as ousfft  # noqa: E402


def crandn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def main():
    rng = np.random.default_rng(20240601)
    # USFFT: n = 20 (N/2 odd -> global sign -1), 12 angles incl. 0
    n, pz, nt = 20, 2, 12
    theta = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    u = crandn(rng, (pz, n, n))
    g = crandn(rng, (nt, pz, n))
    np.savez_compressed(os.path.join(HERE, "usfft_n20.npz"), theta=theta, u=u, g=g,
                        fwd=ousfft.fwd_exact(u, theta).astype(np.complex128),
                        adj=ousfft.adj_exact(g, theta).astype(np.complex128))
    # ptychography: P = 5 < D = 8 (off-centre frame), asymmetric probe,
    # fractional positions (shift phases), one skipped patch
    pt, nz, nx, P, D, S = 2, 18, 21, 5, 8, 7
    psi = crandn(rng, (pt, nz, nx))
    prb = crandn(rng, (P, P))
    scan = np.stack([rng.uniform(0, nx - P, (pt, S)), rng.uniform(0, nz - P, (pt, S))]).astype(np.float32)
    scan[0, 0, 3] = -0.75
    far = crandn(rng, (pt, S, D, D))
    np.savez_compressed(os.path.join(HERE, "ptycho_p5_d8.npz"), psi=psi, prb=prb, scan=scan,
                        far=far, fwd=optycho.fwd_exact(psi, prb, scan, D, D),
                        adj=optycho.adj_exact(far, prb, scan, nz, nx))
    print("written")


if __name__ == "__main__":
    main()
