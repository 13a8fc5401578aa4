"""CPU oracle for the ptycho-tomography ADMM hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker.

Contents
--------
usfft    FP32 line-by-line restatement of the reference USFFT Radon transform
         (``src/cuda/radonusfft.cu`` + ``src/include/kernels.cuh``) and an FP64
         direct non-uniform DFT closed form of the same operator.
ptycho   FP32 restatement of ``src/cuda/ptychofft.cu`` and an FP64 per-patch
         closed form.
solver   FP32 NumPy restatement of ``src/ptychotomo/solver.py`` (operator
         wrappers, CG solvers, line search, ADMM driver, Lagrangian) and an
         FP64 twin used to bound FP32 drift.

Parity pinning (what checks each function)
------------------------------------------
* ``ptycho.fwd/adj``: brute-force DFT-matrix evaluation at small sizes,
  delta-probe and constant-patch special cases, asymmetric probe for the
  transposed detector frame, exact adjointness.
* ``usfft.fwd``: FP64 direct NUDFT at the operator's own sample positions
  (gridding error ~1e-6), theta=0 and theta=pi/2 special cases that reduce to
  plain column / reversed-row sums.
* ``usfft.adj``: core-pair adjointness against ``usfft.fwd`` and FP64 direct
  adjoint NUDFT including the reference's edge-replicating boundary pad.
* ``solver`` pieces: textbook identities (finite-difference adjoint, 2,1-norm
  prox optimality, exact quadratic line-search model).
* ``solver.Solver.admm`` trajectory: pinned only by the FP64 twin (drift
  bound); the reference's own CuPy/cuFFT rounding cannot be reproduced here,
  so the trajectory is "parity unpinned" beyond that bound.
"""
