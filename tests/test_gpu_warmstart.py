"""The trained warm start against plain ADMM on a held-out 128^3 phantom
(the committed weights, `ptychotomo_amd/weights/unet_warmstart_128.npz`): the
network improves the short run it is given, and ADMM continued from its output
reaches the plain run's converged object error in fewer iterations than the
plain run takes.  (The plain run's Lagrangian still oscillates after 100
iterations, so its 1e-3 Lagrangian criterion is reported, not required.)"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "ptychotomo_amd", "weights",
                       "unet_warmstart_128.npz")


def test_warm_start_gains_over_plain():
    from ptychotomo_amd import warmstart
    print("warm start 128^3", flush=True)
    net = warmstart.load(WEIGHTS)
    w = warmstart.workload(128, seed=10_128, photons=1e3)
    r = warmstart.compare(net, w, n_ref=100, n_max=100, tol=1e-3)
    oe = r["object_error"]
    print({k: oe[k] for k in ("plain_level", "plain_iters", "after_short", "network",
                              "warm_iters", "speedup")}, r["lagrangian"]["reached"])
    assert oe["network"] < oe["after_short"]        # the network improves its input
    assert oe["warm_iters"] is not None             # the warm run reaches the plain level
    assert oe["warm_iters"] < oe["plain_iters"]     # in fewer ADMM iterations
