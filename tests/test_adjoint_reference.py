"""CPU checks of the gradient reference used by the GPU adjoint tests.

The torch FP64 restatement (oracle/hb_autodiff.py) is pinned twice, with no
GPU: its forward loss equals the loss computed from the C oracle's FP64 twin,
and its autograd gradient equals central FP64 finite differences of that
oracle loss. A gradient with a dropped term (G's contribution halved) must
fail that finite-difference bar.
"""
import numpy as np
import pytest
import torch

from adjoint_cases import bold_mse_case, fd_grads, kuramoto_case, oracle_loss, rel_err
from hyburn_amd import _lib as L

CASES = {
    "kuramoto_order_rk4": lambda orc: kuramoto_case(),
    "kuramoto_order_euler": lambda orc: kuramoto_case(integ=L.EULER, seed=4),
    "g2do_bold_mse_heun": lambda orc: bold_mse_case(L.G2DO, orc, node_param_idx=2),
    "jr_bold_mse_rk4": lambda orc: bold_mse_case(L.JANSEN_RIT, orc, integ=L.RK4, G=0.5, node_param_idx=12),
}


@pytest.fixture(scope="module")
def ad():
    from oracle import hb_autodiff
    return hb_autodiff


@pytest.mark.parametrize("name", list(CASES))
def test_torch_forward_matches_oracle_fp64(orc, ad, name):
    I, kind, target = CASES[name](orc)
    lo = oracle_loss(orc, I, kind, target)
    G = torch.tensor(float(np.float32(I.cparams[0])), dtype=torch.float64)
    lt = float(ad.forward(I, G, None, None, kind, target))
    assert abs(lt - lo) <= 1e-8 * max(abs(lo), 1e-30), (lt, lo)


@pytest.mark.parametrize("name", list(CASES))
def test_torch_autograd_matches_fd(orc, ad, name):
    I, kind, target = CASES[name](orc)
    loss, gG, gP, _ = ad.gradients(I, kind, target)
    fG, fP = fd_grads(orc, I, kind, target)
    # central differences at h = 1e-4 carry an O(h^2) truncation error of ~1e-6
    assert rel_err(gG, fG) < 1e-5, (gG, fG)
    if fP is not None:
        assert rel_err(gP, fP) < 1e-5, (gP, fP)


def test_fd_detects_a_wrong_gradient(orc, ad):
    """The finite-difference check is sensitive: a gradient with a dropped
    coupling term (G's contribution halved) fails the bar above."""
    I, kind, target = kuramoto_case()
    _, gG, _, _ = ad.gradients(I, kind, target)
    fG, _ = fd_grads(orc, I, kind, target)
    assert rel_err(0.5 * gG, fG) > 1e-2
