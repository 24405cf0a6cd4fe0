"""The learned warm start (GPU): trained briefly on projector-generated scans by the
network's own kernels, it must move held-out scans much closer to their true geometry
than the nominal start, never raise the reprojection residual (the plain start is the
fallback), and let gradient calibration reach a lower error in the same number of steps."""

import pytest
import torch

from cbct_backprojector import calibration, warmstart

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def trained():
    fam = warmstart.ScanFamily()
    data = warmstart.TrainingData(fam, pool=8, seed=1000, device=DEV)
    net = warmstart.WarmStartNet(seed=0, device=DEV)
    hist = warmstart.train(net, data, 300, P=2, lr=1e-3, lr_final=1e-5)
    return fam, net, hist


def _err(p):
    return calibration.reprojection_error_px(p.geometry, p.true_geometry, tuple(p.volume.shape), p.voxel_size,
                                             p.pixel_size)


def test_training_lowers_the_reprojection_loss(trained):
    _, _, hist = trained
    first, last = sum(hist[:20]) / 20, sum(hist[-20:]) / 20
    print(f"training loss (mean squared detector displacement, px^2): {first:.3f} -> {last:.3f}")
    assert last < 0.25 * first


@pytest.mark.parametrize("seed", [6001, 6002])
def test_warm_start_gains_over_the_plain_start(trained, seed):
    fam, net, _ = trained
    p = calibration.synthetic_problem(fam.spec(seed), DEV)
    g_nominal = p.geometry.clone()
    e_nominal = _err(p)
    residuals = warmstart.warm_start(p, net, fam, iterations=3)
    e_warm = _err(p)
    assert all(b < a for a, b in zip(residuals, residuals[1:]))      # only improving updates are kept
    g_warm = p.geometry.clone()
    # ten Adam steps from each start
    calibration.calibrate([p], steps=10)
    e_warm10 = _err(p)
    p.geometry.copy_(g_nominal)
    p.adam_m.zero_()
    p.adam_v.zero_()
    calibration.calibrate([p], steps=10)
    e_plain10 = _err(p)
    print(f"scan {seed}: nominal {e_nominal:.3f} px, warm start {e_warm:.3f} px; after 10 Adam steps: "
          f"plain {e_plain10:.3f} px, warm-started {e_warm10:.3f} px")
    assert e_warm < 0.5 * e_nominal
    assert e_warm10 < 0.7 * e_plain10
    assert not torch.equal(g_warm, g_nominal)
