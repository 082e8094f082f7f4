"""The learned warm start end to end on small cases (GPU): training data
from plain DSMC through hs_dsmc_evolve_batch (averages of 200 to 1,600
steps cut from one run per case), the cross-fitted surrogates and the
denoiser trained on the bfloat16 MFMA kernels (their losses fall), the nets
saved and reloaded unchanged, and the plain and warm paths timed against a
long_avg on a held-out case.  How much faster the warm path is at full size is
measured by bench.py; here the pipeline runs and its outputs are well
formed."""
import numpy as np
import pytest

from hystrath_amd import warmstart as W

pytestmark = pytest.mark.gpu


def test_pipeline_trains_saves_and_times_both_paths(tmp_path):
    import torch
    torch.zeros(1, device="cuda")
    cs = W.make_cases(4, seed=123, grid=64)
    train, test = cs[:3], cs[3:]
    gen = W.generate(cs, develop=400, average=1600, window=200, log=lambda *a: None)
    # 1,600 sampled steps: 200-step averages at every other boundary (4), 400 (4), 800 (2), 1,600 (1)
    assert [sorted(n for n, _ in w) for w, _ in gen] == [[200] * 4 + [400] * 4 + [800] * 2 + [1600]] * 4
    surrogate, priors, hs = W.train_surrogates(train, gen[:3], epochs=40, batch=4, lr=2e-3, log=lambda *a: None)
    assert hs[-1] < 0.5 * hs[0], hs
    assert len(priors) == 3 and all(p.shape == (64, 64, 5) for p in priors)
    X, Y, M = W.denoiser_pairs(train, gen[:3], priors)
    assert tuple(X.shape) == (66, 64, 64, 32) and tuple(Y.shape) == (66, 64, 64, 16)  # and their mirror images
    denoiser, hd = W.train(X, Y, M, epochs=6, batch=11, lr=2e-3, log=lambda *a: None, seed=1,
                           layers=W.ws.DENOISER_LAYERS)
    assert hd[-1] < hd[0], hd
    path = str(tmp_path / "nets.npz")
    W.save_nets((surrogate, denoiser), path)
    s2, d2 = W.load_nets(path)
    img = W.freestream_image(test[0])
    np.testing.assert_array_equal(W.predict(surrogate, img, test[0]), W.predict(s2, img, test[0]))
    p0 = W.predict(s2, img, test[0])
    np.testing.assert_array_equal(W.predict(denoiser, img, test[0], prior=p0, nsteps=400),
                                  W.predict(d2, img, test[0], prior=p0, nsteps=400))
    long_avg = W.long_average_fields(test[0], gen[3][1])
    plain = W.plain_time_to_tol(test[0], long_avg, develop=400, check=200, max_sampled=600)
    warm = W.warm_time_to_tol((s2, d2), test[0], long_avg, relax=100, check=200, max_sampled=600)
    for r in (plain, warm):
        assert r["steps"] > 0 and set(r["errors"]) >= {"rho", "Ux", "Uy", "Ttra"}
    assert warm["inference_s"] > 0
    assert warm["used"] in ("denoised running average", "plain running average", "none within the tolerance")
