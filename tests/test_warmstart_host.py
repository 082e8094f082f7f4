"""Host-side pieces of the learned warm start (no GPU): the field encoding,
the input planes and the training-target layout.  No counterpart in the
reference; the fields are dsmcVolFields' rho_N, U and T_tr
(dsmcVolFields.C:1242-1290) scaled by the freestream."""
import math

import numpy as np

from hystrath_amd import warmstart as W


def _case():
    return W.make_cases(1, seed=11, grid=64)[0]


def test_encode_decode_round_trip_and_mask():
    c = _case()
    rng = np.random.default_rng(0)
    nc = c.ncells
    f = np.stack([c.n_inf * np.exp(rng.normal(0, 1, nc)), c.U_inf * rng.normal(0, 1, nc),
                  c.U_inf * rng.normal(0, 1, nc), c.T_inf * np.exp(rng.normal(0, 1, nc)), np.zeros(nc)])
    img = W.encode(f, c)
    back = W.decode(img, c)
    m = W.fluid_mask(c).reshape(-1) > 0
    assert (~m).sum() > 0  # the cylinder's cells
    np.testing.assert_allclose(back[:, m], f[:, m], rtol=2e-6, atol=1e-6 * c.U_inf)
    assert np.all(back[:, ~m] == 0.0) and np.all(img.reshape(-1, 5)[~m] == 0.0)
    # the freestream encodes as (0, 1, 0, 0, 0)
    fs = np.stack([np.full(nc, c.n_inf), np.full(nc, c.U_inf), np.zeros(nc), np.full(nc, c.T_inf), np.zeros(nc)])
    np.testing.assert_allclose(W.encode(fs, c).reshape(-1, 5)[m], np.tile([0, 1, 0, 0, 0], (m.sum(), 1)),
                               atol=1e-6)
    np.testing.assert_allclose(W.encode(fs, c), W.freestream_image(c), atol=1e-6)


def test_geometry_planes_measure_the_cylinder_and_the_domain():
    c = _case()
    g = W.geometry_planes(c)
    cfg = c.cfg
    m = W.fluid_mask(c) > 0
    # cell (i, j) centre, distance from the surface in radii / 10, by hand
    for (j, i) in [(0, 0), (31, 5), (40, 50), (63, 63)]:
        x = cfg.origin[0] + (i + 0.5) * cfg.delta[0]
        y = cfg.origin[1] + (j + 0.5) * cfg.delta[1]
        d = (math.hypot(x - cfg.cyl_center[0], y - cfg.cyl_center[1]) - cfg.cyl_radius) / cfg.cyl_radius / 10
        if m[j, i]:
            assert abs(g[j, i, 0] - min(max(d, 0.0), 1.0)) < 1e-6
            assert abs(g[j, i, 1] - (2 * (i + 0.5) / cfg.nx - 1)) < 1e-6
            assert abs(g[j, i, 2] - (2 * (j + 0.5) / cfg.ny - 1)) < 1e-6
    assert np.all(g[~m] == 0.0)
    assert g[..., 1:].min() > -1 and g[..., 1:].max() < 1


def test_net_input_layout_and_residual_targets():
    c = _case()
    img = W.freestream_image(c)
    x = W.net_input(img, c)
    m = W.fluid_mask(c) > 0
    assert x.shape == (64, 64, 32)
    np.testing.assert_array_equal(x[..., :5], img)
    np.testing.assert_allclose(x[m][:, 5:8], np.tile(W.conditions(c), (m.sum(), 1)))
    np.testing.assert_array_equal(x[..., 8:11], W.geometry_planes(c))
    assert np.all(x[..., 11:] == 0.0) and np.all(x[~m] == 0.0)
    # the denoiser's input: a prior and the noise level of an 800-step average
    prior = np.random.default_rng(1).normal(size=img.shape).astype(np.float32)
    xd = W.net_input(img, c, prior=prior, nsteps=800)
    np.testing.assert_array_equal(xd[..., 11:16][m], prior[m])
    np.testing.assert_allclose(xd[..., 16][m], 0.5)
    assert np.all(xd[..., 17:] == 0.0) and np.all(xd[~m] == 0.0)


def test_mirror_maps_the_domain_onto_itself():
    """The mesh and cylinder are symmetric about y = 0: a mirrored input
    image is the input image of the mirrored fields (Uy, the y plane and the
    prior's Uy change sign), so mirrored pairs are valid training pairs."""
    c = _case()
    rng = np.random.default_rng(3)
    m = W.fluid_mask(c)
    img = rng.normal(size=(64, 64, 5)).astype(np.float32) * m[..., None]
    prior = rng.normal(size=(64, 64, 5)).astype(np.float32) * m[..., None]
    np.testing.assert_array_equal(m[::-1], m)
    x = W.net_input(img, c, prior=prior, nsteps=400, with_conditions=False)
    xm = W.net_input(W.mirror(img, (2,)), c, prior=W.mirror(prior, (2,)), nsteps=400, with_conditions=False)
    np.testing.assert_allclose(W.mirror(x, W.MIRROR_SIGN_IN), xm, atol=1e-6)
    np.testing.assert_array_equal(W.mirror(W.mirror(x, W.MIRROR_SIGN_IN), W.MIRROR_SIGN_IN), x)
    # the denoiser's input has no condition planes
    assert np.all(x[..., 5:8] == 0.0)


def test_loss_weights_follow_the_density_with_mean_one():
    c = _case()
    m = W.fluid_mask(c)
    target = np.zeros((64, 64, 5), np.float32)
    target[..., 0] = np.log(np.linspace(0.01, 50.0, 64))[None, :]   # rho / rho_inf across x
    w = W.loss_weights(target, c)
    assert abs(w[m > 0].mean() - 1.0) < 1e-5 and np.all(w[m == 0] == 0.0)
    ratio = w[5, 63] / w[5, 10]
    want = min(50.0, 20.0) / max(np.linspace(0.01, 50.0, 64)[10], 0.05)
    assert abs(ratio / want - 1) < 1e-4
