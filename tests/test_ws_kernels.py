"""The warm-start network's bfloat16 MFMA kernels against float32 PyTorch.

Every product in the kernels is bfloat16 x bfloat16 accumulated in FP32, so
they are compared with float32 PyTorch convolutions of the same
bfloat16-rounded operands: the only difference left is the order of the FP32
sums, a relative 1e-4 at most for these sizes (the bound in the tests).
Results stored in bfloat16 are compared within one bfloat16 rounding (2^-8
relative).  The MFMA lane layouts the kernels rely on are checked first.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    import torch
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)


def _round(a):
    import torch
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).float().numpy()


def test_mfma_32x32x16_lane_layout():
    from hystrath_amd import ws
    rng = np.random.default_rng(0)
    A = _round(rng.normal(size=(32, 16)))
    B = _round(rng.normal(size=(16, 32)))
    a = np.zeros((64, 8), np.float32)
    b = np.zeros((64, 8), np.float32)
    for l in range(64):
        for i in range(8):
            a[l, i] = A[l % 32, 8 * (l // 32) + i]
            b[l, i] = B[8 * (l // 32) + i, l % 32]
    d = ws.mfma_tile(0, _bits(a), _bits(b)).reshape(64, 16)
    D = A.astype(np.float64) @ B.astype(np.float64)
    got = np.zeros((32, 32))
    for l in range(64):
        for v in range(16):
            got[8 * (v // 4) + 4 * (l // 32) + v % 4, l % 32] = d[l, v]
    np.testing.assert_allclose(got, D, rtol=1e-5, atol=1e-5)


def test_mfma_16x16x32_lane_layout():
    from hystrath_amd import ws
    rng = np.random.default_rng(1)
    A = _round(rng.normal(size=(16, 32)))
    B = _round(rng.normal(size=(32, 16)))
    a = np.zeros((64, 8), np.float32)
    b = np.zeros((64, 8), np.float32)
    for l in range(64):
        for i in range(8):
            a[l, i] = A[l % 16, 8 * (l // 16) + i]
            b[l, i] = B[8 * (l // 16) + i, l % 16]
    d = ws.mfma_tile(1, _bits(a), _bits(b)).reshape(64, 4)
    got = np.zeros((16, 16))
    for l in range(64):
        for v in range(4):
            got[4 * (l // 16) + v, l % 16] = d[l, v]
    np.testing.assert_allclose(got, A.astype(np.float64) @ B.astype(np.float64), rtol=1e-5, atol=1e-5)


def _case(cin, cout, n=2, h=24, w=64, seed=0):
    import torch
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(n, h, w, cin, generator=g).to(torch.bfloat16)
    wm = torch.randn(9, cin, cout, generator=g) * (2.0 / (9 * cin)) ** 0.5
    b = torch.randn(cout, generator=g) * 0.1
    return x, wm, b


def _torch_conv(x_nhwc, wm, b=None, dil=1):
    """float32 conv2d (dilation dil, padding dil) of NHWC input with [9][cin][cout] weights."""
    import torch
    cin, cout = wm.shape[1], wm.shape[2]
    wt = wm.reshape(3, 3, cin, cout).permute(3, 2, 0, 1).contiguous()  # OIHW
    y = torch.nn.functional.conv2d(x_nhwc.permute(0, 3, 1, 2).float(), wt.float(), b, padding=dil, dilation=dil)
    return y.permute(0, 2, 3, 1).contiguous()


CONV_SHAPES = [(16, 32, 1), (32, 64, 1), (64, 64, 1), (64, 32, 1), (32, 16, 1), (16, 32, 4), (64, 64, 16), (32, 16, 2)]


@pytest.mark.parametrize("cin,cout,dil", CONV_SHAPES)
def test_conv_forward_matches_float32(cin, cout, dil):
    import torch
    from hystrath_amd import ws
    x, wm, b = _case(cin, cout)
    n, h, w, _ = x.shape
    dev = "cuda"
    wf, _ = ws.pack(wm.to(dev), cin, cout)
    wq = wm.to(torch.bfloat16).float()
    want = _torch_conv(x, wq, b, dil)
    y, y32 = ws.conv3x3(x.to(dev).view(torch.int16), wf, b.to(dev), n, h, w, cin, cout, out_f32=True, dilation=dil)
    got32 = y32.cpu()
    scale = want.abs().max().item()
    assert (got32 - want).abs().max().item() < 1e-4 * scale
    got = ws.from_bf16_bits(y.cpu())
    np.testing.assert_allclose(got.numpy(), want.to(torch.bfloat16).float().numpy(), rtol=2 ** -7, atol=1e-3 * scale)
    # ReLU, and the mask of the input-gradient pass
    yr, _ = ws.conv3x3(x.to(dev).view(torch.int16), wf, b.to(dev), n, h, w, cin, cout, relu=True, dilation=dil)
    np.testing.assert_array_equal(ws.from_bf16_bits(yr.cpu()).numpy(), np.maximum(got.numpy(), 0.0))
    mask = torch.randn(n, h, w, cout).to(torch.bfloat16)
    ym, _ = ws.conv3x3(x.to(dev).view(torch.int16), wf, b.to(dev), n, h, w, cin, cout,
                       mask_bits=mask.to(dev).view(torch.int16), dilation=dil)
    np.testing.assert_array_equal(ws.from_bf16_bits(ym.cpu()).numpy(),
                                  np.where(mask.float().numpy() > 0, got.numpy(), 0.0))


@pytest.mark.parametrize("cin,cout,dil", CONV_SHAPES)
def test_input_and_weight_gradients_match_autograd(cin, cout, dil):
    import torch
    from hystrath_amd import ws
    x, wm, b = _case(cin, cout, seed=3)
    n, h, w, _ = x.shape
    dev = "cuda"
    wq = wm.to(torch.bfloat16).float()
    dy = torch.randn(n, h, w, cout).to(torch.bfloat16)
    xf = x.float().requires_grad_(True)
    wt = wq.clone().requires_grad_(True)
    bt = b.clone().requires_grad_(True)
    y = _torch_conv(xf, wt, bt, dil)
    y.backward(dy.float())
    _, wb = ws.pack(wm.to(dev), cin, cout)
    _, dx32 = ws.conv3x3(dy.to(dev).view(torch.int16), wb, None, n, h, w, cout, cin, out_bf16=False, out_f32=True,
                         dilation=dil)
    gx = xf.grad
    assert (dx32.cpu() - gx).abs().max().item() < 1e-4 * gx.abs().max().item()
    dw = torch.zeros(9, cin, cout, device=dev)
    db = torch.zeros(cout, device=dev)
    ws.wgrad(x.to(dev).view(torch.int16), dy.to(dev).view(torch.int16), dw, db, n, h, w, cin, cout, dilation=dil)
    gw = wt.grad
    assert (dw.cpu() - gw).abs().max().item() < 1e-4 * gw.abs().max().item()
    assert (db.cpu() - bt.grad).abs().max().item() < 1e-4 * bt.grad.abs().max().item()


def test_network_forward_backward_and_training():
    """The six-layer network's forward pass and every layer's weight and bias
    gradients against float32 PyTorch that rounds to bfloat16 where the
    kernels store it; then Adam on a small fixed problem at least halves the
    loss."""
    import torch
    from hystrath_amd import ws
    dev = "cuda"
    net = ws.ConvNet(dev, seed=1)
    n, h, w = 2, 32, 64
    g = torch.Generator().manual_seed(5)
    x = torch.zeros(n, h, w, ws.C_IN)
    x[..., :ws.N_IN] = torch.randn(n, h, w, ws.N_IN, generator=g)
    xb = x.to(torch.bfloat16)
    out, acts = net.forward(xb.to(dev).view(torch.int16), keep=True)
    a = xb.float()
    for k, (ti, to, si, so, dl) in enumerate(ws.LAYERS):
        z = _torch_conv(a, net.w[k].cpu().to(torch.bfloat16).float(), net.b[k].cpu(), dl)
        if k < len(ws.LAYERS) - 1:
            a = torch.relu(z).to(torch.bfloat16).float()
        else:
            a = z
    # an FP32 sum order that differs can move an intermediate activation across
    # a bfloat16 rounding boundary (2^-8 relative); a few such flips, not more
    err, top = (out.cpu() - a).abs().max().item(), a.abs().max().item()
    assert err < 1e-2 * top, (err, top)
    # the masked MSE and its gradient; a target the network can represent: a
    # pointwise function of the input
    target = torch.zeros(n, h, w, ws.C_OUT)
    target[..., :ws.N_FIELDS] = 0.3 * torch.tanh(xb.float()[..., 0:5] + xb.float()[..., 1:6])
    mask = (torch.rand(n, h, w, generator=g) > 0.2).float()
    loss, dy = ws.mse(out, target.to(dev), mask.to(dev))
    d = (out.cpu() - target)[..., :ws.N_FIELDS]
    want = (mask[..., None] * d * d).sum() / (ws.N_FIELDS * (mask != 0).sum())
    assert abs(loss.item() / want.item() - 1) < 1e-4
    grads = net.backward(acts, dy)
    # the same backward pass in float32 PyTorch, rounding where the kernels
    # store bfloat16: the layer inputs (the forward activations) and the
    # gradient handed from layer to layer.  What is left is the order of the
    # FP32 sums (and the odd bfloat16 rounding it flips).
    from torch.nn.grad import conv2d_input, conv2d_weight
    d = ws.from_bf16_bits(dy.cpu()).permute(0, 3, 1, 2)
    for k in range(len(ws.LAYERS) - 1, -1, -1):
        ti, to, si, so, dl = ws.LAYERS[k]
        a_k = ws.from_bf16_bits(acts[k].cpu()).permute(0, 3, 1, 2)
        w_k = net.w[k].cpu().to(torch.bfloat16).float().reshape(3, 3, si, so).permute(3, 2, 0, 1).contiguous()
        gref = conv2d_weight(a_k, w_k.shape, d, padding=dl, dilation=dl).permute(2, 3, 1, 0).reshape(9, si, so)
        gw = grads[k].cpu()
        rel = ((gw - gref).norm() / (gref.norm() + 1e-30)).item()
        assert rel < 1e-2, (k, rel)
        gb, bref = grads[len(ws.LAYERS) + k].cpu(), d.sum(dim=(0, 2, 3))
        assert ((gb - bref).norm() / (bref.norm() + 1e-30)).item() < 2e-3, k
        if k > 0:
            dx = conv2d_input(a_k.shape, w_k, d, padding=dl, dilation=dl) * (a_k > 0)
            d = dx.to(torch.bfloat16).float()
    # training on a fixed small problem
    losses = []
    for it in range(300):
        out, acts = net.forward(xb.to(dev).view(torch.int16), keep=True)
        loss, dy = ws.mse(out, target.to(dev), mask.to(dev))
        losses.append(loss.item())
        net.adam(net.backward(acts, dy), lr=3e-3)
    assert losses[-1] < losses[0] / 2, (losses[0], losses[-1])


def test_adam_update():
    import torch
    from hystrath_amd import ws
    dev = "cuda"
    g = torch.Generator().manual_seed(2)
    p0 = torch.randn(1000, generator=g)
    p = p0.clone().to(dev)
    m = torch.zeros(1000, device=dev)
    v = torch.zeros(1000, device=dev)
    grads = [torch.randn(1000, generator=g) for _ in range(3)]
    mr, vr, pr = np.zeros(1000), np.zeros(1000), p0.numpy().astype(np.float64)
    for t, gr in enumerate(grads, 1):
        ws._ck(ws.lib().hs_ws_adam(ws._p(p), ws._p(gr.to(dev)), ws._p(m), ws._p(v), 1000, 1e-2, 0.9, 0.999, 1e-8,
                                   t, ws._stream()), "adam")
        gg = gr.numpy().astype(np.float64)
        mr = 0.9 * mr + 0.1 * gg
        vr = 0.999 * vr + 0.001 * gg * gg
        pr = pr - 1e-2 * (mr / (1 - 0.9 ** t)) / (np.sqrt(vr / (1 - 0.999 ** t)) + 1e-8)
    np.testing.assert_allclose(p.cpu().numpy(), pr, rtol=1e-5, atol=1e-6)
