"""The learned warm-start's network on the gfx950 kernels of include/hystrath_ws.h.

No counterpart in the reference.  A six-layer 3x3 convolutional net
(17 -> 32 -> 64 -> 64 -> 64 -> 32 -> 5 channels, ReLU between layers, same
resolution throughout, dilations 1, 2, 4, 8, 16, 1: each output pixel sees a
65-pixel square) maps five field channels, three condition planes, three
geometry planes, five prior-field channels and a noise-level plane to a
correction of the five fields.  Activations are bfloat16 NHWC images, weights FP32 masters with
bfloat16 packings, and every convolution, input gradient and weight gradient
is a bfloat16 MFMA kernel accumulating in FP32.  torch provides device memory
and the stream only.
"""
import ctypes as C

import numpy as np

from . import engine

# (true in, true out, stored in, stored out) channels and the dilation of each layer:
# the surrogate's (a 65-pixel receptive field) and the denoiser's (17 pixels)
LAYERS = [(17, 32, 32, 32, 1), (32, 64, 32, 64, 2), (64, 64, 64, 64, 4), (64, 64, 64, 64, 8), (64, 32, 64, 32, 16),
          (32, 5, 32, 16, 1)]
DENOISER_LAYERS = [(17, 32, 32, 32, 1), (32, 64, 32, 64, 1), (64, 64, 64, 64, 2), (64, 64, 64, 64, 2),
                   (64, 32, 64, 32, 1), (32, 5, 32, 16, 1)]
N_FIELDS = 5
N_IN = 17
C_IN = 32    # stored channels of the input image
C_OUT = 16   # stored channels of the output image

_sig_done = False


def lib():
    global _sig_done
    L = engine.load_library()
    if not _sig_done:
        P, vp = C.POINTER, C.c_void_p
        i32 = C.c_int32
        L.hs_ws_pack.argtypes = [vp, i32, i32, vp, vp, vp]
        L.hs_ws_conv3x3.argtypes = [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, i32, vp]
        L.hs_ws_conv3x3_wgrad.argtypes = [vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp]
        L.hs_ws_mse.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, C.c_double, vp]
        L.hs_ws_adam.argtypes = [vp, vp, vp, vp, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, i32, vp]
        L.hs_ws_mfma_tile.argtypes = [i32, P(C.c_uint16), P(C.c_uint16), P(C.c_float)]
        for f in ("hs_ws_pack", "hs_ws_conv3x3", "hs_ws_conv3x3_wgrad", "hs_ws_mse", "hs_ws_adam", "hs_ws_mfma_tile"):
            getattr(L, f).restype = C.c_int
        _sig_done = True
    return L


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ck(rc, what):
    if rc != 0:
        raise engine.EngineError(f"{what} failed ({rc})")


def to_bf16_bits(t):
    """A float tensor rounded to bfloat16, as its int16 bit patterns."""
    import torch
    return t.to(torch.bfloat16).view(torch.int16)


def from_bf16_bits(b):
    import torch
    return b.view(torch.bfloat16).float()


def conv3x3(x_bits, w_pack, bias, n, h, w, cin, cout, relu=False, mask_bits=None, out_bf16=True, out_f32=False,
            dilation=1):
    """One (dilated) 3x3 convolution on bfloat16 NHWC images (int16 bit tensors)."""
    import torch
    dev = x_bits.device
    y = torch.empty((n, h, w, cout), dtype=torch.int16, device=dev) if out_bf16 else None
    y32 = torch.empty((n, h, w, cout), dtype=torch.float32, device=dev) if out_f32 else None
    _ck(lib().hs_ws_conv3x3(_p(x_bits), _p(w_pack), _p(bias), _p(mask_bits), _p(y), _p(y32), n, h, w, cin, cout,
                            1 if relu else 0, dilation, _stream()), "hs_ws_conv3x3")
    return y, y32


def pack(wm, cin, cout):
    import torch
    dev = wm.device
    cout_pad, cin_pad = (cout + 31) // 32 * 32, (cin + 31) // 32 * 32
    wf = torch.empty(9 * cout_pad * cin, dtype=torch.int16, device=dev)
    wb = torch.empty(9 * cin_pad * cout, dtype=torch.int16, device=dev)
    _ck(lib().hs_ws_pack(_p(wm), cin, cout, _p(wf), _p(wb), _stream()), "hs_ws_pack")
    return wf, wb


def wgrad(x_bits, dy_bits, dw, db, n, h, w, cin, cout, dilation=1):
    _ck(lib().hs_ws_conv3x3_wgrad(_p(x_bits), _p(dy_bits), _p(dw), _p(db), n, h, w, cin, cout, dilation, _stream()),
        "hs_ws_conv3x3_wgrad")


class ConvNet:
    """The network's parameters (FP32 masters, stored-channel shapes) and its
    forward and backward passes on the device."""

    def __init__(self, device="cuda", seed=0, layers=None):
        import torch
        self.layers = [tuple(r) for r in (layers or LAYERS)]
        if self.layers[0][2] != C_IN or self.layers[-1][3] != C_OUT or len(self.layers) < 2:
            raise ValueError("the first layer takes C_IN stored channels and the last gives C_OUT")
        g = torch.Generator().manual_seed(seed)
        self.w, self.b = [], []
        for k, (ti, to, si, so, _) in enumerate(self.layers):
            std = (2.0 / (9 * ti)) ** 0.5 * (0.1 if k == len(self.layers) - 1 else 1.0)
            wm = torch.zeros(9, si, so)
            wm[:, :ti, :to] = torch.randn(9, ti, to, generator=g) * std
            self.w.append(wm.to(device).contiguous())
            self.b.append(torch.zeros(so, device=device))
        self.m = [torch.zeros_like(t) for t in self.w + self.b]
        self.v = [torch.zeros_like(t) for t in self.w + self.b]
        self.t = 0
        self.repack()

    def repack(self):
        self.wf, self.wb = [], []
        for (ti, to, si, so, _), wm in zip(self.layers, self.w):
            f, b = pack(wm, si, so)
            self.wf.append(f)
            self.wb.append(b)

    def params(self):
        return self.w + self.b

    def forward(self, x_bits, keep=False):
        """x: [n, h, w, 16] bf16 bits.  Returns the FP32 output [n, h, w, 16]
        (5 true channels) and, with keep, the activations for backward."""
        n, h, w, _ = x_bits.shape
        acts = [x_bits]
        a = x_bits
        out = None
        for k, (ti, to, si, so, dl) in enumerate(self.layers):
            last = k == len(self.layers) - 1
            y, y32 = conv3x3(a, self.wf[k], self.b[k], n, h, w, si, so, relu=not last, out_bf16=not last,
                             out_f32=last, dilation=dl)
            if last:
                out = y32
            else:
                a = y
                acts.append(a)
        return (out, acts) if keep else out

    def backward(self, acts, dy_bits):
        """Gradients of every parameter given dL/d(output) (bf16 bits, 16
        channels); returns the list matching params()."""
        import torch
        n, h, w, _ = acts[0].shape
        gw = [torch.zeros_like(t) for t in self.w]
        gb = [torch.zeros_like(t) for t in self.b]
        d = dy_bits
        for k in range(len(self.layers) - 1, -1, -1):
            ti, to, si, so, dl = self.layers[k]
            wgrad(acts[k], d, gw[k], gb[k], n, h, w, si, so, dilation=dl)
            if k > 0:
                # dL/d(input of layer k) through the ReLU of layer k-1: masked by its output
                d, _ = conv3x3(d, self.wb[k], None, n, h, w, so, si, mask_bits=acts[k], dilation=dl)
        return gw + gb

    def adam(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        for p, g, m, v in zip(self.params(), grads, self.m, self.v):
            _ck(lib().hs_ws_adam(_p(p), _p(g), _p(m), _p(v), p.numel(), lr, b1, b2, eps, self.t, _stream()),
                "hs_ws_adam")
        self.repack()

    def state(self):
        return {"w": [t.cpu().numpy() for t in self.w], "b": [t.cpu().numpy() for t in self.b], "t": self.t}


def mse(pred, target, mask, nf=N_FIELDS):
    """Masked MSE (FP32 scalar tensor) and its gradient (bf16 bits)."""
    import torch
    n, h, w, c = pred.shape
    dy = torch.empty((n, h, w, c), dtype=torch.int16, device=pred.device)
    loss = torch.zeros(1, dtype=torch.float32, device=pred.device)
    count = float((mask != 0).sum().item())
    _ck(lib().hs_ws_mse(_p(pred), _p(target), _p(mask), _p(dy), _p(loss), n, h, w, c, nf, max(count, 1.0),
                        _stream()), "hs_ws_mse")
    return loss, dy


def mfma_tile(kind, a_bits, b_bits):
    """Runs one MFMA on operands given as each lane holds them (test hook)."""
    a = np.ascontiguousarray(a_bits, np.uint16)
    b = np.ascontiguousarray(b_bits, np.uint16)
    d = np.zeros(64 * (16 if kind == 0 else 4), np.float32)
    _ck(lib().hs_ws_mfma_tile(kind, a.ctypes.data_as(C.POINTER(C.c_uint16)), b.ctypes.data_as(C.POINTER(C.c_uint16)),
                              d.ctypes.data_as(C.POINTER(C.c_float))), "hs_ws_mfma_tile")
    return d
