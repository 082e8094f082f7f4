"""Learned warm start and denoiser for cylinder flows on 256 x 256 meshes.

No counterpart in the reference.  The plain DSMC path reaches converged
fields by developing the flow from a uniform fill and then averaging for
thousands of steps.  This module shortens both:

- two nets of one architecture (hystrath_amd.ws.ConvNet, bfloat16 MFMA
  kernels) map an image of fields to the long-average fields of the same
  case: the surrogate from the freestream state alone (the developed flow
  from the case's conditions and geometry), the denoiser from a short
  running average of a developed flow together with the surrogate's fields
  (its prior) and the average's length;
- the warm path fills the domain from the surrogate's fields
  (hs_dsmc_fill_cells), relaxes briefly, then samples, and every check
  interval denoises the running average; it stops at the first check whose
  denoised (or plain) fields are within the tolerance;
- the plain long-average DSMC stays the yardstick of every prediction's error,
  per field.

The denoiser learns from averages of 200 to 1,600 steps cut from the plain
runs, with priors from surrogates that did not see the case (two-fold
cross-fitting), so it learns how far to trust a prior it has not been
fitted to.  It sees a 17-pixel square and no condition planes (the
surrogate sees 65 pixels and the conditions), so it cannot recognise a
training case and recall its average; every training pair is also used
mirrored about the symmetry line y = 0 (Uy and the y plane change sign).

Image channels (pixel = cell (i, j) of the mesh, row j, column i):
  0 ln(rho / rho_inf), 1 Ux / U_inf, 2 Uy / U_inf, 3 ln(T_tr / T_inf),
  4 internal temperature (zero for argon), 5 Mach / 25, 6 log10(Kn) + 1,
  7 T_wall / T_inf / 5, 8 distance of the cell centre from the cylinder's
  surface in radii / 10 (clipped to [0, 1]), 9 and 10 the cell centre's x
  and y scaled to [-1, 1] over the domain, 11..15 the prior's channels 0..4
  (zero for the surrogate), 16 the noise level sqrt(200 / averaged steps)
  (zero for the surrogate); channels 17..31 zero.  The denoiser's condition
  planes are zero.  Cells inside the cylinder are zero and masked out of the
  loss.
"""
import math
import time

import numpy as np

from . import abi, cases, convergence, fields, ws
from .engine import Engine, evolve_batch

GRID = 256
WINDOW = 200
WINDOW_MULTS = (1, 2, 4, 8)   # training averages of 200, 400, 800 and 1,600 steps
RHO_FLOOR = 1e-3   # of rho_inf, before the log
T_FLOOR = 1e-2     # of T_inf


# ---------------------------------------------------------------------------
# cases
# ---------------------------------------------------------------------------
def make_cases(count, seed, grid=GRID, ppc=10.0):
    """Argon cylinder cases on grid x grid meshes over the batch ranges:
    Mach uniform in [5, 25], Kn log-uniform in [0.01, 1], wall temperature
    uniform in [300, 1000] K."""
    rng = np.random.default_rng(seed)
    out = []
    d_ref = 4.17e-10
    for k in range(count):
        mach = float(rng.uniform(5.0, 25.0))
        kn = float(10.0 ** rng.uniform(-2.0, 0.0))
        tw = float(rng.uniform(300.0, 1000.0))
        n_inf = 1.0 / (math.sqrt(2.0) * math.pi * d_ref * d_ref * kn * 0.2)
        ar = cases.make_species("Ar")
        U = mach * math.sqrt(5.0 / 3.0 * abi.HS_KB * 200.0 / ar.mass)
        dt = 0.25 * (1.0 / grid) / U
        c = cases.cylinder_flow(nx=grid, ny=grid, n_inf=n_inf, T_inf=200.0, mach=mach, T_wall=tw, ppc=ppc, dt=dt,
                                seed=int(rng.integers(1, 2 ** 31)), capacity=int(grid * grid * ppc * 3.0) + (1 << 16),
                                flags=abi.HS_FLAG_SAMPLE)
        c.ws_params = dict(mach=mach, kn=kn, T_wall=tw)
        out.append(c)
    return out


def fill_freestream(eng, case, fill_id=0):
    eng.fill([case.n_inf], case.T_inf, u=(case.U_inf, 0.0, 0.0), fill_id=fill_id)


# ---------------------------------------------------------------------------
# fields <-> images
# ---------------------------------------------------------------------------
def case_fields(samples, nsteps, case):
    """rho, Ux, Uy, T_tr and T_int (zero for argon) per cell."""
    f = convergence.field_set(samples, nsteps, case)
    return np.stack([f["rho"], f["Ux"], f["Uy"], f["Ttra"], f.get("Tint", np.zeros_like(f["rho"]))])


def fluid_mask(case):
    return (case.volumes > 0).astype(np.float32).reshape(case.cfg.ny, case.cfg.nx)


def encode(fields5, case):
    """[5, ncells] physical fields -> [grid, grid, 5] network variables."""
    ny, nx = case.cfg.ny, case.cfg.nx
    rho = np.maximum(fields5[0] / case.n_inf, RHO_FLOOR)
    T = np.maximum(fields5[3] / case.T_inf, T_FLOOR)
    Ti = np.maximum(fields5[4] / case.T_inf, 0.0)
    out = np.stack([np.log(rho), fields5[1] / case.U_inf, fields5[2] / case.U_inf, np.log(T), Ti], axis=-1)
    out = out.reshape(ny, nx, 5).astype(np.float32)
    return out * fluid_mask(case)[..., None]


def decode(img5, case):
    """[grid, grid, 5] network variables -> [5, ncells] physical fields."""
    x = img5.reshape(-1, 5).astype(np.float64)
    m = fluid_mask(case).reshape(-1) > 0
    # predictions are bounded to the range the encoding can hold
    rho = np.where(m, np.exp(np.clip(x[:, 0], np.log(RHO_FLOOR), np.log(1e4))) * case.n_inf, 0.0)
    T = np.where(m, np.exp(np.clip(x[:, 3], np.log(T_FLOOR), np.log(1e4))) * case.T_inf, 0.0)
    return np.stack([rho, x[:, 1] * case.U_inf * m, x[:, 2] * case.U_inf * m, T, x[:, 4] * case.T_inf * m])


def conditions(case):
    p = case.ws_params
    return np.array([p["mach"] / 25.0, math.log10(p["kn"]) + 1.0, p["T_wall"] / case.T_inf / 5.0], np.float32)


def geometry_planes(case):
    """[ny, nx, 3]: distance from the cylinder's surface in radii / 10
    (clipped to [0, 1]) and the cell centre scaled to [-1, 1], fluid cells
    only."""
    c = case.cfg
    ny, nx = c.ny, c.nx
    xc = c.origin[0] + (np.arange(nx) + 0.5) * c.delta[0]
    yc = c.origin[1] + (np.arange(ny) + 0.5) * c.delta[1]
    X, Y = np.meshgrid(xc, yc)
    r = np.hypot(X - c.cyl_center[0], Y - c.cyl_center[1])
    d = np.clip((r - c.cyl_radius) / c.cyl_radius / 10.0, 0.0, 1.0)
    sx = 2.0 * (X - c.origin[0]) / (nx * c.delta[0]) - 1.0
    sy = 2.0 * (Y - c.origin[1]) / (ny * c.delta[1]) - 1.0
    return np.stack([d, sx, sy], axis=-1).astype(np.float32) * fluid_mask(case)[..., None]


def net_input(img5, case, prior=None, nsteps=0, with_conditions=True):
    """The input image (ws.C_IN channels): the 5 field channels, 3 condition
    planes (zero without with_conditions), 3 geometry planes, the prior's 5
    channels and the noise level of an average of nsteps steps (both zero
    when not given)."""
    ny, nx = case.cfg.ny, case.cfg.nx
    mask = fluid_mask(case)
    x = np.zeros((ny, nx, ws.C_IN), np.float32)
    x[..., :5] = img5
    if with_conditions:
        x[..., 5:8] = conditions(case)[None, None, :] * mask[..., None]
    x[..., 8:11] = geometry_planes(case)
    if prior is not None:
        x[..., 11:16] = prior * mask[..., None]
    if nsteps:
        x[..., 16] = math.sqrt(WINDOW / nsteps) * mask
    return x


def freestream_image(case):
    """The network variables of the undisturbed freestream."""
    ny, nx = case.cfg.ny, case.cfg.nx
    img = np.zeros((ny, nx, 5), np.float32)
    img[..., 1] = 1.0
    return img * fluid_mask(case)[..., None]


# ---------------------------------------------------------------------------
# training pairs from plain DSMC
# ---------------------------------------------------------------------------
def generate(case_list, develop=3000, average=10000, window=WINDOW, mults=WINDOW_MULTS, log=print):
    """Steps every case together (hs_dsmc_evolve_batch): a uniform fill,
    develop steps without sampling, then average steps sampled, read every
    window steps.  Returns per case: (list of (steps, image) averages,
    long-average image).  The averages are differences of the cumulative
    sums over m * window steps for m in mults (non-overlapping; of the
    shortest only every other one), so all come from the one run whose
    average is the target."""
    engs = []
    try:
        flags = []
        for c in case_list:
            e = Engine(c)
            fill_freestream(e, c)
            flags.append(c.cfg.flags)
            e.set_flags(c.cfg.flags & ~abi.HS_FLAG_SAMPLE)
            engs.append(e)
        t0 = time.time()
        done = 0
        while done < develop:
            k = min(250, develop - done)
            evolve_batch(engs, k)
            done += k
        for e, f in zip(engs, flags):
            e.synchronize()
            e.set_flags(f)
            e.reset_samples()
        log(f"[ws] {len(engs)} cases developed for {develop} steps in {time.time() - t0:.1f} s")
        ring = [{0: None} for _ in engs]   # cumulative sums at the last boundaries (None: zero)
        wins = [[] for _ in engs]
        keep = max(mults)
        for j in range(1, average // window + 1):
            evolve_batch(engs, window)
            for k, (e, c) in enumerate(zip(engs, case_list)):
                s, n = e.samples()
                ring[k][j] = s
                for m in mults:
                    if j % m or (m == 1 and len(mults) > 1 and j % 2):
                        continue
                    base = ring[k][j - m]
                    ws_ = s - base if base is not None else s
                    wins[k].append((m * window, encode(case_fields(ws_, m * window, c), c)))
                ring[k].pop(j - keep, None)
        out = []
        for k, (e, c) in enumerate(zip(engs, case_list)):
            s, n = e.samples()
            out.append((wins[k], encode(case_fields(s, n, c), c)))
            c_ = e.counters()
            if c_["track_overflow"] or c_["tile_overflow"] or c_["capacity_overflow"]:
                raise RuntimeError(f"case {k}: device-side bound tripped while generating data: {c_}")
        log(f"[ws] {average} averaged steps, {len(wins[0]) if wins else 0} averages a case, "
            f"{time.time() - t0:.1f} s in all")
        return out
    finally:
        for e in engs:
            e.close()


# channels that change sign when an image is mirrored about y = 0: Uy, the
# y plane, the prior's Uy
MIRROR_SIGN_IN = (2, 10, 13)
MIRROR_SIGN_OUT = (2,)


def mirror(img, sign_channels):
    """The image mirrored about the domain's symmetry line y = 0 (its rows
    reversed; the listed channels change sign)."""
    out = img[..., ::-1, :, :].copy() if img.ndim == 4 else img[::-1].copy()
    for ch in sign_channels:
        out[..., ch] = -out[..., ch]
    return out


def loss_weights(target, case):
    """Per-pixel weights of the training loss: the long-average density over
    the freestream's (clipped to [0.05, 20]), mean 1 over the fluid cells.
    The long average's relative L2 errors of rho and T weigh each cell by the
    field's square, so the dense shock layer dominates them; uniform weights
    in the log variables would spend the net on the sparse wake."""
    m = fluid_mask(case)
    w = np.clip(np.exp(target[..., 0]), 0.05, 20.0) * m
    return (w / max(w.sum() / max(m.sum(), 1.0), 1e-30)).astype(np.float32)


def _tensors(case_list, per_case, device, mirrored=False):
    """per_case: for each case (list of input images, list of the field
    images the residual is taken from, long-average image); mirrored adds
    every pair mirrored about y = 0.  The masks carry the loss weights."""
    import torch
    X, Y, M = [], [], []
    for c, (ins, bases, target) in zip(case_list, per_case):
        fluid = fluid_mask(c)
        wgt = loss_weights(target, c)
        xs = np.stack(ins)
        ys = np.zeros((len(ins),) + fluid.shape + (ws.C_OUT,), np.float32)
        ys[..., :5] = (target[None] - np.stack(bases)) * fluid[None, ..., None]
        ms = np.broadcast_to(wgt, (len(ins),) + wgt.shape)
        if mirrored:
            xs = np.concatenate([xs, mirror(xs, MIRROR_SIGN_IN)])
            ys = np.concatenate([ys, mirror(ys, MIRROR_SIGN_OUT)])
            ms = np.concatenate([ms, np.broadcast_to(wgt[::-1], ms.shape)])
        X.append(torch.as_tensor(xs).to(device).to(torch.bfloat16).view(torch.int16))
        Y.append(torch.as_tensor(ys).to(device))
        M.append(torch.as_tensor(np.array(ms, dtype=np.float32)).to(device))
    return torch.cat(X), torch.cat(Y), torch.cat(M)


def surrogate_pairs(case_list, generated, device="cuda"):
    """Device tensors (inputs [N, g, g, C_IN] bf16 bits, residual targets
    [N, g, g, C_OUT] FP32, masks [N, g, g] FP32) with one input per case:
    its freestream image."""
    per = [([net_input(freestream_image(c), c)], [freestream_image(c)], t) for c, (_, t) in zip(case_list, generated)]
    return _tensors(case_list, per, device)


def denoiser_pairs(case_list, generated, priors, device="cuda"):
    """The same tensors for every average of every case, with the case's
    prior (surrogate fields) and the average's length as inputs."""
    per = []
    for c, (wins, t), pr in zip(case_list, generated, priors):
        per.append(([net_input(img, c, prior=pr, nsteps=n, with_conditions=False) for n, img in wins],
                    [img for _, img in wins], t))
    return _tensors(case_list, per, device, mirrored=True)


def train_surrogates(case_list, generated, epochs=400, batch=16, lr=2e-3, log=print, device="cuda"):
    """The surrogate on every case, and each case's prior from a surrogate
    trained on the other half of the cases (two folds): returns (surrogate,
    priors [case][g, g, 5], per-epoch losses of the full surrogate)."""
    priors = [None] * len(case_list)
    idx = np.arange(len(case_list))
    for f in (0, 1):
        tr, te = idx[idx % 2 != f], idx[idx % 2 == f]
        X, Y, M = surrogate_pairs([case_list[i] for i in tr], [generated[i] for i in tr], device)
        net, _ = train(X, Y, M, epochs=epochs, batch=batch, lr=lr, log=log, device=device, what=f"surrogate fold {f}",
                       seed=10 + f)
        del X, Y, M
        for i in te:
            priors[i] = predict(net, freestream_image(case_list[i]), case_list[i], device)
    X, Y, M = surrogate_pairs(case_list, generated, device)
    net, hist = train(X, Y, M, epochs=epochs, batch=batch, lr=lr, log=log, device=device, what="surrogate")
    return net, priors, hist


# ---------------------------------------------------------------------------
# training and prediction
# ---------------------------------------------------------------------------
def train(X, Y, M, epochs=20, batch=16, lr=2e-3, seed=0, log=print, device="cuda", what="net", layers=None):
    """Adam on the masked MSE of the residual (device tensors from
    denoiser_pairs or surrogate_pairs), cosine learning-rate decay; returns
    (net, per-epoch mean losses)."""
    import torch
    from . import ws
    net = ws.ConvNet(device, seed=seed, layers=layers)
    n = len(X)
    batch = min(batch, n)
    Xd, Yd, Md = X, Y, M
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        order = rng.permutation(n)
        tot = 0.0
        lr_ep = lr * 0.5 * (1.0 + math.cos(math.pi * ep / epochs))
        for b0 in range(0, n - batch + 1, batch):
            idx = torch.as_tensor(order[b0:b0 + batch], device=device)
            xb, yb, mb = Xd[idx].contiguous(), Yd[idx].contiguous(), Md[idx].contiguous()
            out, acts = net.forward(xb, keep=True)
            loss, dy = ws.mse(out, yb, mb)
            net.adam(net.backward(acts, dy), lr=lr_ep)
            tot += loss.item()
        hist.append(tot / max(1, n // batch))
        if ep == epochs - 1 or (ep + 1) % max(1, epochs // 10) == 0:
            log(f"[ws] {what} epoch {ep + 1}/{epochs}: loss {hist[-1]:.5f} ({time.time() - t0:.1f} s)")
    return net, hist


def predict(net, img5, case, device="cuda", prior=None, nsteps=0):
    """Long-average network variables [g, g, 5] from an input image; with a
    prior (the denoiser) also from the prior and the averaged steps, without
    condition planes."""
    import torch
    xin = net_input(img5, case, prior, nsteps, with_conditions=prior is None)
    x = torch.as_tensor(xin[None]).to(device).to(torch.bfloat16).view(torch.int16)
    out = net.forward(x).cpu().numpy()[0]
    return (img5 + out[..., :5]) * fluid_mask(case)[..., None]


def warm_fill(eng, case, img5, fill_id=7):
    """A population from predicted fields: each fluid cell filled at its
    predicted density, velocity and temperature."""
    f = decode(img5, case)
    nc = case.ncells
    n_cell = np.ascontiguousarray(np.maximum(f[0], 0.0))
    u_cell = np.ascontiguousarray(np.concatenate([f[1], f[2], np.zeros(nc)]))
    t_cell = np.ascontiguousarray(np.maximum(f[3], 1.0))
    eng.fill_cells(n_cell, u_cell, t_cell, fill_id=fill_id)


# ---------------------------------------------------------------------------
# time to converged fields: plain path against warm path
# ---------------------------------------------------------------------------
def long_average_fields(case, long_image):
    """The long average's fields (convergence.field_set layout) from its image."""
    f = decode(long_image, case)
    m = fluid_mask(case).reshape(-1) > 0
    return {"rho": f[0], "Ux": f[1], "Uy": f[2], "Ttra": f[3], "nmean": np.where(m, 1.0, 0.0)}


def _errors_of_image(img5, case, long_avg):
    f = decode(img5, case)
    run = {"rho": f[0], "Ux": f[1], "Uy": f[2], "Ttra": f[3]}
    return convergence.errors(run, long_avg)


def plain_time_to_tol(case, long_avg, develop=3000, check=250, max_sampled=20000, tol=0.02, seed_offset=1000):
    """Steps of the plain path: uniform fill, develop steps unsampled, then
    sampling until the running average is within tol of the long average."""
    c2 = _reseeded(case, seed_offset)
    with Engine(c2) as e:
        fill_freestream(e, c2)
        flags = c2.cfg.flags
        e.set_flags(flags & ~abi.HS_FLAG_SAMPLE)
        e.evolve(develop)
        e.set_flags(flags)
        e.reset_samples()
        done, err = 0, None
        t0 = time.perf_counter()
        while done < max_sampled:
            e.evolve(check)
            done += check
            s, n = e.samples()
            err = _errors_of_image(encode(case_fields(s, n, c2), c2), c2, long_avg)
            if convergence.converged(err, tol):
                break
        e.synchronize()
    return dict(steps=develop + done, develop=develop, sampled=done, reached=convergence.converged(err, tol),
                errors=err)


def warm_time_to_tol(nets, case, long_avg, relax=500, check=WINDOW, max_sampled=20000, tol=0.02, seed_offset=2000,
                     device="cuda"):
    """Steps of the warm path: the surrogate's fields (nets[0], from the
    freestream image) fill the domain and relax steps run unsampled; then
    every check steps the running average is denoised (nets[1], with the
    surrogate's fields as its prior), and the path stops at the first check
    whose denoised fields -- or, failing that, whose plain running average
    -- are within tol of the long average."""
    import torch
    surrogate, denoiser = nets
    c2 = _reseeded(case, seed_offset)
    t_inf = 0.0
    with Engine(c2) as e:
        t = time.perf_counter()
        img0 = predict(surrogate, freestream_image(c2), c2, device)
        torch.cuda.synchronize()
        t_inf += time.perf_counter() - t
        warm_fill(e, c2, img0)
        flags = c2.cfg.flags
        e.set_flags(flags & ~abi.HS_FLAG_SAMPLE)
        if relax:
            e.evolve(relax)
        e.set_flags(flags)
        e.reset_samples()
        err_surr = _errors_of_image(img0, c2, long_avg)
        done, err, used = 0, None, "none within the tolerance"
        first = None
        while done < max_sampled:
            e.evolve(check)
            done += check
            s, n = e.samples()
            raw = encode(case_fields(s, n, c2), c2)
            t = time.perf_counter()
            den = predict(denoiser, raw, c2, device, prior=img0, nsteps=n)
            torch.cuda.synchronize()
            t_inf += time.perf_counter() - t
            err_d = _errors_of_image(den, c2, long_avg)
            err_r = _errors_of_image(raw, c2, long_avg)
            if first is None:
                first = (err_r, err_d)
            if convergence.converged(err_d, tol):
                err, used = err_d, "denoised running average"
                break
            if convergence.converged(err_r, tol):
                err, used = err_r, "plain running average"
                break
            err = err_d
        e.synchronize()
    return dict(steps=relax + done, relax=relax, sampled=done, reached=convergence.converged(err, tol),
                errors=err, errors_surrogate=err_surr, errors_window=first[0], errors_denoised=first[1],
                used=used, inference_s=t_inf)


def _reseeded(case, offset):
    """The same case with another random stream (an independent run)."""
    import copy
    c2 = copy.copy(case)
    c2.cfg = abi.hs_config.from_buffer_copy(case.cfg)
    c2.cfg.seed = case.cfg.seed + offset
    c2.cfg.cell_volume = abi.ptr(case.volumes, abi.C.c_double)
    return c2


def save_nets(nets, path):
    """Both nets (surrogate, denoiser) in one .npz, each with its layer table."""
    arrays = {}
    for tag, net in zip(("s", "d"), nets):
        arrays[f"{tag}_layers"] = np.array(net.layers, np.int32)
        st = net.state()
        arrays.update({f"{tag}_w{k}": a for k, a in enumerate(st["w"])})
        arrays.update({f"{tag}_b{k}": a for k, a in enumerate(st["b"])})
        arrays[f"{tag}_t"] = np.array(st["t"])
    np.savez_compressed(path, **arrays)


def load_nets(path, device="cuda"):
    """(surrogate, denoiser) from save_nets; refuses a file written for
    another input or output layout."""
    import torch
    d = np.load(path)
    nets = []
    for tag in ("s", "d"):
        if f"{tag}_layers" not in d:
            raise ValueError(f"{path}: no layer table for the {'surrogate' if tag == 's' else 'denoiser'}")
        layers = [tuple(int(v) for v in r) for r in d[f"{tag}_layers"]]
        net = ws.ConvNet(device, layers=layers)
        for k in range(len(layers)):
            net.w[k].copy_(torch.as_tensor(d[f"{tag}_w{k}"]))
            net.b[k].copy_(torch.as_tensor(d[f"{tag}_b{k}"]))
        net.t = int(d[f"{tag}_t"])
        net.repack()
        nets.append(net)
    return tuple(nets)


def main(argv=None):
    """Generates training pairs from plain DSMC, trains the net and compares
    the time to 2% fields of the plain and the warm path on held-out cases.
    Prints one JSON object."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__.splitlines()[0])
    ap.add_argument("--train-cases", type=int, default=40)
    ap.add_argument("--test-cases", type=int, default=4)
    ap.add_argument("--develop", type=int, default=3000)
    ap.add_argument("--average", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=20, help="denoiser epochs")
    ap.add_argument("--epochs-surrogate", type=int, default=400)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--relax", type=int, default=500)
    ap.add_argument("--save", default="")
    ap.add_argument("--seed", type=int, default=2025)
    a = ap.parse_args(argv)

    def log(*x):
        import sys
        print(*x, file=sys.stderr, flush=True)

    import torch
    # the network's memory comes from torch: its runtime starts before the
    # engines' so both see the device
    torch.cuda.init()
    torch.zeros(1, device="cuda")
    t0 = time.time()
    train_cases = make_cases(a.train_cases, a.seed)
    test_cases = make_cases(a.test_cases, a.seed + 1)
    gen = generate(train_cases + test_cases, develop=a.develop, average=a.average, log=log)
    t_gen = time.time() - t0
    t1 = time.time()
    surrogate, priors, hist_s = train_surrogates(train_cases, gen[:a.train_cases], epochs=a.epochs_surrogate,
                                                 batch=a.batch, lr=a.lr, log=log)
    X, Y, M = denoiser_pairs(train_cases, gen[:a.train_cases], priors)
    log(f"[ws] {len(X)} denoiser pairs")
    denoiser, hist = train(X, Y, M, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log, what="denoiser", seed=1,
                           layers=ws.DENOISER_LAYERS)
    t_train = time.time() - t1
    del X, Y, M
    nets = (surrogate, denoiser)
    if a.save:
        save_nets(nets, a.save)
    results = []
    for c, (wins, target) in zip(test_cases, gen[a.train_cases:]):
        long_avg = long_average_fields(c, target)
        plain = plain_time_to_tol(c, long_avg, develop=a.develop)
        warm = warm_time_to_tol(nets, c, long_avg, relax=a.relax)
        ratio = plain["steps"] / max(warm["steps"], 1)
        log(f"[ws] case {c.ws_params}: plain {plain['steps']} steps, warm {warm['steps']} steps "
            f"({warm['used']}), x{ratio:.2f}; surrogate {warm['errors_surrogate']}; window {warm['errors_window']}; "
            f"denoised {warm['errors_denoised']}")
        results.append(dict(params=c.ws_params, plain=plain, warm=warm, step_ratio=ratio))
    out = dict(generate_s=round(t_gen, 1), train_s=round(t_train, 1), train_cases=a.train_cases,
               loss_surrogate=[round(h, 6) for h in hist_s], loss_denoiser=[round(h, 6) for h in hist],
               cases=results)
    print(json.dumps(out, default=float))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
