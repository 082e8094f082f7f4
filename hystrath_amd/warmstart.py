"""Learned warm start and denoiser for cylinder flows on 256 x 256 meshes.

No counterpart in the reference.  The plain DSMC path reaches converged
fields by developing the flow from a uniform fill and then averaging for
thousands of steps.  This module shortens both:

- a net (hystrath_amd.ws.ConvNet, bfloat16 MFMA kernels) maps an image of
  fields to the long-average fields of the same case: from the freestream
  state alone (a surrogate of the developed flow) or from a 200-step
  sampling window (a denoiser);
- the warm path fills the domain from the surrogate's fields
  (hs_dsmc_fill_cells), relaxes briefly, samples one window and denoises it;
- the plain long-average DSMC stays the judge of every prediction's error,
  per field, and the fallback: the warm path keeps sampling the particles
  it started if the prediction is not within the tolerance.

Image channels (pixel = cell (i, j) of the mesh, row j, column i):
  0 ln(rho / rho_inf), 1 Ux / U_inf, 2 Uy / U_inf, 3 ln(T_tr / T_inf),
  4 internal temperature (zero for argon), 5 Mach / 25, 6 log10(Kn) + 1,
  7 T_wall / T_inf / 5; channels 8..15 zero.  Cells inside the cylinder are
  zero and masked out of the loss.
"""
import math
import time

import numpy as np

from . import abi, cases, convergence, fields
from .engine import Engine, evolve_batch

GRID = 256
WINDOW = 200
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


def net_input(img5, case):
    """The 16-channel input image: the 5 field channels, 3 condition planes."""
    ny, nx = case.cfg.ny, case.cfg.nx
    x = np.zeros((ny, nx, 16), np.float32)
    x[..., :5] = img5
    x[..., 5:8] = conditions(case)[None, None, :] * fluid_mask(case)[..., None]
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
def generate(case_list, develop=3000, average=10000, window=WINDOW, log=print):
    """Steps every case together (hs_dsmc_evolve_batch): a uniform fill,
    develop steps without sampling, then average steps sampled, read every
    window steps.  Returns per case: (list of window images, long-average
    image).  The window images are differences of the cumulative sums, so
    all come from the one run whose average is the target."""
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
        prev = [None] * len(engs)
        wins = [[] for _ in engs]
        for w0 in range(0, average, window):
            evolve_batch(engs, window)
            for k, (e, c) in enumerate(zip(engs, case_list)):
                s, n = e.samples()
                ws_ = s - prev[k] if prev[k] is not None else s.copy()
                prev[k] = s
                wins[k].append(encode(case_fields(ws_, window, c), c))
        out = []
        for k, (e, c) in enumerate(zip(engs, case_list)):
            s, n = e.samples()
            out.append((wins[k], encode(case_fields(s, n, c), c)))
            c_ = e.counters()
            if c_["track_overflow"] or c_["tile_overflow"] or c_["capacity_overflow"]:
                raise RuntimeError(f"case {k}: device-side bound tripped while generating data: {c_}")
        log(f"[ws] {average} averaged steps, {average // window} windows a case, {time.time() - t0:.1f} s in all")
        return out
    finally:
        for e in engs:
            e.close()


def pairs_from(case_list, generated, freestream_every=5, device="cuda"):
    """Device tensors (inputs [N, g, g, 16] bf16 bits, residual targets
    [N, g, g, 16] FP32, masks [N, g, g] FP32) from generated data: every
    window of every case, and the freestream image of each case repeated once
    per freestream_every windows (the surrogate task)."""
    import torch
    X, Y, M = [], [], []
    for c, (wins, target) in zip(case_list, generated):
        mask = fluid_mask(c)
        inputs = list(wins) + [freestream_image(c)] * max(1, len(wins) // freestream_every)
        xs = np.stack([net_input(img, c) for img in inputs])
        ys = np.zeros_like(xs)
        ys[..., :5] = (target[None] - np.stack(inputs)) * mask[None, ..., None]
        X.append(torch.as_tensor(xs).to(device).to(torch.bfloat16).view(torch.int16))
        Y.append(torch.as_tensor(ys).to(device))
        M.append(torch.as_tensor(np.broadcast_to(mask, (len(inputs),) + mask.shape).copy()).to(device))
    return torch.cat(X), torch.cat(Y), torch.cat(M)


# ---------------------------------------------------------------------------
# training and prediction
# ---------------------------------------------------------------------------
def train(X, Y, M, epochs=20, batch=16, lr=2e-3, seed=0, log=print, device="cuda"):
    """Adam on the masked MSE of the residual (device tensors from
    pairs_from), cosine learning-rate decay; returns (net, per-epoch mean
    losses)."""
    import torch
    from . import ws
    net = ws.ConvNet(device, seed=seed)
    n = len(X)
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
        log(f"[ws] epoch {ep + 1}/{epochs}: loss {hist[-1]:.5f} ({time.time() - t0:.1f} s)")
    return net, hist


def predict(net, img5, case, device="cuda"):
    """Long-average network variables [g, g, 5] from an input image."""
    import torch
    x = torch.as_tensor(net_input(img5, case)[None]).to(device).to(torch.bfloat16).view(torch.int16)
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
def judge_fields(case, long_image):
    """The judge's fields (convergence.field_set layout) from its image."""
    f = decode(long_image, case)
    m = fluid_mask(case).reshape(-1) > 0
    return {"rho": f[0], "Ux": f[1], "Uy": f[2], "Ttra": f[3], "nmean": np.where(m, 1.0, 0.0)}


def _errors_of_image(img5, case, judge):
    f = decode(img5, case)
    run = {"rho": f[0], "Ux": f[1], "Uy": f[2], "Ttra": f[3]}
    return convergence.errors(run, judge)


def plain_time_to_tol(case, judge, develop=3000, check=250, max_sampled=20000, tol=0.02, seed_offset=1000):
    """Steps of the plain path: uniform fill, develop steps unsampled, then
    sampling until the running average is within tol of the judge."""
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
            err = _errors_of_image(encode(case_fields(s, n, c2), c2), c2, judge)
            if convergence.converged(err, tol):
                break
        e.synchronize()
    return dict(steps=develop + done, develop=develop, sampled=done, reached=convergence.converged(err, tol),
                errors=err)


def warm_time_to_tol(net, case, judge, relax=300, window=WINDOW, check=250, max_sampled=20000, tol=0.02,
                     seed_offset=2000, device="cuda"):
    """Steps of the warm path: the surrogate's fields from the freestream
    image fill the domain, relax steps unsampled, one sampled window is
    denoised by the net; if that is not within tol of the judge the
    particles keep being sampled (the plain average is the fallback)."""
    import torch
    c2 = _reseeded(case, seed_offset)
    t_inf = 0.0
    with Engine(c2) as e:
        t = time.perf_counter()
        img0 = predict(net, freestream_image(c2), c2, device)
        torch.cuda.synchronize()
        t_inf += time.perf_counter() - t
        warm_fill(e, c2, img0)
        flags = c2.cfg.flags
        e.set_flags(flags & ~abi.HS_FLAG_SAMPLE)
        if relax:
            e.evolve(relax)
        e.set_flags(flags)
        e.reset_samples()
        e.evolve(window)
        s, n = e.samples()
        t = time.perf_counter()
        img1 = predict(net, encode(case_fields(s, n, c2), c2), c2, device)
        torch.cuda.synchronize()
        t_inf += time.perf_counter() - t
        err_net = _errors_of_image(img1, c2, judge)
        err_surr = _errors_of_image(img0, c2, judge)
        err_raw = _errors_of_image(encode(case_fields(s, n, c2), c2), c2, judge)
        done, err, used = window, err_net, "denoised window"
        if not convergence.converged(err, tol):
            used = "plain average after the warm fill"
            while done < max_sampled:
                e.evolve(check)
                done += check
                s, n = e.samples()
                err = _errors_of_image(encode(case_fields(s, n, c2), c2), c2, judge)
                if convergence.converged(err, tol):
                    break
        e.synchronize()
    return dict(steps=relax + done, relax=relax, sampled=done, reached=convergence.converged(err, tol),
                errors=err, errors_surrogate=err_surr, errors_window=err_raw, errors_denoised=err_net,
                used=used, inference_s=t_inf)


def _reseeded(case, offset):
    """The same case with another random stream (an independent run)."""
    import copy
    c2 = copy.copy(case)
    c2.cfg = abi.hs_config.from_buffer_copy(case.cfg)
    c2.cfg.seed = case.cfg.seed + offset
    c2.cfg.cell_volume = abi.ptr(case.volumes, abi.C.c_double)
    return c2


def save_net(net, path):
    st = net.state()
    arrays = {f"w{k}": a for k, a in enumerate(st["w"])}
    arrays.update({f"b{k}": a for k, a in enumerate(st["b"])})
    np.savez_compressed(path, t=np.array(st["t"]), **arrays)


def load_net(path, device="cuda"):
    import torch
    from . import ws
    net = ws.ConvNet(device)
    d = np.load(path)
    for k in range(len(ws.LAYERS)):
        net.w[k].copy_(torch.as_tensor(d[f"w{k}"]))
        net.b[k].copy_(torch.as_tensor(d[f"b{k}"]))
    net.t = int(d["t"])
    net.repack()
    return net


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
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--relax", type=int, default=300)
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
    X, Y, M = pairs_from(train_cases, gen[:a.train_cases])
    log(f"[ws] {len(X)} training pairs")
    t1 = time.time()
    net, hist = train(X, Y, M, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
    t_train = time.time() - t1
    del X, Y, M
    if a.save:
        save_net(net, a.save)
    results = []
    for c, (wins, target) in zip(test_cases, gen[a.train_cases:]):
        judge = judge_fields(c, target)
        plain = plain_time_to_tol(c, judge, develop=a.develop)
        warm = warm_time_to_tol(net, c, judge, relax=a.relax)
        ratio = plain["steps"] / max(warm["steps"], 1)
        log(f"[ws] case {c.ws_params}: plain {plain['steps']} steps, warm {warm['steps']} steps "
            f"({warm['used']}), x{ratio:.2f}; denoised {warm['errors_denoised']}")
        results.append(dict(params=c.ws_params, plain=plain, warm=warm, step_ratio=ratio))
    out = dict(generate_s=round(t_gen, 1), train_s=round(t_train, 1), pairs=len(hist) and int(a.train_cases),
               loss=[round(h, 6) for h in hist], cases=results)
    print(json.dumps(out, default=float))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
