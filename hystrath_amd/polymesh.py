"""OpenFOAM polyMesh (constant/polyMesh: points, faces, owner, neighbour,
boundary) read and recognised as the engine's mesh: one uniform Cartesian
block of hex cells numbered as blockMesh numbers them, with one patch on each
side.

The reference reads any polyMesh through OpenFOAM (fvMesh, polyMesh.C); this
engine tracks on a Cartesian block, so a polyMesh is accepted when it is
one, and rejected with the reason otherwise.  The result carries what the
engine's configuration needs: cells per direction, origin, spacing, and the
name and type of the patch on each of the six sides (x-, x+, y-, y+, z-, z+).
"""
import os
import re

import numpy as np

from .lagrangian import _body, _list_items, _uniform

SIDES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def read_points(path):
    n, inner = _list_items(_body(path))
    a = np.array(inner.replace("(", " ").replace(")", " ").split(), np.float64).reshape(-1, 3)
    if len(a) != n:
        raise ValueError(f"{path}: {len(a)} points for a list of {n}")
    return a


def read_faces(path):
    """faceList "k(p0 .. pk-1)" as an (N, 4) array; every face must be a quad."""
    n, inner = _list_items(_body(path))
    faces = []
    for m in re.finditer(r"(\d+)\s*\(([^)]*)\)", inner):
        v = [int(x) for x in m.group(2).split()]
        if int(m.group(1)) != 4 or len(v) != 4:
            raise ValueError(f"{path}: a face of {len(v)} points; only hex cells are read")
        faces.append(v)
    if len(faces) != n:
        raise ValueError(f"{path}: {len(faces)} faces for a list of {n}")
    return np.array(faces, np.int64)


def read_labels(path):
    body = _body(path)
    u = _uniform(body)
    if u:
        return np.full(u[0], int(u[1]), np.int64)
    n, inner = _list_items(body)
    a = np.array(inner.split(), np.int64)
    if len(a) != n:
        raise ValueError(f"{path}: {len(a)} labels for a list of {n}")
    return a


def read_boundary(path):
    """[(name, type, nFaces, startFace)] in file order."""
    n, inner = _list_items(_body(path))
    out = []
    for m in re.finditer(r"(\w+)\s*\{([^}]*)\}", inner):
        d = dict(re.findall(r"(\w+)\s+([^;]+);", m.group(2)))
        out.append((m.group(1), d["type"].strip(), int(d["nFaces"]), int(d["startFace"])))
    if len(out) != n:
        raise ValueError(f"{path}: {len(out)} patches for a list of {n}")
    return out


def _axis_lattice(x, what):
    """(count, origin, spacing) of cell centres x on a uniform lattice.
    Centres are grouped into planes by a tolerance relative to the span of
    the coordinates, and the gaps between planes must agree to a relative
    1e-9 (OpenFOAM writes points with about 1e-10 relative precision)."""
    u = np.sort(np.asarray(x, np.float64))
    span = u[-1] - u[0]
    if span <= 1e-12 * max(abs(u[0]), abs(u[-1]), 1e-300):
        return 1, None, None
    # planes: a gap wider than 1e-7 of the span starts a new plane (uniform
    # lattices of up to ten million cells a side are resolved)
    breaks = np.flatnonzero(np.diff(u) > 1e-7 * span)
    groups = np.split(u, breaks + 1)
    planes = np.array([g.mean() for g in groups])
    for g in groups:
        if g[-1] - g[0] > 1e-7 * span:
            raise ValueError(f"{what}: cell centres do not lie on planes")
    d = np.diff(planes)
    mean = d.mean()
    if np.abs(d - mean).max() > 1e-9 * abs(mean) + 1e-12 * span:
        raise ValueError(f"{what}: cells are not evenly spaced (simpleGrading other than 1)")
    return len(planes), planes[0] - 0.5 * mean, mean


def read_block(case_dir):
    """The Cartesian block of <case_dir>/constant/polyMesh:
    dict(n=(nx, ny, nz), origin=(x0, y0, z0), spacing=(dx, dy, dz),
    sides={side: (patch, type)}).  ValueError names what does not fit."""
    d = os.path.join(case_dir, "constant", "polyMesh")
    pts = read_points(os.path.join(d, "points"))
    faces = read_faces(os.path.join(d, "faces"))
    owner = read_labels(os.path.join(d, "owner"))
    neigh = read_labels(os.path.join(d, "neighbour"))
    patches = read_boundary(os.path.join(d, "boundary"))
    nf, nc = len(faces), int(owner.max()) + 1
    if len(owner) != nf:
        raise ValueError(f"{d}: {len(owner)} owners for {nf} faces")
    fc = pts[faces].mean(axis=1)                       # face centres
    # cell centres: the mean of the six face centres (exact for a box)
    cnt = np.bincount(owner, minlength=nc) + np.bincount(neigh, minlength=nc)
    if (cnt != 6).any():
        raise ValueError(f"{d}: cells with {sorted(set(cnt.tolist()) - {6})} faces; only hex cells are read")
    cc = np.zeros((nc, 3))
    for a in range(3):
        cc[:, a] = (np.bincount(owner, fc[:, a], nc) + np.bincount(neigh, fc[:len(neigh), a], nc)) / 6.0
    n, origin, spacing = [], [], []
    for a, name in enumerate("xyz"):
        k, _, _ = _axis_lattice(cc[:, a], f"{d} ({name})")
        lo_a, hi_a = float(pts[:, a].min()), float(pts[:, a].max())   # exact, from the points
        n.append(k)
        origin.append(lo_a)
        spacing.append((hi_a - lo_a) / k)
    if n[0] * n[1] * n[2] != nc:
        raise ValueError(f"{d}: {nc} cells do not fill a {n[0]}x{n[1]}x{n[2]} block")
    ijk = np.floor((cc - np.array(origin)) / np.array(spacing)).astype(np.int64)
    want = ijk[:, 0] + n[0] * (ijk[:, 1] + n[1] * ijk[:, 2])
    if (want != np.arange(nc)).any():
        raise ValueError(f"{d}: the cell numbering is not blockMesh's i + nx*(j + ny*k)")
    lo = np.array(origin)
    hi = lo + np.array(n) * np.array(spacing)
    sides = {}
    for name, ptype, cnt_p, start in patches:
        if cnt_p == 0:
            continue
        if start < len(neigh) or start + cnt_p > nf:
            raise ValueError(f"{d}: patch {name} holds faces outside the boundary faces")
        c = fc[start:start + cnt_p]
        on_sides = 0
        for a in range(3):
            for s, plane in ((0, lo[a]), (1, hi[a])):
                on = np.abs(c[:, a] - plane) <= 1e-9 * max(1.0, abs(plane)) + 1e-12 * spacing[a]
                if not on.any():
                    continue
                side = SIDES[2 * a + s]
                if side in sides and sides[side][0] != name:
                    raise ValueError(f"{d}: side {side} holds patches {sides[side][0]} and {name}")
                sides[side] = (name, ptype)
                need = nc // n[a]
                if on.sum() != need:
                    raise ValueError(f"{d}: patch {name} covers {on.sum()} of the {need} faces of side {side}")
                on_sides += int(on.sum())
        if on_sides == 0:
            raise ValueError(f"{d}: patch {name} lies on no side of the block (an internal body?)")
        if on_sides != cnt_p:
            raise ValueError(f"{d}: patch {name} has {cnt_p - on_sides} faces inside the block")
    missing = [s for s in SIDES if s not in sides]
    if missing:
        raise ValueError(f"{d}: sides {missing} have no patch")
    return dict(n=tuple(n), origin=tuple(origin), spacing=tuple(spacing), sides=sides, patches=patches)
