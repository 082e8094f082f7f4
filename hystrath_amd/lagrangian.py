"""dsmcFoam+'s lagrangian cloud files (<time>/lagrangian/dsmc/) in OpenFOAM
ASCII format: the data format on either side of the particle engine.

The files a dsmcParcel cloud writes (dsmcParcelIO.C:133-519, Cloud<dsmcParcel>
through IOPosition) are read into the engine's structure-of-arrays parcel
layout and written back from it:

    positions       "(x y z) celli" per parcel (class Cloud<passiveParticle>)
    U               vectorField
    ERot            scalarField
    typeId          labelField (index into constant/dsmcProperties typeIdList)
    vibLevel        labelFieldField: "k(l0 .. lk-1)" per parcel
    newParcel, classification, origProcId, origId   labelFields (written for
                    OpenFOAM's readers; newParcel -1: not injected this step)

Lists may be written in full ("N ( ... )") or uniform ("N{value}").
"""
import os
import re

import numpy as np

_HEADER = """/*--------------------------------*- C++ -*----------------------------------*\\
| =========                 |                                                 |
| \\\\      /  F ield         | OpenFOAM: The Open Source CFD Toolbox           |
|  \\\\    /   O peration     |                                                 |
|   \\\\  /    A nd           |                                                 |
|    \\\\/     M anipulation  |                                                 |
\\*---------------------------------------------------------------------------*/
FoamFile
{{
    version     2.0;
    format      ascii;
    class       {cls};
    location    "{loc}";
    object      {obj};
}}
// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //

"""
_FOOTER = "\n// ************************************************************************* //\n"


def _body(path):
    """The text after the FoamFile header, comments removed."""
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    i = txt.find("FoamFile")
    if i >= 0:
        txt = txt[txt.index("}", i) + 1:]
    return txt.strip()


def _uniform(body):
    """(count, value) of a uniform list "N{value}", else None."""
    m = re.match(r"^(\d+)\s*\{\s*([^}]*)\}", body)
    return (int(m.group(1)), m.group(2).strip()) if m else None


def _list_items(body):
    """count and the text between the list's outer parentheses."""
    m = re.match(r"^(\d+)\s*\(", body)
    if not m:
        raise ValueError("not an OpenFOAM list")
    n = int(m.group(1))
    inner = body[m.end():body.rindex(")")]
    return n, inner


def read_scalars(path, dtype=np.float64):
    body = _body(path)
    u = _uniform(body)
    if u:
        return np.full(u[0], float(u[1]), dtype)
    n, inner = _list_items(body)
    a = np.array(inner.split(), dtype=np.float64).astype(dtype)
    if len(a) != n:
        raise ValueError(f"{path}: {len(a)} values for a list of {n}")
    return a


def read_vectors(path):
    body = _body(path)
    u = _uniform(body)
    if u:
        v = [float(x) for x in u[1].strip("()").split()]
        return np.tile(np.array(v, np.float64), (u[0], 1))
    n, inner = _list_items(body)
    a = np.array(inner.replace("(", " ").replace(")", " ").split(), np.float64).reshape(-1, 3)
    if len(a) != n:
        raise ValueError(f"{path}: {len(a)} vectors for a list of {n}")
    return a


def read_positions(path):
    """positions: (N, 3) coordinates and the cell label of each parcel."""
    n, inner = _list_items(_body(path))
    tok = inner.replace("(", " ").replace(")", " ").split()
    a = np.array(tok, np.float64).reshape(-1, 4)
    if len(a) != n:
        raise ValueError(f"{path}: {len(a)} positions for a list of {n}")
    return a[:, :3].copy(), a[:, 3].astype(np.int32)


def read_label_lists(path, width):
    """A labelFieldField ("k(l ...)" per parcel) as an (N, width) array, each
    list padded with zeros."""
    body = _body(path)
    u = _uniform(body)
    if u:
        vals = [int(x) for x in re.findall(r"-?\d+", u[1])[1:]]
        out = np.zeros((u[0], width), np.int64)
        out[:, :len(vals)] = vals[:width]
        return out
    n, inner = _list_items(body)
    out = np.zeros((n, width), np.int64)
    # each entry is "k(l0 l1 ...)" or OpenFOAM's uniform short form "k{l}"
    k = -1
    for k, m in enumerate(re.finditer(r"(\d+)\s*(?:\(([^)]*)\)|\{\s*(-?\d+)\s*\})", inner)):
        if k >= n:
            raise ValueError(f"{path}: more than {n} entries")
        cnt = int(m.group(1))
        vals = [int(x) for x in m.group(2).split()] if m.group(2) is not None else [int(m.group(3))] * cnt
        if len(vals) != cnt:
            raise ValueError(f"{path}: entry {k} has {len(vals)} levels, not {cnt}")
        out[k, :min(width, len(vals))] = vals[:width]
    if k + 1 != n:
        raise ValueError(f"{path}: {k + 1} entries for a list of {n}")
    return out


def read_cloud(time_dir, vib_modes=1):
    """The cloud of <time_dir>/lagrangian/dsmc as engine arrays: px, py, pz,
    cell, ux, uy, uz, erot, vib (modes x N) and type."""
    d = os.path.join(time_dir, "lagrangian", "dsmc")
    xyz, cell = read_positions(os.path.join(d, "positions"))
    n = len(cell)
    U = read_vectors(os.path.join(d, "U"))
    if len(U) != n:
        raise ValueError(f"{d}/U: {len(U)} entries for {n} parcels")
    out = dict(px=xyz[:, 0].copy(), py=xyz[:, 1].copy(), pz=xyz[:, 2].copy(), cell=cell,
               ux=U[:, 0].astype(np.float32), uy=U[:, 1].astype(np.float32), uz=U[:, 2].astype(np.float32))
    f = os.path.join(d, "ERot")
    out["erot"] = read_scalars(f, np.float32) if os.path.exists(f) else np.zeros(n, np.float32)
    f = os.path.join(d, "typeId")
    out["type"] = read_scalars(f, np.float64).astype(np.uint8) if os.path.exists(f) else np.zeros(n, np.uint8)
    f = os.path.join(d, "vibLevel")
    vib = read_label_lists(f, vib_modes) if os.path.exists(f) else np.zeros((n, vib_modes), np.int64)
    out["vib"] = np.ascontiguousarray(vib.T.astype(np.uint16))
    for k in ("erot", "type"):
        if len(out[k]) != n:
            raise ValueError(f"{d}: {len(out[k])} {k} entries for {n} parcels")
    return out


def _write(path, cls, loc, obj, body):
    with open(path, "w") as f:
        f.write(_HEADER.format(cls=cls, loc=loc, obj=obj) + body + _FOOTER)


def _g(x):
    return repr(float(x))


def write_cloud(time_dir, parcels, time_name=None):
    """Writes the engine arrays (as read_cloud returns them, or as
    Engine.download gives them) to <time_dir>/lagrangian/dsmc.  Positions are
    written with full double precision so that a read gives the same values."""
    d = os.path.join(time_dir, "lagrangian", "dsmc")
    os.makedirs(d, exist_ok=True)
    loc = f"{time_name or os.path.basename(os.path.normpath(time_dir))}/lagrangian/dsmc"
    n = len(parcels["px"])
    px, py, pz, cell = parcels["px"], parcels["py"], parcels["pz"], parcels["cell"]
    lines = "\n".join(f"({_g(px[k])} {_g(py[k])} {_g(pz[k])}) {int(cell[k])}" for k in range(n))
    _write(os.path.join(d, "positions"), "Cloud<passiveParticle>", loc, "positions", f"{n}\n(\n{lines}\n)\n")
    ux, uy, uz = parcels["ux"], parcels["uy"], parcels["uz"]
    lines = "\n".join(f"({_g(ux[k])} {_g(uy[k])} {_g(uz[k])})" for k in range(n))
    _write(os.path.join(d, "U"), "vectorField", loc, "U", f"{n}\n(\n{lines}\n)\n")
    erot = parcels.get("erot", np.zeros(n, np.float32))
    _write(os.path.join(d, "ERot"), "scalarField", loc, "ERot",
           f"{n}\n(\n" + "\n".join(_g(x) for x in erot[:n]) + "\n)\n")
    tp = parcels.get("type", np.zeros(n, np.uint8))
    _write(os.path.join(d, "typeId"), "labelField", loc, "typeId",
           f"{n}\n(\n" + "\n".join(str(int(x)) for x in tp[:n]) + "\n)\n")
    vib = parcels.get("vib")
    if vib is None:
        vib = np.zeros((1, n), np.uint16)
    vib = np.asarray(vib).reshape(-1, n)
    m = vib.shape[0]
    lines = "\n".join(f"{m}(" + " ".join(str(int(vib[q, k])) for q in range(m)) + ")" for k in range(n))
    _write(os.path.join(d, "vibLevel"), "labelFieldField", loc, "vibLevel", f"{n}\n(\n{lines}\n)\n")
    for obj, val in (("newParcel", -1), ("classification", 0), ("origProcId", 0)):
        _write(os.path.join(d, obj), "labelField", loc, obj, f"{n}{{{val}}}\n")
    _write(os.path.join(d, "origId"), "labelField", loc, "origId",
           f"{n}\n(\n" + "\n".join(str(k) for k in range(n)) + "\n)\n")
    return d
