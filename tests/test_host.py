"""The C++ host mirror of dsmcCloud (hystrath_amd/host): dictionary reading,
runtime selection by the reference's model names, FatalError behaviour, and
(on a GPU) the dsmcInitialisePlus -> dsmcFoamPlus pipeline with restart."""
import glob
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from hystrath_amd import casegen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hystrath_amd", "bin")
REF = "/root/reference/run/hyStrath/dsmcFoam+"


TOOLS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tools", "_bin")


def run(app, case, timeout=300):
    return subprocess.run([os.path.join(BIN, app), "-case", case], capture_output=True, text=True, timeout=timeout)


def no_gpu():
    from hystrath_amd import engine
    return engine.device_count() == 0


@pytest.mark.skipif(not os.path.isdir(REF), reason="hyStrath tutorials not present")
def test_reads_every_reference_tutorial_dictionary():
    files = []
    for c in sorted(os.listdir(REF)):
        for rel in ("constant/dsmcProperties", "system/boundariesDict", "system/fieldPropertiesDict",
                    "system/controlDict", "system/dsmcInitialiseDict", "system/blockMeshDict", "system/chemReactDict"):
            p = os.path.join(REF, c, rel)
            if os.path.exists(p):
                files.append(p)
    assert len(files) > 25
    r = subprocess.run([os.path.join(TOOLS, "dictCheck")] + files, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "BinaryCollisionModel" in r.stdout and "dsmcPatchBoundaries" in r.stdout


@pytest.mark.skipif(not os.path.isdir(REF), reason="hyStrath tutorials not present")
def test_reference_case_outcomes(tmp_path):
    """couette_N2-O2 (its mesh read from its constant/polyMesh) and
    heatBath-5species (its twelve QK reactions read from system/chemReactDict)
    are accepted up to the device; the multi-block tutorials stop with the
    reason they are outside this engine."""
    outcomes = {}
    for c in ("couette_N2-O2", "heatBath-5species", "hypersonicCorner", "supersonicFlatPlate"):
        d = tmp_path / c
        shutil.copytree(os.path.join(REF, c, "system"), d / "system")
        shutil.copytree(os.path.join(REF, c, "constant"), d / "constant")
        r = run("dsmcInitialisePlus", str(d))
        outcomes[c] = r.stdout + r.stderr
    assert "one block" in outcomes["hypersonicCorner"] and "one block" in outcomes["supersonicFlatPlate"]
    if no_gpu():
        assert "no HIP device" in outcomes["couette_N2-O2"]
        assert "no HIP device" in outcomes["heatBath-5species"]


def test_reaction_dictionary_errors(tmp_path):
    """chemReactDict entries are selected by reactionModel and checked as the
    reference checks them."""
    case = casegen.heat_bath_case(str(tmp_path / "hb"), fn=1e5)
    p = os.path.join(case, "system/chemReactDict")
    txt = open(p).read()
    open(p, "w").write(txt.replace("reactionModel dissociationQK;", "reactionModel ionisationQK;", 1))
    r = run("dsmcInitialisePlus", case)
    assert r.returncode == 1 and "unknown dsmc reaction model type ionisationQK" in r.stderr
    open(p, "w").write(txt.replace("reactants (O2 N2);", "reactants (O2 Ar);", 1))
    r = run("dsmcInitialisePlus", case)
    assert r.returncode == 1 and "Cannot find type id: Ar" in r.stderr
    open(p, "w").write(txt.replace("dissociationProducts ((O O) (N N));", "dissociationProducts ((O O));", 1))
    r = run("dsmcInitialisePlus", case)
    assert r.returncode == 1 and "two lists of dissociation products" in r.stderr
    open(p, "w").write(txt.replace("dissociationProducts ((O O) (N N));", "dissociationProducts ((O O) (N NO));", 1))
    r = run("dsmcInitialisePlus", case)
    assert r.returncode == 1 and "must be an atom" in r.stderr


@pytest.mark.gpu
def test_heat_bath_pipeline_follows_the_reference_run(tmp_path):
    """dsmcInitialisePlus and dsmcFoamPlus on the heatBath-5species set-up
    (nEquivalentParticles 2000 instead of 100: about 7,700 parcels) for 100
    steps: the written averages of the species number densities and the
    mixture temperature over those steps follow the reference's probe
    series averaged over the same steps (tests/golden/heatbath_probes.npz)."""
    case = casegen.heat_bath_case(str(tmp_path / "hb"), fn=2000.0, end_time=1e-7, write_interval=1e-7)
    r = run("dsmcInitialisePlus", case)
    assert r.returncode == 0, r.stderr
    r = run("dsmcFoamPlus", case)
    assert r.returncode == 0, r.stderr
    last = sorted((d for d in os.listdir(case) if re.match(r"^[0-9.e+-]+$", d) and float(d) > 0), key=float)[-1]
    assert abs(float(last) - 1e-7) < 1e-12
    fx = np.load(os.path.join(ROOT, "tests", "golden", "heatbath_probes.npz"))
    first = fx["step"] <= 100
    for sp, tol in (("N2", 0.05), ("O2", 0.08), ("N", 0.08), ("O", 0.08)):
        got = float(_field_or_uniform(os.path.join(case, last, "rhoN_" + sp)))
        want = float(fx["rhoN_" + sp][first].mean())
        assert abs(got / want - 1.0) < tol, (sp, got, want)
    T = float(_field_or_uniform(os.path.join(case, last, "Ttra_mixture")))
    want = float(fx["Ttra_mixture"][first].mean())
    assert abs(T / want - 1.0) < 0.03, (T, want)


def _field_or_uniform(path):
    txt = open(path).read()
    m = re.search(r"internalField\s+uniform\s+([-0-9.eE+]+)\s*;", txt)
    if m:
        return float(m.group(1))
    return _field(path)[0]


def test_runtime_selection_errors(tmp_path):
    case = casegen.couette_case(str(tmp_path / "c"))
    p = os.path.join(case, "constant/dsmcProperties")
    txt = open(p).read()
    open(p, "w").write(txt.replace("LarsenBorgnakkeVariableHardSphere;", "MaxwellMolecule;"))
    r = run("dsmcInitialisePlus", case)
    assert r.returncode == 1
    assert "Unknown BinaryCollisionModel type MaxwellMolecule" in r.stderr
    assert "VariableSoftSphere" in r.stderr  # lists the valid types, as the reference does
    open(p, "w").write(txt)
    b = os.path.join(case, "system/boundariesDict")
    btxt = open(b).read()
    open(b, "w").write(btxt.replace("dsmcDiffuseWallPatch;", "dsmcCLLWallPatch;", 1))
    r = run("dsmcInitialisePlus", case)
    assert r.returncode == 1 and "unknown dsmcPatchBoundary type dsmcCLLWallPatch" in r.stderr


def test_generated_cases_parse(tmp_path):
    for maker in (casegen.couette_case, casegen.cylinder_case):
        case = maker(str(tmp_path / maker.__name__))
        files = glob.glob(os.path.join(case, "*/*Dict")) + glob.glob(os.path.join(case, "constant/*"))
        r = subprocess.run([os.path.join(TOOLS, "dictCheck")] + files, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout
        if no_gpu():
            r = run("dsmcInitialisePlus", case)
            assert "no HIP device" in r.stderr


def _patch_values(path, patch):
    txt = open(path).read()
    m = re.search(patch + r"\s*\{\s*type\s+calculated;\s*value\s+nonuniform List<(\w+)>\s*\n(\d+)\n\(\n(.*?)\n\)\n;",
                  txt, re.S)
    rows = m.group(3).split("\n")
    if m.group(1) == "vector":
        return np.array([[float(x) for x in r.strip("()").split()] for r in rows])
    return np.array([float(x) for x in rows])


def _field(path):
    txt = open(path).read()
    m = re.search(r"internalField\s+nonuniform List<(\w+)>\s*\n(\d+)\n\(\n(.*?)\n\)\n;", txt, re.S)
    rows = m.group(3).split("\n")
    if m.group(1) == "vector":
        return np.array([[float(x) for x in r.strip("()").split()] for r in rows])
    return np.array([float(x) for x in rows])


@pytest.mark.gpu
def test_couette_pipeline_and_restart(tmp_path):
    case = casegen.couette_case(str(tmp_path / "couette"), end_time=4e-4, write_interval=2e-4, n_terminal=10)
    r = run("dsmcInitialisePlus", case)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(os.path.join(case, "0/lagrangian/dsmc/hystrath.state"))
    r = run("dsmcFoamPlus", case)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("Number of DSMC particles") == 4, r.stdout
    last = os.path.join(case, "0.0004")
    for f in ("rhoN_mixture", "Ttra_mixture", "Trot_N2", "U_mixture", "Tov_O2", "measuredCollisionRate_mixture"):
        assert os.path.exists(os.path.join(last, f)), f
    rho = _field(os.path.join(last, "rhoN_mixture"))
    assert rho.shape == (500,)
    assert abs(rho.mean() / 4.09e19 - 1) < 0.05  # 3.23e19 N2 + 0.86e19 O2, walls conserve parcels
    T = _field(os.path.join(last, "Ttra_mixture"))
    assert 1500 < T.mean() < 3500
    U = _field(os.path.join(last, "U_mixture"))
    assert U.shape == (500, 3)
    # wall values (boundaryMeasurements): 5 faces per wall
    q_up = _patch_values(os.path.join(last, "wallHeatFlux_mixture"), "upperWall")
    q_lo = _patch_values(os.path.join(last, "wallHeatFlux_mixture"), "lowerWall")
    assert q_up.shape == q_lo.shape == (5,)
    assert q_up.mean() < 0          # the 3000 K wall heats the 2000 K gas
    p_lo = _patch_values(os.path.join(last, "p_mixture"), "lowerWall")
    kT = 1.38064852e-23 * 2000.0
    assert 0.5 * 4.09e19 * kT < p_lo.mean() < 2.0 * 4.09e19 * kT
    fD = _patch_values(os.path.join(last, "fD_mixture"), "upperWall")
    assert fD.shape == (5, 3) and (fD[:, 1] > 0).all()  # the gas pushes the upper wall outward (+y)
    # restart from the latest time and continue to a later end time
    cd = os.path.join(case, "system/controlDict")
    ctxt = open(cd).read()
    assert "endTime 0.0004;" in ctxt
    open(cd, "w").write(ctxt.replace("endTime 0.0004;", "endTime 0.0006;"))
    r = run("dsmcFoamPlus", case)
    assert r.returncode == 0, r.stderr
    assert "from" in r.stdout and "0.0004" in r.stdout
    assert os.path.exists(os.path.join(case, "0.0006", "Ttra_mixture"))


COUETTE_PATCHES = [("periodicX_half0", "cyclic", ["xmin"]), ("periodicX_half1", "cyclic", ["xmax"]),
                   ("lowerWall", "wall", ["ymin"]), ("upperWall", "wall", ["ymax"]),
                   ("frontAndBack", "empty", ["zmin", "zmax"])]


def _couette_from_polymesh(root, **kw):
    """The generated couette case with its mesh given as constant/polyMesh
    only (blockMeshDict removed), patches in the blockMeshDict's order."""
    from test_polymesh import write_block_mesh
    case = casegen.couette_case(root, **kw)
    write_block_mesh(case, (5, 100, 1), (0.0, 0.0, 0.0), (0.05, 1.0, 0.01), COUETTE_PATCHES)
    os.remove(os.path.join(case, "system", "blockMeshDict"))
    return case


def test_mesh_is_read_from_constant_polymesh(tmp_path):
    """With constant/polyMesh present the host reads the mesh from it (as
    fvMesh does); a polyMesh that is not one uniform block stops the run with
    the reason."""
    case = _couette_from_polymesh(str(tmp_path / "c"))
    r = run("dsmcInitialisePlus", case)
    if no_gpu():
        assert r.returncode == 1 and "no HIP device" in r.stderr, r.stdout + r.stderr
    else:
        assert r.returncode == 0, r.stdout + r.stderr
    from test_polymesh import write_block_mesh
    bad = casegen.couette_case(str(tmp_path / "g"))
    xs = np.concatenate([[0.0], np.cumsum([0.005, 0.01, 0.01, 0.01, 0.015])])
    write_block_mesh(bad, (5, 100, 1), (0.0, 0.0, 0.0), (0.05, 1.0, 0.01), COUETTE_PATCHES, xs=xs)
    r = run("dsmcInitialisePlus", bad)
    assert r.returncode == 1 and "constant/polyMesh (x): cells are not evenly spaced" in r.stderr, r.stderr


@pytest.mark.gpu
def test_polymesh_and_blockmeshdict_runs_are_identical(tmp_path):
    """The couette case run from its blockMeshDict and from the equivalent
    polyMesh writes identical fields and parcels."""
    a = casegen.couette_case(str(tmp_path / "a"), end_time=2e-4, write_interval=1e-4, n_terminal=10)
    b = _couette_from_polymesh(str(tmp_path / "b"), end_time=2e-4, write_interval=1e-4, n_terminal=10)
    for c in (a, b):
        for app in ("dsmcInitialisePlus", "dsmcFoamPlus"):
            r = run(app, c)
            assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    names = sorted(f for f in os.listdir(os.path.join(a, "0.0002")) if os.path.isfile(os.path.join(a, "0.0002", f)))
    assert {"Ttra_mixture", "rhoN_N2", "U_mixture", "wallHeatFlux_mixture"} <= set(names)
    for f in names:
        assert open(os.path.join(a, "0.0002", f)).read() == open(os.path.join(b, "0.0002", f)).read(), f
    sa = _read_state(os.path.join(a, "0.0002/lagrangian/dsmc/hystrath.state"))
    sb = _read_state(os.path.join(b, "0.0002/lagrangian/dsmc/hystrath.state"))
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


@pytest.mark.gpu
def test_zv_2008_case_runs_with_the_first_fields_temperature(tmp_path):
    """A couette case with inverseZvFormulation "2008": the host passes the
    first field's overall temperature to the engine at each write, and the run
    differs from the pre-2008 one only after the first write."""
    runs = {}
    for form in ("pre-2008", "2008"):
        case = casegen.couette_case(str(tmp_path / form), end_time=4e-4, write_interval=2e-4, n_terminal=10)
        p = os.path.join(case, "constant", "dsmcProperties")
        txt = open(p).read()
        assert 'inverseZvFormulation "pre-2008";' in txt
        open(p, "w").write(txt.replace('"pre-2008"', f'"{form}"'))
        for app in ("dsmcInitialisePlus", "dsmcFoamPlus"):
            r = run(app, case)
            assert r.returncode == 0, r.stdout[-2000:] + r.stderr
        runs[form] = case
    first = [open(os.path.join(runs[f], "0.0002", "Tvib_N2")).read() for f in runs]
    second = [open(os.path.join(runs[f], "0.0004", "Tvib_N2")).read() for f in runs]
    assert first[0] == first[1]
    assert second[0] != second[1]


@pytest.mark.gpu
def test_restart_continues_the_average(tmp_path):
    """A run of 60 steps written in one go equals 40 steps, a restart from
    the checkpoint and 20 more: the particles and the cumulative cell and
    wall sums (resetAtOutputUntilTime 0: one average from the start) carry
    over, so every field file of the last write is identical."""
    one = casegen.couette_case(str(tmp_path / "one"), end_time=6e-4, write_interval=2e-4, n_terminal=10)
    two = casegen.couette_case(str(tmp_path / "two"), end_time=4e-4, write_interval=2e-4, n_terminal=10)
    for c in (one, two):
        r = run("dsmcInitialisePlus", c)
        assert r.returncode == 0, r.stderr
        r = run("dsmcFoamPlus", c)
        assert r.returncode == 0, r.stderr
    cd = os.path.join(two, "system/controlDict")
    ctxt = open(cd).read()
    assert "endTime 0.0004;" in ctxt
    with open(cd, "w") as f:
        f.write(ctxt.replace("endTime 0.0004;", "endTime 0.0006;"))
    r = run("dsmcFoamPlus", two)
    assert r.returncode == 0, r.stderr
    names = sorted(f for f in os.listdir(os.path.join(one, "0.0006")) if os.path.isfile(os.path.join(one, "0.0006", f)))
    assert {"Ttra_mixture", "rhoN_N2", "U_mixture", "wallHeatFlux_mixture", "Tvib_O2"} <= set(names)
    for f in names:
        a = open(os.path.join(one, "0.0006", f)).read()
        b = open(os.path.join(two, "0.0006", f)).read()
        assert a == b, f
    sa = _read_state(os.path.join(one, "0.0006/lagrangian/dsmc/hystrath.state"))
    sb = _read_state(os.path.join(two, "0.0006/lagrangian/dsmc/hystrath.state"))
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


def _read_state(path):
    """Parcels of an engine checkpoint (hystrath.state): magic, eight int64
    (n, cells, species, vib modes, rotational, inflow faces, step, samples),
    the time, then the parcel arrays in structure-of-arrays order."""
    b = open(path, "rb").read()
    assert b[:8] == b"HSDSMC02"
    hdr = np.frombuffer(b, np.int64, 8, 8)
    n, ns, modes, rot = int(hdr[0]), int(hdr[2]), int(hdr[3]), int(hdr[4])
    off = 8 + 64 + 8
    out = {}

    def take(name, dt, count):
        nonlocal off
        out[name] = np.frombuffer(b, dt, count, off)
        off += np.dtype(dt).itemsize * count

    for k in ("px", "py", "pz"):
        take(k, np.float64, n)
    for k in ("ux", "uy", "uz"):
        take(k, np.float32, n)
    take("cell", np.int32, n)
    if rot:
        take("erot", np.float32, n)
    if modes:
        take("vib", np.uint16, modes * n)
        out["vib"] = out["vib"].reshape(modes, n)
    if ns > 1:
        take("type", np.uint8, n)
    return out


@pytest.mark.gpu
def test_dsmcfoam_starts_from_a_lagrangian_cloud(tmp_path):
    """dsmcFoamPlus started from OpenFOAM cloud files (the first 8,192 parcels
    of the couette_N2-O2 tutorial's snapshot, written by the Python writer) and
    a dsmcSigmaTcRMax field: the walls and cyclic sides keep every parcel and
    both species' counts, the cloud written at each write time holds exactly
    the parcels of the engine checkpoint beside it, every parcel lies in the
    cell it is filed under, and a cloud newer than the latest checkpoint is
    the one a restart reads."""
    from hystrath_amd import lagrangian as lg
    case = casegen.couette_case(str(tmp_path / "couette"), end_time=2e-4, write_interval=1e-4, n_terminal=10)
    h = np.load(os.path.join(ROOT, "tests", "golden", "couette_cloud_head.npz"))
    xyz = h["xyz"]
    parcels = dict(px=xyz[:, 0], py=xyz[:, 1], pz=xyz[:, 2], cell=h["cell"].astype(np.int32),
                   ux=h["U"][:, 0], uy=h["U"][:, 1], uz=h["U"][:, 2], erot=h["erot"], type=h["type"],
                   vib=h["vib"].reshape(1, -1))
    n = len(parcels["px"])
    lg.write_cloud(os.path.join(case, "0"), parcels, "0")
    with open(os.path.join(case, "0", "dsmcSigmaTcRMax"), "w") as f:
        f.write("FoamFile\n{\n    version 2.0;\n    format ascii;\n    class volScalarField;\n"
                "    object dsmcSigmaTcRMax;\n}\ndimensions [0 3 -1 0 0 0 0];\n"
                "internalField   uniform 1.2e-15;\nboundaryField\n{\n}\n")
    r = run("dsmcFoamPlus", case)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert f"Read {n} parcels at time 0 from" in r.stdout and "lagrangian/dsmc" in r.stdout, r.stdout[:2000]
    n_type = np.bincount(parcels["type"], minlength=2)
    for t in ("0.0001", "0.0002"):
        d = os.path.join(case, t)
        c = lg.read_cloud(d, vib_modes=1)
        assert len(c["px"]) == n, t
        np.testing.assert_array_equal(np.bincount(c["type"], minlength=2), n_type)
        s = _read_state(os.path.join(d, "lagrangian", "dsmc", "hystrath.state"))
        for k in ("px", "py", "pz", "cell", "ux", "uy", "uz", "erot", "vib", "type"):
            np.testing.assert_array_equal(c[k], s[k], err_msg=f"{t}/{k}")
        i = np.floor(c["px"] / 0.01).astype(int)
        j = np.floor(c["py"] / 0.01).astype(int)
        np.testing.assert_array_equal(c["cell"], i + 5 * j)
        sig = _field(os.path.join(d, "dsmcSigmaTcRMax"))
        assert sig.shape == (500,) and (sig >= 1.2e-15 * (1 - 1e-9)).all()
    # a cloud at a later time than the latest checkpoint is where a restart starts
    later = os.path.join(case, "0.00025")
    shutil.copytree(os.path.join(case, "0.0002", "lagrangian"), os.path.join(later, "lagrangian"),
                    ignore=shutil.ignore_patterns("hystrath.state"))
    shutil.copy(os.path.join(case, "0.0002", "dsmcSigmaTcRMax"), later)
    cd = os.path.join(case, "system/controlDict")
    # write times are counted from the start time: 0.00025 + 1e-4
    ctxt = open(cd).read()
    assert "endTime 0.0002;" in ctxt
    open(cd, "w").write(ctxt.replace("endTime 0.0002;", "endTime 0.00035;"))
    r = run("dsmcFoamPlus", case)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert f"Read {n} parcels at time 0.00025 from" in r.stdout, r.stdout[:2000]
    assert len(lg.read_cloud(os.path.join(case, "0.00035"), 1)["px"]) == n
