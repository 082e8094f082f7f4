"""Writes dsmcFoam+ case directories (OpenFOAM dictionary format) for the
engine's host applications (dsmcInitialisePlus, dsmcFoamPlus).

The files use the reference's dictionary names and keywords:
constant/dsmcProperties, system/{controlDict, blockMeshDict, boundariesDict,
fieldPropertiesDict, dsmcInitialiseDict, chemReactDict} plus
system/embeddedCylinderDict for a cut-cell cylinder body.
"""
import os

HEADER = """FoamFile
{{
    version     2.0;
    format      ascii;
    class       dictionary;
    object      {obj};
}}

"""


def _w(path, obj, body):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(HEADER.format(obj=obj) + body)


def _species_block(name, p):
    lines = [f"    {name}", "    {"]
    for k in ("mass", "diameter", "omega", "alpha", "rotationalDegreesOfFreedom", "nVibrationalModes",
              "dissociationTemperature", "charge"):
        if k in p:
            lines.append(f"        {k:<40} {p[k]};")
    for k in ("characteristicVibrationalTemperature", "Zref", "referenceTempForZref"):
        if k in p:
            lines.append(f"        {k:<40} ({' '.join(str(x) for x in p[k])});")
    lines.append("    }")
    return "\n".join(lines)


def _reactions_text(reactions):
    """chemReactDict entries from (reactionModel, reactants,
    dissociationProducts, exchange) tuples (see cases.HEATBATH_REACTIONS)."""
    out = ""
    for k, (model, reac, diss, exch) in enumerate(reactions):
        body = f"        reactionModel {model};\n        reactants ({' '.join(reac)});\n        allowSplitting yes;\n"
        if model != "exchangeQK":
            lists = " ".join("(" + " ".join(d) + ")" for d in diss) if len(diss) == 2 else \
                "(" + " ".join(diss[0]) + ") ()"
            body += f"        dissociationQKProperties\n        {{\n            dissociationProducts ({lists});\n        }}\n"
        if exch is not None:
            prods, heat, a, b = exch
            body += (f"        exchangeQKProperties\n        {{\n            exchangeProducts ({' '.join(prods)});\n"
                     f"            heatOfReactionExchange {heat!r};\n            aCoeff {a!r};\n"
                     f"            bCoeff {b!r};\n        }}\n")
        out += f"    reaction{k}\n    {{\n{body}    }}\n"
    return out


def write_case(root, *, species, n, lo, hi, patches, fn, dt, end_time, write_interval, model="VariableHardSphere",
               coeffs=None, seed=1, fields=None, init=None, inflow=None, cylinder=None, n_terminal=50,
               reactions=None):
    """patches: {side: (name, type, model, props)} for side in xmin..zmax;
    type is the polyPatch type (wall, patch, cyclic, empty), model the
    dsmcPatchBoundaries model name or None; reactions: chemReactDict entries
    as in cases.HEATBATH_REACTIONS."""
    sp_list = " ".join(species)
    mol = "\n".join(_species_block(k, v) for k, v in species.items())
    coeff_txt = ""
    if coeffs:
        coeff_txt = f"{model}Coeffs\n{{\n" + "".join(f"    {k} {v};\n" for k, v in coeffs.items()) + "}\n\n"
    _w(os.path.join(root, "constant/dsmcProperties"), "dsmcProperties",
       f"nEquivalentParticles {fn!r};\n\nBinaryCollisionModel {model};\n\n{coeff_txt}"
       f"collisionPartnerSelectionModel noTimeCounter;\n\nseedNumber {seed};\n\n"
       f"typeIdList ({sp_list});\n\nmoleculeProperties\n{{\n{mol}\n}}\n")
    _w(os.path.join(root, "system/controlDict"), "controlDict",
       f"application dsmcFoam+;\nstartFrom latestTime;\nstartTime 0;\nstopAt endTime;\nendTime {end_time!r};\n"
       f"deltaT {dt!r};\nwriteControl runTime;\nwriteInterval {write_interval!r};\nnTerminalOutputs {n_terminal};\n"
       f"writePrecision 10;\n")
    v = [(lo[0], lo[1], lo[2]), (hi[0], lo[1], lo[2]), (hi[0], hi[1], lo[2]), (lo[0], hi[1], lo[2]),
         (lo[0], lo[1], hi[2]), (hi[0], lo[1], hi[2]), (hi[0], hi[1], hi[2]), (lo[0], hi[1], hi[2])]
    faces = {"xmin": "(0 4 7 3)", "xmax": "(1 2 6 5)", "ymin": "(0 1 5 4)", "ymax": "(3 7 6 2)",
             "zmin": "(0 3 2 1)", "zmax": "(4 5 6 7)"}
    by_name = {}
    for side, (name, ptype, _, _) in patches.items():
        by_name.setdefault(name, [ptype, []])[1].append(faces[side])
    bnd = ""
    for name, (ptype, fl) in by_name.items():
        bnd += f"    {name}\n    {{\n        type {ptype};\n        faces ({' '.join(fl)});\n    }}\n"
    verts = "\n".join(f"    ({a!r} {b!r} {c!r})" for a, b, c in v)
    nx, ny, nz = n
    _w(os.path.join(root, "system/blockMeshDict"), "blockMeshDict",
       f"convertToMeters 1;\n\nvertices\n(\n{verts}\n);\n\nblocks\n(\n    hex (0 1 2 3 4 5 6 7) ({nx} {ny} {nz}) "
       f"simpleGrading (1 1 1)\n);\n\nboundary\n(\n{bnd});\n")
    pb = ""
    seen = set()
    for side, (name, ptype, model_name, props) in patches.items():
        if not model_name or name in seen:
            continue
        seen.add(name)
        pr = "".join(f"            {k} {v};\n" for k, v in (props or {}).items())
        pb += (f"    boundary\n    {{\n        patchBoundaryProperties\n        {{\n            patchName {name};\n"
               f"        }}\n        boundaryModel {model_name};\n        {model_name}Properties\n        {{\n{pr}"
               f"        }}\n    }}\n")
    gb = ""
    if inflow:
        nd = "".join(f"                {k} {x!r};\n" for k, x in inflow["n"].items())
        for name in inflow["patches"]:
            gb += (f"    boundary\n    {{\n        generalBoundaryProperties\n        {{\n            patchName {name};\n"
                   f"        }}\n        boundaryModel dsmcFreeStreamInflowPatch;\n"
                   f"        dsmcFreeStreamInflowPatchProperties\n        {{\n"
                   f"            typeIds ({' '.join(inflow['n'])});\n            numberDensities\n            {{\n{nd}"
                   f"            }}\n            translationalTemperature {inflow['T']!r};\n"
                   f"            rotationalTemperature {inflow.get('Trot', inflow['T'])!r};\n"
                   f"            vibrationalTemperature {inflow.get('Tvib', inflow['T'])!r};\n"
                   f"            velocity ({' '.join(repr(x) for x in inflow['U'])});\n        }}\n    }}\n")
    _w(os.path.join(root, "system/boundariesDict"), "boundariesDict",
       f"dsmcPatchBoundaries\n(\n{pb});\n\ndsmcCyclicBoundaries\n(\n);\n\ndsmcGeneralBoundaries\n(\n{gb});\n")
    fl = ""
    for fname, ids in (fields or {"mixture": list(species)}).items():
        fl += (f"    field\n    {{\n        fieldModel dsmcVolFields;\n        timeProperties\n        {{\n"
               f"            timeOption write;\n            resetAtOutput on;\n            resetAtOutputUntilTime 0;\n"
               f"        }}\n        dsmcVolFieldsProperties\n        {{\n            fieldName {fname};\n"
               f"            typeIds ({' '.join(ids)});\n            measureMeanFreePath true;\n"
               f"            averagingAcrossManyRuns false;\n        }}\n    }}\n")
    _w(os.path.join(root, "system/fieldPropertiesDict"), "fieldPropertiesDict", f"dsmcFields\n(\n{fl});\n")
    cfgs = ""
    for c in init or []:
        nd = "".join(f"            {k} {x!r};\n" for k, x in c["n"].items())
        cfgs += (f"    configuration\n    {{\n        type dsmcMeshFill;\n        numberDensities\n        {{\n{nd}"
                 f"        }};\n        translationalTemperature {c['T']!r};\n"
                 f"        rotationalTemperature {c.get('Trot', c['T'])!r};\n"
                 f"        vibrationalTemperature {c.get('Tvib', c['T'])!r};\n        electronicTemperature 0;\n"
                 f"        velocity ({' '.join(repr(x) for x in c.get('U', (0, 0, 0)))});\n    }}\n")
    _w(os.path.join(root, "system/dsmcInitialiseDict"), "dsmcInitialiseDict", f"configurations\n(\n{cfgs});\n")
    _w(os.path.join(root, "system/chemReactDict"), "chemReactDict",
       f"reactions\n(\n{_reactions_text(reactions or [])});\n")
    if cylinder:
        props = "".join(f"    {k} {v};\n" for k, v in cylinder.get("props", {}).items())
        _w(os.path.join(root, "system/embeddedCylinderDict"), "embeddedCylinderDict",
           f"centre ({cylinder['centre'][0]!r} {cylinder['centre'][1]!r});\nradius {cylinder['radius']!r};\n"
           f"boundaryModel {cylinder['model']};\n{cylinder['model']}Properties\n{{\n{props}}}\n")
    return root


N2 = dict(mass=46.5e-27, diameter=4.17e-10, omega=0.74, alpha=1.36, rotationalDegreesOfFreedom=2,
          nVibrationalModes=1, characteristicVibrationalTemperature=[3371], dissociationTemperature=113500,
          Zref=[52560], referenceTempForZref=[3371], charge=0)
O2 = dict(mass=53.12e-27, diameter=4.07e-10, omega=0.77, alpha=1.4, rotationalDegreesOfFreedom=2,
          nVibrationalModes=1, characteristicVibrationalTemperature=[2256], dissociationTemperature=59500,
          Zref=[17900], referenceTempForZref=[2256], charge=0)
AR = dict(mass=6.63e-26, diameter=4.17e-10, omega=0.81, alpha=1.0)
# heatBath-5species/constant/dsmcProperties
NO = dict(mass=49.81e-27, diameter=4.2e-10, omega=0.79, alpha=1.0, rotationalDegreesOfFreedom=2,
          nVibrationalModes=1, characteristicVibrationalTemperature=[2719], dissociationTemperature=75500,
          Zref=[1400], referenceTempForZref=[2719], charge=0)
N = dict(mass=23.25e-27, diameter=3.00e-10, omega=0.8, alpha=1.0, charge=0)
O = dict(mass=26.56e-27, diameter=3.00e-10, omega=0.8, alpha=1.0, charge=0)


def heat_bath_case(root, fn=100.0, end_time=1e-5, write_interval=1e-9, seed=31, n_terminal=100):
    """The heatBath-5species tutorial: one (1e-5 m)^3 cell with specular
    walls, LB-VHS with relaxation collision numbers 1, five-species air and
    the twelve QK reactions of its chemReactDict, filled with N2 and O2 at
    30,000 K; fn scales the parcel count (the tutorial's is 100)."""
    from .cases import HEATBATH_REACTIONS

    spec = {"dsmcSpecularWallPatch": None}
    wall = ("fixedWalls", "wall", "dsmcSpecularWallPatch", spec.get("dsmcSpecularWallPatch"))
    L = 1e-5
    return write_case(
        root, species={"N2": N2, "O2": O2, "NO": NO, "N": N, "O": O}, n=(1, 1, 1), lo=(0, 0, 0), hi=(L, L, L),
        patches={s: wall for s in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}, fn=fn, dt=1e-9,
        end_time=end_time, write_interval=write_interval, model="LarsenBorgnakkeVariableHardSphere",
        coeffs={"rotationalRelaxationCollisionNumber": 1.0, "vibrationalRelaxationCollisionNumber": 1.0,
                "electronicRelaxationCollisionNumber": 1.0}, seed=seed,
        fields={"N2": ["N2"], "O2": ["O2"], "NO": ["NO"], "N": ["N"], "O": ["O"],
                "mixture": ["N2", "O2", "NO", "N", "O"]},
        init=[{"n": {"N2": 1.21753030168e22, "O2": 3.23647295384e21, "NO": 0.0, "N": 0.0, "O": 0.0},
               "T": 30000.0}], n_terminal=n_terminal, reactions=HEATBATH_REACTIONS)


def couette_case(root, end_time=2e-4, write_interval=1e-4, seed=11, n_terminal=50):
    """The couette_N2-O2 tutorial set-up (5x100x1 cells, cyclic x, diffuse
    walls at 2000 K and 3000 K moving at 300 m/s, LB-VHS N2/O2)."""
    wall = "wall"
    return write_case(
        root, species={"N2": N2, "O2": O2}, n=(5, 100, 1), lo=(0, 0, 0), hi=(0.05, 1.0, 0.01),
        patches={"xmin": ("periodicX_half0", "cyclic", None, None), "xmax": ("periodicX_half1", "cyclic", None, None),
                 "ymin": ("lowerWall", wall, "dsmcDiffuseWallPatch", {"temperature": 2000.0, "velocity": "(0 0 0)"}),
                 "ymax": ("upperWall", wall, "dsmcDiffuseWallPatch", {"temperature": 3000.0, "velocity": "(300 0 0)"}),
                 "zmin": ("frontAndBack", "empty", None, None), "zmax": ("frontAndBack", "empty", None, None)},
        fn=429576195535.3, dt=1e-5, end_time=end_time, write_interval=write_interval,
        model="LarsenBorgnakkeVariableHardSphere",
        coeffs={"rotationalRelaxationCollisionNumber": 5.0, "inverseZvFormulation": '"pre-2008"'}, seed=seed,
        fields={"N2": ["N2"], "O2": ["O2"], "mixture": ["N2", "O2"]},
        init=[{"n": {"N2": 3.2320494711701479e19, "O2": 8.591523910705456e18}, "T": 2000.0}], n_terminal=n_terminal)


def cylinder_case(root, nx=400, ny=400, end_time=5e-4, write_interval=2.5e-4, seed=7, ppc=25):
    """2D Mach-10 argon over a cylinder of radius 0.1 m (diffuse wall at
    500 K, freestream inflow + deletion on the outer sides)."""
    import math
    kb = 1.38064852e-23
    T, n = 200.0, 4.2e20
    U = 10.0 * math.sqrt(5.0 / 3.0 * kb * T / AR["mass"])
    dx = 1.0 / nx
    fn = n * dx * (1.0 / ny) * dx / ppc
    flow = ("flow", "patch", "dsmcDeletionPatch", {"allSpecies": "yes"})
    return write_case(
        root, species={"Ar": AR}, n=(nx, ny, 1), lo=(-0.35, -0.5, 0.0), hi=(0.65, 0.5, dx),
        patches={"xmin": flow, "xmax": flow, "ymin": flow, "ymax": flow,
                 "zmin": ("frontAndBack", "empty", None, None), "zmax": ("frontAndBack", "empty", None, None)},
        fn=fn, dt=2.5e-7 * 400 / nx, end_time=end_time, write_interval=write_interval, seed=seed,
        fields={"Ar": ["Ar"]}, init=[{"n": {"Ar": n}, "T": T, "U": (U, 0.0, 0.0)}],
        inflow={"patches": ["flow"], "n": {"Ar": n}, "T": T, "U": (U, 0.0, 0.0)},
        cylinder={"centre": (0.0, 0.0), "radius": 0.1, "model": "dsmcDiffuseWallPatch",
                  "props": {"temperature": 500.0, "velocity": "(0 0 0)"}})
