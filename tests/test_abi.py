"""The C ABI: the library exports every entry point include/hystrath_dsmc.h
declares, the ctypes mirror has the C layout, and configuration errors are
reported (as dsmcCloud's FatalError cases are) before any device work."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from hystrath_amd import abi, cases, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hystrath_dsmc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hs_dsmc_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = engine.load_library()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(engine.EXPORTS) == names


def test_struct_layout_matches_c():
    fields = {
        "hs_config": [f for f, _ in abi.hs_config._fields_],
        "hs_species": [f for f, _ in abi.hs_species._fields_],
        "hs_patch": [f for f, _ in abi.hs_patch._fields_],
        "hs_particles": [f for f, _ in abi.hs_particles._fields_],
        "hs_reaction": [f for f, _ in abi.hs_reaction._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} size %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st} {f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            st, f, v = line.split()
            got[(st, f)] = int(v)
    for st, fs in fields.items():
        cls = getattr(abi, st)
        assert got[(st, "size")] == C.sizeof(cls), st
        for f in fs:
            assert got[(st, f)] == getattr(cls, f).offset, (st, f)


def _create(case):
    L = engine.load_library()
    h = C.c_void_p()
    rc = L.hs_dsmc_create(C.byref(case.cfg), C.byref(h))
    return rc, L.hs_dsmc_create_error().decode(), h


@pytest.mark.parametrize("mutate,code,msg", [
    (lambda c: setattr(c.cfg, "collision_model", 9), abi.HS_E_INVALID, "BinaryCollisionModel"),
    (lambda c: setattr(c.cfg, "n_species", 0), abi.HS_E_INVALID, "n_species"),
    (lambda c: setattr(c.cfg.patch[0], "model", abi.HS_PATCH_SPECULAR), abi.HS_E_INVALID, "pairs"),
    (lambda c: setattr(c.cfg, "capacity", 0), abi.HS_E_INVALID, "capacity"),
    (lambda c: setattr(c.cfg, "fn", 0.0), abi.HS_E_INVALID, "nEquivalentParticles"),
    (lambda c: setattr(c.cfg.species[0], "charge", 3), abi.HS_E_INVALID, "Charge"),
])
def test_invalid_configuration_is_rejected(mutate, code, msg):
    case = cases.periodic_box(n_side=2)
    mutate(case)
    rc, err, h = _create(case)
    assert rc == code and msg in err and not h.value


def test_unsupported_option_is_reported():
    case = cases.heat_bath_5species(fn=1e5)
    case.cfg.collision_model = abi.HS_COLL_LB_VSS
    rc, err, _ = _create(case)
    assert rc == abi.HS_E_UNSUPPORTED and "chemical reactions need" in err


def test_zv_formulation_range():
    case = cases.couette_n2o2()
    case.cfg.inv_zv_form = 3
    rc, err, _ = _create(case)
    assert rc == abi.HS_E_INVALID and "inverseZvFormulation" in err


def test_no_cpu_fallback_without_device():
    """With no GPU the engine refuses to start instead of running elsewhere."""
    if engine.device_count() > 0:
        pytest.skip("a GPU is visible")
    rc, err, _ = _create(cases.periodic_box(n_side=2))
    assert rc == abi.HS_E_HIP and "no HIP device" in err
    with pytest.raises(engine.EngineError):
        engine.Engine(cases.periodic_box(n_side=2))
