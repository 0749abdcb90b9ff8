"""CPU checks of libs4f: it loads, exports every symbol the header declares,
its host-side polyMesh generator and geometry agree bit-for-bit with the
oracle's, and host-side validation rejects bad input without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from solids4foam_amd import _abi
from solids4foam_amd.solid import PolyMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "s4f_abi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(s4f_[a-z0-9_]+)\s*\(", txt)))


def test_library_loads_and_exports_header_symbols():
    L = _abi.lib()
    syms = _header_symbols()
    assert len(syms) >= 25
    raw = ctypes.CDLL(_abi.LIB_PATH)
    missing = [s for s in syms if not hasattr(raw, s)]
    assert not missing, missing
    assert set(syms) == set(_abi.EXPORTED)
    assert b"gfx950" in L.s4f_version()


def test_library_is_built_for_gfx950():
    data = open(_abi.LIB_PATH, "rb").read()
    assert b"gfx950" in data


@pytest.mark.parametrize("args", [
    dict(n=(50, 10, 10), L=(1.0, 0.2, 0.2), jitter=0.0, seed=0, tets=False),
    dict(n=(7, 5, 4), L=(1.0, 1.0, 1.0), jitter=0.15, seed=2, tets=False),
    dict(n=(4, 3, 3), L=(1.0, 0.5, 0.7), jitter=0.15, seed=5, tets=True),
    dict(n=(1, 1, 1), L=(1.0, 1.0, 1.0), jitter=0.0, seed=0, tets=True),
])
def test_polymesh_matches_oracle_bitwise(oracle_mod, args):
    om = oracle_mod.Mesh.box(*args["n"], L=args["L"], jitter=args["jitter"], seed=args["seed"], tets=args["tets"])
    pm = PolyMesh.box(*args["n"], L=args["L"], jitter=args["jitter"], seed=args["seed"], tets=args["tets"])
    a = pm.addressing()
    assert (pm.nCells, pm.nFaces, pm.nInternalFaces, pm.nPoints) == (om.nCells, om.nFaces, om.nInternalFaces,
                                                                      om.nPoints)
    assert np.array_equal(a["points"], om.points)
    assert np.array_equal(a["fOff"], om.fOff)
    assert np.array_equal(a["fVerts"], om.fVerts)
    assert np.array_equal(a["owner"], om.owner)
    assert np.array_equal(a["neighbour"], om.neighbour)
    assert np.array_equal(a["patchStart"], om.patchStart)
    assert np.array_equal(a["patchSize"], om.patchSize)
    g = pm.geometry()
    # same OpenFOAM formulas in the same operation order: identical bits
    assert np.array_equal(g["Cf"], om.Cf)
    assert np.array_equal(g["Sf"], om.Sf)
    assert np.array_equal(g["C"], om.C)
    assert np.array_equal(g["V"], om.V)


def test_polymesh_from_arrays_roundtrip(oracle_mod):
    om = oracle_mod.Mesh.box(3, 4, 2, jitter=0.1, seed=9)
    pm = PolyMesh.from_arrays(om.points, om.fOff, om.fVerts, om.owner, om.neighbour, om.nCells, om.patchStart,
                              om.patchSize)
    g = pm.geometry()
    assert np.array_equal(g["C"], om.C) and np.array_equal(g["Sf"], om.Sf)


def test_polymesh_rejects_bad_input(oracle_mod):
    om = oracle_mod.Mesh.box(2, 2, 2)
    bad_owner = om.owner.copy()
    bad_owner[3] = 99
    with pytest.raises(_abi.S4FError) as e:
        PolyMesh.from_arrays(om.points, om.fOff, om.fVerts, bad_owner, om.neighbour, om.nCells, om.patchStart,
                             om.patchSize)
    assert e.value.status == _abi.S4F_ERR_INVALID
    assert "owner" in str(e.value)
    with pytest.raises(_abi.S4FError) as e:
        PolyMesh.box(0, 2, 2)
    assert e.value.status == _abi.S4F_ERR_INVALID


def test_default_opts_match_reference_defaults():
    o = _abi.default_opts()
    # solidModel.C:1147-1177 and the RhieChow default (solidModel.C:1307-1315)
    assert o.nCorrectors == 10000 and o.solutionTolerance == 1e-6 and o.alternativeTolerance == 1e-7
    assert o.rhieChowScale == 0.1 and o.interfaceScale == 0.01
    assert o.pcgRelTol == 0.1 and o.pcgMaxIter == 1000


def test_unknown_solid_model_type_raises():
    from solids4foam_amd import solid_model_new
    with pytest.raises(ValueError, match="linearGeometryTotalDisplacement"):
        solid_model_new("unsLinearGeometry")
