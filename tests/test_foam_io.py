"""OpenFOAM case I/O: dictionary parsing, polyMesh round trip, boundary
conditions, material and solver controls (CPU), and a case run end to end
(GPU)."""
import os

import numpy as np
import pytest

from solids4foam_amd import foam_io

SOLID_PROPS = """
FoamFile { version 2.0; format ascii; class dictionary; object solidProperties; }
// comment
solidModel     linearGeometryTotalDisplacement;
"linearGeometryTotalDisplacementCoeffs|unsLinearGeometryCoeffs"
{
    nCorrectors     100000;
    solutionTolerance 1e-07;
    /* block comment */
    stabilisation { type RhieChow; scaleFactor 0.2; }
}
"""
FV_SOLUTION = """
FoamFile { version 2.0; format ascii; class dictionary; object fvSolution; }
solvers
{
    "D|DD" { tolerance 1e-8; relTol 0.1; maxIter 50; solver PCG; preconditioner FDIC; }
}
"""
MECH = """
FoamFile { version 2.0; format ascii; class dictionary; object mechanicalProperties; }
planeStress no;
mechanical
(
    steel
    {
        type  linearElastic;
        rho   rho [1 -3 0 0 0 0 0] 7800;
        E     E [1 -1 -2 0 0 0 0] 200e+09;
        nu    nu [0 0 0 0 0 0 0] 0.3;
    }
);
"""


def _write(d, rel, text):
    p = os.path.join(d, rel)
    os.makedirs(os.path.dirname(p), exist_ok=True)
    open(p, "w").write(text)


def _d_file(patches):
    body = []
    for name, kind in patches:
        if kind == "fixed":
            body.append(f"{name} {{ type fixedDisplacement; displacement uniform (0 0 0); value uniform (0 0 0); }}")
        elif kind == "load":
            body.append(f"{name} {{ type solidTraction; traction uniform (0 0 1e7); pressure uniform 0; "
                        f"value uniform (0 0 0); }}")
        else:
            body.append(f"{name} {{ type solidTraction; traction uniform (0 0 0); pressure uniform 0; "
                        f"value uniform (0 0 0); }}")
    return ("FoamFile { version 2.0; format ascii; class volVectorField; object D; }\n"
            "dimensions [0 1 0 0 0 0 0];\ninternalField uniform (0 0 0);\nboundaryField\n{\n"
            + "\n".join(body) + "\n}\n")


PATCHES = [("left", "free"), ("right", "free"), ("down", "free"), ("up", "free"), ("back", "fixed"),
           ("front", "load")]


def _box_case(tmp, n=(5, 4, 3), jitter=0.1):
    from solids4foam_amd.solid import PolyMesh
    pm = PolyMesh.box(*n, jitter=jitter, seed=3)
    a = pm.addressing()
    patches = [dict(name=PATCHES[i][0], type="patch", startFace=int(a["patchStart"][i]), nFaces=int(a["patchSize"][i]))
               for i in range(6)]
    foam_io.write_polymesh(tmp, a["points"], a["fOff"], a["fVerts"], a["owner"], a["neighbour"], patches, pm.nCells)
    _write(tmp, "constant/solidProperties", SOLID_PROPS)
    _write(tmp, "constant/mechanicalProperties", MECH)
    _write(tmp, "system/fvSolution", FV_SOLUTION)
    _write(tmp, "0/D", _d_file(PATCHES))
    return pm, a


def test_relaxation_settings(tmp_path):
    # relaxationMethod (solidProperties) and the D field factor (fvSolution)
    _write(str(tmp_path), "constant/solidProperties", SOLID_PROPS.replace("nCorrectors     100000;",
                                                                           "nCorrectors 100000; relaxationMethod Aitken;"))
    _write(str(tmp_path), "system/fvSolution", FV_SOLUTION + "relaxationFactors { equations { } fields { D 0.7; } }\n")
    _, opts = foam_io.read_solid_properties(str(tmp_path))
    assert opts["relaxationMethod"] == 1
    assert foam_io.read_solver_controls(str(tmp_path))["relaxD"] == 0.7
    _write(str(tmp_path), "system/fvSolution", FV_SOLUTION + "relaxationFactors { D 0.5; }\n")
    assert foam_io.read_solver_controls(str(tmp_path))["relaxD"] == 0.5
    _write(str(tmp_path), "constant/solidProperties", SOLID_PROPS.replace("nCorrectors     100000;",
                                                                           "relaxationMethod QuasiNewton;"))
    with pytest.raises(ValueError, match="relaxationMethod"):
        foam_io.read_solid_properties(str(tmp_path))


def test_parse_dictionaries(tmp_path):
    _write(str(tmp_path), "constant/solidProperties", SOLID_PROPS)
    _write(str(tmp_path), "constant/mechanicalProperties", MECH)
    _write(str(tmp_path), "system/fvSolution", FV_SOLUTION)
    model, opts = foam_io.read_solid_properties(str(tmp_path))
    assert model == "linearGeometryTotalDisplacement"
    assert opts["nCorrectors"] == 100000 and opts["solutionTolerance"] == 1e-7
    assert opts["alternativeTolerance"] == 1e-7 and opts["rhieChowScale"] == 0.2 and opts["relaxationMethod"] == 0
    assert foam_io.read_linear_elastic(str(tmp_path)) == (200e9, 0.3)
    sc = foam_io.read_solver_controls(str(tmp_path))
    assert sc == dict(pcgTolerance=1e-8, pcgRelTol=0.1, pcgMaxIter=50)


def test_include_directives(tmp_path):
    # next to the including file, then in the case directory; IfPresent
    # skips a missing file; other directives are still refused
    _write(str(tmp_path), "system/controlDict", 'application solids4Foam;\n#include "times"\n'
           'functions { #include "./system/functions" }\n#includeIfPresent "missing"\nendTime 3;\n')
    _write(str(tmp_path), "system/times", "deltaT 0.5; // a comment\nwriteInterval 2;\n")
    _write(str(tmp_path), "system/functions", 'probe { type probes; fields (D); }\n')
    d = foam_io.read_dict(os.path.join(str(tmp_path), "system/controlDict"))
    assert d["deltaT"] == 0.5 and d["writeInterval"] == 2 and d["endTime"] == 3
    assert d["functions"]["probe"]["type"] == "probes"
    assert foam_io.read_time_controls(str(tmp_path))["deltaT"] == 0.5
    _write(str(tmp_path), "system/bad", '#include "nowhere"\n')
    with pytest.raises(FileNotFoundError):
        foam_io.read_dict(os.path.join(str(tmp_path), "system/bad"))
    _write(str(tmp_path), "system/etc", '#includeEtc "caseDicts/setConstraintTypes"\n')
    with pytest.raises(ValueError, match="not supported"):
        foam_io.read_dict(os.path.join(str(tmp_path), "system/etc"))


def test_polymesh_and_bc_roundtrip(tmp_path):
    pm, a = _box_case(str(tmp_path))
    m = foam_io.read_polymesh(str(tmp_path))
    assert m["nCells"] == pm.nCells
    assert np.array_equal(m["points"], a["points"])  # %.17g round trips doubles
    assert np.array_equal(m["fOff"], a["fOff"]) and np.array_equal(m["fVerts"], a["fVerts"])
    assert np.array_equal(m["owner"], a["owner"]) and np.array_equal(m["neighbour"], a["neighbour"])
    assert [b["startFace"] for b in m["boundary"]] == list(a["patchStart"])
    bcs = foam_io.read_D_boundary(str(tmp_path), m["boundary"])
    assert [b["type"] for b in bcs] == ["solidTraction"] * 4 + ["fixedDisplacement", "solidTraction"]
    assert np.all(bcs[5]["traction"] == [0, 0, 1e7]) and bcs[5]["traction"].shape == (m["boundary"][5]["nFaces"], 3)


def test_unsupported_bc_is_rejected(tmp_path):
    pm, a = _box_case(str(tmp_path))
    m = foam_io.read_polymesh(str(tmp_path))
    _write(str(tmp_path), "0/D", _d_file(PATCHES).replace("type fixedDisplacement",
                                                           "type solidContact"))
    with pytest.raises(ValueError, match="not supported"):
        foam_io.read_D_boundary(str(tmp_path), m["boundary"])
    # zero shear is read (value = the prescribed displacement); its options
    # outside the device path's first-order form are refused
    _write(str(tmp_path), "0/D", _d_file(PATCHES).replace("type fixedDisplacement;",
                                                           "type fixedDisplacementZeroShear;"))
    bcs = foam_io.read_D_boundary(str(tmp_path), m["boundary"])
    assert bcs[4]["type"] == "fixedDisplacementZeroShear" and bcs[4]["value"].shape == (m["boundary"][4]["nFaces"], 3)
    _write(str(tmp_path), "0/D", _d_file(PATCHES).replace("type fixedDisplacement;",
                                                           "type fixedDisplacementZeroShear; secondOrder yes;"))
    with pytest.raises(ValueError, match="not supported"):
        foam_io.read_D_boundary(str(tmp_path), m["boundary"])


@pytest.mark.skipif(not os.path.isdir("/root/reference/tutorials"), reason="reference tutorials not present")
def test_parse_reference_tutorial_dictionaries():
    case = "/root/reference/tutorials/solids/linearElasticity/cantilever2d/segregatedCantilever2d"
    model, opts = foam_io.read_solid_properties(case)
    assert model == "linearGeometryTotalDisplacement" and opts["nCorrectors"] == 100000
    assert foam_io.read_solver_controls(case) == dict(pcgTolerance=1e-8, pcgRelTol=0.1, pcgMaxIter=50)
    assert foam_io.read_linear_elastic(case) == (200e9, 0.3)
    # a controlDict that includes its function objects from system/functions
    tc = foam_io.read_time_controls("/root/reference/tutorials/solids/linearElasticity/contactPatchTest")
    assert tc["endTime"] == 1.0 and tc["deltaT"] == 1.0


@pytest.mark.gpu
def test_run_case_end_to_end(tmp_path):
    from solids4foam_amd import cases, Context
    pm, a = _box_case(str(tmp_path), n=(6, 6, 6), jitter=0.1)
    ctx = Context(0)
    m = foam_io.run_case(str(tmp_path), ctx=ctx, solver_overrides=dict(nCorrectors=20000, alternativeTolerance=1e-10,
                                                                         solutionTolerance=1e-10))
    assert m.report["converged"]
    D = foam_io._field_value(foam_io.read_dict(os.path.join(str(tmp_path), "1", "D"))["internalField"], pm.nCells, 3)
    assert np.allclose(D, m.D, rtol=1e-9, atol=1e-20)
    for name, k in (("pointD", 3), ("epsilonEq", 1), ("epsilon", 6), ("sigmaEq", 1)):
        n = pm.nPoints if name == "pointD" else pm.nCells
        v = foam_io._field_value(foam_io.read_dict(os.path.join(str(tmp_path), "1", name))["internalField"], n, k)
        assert np.all(np.isfinite(v)) and np.abs(v).max() > 0
    # same as the box generator's solve with the same settings
    _, mesh = cases.cube_tension(ctx, n=6, jitter=0.1, seed=3)
    ref = mesh.solve(nCorrectors=20000, alternativeTolerance=1e-10, solutionTolerance=1e-10, rhieChowScale=0.2,
                     pcgTolerance=1e-8, pcgRelTol=0.1, pcgMaxIter=50)
    assert np.linalg.norm(ref["D"] - m.D) / np.linalg.norm(ref["D"]) < 1e-9


G_FILE = """
FoamFile { version 2.0; format ascii; class uniformDimensionedVectorField; object g; }
dimensions [0 1 -2 0 0 0 0];
value (0 -9.81 2);
"""
FV_SCHEMES_EULER = """
FoamFile { version 2.0; format ascii; class dictionary; object fvSchemes; }
d2dt2Schemes { default Euler; }
ddtSchemes { default Euler; }
gradSchemes { default leastSquares; }
"""
CONTROL = """
FoamFile { version 2.0; format ascii; class dictionary; object controlDict; }
application solids4Foam; startFrom startTime; startTime 0; stopAt endTime; endTime 0.003;
deltaT 0.001; writeControl timeStep; writeInterval 1;
"""


def test_dynamic_dictionaries(tmp_path):
    d = str(tmp_path)
    _write(d, "constant/mechanicalProperties", MECH)
    assert foam_io.read_density(d) == 7800.0
    assert foam_io.read_gravity(d) is None and foam_io.read_d2dt2_scheme(d) == "steadyState"
    _write(d, "constant/g", G_FILE)
    _write(d, "system/fvSchemes", FV_SCHEMES_EULER)
    _write(d, "system/controlDict", CONTROL)
    assert np.array_equal(foam_io.read_gravity(d), [0.0, -9.81, 2.0])
    assert foam_io.read_d2dt2_scheme(d) == "Euler"
    tc = foam_io.read_time_controls(d)
    assert tc["deltaT"] == 0.001 and tc["endTime"] == 0.003 and tc["writeControl"] == "timeStep"
    assert foam_io._time_name(0.003) == "0.003" and foam_io._time_name(1.0) == "1"


@pytest.mark.gpu
def test_run_case_transient_free_fall(tmp_path):
    # all faces traction free, gravity, Euler d2dt2: the controlDict loop
    # writes every step and the body falls as D^n = n(n+1)/2 g dt^2
    from solids4foam_amd import Context
    d = str(tmp_path)
    pm, a = _box_case(d, n=(4, 3, 3), jitter=0.1)
    _write(d, "0/D", _d_file([(p, "free") for p, _ in PATCHES]))
    _write(d, "constant/g", G_FILE)
    _write(d, "system/fvSchemes", FV_SCHEMES_EULER)
    _write(d, "system/controlDict", CONTROL)
    ctx = Context(0)
    m = foam_io.run_case(d, ctx=ctx, solver_overrides=dict(nCorrectors=200, solutionTolerance=1e-12,
                                                           alternativeTolerance=1e-12, pcgTolerance=1e-14,
                                                           pcgRelTol=0.0))
    assert m.report["converged"] and abs(m.time - 0.003) < 1e-12
    g = np.array([0.0, -9.81, 2.0])
    for n, name in ((1, "0.001"), (2, "0.002"), (3, "0.003")):
        D = foam_io._field_value(foam_io.read_dict(os.path.join(d, name, "D"))["internalField"], pm.nCells, 3)
        exact = n * (n + 1) / 2 * g * 1e-6
        assert np.abs(D - exact).max() < 1e-9 * np.abs(exact).max()


MECH_TWO = """
FoamFile { version 2.0; format ascii; class dictionary; object mechanicalProperties; }
planeStress no;
mechanical
(
    steel { type linearElastic; rho rho [1 -3 0 0 0 0 0] 7800; E E [1 -1 -2 0 0 0 0] 200e+09; nu nu [0 0 0 0 0 0 0] 0.3; }
    rubber { type linearElastic; rho rho [1 -3 0 0 0 0 0] 1200; E E [1 -1 -2 0 0 0 0] 1e+09; nu nu [0 0 0 0 0 0 0] 0.45; }
);
"""


def _zones_file(zones):
    body = "\n".join(f"{name}\n{{\n    type cellZone;\n    cellLabels List<label> {len(c)}({' '.join(map(str, c))});\n}}"
                     for name, c in zones)
    return ("FoamFile { version 2.0; format ascii; class regIOobject; object cellZones; }\n"
            f"{len(zones)}\n(\n{body}\n)\n")


def test_multi_material_cell_zones(tmp_path):
    # several laws: each acts on the cellZone of its name, per cell
    d = str(tmp_path)
    pm, a = _box_case(d)
    C = pm.nCells
    soft = np.arange(C)[np.arange(C) % 3 == 0]
    hard = np.setdiff1d(np.arange(C), soft)
    _write(d, "constant/mechanicalProperties", MECH_TWO)
    _write(d, "constant/polyMesh/cellZones", _zones_file([("steel", hard), ("rubber", soft)]))
    assert set(foam_io.read_cell_zones(d)) == {"steel", "rubber"}
    E, nu, rho = foam_io.read_materials(d, C)
    assert np.all(E[hard] == 200e9) and np.all(E[soft] == 1e9)
    assert np.all(nu[soft] == 0.45) and np.all(rho[hard] == 7800.0)
    # the law of every cell, for the sub-mesh interfaces
    z = foam_io.read_material_zones(d, C)
    assert z.dtype == np.int32 and np.all(z[hard] == 0) and np.all(z[soft] == 1)
    _write(d, "constant/polyMesh/cellZones", _zones_file([("steel", hard[1:]), ("rubber", soft)]))
    with pytest.raises(ValueError, match="no material"):
        foam_io.read_materials(d, C)
    with pytest.raises(ValueError, match="no material"):
        foam_io.read_material_zones(d, C)
    _write(d, "constant/mechanicalProperties", MECH)
    assert foam_io.read_material_zones(d, C) is None


@pytest.mark.gpu
def test_run_case_multi_material(tmp_path):
    # two laws on two cellZones: the case runs as a per-cell material with
    # one sub-mesh per law and matches the oracle with the same per-cell E
    # and nu and the same zones
    import oracle
    from solids4foam_amd import Context
    d = str(tmp_path)
    n = (6, 5, 4)
    pm, a = _box_case(d, n=n, jitter=0.1)
    C = pm.nCells
    soft = np.arange(C)[np.arange(C) % 3 == 0]
    hard = np.setdiff1d(np.arange(C), soft)
    _write(d, "constant/mechanicalProperties", MECH_TWO)
    _write(d, "constant/polyMesh/cellZones", _zones_file([("steel", hard), ("rubber", soft)]))
    tight = dict(nCorrectors=50000, solutionTolerance=1e-12, alternativeTolerance=1e-12, pcgTolerance=1e-14)
    m = foam_io.run_case(d, ctx=Context(0), solver_overrides=tight)
    assert m.report["converged"]
    E = np.where(np.isin(np.arange(C), soft), 1e9, 200e9)
    nu = np.where(np.isin(np.arange(C), soft), 0.45, 0.3)
    om = oracle.Mesh.box(*n, (1.0, 1.0, 1.0), 0.1, 3)
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(5)] = (0.0, 0.0, 1e7)
    oc = oracle.Case(om, E, nu, [1, 1, 1, 1, 0, 1], bv, rc_scale=0.2)
    zones = foam_io.read_material_zones(d, C)
    assert np.array_equal(zones, np.isin(np.arange(C), soft).astype(np.int32))
    assert oc.set_zones(zones) == 60
    oc.init()
    oc.solve(nCorr=50000, solTol=1e-12, altTol=1e-12, pcgTol=1e-14, rcScale=0.2, precond=oracle.PRECOND_DIC)
    Do = oc.get("D").reshape(-1, 3)
    assert np.linalg.norm(m.D - Do) / np.linalg.norm(Do) < 1e-8


def test_unsupported_law_options_are_refused(tmp_path):
    # solvePressureEqn (mechanicalLaw.C:1374) and sigma0 (linearElastic.C:158-166)
    # change the reference's stress: the case reader refuses them
    d = str(tmp_path)
    _box_case(d)
    assert foam_io.read_materials(d, 60)[0] == 200e9
    _write(d, "constant/mechanicalProperties", MECH.replace("nu    nu", "solvePressureEqn yes;\n        nu    nu"))
    with pytest.raises(ValueError, match="solvePressureEqn"):
        foam_io.read_materials(d, 60)
    with pytest.raises(ValueError, match="solvePressureEqn"):
        foam_io.read_linear_elastic(d)
    _write(d, "constant/mechanicalProperties", MECH.replace("nu    nu", "solvePressureEqn no;\n        nu    nu"))
    assert foam_io.read_materials(d, 60)[1] == 0.3
    _write(d, "constant/mechanicalProperties",
           MECH.replace("nu    nu", "sigma0 sigma0 [1 -1 -2 0 0 0 0] (1e6 0 0 0 0 0);\n        nu    nu"))
    with pytest.raises(ValueError, match="sigma0"):
        foam_io.read_materials(d, 60)
    _write(d, "constant/mechanicalProperties", MECH)
    _write(d, "0/sigma0", "FoamFile { version 2.0; format ascii; class volSymmTensorField; object sigma0; }\n")
    with pytest.raises(ValueError, match="sigma0"):
        foam_io.run_case(d)
