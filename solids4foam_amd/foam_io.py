"""OpenFOAM case I/O for the linearGeometryTotalDisplacement path.

Reads what a solids4foam case gives the reference and writes what it writes:
  * dictionaries (FoamFile subset: keywords, sub-dictionaries, lists,
    ``uniform``/``nonuniform List<...>`` fields, dimensioned scalars, quoted
    regex keys, // and /* */ comments, ``#include`` / ``#includeIfPresent``
    of files; other ``#`` directives are refused);
  * ASCII polyMesh (points, faces, owner, neighbour, boundary), optionally
    gzip-compressed (``*.gz``);
  * the ``0/D`` boundary conditions the device path supports
    (fixedDisplacement, linearSpatialDisplacement, fixedDisplacementZeroShear,
    solidTraction, solidSymmetry, empty), the material of
    ``constant/mechanicalProperties`` (one linearElastic law, its density),
    ``constant/g``, the d2dt2 scheme of ``system/fvSchemes`` and the time
    controls of ``system/controlDict``;
  * ASCII volVectorField / volSymmTensorField / volScalarField output.

Reference: solidModel field I/O (src/solids4FoamModels/solidModels/solidModel/solidModel.C:1017-1124,
1936-1996), solidTractionFvPatchVectorField.C:60-120 and
fixedDisplacementFvPatchVectorField.C:150-190 (dictionary keys), linearElastic.C:62-104.
"""
import gzip
import os
import re

import numpy as np

# ---------------------------------------------------------------------------
# tokenizer / dictionary parser
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r'"(?:[^"\\]|\\.)*"|[{}();\[\]]|[^\s{}();\[\]"]+')


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _open(path):
    if os.path.exists(path):
        return open(path, "rt")
    if os.path.exists(path + ".gz"):
        return gzip.open(path + ".gz", "rt")
    raise FileNotFoundError(path)


def read_text(path):
    with _open(path) as f:
        return f.read()


def _num(tok):
    try:
        if re.fullmatch(r"[-+]?\d+", tok):
            return int(tok)
        return float(tok)
    except ValueError:
        return None


class _Parser:
    def __init__(self, text):
        self.toks = _TOKEN.findall(_strip_comments(text))
        self.i = 0

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else None

    def next(self):
        t = self.toks[self.i]
        self.i += 1
        return t

    def expect(self, t):
        got = self.next()
        if got != t:
            raise ValueError(f"expected {t!r}, got {got!r} (token {self.i})")

    def parse_dict_body(self, end="}"):
        d = {}
        while self.peek() is not None and self.peek() != end:
            key = self.next()
            if key.startswith("#"):
                raise ValueError(f"{key} directives are not supported")
            if key.startswith('"'):
                key = key[1:-1]
            if self.peek() == "{":
                self.next()
                d[key] = self.parse_dict_body("}")
                self.expect("}")
            else:
                vals = []
                while self.peek() != ";":
                    vals.append(self.parse_value())
                self.expect(";")
                d[key] = vals[0] if len(vals) == 1 else vals
        return d

    def parse_value(self):
        t = self.peek()
        if t == "(":
            return self.parse_list()
        if t == "[":  # dimensions
            self.next()
            dims = []
            while self.peek() != "]":
                dims.append(self.next())
            self.next()
            return ("dimensions", dims)
        if t == "{":
            self.next()
            d = self.parse_dict_body("}")
            self.expect("}")
            return d
        self.next()
        n = _num(t)
        if n is not None and self.peek() == "(":  # counted list N ( ... )
            return self.parse_list()
        if t.startswith("List<") and self.peek() is not None:
            return ("List", t[5:-1])
        return n if n is not None else t

    def parse_list(self):
        self.expect("(")
        out = []
        while self.peek() != ")":
            if self.peek() == "(":
                out.append(self.parse_list())
            else:
                t = self.next()
                n = _num(t)
                if n is not None and self.peek() == "(":
                    out.append(self.parse_list())
                elif self.peek() == "{" and n is None:
                    self.next()
                    d = self.parse_dict_body("}")
                    self.expect("}")
                    out.append((t, d))
                else:
                    out.append(n if n is not None else t)
        self.expect(")")
        return out


def parse_dict(text):
    """OpenFOAM dictionary text -> nested dict (FoamFile header kept)."""
    return _Parser(text).parse_dict_body(end=None)


_INCLUDE = re.compile(r'#(includeIfPresent|sinclude|include)\s+"([^"]+)"')


def _expand_includes(text, path, depth=0):
    """Replaces ``#include "file"`` (and ``#includeIfPresent`` / ``#sinclude``,
    which skip a missing file) by the file's text.  A relative name is looked
    up next to the including file, then in the case directory (the parent of
    system/, constant/ or the time directory); $FOAM_CASE and <case> name the
    case directory."""
    if depth > 8:
        raise ValueError(f"{path}: #include nested too deeply")
    here = os.path.dirname(os.path.abspath(path))
    case = os.path.dirname(here)

    def repl(m):
        kind, name = m.group(1), m.group(2)
        name = name.replace("$FOAM_CASE", case).replace("<case>", case)
        for c in ([name] if os.path.isabs(name) else [os.path.join(here, name), os.path.join(case, name)]):
            if os.path.exists(c) or os.path.exists(c + ".gz"):
                return _expand_includes(_strip_comments(read_text(c)), c, depth + 1)
        if kind == "include":
            raise FileNotFoundError(f"{path}: #include {name!r} not found")
        return ""

    return _INCLUDE.sub(repl, text)


def read_dict(path):
    return parse_dict(_expand_includes(_strip_comments(read_text(path)), path))


def lookup(d, key, default=None):
    """dictionary lookup that honours quoted regex keys ("D|DD")"""
    if key in d:
        return d[key]
    for k, v in d.items():
        try:
            if re.fullmatch(k, key):
                return v
        except re.error:
            continue
    return default


def dimensioned(v):
    """value of `E [dims] 200e9`, `E E [dims] 200e9` or a plain number"""
    if isinstance(v, list):
        return float(v[-1])
    return float(v)


# ---------------------------------------------------------------------------
# polyMesh
# ---------------------------------------------------------------------------
def _header_and_body(text):
    text = _strip_comments(text)
    m = re.search(r"FoamFile\s*\{.*?\}", text, flags=re.S)
    return text[m.end():] if m else text


def _read_list_block(text):
    """body of an OpenFOAM list file: N ( entries )"""
    body = _header_and_body(text)
    m = re.search(r"(\d+)\s*\(", body)
    if not m:
        raise ValueError("no list found")
    n = int(m.group(1))
    start = m.end()
    end = body.rfind(")")
    return n, body[start:end]


def read_points(path):
    n, inner = _read_list_block(read_text(path))
    vals = np.array(inner.replace("(", " ").replace(")", " ").split(), dtype=np.float64)
    return vals.reshape(n, 3)


def read_faces(path):
    n, inner = _read_list_block(read_text(path))
    toks = inner.replace("(", " ( ").replace(")", " ) ").split()
    off = [0]
    verts = []
    i = 0
    while i < len(toks):
        k = int(toks[i])
        assert toks[i + 1] == "("
        verts.extend(int(x) for x in toks[i + 2:i + 2 + k])
        i += k + 3
        off.append(len(verts))
    if len(off) - 1 != n:
        raise ValueError(f"faces: expected {n}, read {len(off) - 1}")
    return np.array(off, np.int32), np.array(verts, np.int32)


def read_labels(path):
    text = read_text(path)
    n, inner = _read_list_block(text)
    vals = np.array(inner.split(), dtype=np.int64).astype(np.int32)
    if len(vals) != n:
        raise ValueError(f"{path}: expected {n} labels, read {len(vals)}")
    note = re.search(r'note\s+"([^"]*)"', text)
    return vals, (note.group(1) if note else "")


def read_boundary(path):
    n, inner = _read_list_block(read_text(path))
    entries = _Parser("(" + inner + ")").parse_list()
    out = []
    for name, d in entries:
        out.append(dict(name=name, type=d.get("type"), startFace=int(d["startFace"]), nFaces=int(d["nFaces"])))
    return out


def read_polymesh(case_dir):
    pm = os.path.join(case_dir, "constant", "polyMesh")
    points = read_points(os.path.join(pm, "points"))
    fOff, fVerts = read_faces(os.path.join(pm, "faces"))
    owner, note = read_labels(os.path.join(pm, "owner"))
    neighbour, _ = read_labels(os.path.join(pm, "neighbour"))
    boundary = read_boundary(os.path.join(pm, "boundary"))
    nCells = int(owner.max()) + 1
    m = re.search(r"nCells:\s*(\d+)", note)
    if m:
        nCells = int(m.group(1))
    return dict(points=points, fOff=fOff, fVerts=fVerts, owner=owner, neighbour=neighbour, boundary=boundary,
                nCells=nCells)


def _foam_header(cls, obj, location=None):
    loc = f'    location    "{location}";\n' if location else ""
    return ("FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
            f"    class       {cls};\n{loc}    object      {obj};\n}}\n\n")


def write_polymesh(case_dir, points, fOff, fVerts, owner, neighbour, patches, nCells):
    """patches: list of dict(name, type, startFace, nFaces)"""
    pm = os.path.join(case_dir, "constant", "polyMesh")
    os.makedirs(pm, exist_ok=True)
    with open(os.path.join(pm, "points"), "w") as f:
        f.write(_foam_header("vectorField", "points", "constant/polyMesh"))
        f.write(f"{len(points)}\n(\n")
        f.writelines(f"({p[0]:.17g} {p[1]:.17g} {p[2]:.17g})\n" for p in points)
        f.write(")\n")
    with open(os.path.join(pm, "faces"), "w") as f:
        f.write(_foam_header("faceList", "faces", "constant/polyMesh"))
        f.write(f"{len(fOff) - 1}\n(\n")
        for i in range(len(fOff) - 1):
            v = fVerts[fOff[i]:fOff[i + 1]]
            f.write(f"{len(v)}(" + " ".join(map(str, v)) + ")\n")
        f.write(")\n")
    note = f"nPoints:{len(points)}  nCells:{nCells}  nFaces:{len(owner)}  nInternalFaces:{len(neighbour)}"
    for name, arr in (("owner", owner), ("neighbour", neighbour)):
        with open(os.path.join(pm, name), "w") as f:
            f.write(_foam_header("labelList", name, "constant/polyMesh").replace(
                "    object", f'    note        "{note}";\n    object'))
            f.write(f"{len(arr)}\n(\n")
            f.write("\n".join(map(str, arr)))
            f.write("\n)\n")
    with open(os.path.join(pm, "boundary"), "w") as f:
        f.write(_foam_header("polyBoundaryMesh", "boundary", "constant/polyMesh"))
        f.write(f"{len(patches)}\n(\n")
        for p in patches:
            f.write(f"    {p['name']}\n    {{\n        type            {p.get('type', 'patch')};\n"
                    f"        nFaces          {p['nFaces']};\n        startFace       {p['startFace']};\n    }}\n")
        f.write(")\n")


# ---------------------------------------------------------------------------
# fields and boundary conditions
# ---------------------------------------------------------------------------
def _field_value(v, n, width):
    """uniform x | uniform (x y z) | nonuniform List<T> N ( ... ) -> (n, width) array"""
    if isinstance(v, list) and len(v) >= 2 and v[0] == "uniform":
        val = v[1]
        arr = np.array(val if isinstance(val, list) else [val], dtype=np.float64).reshape(1, -1)
        return np.broadcast_to(arr, (n, width)).copy()
    if isinstance(v, list) and len(v) >= 3 and v[0] == "nonuniform":
        data = v[-1]
        arr = np.array(data, dtype=np.float64).reshape(-1, width)
        if len(arr) != n:
            raise ValueError(f"nonuniform field has {len(arr)} values, patch has {n}")
        return arr
    if isinstance(v, (int, float)):
        return np.full((n, width), float(v))
    raise ValueError(f"unsupported field value {v!r}")


SUPPORTED_D_BCS = ("fixedDisplacement", "fixedDisplacementZeroShear", "linearSpatialDisplacement", "solidTraction",
                   "solidSymmetry", "empty")


def read_D_boundary(case_dir, boundary, time="0", face_centres=None):
    """per patch: dict(type, value (n,3) | traction (n,3), pressure (n,)).
    linearSpatialDisplacement (linearSpatialDisplacementFvPatchVectorField.C:96-106)
    is a fixedDisplacement with value a + (b & Cf); it needs the boundary face
    centres (face_centres: (nBoundaryFaces, 3) in patch order)."""
    d = read_dict(os.path.join(case_dir, time, "D"))
    bf = d["boundaryField"]
    out = []
    for p in boundary:
        e = lookup(bf, p["name"])
        if e is None:
            raise ValueError(f"0/D has no entry for patch {p['name']}")
        t = e.get("type")
        n = p["nFaces"]
        if t == "fixedDisplacement":
            disp = e.get("displacement", e.get("value", ["uniform", [0, 0, 0]]))
            out.append(dict(type=t, value=_field_value(disp, n, 3)))
        elif t == "linearSpatialDisplacement":
            if face_centres is None:
                raise ValueError(f"patch {p['name']}: linearSpatialDisplacement needs the face centres")
            a = np.asarray(e["a"], np.float64).reshape(3)
            b = np.asarray(e["b"], np.float64).reshape(3, 3)
            s0 = p["startFace"] - boundary[0]["startFace"]
            Cf = np.asarray(face_centres)[s0:s0 + n]
            out.append(dict(type="fixedDisplacement", value=a + Cf @ b.T))
        elif t == "fixedDisplacementZeroShear":
            # fixedDisplacementZeroShearFvPatchVectorField.C:81-120 (totalDisp from
            # "value") over solidDirectionMixedFvPatchVectorField.C:79-111 (defaults)
            on = ("yes", "on", "true")
            if "displacementSeries" in e:
                raise ValueError(f"patch {p['name']}: displacementSeries is not supported on the device path")
            if str(e.get("forceZeroShearGrad", "no")) in on or str(e.get("secondOrder", "no")) in on:
                raise ValueError(f"patch {p['name']}: forceZeroShearGrad/secondOrder are not supported on the "
                                 "device path")
            if str(e.get("nonOrthogonalCorrections", "yes")) not in on or float(e.get("limitCoeff", 1.0)) != 1.0:
                raise ValueError(f"patch {p['name']}: only nonOrthogonalCorrections yes and limitCoeff 1 are "
                                 "supported on the device path")
            out.append(dict(type=t, value=_field_value(e.get("value", ["uniform", [0, 0, 0]]), n, 3)))
        elif t == "empty":
            out.append(dict(type="empty"))
        elif t == "solidSymmetry":
            if str(e.get("secondOrder", "no")) in ("yes", "on", "true"):
                raise ValueError(f"patch {p['name']}: solidSymmetry secondOrder is not supported on the device path")
            out.append(dict(type="solidSymmetry"))
        elif t == "solidTraction":
            tr = e.get("traction", ["uniform", [0, 0, 0]])
            pr = e.get("pressure", ["uniform", 0])
            out.append(dict(type=t, traction=_field_value(tr, n, 3), pressure=_field_value(pr, n, 1)[:, 0]))
        else:
            raise ValueError(f"patch {p['name']}: boundary condition {t!r} is not supported on the device path "
                             f"(supported: {', '.join(SUPPORTED_D_BCS)})")
    return out


def read_linear_elastic(case_dir):
    """(E, nu) of the single linearElastic law (linearElastic.C:62-104)"""
    d = read_dict(os.path.join(case_dir, "constant", "mechanicalProperties"))
    laws = d.get("mechanical")
    if isinstance(laws, list):
        if len(laws) != 1:
            raise ValueError("the device path takes one mechanical law")
        _, law = laws[0]
    else:
        law = laws
    E, nu, _ = _law_constants(law)
    return E, nu


def read_plane_stress(case_dir):
    """mechanicalProperties planeStress switch (mechanicalLaw.C:485-497)"""
    d = read_dict(os.path.join(case_dir, "constant", "mechanicalProperties"))
    return str(d.get("planeStress", "no")).lower() in ("yes", "on", "true", "1")


def read_cell_zones(case_dir):
    """constant/polyMesh/cellZones -> {zone name: cell labels} ({} when absent)"""
    path = os.path.join(case_dir, "constant", "polyMesh", "cellZones")
    if not (os.path.exists(path) or os.path.exists(path + ".gz")):
        return {}
    n, inner = _read_list_block(read_text(path))
    out = {}
    for name, d in _Parser("(" + inner + ")").parse_list():
        labels = d.get("cellLabels", [])
        if labels and isinstance(labels[0], tuple):  # List<label> N ( ... )
            labels = labels[1]
        out[name] = np.asarray(labels, np.int64)
    if len(out) != n:
        raise ValueError(f"cellZones: expected {n} zones, read {len(out)}")
    return out


def _check_law_options(law):
    """linearElastic options the device path does not build raise: the
    pressure-equation formulation (mechanicalLaw.C:1374, 1525-1527) and an
    initial stress sigma0 (linearElastic.C:158-166) change the stress"""
    if str(law.get("solvePressureEqn", "no")).lower() in ("yes", "on", "true", "1"):
        raise ValueError("linearElastic solvePressureEqn is not supported on the device path")
    if "sigma0" in law:
        raise ValueError("linearElastic sigma0 (initial stress) is not supported on the device path")


def _law_constants(law):
    if law.get("type") != "linearElastic":
        raise ValueError(f"mechanical law {law.get('type')!r} is not supported (linearElastic)")
    _check_law_options(law)
    if "E" in law and "nu" in law:
        E, nu = dimensioned(law["E"]), dimensioned(law["nu"])
    else:
        mu, K = dimensioned(law["mu"]), dimensioned(law["K"])
        E, nu = 9 * K * mu / (3 * K + mu), (3 * K - 2 * mu) / (2 * (3 * K + mu))
    return E, nu, (dimensioned(law["rho"]) if "rho" in law else None)


def read_materials(case_dir, nCells):
    """(E, nu, rho) of mechanicalProperties: scalars for one law, per-cell
    arrays when there are several laws, each acting on the cellZone of its
    name (the reference's solidSubMeshes assignment, mechanicalModel.C:195-245;
    the device path takes them as a per-cell linearElastic material, and
    the zones of read_material_zones for the sub-mesh interfaces)."""
    d = read_dict(os.path.join(case_dir, "constant", "mechanicalProperties"))
    laws = d.get("mechanical")
    if not isinstance(laws, list):
        return _law_constants(laws)
    if len(laws) == 1:
        return _law_constants(laws[0][1])
    zones = read_cell_zones(case_dir)
    E, nu, rho = np.full(nCells, np.nan), np.full(nCells, np.nan), np.full(nCells, np.nan)
    for name, law in laws:
        if name not in zones:
            raise ValueError(f"mechanical law {name!r} has no cellZone of that name")
        e, v, r = _law_constants(law)
        cells = zones[name]
        if np.any(~np.isnan(E[cells])):
            raise ValueError(f"cellZone {name!r} overlaps another material's zone")
        E[cells], nu[cells] = e, v
        rho[cells] = np.nan if r is None else r
    if np.any(np.isnan(E)):
        raise ValueError("some cells belong to no material's cellZone")
    return E, nu, (None if np.all(np.isnan(rho)) else rho)


def read_material_zones(case_dir, nCells):
    """Material zone of every cell for several mechanical laws (the law's
    index in mechanicalProperties, on the cellZone of its name), None for
    one law.  The reference solves each law on its own sub-mesh
    (mechanicalModel.C:571-615, solidSubMeshes); the device path does the
    same given these zones (s4f_mesh_zones_set)."""
    d = read_dict(os.path.join(case_dir, "constant", "mechanicalProperties"))
    laws = d.get("mechanical")
    if not isinstance(laws, list) or len(laws) == 1:
        return None
    zones = read_cell_zones(case_dir)
    z = np.full(nCells, -1, np.int32)
    for k, (name, _) in enumerate(laws):
        if name not in zones:
            raise ValueError(f"mechanical law {name!r} has no cellZone of that name")
        z[zones[name]] = k
    if np.any(z < 0):
        raise ValueError("some cells belong to no material's cellZone")
    return z


def read_density(case_dir):
    """rho of the single mechanical law (None when absent)"""
    d = read_dict(os.path.join(case_dir, "constant", "mechanicalProperties"))
    laws = d.get("mechanical")
    law = laws[0][1] if isinstance(laws, list) and laws else laws
    return dimensioned(law["rho"]) if isinstance(law, dict) and "rho" in law else None


def read_gravity(case_dir):
    """constant/g (uniformDimensionedVectorField) or None"""
    path = os.path.join(case_dir, "constant", "g")
    if not (os.path.exists(path) or os.path.exists(path + ".gz")):
        return None
    v = read_dict(path).get("value")
    return np.asarray(v, np.float64).reshape(3)


def read_d2dt2_scheme(case_dir, field="D"):
    """d2dt2Schemes entry for d2dt2(D), default steadyState"""
    try:
        fvs = read_dict(os.path.join(case_dir, "system", "fvSchemes"))
    except FileNotFoundError:
        return "steadyState"
    sch = fvs.get("d2dt2Schemes", {}) or {}
    v = lookup(sch, f"d2dt2({field})", None) or sch.get("default", "steadyState")
    return str(v[0] if isinstance(v, list) else v)


def read_time_controls(case_dir):
    """controlDict startTime, endTime, deltaT, writeControl, writeInterval"""
    d = read_dict(os.path.join(case_dir, "system", "controlDict"))
    return dict(startTime=float(d.get("startTime", 0)), endTime=float(d.get("endTime", 1)),
                deltaT=float(d.get("deltaT", 1)), writeControl=str(d.get("writeControl", "timeStep")),
                writeInterval=float(d.get("writeInterval", 1)))


def _time_name(t):
    """OpenFOAM's default time directory name (general format, 6 digits)"""
    return f"{t:.6g}"


def read_solid_properties(case_dir):
    """solidModel type + coefficients with the reference defaults (solidModel.C:1147-1177)"""
    d = read_dict(os.path.join(case_dir, "constant", "solidProperties"))
    model = d["solidModel"]
    coeffs = lookup(d, model + "Coeffs", {}) or {}
    stab = coeffs.get("stabilisation", {"type": "RhieChow", "scaleFactor": 0.1})
    if stab.get("type", "RhieChow") != "RhieChow":
        raise ValueError("only the RhieChow stabilisation is on the device path")
    # solidModel::relaxField (solidModel.C:823-906): fixed or Aitken
    method = str(coeffs.get("relaxationMethod", "fixed"))
    if method not in ("fixed", "Aitken"):
        raise ValueError(f"relaxationMethod {method!r} is not on the device path (fixed, Aitken)")
    predictor = str(coeffs.get("predictor", "no")) in ("yes", "on", "true")
    opts = dict(nCorrectors=int(coeffs.get("nCorrectors", 10000)),
                solutionTolerance=float(coeffs.get("solutionTolerance", 1e-6)),
                alternativeTolerance=float(coeffs.get("alternativeTolerance", 1e-7)),
                rhieChowScale=float(stab.get("scaleFactor", 0.1)),
                interfaceScale=float(stab.get("interfaceScaleFactor", 0.01)))
    prec = coeffs.get("precision", "fp64")
    opts["precision"] = 1 if prec == "mixed" else 0
    opts["relaxationMethod"] = 1 if method == "Aitken" else 0
    opts["predictor"] = predictor
    return model, opts


def read_solver_controls(case_dir, field="D"):
    """fvSolution solvers entry for D (PCG tolerance / relTol / maxIter) and,
    when present, the field relaxation factor of D (relaxationFactors
    { fields { D a; } } or the older flat relaxationFactors { D a; })"""
    d = read_dict(os.path.join(case_dir, "system", "fvSolution"))
    s = lookup(d.get("solvers", {}), field, {}) or {}
    out = dict(pcgTolerance=float(s.get("tolerance", 1e-6)), pcgRelTol=float(s.get("relTol", 0.0)),
               pcgMaxIter=int(s.get("maxIter", 1000)))
    rf = d.get("relaxationFactors", {}) or {}
    fields = rf.get("fields", None)
    a = lookup(fields, field, None) if isinstance(fields, dict) else None
    if a is None and isinstance(rf.get(field, None), (int, float, str)):
        a = rf[field]
    if a is not None:
        out["relaxD"] = float(a)
    return out


def write_vol_field(path, name, cls, values, boundary=None, dims="[0 1 0 0 0 0 0]"):
    """ASCII vol*Field with nonuniform internal values; boundary: list of (patchName, values)"""
    values = np.asarray(values, np.float64)
    kind = {"volScalarField": "scalar", "volVectorField": "vector", "volSymmTensorField": "symmTensor",
            "volTensorField": "tensor"}[cls]

    def fmt(v):
        v = np.atleast_1d(v)
        return f"{v[0]:.10g}" if kind == "scalar" else "(" + " ".join(f"{x:.10g}" for x in v) + ")"

    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(_foam_header(cls, name))
        f.write(f"dimensions      {dims};\n\n")
        f.write(f"internalField   nonuniform List<{kind}>\n{len(values)}\n(\n")
        f.writelines(fmt(v) + "\n" for v in values)
        f.write(")\n;\n\nboundaryField\n{\n")
        for pname, pv in boundary or []:
            pv = np.asarray(pv)
            f.write(f"    {pname}\n    {{\n        type            calculated;\n")
            f.write(f"        value           nonuniform List<{kind}>\n{len(pv)}\n(\n")
            f.writelines(fmt(v) + "\n" for v in pv)
            f.write(")\n;\n    }\n")
        f.write("}\n")


def run_case(case_dir, ctx=None, write_time="1", solver_overrides=None):
    """Reads a solids4foam case (polyMesh, 0/D, mechanicalProperties,
    solidProperties, fvSolution, fvSchemes, constant/g), runs
    linearGeometryTotalDisplacement on the device and writes D, sigma,
    sigmaEq, epsilon, epsilonEq and pointD.  Steady (d2dt2 steadyState): one
    solve written to `write_time`.  Transient (Euler or backward): controlDict's time
    loop from startTime to endTime by deltaT, written at every writeInterval
    (time steps, or simulated time with writeControl runTime/adjustableRunTime)
    to directories named by the time."""
    from . import solid
    model, opts = read_solid_properties(case_dir)
    if model not in ("linearGeometryTotalDisplacement", "linearGeometryTotalDisplacementHIP"):
        raise ValueError(f"solidModel {model!r} is not on the device path")
    opts.update(read_solver_controls(case_dir))
    opts.update(solver_overrides or {})
    predictor = opts.pop("predictor", False)
    mesh = read_polymesh(case_dir)
    bnd = mesh["boundary"]
    pm = solid.PolyMesh.from_arrays(mesh["points"], mesh["fOff"], mesh["fVerts"], mesh["owner"], mesh["neighbour"],
                                    mesh["nCells"], [p["startFace"] for p in bnd], [p["nFaces"] for p in bnd])
    Cf = pm.geometry()["Cf"][pm.nInternalFaces:]
    bcs = read_D_boundary(case_dir, bnd, face_centres=Cf)
    E, nu, rho = read_materials(case_dir, pm.nCells)
    zones = read_material_zones(case_dir, pm.nCells)
    if os.path.exists(os.path.join(case_dir, "0", "sigma0")):  # the initial-stress field (mechanicalLaw.C:157-230)
        raise ValueError("0/sigma0 (initial stress field) is not supported on the device path")
    patches = [dict(type=b["type"], value=b.get("value"), traction=b.get("traction"), pressure=b.get("pressure"))
               for b in bcs]
    ctx = ctx or solid.Context(0)
    grad_scheme = ""
    try:
        fvs = read_dict(os.path.join(case_dir, "system", "fvSchemes"))
        grad_scheme = str(lookup(fvs.get("gradSchemes", {}), "grad(D)", "") or "")
    except FileNotFoundError:
        pass
    scheme = read_d2dt2_scheme(case_dir)
    if scheme not in ("steadyState", "Euler", "backward"):
        raise ValueError(f"d2dt2 scheme {scheme!r} is not on the device path (steadyState, Euler, backward)")
    g = read_gravity(case_dir)
    transient = scheme in ("Euler", "backward")
    tc = read_time_controls(case_dir) if transient else None
    if (transient or g is not None) and (rho is None or np.any(np.isnan(rho))):
        raise ValueError("mechanicalProperties gives no rho (needed for d2dt2 and g)")
    if rho is not None and np.any(np.isnan(rho)):
        rho = None
    m = solid.LinGeomTotalDispSolid(ctx, pm, patches, E, nu, solver=opts,
                                    plane_stress=read_plane_stress(case_dir),
                                    ls_full_delta="extendedLeastSquares" in grad_scheme, rho=rho, g=g,
                                    d2dt2=scheme, deltaT=tc["deltaT"] if tc else None, predictor=predictor,
                                    zones=zones)
    if not transient:
        m.evolve()
        _write_state(m, os.path.join(case_dir, write_time))
        return m
    t, step = tc["startTime"], 0
    m.time = t
    eps = 1e-9 * tc["deltaT"]
    while t < tc["endTime"] - eps:
        m.evolve()
        t, step = m.time, step + 1
        if tc["writeControl"] == "timeStep":
            write = step % max(1, int(round(tc["writeInterval"]))) == 0
        else:
            k = round(t / tc["writeInterval"])
            write = abs(t - k * tc["writeInterval"]) < eps
        if write or t >= tc["endTime"] - eps:
            _write_state(m, os.path.join(case_dir, _time_name(t)))
    return m


def _write_state(m, out):
    """D, sigma, sigmaEq, epsilon, epsilonEq and pointD of the last evolve()"""
    write_vol_field(os.path.join(out, "D"), "D", "volVectorField", m.D)
    write_vol_field(os.path.join(out, "sigma"), "sigma", "volSymmTensorField", m.sigma, dims="[1 -1 -2 0 0 0 0]")
    f = m.write_fields()
    write_vol_field(os.path.join(out, "sigmaEq"), "sigmaEq", "volScalarField", f["sigmaEq"],
                    dims="[1 -1 -2 0 0 0 0]")
    write_vol_field(os.path.join(out, "epsilon"), "epsilon", "volSymmTensorField", f["epsilon"], dims="[0 0 0 0 0 0 0]")
    write_vol_field(os.path.join(out, "epsilonEq"), "epsilonEq", "volScalarField", f["epsilonEq"],
                    dims="[0 0 0 0 0 0 0]")
    write_point_field(os.path.join(out, "pointD"), "pointD", f["pointD"])
    if getattr(m, "transient", False):
        write_vol_field(os.path.join(out, "U"), "U", "volVectorField", m.U, dims="[0 1 -1 0 0 0 0]")


def write_point_field(path, name, values, dims="[0 1 0 0 0 0 0]"):
    """ASCII pointVectorField with nonuniform internal values (calculated patches)"""
    values = np.asarray(values, np.float64)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(_foam_header("pointVectorField", name))
        f.write(f"dimensions      {dims};\n\n")
        f.write(f"internalField   nonuniform List<vector>\n{len(values)}\n(\n")
        f.writelines("(" + " ".join(f"{x:.10g}" for x in v) + ")\n" for v in values)
        f.write(")\n;\n\nboundaryField\n{\n    \".*\"\n    {\n        type            calculated;\n    }\n}\n")
