"""Writing OpenFOAM case directories for the ingest tests.

* ``write_fixture_case(dest)`` rebuilds the flowOverHeatedPlate tutorial case
  (monolithic set-up) from tests/golden/flowOverHeatedPlate.npz: ASCII
  polyMesh of both regions and the dictionaries the temperature equation
  reads, with the values the fixture holds.
* ``write_box_case(dest, case)`` writes a structured two-region box case of
  ``multiregionfoam_amd.cases`` as a polyMesh case, cells and faces numbered
  as the native box builder numbers them, so the two assemblies can be
  compared entry by entry.

Test infrastructure: used by tests/ only.
"""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "flowOverHeatedPlate.npz")

HEADER = """FoamFile
{
    version     2.0;
    format      ascii;
    class       %s;
    object      %s;
}
"""


def _write(path, cls, obj, body):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(HEADER % (cls, obj))
        f.write(body)


def _num(x):
    return repr(float(x))


def write_poly_mesh(d, points, face_sizes, face_verts, owner, neighbour, patches):
    """patches: list of dict(name, type, nFaces, startFace[, extra dict text])."""
    pts = "\n".join("(%s %s %s)" % (_num(p[0]), _num(p[1]), _num(p[2])) for p in points)
    _write(os.path.join(d, "points"), "vectorField", "points", "%d\n(\n%s\n)\n" % (len(points), pts))
    off = np.concatenate([[0], np.cumsum(face_sizes)])
    faces = "\n".join("%d(%s)" % (face_sizes[i], " ".join(str(v) for v in face_verts[off[i]:off[i + 1]]))
                      for i in range(len(face_sizes)))
    _write(os.path.join(d, "faces"), "faceList", "faces", "%d\n(\n%s\n)\n" % (len(face_sizes), faces))
    for name, arr in (("owner", owner), ("neighbour", neighbour)):
        _write(os.path.join(d, name), "labelList", name,
               "%d\n(\n%s\n)\n" % (len(arr), "\n".join(str(int(v)) for v in arr)))
    items = []
    for p in patches:
        items.append("    %s\n    {\n        type %s;\n        nFaces %d;\n        startFace %d;\n%s    }"
                     % (p["name"], p["type"], p["nFaces"], p["startFace"], p.get("extra", "")))
    _write(os.path.join(d, "boundary"), "polyBoundaryMesh", "boundary",
           "%d\n(\n%s\n)\n" % (len(patches), "\n".join(items)))


def _value_text(v, n=None):
    if isinstance(v, np.ndarray) or (isinstance(v, list) and v and isinstance(v[0], (list, tuple, np.ndarray))):
        arr = np.asarray(v, dtype=np.float64)
        if arr.ndim == 1:
            inner = "\n".join(_num(x) for x in arr)
            return "nonuniform List<scalar> %d\n(\n%s\n)" % (len(arr), inner)
        inner = "\n".join("(%s)" % " ".join(_num(x) for x in row) for row in arr)
        return "nonuniform List<vector> %d\n(\n%s\n)" % (len(arr), inner)
    if isinstance(v, (list, tuple)):
        return "uniform (%s)" % " ".join(_num(x) for x in v)
    return "uniform %s" % _num(v)


def write_field(path, cls, internal, patches):
    """patches: {name: dict(type=..., value=...)}; values as uniform scalars /
    vectors or per-face arrays."""
    lines = ["dimensions [0 0 0 0 0 0 0];", "", "internalField %s;" % _value_text(internal), "",
             "boundaryField", "{"]
    for name, p in patches.items():
        lines.append("    %s" % name)
        lines.append("    {")
        lines.append("        type %s;" % p["type"])
        if "value" in p and p["type"] != "empty":
            lines.append("        value %s;" % _value_text(p["value"]))
        for key, v in p.items():
            if key not in ("type", "value"):
                lines.append("        %s %s;" % (key, v))
        lines.append("    }")
    lines.append("}")
    _write(path, cls, os.path.basename(path), "\n".join(lines) + "\n")


def write_dict(path, text):
    _write(path, "dictionary", os.path.basename(path), text)


# ---------------------------------------------------------------------------
# the tutorial fixture
# ---------------------------------------------------------------------------

def load_fixture():
    z = np.load(FIXTURE)
    meta = json.loads(str(z["meta_json"]))
    meshes = {}
    for region, _ in meta["regions"]:
        pts = np.stack([z["%s_p%d_values" % (region, a)][z["%s_p%d_index" % (region, a)]] for a in range(3)],
                       axis=1)
        sizes = z[region + "_face_sizes"].astype(np.int64)
        verts = np.cumsum(z[region + "_face_verts_delta"].astype(np.int64))
        owner = np.cumsum(z[region + "_owner_delta"].astype(np.int64))
        neighbour = np.cumsum(z[region + "_neighbour_delta"].astype(np.int64))
        meshes[region] = dict(points=pts, face_sizes=sizes, face_verts=verts, owner=owner, neighbour=neighbour,
                              patches=meta["patches"][region])
    return meta, meshes


DNA_T = ("DNA\n{\n    T\n    {\n        maxCoupleIter %d;\n        residualControl\n        {\n"
         "            maxJumpRes %s;\n            maxFluxRes %s;\n        }\n    }\n}\n")


# the tutorial's convection schemes (system/{fluid,solid}/fvSchemes): the
# fluid's temperature is convected with linearUpwind, which the reader
# refuses unless told to assemble upwind in its place (div="upwind")
DIV_SCHEMES = {
    "fluid": "divSchemes { default Gauss linear; div(phi,U) Gauss upwind; "
             "div(phi,T) Gauss linearUpwind Gauss linear; }\n",
    "solid": "divSchemes { default none; }\n",
}


def write_fixture_case(dest, params=None, ddt=None, partitioned=None):
    """params: the test suite's parameter map (genCases.py update_params),
    dict(kf=..., cp=...): the fluid's conductivity (transportProperties and
    0/fluid/k) and specific heat; optionally U (the fluid's uniform internal
    velocity) and T (dict region -> uniform internal temperature).  ddt
    replaces every region's ddt scheme.  partitioned: the tutorial's
    partitioned set-up instead of the monolithic one (what Allrun -p copies:
    regionInterfacePropertiesPartitioned, 0/*/orig/partitioned/T, the DNA
    controls of multiRegionProperties), dict(accType=..., relax=...,
    maxCoupleIter=20, maxJumpRes=1e-7, maxFluxRes=1e-7, K=1)."""
    meta, meshes = load_fixture()
    params = params or {}
    if "kf" in params:
        meta["transportProperties"]["fluid"]["k"] = float(params["kf"])
        meta["fields"]["fluid"]["k"]["internal"] = float(params["kf"])
    if "cp" in params:
        meta["transportProperties"]["fluid"]["cp"] = float(params["cp"])
    if "U" in params:
        meta["fields"]["fluid"]["U"]["internal"] = [float(x) for x in params["U"]]
    for region, value in params.get("T", {}).items():
        meta["fields"][region]["T"]["internal"] = float(value)
    if ddt is not None:
        for region in meta["ddt"]:
            meta["ddt"][region] = ddt
    regions_txt = " ".join("(%s (%s))" % (r, " ".join(t)) for r, t in meta["regions"])
    mrp = "regions (%s);\n" % regions_txt
    if partitioned is not None:
        pt = dict(accType="fixed", relax=1.0, maxCoupleIter=20, maxJumpRes=1e-7, maxFluxRes=1e-7, K=1.0)
        pt.update(partitioned)
        mrp += DNA_T % (pt["maxCoupleIter"], _num(pt["maxJumpRes"]), _num(pt["maxFluxRes"]))
        pairs = " ".join(
            "%s%s { interfaceType heatTransferInterface; coupledPatchPair ((%s %s) (%s %s)); coupledFields (T); "
            "heatTransferInterfaceCoeffs { K K [0 0 0 0 0 0 0] %s; } }" % (a[0], b[0], a[0], a[1], b[0], b[1],
                                                                         _num(pt["K"]))
            for a, b in meta["coupledPatchPairs"])
        write_dict(os.path.join(dest, "constant", "regionInterfaceProperties"),
                   "partitionedCoupledPatches ( %s );\nmonolithicCoupledPatches ( );\ninterfaceTransferMethod %s;\n"
                   % (pairs, meta["interfaceTransferMethod"]))
        for (ra, pa), (rb, pb) in meta["coupledPatchPairs"]:
            meta["fields"][ra]["T"]["patches"][pa] = dict(
                type="regionCoupledTemperatureJump", value=meta["fields"][ra]["T"]["internal"],
                interfaceType="heatTransferInterface", k='"k"', accType=pt["accType"], relax=_num(pt["relax"]),
                neighbourRegionName=rb, neighbourPatchName=pb, neighbourFieldName="T", secondOrder="false",
                nonOrthCorr="true", **({"couplingReuse": int(pt["couplingReuse"])} if "couplingReuse" in pt else {}))
            meta["fields"][rb]["T"]["patches"][pb] = dict(
                type="regionCoupledHeatFlux", value=meta["fields"][rb]["T"]["internal"],
                interfaceType="heatTransferInterface", k='"k"', neighbourRegionName=ra, neighbourPatchName=pa,
                neighbourFieldName="T", secondOrder="false", nonOrthCorr="true")
    else:
        pairs = " ".join("%s%s { coupledPatchPair ((%s %s) (%s %s)); }" % (a[0], b[0], a[0], a[1], b[0], b[1])
                         for a, b in meta["coupledPatchPairs"])
        write_dict(os.path.join(dest, "constant", "regionInterfaceProperties"),
                   "monolithicCoupledPatches ( %s );\ninterfaceTransferMethod %s;\n"
                   % (pairs, meta["interfaceTransferMethod"]))
    write_dict(os.path.join(dest, "constant", "multiRegionProperties"), mrp)
    write_dict(os.path.join(dest, "system", "controlDict"), "deltaT %s;\n" % _num(meta["deltaT"]))
    for region, types in meta["regions"]:
        m = meshes[region]
        write_poly_mesh(os.path.join(dest, "constant", region, "polyMesh"), m["points"], m["face_sizes"],
                        m["face_verts"], m["owner"], m["neighbour"], m["patches"])
        tp = meta["transportProperties"][region]
        write_dict(os.path.join(dest, "constant", region, "transportProperties"),
                   "".join("%s %s [0 0 0 0 0 0 0] %s;\n" % (k, k, _num(v)) for k, v in sorted(tp.items())))
        write_dict(os.path.join(dest, "system", region, "fvSchemes"),
                   "ddtSchemes { default %s; }\n%s" % (meta["ddt"][region], DIV_SCHEMES[region]))
        f = meta["fields"][region]
        for name in ("T", "k", "U"):
            if name not in f:
                continue
            cls = "volVectorField" if name == "U" else "volScalarField"
            pats = {p["name"]: f[name]["patches"][p["name"]] for p in m["patches"]}
            write_field(os.path.join(dest, "0", region, name), cls, f[name]["internal"], pats)
    return meta


# ---------------------------------------------------------------------------
# structured box cases in the native box builder's numbering
# ---------------------------------------------------------------------------

SIDES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def _tangential(a):
    return (1, 2) if a == 0 else (0, 2) if a == 1 else (0, 1)


def box_mesh(n, origin, length):
    """Points, faces, owner, neighbour and per-side patch faces of a box of
    n[0] x n[1] x n[2] uniform hex cells: cells i fastest, internal faces
    owner-ascending with neighbours +1, +nx, +nx*ny, side faces first
    tangential axis fastest (xmin, xmax, ymin, ymax, zmin, zmax)."""
    nx, ny, nz = n
    ax = [origin[a] + length[a] * np.arange(n[a] + 1) / n[a] for a in range(3)]
    P = lambda I, J, K: I + (nx + 1) * (J + (ny + 1) * K)
    points = [(ax[0][I], ax[1][J], ax[2][K]) for K in range(nz + 1) for J in range(ny + 1) for I in range(nx + 1)]

    def quad(a, I, J, K, outward_plus):
        # the face normal to axis a at lattice point (I, J, K), counter-
        # clockwise about +a when outward_plus
        e = [(0, 0, 0)] * 4
        if a == 0:
            e = [(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)]
        elif a == 1:
            e = [(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)]
        else:
            e = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
        v = [P(I + d[0], J + d[1], K + d[2]) for d in e]
        return v if outward_plus else v[::-1]

    faces, owner, neighbour = [], [], []
    cell = lambda i, j, k: i + nx * (j + ny * k)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = cell(i, j, k)
                if i + 1 < nx:
                    faces.append(quad(0, i + 1, j, k, True)); owner.append(c); neighbour.append(c + 1)
                if j + 1 < ny:
                    faces.append(quad(1, i, j + 1, k, True)); owner.append(c); neighbour.append(c + nx)
                if k + 1 < nz:
                    faces.append(quad(2, i, j, k + 1, True)); owner.append(c); neighbour.append(c + nx * ny)
    sides = {}
    for s, name in enumerate(SIDES):
        a, hi = s // 2, s % 2
        b, c2 = _tangential(a)
        lst = []
        for fc in range(n[c2]):
            for fb in range(n[b]):
                idx = [0, 0, 0]
                idx[a] = n[a] - 1 if hi else 0
                idx[b], idx[c2] = fb, fc
                lat = list(idx)
                lat[a] = n[a] if hi else 0
                lst.append((quad(a, lat[0], lat[1], lat[2], bool(hi)), cell(*idx)))
        sides[name] = lst
    return np.array(points), faces, owner, neighbour, sides


def rotation(axis, degrees):
    """Rotation matrix about `axis` (Rodrigues)."""
    k = np.asarray(axis, dtype=np.float64)
    k = k / np.linalg.norm(k)
    t = np.radians(degrees)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * (K @ K)


def write_box_case(dest, case, deltaT=1.0, rotate=None):
    """Write a case of multiregionfoam_amd.cases (uniform k, fixedValue /
    zeroGradient / empty / monolithicTemperature patches, steady) as a
    polyMesh case directory; `rotate` = (axis, degrees) turns the whole case
    (points and velocities) about the origin."""
    Rm = rotation(*rotate) if rotate else None
    regions_txt = " ".join("(%s (%s))" % (r["name"], r["type"]) for r in case["regions"])
    write_dict(os.path.join(dest, "constant", "multiRegionProperties"), "regions (%s);\n" % regions_txt)
    pairs = " ".join("%s { coupledPatchPair ((%s %s) (%s %s)); }" % (i["name"], *i["A"], *i["B"])
                     for i in case["interfaces"])
    write_dict(os.path.join(dest, "constant", "regionInterfaceProperties"),
               "monolithicCoupledPatches ( %s );\n" % pairs)
    write_dict(os.path.join(dest, "system", "controlDict"), "deltaT %s;\n" % _num(deltaT))
    for r in case["regions"]:
        pts, faces, owner, neighbour, sides = box_mesh(r["nCells"], r["origin"], r["length"])
        if Rm is not None:
            pts = pts @ Rm.T
        patches, allf, start = [], list(faces), len(faces)
        for name in SIDES:
            p = r["patches"][name]
            lst = sides[name] if p["type"] != "empty" else sides[name]
            patches.append(dict(name=name, type="empty" if p["type"] == "empty" else "patch",
                                nFaces=len(lst), startFace=start))
            for q, c in lst:
                allf.append(q)
                owner.append(c)
            start += len(lst)
        write_poly_mesh(os.path.join(dest, "constant", r["name"], "polyMesh"), pts,
                        np.array([len(q) for q in allf]), np.concatenate(allf), owner, neighbour, patches)
        transport = r["type"] == "transportTemperature"
        k = float(r["k"]["uniform"])
        write_dict(os.path.join(dest, "constant", r["name"], "transportProperties"),
                   "rho rho [1 -3 0 0 0 0 0] 1.0;\n%s %s [0 2 -2 -1 0 0 0] %s;\nk k [1 1 -3 -1 0 0 0] %s;\n"
                   % ("cp" if transport else "cv", "cp" if transport else "cv",
                      _num(r["rhoCp"] if transport else 1.0), _num(k)))
        write_dict(os.path.join(dest, "system", r["name"], "fvSchemes"), "ddtSchemes { default steadyState; }\n")
        Tp, kp = {}, {}
        for name in SIDES:
            p = r["patches"][name]
            t = p["type"]
            Tp[name] = {"type": t}
            if t == "fixedValue":
                Tp[name]["value"] = float(p["value"])
            kp[name] = {"type": "empty" if t == "empty" else "zeroGradient"}
        write_field(os.path.join(dest, "0", r["name"], "T"), "volScalarField", float(r["T0"]), Tp)
        write_field(os.path.join(dest, "0", r["name"], "k"), "volScalarField", k, kp)
        if transport:
            v = r["velocity"]
            n = r["nCells"]
            prof = lambda idx: np.prod([4.0 * ((idx[a] + 0.5) / n[a]) * (1.0 - (idx[a] + 0.5) / n[a])
                                        for a in v.get("profile", ())]) if v.get("profile") else 1.0
            U = np.zeros((int(np.prod(n)), 3))
            for kk in range(n[2]):
                for j in range(n[1]):
                    for i in range(n[0]):
                        U[i + n[0] * (j + n[1] * kk), v["axis"]] = v["Umax"] * prof((i, j, kk))
            Up = {}
            for s, name in enumerate(SIDES):
                p = r["patches"][name]
                if p["type"] == "empty":
                    Up[name] = {"type": "empty"}
                elif p.get("flux") == "inlet":
                    Ub = np.zeros((len(sides[name]), 3))
                    for f, (_, c) in enumerate(sides[name]):
                        Ub[f] = U[c]
                    Up[name] = {"type": "fixedValue", "value": Ub}
                elif p.get("flux") == "outlet":
                    Up[name] = {"type": "zeroGradient"}
                else:
                    Up[name] = {"type": "fixedValue", "value": [0.0, 0.0, 0.0]}
            if Rm is not None:
                U = U @ Rm.T
                for p in Up.values():
                    if "value" in p:
                        p["value"] = (np.atleast_2d(np.asarray(p["value"], dtype=np.float64)) @ Rm.T)
                        if p["value"].shape[0] == 1:
                            p["value"] = list(p["value"][0])
            write_field(os.path.join(dest, "0", r["name"], "U"), "volVectorField", U, Up)
