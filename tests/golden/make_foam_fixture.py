"""Write tests/golden/flowOverHeatedPlate.npz from the reference tutorial case
tutorials/conjugateHeatTransfer/flowOverHeatedPlate (monolithic set-up: the
files Allrun -m selects).

The fixture is data: both regions' polyMesh (points, faces, owner, neighbour,
boundary patches) in a lossless compact encoding (points as per-axis unique
coordinates plus indices, face vertex lists and owner / neighbour delta
coded), and the values of the case dictionaries the temperature equation
reads (fields, transport properties, time step, ddt scheme, region types,
coupled patch pair), stored as JSON.  tests/foam_fixture.py writes the case
back out as an OpenFOAM case directory.

Run in the build container (the reference is not on the GPU box):
    python tests/golden/make_foam_fixture.py
"""
import json
import os
import re

import numpy as np

CASE = "/root/reference/tutorials/conjugateHeatTransfer/flowOverHeatedPlate"
SUITE_REFS = "/root/reference/tutorials/conjugateHeatTransfer/flowOverHeatedPlate_testSuite/references"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flowOverHeatedPlate.npz")
CSV_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vynnycky")


def strip(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    h = text.find("FoamFile")
    if h >= 0:
        text = text[text.index("}", h) + 1:]
    return text


def list_body(path):
    t = strip(open(path).read())
    i = t.index("(")
    return int(t[:i].split()[-1]), t[i + 1:t.rindex(")")]


def blocks(text):
    """Top-level 'name { ... }' blocks of a text (no nested braces)."""
    return {m.group(1): m.group(2) for m in re.finditer(r"([\w\-]+)\s*\{([^{}]*)\}", text)}


def entry(body, key):
    m = re.search(r"(?:^|[\s;{])" + re.escape(key) + r"\s+([^;]*);", body)
    return m.group(1).strip() if m else None


def value_of(text, raw, internal):
    raw = raw.strip()
    if raw == "$internalField":
        raw = internal
    m = re.fullmatch(r"uniform\s+\(([^)]*)\)", raw)
    if m:
        return [float(x) for x in m.group(1).split()]
    m = re.fullmatch(r"uniform\s+(\S+)", raw)
    if m:
        return float(m.group(1))
    raise ValueError("unsupported value '%s'" % raw)


def field(path):
    t = strip(open(path).read())
    internal = entry(t, "internalField")
    bf = t[t.index("boundaryField"):]
    bf = bf[bf.index("{") + 1:bf.rindex("}")]
    patches = {}
    for name, body in blocks(bf).items():
        p = {"type": entry(body, "type")}
        v = entry(body, "value")
        if v is not None and p["type"] != "empty":
            p["value"] = value_of(t, v, internal)
        patches[name] = p
    return {"internal": value_of(t, internal, internal), "patches": patches}


def props(path, keys):
    t = strip(open(path).read())
    out = {}
    for k in keys:
        m = re.search(r"(?:^|\s)" + k + r"\s+[^;]*?(\S+)\s*;", t)
        out[k] = float(m.group(1))
    return out


def mesh(region, arrays, meta):
    d = os.path.join(CASE, "constant", region, "polyMesh")
    n, b = list_body(os.path.join(d, "points"))
    pts = np.array(b.replace("(", " ").replace(")", " ").split(), dtype=np.float64).reshape(n, 3)
    for a in range(3):
        u, inv = np.unique(pts[:, a], return_inverse=True)
        assert np.array_equal(u[inv], pts[:, a])
        arrays["%s_p%d_values" % (region, a)] = u
        arrays["%s_p%d_index" % (region, a)] = inv.astype(np.int32)
    n, b = list_body(os.path.join(d, "faces"))
    tok = np.array(b.replace("(", " ").replace(")", " ").split(), dtype=np.int64)
    sizes, verts, i = [], [], 0
    while i < len(tok):
        k = int(tok[i])
        sizes.append(k)
        verts.extend(tok[i + 1:i + 1 + k])
        i += 1 + k
    assert len(sizes) == n
    arrays[region + "_face_sizes"] = np.array(sizes, dtype=np.int32)
    arrays[region + "_face_verts_delta"] = np.diff(np.array(verts, dtype=np.int64), prepend=0).astype(np.int32)
    for f in ("owner", "neighbour"):
        n, b = list_body(os.path.join(d, f))
        x = np.array(b.split(), dtype=np.int64)
        assert len(x) == n
        arrays["%s_%s_delta" % (region, f)] = np.diff(x, prepend=0).astype(np.int32)
    t = strip(open(os.path.join(d, "boundary")).read())
    body = t[t.index("(") + 1:t.rindex(")")]
    meta["patches"][region] = [
        {"name": name, "type": entry(pb, "type"), "nFaces": int(entry(pb, "nFaces")),
         "startFace": int(entry(pb, "startFace"))}
        for name, pb in blocks(body).items()]


def main():
    arrays, meta = {}, {"patches": {}}
    t = strip(open(os.path.join(CASE, "constant", "multiRegionProperties")).read())
    meta["regions"] = [[m.group(1), m.group(2).split()]
                       for m in re.finditer(r"\(\s*(\w+)\s*\(([^()]*)\)\s*\)", t)]
    for region, _ in meta["regions"]:
        mesh(region, arrays, meta)
    meta["fields"] = {}
    for region, types in meta["regions"]:
        f = {"T": field(os.path.join(CASE, "0", region, "orig", "monolithic", "T")),
             "k": field(os.path.join(CASE, "0", region, "orig", "monolithic", "k"))}
        if "transportTemperature" in types:
            f["U"] = field(os.path.join(CASE, "0", region, "U"))
        meta["fields"][region] = f
    meta["transportProperties"] = {
        "fluid": props(os.path.join(CASE, "constant", "fluid", "transportProperties"), ["rho", "cp", "k"]),
        "solid": props(os.path.join(CASE, "constant", "solid", "transportProperties"), ["rho", "cv", "k"])}
    meta["deltaT"] = float(entry(strip(open(os.path.join(CASE, "system", "controlDict")).read()), "deltaT"))
    meta["ddt"] = {}
    for region, _ in meta["regions"]:
        t = strip(open(os.path.join(CASE, "system", region, "fvSchemes")).read())
        meta["ddt"][region] = entry(blocks(t)["ddtSchemes"], "default")
    t = strip(open(os.path.join(CASE, "constant", "regionInterfacePropertiesMonolithic")).read())
    pairs = re.findall(r"coupledPatchPair\s*\(\s*\(\s*(\w+)\s+(\w+)\s*\)\s*\(\s*(\w+)\s+(\w+)\s*\)\s*\)", t)
    meta["coupledPatchPairs"] = [[[a, b], [c, d]] for a, b, c, d in pairs]
    meta["interfaceTransferMethod"] = entry(t, "interfaceTransferMethod")
    # the coupled solver entry of region 0's fvSolution (multiRegionSystem.C:150-153)
    t = strip(open(os.path.join(CASE, "system", "fluid", "fvSolution")).read())
    i = t.index("Tcoupled")
    i = t.index("{", i)
    depth, j = 0, i
    while True:
        depth += {"{": 1, "}": -1}.get(t[j], 0)
        if depth == 0:
            break
        j += 1
    body = t[i + 1:j]
    pre = re.search(r"preconditioner\s*\{([^{}]*)\}", body)
    flat = body[:pre.start()] + body[pre.end():] if pre else body
    meta["Tcoupled"] = {k: entry(flat, k) for k in ("solver", "tolerance", "relTol", "minIter", "maxIter")}
    meta["Tcoupled"]["preconditioner"] = entry(pre.group(1), "preconditioner") if pre else entry(flat, "preconditioner")
    meta["source"] = "tutorials/conjugateHeatTransfer/flowOverHeatedPlate (monolithic)"
    arrays["meta_json"] = np.array(json.dumps(meta, sort_keys=True))
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    # the test suite's reference curves theta(x) (Vynnycky et al. 1998),
    # data files copied as they are
    os.makedirs(CSV_DIR, exist_ok=True)
    for f in sorted(os.listdir(SUITE_REFS)):
        if f.endswith(".csv"):
            with open(os.path.join(SUITE_REFS, f)) as src, open(os.path.join(CSV_DIR, f), "w") as dst:
                dst.write(src.read())
    print("copied the reference curves to", CSV_DIR)


if __name__ == "__main__":
    main()
