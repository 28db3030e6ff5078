"""Case descriptions for the coupled conjugate-heat-transfer solve.

A case is plain Python data (regions + monolithic interfaces) that renders to
the OpenFOAM-style dictionary the native case builder reads.  The dictionary
keeps the reference's vocabulary: region types ``transportTemperature`` /
``conductTemperature`` (src/regions/*), patch type ``monolithicTemperature``
for coupled patches (0/*/orig/monolithic/T) and ``monolithicCoupledPatches``
with ``coupledPatchPair`` (constant/regionInterfacePropertiesMonolithic).

The named workloads:

* ``plate2d`` -- 2-D steady conjugate heat transfer: fluid 200x50 over a
  solid 200x20, conformal (directMap) interface.
* ``duct3d``  -- 3-D fluid duct 160x80x80 (1.024M cells) over a solid slab
  200x100x19 (380k cells) whose face grid is 1.25x refined and offset by
  0.37 fluid cells: a non-conformal GGI interface, 1.404M cells total.
* ``batch``   -- many independent two-region systems of seeded uneven size.
* ``exchanger3`` -- three regions (fluid / wall / fluid), two GGI interfaces.
* ``serving_stream`` -- Poisson arrivals of two-region 3-D box solves of
  log-uniform size, half of them with a GGI interface.
"""
from __future__ import annotations

import math
import copy

SIDES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def _fmt(v):
    if isinstance(v, float):
        return repr(v)
    return str(v)


def region(name, rtype, n, origin, length, *, k, T0, patches, rhoCp=0.0,
           velocity=None):
    """One structured hex box region.

    ``patches`` maps side name -> dict(type=..., value=..., flux=...), where
    type is fixedValue / zeroGradient / empty / monolithicTemperature and flux
    is wall / inlet / outlet (fluid only).
    """
    pat = {}
    for s in SIDES:
        p = dict(patches.get(s, {"type": "zeroGradient"}))
        p.setdefault("flux", "wall")
        pat[s] = p
    return dict(name=name, type=rtype, nCells=tuple(int(x) for x in n),
                origin=tuple(float(x) for x in origin),
                length=tuple(float(x) for x in length), k=dict(k), T0=float(T0),
                rhoCp=float(rhoCp), velocity=dict(velocity) if velocity else None,
                patches=pat)


def interface(name, a, b, method="GGI"):
    return dict(name=name, A=tuple(a), B=tuple(b), method=method)


def case(regions, interfaces):
    return dict(regions=list(regions), interfaces=list(interfaces))


def to_dict_text(c) -> str:
    """Render a case as an OpenFOAM-style dictionary."""
    out = ["regions", "("]
    for r in c["regions"]:
        out.append("    %s" % r["name"])
        out.append("    {")
        out.append("        type %s;" % r["type"])
        out.append("        nCells (%d %d %d);" % r["nCells"])
        out.append("        origin (%s);" % " ".join(_fmt(x) for x in r["origin"]))
        out.append("        length (%s);" % " ".join(_fmt(x) for x in r["length"]))
        out.append("        rhoCp %s;" % _fmt(r["rhoCp"]))
        k = r["k"]
        if "uniform" in k:
            out.append("        k { type uniform; value %s; }" % _fmt(float(k["uniform"])))
        else:
            out.append("        k { type blockLogUniform; blockSize %d; min %s; max %s; seed %d; }"
                       % (k["block"], _fmt(float(k["min"])), _fmt(float(k["max"])), k["seed"]))
        v = r["velocity"]
        if v:
            out.append("        U { flowAxis %d; profileAxes (%s); Umax %s; noise %s; seed %d; }"
                       % (v["axis"], " ".join(str(a) for a in v.get("profile", ())),
                          _fmt(float(v["Umax"])), _fmt(float(v.get("noise", 0.0))),
                          int(v.get("seed", 0))))
        out.append("        T0 %s;" % _fmt(r["T0"]))
        out.append("        boundary")
        out.append("        {")
        for s in SIDES:
            p = r["patches"][s]
            items = ["type %s;" % p["type"]]
            if p["type"] == "fixedValue":
                items.append("value %s;" % _fmt(float(p["value"])))
            items.append("flux %s;" % p["flux"])
            out.append("            %s { %s }" % (s, " ".join(items)))
        out.append("        }")
        out.append("    }")
    out.append(");")
    out.append("monolithicCoupledPatches")
    out.append("(")
    for i in c["interfaces"]:
        out.append("    %s" % i["name"])
        out.append("    {")
        out.append("        coupledPatchPair ((%s %s) (%s %s));" % (i["A"] + i["B"]))
        out.append("        interfaceTransferMethod %s;" % i["method"])
        out.append("    }")
    out.append(");")
    return "\n".join(out) + "\n"


# --------------------------------------------------------------------------
# Named workloads
# --------------------------------------------------------------------------

def plate2d(nx=200, nyf=50, nys=20):
    """2-D steady CHT: fluid 1x0.25 m over solid 1x0.1 m, one cell deep."""
    depth = 1.0 / nx
    fluid = region(
        "fluid", "transportTemperature", (nx, nyf, 1), (0.0, 0.0, 0.0), (1.0, 0.25, depth),
        k=dict(uniform=0.6), T0=300.0, rhoCp=4.18e6,
        velocity=dict(axis=0, profile=(1,), Umax=1.0, noise=0.0, seed=0),
        patches=dict(
            xmin=dict(type="fixedValue", value=300.0, flux="inlet"),
            xmax=dict(type="zeroGradient", flux="outlet"),
            ymin=dict(type="monolithicTemperature"),
            ymax=dict(type="zeroGradient"),
            zmin=dict(type="empty"), zmax=dict(type="empty")))
    solid = region(
        "solid", "conductTemperature", (nx, nys, 1), (0.0, -0.1, 0.0), (1.0, 0.1, depth),
        k=dict(uniform=16.0), T0=350.0,
        patches=dict(
            ymin=dict(type="fixedValue", value=350.0),
            ymax=dict(type="monolithicTemperature"),
            zmin=dict(type="empty"), zmax=dict(type="empty")))
    return case([fluid, solid], [interface("fluidsolid", ("fluid", "ymin"), ("solid", "ymax"),
                                           method="directMap")])


def duct3d(nf=(160, 80, 80), ns=(200, 100, 19), offset=0.37, seed=20260101, conformal=False):
    """3-D two-region coupled temperature solve with a non-conformal GGI
    interface.  Fluid duct 1.0 x 0.5 x 0.5 m (cells 6.25 mm) over a solid
    slab whose face grid is 1.25x refined (cells 5 mm, 19 layers) and offset
    by `offset` fluid cells in x and y.  Solid k is log-uniform in [1,200] per
    10x10x10-cell block; fluid velocity is a duct profile plus 5% noise."""
    dxf = 1.0 / 160.0
    Lf = (nf[0] * dxf, nf[1] * dxf, nf[2] * dxf)
    if conformal:
        ds = dxf
        off = 0.0
    else:
        ds = 0.8 * dxf
        off = offset * dxf
    Ls = (ns[0] * ds, ns[1] * ds, ns[2] * ds)
    fluid = region(
        "fluid", "transportTemperature", nf, (0.0, 0.0, 0.0), Lf,
        k=dict(uniform=0.6), T0=300.0, rhoCp=4.18e6,
        velocity=dict(axis=0, profile=(1, 2), Umax=1.0, noise=0.05, seed=seed + 1),
        patches=dict(
            xmin=dict(type="fixedValue", value=300.0, flux="inlet"),
            xmax=dict(type="zeroGradient", flux="outlet"),
            zmin=dict(type="monolithicTemperature")))
    solid = region(
        "solid", "conductTemperature", ns, (off, off, -Ls[2]), Ls,
        k=dict(block=10, min=1.0, max=200.0, seed=seed + 2), T0=350.0,
        patches=dict(
            zmin=dict(type="fixedValue", value=350.0),
            zmax=dict(type="monolithicTemperature")))
    return case([fluid, solid], [interface("fluidsolid", ("fluid", "zmin"), ("solid", "zmax"),
                                           method="directMap" if conformal else "GGI")])


def _rng(seed):
    import random
    return random.Random(seed)


def box_pair(cells_fluid, cells_solid, kratio, seed, ggi=False, Umax=1.0):
    """A small two-region 3-D box system of about the given cell counts."""
    # fluid: aspect 2:1:1, solid: same footprint, thickness from cell count
    nfx = max(4, int(round((cells_fluid * 4.0) ** (1.0 / 3.0))))
    nfy = max(2, nfx // 2)
    nfz = max(2, int(round(cells_fluid / float(nfx * nfy))))
    d = 1.0 / nfx
    if ggi:
        ds = 0.8 * d
        nsx, nsy = int(round(nfx * 1.25)), int(round(nfy * 1.25))
        off = 0.37 * d
    else:
        ds, nsx, nsy, off = d, nfx, nfy, 0.0
    nsz = max(2, int(round(cells_solid / float(nsx * nsy))))
    kf = 0.6
    fluid = region(
        "fluid", "transportTemperature", (nfx, nfy, nfz), (0.0, 0.0, 0.0),
        (nfx * d, nfy * d, nfz * d), k=dict(uniform=kf), T0=300.0, rhoCp=4.18e6,
        velocity=dict(axis=0, profile=(1, 2), Umax=Umax, noise=0.0, seed=seed + 1),
        patches=dict(
            xmin=dict(type="fixedValue", value=300.0, flux="inlet"),
            xmax=dict(type="zeroGradient", flux="outlet"),
            zmin=dict(type="monolithicTemperature")))
    solid = region(
        "solid", "conductTemperature", (nsx, nsy, nsz), (off, off, -nsz * ds),
        (nsx * ds, nsy * ds, nsz * ds), k=dict(uniform=kf * kratio), T0=350.0,
        patches=dict(
            zmin=dict(type="fixedValue", value=350.0),
            zmax=dict(type="monolithicTemperature")))
    return case([fluid, solid], [interface("fluidsolid", ("fluid", "zmin"), ("solid", "zmax"),
                                           method="GGI" if ggi else "directMap")])


def batch(n_systems=64, lo=10_000, hi=500_000, seed=20260103):
    """Independent two-region systems, cells per region log-uniform in
    [lo, hi], conductivity ratio k_s/k_f log-uniform in [1, 1000]."""
    rng = _rng(seed)
    out = []
    for s in range(n_systems):
        cf = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        cs = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        kr = math.exp(rng.uniform(0.0, math.log(1000.0)))
        out.append(box_pair(cf, cs, kr, seed + 17 * s))
    return out


def exchanger3(n=(200, 120, 100), seed=20260105):
    """Three-region counter-flow exchanger: fluid A / solid wall / fluid B,
    each n cells, two non-conformal GGI interfaces (wall grid offset)."""
    rng = _rng(seed)
    d = 1.0 / n[0]
    L = (n[0] * d, n[1] * d, n[2] * d)
    TA, TB = rng.uniform(290.0, 370.0), rng.uniform(290.0, 370.0)
    off = 0.37 * d
    fa = region(
        "fluidA", "transportTemperature", n, (0.0, 0.0, L[2]), L,
        k=dict(uniform=0.6), T0=TA, rhoCp=4.18e6,
        velocity=dict(axis=0, profile=(1, 2), Umax=0.5, noise=0.0, seed=seed + 1),
        patches=dict(xmin=dict(type="fixedValue", value=TA, flux="inlet"),
                     xmax=dict(type="zeroGradient", flux="outlet"),
                     zmin=dict(type="monolithicTemperature")))
    wall = region(
        "wall", "conductTemperature", n, (off, off, 0.0), L,
        k=dict(block=10, min=1.0, max=200.0, seed=seed + 2), T0=0.5 * (TA + TB),
        patches=dict(zmin=dict(type="monolithicTemperature"),
                     zmax=dict(type="monolithicTemperature")))
    fb = region(
        "fluidB", "transportTemperature", n, (0.0, 0.0, -L[2]), L,
        k=dict(uniform=0.6), T0=TB, rhoCp=4.18e6,
        velocity=dict(axis=0, profile=(1, 2), Umax=-1.5, noise=0.0, seed=seed + 3),
        patches=dict(xmax=dict(type="fixedValue", value=TB, flux="inlet"),
                     xmin=dict(type="zeroGradient", flux="outlet"),
                     zmax=dict(type="monolithicTemperature")))
    return case([fa, wall, fb], [
        interface("fluidAwall", ("fluidA", "zmin"), ("wall", "zmax")),
        interface("wallfluidB", ("wall", "zmin"), ("fluidB", "zmax"))])


def scaled(c, factor):
    """Return a copy with every region's cell counts scaled (for miniatures)."""
    c = copy.deepcopy(c)
    for r in c["regions"]:
        r["nCells"] = tuple(max(1, int(round(n * factor))) if n > 1 else 1 for n in r["nCells"])
    return c


def serving_case(total_cells, ggi, seed):
    """A two-region 3-D duct-over-slab solve of about `total_cells` cells
    (fluid 73 %, solid 27 %, as in duct3d), GGI (1.25x refined, offset solid
    face grid) or conformal, seeded velocity noise and solid conductivity
    log-uniform in [1, 200] per 10-cell block."""
    m = max(3, int(round((0.73 * total_cells / 2.0) ** (1.0 / 3.0))))
    nf = (2 * m, m, m)
    if ggi:
        nsx, nsy = int(round(2 * m * 1.25)), int(round(m * 1.25))
    else:
        nsx, nsy = 2 * m, m
    nsz = max(2, int(round(0.27 * total_cells / float(nsx * nsy))))
    return duct3d(nf=nf, ns=(nsx, nsy, nsz), offset=0.37, seed=seed, conformal=not ggi)


def serving_stream(n_requests=100, rate=50.0, lo=20_000, hi=2_000_000, ggi_fraction=0.5, seed=20260104):
    """[(arrival time in s, case)] for Poisson arrivals at `rate` per second,
    total cells log-uniform in [lo, hi]."""
    rng = _rng(seed)
    t = 0.0
    out = []
    for i in range(n_requests):
        t += rng.expovariate(rate)
        cells = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        ggi = rng.random() < ggi_fraction
        out.append((t, serving_case(cells, ggi, seed + 31 * i)))
    return out
