"""Reading OpenFOAM case directories (mrfg_case_read_foam): the reference
tutorial flowOverHeatedPlate (monolithic set-up) from the committed fixture,
and structured box cases written as polyMesh.

Pinning of the ingest, independent of its code:
* the fixture rebuilds the tutorial exactly (against the reference's own
  case directory where it is present);
* geometry from bounding boxes of the hex cells (numpy, independent of the
  pyramid / triangle-fan decomposition the reader uses): cell volumes in the
  Euler ddt terms, the Gauss-linear Laplacian coefficients k |S| / |d|, the
  upwind convection split rho*cp*phi with phi = U . S for the uniform U;
* the interface coefficient's closed form, series resistance
  |S| / (d_A / k_A + d_B / k_B);
* a box case written as polyMesh assembles to the native box builder's
  system (itself pinned against the oracle) entry by entry;
* the oracle's solve of the read tutorial equals a direct sparse solve.
The GPU solve of the tutorial, with the tutorial's own Tcoupled solver entry,
is held to the oracle's (bit-identical FP64, as for every other case).
"""
import os

import numpy as np
import pytest

import multiregionfoam_amd as mrf
from multiregionfoam_amd import cases
import foam_fixture as ff
import oracle_bind as ob

REF_CASE = "/root/reference/tutorials/conjugateHeatTransfer/flowOverHeatedPlate"


def _lib_or_skip():
    try:
        mrf.lib()
    except ImportError as e:
        pytest.skip("libmrfg.so not built: %s" % e)


@pytest.fixture(scope="module")
def tutorial_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("flowOverHeatedPlate")
    ff.write_fixture_case(str(d))
    return str(d)


def _arrays(case):
    return case.regions(), case.interfaces()


def _hex_geometry(mesh):
    """Per-cell bounding boxes and per-face bounding boxes of an axis-aligned
    hex mesh, from the points alone."""
    P = mesh["points"]
    sizes, verts = mesh["face_sizes"], mesh["face_verts"]
    off = np.concatenate([[0], np.cumsum(sizes)])
    fid = np.repeat(np.arange(len(sizes)), sizes)
    flo = np.full((len(sizes), 3), np.inf)
    fhi = np.full((len(sizes), 3), -np.inf)
    np.minimum.at(flo, fid, P[verts])
    np.maximum.at(fhi, fid, P[verts])
    owner, nei = mesh["owner"], mesh["neighbour"]
    n = int(max(owner.max(), nei.max())) + 1
    clo = np.full((n, 3), np.inf)
    chi = np.full((n, 3), -np.inf)
    for cells, faces in ((owner, np.arange(len(owner))), (nei, np.arange(len(nei)))):
        np.minimum.at(clo, cells, flo[faces])
        np.maximum.at(chi, cells, fhi[faces])
    return flo, fhi, clo, chi


def test_fixture_rebuilds_reference_case(tutorial_dir):
    """The fixture's case directory assembles to exactly the system of the
    reference's own case directory."""
    _lib_or_skip()
    if not os.path.isdir(REF_CASE):
        pytest.skip("reference case not present")
    a = mrf.Case.from_foam(REF_CASE, div="upwind")
    b = mrf.Case.from_foam(tutorial_dir, div="upwind")
    for x, y in zip(a.regions() + a.interfaces(), b.regions() + b.interfaces()):
        assert x.keys() == y.keys()
        for k in x:
            assert np.array_equal(np.asarray(x[k]), np.asarray(y[k])), k


def test_tutorial_structure(tutorial_dir):
    _lib_or_skip()
    meta, meshes = ff.load_fixture()
    c = mrf.Case.from_foam(tutorial_dir, div="upwind")
    R, I = _arrays(c)
    assert [r["nCells"] for r in R] == [13612, 8200]
    assert [r["nFaces"] for r in R] == [len(meshes["fluid"]["neighbour"]), len(meshes["solid"]["neighbour"])]
    assert not R[0]["symmetric"] and R[1]["symmetric"]
    assert len(I) == 1 and I[0]["nA"] == 200 and I[0]["nB"] == 200
    # conformal (directMap): each face gathers exactly one partner with weight 1
    for s in "AB":
        assert np.array_equal(np.diff(I[0]["off" + s]), np.ones(200, np.int32))
        assert np.array_equal(I[0]["w" + s], np.ones(200))
    # the face pairing is a bijection
    assert sorted(I[0]["idxA"]) == list(range(200))
    assert np.array_equal(I[0]["idxB"][I[0]["idxA"]], np.arange(200))


def test_tutorial_geometry_and_coefficients(tutorial_dir):
    """Euler ddt volumes, Laplacian and convection coefficients of the read
    tutorial against bounding-box geometry of the fixture mesh."""
    _lib_or_skip()
    meta, meshes = ff.load_fixture()
    c = mrf.Case.from_foam(tutorial_dir, div="upwind")
    R, I = _arrays(c)
    dt = meta["deltaT"]
    tp = meta["transportProperties"]
    for g, region in enumerate(("fluid", "solid")):
        m = meshes[region]
        flo, fhi, clo, chi = _hex_geometry(m)
        V = np.prod(chi - clo, axis=1)
        C = 0.5 * (clo + chi)
        rhoC = tp[region]["rho"] * (tp[region]["cp"] if region == "fluid" else tp[region]["cv"])
        k = tp[region]["k"]
        T0 = meta["fields"][region]["T"]["internal"]
        # cells whose faces are all internal or zeroGradient / empty: source
        # is the Euler term rho*c*V*T_old/dt alone
        patches = m["patches"]
        bc = meta["fields"][region]["T"]["patches"]
        touched = np.zeros(R[g]["nCells"], bool)
        for p in patches:
            if bc[p["name"]]["type"] in ("fixedValue",):
                touched[m["owner"][p["startFace"]:p["startFace"] + p["nFaces"]]] = True
        src = R[g]["source"]
        np.testing.assert_allclose(src[~touched], rhoC * V[~touched] * T0 / dt, rtol=1e-12)
        # Laplacian: -(k |S| / |d|) in the upper coefficient (solid), and the
        # same plus the upwind split in the fluid
        nf = R[g]["nFaces"]
        ext = fhi[:nf] - flo[:nf]
        normal = np.argmin(ext, axis=1)
        area = np.prod(np.where(np.eye(3, dtype=bool)[normal], 1.0, ext), axis=1)
        o, n = m["owner"][:nf], m["neighbour"]
        dist = np.abs(C[n, normal] - C[o, normal])
        lap = k * area / dist
        if region == "solid":
            np.testing.assert_allclose(-R[g]["upper"], lap, rtol=1e-12)
            np.testing.assert_allclose(R[g]["lower"], R[g]["upper"], rtol=0)
        else:
            # U uniform (1 0 0): phi = S_x, signed owner -> neighbour
            U = np.asarray(meta["fields"]["fluid"]["U"]["internal"])
            phi = np.where(normal == 0, area * np.sign(C[n, 0] - C[o, 0]) * U[0], 0.0)
            np.testing.assert_allclose(R[g]["upper"] - R[g]["lower"], rhoC * phi, rtol=1e-12, atol=1e-9)
            np.testing.assert_allclose(-np.minimum(R[g]["upper"], R[g]["lower"]),
                                       lap + rhoC * np.abs(phi) * 1.0, rtol=1e-12)
    # interface coefficient: series resistance of the two half cells
    it = I[0]
    mf, ms = meshes["fluid"], meshes["solid"]
    gf, gs = _hex_geometry(mf), _hex_geometry(ms)
    pf = next(p for p in mf["patches"] if p["name"] == "interface")
    ps = next(p for p in ms["patches"] if p["name"] == "top")
    ff_ = np.arange(pf["startFace"], pf["startFace"] + pf["nFaces"])
    fs_ = np.arange(ps["startFace"], ps["startFace"] + ps["nFaces"])
    dA = np.abs(gf[0][ff_, 1] - 0.5 * (gf[2][mf["owner"][ff_], 1] + gf[3][mf["owner"][ff_], 1]))
    dB = np.abs(gs[0][fs_, 1] - 0.5 * (gs[2][ms["owner"][fs_], 1] + gs[3][ms["owner"][fs_], 1]))
    areaA = (gf[1][ff_, 0] - gf[0][ff_, 0]) * (gf[1][ff_, 2] - gf[0][ff_, 2])
    kA, kB = tp["fluid"]["k"], tp["solid"]["k"]
    expect = areaA / (dA / kA + dB[it["idxA"]] / kB)
    np.testing.assert_allclose(it["coeffA"], expect, rtol=1e-12)


def _plate_small():
    return cases.plate2d(nx=12, nyf=5, nys=4)


def _duct_small():
    """3-D two-box case with a non-conformal GGI interface (offset, refined
    solid faces), uniform k and a two-axis parabolic inflow."""
    fluid = cases.region(
        "fluid", "transportTemperature", (6, 5, 4), (0.0, 0.0, 0.0), (0.6, 0.5, 0.4),
        k=dict(uniform=0.6), T0=300.0, rhoCp=4.18e6,
        velocity=dict(axis=0, profile=(1, 2), Umax=0.2, noise=0.0, seed=0),
        patches=dict(xmin=dict(type="fixedValue", value=300.0, flux="inlet"),
                     xmax=dict(type="zeroGradient", flux="outlet"),
                     ymin=dict(type="monolithicTemperature")))
    solid = cases.region(
        "solid", "conductTemperature", (7, 3, 6), (0.037, -0.2, -0.02), (0.63, 0.2, 0.45),
        k=dict(uniform=16.0), T0=350.0,
        patches=dict(ymin=dict(type="fixedValue", value=350.0), ymax=dict(type="monolithicTemperature")))
    return cases.case([fluid, solid], [cases.interface("fluidsolid", ("fluid", "ymin"), ("solid", "ymax"))])


@pytest.mark.parametrize("make", [_plate_small, _duct_small], ids=["plate2d", "duct_ggi"])
def test_box_case_matches_native_builder(tmp_path, make):
    """A box case written as an OpenFOAM polyMesh case assembles to the
    native box builder's system: addressing exactly, coefficients to
    rounding."""
    _lib_or_skip()
    case = make()
    ff.write_box_case(str(tmp_path), case)
    a = mrf.Case(case)
    b = mrf.Case.from_foam(str(tmp_path), div="upwind")
    for x, y in zip(a.regions(), b.regions()):
        for k in ("nCells", "nFaces", "symmetric"):
            assert x[k] == y[k], k
        for k in ("lAddr", "uAddr"):
            assert np.array_equal(x[k], y[k]), k
        scale = np.abs(x["diag"]).max()
        for k in ("diag", "lower", "upper", "source", "psi"):
            np.testing.assert_allclose(y[k], x[k], rtol=1e-12, atol=1e-13 * scale, err_msg=k)
    for x, y in zip(a.interfaces(), b.interfaces()):
        for k in ("regionA", "regionB", "nA", "nB"):
            assert x[k] == y[k], k
        for k in ("fcA", "fcB", "offA", "offB", "idxA", "idxB"):
            assert np.array_equal(x[k], y[k]), k
        for k in ("wA", "wB", "coeffA", "coeffB"):
            np.testing.assert_allclose(y[k], x[k], rtol=1e-12, err_msg=k)


def test_tutorial_oracle_solve_matches_direct(tutorial_dir):
    """The oracle's DILU-BiCGStab on the read tutorial system against a
    direct sparse solve."""
    _lib_or_skip()
    import scipy.sparse.linalg as sla
    c = mrf.Case.from_foam(tutorial_dir, div="upwind")
    O = ob.OracleSystem.from_arrays(*_arrays(c))
    psi, perf = O.solve(tol=1e-13, maxIter=1000)
    assert perf["converged"]
    A, b = O.to_scipy(), np.concatenate([r["source"] for r in O.regions])
    x = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(np.concatenate(psi) - x) / np.linalg.norm(x) < 1e-10


def test_read_errors(tmp_path):
    _lib_or_skip()
    with pytest.raises(mrf.MrfgError, match="multiRegionProperties"):
        mrf.Case.from_foam(str(tmp_path), div="upwind")
    ff.write_box_case(str(tmp_path), _plate_small())
    with pytest.raises(mrf.MrfgError, match="ddt scheme"):
        mrf.Case.from_foam(str(tmp_path), ddt="CrankNicolson", div="upwind")


@pytest.mark.gpu
def test_tutorial_gpu_matches_oracle(gpu, tutorial_dir):
    """The tutorial's coupled T system, read from disk and solved on the GPU
    with the tutorial's own Tcoupled entry (BiCGStab, Cholesky, tolerance
    1e-15, maxIter 200), equals the oracle's solve."""
    meta, _ = ff.load_fixture()
    t = meta["Tcoupled"]
    fvs = ("solvers { Tcoupled { solver %s; preconditioner { preconditioner %s; } tolerance %s; relTol %s; "
           "minIter %s; maxIter %s; } }" % (t["solver"], t["preconditioner"], t["tolerance"], t["relTol"],
                                              t["minIter"], t["maxIter"]))
    c = gpu.Case.from_foam(tutorial_dir, div="upwind")
    A = gpu.CoupledFvMatrix(0)
    A.add_case(c)
    A.commit()
    perf = A.solve(fv_solution=fvs, entry="Tcoupled")[0]
    O = ob.OracleSystem.from_arrays(*_arrays(c))
    psi_ref, pref = O.solve(tol=float(t["tolerance"]), relTol=float(t["relTol"]), minIter=int(t["minIter"]),
                            maxIter=int(t["maxIter"]))
    assert perf.n_iterations == pref["nIterations"]
    assert perf.final_residual == pref["finalResidual"]
    assert np.array_equal(np.concatenate(A.psi(0)), np.concatenate(psi_ref))


# ---------------------------------------------------------------------------
# time stepping (ddt terms rebuilt per step; fixed flow)
# ---------------------------------------------------------------------------

def _ddt(case):
    return [({k: R[k] for k in ("ddtCoeff", "diagBase", "sourceBase")} if "ddtCoeff" in R else None)
            for R in case.regions()]


def _conduct_pair():
    """Two conducting boxes with a conformal interface and insulated outer
    walls: the coupled matrix without the ddt terms has zero column sums."""
    a = cases.region("hot", "conductTemperature", (8, 5, 1), (0.0, 0.0, 0.0), (0.8, 0.5, 0.1),
                     k=dict(uniform=3.0), T0=350.0,
                     patches=dict(ymin=dict(type="monolithicTemperature"), zmin=dict(type="empty"),
                                  zmax=dict(type="empty")))
    b = cases.region("cold", "conductTemperature", (8, 4, 1), (0.0, -0.4, 0.0), (0.8, 0.4, 0.1),
                     k=dict(uniform=40.0), T0=300.0,
                     patches=dict(ymax=dict(type="monolithicTemperature"), zmin=dict(type="empty"),
                                  zmax=dict(type="empty")))
    return cases.case([a, b], [cases.interface("hotcold", ("hot", "ymin"), ("cold", "ymax"), method="directMap")])


def test_ddt_split(tutorial_dir):
    """diag = diagBase + a and source = sourceBase + a*T0 exactly, with
    a = rho*c*V/dt from bounding-box cell volumes."""
    _lib_or_skip()
    meta, meshes = ff.load_fixture()
    c = mrf.Case.from_foam(tutorial_dir, div="upwind")
    tp = meta["transportProperties"]
    for g, (region, R) in enumerate(zip(("fluid", "solid"), c.regions())):
        assert np.array_equal(R["diag"], R["diagBase"] + R["ddtCoeff"])
        assert np.array_equal(R["source"], R["sourceBase"] + R["ddtCoeff"] * R["psi"])
        _, _, clo, chi = _hex_geometry(meshes[region])
        rhoC = tp[region]["rho"] * (tp[region]["cp"] if region == "fluid" else tp[region]["cv"])
        np.testing.assert_allclose(R["ddtCoeff"], rhoC * np.prod(chi - clo, axis=1) / meta["deltaT"], rtol=1e-12)
    s = mrf.Case.from_foam(tutorial_dir, ddt="steadyState", div="upwind")
    assert all("ddtCoeff" not in R for R in s.regions())


@pytest.mark.parametrize("scheme", ["Euler", "backward"])
def test_time_steps_conserve_energy(tmp_path, scheme):
    """Insulated conduction pair: Euler keeps sum(a T) and backward keeps
    sum(a (3 T_n+1 - T_n) / 2) from step to step, to the solver tolerance."""
    _lib_or_skip()
    ff.write_box_case(str(tmp_path), _conduct_pair(), deltaT=0.05)
    c = mrf.Case.from_foam(str(tmp_path), ddt=scheme, div="upwind")
    O = ob.OracleSystem.from_arrays(*_arrays(c))
    ddt = _ddt(c)
    a = np.concatenate([D["ddtCoeff"] for D in ddt])
    T0 = np.concatenate([R["psi"] for R in c.regions()])
    steps = O.time_steps(ddt, 6, scheme=scheme, tol=1e-15, maxIter=2000)
    T = [T0] + [np.concatenate(p) for p, _ in steps]
    assert np.abs(T[-1] - T[0]).max() > 1.0           # heat did flow
    if scheme == "Euler":
        e = [np.dot(a, t) for t in T]
    else:
        e = [np.dot(a, T[1])] + [np.dot(a, 1.5 * T[n + 1] - 0.5 * T[n]) for n in range(1, len(T) - 1)]
    np.testing.assert_allclose(e, e[0], rtol=1e-12)


def test_time_steps_reach_steady_state(tutorial_dir):
    """With a very large time step the ddt terms vanish: Euler steps of the
    tutorial approach the steady system's solution, the first one to 1e-6."""
    _lib_or_skip()
    import scipy.sparse.linalg as sla
    s = mrf.Case.from_foam(tutorial_dir, ddt="steadyState", div="upwind")
    Os = ob.OracleSystem.from_arrays(*_arrays(s))
    Ts = sla.spsolve(Os.to_scipy().tocsc(), np.concatenate([r["source"] for r in Os.regions]))
    c = mrf.Case.from_foam(tutorial_dir, ddt="Euler", deltaT=1e6, div="upwind")
    O = ob.OracleSystem.from_arrays(*_arrays(c))
    steps = O.time_steps(_ddt(c), 2, scheme="Euler", tol=1e-15, maxIter=2000)
    err = [np.linalg.norm(np.concatenate(p) - Ts) / np.linalg.norm(Ts) for p, _ in steps]
    assert err[0] < 1e-6 and err[1] < err[0]


@pytest.mark.gpu
def test_time_steps_gpu_match_oracle(gpu, tutorial_dir):
    """Backward time steps of the tutorial (the first Euler) on the GPU,
    ddt terms rebuilt on the device, equal the oracle's step by step; after
    reset_psi a run starts anew."""
    meta, _ = ff.load_fixture()
    t = meta["Tcoupled"]
    kw = dict(tol=float(t["tolerance"]), relTol=float(t["relTol"]), minIter=int(t["minIter"]),
              maxIter=int(t["maxIter"]))
    c = gpu.Case.from_foam(tutorial_dir, div="upwind")
    A = gpu.CoupledFvMatrix(0)
    A.add_case(c)
    A.commit()
    ctrl = gpu.controls(solver=t["solver"], tolerance=kw["tol"], relTol=kw["relTol"], minIter=kw["minIter"],
                        maxIter=kw["maxIter"])
    perf = A.time_steps(4, ddt="backward", ctrl=ctrl)
    O = ob.OracleSystem.from_arrays(*_arrays(c))
    ref = O.time_steps(_ddt(c), 4, scheme="backward", **kw)
    for (p,), (_, pr) in zip(perf, ref):
        assert p.n_iterations == pr["nIterations"] and p.final_residual == pr["finalResidual"]
    assert np.array_equal(np.concatenate(A.psi(0)), np.concatenate(ref[-1][0]))
    A.reset_psi()
    again = A.time_steps(2, ddt="backward", ctrl=ctrl)
    assert [q[0].final_residual for q in again] == [q[0].final_residual for q in perf[:2]]


# ---------------------------------------------------------------------------
# GGI on planar interfaces of any orientation (convex polygon overlaps)
# ---------------------------------------------------------------------------

def _covering_pair():
    """Fluid over solid with one interface outline and non-conformal faces
    (6 x 4 against 7 x 5 per side, the solid's faces also narrower)."""
    fluid = cases.region(
        "fluid", "transportTemperature", (6, 3, 4), (0.0, 0.0, 0.0), (0.6, 0.3, 0.4),
        k=dict(uniform=0.6), T0=300.0, rhoCp=4.18e6,
        velocity=dict(axis=0, profile=(1, 2), Umax=0.1, noise=0.0, seed=0),
        patches=dict(xmin=dict(type="fixedValue", value=300.0, flux="inlet"),
                     xmax=dict(type="zeroGradient", flux="outlet"),
                     ymin=dict(type="monolithicTemperature")))
    solid = cases.region(
        "solid", "conductTemperature", (7, 2, 5), (0.0, -0.2, 0.0), (0.6, 0.2, 0.4),
        k=dict(uniform=16.0), T0=350.0,
        patches=dict(ymin=dict(type="fixedValue", value=350.0), ymax=dict(type="monolithicTemperature")))
    return cases.case([fluid, solid], [cases.interface("fluidsolid", ("fluid", "ymin"), ("solid", "ymax"))])


@pytest.mark.parametrize("make", [_covering_pair, _duct_small], ids=["covering", "duct_ggi"])
def test_rotated_case_assembles_the_same(tmp_path, make):
    """The whole case turned by 37 degrees about a skew axis (points and
    velocities): every coefficient, GGI weight and address of the assembled
    system is unchanged to rounding, so geometry and the polygon GGI do not
    depend on orientation."""
    _lib_or_skip()
    case = make()
    ff.write_box_case(str(tmp_path / "a"), case)
    ff.write_box_case(str(tmp_path / "b"), case, rotate=((1.0, 2.0, 3.0), 37.0))
    a = mrf.Case.from_foam(str(tmp_path / "a"), div="upwind")
    b = mrf.Case.from_foam(str(tmp_path / "b"), div="upwind")
    for x, y in zip(a.regions(), b.regions()):
        for k in ("lAddr", "uAddr"):
            assert np.array_equal(x[k], y[k])
        scale = np.abs(x["diag"]).max()
        for k in ("diag", "lower", "upper", "source"):
            np.testing.assert_allclose(y[k], x[k], rtol=1e-10, atol=1e-12 * scale, err_msg=k)
    for x, y in zip(a.interfaces(), b.interfaces()):
        for k in ("fcA", "fcB", "offA", "offB", "idxA", "idxB"):
            assert np.array_equal(x[k], y[k]), k
        for k in ("wA", "wB", "coeffA", "coeffB"):
            np.testing.assert_allclose(y[k], x[k], rtol=1e-10, err_msg=k)


@pytest.mark.parametrize("rotate", [None, ((1.0, 2.0, 3.0), 37.0)], ids=["axis", "skew"])
def test_ggi_area_identity(tmp_path, rotate):
    """Patches that cover each other: the area each side's faces send to a
    face of the other side adds up to that face's area,
    sum_i |S_i| w_ij = |S_j|, and every face's weights sum to 1."""
    _lib_or_skip()
    case = _covering_pair()
    ff.write_box_case(str(tmp_path), case, rotate=rotate)
    I = mrf.Case.from_foam(str(tmp_path), div="upwind").interfaces()[0]
    area = {"A": 0.1 * 0.1, "B": (0.6 / 7) * (0.4 / 5)}   # fluid ymin faces, solid ymax faces
    for s, o in (("A", "B"), ("B", "A")):
        off, idx, w = I["off" + s], I["idx" + s], I["w" + s]
        np.testing.assert_allclose(np.add.reduceat(w, off[:-1]), 1.0, rtol=1e-13)
        sent = np.zeros(I["n" + o])
        np.add.at(sent, idx, area[s] * w)
        np.testing.assert_allclose(sent, area[o], rtol=1e-12)
    assert np.diff(I["offA"]).max() > 1 and np.diff(I["offB"]).max() > 1   # non-conformal


@pytest.mark.gpu
@pytest.mark.parametrize("make", [_covering_pair, _duct_small], ids=["covering", "duct_ggi"])
@pytest.mark.parametrize("rotate", [None, ((1.0, 2.0, 3.0), 37.0)], ids=["axis", "skew"])
def test_nonconformal_case_read_from_disk_gpu_matches_oracle(gpu, tmp_path, make, rotate):
    """A non-conformal case read from disk (polyMesh, fields, dictionaries;
    the GGI weights built from the polygon overlaps of the faces as read,
    axis-aligned and turned about a skew axis) solved on the GPU, and
    re-assembled there from its fields: both equal the oracle's solve of the
    read system (iterations, residual and field bit for bit)."""
    case = make()
    ff.write_box_case(str(tmp_path), case, rotate=rotate)
    c = gpu.Case.from_foam(str(tmp_path), div="upwind")
    A = gpu.CoupledFvMatrix(0)
    A.add_case(c)
    A.commit()
    perf = A.solve(tolerance=1e-13, maxIter=500)[0]
    O = ob.OracleSystem.from_arrays(*_arrays(c))
    psi_ref, pref = O.solve(tol=1e-13, maxIter=500)
    assert perf.converged
    assert perf.n_iterations == pref["nIterations"]
    assert perf.final_residual == pref["finalResidual"]
    assert np.array_equal(np.concatenate(A.psi(0)), np.concatenate(psi_ref))
    # the device assembly from the read fields gives the same system
    A.reset_psi()
    A.assemble()
    for r, R in enumerate(c.regions()):
        M = A.region_matrix(0, r)
        lower = R["upper"] if R["symmetric"] else R["lower"]
        for key, ref in (("diag", R["diag"]), ("upper", R["upper"]), ("lower", lower), ("source", R["source"])):
            assert np.array_equal(M[key], ref), (r, key)
    I = c.interfaces()[0]
    assert np.array_equal(A.interface_coeff(0, 0, 0), I["coeffA"])
    assert np.array_equal(A.interface_coeff(0, 0, 1), I["coeffB"])
    perf2 = A.solve(tolerance=1e-13, maxIter=500)[0]
    assert perf2.n_iterations == pref["nIterations"]
    assert np.array_equal(np.concatenate(A.psi(0)), np.concatenate(psi_ref))


def test_convection_scheme_is_read_not_assumed(tutorial_dir, tmp_path):
    """The tutorial convects T with `Gauss linearUpwind Gauss linear`
    (system/fluid/fvSchemes): its explicit correction is not assembled, so
    the reader refuses the case rather than assembling upwind silently; the
    caller may ask for upwind in its place (div="upwind"), and a case naming
    Gauss upwind, or no convection scheme, reads as it did."""
    _lib_or_skip()
    with pytest.raises(mrf.MrfgError, match="linearUpwind"):
        mrf.Case.from_foam(tutorial_dir)
    if os.path.isdir(REF_CASE):
        with pytest.raises(mrf.MrfgError, match="linearUpwind"):
            mrf.Case.from_foam(REF_CASE)
    ref = mrf.Case.from_foam(tutorial_dir, div="upwind").regions()
    d = tmp_path / "upwind"
    ff.write_fixture_case(str(d))
    fs = d / "system" / "fluid" / "fvSchemes"
    fs.write_text(fs.read_text().replace("Gauss linearUpwind Gauss linear", "Gauss upwind"))
    got = mrf.Case.from_foam(str(d)).regions()
    for a, b in zip(got, ref):
        for k in ("diag", "upper", "lower", "source"):
            assert np.array_equal(a[k], b[k]), k
    fs.write_text("ddtSchemes { default backward; }\n")
    got = mrf.Case.from_foam(str(d)).regions()
    assert np.array_equal(got[0]["upper"], ref[0]["upper"])
    fs.write_text("ddtSchemes { default backward; }\ndivSchemes { default Gauss linear; }\n")
    with pytest.raises(mrf.MrfgError, match="Gauss linear"):
        mrf.Case.from_foam(str(d))
    with pytest.raises(mrf.MrfgError, match="div"):
        mrf.Case.from_foam(str(d), div="linear")
