"""Pins the CPU oracle (oracle/mrf_oracle.c) against what the mathematics
fixes: closed forms, invariants, brute force and direct solves.

The reference's own tests only check that its tutorials run to completion
(flowOverHeatedPlate_testSuite/baseCase/test_template.py:6-8) and foam-extend,
which holds the solver arithmetic, is not part of the reference tree, so the
oracle cannot be compared with reference outputs.  Each check here is chosen
so that a plausible mistake in the restatement fails it: a dropped or
mis-signed term, the wrong interpolation or upwind side, a transposed
interface coefficient, a DILU recurrence off by one face, a wrong normFactor.
All run on the CPU in a few seconds.
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from multiregionfoam_amd import cases
import oracle_bind as ob


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def box(name, rtype, n, origin, length, k, patches, T0=300.0, **kw):
    return cases.region(name, rtype, n, origin, length, k=k, T0=T0, patches=patches, **kw)


def geometry(r):
    d = [r["length"][a] / r["nCells"][a] for a in range(3)]
    area = [d[1] * d[2], d[0] * d[2], d[0] * d[1]]
    return d, area


def faces_by_axis(R, n):
    """Axis of each face of an LDU box region (neighbour offsets 1, nx, nx*ny)."""
    diff = R["uAddr"] - R["lAddr"]
    assert n[0] > 1 and n[1] > 1
    return np.where(diff == 1, 0, np.where(diff == n[0], 1, 2))


# --------------------------------------------------------------------------
# a2: laplacian assembly (conductTemperature.C:135-152)
# --------------------------------------------------------------------------

def test_laplacian_face_coefficients_closed_form():
    n = (5, 4, 3)
    r = box("solid", "conductTemperature", n, (0, 0, 0), (0.5, 0.2, 0.3),
            dict(block=2, min=1.0, max=100.0, seed=7), {})
    O = ob.OracleSystem(cases.case([r], []))
    R = O.regions[0]
    d, area = geometry(r)
    axis = faces_by_axis(R, n)
    k = R["k"]
    kf = 0.5 * k[R["lAddr"]] + 0.5 * k[R["uAddr"]]          # linear interpolation
    a_f = kf * np.array(area)[axis] / np.array(d)[axis]      # k |S| / |d|
    assert R["symmetric"]
    assert np.allclose(R["upper"], -a_f, rtol=1e-14, atol=0)
    assert np.array_equal(R["lower"], R["upper"])
    # zeroGradient everywhere: every row sums to zero, no source
    A = O.to_scipy().toarray()
    assert np.allclose(A.sum(axis=1), 0.0, atol=1e-12 * np.abs(A).max())
    assert np.all(R["source"] == 0.0)
    assert len(np.unique(k)) > 1                            # the interpolation is exercised


def test_fixed_value_boundary_closed_form():
    n = (4, 3, 2)
    kval, Tb = 3.0, 410.0
    r = box("solid", "conductTemperature", n, (0, 0, 0), (0.4, 0.3, 0.2), dict(uniform=kval),
            dict(xmin=dict(type="fixedValue", value=Tb)))
    O = ob.OracleSystem(cases.case([r], []))
    R = O.regions[0]
    d, area = geometry(r)
    A = O.to_scipy().toarray()
    rowsum = A.sum(axis=1)
    ab = kval * area[0] / (0.5 * d[0])                     # 2 k |S| / dx
    i = np.arange(R["nCells"]) % n[0]
    assert np.allclose(rowsum[i == 0], ab, rtol=1e-13)
    assert np.allclose(rowsum[i > 0], 0.0, atol=1e-12 * ab)
    assert np.allclose(R["source"][i == 0], ab * Tb, rtol=1e-13)
    assert np.all(R["source"][i > 0] == 0.0)


# --------------------------------------------------------------------------
# a3: upwind convection with the negSumDiag diagonal (transportTemperature.C:129-150)
# --------------------------------------------------------------------------

def test_upwind_convection_closed_form():
    n = (6, 3, 2)
    U, rhoCp, kval = 0.4, 1000.0, 0.5
    r = box("fluid", "transportTemperature", n, (0, 0, 0), (0.6, 0.3, 0.2), dict(uniform=kval),
            dict(xmin=dict(type="fixedValue", value=300.0, flux="inlet"),
                 xmax=dict(type="zeroGradient", flux="outlet")),
            rhoCp=rhoCp, velocity=dict(axis=0, profile=(), Umax=U, noise=0.0, seed=0))
    O = ob.OracleSystem(cases.case([r], []))
    R = O.regions[0]
    d, area = geometry(r)
    axis = faces_by_axis(R, n)
    a_f = kval * np.array(area)[axis] / np.array(d)[axis]
    phi = np.where(axis == 0, U * area[0], 0.0)              # uniform flow along +x
    assert not R["symmetric"]
    # lower = -rhoCp max(phi, 0) - a ; upper = -rhoCp max(-phi, 0) - a
    assert np.allclose(R["lower"], -rhoCp * np.maximum(phi, 0) - a_f, rtol=1e-13)
    assert np.allclose(R["upper"], -rhoCp * np.maximum(-phi, 0) - a_f, rtol=1e-13)
    # negSumDiag: an interior cell's diagonal is its outflow plus its diffusion
    i, j, kk = np.arange(R["nCells"]) % n[0], (np.arange(R["nCells"]) // n[0]) % n[1], \
        np.arange(R["nCells"]) // (n[0] * n[1])
    interior = (i > 0) & (i < n[0] - 1) & (j > 0) & (j < n[1] - 1) & (kk > 0) & (kk < n[2] - 1)
    diff_sum = 2 * kval * (area[0] / d[0] + area[1] / d[1] + area[2] / d[2])
    if interior.any():
        assert np.allclose(R["diag"][interior], rhoCp * U * area[0] + diff_sum, rtol=1e-13)
    # the discrete operator conserves: a constant field is an exact steady
    # solution away from boundaries (zero row sum on interior cells)
    A = O.to_scipy().toarray()
    assert np.allclose(A.sum(axis=1)[interior], 0.0, atol=1e-10 * rhoCp * U * area[0])


# --------------------------------------------------------------------------
# a4-a7, a18: interface conductance (monolithicThermalDiffusivity...C:167-277)
# --------------------------------------------------------------------------

def two_slabs(nA, nB, kA, kB, LA, LB, TA, TB, nxy=(1, 1)):
    a = box("a", "conductTemperature", (nxy[0], nxy[1], nA), (0, 0, 0), (1.0, 1.0, LA), dict(uniform=kA),
            dict(zmin=dict(type="fixedValue", value=TA), zmax=dict(type="monolithicTemperature")))
    b = box("b", "conductTemperature", (nxy[0], nxy[1], nB), (0, 0, LA), (1.0, 1.0, LB), dict(uniform=kB),
            dict(zmin=dict(type="monolithicTemperature"), zmax=dict(type="fixedValue", value=TB)))
    return cases.case([a, b], [cases.interface("ab", ("a", "zmax"), ("b", "zmin"), method="directMap")])


def test_interface_coefficient_series_resistance():
    nA, nB, kA, kB, LA, LB = 4, 6, 2.5, 40.0, 0.2, 0.3
    O = ob.OracleSystem(two_slabs(nA, nB, kA, kB, LA, LB, 300.0, 400.0, nxy=(3, 2)))
    I = O.interfaces[0]
    dA, dB = LA / nA, LB / nB
    magSf = (1.0 / 3) * (1.0 / 2)
    expect = magSf / (0.5 * dA / kA + 0.5 * dB / kB)
    assert np.allclose(I["coeffA"], expect, rtol=1e-13)
    assert np.allclose(I["coeffB"], expect, rtol=1e-13)
    # conformal faces gather exactly one partner with weight 1
    assert np.array_equal(np.diff(I["offA"]), np.ones(I["nA"], np.int32))
    assert np.all(I["wA"] == 1.0) and np.all(I["wB"] == 1.0)
    # the coupled face cell's diagonal carries the coefficient (internalCoeffs)
    A = O.to_scipy().toarray()
    assert np.allclose(A.sum(axis=1)[O.offsets()[0] + I["fcA"]], 0.0, atol=1e-12 * expect)


def test_two_slab_conduction_exact_solution():
    """1-D steady conduction through two slabs: the FV solution is exact for
    the piecewise-linear profile (half-cell boundary and interface distances)."""
    nA, nB, kA, kB, LA, LB, TA, TB = 7, 5, 1.5, 30.0, 0.35, 0.25, 290.0, 370.0
    O = ob.OracleSystem(two_slabs(nA, nB, kA, kB, LA, LB, TA, TB))
    psi, perf = O.solve(tol=1e-15, maxIter=200)
    q = (TA - TB) / (LA / kA + LB / kB)
    zA = (np.arange(nA) + 0.5) * LA / nA
    zB = (np.arange(nB) + 0.5) * LB / nB
    exactA = TA - q * zA / kA
    exactB = TA - q * LA / kA - q * zB / kB
    assert np.allclose(psi[0], exactA, rtol=1e-12)
    assert np.allclose(psi[1], exactB, rtol=1e-12)


# --------------------------------------------------------------------------
# a17: GGI weights (ggiInterfaceToInterfaceMapping.C:62-77, rescaled)
# --------------------------------------------------------------------------

def overlap(a0, a1, b0, b1):
    return max(0.0, min(a1, b1) - max(a0, b0))


@pytest.mark.parametrize("offset", [0.37, 0.0])
def test_ggi_weights_exact_overlaps(offset):
    c = cases.duct3d(nf=(4, 3, 2), ns=(5, 4, 2), offset=offset)
    O = ob.OracleSystem(c)
    I = O.interfaces[0]
    f_, s_ = c["regions"]
    # faces on z planes: index = ix + nx*iy for both sides
    def faces(r):
        nx, ny = r["nCells"][0], r["nCells"][1]
        dx, dy = r["length"][0] / nx, r["length"][1] / ny
        x0, y0 = r["origin"][0], r["origin"][1]
        return [(x0 + ix * dx, x0 + (ix + 1) * dx, y0 + iy * dy, y0 + (iy + 1) * dy)
                for iy in range(ny) for ix in range(nx)]
    FA, FB = faces(f_), faces(s_)
    for side, (Fo, Fn, off, idx, w) in enumerate(((FA, FB, I["offA"], I["idxA"], I["wA"]),
                                                  (FB, FA, I["offB"], I["idxB"], I["wB"]))):
        for i, fo in enumerate(Fo):
            ov = {j: overlap(fo[0], fo[1], fn[0], fn[1]) * overlap(fo[2], fo[3], fn[2], fn[3])
                  for j, fn in enumerate(Fn)}
            ov = {j: a for j, a in ov.items() if a > 0}
            got = dict(zip(idx[off[i]:off[i + 1]].tolist(), w[off[i]:off[i + 1]].tolist()))
            if not ov:
                assert not got
                continue
            tot = sum(ov.values())
            assert set(got) == set(ov), (side, i)
            for j in ov:
                assert got[j] == pytest.approx(ov[j] / tot, rel=1e-12)
            assert sum(got.values()) == pytest.approx(1.0, rel=1e-14)


# --------------------------------------------------------------------------
# a8, a9, a15: A x with the interface gather (lduMatrix::Amul + updateInterfaceMatrix)
# --------------------------------------------------------------------------

SMALL = {
    "plate": cases.plate2d(nx=20, nyf=6, nys=4),
    "duct_ggi": cases.duct3d(nf=(8, 5, 4), ns=(10, 6, 3)),
    "exchanger3": cases.exchanger3(n=(6, 4, 3)),
}


@pytest.mark.parametrize("name", list(SMALL))
def test_amul_matches_assembled_matrix(name):
    O = ob.OracleSystem(SMALL[name])
    rng = np.random.default_rng(4)
    x = [rng.uniform(250, 400, r["nCells"]) for r in O.regions]
    y = np.concatenate(O.amul(x))
    y_ref = O.to_scipy() @ np.concatenate(x)
    assert rel(y, y_ref) < 1e-14


# --------------------------------------------------------------------------
# a14: Cholesky / DILU (D-ILU, Barrett et al., Templates 1994, 3.4.1)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("name", list(SMALL))
def test_dilu_dense_identities(name):
    O = ob.OracleSystem(SMALL[name])
    rD = O.dilu_factor()
    rng = np.random.default_rng(5)
    r = [rng.standard_normal(R["nCells"]) for R in O.regions]
    w = O.dilu_precondition(rD, r)
    for g, R in enumerate(O.regions):
        n = R["nCells"]
        A = np.zeros((n, n))
        A[np.arange(n), np.arange(n)] = R["diag"]
        A[R["uAddr"], R["lAddr"]] = R["lower"] if not R["symmetric"] else R["upper"]
        A[R["lAddr"], R["uAddr"]] = R["upper"]
        D = np.diag(1.0 / rD[g])
        L, U = np.tril(A, -1), np.triu(A, 1)
        M = (D + L) @ np.diag(rD[g]) @ (D + U)
        # the DILU diagonal makes the preconditioner agree with A on the diagonal
        assert np.allclose(np.diag(M), np.diag(A), rtol=1e-12)
        # and the sweeps apply M^-1
        assert rel(M @ w[g], r[g]) < 1e-12


# --------------------------------------------------------------------------
# a16: sumA and normFactor (lduMatrix::sumA, lduMatrix::solver::normFactor)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("name", list(SMALL))
def test_sumA_and_norm_factor(name):
    O = ob.OracleSystem(SMALL[name])
    A = O.to_scipy()
    sA = np.concatenate(O.sumA())
    assert np.allclose(sA, np.asarray(A.sum(axis=1)).ravel(), rtol=1e-13, atol=1e-12 * abs(A).max())
    # for a constant field A x = sumA * xRef, so normFactor = sum|b - sumA xRef| + small
    c = 317.25
    x = [np.full(R["nCells"], c) for R in O.regions]
    b = [R["source"] for R in O.regions]
    nf = O.norm_factor(x, b, O.amul(x))
    expect = np.abs(np.concatenate(b) - sA * c).sum() + 1e-20
    assert nf == pytest.approx(expect, rel=1e-11)


# --------------------------------------------------------------------------
# a13: BiCGStab / PCG converge to the direct solution
# --------------------------------------------------------------------------

@pytest.mark.parametrize("name", list(SMALL))
def test_bicgstab_matches_direct_solve(name):
    O = ob.OracleSystem(SMALL[name])
    x_ref = spla.spsolve(O.to_scipy().tocsc(), O.source_vector())
    psi, perf = O.solve(tol=1e-14, maxIter=5000)
    assert perf["converged"]
    assert rel(np.concatenate(psi), x_ref) < 1e-11


def test_pcg_matches_direct_solve():
    c = two_slabs(6, 4, 2.0, 9.0, 0.3, 0.2, 300.0, 350.0, nxy=(4, 3))
    O = ob.OracleSystem(c)
    x_ref = spla.spsolve(O.to_scipy().tocsc(), O.source_vector())
    psi, perf = O.solve(solver="PCG", tol=1e-14, maxIter=2000)
    assert perf["converged"]
    assert rel(np.concatenate(psi), x_ref) < 1e-11


def test_stop_test_and_residual_definition():
    """initialResidual = sum|b - A x0| / normFactor; nIterations counts the
    Krylov loop; maxIter bounds it (lduSolver::stop)."""
    O = ob.OracleSystem(SMALL["duct_ggi"])
    x0 = [R["psi"] for R in O.regions]
    b = [R["source"] for R in O.regions]
    Ax = O.amul(x0)
    nf = O.norm_factor(x0, b, Ax)
    r0 = np.abs(np.concatenate(b) - np.concatenate(Ax)).sum()
    _, perf = O.solve(tol=1e-10, maxIter=3)
    assert perf["initialResidual"] == pytest.approx(r0 / nf, rel=1e-12)
    assert perf["nIterations"] == 3 and not perf["converged"]
    _, perf = O.solve(tol=1e3, maxIter=100)              # already converged: no iteration
    assert perf["nIterations"] == 0 and perf["converged"]


# --------------------------------------------------------------------------
# interface conservation and the monolithicTemperature source
# --------------------------------------------------------------------------

def test_interface_flux_conservation_and_zero_source():
    c = cases.plate2d(nx=30, nyf=8, nys=5)
    O = ob.OracleSystem(c)
    psi, _ = O.solve(tol=1e-14, maxIter=2000)
    QA, QB, sA, sB = O.interface_flux(0, psi)
    assert abs(QA + QB) <= 1e-9 * abs(QA)
    assert abs(sA) <= 1e-9 * abs(QA) and abs(sB) <= 1e-9 * abs(QA)


# --------------------------------------------------------------------------
# the native case builder (the product's assembly) equals the oracle
# --------------------------------------------------------------------------

@pytest.mark.parametrize("name", list(SMALL) + ["duct_full_geometry"])
def test_native_builder_matches_oracle(name):
    import multiregionfoam_amd as mrf
    c = SMALL.get(name) or cases.duct3d(nf=(16, 8, 8), ns=(20, 10, 2))
    O = ob.OracleSystem(c)
    N = mrf.Case(c)
    for R, Rn in zip(O.regions, N.regions()):
        for key in ("lAddr", "uAddr", "diag", "upper", "source", "psi"):
            assert np.array_equal(R[key], Rn[key]), key
        if not R["symmetric"]:
            assert np.array_equal(R["lower"], Rn["lower"])
    for I, In in zip(O.interfaces, N.interfaces()):
        for key in ("fcA", "coeffA", "offA", "idxA", "wA", "fcB", "coeffB", "offB", "idxB", "wB"):
            assert np.array_equal(I[key], In[key]), key


def test_golden_direct_solves():
    """Committed fixtures: direct solves of the assembled plate and GGI duct
    miniature (tests/golden/make_golden.py); the oracle's converged fields
    match them."""
    import json
    import os
    here = os.path.join(os.path.dirname(__file__), "golden")
    meta = json.load(open(os.path.join(here, "golden.json")))
    for entry in meta["fixtures"]:
        c = getattr(cases, entry["builder"])(**entry["kwargs"])
        assert cases.to_dict_text(c) == open(os.path.join(here, entry["case_file"])).read()
        ref = np.load(os.path.join(here, entry["field_file"]))["psi"]
        O = ob.OracleSystem(c)
        psi, perf = O.solve(tol=1e-14, maxIter=5000)
        assert perf["converged"]
        assert rel(np.concatenate(psi), ref) < 1e-10, entry["name"]


@pytest.mark.parametrize("name", ["plate2d", "duct_ggi"])
def test_multithreaded_solve_matches_sequential(name):
    """orc_solve_mt (all host cores; level-scheduled DILU, row-parallel A*x)
    restates orc_solve: the same iterations (+-1, from the last bit of the
    merged sums) and the same field."""
    case = {"plate2d": cases.plate2d(), "duct_ggi": cases.duct3d(nf=(24, 12, 12), ns=(30, 15, 5))}[name]
    O = ob.OracleSystem(case)
    psi, perf = O.solve(tol=1e-12, maxIter=2000)
    psi_mt, perf_mt, threads = O.solve_mt(tol=1e-12, maxIter=2000)
    assert threads >= 1 and perf_mt["converged"]
    assert abs(perf_mt["nIterations"] - perf["nIterations"]) <= 1
    a, b = np.concatenate(psi_mt), np.concatenate(psi)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-12
