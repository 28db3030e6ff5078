"""Counterpart of the reference's flowOverHeatedPlate test suite
(tutorials/conjugateHeatTransfer/flowOverHeatedPlate_testSuite): the nine
parameter cases of genCases.py, monolithic coupling, on the tutorial's mesh.

The reference runs each case through the full transient solver (flow and
energy) and asserts only that it completes (baseCase/test_template.py:6-8);
its getResults.py then plots theta = (T - 300) / (310 - 300) sampled along
the fluid side of the plate (sampleDict: start (0.01 0 0), end (0.99 0 0),
20 points) against the reference curves of Vynnycky et al. (1998) in
references/*.csv.  Here the flow is the case's own fixed uniform U (the flow
solve is outside this library) and the coupled energy system is solved to
steady state:
* every case converges, on the CPU oracle and, as one batch of nine systems,
  on the GPU, where the result equals the oracle's bit for bit;
* theta stays within [0, 1] (a maximum principle of the upwind,
  M-matrix discretisation) and orders with the conductivity ratio k: a
  poorer fluid conductor (larger k) leaves the interface hotter;
* at Pr = 0.01 the temperature field is set by conduction in a thermal layer
  much thicker than the velocity boundary layer, so the slug (uniform-flow)
  model is the one the analytical curves of the paper assume and theta
  matches them within 0.03 (measured: 0.005 - 0.019), and the paper's
  numerical curves at Re = 500 within 0.07 (measured: 0.04 - 0.06).  At
  Pr = 100 the thermal layer sits inside the velocity boundary layer, which
  a uniform flow lacks; those curves are not compared.
The reference curves are committed as data under tests/golden/vynnycky
(copied by tests/golden/make_foam_fixture.py).
"""
import os

import numpy as np
import pytest

import multiregionfoam_amd as mrf
import foam_fixture as ff
import oracle_bind as ob

CURVES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vynnycky")

# genCases.py:9-13, caseStructure[1]
CASES = [(0.01, 500, 1), (0.01, 500, 5), (0.01, 500, 20),
         (0.01, 10000, 1), (0.01, 10000, 5), (0.01, 10000, 20),
         (100, 500, 1), (100, 500, 5), (100, 500, 20)]

# the sampled line, system/fluid/sampleDict of the test suite's baseCase
SAMPLE_X = np.linspace(0.01, 0.99, 20)

SOLVE = dict(tol=1e-12, relTol=0.0, minIter=0, maxIter=1000)
FV_SOLUTION = ("solvers { Tcoupled { solver BiCGStab; preconditioner DILU; tolerance 1e-12; relTol 0; "
               "minIter 0; maxIter 1000; } }")


def suite_params(Pr, Re, k):
    """genCases.py:43-60 update_params: L = rho_f = U_inf = 1, k_s = 100,
    mu = rho_f U_inf L / Re, k_f = k_s / k, c_p = k_f Pr / mu."""
    L, rhof, ks, Uinf = 1.0, 1.0, 100.0, 1.0
    mu = rhof * Uinf * L / int(Re)
    kf = ks / k
    return dict(kf=kf, cp=kf * Pr / mu)


def _tag(Pr):
    return "0.01" if Pr == 0.01 else "100"


def _curve(Pr, Re, k, kind):
    f = os.path.join(CURVES, "Pr%sRe%sk%s%s.csv" % (_tag(Pr), Re, k, kind))
    return np.loadtxt(f, delimiter=",") if os.path.exists(f) else None


def _read(tmp_path_factory, Pr, Re, k):
    d = tmp_path_factory.mktemp("Pr%sRe%sk%s" % (_tag(Pr), Re, k))
    ff.write_fixture_case(str(d), params=suite_params(Pr, Re, k))
    return mrf.Case.from_foam(str(d), ddt="steadyState", div="upwind")


def _sample_cells():
    """Fluid cells on the plate (y = 0) whose x-extent holds each sample
    point, from the points of the hex cells alone."""
    _, meshes = ff.load_fixture()
    m = meshes["fluid"]
    P, sizes, verts = m["points"], m["face_sizes"], m["face_verts"]
    fid = np.repeat(np.arange(len(sizes)), sizes)
    flo = np.full((len(sizes), 3), np.inf)
    fhi = np.full((len(sizes), 3), -np.inf)
    np.minimum.at(flo, fid, P[verts])
    np.maximum.at(fhi, fid, P[verts])
    n = int(max(m["owner"].max(), m["neighbour"].max())) + 1
    clo = np.full((n, 3), np.inf)
    chi = np.full((n, 3), -np.inf)
    for cells, faces in ((m["owner"], np.arange(len(m["owner"]))),
                         (m["neighbour"], np.arange(len(m["neighbour"])))):
        np.minimum.at(clo, cells, flo[faces])
        np.maximum.at(chi, cells, fhi[faces])
    bottom = np.where(np.abs(clo[:, 1]) < 1e-12)[0]
    out = []
    for x in SAMPLE_X:
        hit = bottom[(clo[bottom, 0] <= x) & (chi[bottom, 0] > x)]
        assert len(hit) == 1
        out.append(int(hit[0]))
    return np.array(out)


def _theta(psi_fluid, cells):
    return (psi_fluid[cells] - 300.0) / (310.0 - 300.0)


@pytest.fixture(scope="module")
def suite(tmp_path_factory):
    """The nine cases read from disk and solved by the oracle."""
    out = {}
    for c in CASES:
        case = _read(tmp_path_factory, *c)
        O = ob.OracleSystem.from_arrays(case.regions(), case.interfaces())
        psi, perf = O.solve(**SOLVE)
        out[c] = (case, psi, perf)
    return out


@pytest.fixture(scope="module")
def cells():
    return _sample_cells()


def test_parameter_map():
    """genCases.py's formulas, on two of its cases."""
    p = suite_params(0.01, 500, 1)
    assert p["kf"] == 100.0 and p["cp"] == pytest.approx(500.0, rel=1e-15)
    p = suite_params(100, 500, 20)
    assert p["kf"] == 5.0 and p["cp"] == pytest.approx(5.0 * 100 * 500, rel=1e-15)


def test_parameters_enter_the_fluid_matrix_linearly(tmp_path_factory):
    """The fluid's face coefficients and source are k_f (Laplacian) plus
    rho c_p (upwind convection) terms, so they add up over parameter sets;
    the solid's do not move with the fluid's parameters."""
    d = {}
    for kf, cp in ((2.0, 3.0), (5.0, 7.0), (7.0, 10.0)):
        p = tmp_path_factory.mktemp("lin")
        ff.write_fixture_case(str(p), params=dict(kf=kf, cp=cp))
        d[kf] = mrf.Case.from_foam(str(p), ddt="steadyState", div="upwind").regions()
    a, b, s = d[2.0], d[5.0], d[7.0]
    for key in ("lower", "upper", "source"):
        np.testing.assert_allclose(a[0][key] + b[0][key], s[0][key], rtol=1e-12, atol=1e-10)
    for key in ("lower", "upper"):
        assert np.array_equal(a[1][key], s[1][key])
    assert not np.allclose(a[0]["upper"], s[0]["upper"])


@pytest.mark.parametrize("Pr,Re,k", CASES)
def test_case_completes(suite, cells, Pr, Re, k):
    """Each case converges, and theta on the plate is within [0, 1]."""
    case, psi, perf = suite[(Pr, Re, k)]
    assert perf["converged"] and perf["finalResidual"] < 1e-12
    th = _theta(psi[0], cells)
    assert np.all(th >= 0.0) and np.all(th <= 1.0)
    assert np.all(np.isfinite(np.concatenate(psi)))


@pytest.mark.parametrize("Pr,Re", [(0.01, 500), (0.01, 10000), (100, 500)])
def test_theta_orders_with_conductivity_ratio(suite, cells, Pr, Re):
    t = [_theta(suite[(Pr, Re, k)][1][0], cells) for k in (1, 5, 20)]
    assert np.all(t[0] < t[1]) and np.all(t[1] < t[2])


def _compared_curves(Pr, Re, k):
    """(kind, tolerance) pairs held for a case: the analytical curves at
    Pr = 0.01, the numerical ones at Pr = 0.01, Re = 500."""
    if Pr != 0.01:
        return []
    out = [("analytical", 0.03)] if _curve(Pr, Re, k, "analytical") is not None else []
    if Re == 500 and _curve(Pr, Re, k, "numerical") is not None:
        out.append(("numerical", 0.07))
    return out


@pytest.mark.parametrize("Pr,Re,k", [c for c in CASES if _compared_curves(*c)])
def test_theta_matches_vynnycky_low_prandtl(suite, cells, Pr, Re, k):
    th = _theta(suite[(Pr, Re, k)][1][0], cells)
    for kind, tol in _compared_curves(Pr, Re, k):
        ref = _curve(Pr, Re, k, kind)
        inside = (SAMPLE_X >= ref[:, 0].min()) & (SAMPLE_X <= ref[:, 0].max())
        assert inside.sum() >= 15
        err = np.abs(th - np.interp(SAMPLE_X, ref[:, 0], ref[:, 1]))[inside]
        assert err.max() < tol, (kind, err.max())


def test_vynnycky_curves_present():
    """The committed curves: the 15 files of the reference's references/."""
    files = sorted(f for f in os.listdir(CURVES) if f.endswith(".csv"))
    assert len(files) == 15
    for f in files:
        d = np.loadtxt(os.path.join(CURVES, f), delimiter=",")
        assert d.ndim == 2 and d.shape[1] == 2 and len(d) >= 10
        assert np.all((d[:, 1] > -0.05) & (d[:, 1] < 1.05))


@pytest.mark.gpu
def test_suite_batch_gpu_matches_oracle(gpu, suite):
    """All nine cases as one batch on the GPU: per system, the iteration
    count, final residual and field equal the oracle's."""
    A = gpu.CoupledFvMatrix(0)
    for c in CASES:
        A.add_case(suite[c][0])
    A.commit()
    perfs = A.solve(fv_solution=FV_SOLUTION, entry="Tcoupled")
    for s, c in enumerate(CASES):
        _, psi, perf = suite[c]
        assert perfs[s].n_iterations == perf["nIterations"], c
        assert perfs[s].final_residual == perf["finalResidual"], c
        assert np.array_equal(np.concatenate(A.psi(s)), np.concatenate(psi)), c
    A.close()


@pytest.mark.gpu
def test_suite_assembled_on_device_matches_oracle(gpu, suite, tmp_path_factory):
    """The nine cases as nine copies of the tutorial's mesh whose fluid k
    and rho*c_p are set as fields and assembled on the device (no matrix
    upload): per system, the same iterations, residual and field as the
    oracle's solve of the case read from disk."""
    d = tmp_path_factory.mktemp("base")
    ff.write_fixture_case(str(d))
    base = gpu.Case.from_foam(str(d), ddt="steadyState", div="upwind")
    n = base.regions()[0]["nCells"]
    A = gpu.CoupledFvMatrix(0)
    for _ in CASES:
        A.add_case(base)
    A.commit()
    for s, c in enumerate(CASES):
        p = suite_params(*c)
        A.set_field(s, 0, "k", np.full(n, p["kf"]))
        A.set_field(s, 0, "rhoCp", p["cp"])
    A.assemble()
    perfs = A.solve(fv_solution=FV_SOLUTION, entry="Tcoupled")
    for s, c in enumerate(CASES):
        _, psi, perf = suite[c]
        assert perfs[s].n_iterations == perf["nIterations"], c
        assert perfs[s].final_residual == perf["finalResidual"], c
        assert np.array_equal(np.concatenate(A.psi(s)), np.concatenate(psi)), c
    A.close()
