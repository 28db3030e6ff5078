"""Finite-volume assembly on the device (mrfg_set_field, mrfg_assemble): the
region equations of setCoupledEqns (transportTemperature.C:129-150), the
monolithic interface coefficients (calcThermalDiffusivity) and the ddt
terms, assembled in HBM from the fields k, U, T and rho*cp and put in
place of the committed matrices.

Parity: the device-assembled region matrices and interface coefficients
equal, bit for bit, the host reader's assembly of a case written with the
same fields (the reader itself is pinned against closed forms in
test_foam_case.py), and every solve or time-step run after a device
assembly equals the oracle's run on the host-read case (iterations,
residuals and fields).  The mixed-precision path, which derives its FP32
copies from the device-assembled values, equals the mixed solve of the
host-read case.
"""
import numpy as np
import pytest

from multiregionfoam_amd import cases
import foam_fixture as ff
import oracle_bind as ob

pytestmark = pytest.mark.gpu

SOLVE = dict(tol=1e-12, relTol=0.0, minIter=0, maxIter=1000)
FV_SOLUTION = ("solvers { Tcoupled { solver BiCGStab; preconditioner DILU; tolerance 1e-12; relTol 0; "
               "minIter 0; maxIter 1000; } }")


def _read(gpu, tmp_path_factory, params=None, ddt="steadyState"):
    d = tmp_path_factory.mktemp("case")
    ff.write_fixture_case(str(d), params=params, ddt=None if ddt == "steadyState" else ddt)
    return gpu.Case.from_foam(str(d), ddt=ddt, div="upwind")


def _assert_matrices_equal(A, sys, case):
    for r, R in enumerate(case.regions()):
        M = A.region_matrix(sys, r)
        lower = R["upper"] if R["symmetric"] else R["lower"]
        for key, ref in (("diag", R["diag"]), ("upper", R["upper"]), ("lower", lower), ("source", R["source"])):
            assert np.array_equal(M[key], ref), (r, key, np.abs(M[key] - ref).max())
    for i, I in enumerate(case.interfaces()):
        assert np.array_equal(A.interface_coeff(sys, i, 0), I["coeffA"]), i
        assert np.array_equal(A.interface_coeff(sys, i, 1), I["coeffB"]), i


def _assert_solve_matches_oracle(A, sys, perf, case):
    O = ob.OracleSystem.from_arrays(case.regions(), case.interfaces())
    psi, ref = O.solve(**SOLVE)
    assert perf.n_iterations == ref["nIterations"]
    assert perf.final_residual == ref["finalResidual"]
    assert np.array_equal(np.concatenate(A.psi(sys)), np.concatenate(psi))


def test_assembly_reproduces_the_reader(gpu, tmp_path_factory):
    """Unchanged fields: the device assembly equals the reader's, and the
    solve after it equals the oracle's."""
    c = _read(gpu, tmp_path_factory)
    A = gpu.CoupledFvMatrix(0)
    A.add_case(c)
    A.commit()
    A.assemble()
    _assert_matrices_equal(A, 0, c)
    perf = A.solve(fv_solution=FV_SOLUTION, entry="Tcoupled")[0]
    _assert_solve_matches_oracle(A, 0, perf, c)
    A.close()


def test_conductivity_and_heat_capacity_fields(gpu, tmp_path_factory):
    """New k and rho*cp of the fluid, set as fields and assembled on the
    device, equal the case written with them; two assemblies in a row, the
    second back to other values, leave nothing of the first behind."""
    base = _read(gpu, tmp_path_factory)
    A = gpu.CoupledFvMatrix(0)
    A.add_case(base)
    A.commit()
    nf = base.regions()[0]["nCells"]
    for kf, cp in ((20.0, 1000.0), (1.25, 7.5)):
        A.set_field(0, 0, "k", np.full(nf, kf))
        A.set_field(0, 0, "rhoCp", cp)
        A.assemble()
        ref = _read(gpu, tmp_path_factory, params=dict(kf=kf, cp=cp))
        _assert_matrices_equal(A, 0, ref)
        A.reset_psi()
        perf = A.solve(fv_solution=FV_SOLUTION, entry="Tcoupled")[0]
        _assert_solve_matches_oracle(A, 0, perf, ref)
    A.close()


def test_velocity_and_temperature_fields_time_steps(gpu, tmp_path_factory):
    """A transient (Euler) case: a new fluid velocity (phi formed on the
    device) and new initial temperatures; the assembled system, its ddt
    terms and three backward time steps equal the host-read case's."""
    params = dict(U=(2.0, 0.0, 0.0), T=dict(fluid=305.0, solid=308.0))
    base = _read(gpu, tmp_path_factory, ddt="Euler")
    A = gpu.CoupledFvMatrix(0)
    A.add_case(base)
    A.commit()
    nf, ns = base.regions()[0]["nCells"], base.regions()[1]["nCells"]
    A.set_field(0, 0, "U", np.tile([2.0, 0.0, 0.0], (nf, 1)))
    A.set_field(0, 0, "T", np.full(nf, 305.0))
    A.set_field(0, 1, "T", np.full(ns, 308.0))
    A.assemble()
    ref = _read(gpu, tmp_path_factory, params=params, ddt="Euler")
    _assert_matrices_equal(A, 0, ref)
    steps = A.time_steps(3, ddt="backward", fv_solution=FV_SOLUTION, entry="Tcoupled")
    O = ob.OracleSystem.from_arrays(ref.regions(), ref.interfaces())
    ddt = [({k: R[k] for k in ("ddtCoeff", "diagBase", "sourceBase")} if "ddtCoeff" in R else None)
           for R in ref.regions()]
    runs = O.time_steps(ddt, 3, scheme="backward", **SOLVE)
    for n, (perf, (psi, p)) in enumerate(zip(steps, runs)):
        assert perf[0].n_iterations == p["nIterations"], n
        assert perf[0].final_residual == p["finalResidual"], n
    assert np.array_equal(np.concatenate(A.psi(0)), np.concatenate(runs[-1][0]))
    A.close()


def test_mixed_precision_and_flux_after_device_assembly(gpu, tmp_path_factory):
    """The FP32 copies and the interface-flux check follow the device
    assembly: both equal what the host-read case gives."""
    kf, cp = 20.0, 1000.0
    base = _read(gpu, tmp_path_factory)
    ref = _read(gpu, tmp_path_factory, params=dict(kf=kf, cp=cp))
    A = gpu.CoupledFvMatrix(0)
    A.add_case(base)
    A.commit()
    A.set_field(0, 0, "k", np.full(base.regions()[0]["nCells"], kf))
    A.set_field(0, 0, "rhoCp", cp)
    A.assemble()
    B = gpu.CoupledFvMatrix(0)
    B.add_case(ref)
    B.commit()
    mixed = ("solvers { Tcoupled { solver BiCGStab; preconditioner DILU; tolerance 1e-12; relTol 0; "
             "maxIter 1000; precision mixed; } }")
    pa = A.solve(fv_solution=mixed, entry="Tcoupled")[0]
    pb = B.solve(fv_solution=mixed, entry="Tcoupled")[0]
    assert (pa.n_iterations, pa.n_outer, pa.final_residual) == (pb.n_iterations, pb.n_outer, pb.final_residual)
    assert np.array_equal(np.concatenate(A.psi(0)), np.concatenate(B.psi(0)))
    assert A.interface_flux(0) == B.interface_flux(0)
    A.close()
    B.close()


def test_errors(gpu, tmp_path_factory):
    c = _read(gpu, tmp_path_factory)
    A = gpu.CoupledFvMatrix(0)
    A.add_case(c)
    A.add_case(gpu.Case(cases.plate2d(nx=8, nyf=4, nys=2)))
    A.commit()
    n = c.regions()[0]["nCells"]
    with pytest.raises(gpu.MrfgError, match="values"):
        A.set_field(0, 0, "k", np.ones(n - 1))
    with pytest.raises(gpu.MrfgError, match="unknown field"):
        A.set_field(0, 0, "mu", np.ones(n))
    with pytest.raises(gpu.MrfgError, match="velocity"):
        A.set_field(0, 1, "U", np.ones(3 * c.regions()[1]["nCells"]))
    with pytest.raises(gpu.MrfgError, match="positive"):
        A.set_field(0, 0, "k", np.zeros(n))
    with pytest.raises(gpu.MrfgError, match="finite-volume model"):
        A.set_field(1, 0, "k", np.ones(4))
    with pytest.raises(gpu.MrfgError, match="device-assembled"):
        A.region_matrix(0, 0)
    A.close()
