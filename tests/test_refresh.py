"""Value refresh on a committed structure (mrfg_update_region /
mrfg_update_interface): the drop-in time loop commits once per mesh and
hands over each step's re-assembled values.

Each refreshed solve is compared with the oracle's solve of the new system
from the same initial field: bit-identical iterations, residuals and field,
as for a system committed afresh.  The new systems are the same meshes with
other seeded conductivities, velocities and conductivity ratios, so the
addressing and GGI weights are those committed while every coefficient,
diagonal, source and interface coefficient changes.
"""
import numpy as np
import pytest

import multiregionfoam_amd as m
from multiregionfoam_amd import cases
import oracle_bind as ob

pytestmark = pytest.mark.gpu


def variants(kind):
    if kind == "duct_ggi":
        return [cases.duct3d(nf=(37, 23, 29), ns=(41, 27, 7), seed=s) for s in (11, 12, 13, 14)]
    if kind == "box_pair":
        return [cases.box_pair(6000, 4000, kr, 7, ggi=True) for kr in (3.0, 30.0, 300.0, 0.5)]
    if kind == "plate2d":
        out = []
        for k, u in ((0.6, 1.0), (1.2, 0.5), (0.3, 2.0), (0.6, 0.1)):
            c = cases.plate2d()
            c["regions"][0]["k"] = dict(uniform=k)
            c["regions"][0]["velocity"]["Umax"] = u
            out.append(c)
        return out
    raise KeyError(kind)


def _refresh(A, sys, case):
    C = m.Case(case)
    for r, R in enumerate(C.regions()):
        A.update_region(sys, r, R["diag"], R["upper"], R["source"],
                        lower=None if R["symmetric"] else R["lower"])
    for i, I in enumerate(C.interfaces()):
        A.update_interface(sys, i, I["coeffA"], I["coeffB"])
    return C


@pytest.mark.parametrize("kind", ["duct_ggi", "box_pair", "plate2d"])
def test_refresh_equals_oracle_of_new_system(gpu, kind):
    vs = variants(kind)
    A = gpu.CoupledFvMatrix(0)
    A.add_case(vs[0])
    A.commit()
    A.solve(tolerance=1e-10, maxIter=1000)
    for case in vs[1:]:
        C = _refresh(A, 0, case)
        psi0 = A.psi(0)
        O = ob.OracleSystem.from_arrays(C.regions(), C.interfaces())
        psi_ref, perf_ref = O.solve(tol=1e-10, maxIter=1000, psi0=psi0)
        perf = A.solve(tolerance=1e-10, maxIter=1000)[0]
        assert perf.converged and perf_ref["converged"]
        assert perf.n_iterations == perf_ref["nIterations"]
        assert perf.initial_residual == perf_ref["initialResidual"]
        assert perf.final_residual == perf_ref["finalResidual"]
        assert np.array_equal(np.concatenate(A.psi(0)), np.concatenate(psi_ref))
        # the device interface-flux check reads the refreshed coefficients
        for i in range(len(O.interfaces)):
            qa, qb = A.interface_flux(i)
            QA, QB, _, _ = O.interface_flux(i, A.psi(0))
            assert qa == QA and qb == QB


def test_refresh_kernels_and_batch_neighbours(gpu):
    """SpMV and the DILU factor follow the refreshed values; a refreshed
    system in a batch leaves the other systems' results unchanged; mixed
    precision re-derives its FP32 copies."""
    vs = variants("box_pair")
    other = cases.duct3d(nf=(14, 8, 6), ns=(18, 10, 3))
    B = gpu.CoupledFvMatrix(0)
    B.add_case(other)
    B.add_case(vs[0])
    B.commit()
    p_other = B.solve(tolerance=1e-10, maxIter=1000)[0]
    x_other = np.concatenate(B.psi(0))
    C = _refresh(B, 1, vs[2])
    O = ob.OracleSystem.from_arrays(C.regions(), C.interfaces())
    rng = np.random.default_rng(4)
    x = [rng.uniform(250, 400, R["nCells"]) for R in O.regions]
    assert np.array_equal(B.spmv(x, sys=1), np.concatenate(O.amul(x)))
    assert np.array_equal(B.dilu_factor(sys=1), np.concatenate(O.dilu_factor()))
    B.reset_psi()
    perfs = B.solve(tolerance=1e-10, maxIter=1000)
    assert perfs[0].n_iterations == p_other.n_iterations
    assert np.array_equal(np.concatenate(B.psi(0)), x_other)
    psi_ref, perf_ref = O.solve(tol=1e-10, maxIter=1000, psi0=[R["psi"] for R in O.regions])
    assert perfs[1].n_iterations == perf_ref["nIterations"]
    assert np.array_equal(np.concatenate(B.psi(1)), np.concatenate(psi_ref))
    B.reset_psi()
    pm = B.solve(tolerance=1e-10, maxIter=1000, precision="mixed")
    assert pm[1].converged
    err = np.linalg.norm(np.concatenate(B.psi(1)) - np.concatenate(psi_ref)) / np.linalg.norm(
        np.concatenate(psi_ref))
    assert err <= 1e-8, err


def test_refresh_rejects_mismatched_use(gpu):
    A = gpu.CoupledFvMatrix(0)
    A.add_case(cases.duct3d(nf=(12, 6, 6), ns=(15, 8, 3)))
    A.commit()
    C = m.Case(cases.duct3d(nf=(12, 6, 6), ns=(15, 8, 3)))
    R = C.region(1)                      # the solid: committed symmetric
    with pytest.raises(gpu.MrfgError, match="symmetric"):
        A.update_region(0, 1, R["diag"], R["upper"], R["source"], lower=R["upper"])
    with pytest.raises(ValueError):
        A.update_region(0, 1, R["diag"][:-1], R["upper"], R["source"])
