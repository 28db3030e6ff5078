"""Coupled patch values after a solve (mrfg_interface_patch_values): the
patch neighbour field of the monolithic coupling
(monolithicCouplingFvPatchField::patchNeighbourField, the GGI interpolate of
the other side's face-cell values) and the monolithicTemperature patch value
(calcTemperature, radiation off):

    Tw = (kOwn To + kNei Tn) / (kOwn + kNei),
    kOwn = forig / (1 - w) / |delta|,  kNei = interpolate(forig_nei) / w / |delta|.

Pinned by what the mathematics fixes: in 1-D steady conduction through two
slabs the finite-volume solution is the exact piecewise-linear profile, so
Tw must be the exact interface temperature on both sides and Tn the other
side's adjacent cell value; a uniform field gives Tn = Tw = that value on
a GGI interface (the rescaled weights sum to one); the builder's wall
conductances are the closed forms k/d of each side.
"""
import os

import numpy as np
import pytest

from multiregionfoam_amd import cases
import multiregionfoam_amd as m
import oracle_bind as ob
from test_oracle import two_slabs


# --------------------------------------------------------------------------
# oracle and host builder (CPU)
# --------------------------------------------------------------------------

def test_oracle_two_slab_interface_temperature_exact():
    nA, nB, kA, kB, LA, LB, TA, TB = 7, 5, 1.5, 30.0, 0.35, 0.25, 290.0, 370.0
    O = ob.OracleSystem(two_slabs(nA, nB, kA, kB, LA, LB, TA, TB, nxy=(3, 2)))
    psi, _ = O.solve(tol=1e-15, maxIter=200)
    q = (TA - TB) / (LA / kA + LB / kB)
    T_int = TA - q * LA / kA
    for side in (0, 1):
        tn, tw = O.interface_patch_values(0, side, psi)
        assert np.allclose(tw, T_int, rtol=1e-12, atol=0)
        # the neighbour field is the other slab's adjacent cell value
        other = psi[1 - side]
        adj = other[O.interfaces[0]["fcB" if side == 0 else "fcA"]]
        assert np.allclose(tn, adj, rtol=1e-14, atol=0)


def test_oracle_uniform_field_on_ggi():
    c = cases.duct3d(nf=(12, 6, 6), ns=(15, 8, 3))
    O = ob.OracleSystem(c)
    psi = [np.full(R["nCells"], 333.25) for R in O.regions]
    for side in (0, 1):
        tn, tw = O.interface_patch_values(0, side, psi)
        # faces without partners (uncovered edge) have tn = 0; the others
        # gather a convex combination
        I = O.interfaces[0]
        covered = np.diff(I["offB" if side else "offA"]) > 0
        assert np.allclose(tn[covered], 333.25, rtol=1e-14)
        assert np.allclose(tw[covered], 333.25, rtol=1e-14)


def test_builder_wall_conductances_closed_form():
    """The native builder's kOwn / kNei are k_own / d_own and k_nei / d_nei
    (d the half cell sizes normal to the interface), for a GGI interface
    with a uniform conductivity on each side."""
    c = cases.box_pair(6000, 4000, 25.0, 3, ggi=True)
    C = m.Case(c)
    I = C.interface(0)
    fl, so = c["regions"]
    dz_f = fl["length"][2] / fl["nCells"][2]
    dz_s = so["length"][2] / so["nCells"][2]
    kf, ks = fl["k"]["uniform"], so["k"]["uniform"]
    assert np.allclose(I["kOwnA"], kf / (0.5 * dz_f), rtol=1e-13)
    assert np.allclose(I["kOwnB"], ks / (0.5 * dz_s), rtol=1e-13)
    covA = np.diff(I["offA"]) > 0
    covB = np.diff(I["offB"]) > 0
    assert np.allclose(I["kNeiA"][covA], ks / (0.5 * dz_s), rtol=1e-13)
    assert np.allclose(I["kNeiB"][covB], kf / (0.5 * dz_f), rtol=1e-13)


# --------------------------------------------------------------------------
# GPU
# --------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("name", ["plate2d", "duct_ggi", "exchanger3", "two_slabs"])
def test_gpu_patch_values_match_oracle(gpu, name):
    case = {"plate2d": cases.plate2d(), "duct_ggi": cases.duct3d(nf=(24, 12, 12), ns=(30, 15, 5)),
            "exchanger3": cases.exchanger3(n=(20, 8, 6)),
            "two_slabs": two_slabs(7, 5, 1.5, 30.0, 0.35, 0.25, 290.0, 370.0, nxy=(3, 2))}[name]
    O = ob.OracleSystem(case)
    A = gpu.CoupledFvMatrix(0)
    A.add_case(case)
    A.commit()
    A.solve(tolerance=1e-12, maxIter=2000)
    psi = A.psi(0)
    for i in range(len(O.interfaces)):
        for side in (0, 1):
            tn, tw = A.interface_patch_values(i, side)
            TN, TW = O.interface_patch_values(i, side, psi)
            assert np.array_equal(tn, TN), (i, side)
            assert np.array_equal(tw, TW), (i, side)


@pytest.mark.gpu
def test_gpu_patch_values_of_ldu_defined_system(gpu):
    """A system handed over as LDU arrays with its wall conductances
    (mrfg_set_interface_wall) gives the patch values of the same case added
    whole; without them the call fails cleanly; after a value refresh of
    the wall conductances the values follow."""
    case = cases.duct3d(nf=(14, 8, 6), ns=(18, 10, 3))
    C = m.Case(case)
    A = gpu.CoupledFvMatrix(0)
    A.add_case(C)
    A.commit()
    A.solve(tolerance=1e-12, maxIter=2000)
    regs, ifs = C.regions(), C.interfaces()
    B = gpu.CoupledFvMatrix(0)
    B.add_system(regs, ifs)
    B.commit()
    B.solve(tolerance=1e-12, maxIter=2000)
    for side in (0, 1):
        a, b = A.interface_patch_values(0, side), B.interface_patch_values(0, side)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    bare = [{k: v for k, v in I.items() if not k.startswith("kOwn") and not k.startswith("kNei")} for I in ifs]
    D = gpu.CoupledFvMatrix(0)
    D.add_system(regs, bare)
    D.commit()
    with pytest.raises(gpu.MrfgError, match="wall"):
        D.interface_patch_values(0, 0)
    I = ifs[0]
    D.set_interface_wall(0, 0, 2.0 * I["kOwnA"], I["kNeiA"], I["kOwnB"], I["kNeiB"])
    D.solve(tolerance=1e-12, maxIter=2000)
    tn, tw = D.interface_patch_values(0, 0)
    To = D.psi(0)[0][I["fcA"]]
    expect = (2.0 * I["kOwnA"] * To + I["kNeiA"] * tn) / (2.0 * I["kOwnA"] + I["kNeiA"])
    assert np.array_equal(tw, expect)


@pytest.mark.gpu
def test_gpu_patch_values_follow_device_assembly(gpu, tmp_path_factory):
    """A case read from disk, re-assembled on the device with a new fluid
    conductivity: the wall values use the re-assembled conductances and
    equal those of the case read with that conductivity."""
    import foam_fixture as ff
    d0 = tmp_path_factory.mktemp("pv0")
    ff.write_fixture_case(str(d0))
    base = gpu.Case.from_foam(str(d0), ddt="steadyState", div="upwind")
    d1 = tmp_path_factory.mktemp("pv1")
    ff.write_fixture_case(str(d1), params=dict(kf=2.5, cp=1000.0))
    ref = gpu.Case.from_foam(str(d1), ddt="steadyState", div="upwind")
    A = gpu.CoupledFvMatrix(0)
    A.add_case(base)
    A.commit()
    n = base.regions()[0]["nCells"]
    A.set_field(0, 0, "k", np.full(n, 2.5))
    A.set_field(0, 0, "rhoCp", 1000.0)
    A.assemble()
    A.solve(tolerance=1e-12, maxIter=2000)
    R = gpu.CoupledFvMatrix(0)
    R.add_case(ref)
    R.commit()
    R.solve(tolerance=1e-12, maxIter=2000)
    assert np.array_equal(np.concatenate(A.psi(0)), np.concatenate(R.psi(0)))
    for side in (0, 1):
        a, b = A.interface_patch_values(0, side), R.interface_patch_values(0, side)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), side
