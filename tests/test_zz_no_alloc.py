"""The allocation-free interface heat-flux check
(mrfg_interface_flux_noalloc), called through the C ABI: the same results
as mrfg_interface_flux bit for bit, on two calls in a row of which the
second has the smaller input (the process-lifetime scratch keeps the first
call's contents), and a clear error for an interface side larger than the
scratch holds.
"""
import ctypes as C

import numpy as np
import pytest

from multiregionfoam_amd import cases

pytestmark = pytest.mark.gpu


def _flux(gpu, A, fn, sys=0, iface=0):
    qa, qb = C.c_double(), C.c_double()
    rc = getattr(gpu.lib(), fn)(A._h, sys, iface, C.byref(qa), C.byref(qb))
    return rc, qa.value, qb.value


def _solved(gpu, case):
    A = gpu.CoupledFvMatrix(0)
    A.add_case(case)
    A.commit()
    A.solve(tolerance=1e-12, maxIter=2000)
    return A


def test_noalloc_flux_equals_allocating_entry_point(gpu):
    big = _solved(gpu, cases.duct3d(nf=(48, 24, 24), ns=(60, 30, 6)))
    small = _solved(gpu, cases.duct3d(nf=(12, 6, 6), ns=(15, 8, 3)))
    for A in (big, small, big):        # the second call has the smaller input
        rc, qa, qb = _flux(gpu, A, "mrfg_interface_flux_noalloc")
        assert rc == 0
        rc2, QA, QB = _flux(gpu, A, "mrfg_interface_flux")
        assert rc2 == 0
        assert qa == QA and qb == QB
        assert np.isfinite(qa) and qa != 0.0
    # a three-region system: each interface in turn through the same scratch
    E = _solved(gpu, cases.exchanger3(n=(20, 8, 6)))
    for i in (1, 0):
        assert _flux(gpu, E, "mrfg_interface_flux_noalloc", iface=i) == _flux(gpu, E, "mrfg_interface_flux", iface=i)


def test_noalloc_flux_rejects_an_interface_beyond_its_scratch(gpu):
    """1050 x 1000 = 1,050,000 faces per side, more than the 1,048,576 the
    scratch holds: a clear error, and the allocating entry point still
    answers."""
    a = cases.region("a", "conductTemperature", (1050, 1000, 1), (0, 0, 0), (1.05, 1.0, 0.01), k=dict(uniform=2.0),
                     T0=300.0, patches=dict(zmin=dict(type="fixedValue", value=300.0),
                                            zmax=dict(type="monolithicTemperature")))
    b = cases.region("b", "conductTemperature", (1050, 1000, 1), (0, 0, 0.01), (1.05, 1.0, 0.01),
                     k=dict(uniform=5.0), T0=350.0, patches=dict(zmin=dict(type="monolithicTemperature"),
                                                                 zmax=dict(type="fixedValue", value=400.0)))
    A = gpu.CoupledFvMatrix(0)
    A.add_case(cases.case([a, b], [cases.interface("ab", ("a", "zmax"), ("b", "zmin"), method="directMap")]))
    A.commit()
    rc, _, _ = _flux(gpu, A, "mrfg_interface_flux_noalloc")
    assert rc == -1
    assert b"1048576" in gpu.lib().mrfg_last_error(A._h)
    rc2, qa, qb = _flux(gpu, A, "mrfg_interface_flux")
    assert rc2 == 0 and np.isfinite(qa)
