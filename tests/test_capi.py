"""The C-ABI library loads and exports every symbol include/mrfg.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

import multiregionfoam_amd as m

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mrfg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mrfg_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "mrfg_solve" in syms and "mrfg_set_interface" in syms
    assert set(syms) == set(m.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(m.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_errors_without_context():
    L = m.lib()
    assert b"gfx950" in L.mrfg_version()
    assert L.mrfg_commit(None) == -1
    assert L.mrfg_solve(None, None, None) == -1


def test_parse_controls_reference_syntax():
    fvs = """
    FoamFile { version 2.0; format ascii; class dictionary; object fvSolution; }
    solvers
    {
        Tcoupled
        {
            solver           BiCGStab;
            preconditioner
            {
                preconditioner  Cholesky;
            }
            tolerance        1e-15;
            relTol           0;
            minIter          0;
            maxIter          200;
        };
        /* a commented-out entry */
        p { solver PCG; preconditioner DIC; tolerance 1e-12; relTol 0.01; }
    }
    """
    c = m.parse_controls(fvs, "Tcoupled")
    assert c.solver == 0 and c.precision == 0
    assert c.tolerance == 1e-15 and c.rel_tol == 0.0 and c.max_iter == 200 and c.min_iter == 0
    p = m.parse_controls(fvs, "p")
    assert p.solver == 1 and p.rel_tol == 0.01


def test_parse_controls_gpu_keys():
    c = m.controls(solver="hipBiCGStab", tolerance=1e-10, precision="mixed", innerRelTol=1e-3, maxOuter=7)
    assert c.precision == 1 and c.inner_rel_tol == 1e-3 and c.max_outer == 7
    # the inner preconditioner: tile-local DILU by default (also the zero value)
    assert c.inner_preconditioner == m.MRFG_INNER_TILE_DILU == 0
    d = m.controls(precision="mixed", innerPreconditioner="DILU")
    assert d.inner_preconditioner == m.MRFG_INNER_DILU
    e = m.parse_controls("solvers { T { solver BiCGStab; precision mixed; } }", "T")
    assert e.inner_preconditioner == m.MRFG_INNER_TILE_DILU


@pytest.mark.parametrize("bad,msg", [
    ("solvers { T { solver GAMG; } }", "Unknown coupled solver"),
    ("solvers { T { solver BiCGStab; preconditioner { preconditioner FDIC; } } }", "Unknown coupled preconditioner"),
    ("solvers { T { solver BiCGStab; precision half; } }", "Unknown precision"),
    ("solvers { T { solver BiCGStab; refinement { innerPreconditioner ILU; } } }", "Unknown innerPreconditioner"),
    ("solvers { T { solver BiCGStab tolerance 1e-6; }", "unterminated"),
    ("solvers { }", "undefined"),
])
def test_parse_controls_errors(bad, msg):
    with pytest.raises(m.MrfgError) as e:
        m.parse_controls(bad, "T")
    assert msg in str(e.value)
