"""MI355X-native monolithic interface-coupled solve (multiRegionFoam hot path).

Python mirror of the reference's coupled-solve interface over the C ABI of
``libmrfg.so`` (include/mrfg.h):

* :class:`CoupledFvMatrix` plays ``coupledFvMatrix<scalar>`` -- one or more
  coupled systems resident on a GPU, ``solve(dict)`` with the controls of an
  fvSolution solver entry (multiRegionSystem.C:150-153).
* :class:`Case` assembles a case dictionary natively (regionType
  ``setCoupledEqns`` and the monolithic patch coefficients).
* :func:`parse_controls` reads an fvSolution solver entry.

The HIP library is required: every entry point raises if ``libmrfg.so`` is
missing or no GPU is present; nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import cases  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmrfg.so")

MRFG_OK = 0
MRFG_INNER_TILE_DILU, MRFG_INNER_DILU = 0, 1   # the mixed solver's inner preconditioners
ERRORS = {-1: "argument", -2: "device", -3: "dictionary", -4: "state"}


class MrfgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mrfg %s error (%d): %s" % (ERRORS.get(code, "unknown"), code, msg))
        self.code = code


class Perf(C.Structure):
    _fields_ = [("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("n_iterations", C.c_int), ("converged", C.c_int), ("singular", C.c_int),
                ("n_outer", C.c_int), ("seconds", C.c_double), ("bytes_moved", C.c_double)]


class DnaPerf(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("region_solves", C.c_int),
                ("solver_iterations", C.c_int), ("jump_residual", C.c_double), ("flux_residual", C.c_double),
                ("seconds", C.c_double)]


class Controls(C.Structure):
    _fields_ = [("solver", C.c_int), ("precision", C.c_int), ("tolerance", C.c_double),
                ("rel_tol", C.c_double), ("min_iter", C.c_int), ("max_iter", C.c_int),
                ("inner_rel_tol", C.c_double), ("inner_max_iter", C.c_int),
                ("max_outer", C.c_int), ("check_interval", C.c_int), ("launch_mode", C.c_int),
                ("tuning", C.c_int), ("inner_preconditioner", C.c_int)]


_lib = None

I32P = C.POINTER(C.c_int32)
F64P = C.POINTER(C.c_double)
F32P = C.POINTER(C.c_float)


def lib():
    """Load libmrfg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmrfg.so is not built (run __graft_entry__.build() or make -C "
                          "multiregionfoam_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "mrfg_version": ([], C.c_char_p),
        "mrfg_create": ([C.POINTER(vp), C.c_int], C.c_int),
        "mrfg_destroy": ([vp], None),
        "mrfg_last_error": ([vp], C.c_char_p),
        "mrfg_begin_system": ([vp, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "mrfg_set_region": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, I32P, I32P, F64P, F64P, F64P, F64P, F64P],
                            C.c_int),
        "mrfg_set_interface": ([vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_int, I32P, F64P, I32P, I32P, F64P,
                                C.c_int, I32P, F64P, I32P, I32P, F64P], C.c_int),
        "mrfg_set_region_ddt": ([vp, C.c_int, C.c_int, F64P, F64P, F64P], C.c_int),
        "mrfg_commit": ([vp], C.c_int),
        "mrfg_update_region": ([vp, C.c_int, C.c_int, F64P, F64P, F64P, F64P], C.c_int),
        "mrfg_update_interface": ([vp, C.c_int, C.c_int, F64P, F64P], C.c_int),
        "mrfg_set_schedule": ([vp, C.c_int], C.c_int),
        "mrfg_solve": ([vp, C.POINTER(Controls), C.POINTER(Perf)], C.c_int),
        "mrfg_solve_dict": ([vp, C.c_char_p, C.c_char_p, C.POINTER(Perf)], C.c_int),
        "mrfg_time_steps": ([vp, C.POINTER(Controls), C.c_char_p, C.c_int, C.POINTER(Perf)], C.c_int),
        "mrfg_dna_solve": ([vp, C.POINTER(Controls), C.POINTER(DnaPerf)], C.c_int),
        "mrfg_dna_history": ([vp, C.c_int, C.c_int, F64P, F64P, F64P, C.POINTER(C.c_int)], C.c_int),
        "mrfg_dna_time_steps": ([vp, C.POINTER(Controls), C.c_char_p, C.c_int, C.POINTER(DnaPerf)], C.c_int),
        "mrfg_parse_controls": ([C.c_char_p, C.c_char_p, C.POINTER(Controls), C.c_char_p, C.c_int], C.c_int),
        "mrfg_get_psi": ([vp, C.c_int, C.c_int, F64P], C.c_int),
        "mrfg_reset_psi": ([vp], C.c_int),
        "mrfg_set_psi": ([vp, C.c_int, C.c_int, F64P], C.c_int),
        "mrfg_case_build": ([C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_int], C.c_int),
        "mrfg_case_free": ([vp], None),
        "mrfg_case_read_foam": ([C.c_char_p, C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_int], C.c_int),
        "mrfg_case_sizes": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "mrfg_case_region_sizes": ([vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
                                   C.c_int),
        "mrfg_case_region": ([vp, C.c_int, I32P, I32P, F64P, F64P, F64P, F64P, F64P], C.c_int),
        "mrfg_case_region_ddt": ([vp, C.c_int, C.POINTER(C.c_int), F64P, F64P, F64P], C.c_int),
        "mrfg_case_dna_controls": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int), F64P, F64P], C.c_int),
        "mrfg_case_dna_side_sizes": ([vp, C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 4 + [F64P, F64P], C.c_int),
        "mrfg_case_dna_side": ([vp, C.c_int, C.c_int, I32P, F64P, F64P, F64P, F64P, I32P, I32P, F64P], C.c_int),
        "mrfg_case_region_source_base": ([vp, C.c_int, F64P], C.c_int),
        "mrfg_case_interface_sizes": ([vp, C.c_int] + [C.POINTER(C.c_int)] * 6, C.c_int),
        "mrfg_case_interface": ([vp, C.c_int, I32P, F64P, I32P, I32P, F64P, I32P, F64P, I32P, I32P, F64P],
                                C.c_int),
        "mrfg_add_case": ([vp, vp, C.POINTER(C.c_int)], C.c_int),
        "mrfg_case_interface_wall": ([vp, C.c_int, F64P, F64P, F64P, F64P], C.c_int),
        "mrfg_set_interface_wall": ([vp, C.c_int, C.c_int, F64P, F64P, F64P, F64P], C.c_int),
        "mrfg_interface_patch_values": ([vp, C.c_int, C.c_int, C.c_int, F64P, F64P], C.c_int),
        "mrfg_spmv": ([vp, C.c_int, F64P, F64P], C.c_int),
        "mrfg_dilu_factor": ([vp, C.c_int, F64P], C.c_int),
        "mrfg_precondition": ([vp, C.c_int, F64P, F64P], C.c_int),
        "mrfg_precondition_fp32": ([vp, C.c_int, F32P, F32P], C.c_int),
        "mrfg_precondition_inner": ([vp, C.c_int, C.c_int, F32P, F32P], C.c_int),
        "mrfg_dilu_tiles": ([vp, C.c_int, I32P], C.c_int),
        "mrfg_time_kernel": ([vp, C.c_int, C.c_int, F64P], C.c_int),
        "mrfg_layout_info": ([vp, C.c_int, C.POINTER(C.c_longlong)], C.c_int),
        "mrfg_tuning": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), F64P, F64P], C.c_int),
        "mrfg_set_tuning": ([vp, C.c_int, C.c_int], C.c_int),
        "mrfg_trace_dilu": ([vp, C.POINTER(C.c_longlong), C.c_int, C.POINTER(C.c_int)], C.c_int),
        "mrfg_bytes": ([vp, C.c_int, C.c_int, F64P, F64P, F64P], C.c_int),
        "mrfg_copy_probe": ([vp, C.c_longlong, C.c_int, F64P], C.c_int),
        "mrfg_interface_flux": ([vp, C.c_int, C.c_int, F64P, F64P], C.c_int),
        "mrfg_interface_flux_noalloc": ([vp, C.c_int, C.c_int, F64P, F64P], C.c_int),
        "mrfg_set_field": ([vp, C.c_int, C.c_int, C.c_char_p, F64P, C.c_longlong], C.c_int),
        "mrfg_assemble": ([vp], C.c_int),
        "mrfg_assembly_bytes": ([vp, F64P, F64P], C.c_int),
        "mrfg_get_region_matrix": ([vp, C.c_int, C.c_int, F64P, F64P, F64P, F64P], C.c_int),
        "mrfg_get_interface_coeff": ([vp, C.c_int, C.c_int, C.c_int, F64P], C.c_int),
        "mrfg_server_create": ([C.POINTER(vp), C.c_int, C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                C.c_char_p, C.c_int], C.c_int),
        "mrfg_server_destroy": ([vp], None),
        "mrfg_server_submit": ([vp, C.c_char_p, C.POINTER(C.c_int)], C.c_int),
        "mrfg_server_wait_all": ([vp], C.c_int),
        "mrfg_server_result": ([vp, C.c_int, C.c_void_p, F64P, C.POINTER(C.c_int)], C.c_int),
        "mrfg_server_get_psi": ([vp, C.c_int, C.c_int, F64P], C.c_int),
        "mrfg_server_times": ([vp, C.c_int, F64P], C.c_int),
        "mrfg_server_last_error": ([vp], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


EXPORTED = (
    "mrfg_version mrfg_create mrfg_destroy mrfg_last_error mrfg_begin_system mrfg_set_region mrfg_set_region_ddt "
    "mrfg_set_interface mrfg_set_schedule mrfg_commit mrfg_update_region mrfg_update_interface mrfg_solve mrfg_solve_dict mrfg_time_steps mrfg_parse_controls "
    "mrfg_dna_solve mrfg_dna_history mrfg_dna_time_steps "
    "mrfg_get_psi mrfg_reset_psi "
    "mrfg_set_psi mrfg_case_build mrfg_case_read_foam mrfg_case_free mrfg_case_sizes mrfg_case_region_sizes "
    "mrfg_case_region mrfg_case_region_ddt mrfg_case_dna_controls mrfg_case_dna_side_sizes mrfg_case_dna_side "
    "mrfg_case_region_source_base "
    "mrfg_case_interface_sizes mrfg_case_interface mrfg_case_interface_wall mrfg_set_interface_wall "
    "mrfg_interface_patch_values mrfg_add_case mrfg_spmv mrfg_dilu_factor "
    "mrfg_precondition mrfg_precondition_fp32 mrfg_precondition_inner mrfg_dilu_tiles mrfg_time_kernel mrfg_trace_dilu mrfg_layout_info mrfg_tuning mrfg_set_tuning mrfg_bytes mrfg_copy_probe mrfg_interface_flux mrfg_interface_flux_noalloc "
    "mrfg_set_field mrfg_assemble mrfg_get_region_matrix mrfg_get_interface_coeff mrfg_assembly_bytes "
    "mrfg_server_create mrfg_server_destroy mrfg_server_submit mrfg_server_wait_all mrfg_server_result "
    "mrfg_server_get_psi mrfg_server_times mrfg_server_last_error").split()


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def _text(s):
    return s.encode() if isinstance(s, str) else s


# ---------------------------------------------------------------------------
# Controls
# ---------------------------------------------------------------------------

SOLVER_IDS = {"BiCGStab": 0, "hipBiCGStab": 0, "PBiCGStab": 0, "PCG": 1, "CG": 1, "hipPCG": 1}


def parse_controls(fv_solution: str, entry: str) -> Controls:
    """Read ``solvers { <entry> { ... } }`` of an fvSolution text."""
    c = Controls()
    err = C.create_string_buffer(512)
    rc = lib().mrfg_parse_controls(_text(fv_solution), _text(entry), C.byref(c), err, 512)
    if rc != MRFG_OK:
        raise MrfgError(rc, err.value.decode())
    return c


def _controls_text(solver="BiCGStab", tolerance=1e-10, relTol=0.0, minIter=0, maxIter=1000,
                   precision="fp64", innerRelTol=1e-4, innerMaxIter=1000, maxOuter=50, checkInterval=8,
                   graphs=True, autotune=True, innerPreconditioner="tileDILU"):
    """fvSolution text with one solver entry "X" holding the given keys."""
    return ("solvers { X { solver %s; preconditioner { preconditioner Cholesky; } tolerance %r; "
            "relTol %r; minIter %d; maxIter %d; precision %s; checkInterval %d; graphs %s; autotune %s; "
            "refinement { innerRelTol %r; innerMaxIter %d; maxOuter %d; innerPreconditioner %s; } } }"
            % (solver, float(tolerance), float(relTol), minIter, maxIter, precision, checkInterval,
               "on" if graphs else "off", "on" if autotune else "off", float(innerRelTol), innerMaxIter,
               maxOuter, innerPreconditioner))


def controls(**kw):
    """Build a Controls record from keyword values (same keys as fvSolution)."""
    return parse_controls(_controls_text(**kw), "X")


@dataclass
class SolverPerformance:
    """lduSolverPerformance of one coupled solve.  ``singular``: the Krylov
    recurrence broke down; ``diverged``: a residual became NaN or Inf (the
    system was stopped at once)."""
    initial_residual: float
    final_residual: float
    n_iterations: int
    converged: bool
    singular: bool
    n_outer: int
    seconds: float
    bytes_moved: float
    diverged: bool = False

    @classmethod
    def from_c(cls, p):
        return cls(p.initial_residual, p.final_residual, p.n_iterations, bool(p.converged), p.singular == 1,
                   p.n_outer, p.seconds, p.bytes_moved, p.singular == 2)

    def line(self, solver="BiCGStab", field="T"):
        return ("%s:  Solving for %s, Initial residual = %.6g, Final residual = %.6g, No Iterations %d"
                % (solver, field, self.initial_residual, self.final_residual, self.n_iterations))


# ---------------------------------------------------------------------------
# Case (native assembly)
# ---------------------------------------------------------------------------

class Case:
    """A case dictionary assembled by the native case builder, or (through
    ``Case.from_foam``) an OpenFOAM case directory read from disk."""

    def __init__(self, case=None, _handle=None):
        if _handle is not None:
            self._h = _handle
            return
        text = case if isinstance(case, str) else cases.to_dict_text(case)
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().mrfg_case_build(_text(text), C.byref(h), err, 1024)
        if rc != MRFG_OK:
            raise MrfgError(rc, err.value.decode())
        self._h = h

    @classmethod
    def from_foam(cls, case_dir, fieldDir=None, interfaceProperties=None, ddt=None, deltaT=None, div=None):
        """Read and assemble a multi-region case directory (polyMesh, 0/,
        constant/, system/), as multiRegionFoam sets up its coupled T
        equation; keyword options override what the case files select
        (``div="upwind"``: assemble upwind convection where the case names a
        div(phi,T) scheme the reader does not assemble, e.g. linearUpwind)."""
        opts = []
        for k, v in (("fieldDir", fieldDir), ("interfaceProperties", interfaceProperties), ("ddt", ddt),
                     ("deltaT", deltaT), ("div", div)):
            if v is not None:
                opts.append("%s %s;" % (k, repr(float(v)) if k == "deltaT" else v))
        h = C.c_void_p()
        err = C.create_string_buffer(2048)
        rc = lib().mrfg_case_read_foam(_text(os.fspath(case_dir)), _text(" ".join(opts)), C.byref(h), err, 2048)
        if rc != MRFG_OK:
            raise MrfgError(rc, err.value.decode())
        return cls(_handle=h)

    def __del__(self):
        try:
            if self._h:
                lib().mrfg_case_free(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def sizes(self):
        nr, ni = C.c_int(), C.c_int()
        lib().mrfg_case_sizes(self._h, C.byref(nr), C.byref(ni))
        return nr.value, ni.value

    def region(self, r):
        n, f, sym = C.c_int(), C.c_int(), C.c_int()
        if lib().mrfg_case_region_sizes(self._h, r, C.byref(n), C.byref(f), C.byref(sym)) != MRFG_OK:
            raise IndexError(r)
        n, f = n.value, f.value
        d = dict(nCells=n, nFaces=f, symmetric=bool(sym.value),
                 lAddr=np.zeros(f, np.int32), uAddr=np.zeros(f, np.int32), diag=np.zeros(n),
                 lower=np.zeros(f), upper=np.zeros(f), source=np.zeros(n), psi=np.zeros(n))
        lib().mrfg_case_region(self._h, r, _p(d["lAddr"], C.c_int32), _p(d["uAddr"], C.c_int32),
                               _p(d["diag"], C.c_double), _p(d["lower"], C.c_double),
                               _p(d["upper"], C.c_double), _p(d["source"], C.c_double),
                               _p(d["psi"], C.c_double))
        # transient regions: the Euler ddt coefficient and the diagonal and
        # source without the ddt term
        tr = C.c_int()
        lib().mrfg_case_region_ddt(self._h, r, C.byref(tr), None, None, None)
        if tr.value:
            d["ddtCoeff"], d["diagBase"], d["sourceBase"] = np.zeros(n), np.zeros(n), np.zeros(n)
            lib().mrfg_case_region_ddt(self._h, r, C.byref(tr), _p(d["ddtCoeff"], C.c_double),
                                       _p(d["diagBase"], C.c_double), _p(d["sourceBase"], C.c_double))
        return d

    def interface(self, i):
        v = [C.c_int() for _ in range(6)]
        if lib().mrfg_case_interface_sizes(self._h, i, *[C.byref(x) for x in v]) != MRFG_OK:
            raise IndexError(i)
        ra, rb, na, nb, pa, pb = [x.value for x in v]
        d = dict(regionA=ra, regionB=rb, nA=na, nB=nb,
                 fcA=np.zeros(na, np.int32), coeffA=np.zeros(na), offA=np.zeros(na + 1, np.int32),
                 idxA=np.zeros(pa, np.int32), wA=np.zeros(pa),
                 fcB=np.zeros(nb, np.int32), coeffB=np.zeros(nb), offB=np.zeros(nb + 1, np.int32),
                 idxB=np.zeros(pb, np.int32), wB=np.zeros(pb))
        lib().mrfg_case_interface(self._h, i, _p(d["fcA"], C.c_int32), _p(d["coeffA"], C.c_double),
                                  _p(d["offA"], C.c_int32), _p(d["idxA"], C.c_int32), _p(d["wA"], C.c_double),
                                  _p(d["fcB"], C.c_int32), _p(d["coeffB"], C.c_double),
                                  _p(d["offB"], C.c_int32), _p(d["idxB"], C.c_int32), _p(d["wB"], C.c_double))
        # the wall conductances of the monolithicTemperature patch value
        wall = dict(kOwnA=np.zeros(na), kNeiA=np.zeros(na), kOwnB=np.zeros(nb), kNeiB=np.zeros(nb))
        if lib().mrfg_case_interface_wall(self._h, i, *[_p(wall[k], C.c_double)
                                                        for k in ("kOwnA", "kNeiA", "kOwnB", "kNeiB")]) == MRFG_OK:
            d.update(wall)
        return d

    def partitioned(self):
        """The partitioned (Dirichlet-Neumann) coupling of a case read from
        disk, or None for a monolithic case: DNA controls and, per pair, the
        Dirichlet ("dir") and Neumann ("neu") sides."""
        n, it = C.c_int(), C.c_int()
        jr, fr = C.c_double(), C.c_double()
        lib().mrfg_case_dna_controls(self._h, C.byref(n), C.byref(it), C.byref(jr), C.byref(fr))
        if n.value == 0:
            return None
        pairs = []
        for q in range(n.value):
            pair = {}
            for side, key in ((0, "dir"), (1, "neu")):
                reg, nf, nw, acc = C.c_int(), C.c_int(), C.c_int(), C.c_int()
                relax, K = C.c_double(), C.c_double()
                lib().mrfg_case_dna_side_sizes(self._h, q, side, C.byref(reg), C.byref(nf), C.byref(nw),
                                               C.byref(acc), C.byref(relax), C.byref(K))
                d = dict(region=reg.value, accType=("fixed", "aitken", "IQN-ILS")[acc.value], relax=relax.value,
                         fc=np.zeros(nf.value, np.int32), coeff=np.zeros(nf.value), kb=np.zeros(nf.value),
                         dc=np.zeros(nf.value), value0=np.zeros(nf.value), off=np.zeros(nf.value + 1, np.int32),
                         idx=np.zeros(nw.value, np.int32), w=np.zeros(nw.value))
                lib().mrfg_case_dna_side(self._h, q, side, *[_p(d[k], C.c_int32 if d[k].dtype == np.int32
                                                                 else C.c_double)
                                                              for k in ("fc", "coeff", "kb", "dc", "value0", "off",
                                                                        "idx", "w")])
                pair[key] = d
                pair["K"] = K.value
            pairs.append(pair)
        nreg = len(self.regions())
        base = []
        for r in range(nreg):
            b = np.zeros(self.region(r)["nCells"])
            lib().mrfg_case_region_source_base(self._h, r, _p(b, C.c_double))
            base.append(b)
        return dict(pairs=pairs, maxCoupleIter=it.value, maxJumpRes=jr.value, maxFluxRes=fr.value,
                    sourceBase=base)

    def regions(self):
        return [self.region(r) for r in range(self.sizes()[0])]

    def interfaces(self):
        return [self.interface(i) for i in range(self.sizes()[1])]


# ---------------------------------------------------------------------------
# Coupled matrix (GPU-resident systems)
# ---------------------------------------------------------------------------

class CoupledFvMatrix:
    """One or more coupled block systems resident on a GPU.

    ``add_case`` / ``add_system`` define systems, ``commit`` uploads them,
    ``solve`` runs the coupled solver on all of them at once.
    """

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().mrfg_create(C.byref(h), device)
        if rc != MRFG_OK:
            raise MrfgError(rc, "cannot create a context on device %d (is a GPU present?)" % device)
        self._h = h
        self._sizes = []
        self._faces = []
        self._ifaces = []

    def __del__(self):
        try:
            if self._h:
                lib().mrfg_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def close(self):
        self.__del__()

    def _check(self, rc):
        if rc != MRFG_OK:
            raise MrfgError(rc, lib().mrfg_last_error(self._h).decode())

    def add_case(self, case):
        if not isinstance(case, Case):
            case = Case(case)
        sid = C.c_int()
        self._check(lib().mrfg_add_case(self._h, case.handle, C.byref(sid)))
        regions = case.regions()
        self._sizes.append([r["nCells"] for r in regions])
        self._faces.append([r["nFaces"] for r in regions])
        self._ifaces.append([(I["nA"], I["nB"]) for I in case.interfaces()])
        return sid.value

    def add_system(self, regions, interfaces):
        """Define a system from LDU arrays (dicts with the keys of Case.region
        / Case.interface)."""
        sid = C.c_int()
        self._check(lib().mrfg_begin_system(self._h, len(regions), len(interfaces), C.byref(sid)))
        s = sid.value
        keep = []
        for r, R in enumerate(regions):
            arr = {k: np.ascontiguousarray(R[k], dtype=np.int32 if k in ("lAddr", "uAddr") else np.float64)
                   for k in ("lAddr", "uAddr", "diag", "upper", "source", "psi")}
            lower = None if R.get("symmetric", False) else np.ascontiguousarray(R["lower"], dtype=np.float64)
            keep.append((arr, lower))
            self._check(lib().mrfg_set_region(
                self._h, s, r, int(R["nCells"]), int(R["nFaces"]), _p(arr["lAddr"], C.c_int32),
                _p(arr["uAddr"], C.c_int32), _p(arr["diag"], C.c_double), _p(lower, C.c_double),
                _p(arr["upper"], C.c_double), _p(arr["source"], C.c_double), _p(arr["psi"], C.c_double)))
            if R.get("ddtCoeff") is not None:
                dd = [np.ascontiguousarray(R[k], dtype=np.float64) for k in ("ddtCoeff", "diagBase", "sourceBase")]
                keep.append(dd)
                self._check(lib().mrfg_set_region_ddt(self._h, s, r, *[_p(x, C.c_double) for x in dd]))
        for i, I in enumerate(interfaces):
            a = {k: np.ascontiguousarray(I[k], dtype=np.int32 if k[:2] in ("fc", "of", "id") else np.float64)
                 for k in ("fcA", "coeffA", "offA", "idxA", "wA", "fcB", "coeffB", "offB", "idxB", "wB")}
            keep.append(a)
            self._check(lib().mrfg_set_interface(
                self._h, s, i, int(I["regionA"]), int(I["regionB"]),
                int(I["nA"]), _p(a["fcA"], C.c_int32), _p(a["coeffA"], C.c_double), _p(a["offA"], C.c_int32),
                _p(a["idxA"], C.c_int32), _p(a["wA"], C.c_double),
                int(I["nB"]), _p(a["fcB"], C.c_int32), _p(a["coeffB"], C.c_double), _p(a["offB"], C.c_int32),
                _p(a["idxB"], C.c_int32), _p(a["wB"], C.c_double)))
            if I.get("kOwnA") is not None:
                wl = [np.ascontiguousarray(I[k], dtype=np.float64) for k in ("kOwnA", "kNeiA", "kOwnB", "kNeiB")]
                keep.append(wl)
                self._check(lib().mrfg_set_interface_wall(self._h, s, i, *[_p(x, C.c_double) for x in wl]))
        self._sizes.append([int(R["nCells"]) for R in regions])
        self._faces.append([int(R["nFaces"]) for R in regions])
        self._ifaces.append([(int(I["nA"]), int(I["nB"])) for I in interfaces])
        return s

    def commit(self, schedule="auto"):
        """Build the device layout; schedule "auto" (line tiles on structured
        hex regions) or "levels" (level schedule everywhere)."""
        self._check(lib().mrfg_set_schedule(self._h, {"auto": 0, "levels": 1}[schedule]))
        self._check(lib().mrfg_commit(self._h))

    @property
    def n_systems(self):
        return len(self._sizes)

    def update_region(self, sys, region, diag, upper, source, lower=None):
        """New values of one committed region on its committed addressing
        (the next time step's fvMatrix); ``lower`` None for a region
        committed symmetric.  Placed on the device before the next solve."""
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        d, lo, up, src = f64(diag), f64(lower), f64(upper), f64(source)
        if d.size != self._sizes[sys][region] or src.size != d.size or up.size != self._faces[sys][region] \
                or (lo is not None and lo.size != up.size):
            raise ValueError("value arrays do not match the committed region sizes")
        self._check(lib().mrfg_update_region(self._h, sys, region, _p(d, C.c_double), _p(lo, C.c_double),
                                             _p(up, C.c_double), _p(src, C.c_double)))

    def update_interface(self, sys, iface, coeffA=None, coeffB=None):
        """New boundary coefficients of one interface side or both (None:
        unchanged); the GGI weights stay those committed."""
        na, nb = self._ifaces[sys][iface]
        a = None if coeffA is None else np.ascontiguousarray(coeffA, dtype=np.float64)
        b = None if coeffB is None else np.ascontiguousarray(coeffB, dtype=np.float64)
        if (a is not None and a.size != na) or (b is not None and b.size != nb):
            raise ValueError("coefficient arrays do not match the interface sizes")
        self._check(lib().mrfg_update_interface(self._h, sys, iface, _p(a, C.c_double), _p(b, C.c_double)))

    def solve(self, ctrl=None, fv_solution=None, entry="Tcoupled", **kw):
        """Solve every system.  ``ctrl``: Controls; or ``fv_solution`` text +
        ``entry``; or keyword controls (see :func:`controls`)."""
        perf = (Perf * max(1, self.n_systems))()
        if fv_solution is not None:
            self._check(lib().mrfg_solve_dict(self._h, _text(fv_solution), _text(entry), perf))
        else:
            if ctrl is None:
                ctrl = controls(**kw)
            self._check(lib().mrfg_solve(self._h, C.byref(ctrl), perf))
        return [SolverPerformance.from_c(p) for p in perf[:self.n_systems]]

    def time_steps(self, n_steps, ddt="backward", ctrl=None, fv_solution=None, entry="Tcoupled", **kw):
        """``n_steps`` time steps of the transient systems (the ddt terms
        rebuilt on the device each step; Euler, or backward whose first step
        in a run is Euler).  Returns, per step, the per-system
        SolverPerformance list."""
        if fv_solution is not None:
            ctrl = parse_controls(fv_solution, entry)
        elif ctrl is None:
            ctrl = controls(**kw)
        n = max(1, self.n_systems)
        perf = (Perf * (n * n_steps))()
        self._check(lib().mrfg_time_steps(self._h, C.byref(ctrl), _text(ddt), int(n_steps), perf))
        out = []
        for k in range(n_steps):
            out.append([SolverPerformance.from_c(p) for p in perf[k * n:(k + 1) * n]])
        return out

    # ---- finite-volume assembly on the device (cases read from disk) --------
    def set_field(self, sys, region, name, values):
        """Replace a field of a system read from disk, for the next
        :meth:`assemble`: "k" (per cell), "U" (n_cells x 3, transport
        regions), "T" (per cell; also the initial guess), "rhoCp" (scalar)."""
        a = np.ascontiguousarray(np.asarray(values, dtype=np.float64).ravel())
        self._check(lib().mrfg_set_field(self._h, sys, region, _text(name), _p(a, C.c_double), a.size))

    def assemble(self):
        """Assemble every system read from disk from its current fields on
        the device (setCoupledEqns + interface coefficients + ddt), in place
        of the committed matrices."""
        self._check(lib().mrfg_assemble(self._h))

    def assembly_bytes(self):
        """Bytes read and written by the two passes of :meth:`assemble`
        (assembly, placement), every array once."""
        a, b = C.c_double(), C.c_double()
        self._check(lib().mrfg_assembly_bytes(self._h, C.byref(a), C.byref(b)))
        return dict(assemble=a.value, place=b.value)

    def region_matrix(self, sys, region):
        """The device-assembled LDU arrays of one region (after assemble)."""
        n, nf = self._sizes[sys][region], self._faces[sys][region]
        d, lo, up, src = np.zeros(n), np.zeros(nf), np.zeros(nf), np.zeros(n)
        self._check(lib().mrfg_get_region_matrix(self._h, sys, region, _p(d, C.c_double), _p(lo, C.c_double),
                                                 _p(up, C.c_double), _p(src, C.c_double)))
        return dict(diag=d, lower=lo, upper=up, source=src)

    def interface_coeff(self, sys, iface, side):
        """The device-assembled coefficients of one interface side (0 = a, 1 = b)."""
        c = np.zeros(self._ifaces[sys][iface][side])
        self._check(lib().mrfg_get_interface_coeff(self._h, sys, iface, side, _p(c, C.c_double)))
        return c

    # ---- partitioned (Dirichlet-Neumann) coupling ---------------------------
    def dna_solve(self, ctrl=None, fv_solution=None, entry="T", **kw):
        """Partitioned coupling of every committed system (cases read with
        regionCoupledTemperatureJump / regionCoupledHeatFlux patches); the
        region solves use ``ctrl`` / the fvSolution ``entry`` / keyword
        controls.  Returns one dict per system."""
        if fv_solution is not None:
            ctrl = parse_controls(fv_solution, entry)
        elif ctrl is None:
            ctrl = controls(**kw)
        perf = (DnaPerf * max(1, self.n_systems))()
        self._check(lib().mrfg_dna_solve(self._h, C.byref(ctrl), perf))
        return [dict(iterations=p.iterations, converged=bool(p.converged), region_solves=p.region_solves,
                     solver_iterations=p.solver_iterations, jump=p.jump_residual, flux=p.flux_residual,
                     seconds=p.seconds) for p in perf[:self.n_systems]]

    def dna_time_steps(self, n_steps, ddt="backward", ctrl=None, fv_solution=None, entry="T", **kw):
        """``n_steps`` time steps of the committed partitioned cases, each a
        DNA loop; returns, per step, one dict per system."""
        if fv_solution is not None:
            ctrl = parse_controls(fv_solution, entry)
        elif ctrl is None:
            ctrl = controls(**kw)
        n = max(1, self.n_systems)
        perf = (DnaPerf * (n * n_steps))()
        self._check(lib().mrfg_dna_time_steps(self._h, C.byref(ctrl), _text(ddt), int(n_steps), perf))
        conv = lambda p: dict(iterations=p.iterations, converged=bool(p.converged), region_solves=p.region_solves,
                              solver_iterations=p.solver_iterations, jump=p.jump_residual, flux=p.flux_residual,
                              seconds=p.seconds)
        return [[conv(p) for p in perf[k * n:(k + 1) * n]] for k in range(n_steps)]

    def dna_history(self, sys=0, max_iter=1000):
        """Per DNA iteration of the last dna_solve: (jump, flux, omega) arrays."""
        j, f, w = np.zeros(max_iter), np.zeros(max_iter), np.zeros(max_iter)
        n = C.c_int()
        self._check(lib().mrfg_dna_history(self._h, sys, max_iter, _p(j, C.c_double), _p(f, C.c_double),
                                           _p(w, C.c_double), C.byref(n)))
        k = min(n.value, max_iter)
        return j[:k], f[:k], w[:k]

    def psi(self, sys=0):
        out = []
        for r, n in enumerate(self._sizes[sys]):
            a = np.zeros(n)
            self._check(lib().mrfg_get_psi(self._h, sys, r, _p(a, C.c_double)))
            out.append(a)
        return out

    def reset_psi(self):
        """Reset every system's field to its initial guess (device copy)."""
        self._check(lib().mrfg_reset_psi(self._h))

    def set_psi(self, sys, psi):
        for r, a in enumerate(psi):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._check(lib().mrfg_set_psi(self._h, sys, r, _p(a, C.c_double)))

    # ---- kernel-level -----------------------------------------------------
    def spmv(self, x, sys=0):
        x = np.ascontiguousarray(np.concatenate(x) if isinstance(x, (list, tuple)) else x, dtype=np.float64)
        y = np.zeros_like(x)
        self._check(lib().mrfg_spmv(self._h, sys, _p(x, C.c_double), _p(y, C.c_double)))
        return y

    def dilu_factor(self, sys=0):
        n = sum(self._sizes[sys])
        rD = np.zeros(n)
        self._check(lib().mrfg_dilu_factor(self._h, sys, _p(rD, C.c_double)))
        return rD

    def precondition(self, r, sys=0):
        r = np.ascontiguousarray(np.concatenate(r) if isinstance(r, (list, tuple)) else r, dtype=np.float64)
        w = np.zeros_like(r)
        self._check(lib().mrfg_precondition(self._h, sys, _p(r, C.c_double), _p(w, C.c_double)))
        return w

    def precondition_fp32(self, r, sys=0):
        """w = M^-1 r with the FP32 records of the mixed-precision solver."""
        r = np.ascontiguousarray(np.concatenate(r) if isinstance(r, (list, tuple)) else r, dtype=np.float32)
        w = np.zeros_like(r)
        self._check(lib().mrfg_precondition_fp32(self._h, sys, _p(r, C.c_float), _p(w, C.c_float)))
        return w

    def precondition_inner(self, r, inner="tileDILU", sys=0):
        """w = M^-1 r with the mixed solver's FP32 inner preconditioner:
        "tileDILU" (each box tile's DILU alone) or "DILU" (the full DILU)."""
        mode = {"tileDILU": MRFG_INNER_TILE_DILU, "DILU": MRFG_INNER_DILU}[inner]
        r = np.ascontiguousarray(np.concatenate(r) if isinstance(r, (list, tuple)) else r, dtype=np.float32)
        w = np.zeros_like(r)
        self._check(lib().mrfg_precondition_inner(self._h, sys, mode, _p(r, C.c_float), _p(w, C.c_float)))
        return w

    def dilu_tiles(self, sys=0):
        """Per cell (concatenated regions): the DILU schedule's box tile
        holding it, or -1 for a cell on general waves."""
        out = np.zeros(sum(self._sizes[sys]), dtype=np.int32)
        self._check(lib().mrfg_dilu_tiles(self._h, sys, _p(out, C.c_int32)))
        return out

    def interface_flux(self, iface=0, sys=0):
        """(Q_a, Q_b): heat flow into each side of a monolithic interface from
        the current FP64 field, computed on the device."""
        qa, qb = C.c_double(), C.c_double()
        self._check(lib().mrfg_interface_flux(self._h, sys, iface, C.byref(qa), C.byref(qb)))
        return qa.value, qb.value

    def interface_patch_values(self, iface=0, side=0, sys=0):
        """(t_neighbour, t_wall) of one interface side from the current FP64
        field: the GGI-interpolated neighbour field and the
        monolithicTemperature patch value, computed on the device."""
        n = self._ifaces[sys][iface][side]
        tn, tw = np.zeros(n), np.zeros(n)
        self._check(lib().mrfg_interface_patch_values(self._h, sys, iface, side, _p(tn, C.c_double),
                                                      _p(tw, C.c_double)))
        return tn, tw

    def set_interface_wall(self, sys, iface, kOwnA, kNeiA, kOwnB, kNeiB):
        """Wall conductances of the patch values (per face of each side)."""
        wl = [np.ascontiguousarray(x, dtype=np.float64) for x in (kOwnA, kNeiA, kOwnB, kNeiB)]
        self._check(lib().mrfg_set_interface_wall(self._h, sys, iface, *[_p(x, C.c_double) for x in wl]))

    def copy_probe(self, nbytes, reps=10):
        """Average ms of a plain device copy of nbytes (reads + writes nbytes)."""
        ms = C.c_double()
        self._check(lib().mrfg_copy_probe(self._h, int(nbytes), reps, C.byref(ms)))
        return ms.value

    def time_kernel(self, which, reps=20):
        ms = C.c_double()
        self._check(lib().mrfg_time_kernel(self._h, which, reps, C.byref(ms)))
        return ms.value

    def trace_dilu(self, max_waves=100000):
        """Per-wave [start_ns, end_ns, slow_waits, steps] of one forward and
        one backward DILU sweep (arrays of shape (n_waves, 4))."""
        out = np.zeros(8 * max_waves, dtype=np.int64)
        n = C.c_int()
        self._check(lib().mrfg_trace_dilu(self._h, out.ctypes.data_as(C.POINTER(C.c_longlong)), max_waves,
                                          C.byref(n)))
        m = min(n.value, max_waves)
        return out[:4 * m].reshape(m, 4), out[4 * m:8 * m].reshape(m, 4)

    def layout_info(self, sys=0):
        info = (C.c_longlong * 6)()
        self._check(lib().mrfg_layout_info(self._h, sys, info))
        keys = ("cells", "faces", "pairs", "rows", "slices", "waves")
        return dict(zip(keys, [int(v) for v in info]))

    def tuning(self):
        """The autotuner's launch configuration for the batch (after a solve)."""
        t, th, ch = C.c_int(), C.c_int(), C.c_int()
        a, b = C.c_double(), C.c_double()
        self._check(lib().mrfg_tuning(self._h, C.byref(t), C.byref(th), C.byref(ch), C.byref(a), C.byref(b)))
        return dict(tuned=bool(t.value), block_threads=th.value, ell_chunk=ch.value, spmv_ms=a.value,
                    spmv_ms_default=b.value)

    def set_tuning(self, block_threads, ell_chunk):
        """Force a launch configuration for the following solves."""
        self._check(lib().mrfg_set_tuning(self._h, int(block_threads), int(ell_chunk)))

    def bytes_model(self, sys=0, value_size=8):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self._check(lib().mrfg_bytes(self._h, sys, value_size, C.byref(a), C.byref(b), C.byref(c)))
        return dict(iter=a.value, spmv=b.value, dilu=c.value)


# ---------------------------------------------------------------------------
# Serving (continuous batching over a request stream)
# ---------------------------------------------------------------------------

@dataclass
class ServeResult:
    perf: SolverPerformance
    t_submit: float
    t_start: float
    t_done: float
    batch_size: int

    @property
    def latency(self):
        return self.t_done - self.t_submit


class Server:
    """Request loop over coupled solves with continuous batching: host
    preparation threads (assembly, layout), device slots with their own
    resident batches and streams, HBM-budget admission in arrival order,
    each request retired as soon as it has converged."""

    def __init__(self, device=0, hbm_budget=64 << 30, workers=4, max_batch=8, fv_solution=None,
                 entry="Tcoupled", keep_fields=False, prep_threads=0, **controls_kw):
        if fv_solution is None:
            kw = dict(solver="BiCGStab", tolerance=1e-10, maxIter=1000)
            kw.update(controls_kw)
            fv_solution = _controls_text(**kw)
            entry = "X"
        h = C.c_void_p()
        rc = lib().mrfg_server_create(C.byref(h), device, int(hbm_budget), workers, max_batch, int(prep_threads),
                                      _text(fv_solution), _text(entry), 1 if keep_fields else 0)
        if rc != MRFG_OK:
            raise MrfgError(rc, "cannot create the server")
        self._h = h
        self._sizes = []

    def __del__(self):
        try:
            if self._h:
                lib().mrfg_server_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def submit(self, case):
        text = case if isinstance(case, str) else cases.to_dict_text(case)
        rid = C.c_int()
        rc = lib().mrfg_server_submit(self._h, _text(text), C.byref(rid))
        if rc != MRFG_OK:
            raise MrfgError(rc, "submit failed")
        if not isinstance(case, str):
            self._sizes.append([int(np.prod(r["nCells"])) for r in case["regions"]])
        else:
            self._sizes.append(None)
        return rid.value

    def wait_all(self):
        lib().mrfg_server_wait_all(self._h)

    def result(self, rid):
        perf = Perf()
        t = np.zeros(3)
        bs = C.c_int()
        rc = lib().mrfg_server_result(self._h, rid, C.byref(perf), _p(t, C.c_double), C.byref(bs))
        if rc != MRFG_OK:
            raise MrfgError(rc, lib().mrfg_server_last_error(self._h).decode())
        sp = SolverPerformance.from_c(perf)
        return ServeResult(sp, float(t[0]), float(t[1]), float(t[2]), bs.value)

    def times(self, rid):
        """dict(submit, prepared, admitted, solve, done) of a finished request
        (seconds since the server started)."""
        t = np.zeros(5)
        rc = lib().mrfg_server_times(self._h, rid, _p(t, C.c_double))
        if rc != MRFG_OK:
            raise MrfgError(rc, "request %d has not finished" % rid)
        return dict(zip(("submit", "prepared", "admitted", "solve", "done"), t.tolist()))

    def psi(self, rid):
        sizes = self._sizes[rid]
        if sizes is None:
            raise ValueError("fields of requests submitted as text: sizes unknown")
        out = []
        for r, n in enumerate(sizes):
            a = np.zeros(n)
            rc = lib().mrfg_server_get_psi(self._h, rid, r, _p(a, C.c_double))
            if rc != MRFG_OK:
                raise MrfgError(rc, "no field kept for request %d" % rid)
            out.append(a)
        return out
