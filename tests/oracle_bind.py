"""ctypes binding to the CPU oracle (oracle/_build/libmrf_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, as the checker.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "libmrf_oracle.so")

SIDES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
PTYPE = {"zeroGradient": 0, "fixedValue": 1, "empty": 2, "monolithicTemperature": 3}
FLUX = {"wall": 0, "inlet": 1, "outlet": 2}


class RegionSpec(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3), ("origin", C.c_double * 3), ("length", C.c_double * 3),
        ("transport", C.c_int), ("rhoCp", C.c_double),
        ("k_uniform", C.c_double), ("k_block", C.c_int), ("k_min", C.c_double),
        ("k_max", C.c_double), ("k_seed", C.c_uint64),
        ("flow_axis", C.c_int), ("profile_mask", C.c_int), ("Umax", C.c_double),
        ("noise", C.c_double), ("vel_seed", C.c_uint64),
        ("patch_type", C.c_int * 6), ("patch_value", C.c_double * 6),
        ("patch_flux", C.c_int * 6), ("T0", C.c_double),
    ]


class IfaceSpec(C.Structure):
    _fields_ = [("regionA", C.c_int), ("sideA", C.c_int), ("regionB", C.c_int), ("sideB", C.c_int)]


class Ldu(C.Structure):
    _fields_ = [
        ("nCells", C.c_int), ("nFaces", C.c_int),
        ("lAddr", C.POINTER(C.c_int32)), ("uAddr", C.POINTER(C.c_int32)),
        ("diag", C.POINTER(C.c_double)), ("lower", C.POINTER(C.c_double)),
        ("upper", C.POINTER(C.c_double)), ("source", C.POINTER(C.c_double)),
        ("psi", C.POINTER(C.c_double)), ("symmetric", C.c_int),
        ("k", C.POINTER(C.c_double)), ("cellSize", C.POINTER(C.c_double)),
    ]


class Iface(C.Structure):
    _fields_ = [
        ("regionA", C.c_int), ("regionB", C.c_int), ("nA", C.c_int), ("nB", C.c_int),
        ("fcA", C.POINTER(C.c_int32)), ("fcB", C.POINTER(C.c_int32)),
        ("coeffA", C.POINTER(C.c_double)), ("coeffB", C.POINTER(C.c_double)),
        ("offA", C.POINTER(C.c_int32)), ("idxA", C.POINTER(C.c_int32)),
        ("wA", C.POINTER(C.c_double)),
        ("offB", C.POINTER(C.c_int32)), ("idxB", C.POINTER(C.c_int32)),
        ("wB", C.POINTER(C.c_double)),
        ("magSfA", C.POINTER(C.c_double)), ("magSfB", C.POINTER(C.c_double)),
    ]


class System(C.Structure):
    _fields_ = [("nRegions", C.c_int), ("nInterfaces", C.c_int),
                ("regions", C.POINTER(Ldu)), ("interfaces", C.POINTER(Iface))]


class Perf(C.Structure):
    _fields_ = [("initialResidual", C.c_double), ("finalResidual", C.c_double),
                ("nIterations", C.c_int), ("converged", C.c_int), ("singular", C.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_build.argtypes = [C.c_int, C.POINTER(RegionSpec), C.c_int, C.POINTER(IfaceSpec),
                                C.POINTER(System)]
        L.orc_build.restype = C.c_int
        L.orc_free.argtypes = [C.POINTER(System)]
        L.orc_solve.argtypes = [C.POINTER(System), C.c_int, C.c_double, C.c_double, C.c_int,
                                C.c_int, C.POINTER(Perf)]
        PP = C.POINTER(C.POINTER(C.c_double))
        L.orc_amul.argtypes = [C.POINTER(System), PP, PP]
        L.orc_sumA.argtypes = [C.POINTER(System), PP]
        L.orc_norm_factor.argtypes = [C.POINTER(System), PP, PP, PP, PP]
        L.orc_norm_factor.restype = C.c_double
        L.orc_dilu_factor.argtypes = [C.POINTER(System), PP]
        L.orc_dilu_precondition.argtypes = [C.POINTER(System), PP, PP, PP]
        L.orc_interface_flux.argtypes = [C.POINTER(System), C.c_int, PP] + [C.POINTER(C.c_double)] * 4
        L.orc_interface_patch_values.argtypes = [C.POINTER(System), C.c_int, C.c_int, PP,
                                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_solve_mt.argtypes = [C.POINTER(System), C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(Perf)]
        L.orc_threads.restype = C.c_int
        L.orc_u01.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_u01.restype = C.c_double
        _lib = L
    return _lib


def _spec(r):
    s = RegionSpec()
    for a in range(3):
        s.n[a] = r["nCells"][a]
        s.origin[a] = r["origin"][a]
        s.length[a] = r["length"][a]
    s.transport = 1 if r["type"] == "transportTemperature" else 0
    s.rhoCp = r["rhoCp"]
    k = r["k"]
    if "uniform" in k:
        s.k_uniform = float(k["uniform"])
        s.k_block = 0
    else:
        s.k_block = int(k["block"])
        s.k_min, s.k_max, s.k_seed = float(k["min"]), float(k["max"]), int(k["seed"])
    v = r["velocity"]
    if v:
        s.flow_axis = int(v["axis"])
        s.profile_mask = sum(1 << a for a in v.get("profile", ()))
        s.Umax = float(v["Umax"])
        s.noise = float(v.get("noise", 0.0))
        s.vel_seed = int(v.get("seed", 0))
    for i, side in enumerate(SIDES):
        p = r["patches"][side]
        s.patch_type[i] = PTYPE[p["type"]]
        s.patch_value[i] = float(p.get("value", 0.0) or 0.0)
        s.patch_flux[i] = FLUX[p.get("flux", "wall")]
    s.T0 = r["T0"]
    return s


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class OracleSystem:
    """An assembled coupled system held by the oracle library."""

    def __init__(self, case):
        L = lib()
        names = [r["name"] for r in case["regions"]]
        specs = (RegionSpec * len(names))(*[_spec(r) for r in case["regions"]])
        ifs = (IfaceSpec * max(1, len(case["interfaces"])))()
        for i, it in enumerate(case["interfaces"]):
            ifs[i].regionA = names.index(it["A"][0])
            ifs[i].sideA = SIDES.index(it["A"][1])
            ifs[i].regionB = names.index(it["B"][0])
            ifs[i].sideB = SIDES.index(it["B"][1])
        self.sys = System()
        rc = L.orc_build(len(names), specs, len(case["interfaces"]), ifs, C.byref(self.sys))
        if rc != 0:
            raise RuntimeError("orc_build failed (%d)" % rc)
        self.names = names
        self._load()

    def _load(self):
        s = self.sys
        self.regions = []
        for r in range(s.nRegions):
            R = s.regions[r]
            n, f = R.nCells, R.nFaces
            self.regions.append(dict(
                nCells=n, nFaces=f, lAddr=_arr(R.lAddr, f, np.int32), uAddr=_arr(R.uAddr, f, np.int32),
                diag=_arr(R.diag, n, np.float64), lower=_arr(R.lower, f, np.float64),
                upper=_arr(R.upper, f, np.float64), source=_arr(R.source, n, np.float64),
                psi=_arr(R.psi, n, np.float64), symmetric=bool(R.symmetric),
                k=_arr(R.k, n, np.float64), cellSize=_arr(R.cellSize, 3, np.float64)))
        self.interfaces = []
        for i in range(s.nInterfaces):
            I = s.interfaces[i]
            nA, nB = I.nA, I.nB
            offA = _arr(I.offA, nA + 1, np.int32)
            offB = _arr(I.offB, nB + 1, np.int32)
            self.interfaces.append(dict(
                regionA=I.regionA, regionB=I.regionB, nA=nA, nB=nB,
                fcA=_arr(I.fcA, nA, np.int32), fcB=_arr(I.fcB, nB, np.int32),
                coeffA=_arr(I.coeffA, nA, np.float64), coeffB=_arr(I.coeffB, nB, np.float64),
                offA=offA, idxA=_arr(I.idxA, int(offA[-1]), np.int32), wA=_arr(I.wA, int(offA[-1]), np.float64),
                offB=offB, idxB=_arr(I.idxB, int(offB[-1]), np.int32), wB=_arr(I.wB, int(offB[-1]), np.float64),
                magSfA=_arr(I.magSfA, nA, np.float64), magSfB=_arr(I.magSfB, nB, np.float64)))

    @classmethod
    def from_arrays(cls, regions, interfaces):
        """An oracle system over already assembled arrays (dicts with the
        fields of Case.region / Case.interface), e.g. a case read from disk.
        The arrays stay owned by Python."""
        self = cls.__new__(cls)
        self._keep = []

        def ptr(a, ctype, dtype):
            a = np.ascontiguousarray(a, dtype=dtype).copy()
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        s = System()
        s.nRegions, s.nInterfaces = len(regions), len(interfaces)
        L = (Ldu * max(1, len(regions)))()
        for r, R in enumerate(regions):
            L[r].nCells, L[r].nFaces = R["nCells"], R["nFaces"]
            L[r].lAddr = ptr(R["lAddr"], C.c_int32, np.int32)
            L[r].uAddr = ptr(R["uAddr"], C.c_int32, np.int32)
            L[r].diag = ptr(R["diag"], C.c_double, np.float64)
            L[r].lower = ptr(R["lower"], C.c_double, np.float64)
            L[r].upper = ptr(R["upper"], C.c_double, np.float64)
            L[r].source = ptr(R["source"], C.c_double, np.float64)
            L[r].psi = ptr(R["psi"], C.c_double, np.float64)
            L[r].symmetric = int(bool(R["symmetric"]))
            L[r].k = ptr(np.ones(R["nCells"]), C.c_double, np.float64)
            L[r].cellSize = ptr(np.ones(3), C.c_double, np.float64)
        I = (Iface * max(1, len(interfaces)))()
        for i, It in enumerate(interfaces):
            I[i].regionA, I[i].regionB, I[i].nA, I[i].nB = It["regionA"], It["regionB"], It["nA"], It["nB"]
            for side in "AB":
                setattr(I[i], "fc" + side, ptr(It["fc" + side], C.c_int32, np.int32))
                setattr(I[i], "coeff" + side, ptr(It["coeff" + side], C.c_double, np.float64))
                setattr(I[i], "off" + side, ptr(It["off" + side], C.c_int32, np.int32))
                setattr(I[i], "idx" + side, ptr(It["idx" + side], C.c_int32, np.int32))
                setattr(I[i], "w" + side, ptr(It["w" + side], C.c_double, np.float64))
                setattr(I[i], "magSf" + side, ptr(np.ones(It["n" + side]), C.c_double, np.float64))
        self._keep += [L, I]
        s.regions, s.interfaces = L, I
        self.sys = s
        self.names = [R.get("name", "region%d" % r) for r, R in enumerate(regions)]
        self._load()
        return self

    def __del__(self):
        if getattr(self, "_keep", None) is not None:
            return   # arrays owned by Python
        try:
            lib().orc_free(C.byref(self.sys))
        except Exception:
            pass

    # ---- vector helpers: a coupled vector is a list of per-region arrays ----
    def _pp(self, vecs):
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in vecs]
        ptrs = (C.POINTER(C.c_double) * len(arrs))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
        return arrs, ptrs

    def zeros(self):
        return [np.zeros(r["nCells"]) for r in self.regions]

    def amul(self, x):
        xa, xp = self._pp(x)
        ya, yp = self._pp(self.zeros())
        lib().orc_amul(C.byref(self.sys), xp, yp)
        return ya

    def sumA(self):
        sa, sp = self._pp(self.zeros())
        lib().orc_sumA(C.byref(self.sys), sp)
        return sa

    def norm_factor(self, x, b, Ax):
        xa, xp = self._pp(x)
        ba, bp = self._pp(b)
        aa, ap = self._pp(Ax)
        ta, tp = self._pp(self.zeros())
        return lib().orc_norm_factor(C.byref(self.sys), xp, bp, ap, tp)

    def dilu_factor(self):
        ra, rp = self._pp(self.zeros())
        lib().orc_dilu_factor(C.byref(self.sys), rp)
        return ra

    def dilu_precondition(self, rD, r):
        da, dp = self._pp(rD)
        ra, rp = self._pp(r)
        wa, wp = self._pp(self.zeros())
        lib().orc_dilu_precondition(C.byref(self.sys), dp, wp, rp)
        return wa

    def interface_flux(self, i, psi):
        pa, pp = self._pp(psi)
        out = [C.c_double() for _ in range(4)]
        lib().orc_interface_flux(C.byref(self.sys), i, pp, *[C.byref(o) for o in out])
        return tuple(o.value for o in out)

    def interface_patch_values(self, i, side, psi):
        """(tn, tw) of side 0 (A) / 1 (B) of interface i: the GGI neighbour
        field and the monolithicTemperature wall value per face."""
        pa, pp = self._pp(psi)
        n = self.interfaces[i]["nB" if side else "nA"]
        tn, tw = np.zeros(n), np.zeros(n)
        lib().orc_interface_patch_values(C.byref(self.sys), i, side, pp, tn.ctypes.data_as(C.POINTER(C.c_double)),
                                         tw.ctypes.data_as(C.POINTER(C.c_double)))
        return tn, tw

    def solve(self, solver="BiCGStab", tol=1e-10, relTol=0.0, minIter=0, maxIter=1000, psi0=None):
        """Solve in place from psi0 (default: the case's T0); returns
        (psi list, perf dict)."""
        s = self.sys
        for r in range(s.nRegions):
            src = psi0[r] if psi0 is not None else self.regions[r]["psi"]
            dst = np.ctypeslib.as_array(s.regions[r].psi, shape=(s.regions[r].nCells,))
            dst[:] = src
        perf = Perf()
        kind = 1 if solver in ("PCG", "CG", "hipPCG") else 0
        lib().orc_solve(C.byref(s), kind, tol, relTol, minIter, maxIter, C.byref(perf))
        psi = [_arr(s.regions[r].psi, s.regions[r].nCells, np.float64) for r in range(s.nRegions)]
        return psi, dict(initialResidual=perf.initialResidual, finalResidual=perf.finalResidual,
                         nIterations=perf.nIterations, converged=bool(perf.converged),
                         singular=perf.singular == 1, diverged=perf.singular == 2)

    def set_values(self, r, diag=None, source=None):
        """Overwrite region r's diagonal and / or source (in the arrays the
        oracle solves with, and in self.regions)."""
        R = self.sys.regions[r]
        n = R.nCells
        if diag is not None:
            np.ctypeslib.as_array(R.diag, shape=(n,))[:] = diag
            self.regions[r]["diag"] = np.array(diag, dtype=np.float64)
        if source is not None:
            np.ctypeslib.as_array(R.source, shape=(n,))[:] = source
            self.regions[r]["source"] = np.array(source, dtype=np.float64)

    def time_steps(self, ddt, n_steps, scheme="backward", psi0=None, **solve_kw):
        """The time loop for a fixed flow: per step the ddt terms of the
        transient regions (ddt[r] = dict(ddtCoeff, diagBase, sourceBase), or
        None for a steady region) are rebuilt from the current and previous
        fields -- diag = diagBase + c*a with c = 1 (Euler) or 3/2 (backward),
        source = sourceBase + a*T (Euler) or a*(2 T - T_prev / 2) (backward,
        constant time step; the first step is Euler, as backwardDdtScheme's)
        -- and the system is solved from the current field.  Returns the
        (psi, perf) of every step."""
        x = [np.array(p, dtype=np.float64) for p in (psi0 or [R["psi"] for R in self.regions])]
        prev, out = None, []
        for k in range(n_steps):
            bwd = scheme == "backward" and k >= 1
            for r, D in enumerate(ddt):
                if D is None:
                    continue
                a = D["ddtCoeff"]
                t = 2.0 * x[r] - 0.5 * prev[r] if bwd else x[r]
                self.set_values(r, diag=D["diagBase"] + (1.5 if bwd else 1.0) * a, source=D["sourceBase"] + a * t)
            psi, perf = self.solve(psi0=x, **solve_kw)
            prev, x = x, psi
            out.append((psi, perf))
        return out

    def solve_mt(self, tol=1e-10, relTol=0.0, minIter=0, maxIter=1000, psi0=None):
        """BiCGStab on all host cores (orc_solve_mt); returns (psi list,
        perf dict, threads)."""
        s = self.sys
        for r in range(s.nRegions):
            src = psi0[r] if psi0 is not None else self.regions[r]["psi"]
            np.ctypeslib.as_array(s.regions[r].psi, shape=(s.regions[r].nCells,))[:] = src
        perf = Perf()
        lib().orc_solve_mt(C.byref(s), tol, relTol, minIter, maxIter, C.byref(perf))
        psi = [_arr(s.regions[r].psi, s.regions[r].nCells, np.float64) for r in range(s.nRegions)]
        return psi, dict(initialResidual=perf.initialResidual, finalResidual=perf.finalResidual,
                         nIterations=perf.nIterations, converged=bool(perf.converged),
                         singular=perf.singular == 1, diverged=perf.singular == 2), lib().orc_threads()

    # ---- global (concatenated) sparse matrix, for brute-force checks ----
    def offsets(self):
        off = [0]
        for r in self.regions:
            off.append(off[-1] + r["nCells"])
        return off

    def to_scipy(self):
        import scipy.sparse as sp
        off = self.offsets()
        rows, cols, vals = [], [], []
        for g, R in enumerate(self.regions):
            o = off[g]
            n = R["nCells"]
            rows.append(np.arange(n) + o); cols.append(np.arange(n) + o); vals.append(R["diag"])
            rows.append(R["uAddr"] + o); cols.append(R["lAddr"] + o); vals.append(R["lower"])
            rows.append(R["lAddr"] + o); cols.append(R["uAddr"] + o); vals.append(R["upper"])
        for I in self.interfaces:
            for side in (0, 1):
                n = I["nB"] if side else I["nA"]
                off_ = I["offB"] if side else I["offA"]
                idx = I["idxB"] if side else I["idxA"]
                w = I["wB"] if side else I["wA"]
                fc = I["fcB"] if side else I["fcA"]
                fcN = I["fcA"] if side else I["fcB"]
                coeff = I["coeffB"] if side else I["coeffA"]
                rO = off[I["regionB"] if side else I["regionA"]]
                rN = off[I["regionA"] if side else I["regionB"]]
                cnt = np.diff(off_)
                face = np.repeat(np.arange(n), cnt)
                rows.append(fc[face] + rO)
                cols.append(fcN[idx] + rN)
                vals.append(-coeff[face] * w)
        N = off[-1]
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))

    def source_vector(self):
        return np.concatenate([r["source"] for r in self.regions])

    def split(self, v):
        off = self.offsets()
        return [np.array(v[off[g]:off[g + 1]]) for g in range(len(self.regions))]


# ---------------------------------------------------------------------------
# Partitioned (Dirichlet-Neumann) coupling: the DNA loop restated
# ---------------------------------------------------------------------------

SMALL = 1e-15   # foam-extend's doubleScalarSMALL


def _fsum(v):
    return math.fsum(np.asarray(v, dtype=np.float64).tolist())


def _gather(side, vals):
    """interpolateFromNbrField through the GGI weights: out_i = sum_p w_p vals[idx_p]
    (sequential per face)."""
    out = np.zeros(len(side["fc"]))
    for i in range(len(out)):
        acc = 0.0
        for p in range(side["off"][i], side["off"][i + 1]):
            acc += side["w"][p] * vals[side["idx"][p]]
        out[i] = acc
    return out


def _iqnils_update(V, W, res, cur):
    """IQNILS<scalar>::updateIQNILS (IQNILSTemplates.C:37-125): QR of the
    columns of V (newest first) by modified Gram-Schmidt, Q^T (-res), the
    triangular solve with columns below 1e-10 * max column sum dropped, and
    cur + sum_i W_i c_(cols-1-i)."""
    cols = len(V)
    R = [[0.0] * cols for _ in range(cols)]
    C = [0.0] * cols
    Q = [np.array(V[cols - 1 - i], dtype=np.float64) for i in range(cols)]
    for i in range(cols):
        R[i][i] = math.sqrt(_fsum(Q[i] * Q[i]))
        Q[i] = Q[i] / R[i][i]
        for j in range(i + 1, cols):
            R[i][j] = _fsum(Q[i] * Q[j])
            Q[j] = Q[j] - R[i][j] * Q[i]
        C[i] = _fsum(Q[i] * (-res))
    colsum = []
    for j in range(cols):
        t = 0.0
        for i in range(j + 1):
            t += abs(R[i][j])
        colsum.append(t)
    eps = 1.0e-10 * max(colsum)
    for i in range(cols):
        if abs(R[i][i]) > eps:
            for j in range(i + 1, cols):
                R[i][j] = R[i][j] / R[i][i]
            C[i] = C[i] / R[i][i]
            R[i][i] = 1.0
    for j in range(cols - 1, -1, -1):
        if abs(R[j][j]) > eps:
            for i in range(j):
                C[i] = C[i] - C[j] * R[i][j]
        else:
            C[j] = 0.0
    out = np.array(cur, dtype=np.float64)
    for i in range(cols):
        out = out + W[i] * C[cols - 1 - i]
    return out


class _Acceleration:
    """accelerationModel / fixedRelaxation / aitkenRelaxation / IQNILS
    (acceleration/accelerationModel/accelerationModel.C:72-86,
    fixedRelaxation.C:76-106, aitkenRelaxation.C:83-110,
    aitkenRelaxationTemplates.C:111-168, IQNILS.C:70-170): prevFld starts at
    the patch's initial values (initialize(*this) of both patch types); the
    Aitken factor keeps the initial relaxation for the first two corrections
    of a time step; IQN-ILS stores the second correction's field and residual
    as references, appends residual and field differences from the third
    on, and takes the quasi-Newton step once it holds two columns (a steady
    run is one time step, so couplingReuse drops nothing)."""

    def __init__(self, kind, relax, init):
        self.kind, self.init = kind, float(relax)
        self.omega = float(relax)
        self.prev = np.array(init, dtype=np.float64)
        self.res = None
        self.prev_res = None
        self.corr = 0
        self.V, self.W = [], []

    def relax(self, cur):
        if self.corr == 0:
            self.corr = 1
        self.prev_res = self.res
        self.res = cur - self.prev
        if self.kind == "IQN-ILS":
            if self.corr == 2:
                self.fld_ref, self.res_ref = np.array(cur), np.array(self.res)
            elif self.corr > 2:
                self.V.append(self.res - self.res_ref)
                self.W.append(cur - self.fld_ref)
            if len(self.V) > 1:
                out = _iqnils_update(self.V, self.W, self.res, cur)
            else:
                out = self.prev + self.init * self.res
            self.prev = out
            self.corr += 1
            return out
        if self.kind == "aitken":
            if self.corr < 3:
                self.omega = self.init if self.omega < self.init else self.omega
            else:
                d = self.res - self.prev_res
                w = -self.omega * (_fsum(self.prev_res * d) / (_fsum(d * d) + SMALL))
                w = abs(w)
                if w > 1.0:
                    w = 1.0
                elif w < SMALL:
                    w = self.init
                self.omega = w
            out = self.prev + self.omega * self.res
        else:
            out = self.prev + self.init * self.res
        self.prev = out
        self.corr += 1
        return out


def _norm_res(r, a, b):
    n = max(min(math.sqrt(_fsum(a * a)), math.sqrt(_fsum(b * b))), SMALL)
    return math.sqrt(_fsum(r * r)) / n


def dna_solve(case_regions, part, **solve_kw):
    """The partitioned coupling loop of multiRegionSystem::solve
    (multiRegionSystem.C:571-610) under dnaControl::loop (dnaControl.C:993-1061)
    for heatTransferInterface pairs, steady: per DNA iteration every region in
    order updates its partitioned patches -- the Dirichlet side
    (genericRegionCoupledJumpFvPatchField::updateCoeffs, :165-217) takes
    K * (the other side's patch values), relaxed; the Neumann side
    (genericRegionCoupledFluxFvPatchField::updateCoeffs, :173-239) takes
    gradient = -(the other side's flux k snGrad T) / k, relaxed -- and is
    solved from its current field; a Neumann patch then evaluates to
    T_P + g / deltaCoeffs (genericRegionCoupledFluxFvPatchFields.C:55-136).
    After the first iteration, the loop stops when every pair's jump residual
    (genericRegionCoupledJumpFvPatchField::normResidual, :316-366) and flux
    residual (genericRegionCoupledFluxFvPatchField::normResidual) are within
    the limits, or after maxCoupleIter iterations.  Region sources are
    sourceBase + coeff * value per partitioned face, in pair order, Dirichlet
    side first.  Returns dict(psi, iterations, converged, history, solves)."""
    return dna_time_steps(case_regions, part, None, 1, **solve_kw)[0]


def dna_time_steps(case_regions, part, ddt, n_steps, scheme="backward", **solve_kw):
    """n_steps time steps, each a DNA loop as in dna_solve, for regions with
    ddt terms (ddt[r] = dict(ddtCoeff, diagBase, sourceBase), None for a
    steady region; ddt None: one steady loop).  Per step the ddt terms are
    rebuilt from the current and previous fields as in OracleSystem.time_steps
    (Euler, or backward with a first Euler step) and the region sources become
    sourceBase + a * t, then the partitioned terms.  The acceleration models
    start a new time step (accelerationModel corr_ reset on a new time value;
    IQN-ILS with couplingReuse 0 drops its columns) and keep their previous
    field and Aitken factor.  Returns one dna_solve result per step."""
    nreg = len(case_regions)
    systems = [OracleSystem.from_arrays([R], []) for R in case_regions]
    x = [np.array(R["psi"], dtype=np.float64) for R in case_regions]
    pairs = part["pairs"]
    st = []
    for P in pairs:
        D, N = P["dir"], P["neu"]
        st.append(dict(TD=np.array(D["value0"]), g=np.zeros(len(N["fc"])), TN=np.array(N["value0"]),
                       accD=_Acceleration(D["accType"], D["relax"], D["value0"]),
                       accN=_Acceleration(N["accType"], N["relax"], N["value0"])))

    def flux_dir(q):   # k snGrad T of the Dirichlet side
        D = pairs[q]["dir"]
        return ((st[q]["TD"] - x[D["region"]][D["fc"]]) * D["dc"]) * D["kb"]

    def residuals():
        jr = fr = 0.0
        for q, P in enumerate(pairs):
            D, N = P["dir"], P["neu"]
            nbr = _gather(D, st[q]["TN"])
            jump = nbr * (P["K"] - 1.0)
            jr = max(jr, _norm_res((st[q]["TD"] - nbr) - jump, st[q]["TD"], nbr + jump))
            fown = st[q]["g"] * N["kb"]
            fnbr = _gather(N, flux_dir(q))
            fr = max(fr, _norm_res((fown + fnbr) - 0.0, fown, -1.0 * fnbr + 0.0))
        return jr, fr

    def loop(base):
        control = part["maxJumpRes"] >= 0.0
        history, solves, corr, converged = [], [], 0, False
        res = (None, None)
        while True:
            corr += 1
            if corr > 1 and control and res[0] <= part["maxJumpRes"] and res[1] <= part["maxFluxRes"]:
                converged = True
                break
            if corr > part["maxCoupleIter"]:
                break
            for r in range(nreg):
                for q, P in enumerate(pairs):
                    if P["dir"]["region"] == r:
                        nbr = _gather(P["dir"], st[q]["TN"])
                        cur = nbr + nbr * (P["K"] - 1.0)
                        st[q]["TD"] = st[q]["accD"].relax(cur)
                    if P["neu"]["region"] == r:
                        fl = _gather(P["neu"], flux_dir(q))
                        fl = fl * -1.0
                        fl = fl + 0.0
                        st[q]["g"] = st[q]["accN"].relax(fl / P["neu"]["kb"])
                src = np.array(base[r], dtype=np.float64)
                for q, P in enumerate(pairs):
                    for key, vals in (("dir", st[q]["TD"]), ("neu", st[q]["g"])):
                        S = P[key]
                        if S["region"] != r:
                            continue
                        for i, c in enumerate(S["fc"]):
                            src[c] = src[c] + S["coeff"][i] * vals[i]
                systems[r].set_values(0, source=src)
                psi, perf = systems[r].solve(psi0=[x[r]], **solve_kw)
                x[r] = psi[0]
                solves.append(perf)
                for q, P in enumerate(pairs):
                    N = P["neu"]
                    if N["region"] == r:
                        st[q]["TN"] = x[r][N["fc"]] + st[q]["g"] / N["dc"]
            res = residuals()
            history.append(dict(jump=res[0], flux=res[1], omega=[s_["accD"].omega for s_ in st]))
        return dict(psi=[a.copy() for a in x], iterations=corr - 1, converged=converged, history=history,
                    solves=solves, boundary=[dict(TD=s_["TD"], g=s_["g"], TN=s_["TN"]) for s_ in st])

    out = []
    prev = None
    for k in range(n_steps):
        base = [np.array(b, dtype=np.float64) for b in part["sourceBase"]]
        if ddt is not None:
            bwd = scheme == "backward" and k >= 1
            for r, D in enumerate(ddt):
                if D is None:
                    continue
                a = D["ddtCoeff"]
                t = 2.0 * x[r] - 0.5 * prev[r] if bwd else x[r]
                systems[r].set_values(0, diag=D["diagBase"] + (1.5 if bwd else 1.0) * a)
                base[r] = D["sourceBase"] + a * t
            prev = [a_.copy() for a_ in x]
            for s_ in st:
                for acc in (s_["accD"], s_["accN"]):
                    acc.corr = 0
                    acc.V, acc.W = [], []
        out.append(loop(base))
    return out
