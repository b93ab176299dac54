"""ctypes binding of the CPU oracle (libpgo.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, always as the checker. The product path
(peregrine_amd) never imports it.

Arrays are numpy views in the device layout: a block array of shape
(ncomp, nk_full, nj_full, ni_full), C-contiguous, i.e. Fortran order
[i, j, k, comp] as the reference's device backend lays it out.
"""

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libpgo.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Arr(C.Structure):
    _fields_ = [("d", _dp), ("sj", C.c_int64), ("sk", C.c_int64), ("sc", C.c_int64)]


class Box(C.Structure):
    _fields_ = [("start", C.c_int * 3), ("extent", C.c_int * 3)]


class Case(C.Structure):
    _fields_ = [
        ("ns", C.c_int), ("ne", C.c_int), ("eos", C.c_int), ("recon", C.c_int), ("limiter", C.c_int),
        ("primary", C.c_int), ("nt", C.c_int), ("ng_", C.c_int),
        ("MW", _dp), ("MWinv", _dp), ("Rsp", _dp), ("cppolyR", _dp), ("h298R", _dp), ("cp0", _dp),
        ("gpoly", _dp), ("h298", _dp),
        ("nr", C.c_int), ("ntb", C.c_int),
        ("kind", _ip), ("reversible", _ip), ("product", _ip), ("tb", _ip), ("troeidx", _ip),
        ("nfwd", _ip), ("fwdsp", _ip), ("nreact", _ip), ("reactsp", _ip), ("nprod", _ip), ("prodsp", _ip),
        ("nnet", _ip), ("netsp", _ip), ("ntbeff", _ip), ("tbeffsp", _ip),
        ("logA", _dp), ("b", _dp), ("EaR", _dp), ("logprA", _dp), ("prb", _dp), ("prEaR", _dp), ("nusum", _dp),
        ("fwdorder", _dp), ("reactnu", _dp), ("prodnu", _dp), ("netnu", _dp), ("netnuMW", _dp),
        ("lfcpoly", _dp), ("tbdefeff", _dp), ("tbeffdev", _dp),
        ("logprefRu", C.c_double), ("prefRu", C.c_double), ("Rupref", C.c_double), ("logcfloor", C.c_double),
        ("expclamp", C.c_double), ("troeuc", C.c_double), ("troeuhinv", C.c_double), ("cfloor", C.c_double),
        ("trans", C.c_int), ("nmu", C.c_int), ("nkappa", C.c_int),
        ("hpoly", _dp), ("mu0", _dp), ("kappa0", _dp), ("lewis", _dp), ("MWqinv", _dp), ("wilkepair", _dp),
        ("mupoly", _dp), ("kappapoly", _dp), ("Ru", C.c_double),
        ("cubic", C.c_double * 4),
        ("cubicsqrtac", _dp), ("cubicb", _dp), ("cubicda", _dp), ("cubicmTc", _dp), ("cubiconepm", _dp),
        ("cubicsqrtacMW", _dp), ("cubicbMW", _dp), ("vshift", _dp), ("vshiftMW", _dp), ("cubicdaMW", _dp),
        ("acentric", _dp),
        ("cppoly", _dp), ("chungsigma", _dp), ("chungepsk", _dp), ("chungvisc", _dp), ("chungcond", _dp),
        ("diffusion", C.c_int), ("npair", C.c_int), ("diffwidth", C.c_int), ("diffpartial", C.c_int),
        ("pairfit", _dp), ("selffit", _dp), ("bathof", _dp),
    ]


class Block(C.Structure):
    _fields_ = [
        ("ncell", C.c_int * 3), ("ng", C.c_int), ("regs", Arr * 4), ("nregs", C.c_int),
        ("state", Arr), ("iF", Arr), ("jF", Arr), ("kF", Arr), ("iS", Arr), ("jS", Arr), ("kS", Arr),
        ("detJinv", Arr), ("integrator", C.c_int), ("chemistry", C.c_int),
        ("subcap", C.c_int), ("subeqtol", C.c_double), ("substeps", C.POINTER(C.c_int32)), ("dt", C.c_double),
        ("viscous", C.c_int), ("alpha", C.c_double),
        ("trstate", Arr), ("hsp", Arr), ("grads", Arr), ("Jinv", Arr), ("xcc", Arr), ("xf", Arr * 3),
        ("bc", C.c_int * 6), ("nstate", C.c_int), ("bcwall", (C.c_double * 3) * 6), ("bcT", C.c_double * 6),
        ("sponge", C.c_int), ("spongeline", C.c_double * 6),
        ("imparts", C.c_int),
    ]


def _load(path=LIB):
    if not path.is_file():
        raise RuntimeError(f"{path} is not built; run `make -C {HERE}`")
    lib = C.CDLL(str(path))
    lib.pgo_tlog.restype = C.c_double
    lib.pgo_tlog.argtypes = [C.c_double]
    lib.pgo_texp.restype = C.c_double
    lib.pgo_texp.argtypes = [C.c_double]
    lib.pgo_elog.restype = C.c_double
    lib.pgo_elog.argtypes = [C.c_double]
    lib.pgo_dcbrt.restype = C.c_double
    lib.pgo_dcbrt.argtypes = [C.c_double]
    lib.pgo_threads.restype = C.c_int
    P = C.POINTER
    lib.pgo_fromQ.argtypes = [P(Case), Arr, Arr, Box]
    lib.pgo_fromprims.argtypes = [P(Case), Arr, Arr, Arr, Box]
    lib.pgo_flux.argtypes = [P(Case), C.c_int, Arr, Arr, Arr, Arr, Box]
    lib.pgo_applyflux.argtypes = [C.c_int, Arr, Arr, Arr, Arr, Arr, C.c_int, Box]
    lib.pgo_production.argtypes = [P(Case), C.c_double, _dp, C.c_double, _dp]
    lib.pgo_source.argtypes = [P(Case), Arr, Arr, Arr, Box]
    lib.pgo_combine.argtypes = [C.c_int, Arr, C.c_int, C.c_double, C.c_int, P(Arr), _dp, C.c_double, C.c_int, Box]
    lib.pgo_fill_periodic.argtypes = [Arr, C.c_int, P(C.c_int), C.c_int, C.c_int, C.c_int]
    lib.pgo_transport.argtypes = [P(Case), Arr, Arr, Arr, Arr, Box]
    lib.pgo_viscsponge.argtypes = [Arr, Arr, _dp, Box]
    lib.pgo_wall_mirror.argtypes = [P(Case), C.c_int, C.c_int, C.c_int, _dp, C.c_double, P(C.c_int), C.c_int,
                                    C.c_int, Arr, Arr, Arr]
    lib.pgo_wall_flux.argtypes = [P(Case), C.c_int, C.c_int, C.c_int, _dp, C.c_double, P(C.c_int), C.c_int, C.c_int,
                                  Arr, Arr, Arr, Arr, Arr, Arr, Arr]
    lib.pgo_grads.argtypes = [P(Case), Arr, Arr, Arr, Arr, Box]
    lib.pgo_diffflux.argtypes = [P(Case), C.c_int, C.c_double, Arr, Arr, Arr, Arr, Arr, Arr, Arr, Arr, Arr, Box]
    lib.pgo_cflmax.argtypes = [Arr, Arr, Arr, Arr, Arr, Arr, Box, _dp]
    lib.pgo_sources.argtypes = [P(Case), C.c_double, _dp, C.c_double, _dp, _dp]
    lib.pgo_dlog1p.restype = C.c_double
    lib.pgo_dlog1p.argtypes = [C.c_double]
    lib.pgo_substep.argtypes = [P(Case), Arr, Arr, Arr, C.c_int, P(C.c_int32), C.c_int, C.c_double, C.c_double, Box]
    lib.pgo_implicit.argtypes = [P(Case), Arr, Arr, Arr, C.c_int, C.c_int, C.c_double, Box]
    lib.pgo_implicit_cell.argtypes = [P(Case), C.c_double, _dp, C.c_double, C.c_int]
    lib.pgo_implicit_rhs.argtypes = [P(Case), C.c_double, _dp, _dp]
    lib.pgo_implicit_matrix.argtypes = [P(Case), C.c_double, _dp, C.c_double, _dp]
    lib.pgo_fd_step.restype = C.c_double
    lib.pgo_fd_step.argtypes = [C.c_double, C.c_int]
    lib.pgo_lu_nopivot.argtypes = [C.c_int, _dp]
    lib.pgo_lu_apply.argtypes = [C.c_int, _dp, _dp]
    lib.pgo_step.argtypes = [P(Case), P(Block), C.c_double]
    lib.pgo_stage.argtypes = [P(Case), P(Block), C.c_int, C.c_double]
    lib.pgo_advance.argtypes = [P(Case), P(Block), C.c_int, C.c_double]
    lib.pgo_rhs.argtypes = [P(Case), P(Block), C.c_int, C.c_int]
    return lib


lib = _load()
_libm = None


def libm():
    """The oracle built with libm's log, cbrt and log1p and plain products in place of the specified
    methods (libpgo_libm.so): the same functions, to measure the gap those methods make."""
    global _libm
    if _libm is None:
        _libm = _load(HERE / "libpgo_libm.so")
    return _libm
tlog = lib.pgo_tlog
texp = lib.pgo_texp
elog = lib.pgo_elog
dcbrt = lib.pgo_dcbrt
dlog1p = lib.pgo_dlog1p

EOS = {"cpg": 0, "tpg": 1, "real-gas": 2, "peng-robinson": 2}
TRANS = {None: -1, "constant": 0, "kinetic-theory": 1, "chung": 2}
RECON = {"constant": 0, "muscl": 1}
LIMITER = {None: 0, "koren": 1, "minmod": 2, "mc": 3, "superbee": 4, "vanleer": 5}
PRIMARY = {"kepaec": 0, "hllc": 1, "rusanov": 2}
INTEGRATOR = {"rk1": 0, "rk2": 1, "tvd-rk3": 2, "rk4": 3}


def arr(a):
    """A view of a (ncomp, nk, nj, ni) C-contiguous float64 array."""
    if a.dtype != np.float64 or not a.flags.c_contiguous or a.ndim != 4:
        raise ValueError("oracle arrays are C-contiguous float64 (ncomp, nk, nj, ni)")
    nc, nk, nj, ni = a.shape
    v = Arr(a.ctypes.data_as(_dp), ni, ni * nj, ni * nj * nk)
    # the view holds the array, so a temporary lives as long as the call it is passed to
    v._keep = a
    return v


def box(start, extent):
    return Box((C.c_int * 3)(*start), (C.c_int * 3)(*extent))


def dbls(values):
    return (C.c_double * len(values))(*values)


def ints(values):
    return (C.c_int * len(values))(*values)


class OracleCase:
    """A pgo_case built from a mixture and the flux choices; keeps its arrays alive."""

    def __init__(self, mixture, recon="constant", limiter=None, primary="kepaec"):
        k = mixture.constants()
        self.keep = []

        def d(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
            self.keep.append(a)
            return a.ctypes.data_as(_dp)

        def i(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
            self.keep.append(a)
            return a.ctypes.data_as(_ip)

        ns = mixture.ns
        c = Case()
        c.ns, c.ne = ns, 4 + ns
        c.eos = EOS[mixture.eos.name]
        c.recon, c.limiter, c.primary = RECON[recon], LIMITER[limiter], PRIMARY[primary]
        c.nt, c.ng_ = k["cppolyR"].shape[1], k["gpoly"].shape[1]
        c.MW, c.MWinv, c.Rsp = d(k["MW"]), d(k["MWinv"]), d(k["Rsp"])
        c.cppolyR, c.h298R, c.gpoly, c.h298 = d(k["cppolyR"]), d(k["h298R"]), d(k["gpoly"]), d(k["h298"])
        c.cp0 = d(k.get("cp0", np.zeros(ns)))
        c.hpoly, c.Ru = d(k["hpoly"]), k["Ru"]
        c.trans = TRANS[mixture.trans.name if mixture.trans else None]
        c.cppoly = d(k["cppoly"])
        if c.eos == 2:
            c.cubic = (C.c_double * 4)(*k["cubic"])
            for name in ("cubicsqrtac", "cubicb", "cubicda", "cubicmTc", "cubiconepm", "cubicsqrtacMW", "cubicbMW",
                         "vshift", "vshiftMW", "cubicdaMW", "acentric"):
                setattr(c, name, d(k[name]))
        if c.trans == 2:
            for name in ("chungsigma", "chungepsk", "chungvisc", "chungcond"):
                setattr(c, name, d(k[name]))
        if mixture.trans is not None:
            c.MWqinv, c.wilkepair = d(k["MWqinv"]), d(k["wilkepair"])
            if mixture.diffusion.name == "binary":
                c.diffusion, c.npair = 1, k["pairfit"].shape[-1]
                c.diffwidth, c.diffpartial = int(k["diffmeta"][0]), int(k["diffmeta"][1])
                c.pairfit, c.selffit, c.bathof = d(k["pairfit"]), d(k["selffit"]), d(k["bathof"])
            else:
                c.lewis = d(k["lewis"])
            c.mu0, c.kappa0 = d(k.get("mu0", np.zeros(ns))), d(k.get("kappa0", np.zeros(ns)))
            mup, kap = k.get("mupoly", np.zeros((ns, 1))), k.get("kappapoly", np.zeros((ns, 1)))
            c.nmu, c.nkappa = mup.shape[1], kap.shape[1]
            c.mupoly, c.kappapoly = d(mup), d(kap)
        if mixture.chemistry is not None:
            records, sets, cols, sc = mixture.chemistry.tables(mixture.species)
            c.nr, c.ntb = len(records), len(sets)
            for name in ("kind", "reversible", "product", "tb", "troeidx", "nfwd", "fwdsp", "nreact", "reactsp",
                         "nprod", "prodsp", "nnet", "netsp", "ntbeff", "tbeffsp"):
                setattr(c, name, i(cols[name]))
            for name in ("logA", "b", "EaR", "logprA", "prb", "prEaR", "nusum", "fwdorder", "reactnu", "prodnu",
                         "netnu", "netnuMW", "lfcpoly", "tbdefeff", "tbeffdev"):
                setattr(c, name, d(cols[name]))
            for name in ("logprefRu", "prefRu", "Rupref", "logcfloor", "expclamp", "troeuc", "troeuhinv", "cfloor"):
                setattr(c, name, sc[name])
        self.c = c
        self.ns, self.ne = ns, 4 + ns

    @property
    def ref(self):
        return C.byref(self.c)

    def implicit_rhs(self, rho, y):
        """f of the 0D reactor at constant density: (w_n / rho, dT/dt) at y = (Y, T)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(self.ns + 1)
        lib.pgo_implicit_rhs(self.ref, rho, y.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        return out

    def implicit_matrix(self, rho, y, h):
        """I - h J at y, J by the oracle's forward differences."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        A = np.zeros((self.ns + 1, self.ns + 1))
        lib.pgo_implicit_matrix(self.ref, rho, y.ctypes.data_as(_dp), h, A.ctypes.data_as(_dp))
        return A

    def implicit_cell(self, rho, y, dt, parts=1):
        """y advanced over dt by the FP64 linearly implicit form."""
        y = np.array(y, dtype=np.float64)
        lib.pgo_implicit_cell(self.ref, rho, y.ctypes.data_as(_dp), dt, parts)
        return y

    def production(self, rho, Y, T):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        out = np.zeros(self.ns)
        lib.pgo_production(self.ref, rho, Y.ctypes.data_as(_dp), T, out.ctypes.data_as(_dp))
        return out


def threads():
    return lib.pgo_threads()


def set_threads(n):
    os.environ["OMP_NUM_THREADS"] = str(n)
