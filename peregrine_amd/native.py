"""ctypes binding of libpga.so, the HIP library (include/peregrine_amd.h).

Loading needs only the library file and the HIP runtime library; calling an
entry point that launches needs a GPU. There is no fallback: every compute call
goes to the HIP kernels, and a failure raises.
"""

import ctypes as C
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIBPATH = HERE / "libpga.so"

_i32 = C.c_int32
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class View(C.Structure):
    _fields_ = [("data", C.c_void_p), ("sj", C.c_int64), ("sk", C.c_int64), ("sc", C.c_int64), ("n", _i32 * 4)]


def view(ptr, ni, nj, nk, nc, strides=None):
    """The ABI view of an array with extents (ni, nj, nk, nc), i fastest; strides (sj, sk, sc) in
    elements, contiguous when not given."""
    sj, sk, sc = strides if strides is not None else (ni, ni * nj, ni * nj * nk)
    return View(ptr, sj, sk, sc, (_i32 * 4)(ni, nj, nk, nc))


class Box(C.Structure):
    _fields_ = [("start", _i32 * 3), ("extent", _i32 * 3)]


class Wall(C.Structure):
    _fields_ = [("kind", _i32), ("reserved", _i32), ("u", C.c_double * 3), ("T", C.c_double)]


WALL_KINDS = ("slp-adia-wall", "slp-isot-wall", "no-slp-adia-wall", "no-slp-isot-wall", "mov-adia-wall",
              "mov-isot-wall")


class SubstepOpts(C.Structure):
    _fields_ = [("max_substeps", _i32), ("bulk_substeps", _i32), ("stiff_bins", _i32), ("reserved", _i32),
                ("equilibrium_tolerance", C.c_double)]


class ImplicitOpts(C.Structure):
    _fields_ = [("parts", _i32), ("max_sweeps", _i32), ("tolerance", C.c_double)]


class BatchBlock(C.Structure):
    _fields_ = [
        ("Q", View * 4), ("state", View), ("iS", View), ("jS", View), ("kS", View), ("iF", View), ("jF", View),
        ("kF", View), ("detJinv", View), ("rates", View), ("interior", Box), ("nslab", _i32), ("slab_src", _i32 * 6),
        ("slab_dst", (_i32 * 3) * 6), ("slab_from", (_i32 * 3) * 6), ("slab_extent", (_i32 * 3) * 6),
        ("slab_orient", (_i32 * 3) * 6), ("stage_wi", _i32), ("stage_kc", _i32), ("flux_ki", _i32),
        ("reserved", _i32),
    ]


class CaseDesc(C.Structure):
    _fields_ = [
        ("ns", _i32), ("eos", _i32), ("recon", _i32), ("limiter", _i32), ("primary", _i32), ("mechanism", _i32),
        ("trans", _i32), ("diffusion", _i32), ("reserved", _i32 * 8),
    ]


PGA_EINVAL, PGA_EUNSUP, PGA_ELAUNCH, PGA_ERANGE = -1, -2, -3, -4
ERRORS = {-1: "invalid argument", -2: "unsupported case", -3: "HIP launch failure", -4: "size out of range"}

# name -> (restype, argtypes); every launching entry point takes the stream last
SIGNATURES = {
    "pga_abi_version": (_i32, []),
    "pga_last_error": (C.c_char_p, []),
    "pga_const_offset": (_i32, [_i32, C.c_char_p, _ip]),
    "pga_const_size": (_i32, [_i32]),
    "pga_mechanism_count": (_i32, []),
    "pga_mechanism_signature": (C.c_char_p, [_i32]),
    "pga_mechanism_name": (C.c_char_p, [_i32]),
    "pga_mechanism_table_size": (_i32, [_i32]),
    "pga_mechanism_find": (_i32, [C.c_char_p]),
    "pga_case_create": (_i32, [C.POINTER(CaseDesc), C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "pga_case_destroy": (_i32, [C.c_void_p]),
    "pga_fromQ": (_i32, [C.c_void_p, View, View, C.POINTER(Box), _i32, C.c_void_p]),
    "pga_fromprims": (_i32, [C.c_void_p, View, View, View, Box, C.c_void_p]),
    "pga_flux": (_i32, [C.c_void_p, _i32, View, View, View, View, Box, C.c_void_p]),
    "pga_faces": (_i32, [C.c_void_p, View, View, C.POINTER(View), C.POINTER(View), C.POINTER(Box), _i32, C.c_double,
                         View, View, View, View, C.POINTER(View), C.c_void_p]),
    "pga_applyflux": (_i32, [_i32, View, View, View, View, View, _i32, Box, C.c_void_p]),
    "pga_source": (_i32, [C.c_void_p, View, View, View, Box, C.c_void_p]),
    "pga_production": (_i32, [C.c_void_p, _i32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pga_combine": (_i32, [_i32, View, _i32, C.c_double, _i32, C.POINTER(View), _dp, C.c_void_p, _i32, Box,
                           C.c_void_p]),
    "pga_fill_periodic": (_i32, [View, _i32, _ip, _i32, _i32, _i32, C.c_void_p]),
    "pga_copy_box": (_i32, [View, _ip, View, _ip, _ip, _i32, _ip, _dp, _i32, _i32, C.c_void_p]),
    "pga_copy_slabs": (_i32, [_i32, C.POINTER(View), C.POINTER(View), _ip, _ip, _ip, _i32, _ip, _dp, _i32, _i32,
                              C.c_void_p]),
    "pga_stage_supported": (_i32, [C.c_void_p]),
    "pga_stage": (_i32, [C.c_void_p, View, View, View, View, View, View, View, View, View, View, _i32, _ip, _dp, _ip,
                         C.c_void_p, Box, _i32, C.c_void_p]),
    "pga_stage2": (_i32, [C.c_void_p, View, View, View, View, View, View, View, View, View, View, View, _i32, _ip, _dp,
                          _ip, _i32, _ip, _dp, _ip, C.c_void_p, Box, _i32, C.c_void_p]),
    "pga_stage_march": (_i32, [C.c_void_p, View, View, View, View, View, View, View, View, View, View, _i32, _ip, _dp,
                               _ip, C.c_void_p, Box, _i32, C.c_void_p]),
    "pga_rates": (_i32, [C.c_void_p, View, View, View, Box, C.c_void_p]),
    "pga_halo": (_i32, [C.c_void_p, _i32, C.POINTER(View), C.POINTER(View), _ip, _ip, _ip, _ip, _dp, View, View,
                        C.c_void_p]),
    "pga_transport": (_i32, [C.c_void_p, View, View, View, View, C.POINTER(Box), _i32, C.c_void_p]),
    "pga_grads": (_i32, [C.c_void_p, View, View, View, View, Box, C.c_void_p]),
    "pga_transport_grads": (_i32, [C.c_void_p, View, View, View, View, C.POINTER(Box), _i32, View, View, Box,
                                   C.c_void_p]),
    "pga_viscsponge": (_i32, [View, View, _dp, C.POINTER(Box), _i32, C.c_void_p]),
    "pga_diffflux": (_i32, [C.c_void_p, _i32, C.c_double, View, View, View, View, View, View, View, View, View, Box,
                            C.c_void_p]),
    "pga_wall_mirror": (_i32, [C.c_void_p, _i32, _i32, C.c_void_p, _ip, _i32, View, View, View, C.c_void_p]),
    "pga_wall_flux": (_i32, [C.c_void_p, _i32, _i32, C.c_void_p, _ip, _i32, View, View, View, View, View, View, View,
                             C.c_void_p]),
    "pga_cflmax_worksize": (C.c_int64, [Box]),
    "pga_substep_listsize": (C.c_int64, [C.POINTER(SubstepOpts), Box]),
    "pga_substep": (_i32, [C.c_void_p, C.POINTER(SubstepOpts), View, View, View, _i32, C.c_void_p, C.c_int64,
                           C.c_void_p, C.c_int64, C.c_void_p, Box, C.c_void_p]),
    "pga_cflmax": (_i32, [C.c_void_p, View, View, View, View, View, View, Box, C.c_void_p, C.c_int64, C.c_void_p,
                          C.c_void_p]),
    "pga_cflmax_noalloc": (_i32, [C.c_void_p, View, View, View, View, View, View, Box, C.c_void_p, C.c_void_p]),
    "pga_batch_bytes": (C.c_int64, [_i32, _i32]),
    "pga_batch_create": (_i32, [C.c_void_p, _i32, C.POINTER(BatchBlock), _i32, C.c_void_p, C.c_int64,
                                C.POINTER(C.c_void_p)]),
    "pga_batch_destroy": (_i32, [C.c_void_p]),
    "pga_batch_halo": (_i32, [C.c_void_p, _i32, C.c_void_p]),
    "pga_batch_rates": (_i32, [C.c_void_p, _i32, C.c_void_p]),
    "pga_batch_flux": (_i32, [C.c_void_p, _i32, _i32, C.c_void_p]),
    "pga_batch_stage": (_i32, [C.c_void_p, _i32, _i32, _i32, _ip, _dp, _ip, C.c_void_p, C.c_void_p]),
    "pga_batch_implicit_worksize": (C.c_int64, [C.c_void_p]),
    "pga_batch_implicit": (_i32, [C.c_void_p, C.POINTER(ImplicitOpts), _i32, C.c_void_p, C.c_int64, C.c_void_p,
                                  C.c_int64, C.c_void_p, C.c_void_p]),
    "pga_graph_begin": (_i32, [C.c_void_p]),
    "pga_graph_end": (_i32, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "pga_graph_launch": (_i32, [C.c_void_p, C.c_void_p]),
    "pga_graph_destroy": (_i32, [C.c_void_p]),
    "pga_allfinite": (_i32, [View, _i32, Box, C.c_void_p, C.c_void_p]),
    "pga_implicit_worksize": (C.c_int64, [C.c_void_p, Box]),
    "pga_implicit": (_i32, [C.c_void_p, C.POINTER(ImplicitOpts), View, View, View, _i32, C.c_void_p, C.c_int64,
                            C.c_void_p, C.c_int64, C.c_void_p, Box, C.c_void_p]),
    "pga_lu_worksize": (C.c_int64, [C.c_int64, _i32]),
    "pga_lu_solve": (_i32, [C.c_int64, _i32, C.c_void_p, C.c_void_p, C.c_void_p, _i32, C.c_double, C.c_void_p,
                            C.c_void_p, C.c_int64, C.c_void_p]),
}


class NativeError(RuntimeError):
    pass


_libs = {}


# the library interface these bindings are written against (PGA_ABI_VERSION)
ABI_VERSION = 4


def _load(path):
    path = Path(path)
    if str(path) not in _libs:
        if not path.is_file():
            raise NativeError(f"{path} is not built: run peregrine_amd.build (or __graft_entry__.build())")
        handle = C.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype, fn.argtypes = res, args
        if handle.pga_abi_version() != ABI_VERSION:
            raise NativeError(f"{path} has ABI version {handle.pga_abi_version()}, these bindings {ABI_VERSION}: "
                              "rebuild it")
        _libs[str(path)] = handle
    return _libs[str(path)]


def lib():
    """The main library, loaded; raises when it is not built."""
    return _load(LIBPATH)


def lib_at(path):
    """A library built for one case (peregrine_amd.build.build_variant), loaded with the same
    bindings as the main one."""
    return _load(path)


def check(rc, what, L=None):
    """Raise on a failed call, with the message of the library that ran it (L, else the main one)."""
    if rc != 0:
        msg = (L or lib()).pga_last_error().decode(errors="replace")
        raise NativeError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")


def box(start, extent):
    return Box((_i32 * 3)(*start), (_i32 * 3)(*extent))


def ints(values):
    return (_i32 * len(values))(*values)


def dbls(values):
    return (C.c_double * len(values))(*values)


def mechanisms(L=None):
    L = L or lib()
    return [
        (L.pga_mechanism_name(i).decode(), L.pga_mechanism_signature(i).decode(), L.pga_mechanism_table_size(i))
        for i in range(L.pga_mechanism_count())
    ]


def const_layout(ns, L=None):
    """{section: (offset, count)} of the species constant table, and its size."""
    L = L or lib()
    names = ("MW", "MWinv", "Rsp", "cppolyR", "h298R", "cp0", "gpoly", "h298", "mu0", "kappa0", "lewis", "hpoly",
             "MWqinv", "wilkepair", "mupoly", "kappapoly", "cppoly", "cubic", "cubicsqrtac", "cubicb", "cubicda",
             "cubicmTc", "cubiconepm", "cubicsqrtacMW", "cubicbMW", "vshift", "vshiftMW", "cubicdaMW", "acentric",
             "chungsigma", "chungepsk", "chungvisc", "chungcond", "pairfit", "selffit", "bathof", "diffmeta")
    out = {}
    for name in names:
        count = _i32(0)
        off = L.pga_const_offset(ns, name.encode(), C.byref(count))
        out[name] = (off, count.value)
    return out, L.pga_const_size(ns)
