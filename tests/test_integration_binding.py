"""The reference-side binding written out in the integration notes, run as written (no GPU): its ctypes
structs lay out as the ABI's, its argument lists match the ABI's, and its view() turns a
reference-layout ArrInfo (peregrine/backend/abi.py:38, an Allocation's `info` as
peregrine/backend/allocation.py:20-27 fills it) into the pga_view the package itself makes for the
same array, and refuses an array it cannot view."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from peregrine_amd import native

ROOT = Path(__file__).resolve().parents[1]


class ArrInfo(ctypes.Structure):
    """The reference's ArrInfo as its AbiStruct builds it: data, rank, extent[5], stride[5] (elements)."""

    _fields_ = [("data", ctypes.c_void_p), ("rank", ctypes.c_int), ("extent", ctypes.c_int * 5),
                ("stride", ctypes.c_long * 5)]


def _arrinfo(ptr, shape, strides):
    pad = [0] * (5 - len(shape))
    return ArrInfo(ptr, len(shape), (ctypes.c_int * 5)(*shape, *pad), (ctypes.c_long * 5)(*strides, *pad))


def _fstrides(shape):
    out, step = [], 1
    for n in shape:
        out.append(step)
        step *= n
    return out


@pytest.fixture(scope="module")
def binding():
    """The integration notes' pga.py blocks executed in one namespace, with the library loaded from
    this tree."""
    native.lib()
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```python\n(.*?)```", text, re.S)
    mine = [b for b in blocks if b.startswith("# peregrine/backend/mi355x/pga.py") or b.startswith("# (pga.py, continued)")]
    assert len(mine) == 2
    code = "\n".join(mine).replace('ctypes.CDLL("libpga.so")', f'ctypes.CDLL({str(native.LIBPATH)!r})')
    ns = {}
    exec(compile(code, "INTEGRATION.md", "exec"), ns)
    return ns


def _layout(struct):
    return ctypes.sizeof(struct), [(name, getattr(struct, name).offset, getattr(struct, name).size)
                                   for name, *_ in struct._fields_]


@pytest.mark.parametrize("name,ours", [("View", native.View), ("Box", native.Box), ("CaseDesc", native.CaseDesc),
                                       ("Wall", native.Wall), ("SubstepOpts", native.SubstepOpts),
                                       ("ImplicitOpts", native.ImplicitOpts)])
def test_the_documented_structs_lay_out_as_the_abi(binding, name, ours):
    assert _layout(binding[name]) == _layout(ours)


def test_the_documented_argument_lists_are_the_abi(binding):
    lib = binding["lib"]
    declared = [n for n in native.SIGNATURES if hasattr(lib, n) and getattr(lib, n).argtypes is not None]
    assert len(declared) >= 20
    # every entry point the documented step() calls is among them
    assert {"pga_halo", "pga_rates", "pga_flux", "pga_stage", "pga_graph_launch", "pga_implicit"} <= set(declared)
    for n in declared:
        got = getattr(lib, n).argtypes
        want = native.SIGNATURES[n][1]
        assert len(got) == len(want), n
        for g, w in zip(got, want):
            assert ctypes.sizeof(g) == ctypes.sizeof(w), (n, g, w)


def test_an_arrinfo_becomes_the_packages_view(binding):
    """A padded cell array (i rows of 24 for 20 cells) and a face array, as the reference's
    Allocation describes them, against peregrine_amd.native.view of the same memory."""
    buf = np.zeros(24 * 7 * 6 * 13)
    ptr = buf.ctypes.data
    for shape, strides in (((20, 7, 6, 13), (1, 24, 24 * 7, 24 * 7 * 6)),
                           ((21, 7, 6, 13), _fstrides((21, 7, 6, 13)))):
        v = binding["view"](_arrinfo(ptr, shape, strides))
        want = native.view(ptr, *shape, strides=strides[1:])
        assert bytes(v) == bytes(want)
        assert (v.sj, v.sk, v.sc, list(v.n)) == (strides[1], strides[2], strides[3], list(shape))


def test_an_array_it_cannot_view_is_refused(binding):
    shape = (20, 7, 6, 13)
    c_order = list(reversed(_fstrides(list(reversed(shape)))))  # the host backend's order "C": comp fastest
    with pytest.raises(ValueError):
        binding["view"](_arrinfo(0, shape, c_order))
    with pytest.raises(ValueError):
        binding["view"](_arrinfo(0, (20, 7, 6), _fstrides((20, 7, 6))))
