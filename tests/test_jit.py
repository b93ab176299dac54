"""Libraries built for one case (peregrine_amd.build.build_variant), on the CPU: a case the main
library is not compiled for (its species count, its mechanism's structure) gets a library of its
own, keyed by the sha256 of what it is compiled from, as the reference keys its build cache
(peregrine/backend/base/compiler.py:42-64). These tests build the two libraries the device tests
use, so those find them cached."""

import support as sp
from peregrine_amd import build, native
from peregrine_amd.config import Config
from peregrine_amd.mixture import get_mixture
from peregrine_amd.simulator import case_mechanism


def test_a_variant_is_keyed_by_what_it_is_built_from():
    five = build.variant_texts(5)
    assert "#define PGA_NS_CASES(X) X(5)" in five["variant.hpp"]
    assert "((n) == 0)" in five["variant.hpp"]  # no real gas unless asked
    assert "((n) == 5)" in build.variant_texts(5, realgas=True)["variant.hpp"]
    assert build.variant_key(five) == build.variant_key(build.variant_texts(5))
    keys = {build.variant_key(build.variant_texts(n)) for n in (4, 5, 6)}
    assert len(keys) == 3
    assert build.variant_key(build.variant_texts(5, realgas=True)) != build.variant_key(five)
    # the main library's own generated headers carry no species list of their own
    assert "variant.hpp" not in build.generated()


def test_a_five_species_library_builds_and_binds():
    path = build.build_variant(5)
    assert path.parent.parent == build.JIT
    L = native.lib_at(path)
    assert L is not native.lib()
    assert L.pga_abi_version() == native.ABI_VERSION
    assert L.pga_mechanism_count() == 0
    for name in native.SIGNATURES:
        assert getattr(L, name).argtypes is not None
    # the species constant layout does not depend on what a library is compiled for
    assert native.const_layout(5, L) == native.const_layout(5)
    # found in the cache the second time, not rebuilt
    stamp = path.stat().st_mtime_ns
    assert build.build_variant(5) == path and path.stat().st_mtime_ns == stamp


def _trimmed_mixture(tmp_path):
    cfg = Config()
    m = cfg["mixture"]
    m["species"], m["eos"], m["T-range"] = sp.trimmed_mechanism(tmp_path), "tpg", (300.0, 3500.0)
    m["chemistry"] = "finite-rate"
    return get_mixture(cfg)


def test_a_mechanism_of_its_own_structure_gets_its_library(tmp_path):
    mix = _trimmed_mixture(tmp_path)
    name, records, sets, _, sig = case_mechanism(mix)
    assert len(records) == 21
    assert native.lib().pga_mechanism_find(sig.encode()) < 0
    L = native.lib_at(build.build_variant(mix.ns, mechanism=(name, records, sets)))
    assert L.pga_mechanism_count() == 1
    assert L.pga_mechanism_find(sig.encode()) == 0
    assert L.pga_mechanism_name(0).decode() == name


def test_the_main_library_reports_a_species_count_it_lacks():
    """The check a case takes to its own library: the main library makes a 5-species handle (its
    constant layout covers any count up to 16) but has no stage for it; a 9-species one has."""
    import ctypes

    L = native.lib()
    dummy = (ctypes.c_double * 1)()
    for ns, has in ((5, False), (9, True)):
        desc = native.CaseDesc()
        desc.ns, desc.eos, desc.mechanism, desc.trans = ns, 1, -1, -1
        h = ctypes.c_void_p()
        assert L.pga_case_create(ctypes.byref(desc), ctypes.cast(dummy, ctypes.c_void_p), None, ctypes.byref(h)) == 0
        assert bool(L.pga_stage_supported(h)) == has
        L.pga_case_destroy(h)


def test_a_real_gas_library_at_three_species():
    """The cubic real gas at three species (the main library has it at one and two): its library,
    which the three-species device test takes."""
    path = build.build_variant(3, realgas=True)
    L = native.lib_at(path)
    assert "((n) == 3)" in build.variant_texts(3, realgas=True)["variant.hpp"]
    assert build.variant_key(build.variant_texts(3, realgas=True)) != build.variant_key(build.variant_texts(3))
    import ctypes

    dummy = (ctypes.c_double * 1)()
    desc = native.CaseDesc()
    desc.ns, desc.eos, desc.mechanism, desc.trans = 3, 2, -1, -1
    h = ctypes.c_void_p()
    assert native.lib().pga_case_create(ctypes.byref(desc), ctypes.cast(dummy, ctypes.c_void_p), None,
                                        ctypes.byref(h)) == native.PGA_EUNSUP
    assert L.pga_case_create(ctypes.byref(desc), ctypes.cast(dummy, ctypes.c_void_p), None, ctypes.byref(h)) == 0
    assert L.pga_stage_supported(h)
    L.pga_case_destroy(h)

