"""The C-ABI library: it loads without a GPU, exports every entry point declared
in include/tis_gpu.h, rejects bad arguments before touching the device, and its
Philox4x32-10 matches the published known-answer vectors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import tis_ctypes as T
from conftest import REPO


def declared_functions():
    text = open(os.path.join(REPO, "include", "tis_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(tis_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


# Random123 known-answer vectors for philox4x32-10 (ctr, key -> out)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(lib, ctr, key, want):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.tis_philox4x32_10(c, k, o)
    assert tuple(o) == want


def test_default_options_are_the_reference_settings(lib):
    o = T.default_options(lib)
    assert (o.skin, o.r_sr, o.r_lr, o.lr_every) == (3.0, 15.0, 53.0, 10)
    assert o.precision == T.PREC_FP64 and o.noise_scale == 1.0
    assert (o.pair_lanes, o.pair_block, o.list_waves, o.graph_steps, o.hb_groups) == (4, 256, 4096, 50, 4)


@pytest.mark.parametrize("field,value", [("pair_lanes", 3), ("pair_lanes", 128), ("pair_block", 96),
                                         ("pair_block", 512), ("list_waves", 10), ("graph_steps", -1),
                                         ("hb_groups", -1), ("hb_groups", 17)])
def test_create_rejects_bad_launch_shapes_without_a_device(lib, cfg1, field, value):
    h = T.SystemHandle(cfg1[0])
    o = T.default_options(lib)
    setattr(o, field, value)
    ctx = C.c_void_p()
    assert lib.tis_create(h.ptr(), 1, C.byref(o), C.byref(ctx)) == -1
    assert field.encode() in lib.tis_last_error()


def test_create_rejects_bad_arguments_without_a_device(lib, cfg1):
    ctx = C.c_void_p()
    assert lib.tis_create(None, 1, None, C.byref(ctx)) == -1
    d = dict(cfg1[0])
    d["recentre"] = d["n_beads"] + 5
    h = T.SystemHandle(d)
    assert lib.tis_create(h.ptr(), 1, None, C.byref(ctx)) == -1
    assert b"recentring" in lib.tis_last_error()
    d = dict(cfg1[0])
    bad = d["bond_ij"].copy()
    bad[3] = d["n_beads"] + 1
    d["bond_ij"] = bad
    h = T.SystemHandle(d)
    assert lib.tis_create(h.ptr(), 1, None, C.byref(ctx)) == -1
    assert b"bond" in lib.tis_last_error()


def test_null_context_calls_fail_cleanly(lib):
    assert lib.tis_step(None, 1) == -1
    assert lib.tis_get_state(None, 0, None, None, None) == -1
    e = T.Energies()
    assert lib.tis_get_energies(None, 0, C.byref(e)) == -1
    lib.tis_destroy(None)


def test_observables_entry_points_reject_bad_arguments_without_a_device(lib):
    o = T.Observables()
    assert lib.tis_observe(None, 10, 7.0) == -1
    assert lib.tis_observe_now(None) == -1
    assert lib.tis_get_observables(None, 0, C.byref(o), None, 0, None, 0, None, 0) == -1
    assert lib.tis_autotune(None, 1, None, 100, None) == -1
