"""The C ABI: both libraries load without a GPU and export every function
that include/hptracing.h declares; argument validation needs no device."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from hptracing_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "hptracing.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hpt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ["hpt_scene_create", "hpt_scene_destroy", "hpt_cast", "hpt_render",
                     "hpt_last_error", "hpt_host_load_obj", "hpt_host_describe"]:
        assert required in names


def test_every_declared_symbol_is_exported():
    host, dev = N.host(), N.device()
    missing = [n for n in declared_functions()
               if not (hasattr(host, n) or hasattr(dev, n))]
    assert not missing, missing


def test_host_symbols_in_host_lib_device_symbols_in_device_lib():
    host, dev = N.host(), N.device()
    for n in declared_functions():
        if n.startswith("hpt_host_"):
            assert hasattr(host, n), n
        else:
            assert hasattr(dev, n), n


def test_null_arguments_are_rejected_without_a_device():
    dev = N.device()
    assert dev.hpt_scene_create(None, None) == N.HPT_ERR_INVALID
    assert b"null" in dev.hpt_last_error()
    assert dev.hpt_cast(None, None, None, 0, None, None, None) == N.HPT_ERR_INVALID
    assert dev.hpt_render(None, None, None, None, None) == N.HPT_ERR_INVALID
    host = N.host()
    out = C.c_void_p()
    assert host.hpt_host_load_obj(None, None, 0, C.byref(out)) == N.HPT_ERR_INVALID


@pytest.mark.parametrize("field,value", [("parent", 10**6), ("sibling", 10**6), ("child", -7),
                                         ("parent", -2)])
def test_scene_create_rejects_node_links_out_of_range(cornell, field, value):
    """hpt_scene_create validates every KD link before it follows any: a
    parent, sibling or child outside [-1, n_nodes) is a scene error, found
    on the host before the first device call."""
    import numpy as np
    dev = N.device()
    d = cornell.desc()
    nodes = np.array(cornell.nodes, copy=True)
    nodes[len(nodes) // 2][field] = value
    d.nodes = nodes.ctypes.data
    out = C.c_void_p()
    assert dev.hpt_scene_create(C.byref(d), C.byref(out)) == N.HPT_ERR_SCENE
    assert b"link out of range" in dev.hpt_last_error()
    assert not out.value
