"""The C-ABI library loads and exports every entry point include/cnq.h
declares; host-only entry points behave as documented. No GPU is touched."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from continuum_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "cnq.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cnq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(str(_lib.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers the same set
    assert sorted(_lib.SIGNATURES) == names


def test_default_options_are_the_references():
    o = _lib.default_options()
    assert (o.hierarchy, o.prefilter, o.wide_extent, o.light_cone_cap, o.wide_cone_cap) == (2, 0.0, 8.0, 4096, 512)
    assert o.object_cov == _lib.NONE and o.nearest_first == 1 and o.profile_stage == 0


def test_no_device_is_reported_not_raised():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = _lib.lib().cnq_device_create(0, C.byref(h))
    assert rc == _lib.NO_DEVICE
    assert b"device" in _lib.lib().cnq_last_error()
    from continuum_amd.kernel import MaskConeKernel

    assert MaskConeKernel.new() is None


def test_flat_build_rejects_null_arguments():
    h = C.c_void_p()
    rc = _lib.lib().cnq_flat_build(0, 0, 3, C.byref(h))
    assert rc == _lib.INVALID_ARG


def test_flat_build_of_nothing_pads_like_the_reference():
    """flatten (hierarchy.rs:52-64) pads empty inputs; mesh_hierarchy of no
    meshes is one leaf holding mesh 0 (hierarchy.rs:206-212)."""
    from continuum_amd.kernel import Flat

    f = Flat.build([])
    assert f.tris.shape == (1, 12) and not f.tris.any()
    assert f.ranges.tolist() == [[0, 0, _lib.NONE, _lib.NONE]]
    assert f.blas.shape == (1, 12) and f.blas_items.tolist() == [0]
    assert f.nodes[0][8:10].tolist() == [0, 1] and f.items.tolist() == [0]
