"""The C-ABI library loads, exports every function include/aphi_gpu.h declares,
and maps failures onto the reference's error conventions (no device needed)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

import aphi_amd as A
from aphi_amd import _capi


def declared_functions():
    text = open(os.path.join(ROOT, "include", "aphi_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aphi_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_symbol_has_a_binding():
    assert set(declared_functions()) <= set(_capi.SIGNATURES)


def test_invalid_arguments_raise_value_error():
    with pytest.raises(ValueError):
        A.Problem(A.CubeConfig(n=0))
    with pytest.raises(ValueError):
        A.Problem(A.CubeConfig(n=4, bar=(3, 6, 0, 1)))
    with pytest.raises(ValueError):
        A.Problem(A.CubeConfig(n=4, side=-1.0))
    with pytest.raises(ValueError):
        A.Problem(A.CubeConfig(n=4, ordering=7))
    p = A.Problem(A.CubeConfig(n=2, bar=(0, 1, 0, 1)))
    with pytest.raises(ValueError):
        p.ints("no_such_array")
    assert A.lib().aphi_version().decode().startswith("aphi")


def test_no_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(A.DeviceError):
        A.Context(0)


def test_handles_are_released_once():
    """close() destroys the host handle once; later calls see a closed handle
    (invalid argument), a second close() is a no-op, and a collected object's
    finalizer has run."""
    import gc
    import weakref
    p = A.Problem(A.CubeConfig(n=2, bar=(0, 1, 0, 1)))
    fin = p._fin
    assert fin.alive and p.ints("row_ptr").size == p.n + 1
    p.close()
    assert not fin.alive and p._h is None
    p.close()
    with pytest.raises(ValueError):
        p.ints("row_ptr")
    q = A.Problem(A.CubeConfig(n=2, bar=(0, 1, 0, 1)))
    fin, ref = q._fin, weakref.ref(q)
    del q
    gc.collect()
    assert ref() is None and not fin.alive
