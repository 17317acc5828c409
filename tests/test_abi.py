"""The C ABI library loads and exports every entry point include/bfvm_hip.h
declares; descriptor validation runs without a device.  CPU only."""
import ctypes as C

import numpy as np
import pytest

from physicore_amd import bfvm


def test_library_exports_every_declared_symbol():
    L = bfvm.lib()
    names = bfvm.declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_binding_covers_header():
    assert sorted(bfvm._SIGS) == bfvm.declared_symbols()


@pytest.mark.parametrize("bad", [
    dict(dims=4), dict(dims=0), dict(grid=(10, 10, 0)), dict(grid=(10, 1, 5), dims=1),
    dict(dt=0.0), dict(dt=-1.0), dict(diffusion=[-1.0]), dict(decay=[float("nan")]),
])
def test_invalid_descriptor_rejected(bad):
    args = dict(dims=3, grid=(10, 10, 10), voxel=(20, 20, 20), dt=0.01, diffusion=[1.0],
                decay=[0.1])
    args.update(bad)
    with pytest.raises(bfvm.BfvmError) as e:
        bfvm.Domain(args["dims"], args["grid"], args["voxel"], args["dt"], args["diffusion"],
                    args["decay"], [0.0] * len(args["diffusion"]))
    assert e.value.code == -1


def test_null_output_pointer():
    desc = bfvm.DomainDesc()
    assert bfvm.lib().bfvm_create(C.byref(desc), None) == -1
    assert b"null" in bfvm.lib().bfvm_last_error(None)


def test_device_count_callable():
    assert bfvm.device_count() >= 0
    np.zeros(1)
