"""C ABI checks that need no GPU: the library loads beside torch, exports every
symbol include/mi355nufft.h declares, validates arguments before touching the
device, and the product path fails loudly without a GPU."""

import ctypes
import os
import re

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import _lib
from mrrt.nufft_amd.build import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mi355nufft.h")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="libmi355nufft.so not built")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mi355nufft_\w+)\s*\(", src)))


def test_header_symbols_exported():
    lib = _lib.get_lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED_SYMBOLS) == names


def test_abi_version():
    assert _lib.get_lib().mi355nufft_abi_version() == 2


def test_single_hip_runtime():
    _lib.get_lib()
    maps = {l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64" in l}
    assert len(maps) == 1, maps
    assert "torch" in next(iter(maps))


def _create(ndim=2, Nd=(8, 8), Kd=(16, 16), Jd=(6, 6), L=64, prec=0, phasing=0):
    lib = _lib.get_lib()
    plan = ctypes.c_void_p()
    tabs = [np.zeros(j * L + 1, np.float32) for j in Jd[:ndim]]
    sns = [np.ones(n, np.float64) for n in Nd[:ndim]]
    tp = (ctypes.c_void_p * max(ndim, 1))(*[t.ctypes.data for t in tabs])
    sp = (ctypes.POINTER(ctypes.c_double) * max(ndim, 1))(
        *[s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for s in sns])
    rc = lib.mi355nufft_plan_create(ctypes.byref(plan), ndim, _lib.i64arr(Nd), _lib.i64arr(Kd), _lib.i32arr(Jd), L,
                                    prec, phasing, 0, 1.0, tp, sp, _lib.dblarr([0.0] * 3), None)
    return rc, lib.mi355nufft_last_error().decode()


@pytest.mark.parametrize("kw,msg", [
    (dict(ndim=4, Nd=(8,) * 4, Kd=(16,) * 4, Jd=(6,) * 4), "ndim"),
    (dict(Jd=(17, 6)), "Jd"),
    (dict(Kd=(0, 16)), "Kd"),
    (dict(L=0), "L must"),
    (dict(prec=2), "precision"),
    (dict(phasing=3), "phasing"),
    (dict(Nd=(32, 8)), "Nd"),
])
def test_plan_create_rejects(kw, msg):
    rc, err = _create(**kw)
    assert rc == 1 and msg in err


def test_table1_rejects_without_device_work():
    lib = _lib.get_lib()
    h = (ctypes.c_void_p * 2)(None, None)
    rc = lib.mi355nufft_table1_forward(2, _lib.i64arr((16, 16)), _lib.i32arr((6, 6)), 64, 0, 0, None, h, None, 10,
                                       None, 1, None)
    assert rc == 1 and "h[0]" in lib.mi355nufft_last_error().decode()
    rc = lib.mi355nufft_table1_forward(2, _lib.i64arr((16, 16)), _lib.i32arr((6, 6)), 64, 0, 0, None, None, None, 10,
                                       None, 0, None)
    assert rc == 1 and "nreps" in lib.mi355nufft_last_error().decode()
    rc = lib.mi355nufft_forward(None, None, None, None, 1, None)
    assert rc == 1 and "null plan" in lib.mi355nufft_last_error().decode()


def test_check_raises():
    lib = _lib.get_lib()
    lib.mi355nufft_forward(None, None, None, None, 1, None)
    with pytest.raises(ValueError, match="null plan"):
        _lib.check(1)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_product_path_fails_loudly_without_gpu():
    from mrrt.nufft_amd import NufftBase
    om = np.zeros((4, 2))
    with pytest.raises(RuntimeError, match="GPU"):
        NufftBase((8, 8), om, Jd=6, Kd=(16, 16))


def test_operator_argument_errors():
    """Reference ValueErrors (_nufft.py:309-350, :505-508, :541-542, :556-566)."""
    from mrrt.nufft_amd import NufftBase
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    om = np.zeros((4, 2))
    with pytest.raises(ValueError, match="phasing"):
        NufftBase((8, 8), om, Jd=6, Kd=(16, 16), phasing="bogus", device=dev)
    with pytest.raises(ValueError, match="mode"):
        NufftBase((8, 8), om, Jd=6, Kd=(16, 16), mode="bogus", device=dev)
    with pytest.raises(ValueError, match="exact"):
        NufftBase((8, 8), om, Jd=6, Kd=(16, 16), mode="exact", device=dev)
    with pytest.raises(ValueError, match="cols"):
        NufftBase((8, 8), np.zeros((4, 3)), Jd=6, Kd=(16, 16), device=dev)
    with pytest.raises(ValueError, match="types"):
        NufftBase((8, 8), np.zeros((4, 2), np.int64), Jd=6, Kd=(16, 16), device=dev)
    with pytest.raises(ValueError, match="precision"):
        NufftBase((8, 8), om, Jd=6, Kd=(16, 16), precision="half", device=dev)
    with pytest.raises(ValueError):
        NufftBase((8, 8), om, Jd=(6, 6, 6), Kd=(16, 16), device=dev)


@pytest.mark.parametrize("call,msg", [
    (lambda L: L.mi355nufft_coil_expand(0, None, None, 64, 1, 4, None, None), "required"),
    (lambda L: L.mi355nufft_coil_expand(2, None, None, 64, 1, 4, None, None), "precision"),
    (lambda L: L.mi355nufft_coil_expand(0, None, None, 64, 1, 0, None, None), "ncoils"),
    (lambda L: L.mi355nufft_coil_combine(0, None, None, 64, 20000, 4, None, None), "ncoils"),
    (lambda L: L.mi355nufft_coil_combine(1, None, None, 0, 1, 4, None, None), "npix"),
    (lambda L: L.mi355nufft_weight_samples(0, None, None, 10, 0, None, None), "nreps"),
    (lambda L: L.mi355nufft_weight_samples(0, None, None, 10, 1, None, None), "required"),
])
def test_coil_entry_points_reject_without_device_work(call, msg):
    lib = _lib.get_lib()
    assert call(lib) == 1 and msg in lib.mi355nufft_last_error().decode()


def test_sparse_entry_points_reject_null_plan():
    lib = _lib.get_lib()
    assert lib.mi355nufft_set_sparse(None, None, None, None) == 1
    assert lib.mi355nufft_clear_sparse(None) == 1
    assert lib.mi355nufft_sparse_apply(None, 0, None, None, 1) == 1
