"""C-ABI and host-side logic, CPU only (no kernel launches).

* libnlse_hip.so loads and exports every function that include/nlse_hip.h
  declares, with the ctypes signatures of nlse_amd._lib;
* argument errors come back as error codes plus a message (no GPU call
  happens before the arguments are validated);
* host logic of the drop-in classes: the step-count rule, parameter
  broadcasting, the separable propagator factors;
* the product package never imports the oracle.
"""

import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nlse_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(nlse_\w+)\s*\(", text, flags=re.M)))


def test_header_parsed():
    names = declared_functions()
    assert "nlse_run" in names and "nlse_fft_exec" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from nlse_amd import _lib

    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # and the Python binding declares each of them
    assert set(declared_functions()) == set(_lib.SIGNATURES), (
        set(declared_functions()) ^ set(_lib.SIGNATURES))


def test_argument_errors_without_gpu():
    from nlse_amd import _lib

    lib = _lib.load()
    assert lib.nlse_version() >= 100
    h = C.c_void_p()
    dims = (C.c_int64 * 2)(100, 128)  # 100 is not a power of two
    rc = lib.nlse_fft_plan_create(2, dims, 1, C.byref(h))
    assert rc == -1 and b"unsupported" in lib.nlse_last_error()
    rc = lib.nlse_fft_plan_create(4, dims, 1, C.byref(h))
    assert rc == -1
    desc = _lib.PlanDesc(ndim=4, ny=64, nx=64, nt=1, items=1, ncomp=1)
    assert lib.nlse_plan_create(C.byref(desc), C.byref(h)) == -1
    desc = _lib.PlanDesc(ndim=2, ny=64, nx=64, nt=1, items=1, ncomp=3)
    assert lib.nlse_plan_create(C.byref(desc), C.byref(h)) == -1
    assert lib.nlse_run(None, None, 1, 0, 0, None) == -1
    assert lib.nlse_square_mod(None, None, 4, None) == -1
    bad = _lib.BParam(None, 0)
    assert lib.nlse_nl_prop(None, None, 1, 1, 0.0, bad, None, 0, bad, bad, 0, None) == -1
    with pytest.raises(_lib.NlseError):
        _lib.check(lib.nlse_run(None, None, 1, 0, 0, None), "run")


def test_step_count_rule_host():
    from nlse_amd.nlse import count_steps

    assert count_steps(10e-3, 1e-5) == 1001
    assert count_steps(10e-3, 5e-6) == 2001
    assert count_steps(2e-3, 1e-4) == 20
    assert count_steps(0.0, 1e-5) == 0


def test_parameter_broadcasting_host():
    from nlse_amd.nlse import bcast_items

    assert bcast_items(2.0, (3,)).tolist() == [2.0, 2.0, 2.0]
    a = np.arange(3.0).reshape(3, 1, 1)
    assert bcast_items(a, (3,)).tolist() == [0.0, 1.0, 2.0]
    b = np.arange(2.0).reshape(2, 1, 1, 1)
    assert bcast_items(b, (2, 4)).tolist() == [0, 0, 0, 0, 1, 1, 1, 1]
    assert bcast_items(5.0, ()).tolist() == [5.0]
    with pytest.raises(ValueError):
        bcast_items(np.ones((3, 4, 4)), (3,))


def test_separable_propagator_is_the_reference_table():
    """complex64(px[kx] * py[ky]) in FP64 equals the reference propagator
    exp(-i (kx^2 + ky^2) dz / 2k) cast to complex64 (nlse.py:194-197) on all
    but a vanishing fraction of entries (rounding ties)."""
    N, dX, k, dz = 2048, 4 * 2.23e-3 / 2048, 2 * np.pi / 780e-9, 5e-6
    Kx = 2 * np.pi * np.fft.fftfreq(N, d=dX)
    Kxx, Kyy = np.meshgrid(Kx, Kx)
    ref = np.exp(-1j * 0.5 * (Kxx**2 + Kyy**2) / k * dz).astype(np.complex64)
    px = np.exp(-1j * 0.5 * Kx**2 / k * dz)
    sep = (px[:, None] * px[None, :]).astype(np.complex64)
    assert np.count_nonzero(sep != ref) <= 10
    assert np.abs(sep.astype(np.complex128) - ref).max() < 1e-7


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "nlse_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                text = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", text, flags=re.M), f
