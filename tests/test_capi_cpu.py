"""The C-ABI library builds, loads and exports every symbol the header
declares; error paths that need no device behave like the Go sentinels
(gpu/plan.go:17-29, gpu/errors.go)."""
import ctypes
import os
import re

import numpy as np
import pytest

import algofft_hip as A
from algofft_hip import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "algofft_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(afft_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert sorted(n for n, _, _ in _capi.SIGNATURES) == declared_symbols()


def test_status_strings_match_go_sentinels():
    lib = A.lib()
    assert lib.afft_status_string(1) == b"algofft/gpu: no backend registered"
    assert lib.afft_status_string(6) == b"algofft/gpu: length mismatch"
    assert lib.afft_status_string(7) == b"algo-fft: invalid stride"


def test_backend_info_without_device():
    info = A.backend_info()
    assert info["name"] == "hip"


def test_invalid_length_before_any_device_call():
    # gpu/plan.go:18-20: n < 1 -> ErrInvalidLength, checked first
    h = ctypes.c_void_p()
    rc = A.lib().afft_plan_create(None, 0, 0, None, ctypes.byref(h))
    assert rc in (A.ErrInvalidLength.status,)


def test_null_handles_are_errors():
    lib = A.lib()
    assert lib.afft_plan_forward(None, None, 0, None, 0) == A.ErrNotImplemented.status
    assert lib.afft_stream_synchronize(None) == A.ErrInvalidArgument.status
    assert lib.afft_buffer_upload(None, None, 0) == A.ErrInvalidArgument.status


def test_no_device_means_unavailable():
    # HIP_VISIBLE_DEVICES / no GPU: Available()==false, NewContext fails with
    # ErrBackendUnavailable (gpu/plan.go:27-29)
    if A.available():
        pytest.skip("a GPU is visible")
    with pytest.raises(A.ErrBackendUnavailable):
        A.Context(0)


def test_fast_plan_length_checks_come_first():
    # NewFastPlan (plan_fast.go:47-49): ErrInvalidLength before anything else
    import ctypes
    h = ctypes.c_void_p()
    for n in (0, -4, 1000, 3):
        assert A.lib().afft_fast_plan_create(None, n, 1, A.COMPLEX64, None, ctypes.byref(h)) == 4
    assert A.lib().afft_fast_plan_create(None, 1024, 0, A.COMPLEX64, None, ctypes.byref(h)) == 4
    assert A.lib().afft_fast_plan_create(None, 1024, 1, A.COMPLEX64, None, ctypes.byref(h)) == 10


@pytest.mark.parametrize("kind,prec,call,dst,src", [
    (A.C2C, A.COMPLEX128, "forward", np.complex128, np.complex64),
    (A.C2C, A.COMPLEX64, "inverse", np.complex128, np.complex64),
    (A.R2C, A.COMPLEX64, "forward", np.complex64, np.float64),
    (A.R2C, A.COMPLEX64, "inverse", np.complex64, np.complex64),
    (A.R2C, A.COMPLEX128, "forward_full", np.complex128, np.complex128),
    (A.R2C, A.COMPLEX64, "inverse_full", np.float64, np.complex64),
])
def test_host_slices_of_the_wrong_element_type_are_refused(kind, prec, call, dst, src):
    """The host-slice entry points count elements, not bytes: an array of the
    wrong element type is refused with ErrNotImplemented before any copy
    (a wrong dynamic type in gpu/plan.go's type switch)."""
    d = _capi.PlanDesc()
    d.kind, d.precision, d.rank = kind, prec, 1
    d.dims[0] = 64
    p = A.Plan(None, d)   # the check runs before the library is called
    with pytest.raises(A.ErrNotImplemented):
        getattr(p, call)(np.zeros(64, dst), np.zeros(64, src))
