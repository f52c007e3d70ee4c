"""Python view of the algofft_hip C ABI (include/algofft_hip.h).

Mirrors the reference's `gpu` package surface (gpu/backend.go, gpu/plan.go):
Backend information, Context, Buffer, Stream and Plan with Forward/Inverse,
plus the extended descriptor for batch, strided, real and N-D plans.  The
sentinel errors of gpu/errors.go and errors.go become exception classes.

This module only drives the HIP library; it computes nothing itself.  If the
library is not built, every call raises RuntimeError.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import PlanDesc, PlanOptions

# afft_status codes
OK = 0
STATUS_NAMES = {
    1: "ErrNoBackend", 2: "ErrBackendUnavailable", 3: "ErrNotImplemented", 4: "ErrInvalidLength",
    5: "ErrNilSlice", 6: "ErrLengthMismatch", 7: "ErrInvalidStride", 8: "ErrInvalidSpectrum",
    9: "ErrDevice", 10: "ErrInvalidArgument",
}

COMPLEX64, COMPLEX128 = 0, 1
FORWARD, INVERSE = -1, 1
C2C, R2C, C2R = 0, 1, 2
FLAG_INPLACE, FLAG_MFMA, FLAG_NORMALIZE, FLAG_UNITARY, FLAG_REF_CHIRP = 1, 2, 4, 8, 16
STRATEGY_AUTO, STRATEGY_STOCKHAM, STRATEGY_SIXSTEP, STRATEGY_BLUESTEIN, STRATEGY_MIXED = 0, 2, 3, 5, 16


class AfftError(Exception):
    """Base of the sentinel errors; `.status` is the C status code."""

    status = -1

    def __init__(self, msg=""):
        super().__init__(f"{type(self).__name__}: {msg}")
        self.msg = msg


_ERR_CLASSES = {}
for _code, _name in STATUS_NAMES.items():
    _ERR_CLASSES[_code] = type(_name, (AfftError,), {"status": _code})
    globals()[_name] = _ERR_CLASSES[_code]


def lib():
    return _capi.load()


def check(rc):
    if rc != OK:
        msg = lib().afft_last_error()
        msg = msg.decode() if msg else ""
        raise _ERR_CLASSES.get(rc, AfftError)(msg)
    return rc


def available():
    """Backend.Available (gpu/backend.go:9)."""
    return bool(lib().afft_backend_available())


def backend_info():
    info = _capi.BackendInfo()
    check(lib().afft_backend_get_info(ctypes.byref(info)))
    return {"name": info.name.decode(), "version": info.version.decode(), "description": info.description.decode()}


def devices():
    """Backend.Devices (gpu/backend.go:10)."""
    n = ctypes.c_int(0)
    check(lib().afft_device_count(ctypes.byref(n)))
    out = []
    for i in range(n.value):
        d = _capi.DeviceInfo()
        check(lib().afft_device_get_info(i, ctypes.byref(d)))
        out.append({"name": d.name.decode(), "vendor": d.vendor.decode(), "driver": d.driver.decode(),
                    "memory_mb": d.memory_mb, "compute_cap": d.compute_cap.decode()})
    return out


def _np_prec(dtype):
    dt = np.dtype(dtype)
    if dt in (np.complex64, np.float32):
        return COMPLEX64
    if dt in (np.complex128, np.float64):
        return COMPLEX128
    raise NotImplementedError(f"unsupported dtype {dt}")


class Context:
    """Backend.NewContext / Context (gpu/backend.go:15-24)."""

    def __init__(self, device_index=0):
        h = ctypes.c_void_p()
        check(lib().afft_context_create(device_index, ctypes.byref(h)))
        self.h = h

    def device(self):
        d = _capi.DeviceInfo()
        check(lib().afft_context_device(self.h, ctypes.byref(d)))
        return {"name": d.name.decode(), "memory_mb": d.memory_mb, "compute_cap": d.compute_cap.decode()}

    def close(self):
        if self.h:
            check(lib().afft_context_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Buffer:
    """Context.NewBuffer / Buffer (gpu/backend.go:18, :27-35)."""

    def __init__(self, ctx, elem_count, precision=COMPLEX64):
        h = ctypes.c_void_p()
        check(lib().afft_buffer_create(ctx.h, elem_count, precision, ctypes.byref(h)))
        self.h = h
        self.precision = precision

    def __len__(self):
        return lib().afft_buffer_len(self.h)

    def upload(self, src):
        src = np.ascontiguousarray(src)
        if _np_prec(src.dtype) != self.precision or not np.iscomplexobj(src):
            raise ErrNotImplemented("wrong element type")  # noqa: F821
        check(lib().afft_buffer_upload(self.h, src.ctypes.data_as(ctypes.c_void_p), src.size))

    def download(self, dst):
        if not dst.flags.c_contiguous or _np_prec(dst.dtype) != self.precision:
            raise ErrNotImplemented("wrong element type")  # noqa: F821
        check(lib().afft_buffer_download(self.h, dst.ctypes.data_as(ctypes.c_void_p), dst.size))

    @property
    def ptr(self):
        return lib().afft_buffer_device_ptr(self.h)

    def close(self):
        if self.h:
            check(lib().afft_buffer_destroy(self.h))
            self.h = None


class Stream:
    """Context.NewStream / Stream (gpu/backend.go:20, :38-41)."""

    def __init__(self, ctx):
        h = ctypes.c_void_p()
        check(lib().afft_stream_create(ctx.h, ctypes.byref(h)))
        self.h = h

    def synchronize(self):
        check(lib().afft_stream_synchronize(self.h))

    @property
    def native(self):
        return lib().afft_stream_native(self.h)

    def close(self):
        if self.h:
            check(lib().afft_stream_destroy(self.h))
            self.h = None


class Plan:
    """A plan of the HIP backend (PlanImpl, gpu/backend.go:45-51).

    Build with Plan.new(ctx, n, precision) (Context.NewFFTPlan) or
    Plan.from_desc(ctx, kind=..., dims=..., batch=..., ...) for the extended
    batch / strided / real / N-D plans.
    """

    def __init__(self, handle, desc):
        self.h = handle
        self.desc = desc

    @classmethod
    def new(cls, ctx, n, precision=COMPLEX64, in_place=False):
        h = ctypes.c_void_p()
        opts = PlanOptions(0, 1, int(in_place))
        check(lib().afft_plan_create(ctx.h, n, precision, ctypes.byref(opts), ctypes.byref(h)))
        return cls(h, None)

    @classmethod
    def from_desc(cls, ctx, dims, precision=COMPLEX64, kind=C2C, batch=1, istride=0, idist=0, ostride=0,
                  odist=0, flags=0, strategy=0, split=0):
        if isinstance(dims, int):
            dims = (dims,)
        d = PlanDesc()
        d.kind = kind
        d.precision = precision
        d.rank = len(dims)
        for i, v in enumerate(dims):
            d.dims[i] = v
        d.batch, d.istride, d.idist, d.ostride, d.odist, d.flags = batch, istride, idist, ostride, odist, flags
        d.strategy, d.split = strategy, split
        h = ctypes.c_void_p()
        check(lib().afft_plan_create_desc(ctx.h, ctypes.byref(d), ctypes.byref(h)))
        return cls(h, d)

    def __len__(self):
        return lib().afft_plan_len(self.h)

    @property
    def precision(self):
        return lib().afft_plan_precision(self.h)

    def io_len(self, direction=FORWARD):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(lib().afft_plan_io_len(self.h, direction, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def describe(self):
        buf = ctypes.create_string_buffer(256)
        check(lib().afft_plan_describe(self.h, buf, 256))
        return buf.value.decode()

    def workspace_bytes(self):
        return lib().afft_plan_workspace_bytes(self.h)

    # -- host-slice execution (Forward/Inverse with []complex) --------------
    @property
    def kind(self):
        return self.desc.kind if self.desc is not None else C2C

    def _host(self, fn, dst, src, dst_real, src_real):
        """The C entry points count elements only, so the element type of each
        side is checked here: complex64/float32 for complex64 plans,
        complex128/float64 for complex128 plans, real or complex as the
        plan's kind has it for that direction.  A wrong type is
        ErrNotImplemented, as for a wrong dynamic type in gpu/plan.go."""
        if dst is None or src is None:
            raise ErrNilSlice("dst or src is nil")  # noqa: F821
        if not isinstance(dst, np.ndarray) or not isinstance(src, np.ndarray):
            raise ErrNotImplemented("dst and src must be numpy arrays")  # noqa: F821
        prec = self.desc.precision if self.desc is not None else self.precision
        for arr, real, side in ((dst, dst_real, "dst"), (src, src_real, "src")):
            want = {(COMPLEX64, False): np.complex64, (COMPLEX64, True): np.float32,
                    (COMPLEX128, False): np.complex128, (COMPLEX128, True): np.float64}[(prec, real)]
            if arr.dtype != want:
                raise ErrNotImplemented(f"{side} has element type {arr.dtype}, the plan needs {np.dtype(want)}")  # noqa: F821
        if not dst.flags.c_contiguous:
            raise ValueError("dst must be C-contiguous")
        src = np.ascontiguousarray(src)
        check(fn(self.h, dst.ctypes.data_as(ctypes.c_void_p), dst.size, src.ctypes.data_as(ctypes.c_void_p),
                 src.size))
        return dst

    def forward(self, dst, src):
        real = self.kind in (R2C, C2R)   # real plans: reals in, half spectrum out
        return self._host(lib().afft_plan_forward, dst, src, False, real)

    def inverse(self, dst, src):
        real = self.kind in (R2C, C2R)
        return self._host(lib().afft_plan_inverse, dst, src, real, False)

    def forward_full(self, dst, src):
        """PlanReal2D/3D.ForwardFull: n reals in, the full n-bin spectrum out."""
        return self._host(lib().afft_plan_forward_full, dst, src, False, True)

    def inverse_full(self, dst, src):
        """PlanReal2D/3D.InverseFull: the full n-bin spectrum in, n reals out."""
        return self._host(lib().afft_plan_inverse_full, dst, src, True, False)

    # -- device-resident execution -------------------------------------------
    def exec_device(self, direction, dst_ptr, src_ptr, stream=None):
        check(lib().afft_plan_exec_device(self.h, direction, ctypes.c_void_p(dst_ptr), ctypes.c_void_p(src_ptr),
                                          ctypes.c_void_p(stream) if stream else None))

    def exec_full_device(self, direction, dst_ptr, src_ptr, stream=None):
        check(lib().afft_plan_exec_full_device(self.h, direction, ctypes.c_void_p(dst_ptr),
                                               ctypes.c_void_p(src_ptr), ctypes.c_void_p(stream) if stream else None))

    def exec_buffers(self, direction, dst, src, stream=None):
        check(lib().afft_plan_exec_buffers(self.h, direction, dst.h, src.h, stream.h if stream else None))

    def clone(self):
        """Plan.Clone / NewExecutor: same description, own stream and workspace."""
        h = ctypes.c_void_p()
        check(lib().afft_plan_clone(self.h, ctypes.byref(h)))
        return Plan(h, self.desc)

    def close(self):
        if self.h:
            check(lib().afft_plan_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class FastPlan:
    """FastPlan (plan_fast.go): one pre-resolved launch per direction, device
    pointers, no validation and no lock."""

    def __init__(self, ctx, n, batch=1, precision=COMPLEX64, stream=None):
        h = ctypes.c_void_p()
        check(lib().afft_fast_plan_create(ctx.h, n, batch, precision, ctypes.c_void_p(stream) if stream else None,
                                          ctypes.byref(h)))
        self.h = h

    def __len__(self):
        return lib().afft_fast_plan_len(self.h)

    def forward(self, dst_ptr, src_ptr):
        lib().afft_fast_forward(self.h, ctypes.c_void_p(dst_ptr), ctypes.c_void_p(src_ptr))

    def inverse(self, dst_ptr, src_ptr):
        lib().afft_fast_inverse(self.h, ctypes.c_void_p(dst_ptr), ctypes.c_void_p(src_ptr))

    def close(self):
        if self.h:
            check(lib().afft_fast_plan_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Ragged:
    """A ragged batch: complex transforms of lengths `sizes` packed back to
    back, run by one persistent launch (exec), its captured graph
    (exec_graph), one batched plan per run of equal lengths (exec_per_size)
    or one plan execution per job (exec_per_job)."""

    def __init__(self, ctx, sizes, precision=COMPLEX64, flags=0):
        arr = (ctypes.c_int64 * len(sizes))(*[int(v) for v in sizes])
        h = ctypes.c_void_p()
        check(lib().afft_ragged_create(ctx.h, precision, arr, len(sizes), flags, ctypes.byref(h)))
        self.h = h
        self.sizes = [int(v) for v in sizes]

    def __len__(self):
        return lib().afft_ragged_total_len(self.h)

    def offsets(self):
        out = (ctypes.c_int64 * len(self.sizes))()
        check(lib().afft_ragged_offsets(self.h, out, len(self.sizes)))
        return list(out)

    def _x(self, fn, direction, dst_ptr, src_ptr, stream):
        check(fn(self.h, direction, ctypes.c_void_p(dst_ptr), ctypes.c_void_p(src_ptr),
                 ctypes.c_void_p(stream) if stream else None))

    def exec(self, direction, dst_ptr, src_ptr, stream=None):
        self._x(lib().afft_ragged_exec, direction, dst_ptr, src_ptr, stream)

    def exec_graph(self, direction, dst_ptr, src_ptr, stream=None):
        self._x(lib().afft_ragged_exec_graph, direction, dst_ptr, src_ptr, stream)

    def exec_per_size(self, direction, dst_ptr, src_ptr, stream=None):
        self._x(lib().afft_ragged_exec_per_size, direction, dst_ptr, src_ptr, stream)

    def exec_per_job(self, direction, dst_ptr, src_ptr, stream=None):
        self._x(lib().afft_ragged_exec_per_job, direction, dst_ptr, src_ptr, stream)

    def close(self):
        if self.h:
            check(lib().afft_ragged_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


CONV_LINEAR, CONV_CORRELATE, CONV_REAL = 0, 1, 2


def convolve(ctx, a, b, correlate=False):
    """Convolve / CrossCorrelate (convolve.go:9-59, correlate.go:13-33) of two
    complex host arrays, or ConvolveReal (convolve_real.go:9-80) of two real
    ones, on the device; returns len(a)+len(b)-1 values."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    real = a.dtype in (np.float32, np.float64)
    if real and correlate:
        raise ValueError("correlation is defined for complex sequences")
    prec = _np_prec(a.dtype)
    mode = CONV_REAL if real else (CONV_CORRELATE if correlate else CONV_LINEAR)
    out = np.empty(a.size + b.size - 1, dtype=a.dtype)
    check(lib().afft_convolve_host(ctx.h, prec, mode, out.ctypes.data_as(ctypes.c_void_p), out.size,
                                   a.ctypes.data_as(ctypes.c_void_p), a.size, b.ctypes.data_as(ctypes.c_void_p),
                                   b.size))
    return out
