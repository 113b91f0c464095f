"""ctypes binding of `libsttopt_hip.so` (declarations: `include/sttopt_hip.h`).

This is the only module that touches the library.  It fails loudly: a missing or
unloadable library raises `NativeLibraryError`, and every product entry point
checks that its tensors live on a ROCm device -- there is no CPU path.
"""

from __future__ import annotations

import collections
import ctypes
import pathlib
import threading

import torch

from . import _build

LIB_PATH = _build.LIB_PATH

OK, EINVAL, EHIP, ENOCONV, ENOTSPD = 0, -1, -2, -3, -4
KEFF_FULL, KEFF_UPPER, KEFF_NODE = 0, 1, 2
SEPARATE_SWEEPS = 1  # sttopt_level.flags of level 0
LEVEL_SEPARATE = 2  # sttopt_level.flags: one pass per sweep at this level
TILE_VARIANT_SHIFT = 8  # sttopt_level.flags bits 8-11: the two passes' tile schedule
F64, F32 = 0, 1
MODE_MATVEC, MODE_OPERATOR, MODE_OPERATOR_PROJECTED = 0, 1, 2
MAX_LEVELS = 16


class NativeLibraryError(RuntimeError):
    """The HIP library is missing, failed to load, or returned an error."""


class DeviceError(RuntimeError):
    """A tensor handed to a HIP entry point is not on a ROCm device."""


c_int32, c_int64, c_double, c_void_p = (
    ctypes.c_int32,
    ctypes.c_int64,
    ctypes.c_double,
    ctypes.c_void_p,
)


class Level(ctypes.Structure):
    """`sttopt_level` (include/sttopt_hip.h)."""

    _fields_ = [
        ("nelx", c_int32),
        ("nely", c_int32),
        ("kx", c_int32),
        ("ky", c_int32),
        ("mask", c_void_p),
        ("density", c_void_p),
        ("density_stride", c_int64),
        ("rdiag", c_void_p),
        ("rdiag_stride", c_int64),
        ("keff", c_void_p),
        ("keff_stride", c_int64),
        ("keff_layout", c_int32),
        ("flags", c_int32),
        ("ptable", c_void_p),
        ("pkx", c_int32),
        ("pky", c_int32),
        ("invdiag", c_void_p),
        ("invdiag_stride", c_int64),
        ("cinv", c_void_p),
        ("cinv_stride", c_int64),
        ("vb", c_void_p),
        ("vx", c_void_p),
        ("vx2", c_void_p),
        ("vr", c_void_p),
        ("ve", c_void_p),
        ("vw", c_void_p),
    ]


class MG(ctypes.Structure):
    """`sttopt_mg` (include/sttopt_hip.h)."""

    _fields_ = [
        ("rows", c_int32),
        ("d", c_int32),
        ("nlevels", c_int32),
        ("vdtype", c_int32),
        ("n_smooth", c_int32),
        ("gamma", c_int32),
        ("omega", c_double),
        ("ke", c_double * 64),
        ("density64", c_void_p),
        ("density64_stride", c_int64),
        ("rdiag64", c_void_p),
        ("rdiag64_stride", c_int64),
        ("levels", Level * MAX_LEVELS),
        ("scratch", c_void_p),
        ("cinv_status", c_void_p),
        ("cinv_ready", c_void_p),
    ]


# name -> (restype, argtypes)
_P = c_void_p
_SIGNATURES = {
    "sttopt_last_error": (ctypes.c_char_p, []),
    "sttopt_version": (c_int32, []),
    "sttopt_abi_sizes": (c_int32, [_P, _P]),
    "sttopt_pack_mask": (c_int32, [_P, c_int64, _P, _P]),
    "sttopt_coarsen_mask": (c_int32, [_P, c_int32, c_int32, c_int32, c_int32, c_int32, _P, _P]),
    "sttopt_apply": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, c_int64,
         _P, _P, _P, c_int64, _P, c_int64, _P],
    ),
    "sttopt_diagonal": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, _P, _P, c_int64, _P,
         c_int64, _P],
    ),
    "sttopt_lumped_reaction": (
        c_int32, [c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, c_int64, _P]
    ),
    "sttopt_apply_keff": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, c_int64, _P, _P,
         c_int64, _P],
    ),
    "sttopt_mg_scratch_bytes": (c_int64, [_P]),
    "sttopt_galerkin": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, c_int64, _P,
         _P, _P, c_int64, _P, _P, c_int64, _P],
    ),
    "sttopt_keff_diagonal": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, c_int32, _P, c_int64,
         _P],
    ),
    "sttopt_param_diagonal": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, _P, _P,
         c_int64, _P],
    ),
    "sttopt_invert_diagonal": (
        c_int32, [c_int32, c_int64, c_int32, _P, c_int64, _P, _P, c_int64, _P]
    ),
    "sttopt_spd_inverse_work_count": (c_int64, [c_int32, c_int32]),
    "sttopt_pack_keff": (c_int32, [c_int32, c_int32, c_int64, c_int32, _P, c_int64, _P, c_int64, _P]),
    "sttopt_node_stencil": (c_int32, [c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64,
                                      _P, c_int64, _P]),
    "sttopt_spd_inverse": (c_int32, [c_int32, c_int32, _P, c_int64, _P, _P]),
    "sttopt_spd_inverse_band": (c_int32, [c_int32, c_int32, c_int32, _P, c_int64, _P, _P]),
    "sttopt_spd_inverse_band_async": (
        c_int32, [c_int32, c_int32, c_int32, _P, c_int64, _P, _P, _P]),
    "sttopt_dense_assemble": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, c_int64, _P, c_int64, _P, _P,
         _P, c_int64, _P],
    ),
    "sttopt_restrict": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, _P,
         c_int64, _P],
    ),
    "sttopt_prolong_add": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, _P,
         c_int64, _P],
    ),
    "sttopt_vcycle": (c_int32, [_P, _P, c_int64, _P, c_int64, _P]),
    "sttopt_level_apply": (c_int32, [_P, c_int32, _P, c_int64, _P, c_int64, _P]),
    "sttopt_mgpcg": (
        c_int32,
        [_P, _P, c_int64, _P, c_int64, c_double, c_int32, c_int32, c_int32, c_int32, _P, _P,
         _P, _P, _P],
    ),
    "sttopt_mgpcg_work_count": (c_int64, [_P]),
    "sttopt_mgpcg_batch": (c_int32, [_P, c_int32, c_double, c_int32, c_int32, c_int32, _P, _P]),
    "sttopt_adjoint_grad": (
        c_int32,
        [c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, c_int64, _P, _P, c_int64, _P,
         c_int64, _P],
    ),
    "sttopt_strain_energy": (
        c_int32, [c_int32, c_int32, c_int32, c_int32, _P, c_int64, _P, _P, c_int64, _P]
    ),
    "sttopt_stage_prologue": (
        c_int32,
        [c_int32, c_int32, c_int32, _P, _P, _P, c_double, c_double, c_double, c_double,
         c_double, _P, _P, _P, _P],
    ),
    "sttopt_profile_outer_apply": (c_int32, [c_int32]),
    "sttopt_profile_outer_apply_read": (c_int32, [_P, _P, _P]),
    "sttopt_gauss_gradients": (
        c_int32,
        [c_int32, c_int32, c_int32, _P, _P, c_int64, _P, c_int64, c_double, c_int32, _P,
         c_int64, _P],
    ),
    "sttopt_divergence_load": (
        c_int32,
        [c_int32, c_int32, c_int32, _P, _P, c_int64, _P, c_int64, c_double, _P, c_int64, _P],
    ),
    "sttopt_gauss_dot": (
        c_int32,
        [c_int32, c_int32, c_int32, _P, _P, c_int64, _P, c_int64, c_double, _P, c_int64, _P],
    ),
    "sttopt_direction_field_backward": (
        c_int32,
        [c_int32, c_int32, c_int32, _P, _P, c_int64, _P, c_int64, _P, c_int64, _P],
    ),
    "sttopt_density_filter": (
        c_int32,
        [c_int32, c_int32, c_int32, c_double, _P, c_int64, _P, c_int32, _P, c_int64, _P],
    ),
    "sttopt_csr_matvec": (
        c_int32,
        [c_int64, _P, _P, _P, c_int32, _P, c_int64, _P, c_int32, _P, c_int64, _P],
    ),
    "sttopt_heaviside": (c_int32, [c_int64, c_double, c_double, c_int32, _P, _P, _P, _P]),
    "sttopt_mma_prepare_work_count": (c_int64, [c_int32, c_int32]),
    "sttopt_mma_prepare": (
        c_int32,
        [c_int32, c_int32, c_int32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, c_int64, _P,
         c_int64, _P, c_int64, _P, c_int64, c_double, c_double, c_double, c_double, _P, _P,
         _P, _P, _P, _P, _P, _P, _P],
    ),
    "sttopt_mma_work_count": (c_int64, [c_int32, c_int32]),
    "sttopt_mma_state_offsets": (c_int32, [c_int32, c_int32, _P]),
    "sttopt_mma_subsolv": (
        c_int32,
        [c_int32, c_int32, _P, _P, _P, _P, _P, _P, _P, _P, c_double, _P, _P, _P, _P, _P,
         c_int32, c_double, c_int32, _P],
    ),
    "sttopt_stage_epilogue": (
        c_int32,
        [c_int32, c_int32, c_int32, _P, _P, _P, c_double, c_double, c_double, c_double,
         c_double, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    ),
    "sttopt_stage_scratch_count": (c_int64, [c_int32, c_int32, c_int32]),
    "sttopt_mgpcg_graphs": (c_int32, [c_int32]),
    "sttopt_vcycle_bottom": (c_int32, [c_int32]),
    "sttopt_stage_epilogue_noalloc": (
        c_int32,
        [c_int32, c_int32, c_int32, _P, _P, _P, c_double, c_double, c_double, c_double,
         c_double, _P, _P, _P, _P, _P, _P, _P, _P],
    ),
}

_lib: ctypes.CDLL | None = None


def exported_symbols() -> list[str]:
    """Every entry point the binding declares (all of include/sttopt_hip.h)."""
    return sorted(_SIGNATURES)


def load() -> ctypes.CDLL:
    """The loaded library (loads it on first use).

    :raises NativeLibraryError: if the library is not built or fails to load
    """
    global _lib
    if _lib is not None:
        return _lib
    path = pathlib.Path(LIB_PATH)
    if not path.exists():
        raise NativeLibraryError(
            f"{path} is not built; run __graft_entry__.build() (hipcc, gfx950) first"
        )
    try:
        lib = ctypes.CDLL(str(path))
    except OSError as exc:  # pragma: no cover - depends on the environment
        raise NativeLibraryError(f"cannot load {path}: {exc}") from exc
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().sttopt_last_error().decode(errors="replace")


def check(status: int, what: str) -> None:
    """Raise `NativeLibraryError` for a non-OK status."""
    if status != OK:
        raise NativeLibraryError(f"{what} failed (status {status}): {last_error()}")


def require_device(*tensors) -> torch.device:
    """The common ROCm device of `tensors` (None entries skipped).

    :raises DeviceError: if any tensor is not on a ROCm/CUDA device
    """
    dev = None
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            where = t.device if isinstance(t, torch.Tensor) else type(t).__name__
            raise DeviceError(
                "sttopt_amd runs on the GPU through its HIP kernels only; got a tensor on "
                f"{where}. Move inputs to a ROCm device (there is no CPU path)."
            )
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise DeviceError(f"tensors on different devices: {dev} and {t.device}")
    if dev is None:
        raise DeviceError("no tensor given to infer the device from")
    return dev


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def stream_handle(device: torch.device) -> int:
    """The raw HIP stream torch is currently using on `device`."""
    return torch.cuda.current_stream(device).cuda_stream


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float64:
        return F64
    if dtype == torch.float32:
        return F32
    raise TypeError(f"unsupported dtype {dtype}; the HIP kernels take float64 or float32")


_KE_HOST: "collections.OrderedDict" = collections.OrderedDict()
_KE_HOST_LOCK = threading.Lock()


def ke_array(KE: torch.Tensor) -> ctypes.Array:
    """The element matrix as a host double[64] (row-major, zero padded).  Cached per
    storage, layout and version (reading a device KE back synchronizes the stream), so
    `KE.detach()` and other aliases of the same memory hit the same entry; an entry
    keeps its KE, and with it the memory, alive, so the address cannot be reused while
    cached.  Callers only read the array."""
    key = (str(KE.device), KE.data_ptr(), tuple(KE.shape), tuple(KE.stride()), KE.dtype,
           KE._version)
    with _KE_HOST_LOCK:
        hit = _KE_HOST.get(key)
        if hit is not None:
            return hit[1]
    flat = KE.detach().to("cpu", torch.float64).reshape(-1).tolist()
    if len(flat) not in (16, 64):
        raise ValueError(f"KE must be 4x4 or 8x8, got {tuple(KE.shape)}")
    arr = (c_double * 64)()
    for i, v in enumerate(flat):
        arr[i] = v
    with _KE_HOST_LOCK:
        _KE_HOST[key] = (KE, arr)
        while len(_KE_HOST) > 16:
            _KE_HOST.popitem(last=False)
    return arr
