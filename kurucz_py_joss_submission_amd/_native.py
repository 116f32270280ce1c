"""ctypes binding of libsynthe_hip.so (include/synthe_hip.h).

The library is the only compute path: if it is missing, or no GPU is
visible, every entry point raises NativeUnavailable.  There is no CPU
fallback.
"""

from __future__ import annotations

import ctypes as ct
import os
import threading
from pathlib import Path
from typing import Optional

import numpy as np

# SYNTHE_HIP_LIBRARY points at an alternative build of the same library.
LIB_PATH = Path(os.environ.get("SYNTHE_HIP_LIBRARY",
                               Path(__file__).resolve().parent / "libsynthe_hip.so"))

SYNTHE_OK = 0
SYMBOLS = (
    "synthe_last_error", "synthe_abi_version", "synthe_create", "synthe_destroy",
    "synthe_copy_tables", "synthe_wavelength_grid", "synthe_continuum", "synthe_transp",
    "synthe_asynth", "synthe_lte_spectrum", "synthe_josh_flux_batch", "synthe_source_function",
    "synthe_source_function_bfudge",
    "synthe_flux_to_lambda", "synthe_diag_div_pos", "synthe_set_kernel_timing",
    "synthe_kernel_times", "synthe_set_profile_precision", "synthe_batch_create", "synthe_batch_run",
    "synthe_batch_info", "synthe_batch_destroy", "synthe_asynth_vjp", "synthe_lte_vjp",
    "synthe_lte_spectrum_noalloc", "synthe_transp_asynth", "synthe_hydrogen_lines", "synthe_fort19_lines",
    "synthe_batch_arrays", "synthe_lte_vjp_batch",
)

SYNTHE_BATCH_AUTOTUNE = 1
SYNTHE_BATCH_GRAPH = 2

SYNTHE_NOALLOC_MAX_WL = 65536

SYNTHE_PREC_FP64 = 0
SYNTHE_PREC_FP32 = 1


class NativeUnavailable(RuntimeError):
    """libsynthe_hip.so is missing or no HIP device is usable."""


class SyntheError(RuntimeError):
    """A libsynthe_hip entry point returned a non-zero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


_VP = ct.c_void_p


class Lines(ct.Structure):
    _fields_ = [("n_lines", ct.c_long), ("wl", _VP), ("gf", _VP), ("cgf", _VP),
                ("gamma_rad", _VP), ("gamma_stark", _VP), ("gamma_vdw", _VP), ("elo", _VP),
                ("elem", _VP), ("ion", _VP), ("ltype", _VP), ("process", _VP),
                ("center_t", _VP), ("center_w", _VP)]


class Depths(ct.Structure):
    _fields_ = [("n_depths", ct.c_long), ("temperature", _VP), ("rhox", _VP), ("rho", _VP),
                ("xne", _VP), ("txnxn", _VP), ("hckt", _VP), ("pop", _VP), ("dop", _VP),
                ("n_ion", ct.c_long), ("n_elem", ct.c_long)]


class HLines(ct.Structure):
    """synthe_hlines (host arrays)."""
    _fields_ = [("n_lines", ct.c_int64), ("wavelength", _VP), ("center", _VP), ("gf", _VP), ("elo", _VP),
                ("n_lower", _VP), ("n_upper", _VP)]


class HDepths(ct.Structure):
    """synthe_hdepths (host arrays)."""
    _fields_ = [("n_depths", ct.c_int32), ("temperature", _VP), ("electron_density", _VP), ("mass_density", _VP),
                ("hckt", _VP), ("pop_h1", _VP), ("dop_h1", _VP), ("xnf_he1", _VP), ("xnf_h2", _VP),
                ("xnfph", _VP), ("dopph", _VP), ("turbulent_velocity", _VP)]


class F19Lines(ct.Structure):
    """synthe_f19lines (host arrays)."""
    _fields_ = [("n_lines", ct.c_int64), ("kind", _VP), ("wavelength", _VP), ("center", _VP), ("gf", _VP),
                ("gamma_rad", _VP), ("gamma_stark", _VP), ("gamma_vdw", _VP), ("elem", _VP), ("ion", _VP),
                ("nelion", _VP), ("nlast", _VP), ("cont_val", _VP), ("cat_wavelength", _VP), ("elo", _VP)]


class F19Depths(ct.Structure):
    """synthe_f19depths (host arrays)."""
    _fields_ = [("n_depths", ct.c_int32), ("n_ion", ct.c_int32), ("n_elem", ct.c_int32), ("pop", _VP),
                ("mass_density", _VP), ("electron_density", _VP), ("txnxn", _VP), ("temperature", _VP),
                ("emerge", _VP), ("emerge_h", _VP), ("boltz", _VP), ("hckt", _VP), ("doppler", _VP),
                ("vscale", _VP)]


class Instance(ct.Structure):
    """synthe_instance: one spectrum of a batch (include/synthe_hip.h)."""

    _fields_ = [("lines", Lines), ("depths", Depths), ("n_wl", ct.c_long), ("wl", _VP),
                ("n_edges", ct.c_long), ("wledge", _VP), ("half_edge", _VP), ("delta_edge", _VP),
                ("coeff_abs", _VP), ("coeff_scat", _VP), ("coeff_log10", ct.c_int),
                ("center_order", _VP), ("center_start", _VP), ("wing_order", _VP),
                ("cutoff", ct.c_double), ("flux", _VP), ("cont", _VP), ("iters", _VP)]


class BatchView(ct.Structure):
    """synthe_batch_view: device arrays of one batch instance."""
    _fields_ = [("cont_abs", _VP), ("cont_scat", _VP), ("kapmin", _VP), ("slinec", _VP), ("asynth", _VP),
                ("line_scat", _VP), ("transp", _VP), ("valid", _VP)]


class VjpInstance(ct.Structure):
    """synthe_vjp_instance: one spectrum of a batched RT vjp (device pointers)."""
    _fields_ = [("n_depths", ct.c_long), ("n_wl", ct.c_long)] + [
        (k, _VP) for k in ("wl", "temperature", "rhox", "cont_abs", "cont_scat", "line_abs", "line_scat",
                           "line_source", "v_flux", "g_cont_abs", "g_cont_scat", "g_line_abs", "g_planck",
                           "g_line_source")]


_lib = None
_lib_lock = threading.Lock()


def load_library(path: Optional[Path] = None):
    """Load the shared library and declare prototypes (no GPU needed)."""
    global _lib
    with _lib_lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise NativeUnavailable(f"{p} is not built: run python -m kurucz_py_joss_submission_amd.build")
        lib = ct.CDLL(str(p))
        for name in SYMBOLS:
            getattr(lib, name)  # raises AttributeError if an export is missing
        lib.synthe_last_error.restype = ct.c_char_p
        lib.synthe_abi_version.restype = ct.c_int
        lib.synthe_create.argtypes = [ct.POINTER(_VP)]
        lib.synthe_destroy.argtypes = [_VP]
        lib.synthe_copy_tables.argtypes = [_VP] + [_VP] * 5
        lib.synthe_wavelength_grid.restype = ct.c_long
        lib.synthe_wavelength_grid.argtypes = [ct.c_double, ct.c_double, ct.c_double, _VP, ct.c_long]
        lib.synthe_continuum.argtypes = [_VP, ct.c_long, ct.c_long, _VP, ct.c_long, _VP, _VP, _VP, _VP,
                                         _VP, ct.c_int, _VP, _VP, _VP]
        lib.synthe_transp.argtypes = [_VP, ct.POINTER(Lines), ct.POINTER(Depths), ct.c_long, _VP,
                                      ct.c_double, ct.c_double, _VP, _VP, _VP]
        lib.synthe_asynth.argtypes = [_VP, ct.POINTER(Lines), ct.POINTER(Depths), ct.c_long, _VP, _VP,
                                      _VP, _VP, ct.c_double, _VP, _VP, _VP, _VP, _VP, _VP]
        lib.synthe_transp_asynth.argtypes = [_VP, ct.POINTER(Lines), ct.POINTER(Depths), ct.c_long, _VP, _VP,
                                             ct.c_double, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]
        lib.synthe_hydrogen_lines.argtypes = [_VP, ct.POINTER(HLines), ct.POINTER(HDepths), ct.c_double,
                                              ct.c_double, ct.c_long, _VP, _VP, _VP, _VP, _VP]
        lib.synthe_fort19_lines.argtypes = [_VP, ct.POINTER(F19Lines), ct.POINTER(F19Depths), ct.c_double,
                                            ct.c_long, _VP, _VP, _VP, _VP, _VP]
        lib.synthe_lte_spectrum.argtypes = [_VP, ct.c_long, ct.c_long] + [_VP] * 12
        lib.synthe_lte_spectrum_noalloc.argtypes = [_VP, ct.c_long, ct.c_long] + [_VP] * 12
        lib.synthe_josh_flux_batch.argtypes = [_VP, ct.c_long, ct.c_long] + [_VP] * 10
        lib.synthe_source_function.argtypes = [_VP, ct.c_long, ct.c_long] + [_VP] * 7
        lib.synthe_source_function_bfudge.argtypes = [_VP, ct.c_long, ct.c_long] + [_VP] * 8
        lib.synthe_flux_to_lambda.argtypes = [_VP, ct.c_long, _VP, _VP, _VP]
        lib.synthe_diag_div_pos.argtypes = [ct.c_long, ct.c_int] + [_VP] * 5
        lib.synthe_set_kernel_timing.argtypes = [_VP, ct.c_int]
        lib.synthe_set_profile_precision.argtypes = [_VP, ct.c_int]
        lib.synthe_batch_create.argtypes = [_VP, ct.c_int, ct.POINTER(Instance), ct.c_int, ct.POINTER(_VP)]
        lib.synthe_batch_run.argtypes = [_VP, _VP]
        lib.synthe_batch_info.argtypes = [_VP, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), _VP, ct.c_int,
                                          ct.POINTER(ct.c_int), ct.POINTER(ct.c_double)]
        lib.synthe_batch_destroy.argtypes = [_VP]
        lib.synthe_batch_arrays.argtypes = [_VP, ct.c_int, ct.POINTER(BatchView)]
        lib.synthe_lte_vjp_batch.argtypes = [_VP, ct.c_int, ct.POINTER(VjpInstance), _VP]
        lib.synthe_asynth_vjp.argtypes = [_VP, ct.POINTER(Lines), ct.POINTER(Depths), ct.c_long, _VP, _VP, _VP,
                                          _VP, ct.c_double, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]
        lib.synthe_lte_vjp.argtypes = [_VP, ct.c_long, ct.c_long] + [_VP] * 15
        lib.synthe_kernel_times.argtypes = [_VP, _VP, ct.c_int, ct.POINTER(ct.c_int)]
        for name in SYMBOLS[2:]:
            if name != "synthe_wavelength_grid":
                getattr(lib, name).restype = ct.c_int
        if path is None:
            _lib = lib
        return lib


def check(fn: str, rc: int) -> None:
    if rc != SYNTHE_OK:
        msg = load_library().synthe_last_error().decode(errors="replace")
        raise SyntheError(fn, rc, msg)


class Context:
    """A libsynthe_hip context (device tables + workspace) on the current GPU."""

    def __init__(self):
        import torch

        if not torch.cuda.is_available():
            raise NativeUnavailable("no HIP device is visible: the synthe hot path needs a GPU")
        self.lib = load_library()
        torch.cuda.init()
        h = _VP()
        check("synthe_create", self.lib.synthe_create(ct.byref(h)))
        self.handle = h

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.synthe_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profile_precision(self, precision: str) -> None:
        """'fp64' (default) or 'fp32' far-wing profile terms for later asynth calls."""
        code = {"fp64": SYNTHE_PREC_FP64, "fp32": SYNTHE_PREC_FP32}[precision]
        check("synthe_set_profile_precision", self.lib.synthe_set_profile_precision(self.handle, code))

    def set_kernel_timing(self, enable: bool) -> None:
        check("synthe_set_kernel_timing", self.lib.synthe_set_kernel_timing(self.handle, int(enable)))

    def kernel_times(self) -> np.ndarray:
        """Durations (ms) of the wing-gather launches recorded since timing was enabled."""
        n = ct.c_int(0)
        check("synthe_kernel_times", self.lib.synthe_kernel_times(self.handle, None, 0, ct.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.float32)
        check("synthe_kernel_times", self.lib.synthe_kernel_times(self.handle, out.ctypes.data, n.value,
                                                                  ct.byref(n)))
        return out[:n.value].astype(np.float64)

    def tables(self):
        h0, h1, h2 = (np.zeros(2001) for _ in range(3))
        e, f = np.zeros(1001), np.zeros(1001)
        check("synthe_copy_tables", self.lib.synthe_copy_tables(
            self.handle, h0.ctypes.data, h1.ctypes.data, h2.ctypes.data, e.ctypes.data, f.ctypes.data))
        return h0, h1, h2, e, f


_ctx: Optional[Context] = None
_ctx_lock = threading.Lock()


def context() -> Context:
    global _ctx
    with _ctx_lock:
        if _ctx is None:
            _ctx = Context()
        return _ctx


def stream_handle():
    import torch

    return _VP(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> Optional[int]:
    """Device address of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def wavelength_grid(start: float, end: float, resolution: float) -> np.ndarray:
    """engine/opacity.py:526-547 through the library's host routine."""
    lib = load_library()
    n = lib.synthe_wavelength_grid(start, end, resolution, None, 0)
    if n < 0:
        check("synthe_wavelength_grid", int(n))
    out = np.zeros(int(n), dtype=np.float64)
    lib.synthe_wavelength_grid(start, end, resolution, out.ctypes.data, int(n))
    return out
