"""ctypes bindings of the C-ABI libraries ``libebsd_radon.so`` (include/ebsd_radon.h) and
``libebsd_nlpar.so`` (include/ebsd_nlpar.h).

There is deliberately no fallback: if a library is missing or fails to load, every
entry point raises ``HipLibraryError``.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libebsd_radon.so")
NLPAR_LIB_PATH = os.path.join(HERE, "libebsd_nlpar.so")
TRIPVOTE_LIB_PATH = os.path.join(HERE, "libebsd_tripvote.so")

EBSD_U8, EBSD_U16, EBSD_F32 = 0, 1, 2

#: exported symbols, in declaration order of include/ebsd_radon.h
SYMBOLS = (
    "ebsd_plan_create", "ebsd_plan_destroy", "ebsd_plan_set_background", "ebsd_plan_set_tuning",
    "ebsd_plan_set_stage_pitch", "ebsd_plan_set_row_segments", "ebsd_plan_set_peak_kernel", "ebsd_plan_set_peak_block", "ebsd_plan_set_line_unit", "ebsd_workspace_bytes", "ebsd_find_bands", "ebsd_find_bands_noalloc", "ebsd_radon", "ebsd_peaks", "ebsd_conv",
    "ebsd_plan_info", "ebsd_last_error", "ebsd_batch_workspace_bytes", "ebsd_batch_create",
    "ebsd_batch_run", "ebsd_batch_graph_launch", "ebsd_batch_destroy", "ebsd_band_grad",
    "ebsd_conv_transpose", "ebsd_backproject", "ebsd_gnomonic_bands", "ebsd_radon2pole",
    "ebsd_pack_band_records",
)

#: exported symbols of the NLPAR library, in declaration order of include/ebsd_nlpar.h
NLPAR_SYMBOLS = ("ebsd_nlpar_max", "ebsd_nlpar_workspace_bytes", "ebsd_nlpar_sigma", "ebsd_nlpar",
                 "ebsd_nlpar_set_tiled", "ebsd_nlpar_pair_bytes", "ebsd_nlpar_min", "ebsd_nlpar_prepare", "ebsd_nlpar_colsum",
                 "ebsd_nlpar_finish", "ebsd_nlpar_last_error")

#: exported symbols of the triplet-vote library, in declaration order of include/ebsd_tripvote.h
TRIPVOTE_SYMBOLS = ("ebsd_tripvote", "ebsd_bandindex", "ebsd_tripvote_last_error")


class HipLibraryError(RuntimeError):
    pass


class EbsdBands(ctypes.Structure):
    """Mirror of ``ebsd_bands`` (device pointers of the output fields)."""
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "d_max", "d_normmax", "d_avemax", "d_maxloc", "d_aveloc", "d_width", "d_theta", "d_rho",
        "d_valid", "d_stats")]


_lib = None
_nlpar_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH):
    """Load the library once; raise HipLibraryError if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise HipLibraryError(
                f"HIP library not found at {path}; build it with `python -m pyebsdindex_amd.build`")
        try:
            L = ctypes.CDLL(path)
        except OSError as e:
            raise HipLibraryError(f"failed to load {path}: {e}") from e
        P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        L.ebsd_plan_create.argtypes = [I, I, I, I, P, I, P, P, I, I, I, I, I, I, I, P, P,
                                       ctypes.POINTER(ctypes.c_void_p)]
        L.ebsd_plan_destroy.argtypes = [P]
        L.ebsd_plan_set_background.argtypes = [P, P, I]
        L.ebsd_plan_set_tuning.argtypes = [P, I, I, I, I]
        L.ebsd_plan_set_peak_kernel.argtypes = [P, I]
        L.ebsd_plan_set_peak_block.argtypes = [P, I]
        L.ebsd_plan_set_line_unit.argtypes = [P, I]
        L.ebsd_plan_set_stage_pitch.argtypes = [P, I]
        L.ebsd_plan_set_row_segments.argtypes = [P, I]
        L.ebsd_workspace_bytes.argtypes = [P, I64]
        L.ebsd_workspace_bytes.restype = ctypes.c_size_t
        L.ebsd_find_bands.argtypes = [P, P, I, I64, ctypes.POINTER(EbsdBands), P, ctypes.c_size_t, P]
        L.ebsd_find_bands_noalloc.argtypes = [P, P, I, I64, ctypes.POINTER(EbsdBands), P]
        L.ebsd_radon.argtypes = [P, P, I, I64, P, P]
        L.ebsd_peaks.argtypes = [P, P, I64, ctypes.POINTER(EbsdBands), P]
        L.ebsd_conv.argtypes = [P, P, I64, P, P]
        L.ebsd_plan_info.argtypes = [P, ctypes.POINTER(ctypes.c_int64)]
        L.ebsd_batch_workspace_bytes.argtypes = [I, P, P]
        L.ebsd_batch_workspace_bytes.restype = ctypes.c_size_t
        L.ebsd_batch_create.argtypes = [I, P, P, P, P, P, P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
        L.ebsd_batch_run.argtypes = [P, P]
        L.ebsd_batch_graph_launch.argtypes = [P, P]
        L.ebsd_batch_destroy.argtypes = [P]
        L.ebsd_band_grad.argtypes = [P, P, I64, P, P, P, P, P, P, P]
        L.ebsd_conv_transpose.argtypes = [P, P, I64, P, P]
        L.ebsd_backproject.argtypes = [P, P, I64, P, P]
        L.ebsd_gnomonic_bands.argtypes = [I64, I, P, P, I, I, I, I, P, P, P, P]
        L.ebsd_radon2pole.argtypes = [I64, I, P, P, P, I, I, I, I, P, P, P]
        L.ebsd_pack_band_records.argtypes = [ctypes.POINTER(EbsdBands), I64, I, I, P, P]
        L.ebsd_last_error.argtypes = []
        L.ebsd_last_error.restype = ctypes.c_char_p
        for name in ("ebsd_plan_create", "ebsd_plan_destroy", "ebsd_plan_set_background",
                     "ebsd_plan_set_tuning", "ebsd_plan_set_peak_kernel", "ebsd_plan_set_peak_block", "ebsd_plan_set_line_unit", "ebsd_plan_set_stage_pitch",
                     "ebsd_plan_set_row_segments", "ebsd_find_bands", "ebsd_find_bands_noalloc",
                     "ebsd_radon", "ebsd_peaks",
                     "ebsd_conv", "ebsd_plan_info", "ebsd_batch_create", "ebsd_batch_run",
                     "ebsd_batch_graph_launch", "ebsd_batch_destroy",
                     "ebsd_band_grad", "ebsd_conv_transpose", "ebsd_backproject", "ebsd_gnomonic_bands",
                     "ebsd_radon2pole", "ebsd_pack_band_records"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
        return L


def check(rc: int):
    if rc != 0:
        msg = load().ebsd_last_error().decode(errors="replace")
        raise HipLibraryError(f"ebsd library error {rc}: {msg}")


def load_nlpar(path: str = NLPAR_LIB_PATH):
    """Load the NLPAR library once; raise HipLibraryError if it is not built."""
    global _nlpar_lib
    with _lock:
        if _nlpar_lib is not None:
            return _nlpar_lib
        if not os.path.exists(path):
            raise HipLibraryError(
                f"HIP library not found at {path}; build it with `python -m pyebsdindex_amd.build`")
        try:
            L = ctypes.CDLL(path)
        except OSError as e:
            raise HipLibraryError(f"failed to load {path}: {e}") from e
        P, I, I64, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_float
        L.ebsd_nlpar_max.argtypes = [P, I, I64, I64, P, P]
        L.ebsd_nlpar_workspace_bytes.argtypes = [I64, I64]
        L.ebsd_nlpar_workspace_bytes.restype = ctypes.c_size_t
        L.ebsd_nlpar_sigma.argtypes = [P, I, I, I, I64, P, I, I, F, P, P, P, P, ctypes.c_size_t, P]
        L.ebsd_nlpar.argtypes = [P, I, I, I, I64, P, I, P, I, F, F, F, F, I, I, I, I, P, I, P, P, ctypes.c_size_t, P]
        L.ebsd_nlpar_pair_bytes.argtypes = [I64, I]
        L.ebsd_nlpar_pair_bytes.restype = ctypes.c_size_t
        L.ebsd_nlpar_set_tiled.argtypes = [I]
        L.ebsd_nlpar_set_tiled.restype = ctypes.c_int
        L.ebsd_nlpar_min.argtypes = [P, I, I64, I64, P, P]
        L.ebsd_nlpar_prepare.argtypes = [P, I, I64, I64, I, F, P, P, P]
        L.ebsd_nlpar_colsum.argtypes = [P, I64, I64, P, P]
        L.ebsd_nlpar_finish.argtypes = [P, I64, I64, I, F, F, P, I, P]
        for name in ("ebsd_nlpar_max", "ebsd_nlpar_sigma", "ebsd_nlpar", "ebsd_nlpar_min", "ebsd_nlpar_prepare",
                     "ebsd_nlpar_colsum", "ebsd_nlpar_finish"):
            getattr(L, name).restype = ctypes.c_int
        L.ebsd_nlpar_last_error.argtypes = []
        L.ebsd_nlpar_last_error.restype = ctypes.c_char_p
        _nlpar_lib = L
        return L


def check_nlpar(rc: int):
    if rc != 0:
        msg = load_nlpar().ebsd_nlpar_last_error().decode(errors="replace")
        raise HipLibraryError(f"ebsd NLPAR library error {rc}: {msg}")


_tripvote_lib = None


def load_tripvote(path: str = TRIPVOTE_LIB_PATH):
    """Load the triplet-vote library once; raise HipLibraryError if it is not built."""
    global _tripvote_lib
    with _lock:
        if _tripvote_lib is not None:
            return _tripvote_lib
        if not os.path.exists(path):
            raise HipLibraryError(
                f"HIP library not found at {path}; build it with `python -m pyebsdindex_amd.build`")
        try:
            L = ctypes.CDLL(path)
        except OSError as e:
            raise HipLibraryError(f"failed to load {path}: {e}") from e
        P, I = ctypes.c_void_p, ctypes.c_int
        L.ebsd_tripvote.argtypes = [P, P, ctypes.c_int64, I, P, P, I, I, ctypes.c_double, P, P, I,
                                    ctypes.c_double, P, P, P, P, P, P]
        L.ebsd_tripvote.restype = ctypes.c_int
        L.ebsd_bandindex.argtypes = [P, P, ctypes.c_int64, I, P, P, P, P, I, P, P, P, I, P, P, P, ctypes.c_double,
                                     I, I, I, P, P, P, P, P, P, P, P, P, P, P, P]
        L.ebsd_bandindex.restype = ctypes.c_int
        L.ebsd_tripvote_last_error.argtypes = []
        L.ebsd_tripvote_last_error.restype = ctypes.c_char_p
        _tripvote_lib = L
        return L


def check_tripvote(lib, rc: int):
    if rc != 0:
        msg = lib.ebsd_tripvote_last_error().decode(errors="replace")
        raise HipLibraryError(f"ebsd triplet-vote library error {rc}: {msg}")


def launch_stream(stream=None, device=None):
    """The stream a wrapper launches on: the caller's current stream, or ``stream``.  A
    side stream first waits for the work already queued on the current stream (which may
    have produced the inputs); the results are then ready in ``stream``'s order, and the
    caller synchronises with it before reading them elsewhere."""
    import torch
    cur = torch.cuda.current_stream(device)
    if stream is None or stream == cur:
        return cur
    stream.wait_stream(cur)
    return stream


def keep_for_stream(stream, *tensors):
    """Tell the caching allocator that ``tensors`` (made on the current stream) are used on
    ``stream``, so their memory is not reused before ``stream`` has passed this point."""
    import torch
    if stream is None or stream == torch.cuda.current_stream(stream.device):
        return
    for t in tensors:
        if t is not None and hasattr(t, "record_stream") and t.is_cuda:
            t.record_stream(stream)
