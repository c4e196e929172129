"""ctbangbang_amd — MI355X-native reconstruction hot path of CTBangBang.

Python mirror of the reference program's stage interface over the C-ABI of
``libctbb_hip.so`` (declared in ``include/ctbb_hip.h``).  The reference's
stages are ``adaptive_filter_kk``, ``rebin_filter``, ``backproject`` and
``finalize_image_stack`` (reference include/preprocessing.h:26,
include/rebin_filter.h:27, include/backproject.h:26,
include/finalize_image_stack.h:26), all taking the ``recon_metadata``
struct; the same names are provided here with the same argument and the
reference's error behaviour mapped to :class:`CTBBError`.

There is no CPU fallback: every compute call goes to the HIP kernels, and
importing the stage functions fails loudly when the library has not been
built (``python -m ctbangbang_amd.build``).
"""
import ctypes as C
import os

import numpy as np

from .structs import (BlockInfo, CtData, CtGeom, Flags, Options, PhantomParams, ReconInfo, ReconMetadata,
                      ReconParams, Sizes, check_layout)

__all__ = [
    "CTBBError", "lib", "library_path", "ReconMetadata", "ReconParams", "CtGeom", "ReconInfo", "Options",
    "PhantomParams", "Sizes", "Context", "write_prm", "configure_recon_params", "configure_ct_geom",
    "plan_reconstruction", "plan_block", "block_sizes", "n_slices_final", "load_filter", "ramp_filter",
    "metadata_from_prm", "synth_scan", "free_metadata", "FILTER_F32", "FILTER_BF16X3", "FILTER_F16X3", "FILTER_FFT32",
    "PROJ_F32", "PROJ_BF16", "adaptive_filter_kk", "rebin_filter", "backproject", "finalize_image_stack",
]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
FILTER_F32, FILTER_BF16X3, FILTER_F16X3, FILTER_FFT32 = 0, 1, 2, 3
PROJ_F32, PROJ_BF16 = 0, 1
BATCH_PERSISTENT, BATCH_PIPELINE = 0, 1
NOALLOC_MAX_VIEWS = 16384      # CTBB_NOALLOC_MAX_VIEWS of include/ctbb_hip.h

_STATUS = {
    0: "success", 1: "invalid argument", 2: "file missing, unreadable or unwritable",
    3: "essential prm value missing", 4: "requested reconstruction outside allowed data range",
    5: "could not parse selected scanner", 6: "filter file not found", 7: "geometry not supported by the kernels",
    8: "HIP runtime error", 9: "allocation failed", 10: "raw file type not supported by this build",
    11: "input beyond the fixed capacity of a no-allocation entry point",
    12: "raw data holds NaN or infinite samples",
}


class CTBBError(RuntimeError):
    """A nonzero ctbb_status; ``code`` is the status value."""

    def __init__(self, code, what=""):
        self.code = int(code)
        super().__init__(f"{what}: {_STATUS.get(self.code, 'unknown status')} (status {self.code})")


def library_path():
    return os.path.join(PKG_DIR, "libctbb_hip.so")


_LIB = None
_P = C.POINTER
_MR = _P(ReconMetadata)
_VP = C.c_void_p

_SIGS = {
    "ctbb_strerror": (C.c_char_p, [C.c_int]),
    "ctbb_parse_config": (C.c_int, [C.c_char_p, _P(ReconParams)]),
    "ctbb_empty_config": (C.c_int, [C.c_char_p]),
    "ctbb_configure_recon_params": (C.c_int, [C.c_char_p, _P(ReconParams)]),
    "ctbb_prm_mu_water": (C.c_int, [C.c_char_p, _P(C.c_float)]),
    "ctbb_prm_float": (C.c_int, [C.c_char_p, C.c_char_p, _P(C.c_float)]),
    "ctbb_configure_paths": (C.c_int, [_MR]),
    "ctbb_configure_ct_geom": (C.c_int, [_MR, _P(CtGeom)]),
    "ctbb_configure_reconstruction": (C.c_int, [_MR]),
    "ctbb_plan_reconstruction": (C.c_int, [_MR]),
    "ctbb_update_block_info": (C.c_int, [_MR]),
    "ctbb_plan_block": (C.c_int, [_MR, C.c_int]),
    "ctbb_extract_projections": (C.c_int, [_MR]),
    "ctbb_finish_and_cleanup": (C.c_int, [_MR]),
    "ctbb_array_search": (C.c_int, [C.c_float, _P(C.c_double), C.c_int, C.c_int]),
    "ctbb_n_slices_final": (C.c_int, [_MR]),
    "ctbb_load_filter": (C.c_int, [_MR, _P(C.c_float)]),
    "ctbb_ramp_filter": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, _P(C.c_float)]),
    "ctbb_block_sizes": (C.c_int, [_MR, _P(Sizes)]),
    "ctbb_ctx_create": (C.c_int, [C.c_int, _P(_VP)]),
    "ctbb_ctx_destroy": (None, [_VP]),
    "ctbb_ctx_get_options": (C.c_int, [_VP, _P(Options)]),
    "ctbb_ctx_set_options": (C.c_int, [_VP, _P(Options)]),
    "ctbb_ctx_stream": (_VP, [_VP]),
    "ctbb_ctx_synchronize": (C.c_int, [_VP]),
    "ctbb_ctx_kernel_time": (C.c_int, [_VP, C.c_char_p, _P(C.c_float), _P(C.c_int)]),
    "ctbb_adaptive_filter": (C.c_int, [_VP, _MR]),
    "ctbb_rebin_filter": (C.c_int, [_VP, _MR]),
    "ctbb_backproject": (C.c_int, [_VP, _MR]),
    "ctbb_finalize_image_stack": (C.c_int, [_VP, _MR]),
    "ctbb_adaptive_filter_dev": (C.c_int, [_VP, _MR, _VP, _VP]),
    "ctbb_rebin_filter_dev": (C.c_int, [_VP, _MR, _VP, _VP]),
    "ctbb_rebin_dev": (C.c_int, [_VP, _MR, _VP, _VP]),
    "ctbb_filter_rows_dev": (C.c_int, [_VP, _VP, C.c_int, C.c_size_t, _VP, _VP]),
    "ctbb_backproject_dev": (C.c_int, [_VP, _MR, _VP, _VP]),
    "ctbb_backproject_dev_noalloc": (C.c_int, [_VP, _MR, _VP, _VP]),
    "ctbb_finalize_dev": (C.c_int, [_VP, _MR, _VP, _VP]),
    "ctbb_autotune_backprojection": (C.c_int, [_VP, _MR, C.c_char_p, _P(C.c_int)]),
    "ctbb_hu_to_int16_dev": (C.c_int, [_VP, _VP, C.c_size_t, C.c_float, C.c_float, _VP]),
    "ctbb_write_dicom_series": (C.c_int, [_MR, _VP, C.c_float, C.c_float, C.c_char_p, C.c_char_p]),
    "ctbb_reconstruct_dev": (C.c_int, [_VP, _MR, _VP, C.c_size_t, _VP, _VP]),
    "ctbb_reconstruct_host": (C.c_int, [_VP, _MR, _VP, C.c_size_t, _VP]),
    "ctbb_reconstruct_batch_dev": (C.c_int, [_VP, C.c_int, C.POINTER(_MR), C.POINTER(_VP), C.POINTER(C.c_size_t),
                                             C.POINTER(_VP)]),
    "ctbb_rebin_filter_cpu": (C.c_int, [_MR, C.c_int]),
    "ctbb_backproject_cpu": (C.c_int, [_MR, C.c_int]),
    "ctbb_finalize_image_stack_cpu": (C.c_int, [_MR, C.c_float, C.c_int]),
    "ctbb_plan_batch_dev": (C.c_int, [_VP, C.c_int, C.POINTER(_MR), _VP, C.c_int]),
    "ctbb_synth_scan": (C.c_int, [_MR, _P(PhantomParams), C.c_long, C.c_long, _P(C.c_float)]),
    "ctbb_synth_scan_dev": (C.c_int, [_VP, _MR, _P(PhantomParams), C.c_long, C.c_long, _VP]),
}

EXPORTED = tuple(_SIGS)


def lib():
    """Load libctbb_hip.so (in-tree build); raises if it has not been built."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -m ctbangbang_amd.build` "
                              "(there is no fallback implementation)")
        check_layout()
        l = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _LIB = l
    return _LIB


def _check(st, what):
    if st != 0:
        raise CTBBError(st, what)


def _ptr(x):
    """Device/host address of a torch tensor, numpy array or int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    raise TypeError(f"cannot take the address of {type(x)}")


# --------------------------------------------------------------------------
# prm files and planning


def write_prm(path, tokens):
    """Write a prm file from a dict {token_without_colon: value}."""
    with open(path, "w") as f:
        for k, v in tokens.items():
            f.write(f"{k}:\t{v}\n")
    return path


def configure_recon_params(prm):
    rp = ReconParams()
    _check(lib().ctbb_configure_recon_params(str(prm).encode(), C.byref(rp)), "configure_recon_params")
    return rp


def configure_ct_geom(mr):
    cg = CtGeom()
    _check(lib().ctbb_configure_ct_geom(C.byref(mr), C.byref(cg)), "configure_ct_geom")
    mr.cg = cg
    return cg


def plan_reconstruction(mr):
    _check(lib().ctbb_plan_reconstruction(C.byref(mr)), "plan_reconstruction")


def plan_block(mr, block_idx):
    _check(lib().ctbb_plan_block(C.byref(mr), int(block_idx)), "plan_block")


def pull_start(mr):
    """First reading of the planned block's pull as a signed number.  The
    reference keeps it in a size_t (recon_structs.h), so a pull that starts
    before reading 0 (possible with TableDir in near the start of the data)
    wraps; the device paths zero-pad those readings."""
    v = int(mr.ri.idx_pull_start) & ((1 << 64) - 1)
    return v - (1 << 64) if v >= (1 << 63) else v


def block_sizes(mr):
    s = Sizes()
    _check(lib().ctbb_block_sizes(C.byref(mr), C.byref(s)), "block_sizes")
    return s


def n_slices_final(mr):
    return int(lib().ctbb_n_slices_final(C.byref(mr)))


def write_dicom_series(mr, stack_i16, directory, prefix="slice", slope=1.0, intercept=0.0):
    """Final stack of int16 stored values ([n_slices_final][nx][ny], host
    memory) as a DICOM CT series <directory>/<prefix>_NNNN.dcm."""
    a = np.ascontiguousarray(stack_i16, dtype=np.int16)
    if a.size != n_slices_final(mr) * mr.rp.nx * mr.rp.ny:
        raise ValueError("stack size does not match the final stack of mr")
    _check(lib().ctbb_write_dicom_series(C.byref(mr), a.ctypes.data, float(slope), float(intercept),
                                         os.fsencode(directory), os.fsencode(prefix)), "write_dicom_series")


def load_filter(mr):
    n = 2 * mr.cg.n_channels_oversampled
    out = np.zeros(n, dtype=np.float32)
    _check(lib().ctbb_load_filter(C.byref(mr), out.ctypes.data_as(_P(C.c_float))), "load_filter")
    return out


def ramp_filter(n, ds, c, a):
    out = np.zeros(2 * n, dtype=np.float32)
    _check(lib().ctbb_ramp_filter(int(n), float(ds), float(c), float(a), out.ctypes.data_as(_P(C.c_float))),
           "ramp_filter")
    return out


def metadata_from_prm(prm, install_dir=None, plan=True):
    """recon_metadata for a prm file: parse + geometry (+ plan, without
    touching the raw file).  install_dir defaults to the package directory."""
    mr = ReconMetadata()
    mr.rp = configure_recon_params(prm)
    mr.install_dir = (install_dir or PKG_DIR).encode()
    configure_ct_geom(mr)
    if plan:
        plan_reconstruction(mr)
    return mr


def free_metadata(mr):
    """Release the C allocations held by mr (tube angles, table positions, ...)."""
    mr.ctd.final_image_stack = None
    lib().ctbb_finish_and_cleanup(C.byref(mr))


def synth_scan(mr, pp, first, count):
    """Host evaluation of the synthetic scan, [count][n_rows_raw][n_channels] float32."""
    out = np.zeros((count, mr.cg.n_rows_raw, mr.cg.n_channels), dtype=np.float32)
    _check(lib().ctbb_synth_scan(C.byref(mr), C.byref(pp), int(first), int(count),
                                 out.ctypes.data_as(_P(C.c_float))), "synth_scan")
    return out


# --------------------------------------------------------------------------
# GPU context


class Context:
    """A ctbb_ctx: one HIP stream on one device plus its workspaces."""

    def __init__(self, device=0):
        h = _VP()
        _check(lib().ctbb_ctx_create(int(device), C.byref(h)), "ctx_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().ctbb_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def options(self):
        o = Options()
        _check(lib().ctbb_ctx_get_options(self._h, C.byref(o)), "get_options")
        return o

    def set_options(self, **kw):
        o = self.options()
        for k, v in kw.items():
            setattr(o, k, v)
        _check(lib().ctbb_ctx_set_options(self._h, C.byref(o)), "set_options")
        return o

    def stream(self):
        return lib().ctbb_ctx_stream(self._h)

    def synchronize(self):
        _check(lib().ctbb_ctx_synchronize(self._h), "synchronize")

    def kernel_time(self, family):
        ms = C.c_float()
        n = C.c_int()
        _check(lib().ctbb_ctx_kernel_time(self._h, family.encode(), C.byref(ms), C.byref(n)), "kernel_time")
        return float(ms.value), int(n.value)

    # device-resident stages (pointers: torch tensors or raw addresses)
    def adaptive_filter_dev(self, mr, d_raw, d_out):
        _check(lib().ctbb_adaptive_filter_dev(self._h, C.byref(mr), _ptr(d_raw), _ptr(d_out)), "adaptive_filter_dev")

    def rebin_dev(self, mr, d_raw, d_rebin):
        _check(lib().ctbb_rebin_dev(self._h, C.byref(mr), _ptr(d_raw), _ptr(d_rebin)), "rebin_dev")

    def rebin_filter_dev(self, mr, d_raw, d_rebin):
        _check(lib().ctbb_rebin_filter_dev(self._h, C.byref(mr), _ptr(d_raw), _ptr(d_rebin)), "rebin_filter_dev")

    def filter_rows_dev(self, d_taps, n, rows, d_in, d_out):
        _check(lib().ctbb_filter_rows_dev(self._h, _ptr(d_taps), int(n), int(rows), _ptr(d_in), _ptr(d_out)),
               "filter_rows_dev")

    def backproject_dev(self, mr, d_rebin, d_image_block):
        _check(lib().ctbb_backproject_dev(self._h, C.byref(mr), _ptr(d_rebin), _ptr(d_image_block)),
               "backproject_dev")

    def backproject_dev_noalloc(self, mr, d_rebin, d_image_block):
        """backproject_dev without allocating: the view table goes to a fixed
        per-GPU device array (at most NOALLOC_MAX_VIEWS views per block)."""
        _check(lib().ctbb_backproject_dev_noalloc(self._h, C.byref(mr), _ptr(d_rebin), _ptr(d_image_block)),
               "backproject_dev_noalloc")

    def finalize_dev(self, mr, d_image, d_final):
        _check(lib().ctbb_finalize_dev(self._h, C.byref(mr), _ptr(d_image), _ptr(d_final)), "finalize_dev")

    def autotune_backprojection(self, mr, cache_path=None):
        """Choose bp_tile (16 or 32) by timing the middle block; sets it in the
        context's options and returns it.  cache_path: text cache per shape."""
        t = C.c_int()
        _check(lib().ctbb_autotune_backprojection(self._h, C.byref(mr),
                                                  None if cache_path is None else os.fsencode(cache_path),
                                                  C.byref(t)), "autotune_backprojection")
        return t.value

    def hu_to_int16_dev(self, d_hu, n, d_out, slope=1.0, intercept=0.0):
        """int16 stored values of a HU stack (HU = slope * stored + intercept):
        round half to even, saturated, NaN -> 0."""
        _check(lib().ctbb_hu_to_int16_dev(self._h, _ptr(d_hu), int(n), float(slope), float(intercept),
                                          _ptr(d_out)), "hu_to_int16_dev")

    def reconstruct_dev(self, mr, d_raw, n_raw_readings, d_final, d_image=None):
        _check(lib().ctbb_reconstruct_dev(self._h, C.byref(mr), _ptr(d_raw), int(n_raw_readings), _ptr(d_image),
                                          _ptr(d_final)), "reconstruct_dev")

    def reconstruct_host(self, mr, h_raw, n_raw_readings, h_final):
        """Whole reconstruction from a host-resident scan (page-locked memory
        overlaps the uploads with the compute); h_final is host memory."""
        _check(lib().ctbb_reconstruct_host(self._h, C.byref(mr), _ptr(h_raw), int(n_raw_readings), _ptr(h_final)),
               "reconstruct_host")

    def reconstruct_batch_dev(self, mrs, d_raws, n_raw_readings, d_finals):
        """Batching form: several device-resident scans as one pipeline
        (returns when queued; synchronize() waits)."""
        n = len(mrs)
        assert len(d_raws) == n and len(n_raw_readings) == n and len(d_finals) == n
        m_arr = (_MR * n)(*[C.pointer(m) for m in mrs])
        r_arr = (_VP * n)(*[_ptr(d) for d in d_raws])
        n_arr = (C.c_size_t * n)(*[int(x) for x in n_raw_readings])
        f_arr = (_VP * n)(*[_ptr(d) for d in d_finals])
        _check(lib().ctbb_reconstruct_batch_dev(self._h, n, m_arr, r_arr, n_arr, f_arr), "reconstruct_batch_dev")

    def plan_batch_dev(self, mrs):
        """Block plans of a batch computed on the device (the batching form's
        work list): a structured array with one row per (scan, block)."""
        n = sum(int(m.ri.n_blocks) for m in mrs)
        dt = np.dtype([("scan", "<i4"), ("block", "<i4"), ("idx_pull_start", "<i4"), ("n_proj_pull", "<i4"),
                       ("n_half_turns", "<i4"), ("idx_block_slice_start", "<i4"), ("idx_block_slice_end", "<i4"),
                       ("block_slice_start", "<f4"), ("block_slice_end", "<f4"), ("tube_start", "<f4")])
        out = np.zeros(n, dtype=dt)
        m_arr = (_MR * len(mrs))(*[C.pointer(m) for m in mrs])
        _check(lib().ctbb_plan_batch_dev(self._h, len(mrs), m_arr, out.ctypes.data, n), "plan_batch_dev")
        return out

    def synth_scan_dev(self, mr, pp, first, count, d_out):
        _check(lib().ctbb_synth_scan_dev(self._h, C.byref(mr), C.byref(pp), int(first), int(count), _ptr(d_out)),
               "synth_scan_dev")

    # stages with the reference's host-memory contract
    def adaptive_filter_kk(self, mr):
        _check(lib().ctbb_adaptive_filter(self._h, C.byref(mr)), "adaptive_filter_kk")

    def rebin_filter(self, mr):
        _check(lib().ctbb_rebin_filter(self._h, C.byref(mr)), "rebin_filter")

    def backproject(self, mr):
        _check(lib().ctbb_backproject(self._h, C.byref(mr)), "backproject")

    def finalize_image_stack(self, mr):
        _check(lib().ctbb_finalize_image_stack(self._h, C.byref(mr)), "finalize_image_stack")


_DEFAULT_CTX = None


def _default_ctx():
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(-1)
    return _DEFAULT_CTX


def adaptive_filter_kk(mr):
    """reference include/preprocessing.h:26 — in place on mr.ctd.raw."""
    _default_ctx().adaptive_filter_kk(mr)
    return 0


def rebin_filter(mr):
    """reference include/rebin_filter.h:27 — allocates mr.ctd.d_rebin."""
    _default_ctx().rebin_filter(mr)
    return 0


def backproject(mr):
    """reference include/backproject.h:26 — fills block cb.block_idx-1 of mr.ctd.image."""
    _default_ctx().backproject(mr)
    return 0


def finalize_image_stack(mr):
    """reference include/finalize_image_stack.h:26 — allocates mr.ctd.final_image_stack."""
    _default_ctx().finalize_image_stack(mr)
    return 0
