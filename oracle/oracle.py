"""Python driver of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module; the product never does.  It wraps

  liboracle64.so / liboracle32.so   oracle.c built in FP64 / FP32 (Makefile)
  _ref/libctbb_ref.so               the reference's own CPU twins compiled
                                    from their sources (optional; absent on
                                    machines without the reference tree)

and adds the FFT form of the ramp filter (numpy, FP64), which is how the
reference's CPU path filters (src/rebin_filter_cpu.cu:647-709: zero-padded
circular convolution of length M = 2*2^ceil(log2 N)).

Pinning status (see oracle.c header): backproject and finalize are pinned
against the reference's own CPU code; the filter against numpy.convolve and
its closed-form taps; rebin by closed-form invariants; the adaptive filter by
closed forms and invariants of its own algorithm (the reference's stage is
GPU-only CUDA, so no reference build exists to compare against).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BP_GPU, BP_CPU = 0, 1
# planted mistakes (added to BP_GPU) for tests that must see their checks fail
BP_PLANT_NO_FLIP, BP_PLANT_NO_QSIGN = 2, 4

_libs = {}


def build(quiet=True):
    """Compile liboracle64/32 (and _ref when the reference tree exists)."""
    cmd = ["make", "-C", HERE, "all"]
    subprocess.run(cmd, check=True, capture_output=quiet)
    if os.path.isdir(os.environ.get("CTBB_REFERENCE", "/root/reference")):
        subprocess.run(["make", "-C", HERE, "ref", "REF=" + os.environ.get("CTBB_REFERENCE", "/root/reference")],
                       check=True, capture_output=quiet)


def _load(name):
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    _libs[name] = lib
    return lib


def lib(double=True):
    l = _load("liboracle64.so" if double else "liboracle32.so")
    if not getattr(l, "_typed", False):
        vp = C.c_void_p
        l.orc_rebin.argtypes = [vp, vp, vp]
        l.orc_rebin.restype = C.c_int
        l.orc_filter_direct.argtypes = [C.c_int, vp, C.c_long, vp]
        l.orc_filter_direct.restype = None
        l.orc_backproject.argtypes = [vp, vp, vp, C.c_int, vp, vp, C.c_long]
        l.orc_backproject.restype = C.c_int
        l.orc_finalize.argtypes = [vp, vp, vp, C.c_float]
        l.orc_finalize.restype = C.c_int
        l.orc_adaptive_filter.argtypes = [vp, vp, vp, vp]
        l.orc_adaptive_filter.restype = None
        l.orc_adaptive_thresholds.argtypes = [vp, vp, vp]
        l.orc_adaptive_thresholds.restype = None
        l._typed = True
    return l


def ref_lib():
    """The reference's CPU twins compiled from source, or None."""
    path = os.path.join(HERE, "_ref", "libctbb_ref.so")
    if not os.path.exists(path):
        try:
            build()
        except Exception:
            return None
    if not os.path.exists(path):
        return None
    if "ref" not in _libs:
        _libs["ref"] = C.CDLL(path)
    return _libs["ref"]


def _addr(a):
    return a.ctypes.data


def _mr(mr):
    return C.addressof(mr)


def rebin(mr, raw, double=True):
    """Unfiltered rebinned block [n_proj_out][n_rows][n_channels_oversampled]."""
    cg, ri = mr.cg, mr.ri
    n_proj = ri.n_proj_pull // ri.n_ffs
    n_out = n_proj - 2 * cg.add_projections
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    assert raw.size == ri.n_proj_pull * cg.n_rows_raw * cg.n_channels
    out = np.zeros((n_out, cg.n_rows, cg.n_channels_oversampled), dtype=np.float32)
    st = lib(double).orc_rebin(_mr(mr), _addr(raw), _addr(out))
    assert st == 0, st
    return out


def filter_fft(rows, taps):
    """Ramp filter of rows [..., N] in FP64 via zero-padded FFT (the reference
    CPU twin's method): out[l] = sum_k taps[N + l - k] rows[k]."""
    rows = np.asarray(rows, dtype=np.float64)
    n = rows.shape[-1]
    # any circular length M >= 2N - 1 gives the exact linear convolution (the
    # reference uses M = 2 * 2^ceil(log2 N)); take a fast one
    try:
        from scipy.fft import next_fast_len
        m = next_fast_len(2 * n - 1, real=True)
    except ImportError:
        m = 2 * int(2 ** np.ceil(np.log2(n)))
    f = np.zeros(m)
    taps = np.asarray(taps, dtype=np.float64)
    f[:n] = taps[n:2 * n]
    f[m - n + 1:] = taps[1:n]
    try:
        import scipy.fft as sfft       # row-parallel FFTs (each row on its own)
        kw = {"workers": -1}
    except ImportError:
        sfft, kw = np.fft, {}
    F = sfft.rfft(f)
    flat = rows.reshape(-1, n)
    out = np.empty_like(flat)
    step = 8192
    for i in range(0, flat.shape[0], step):
        R = sfft.rfft(flat[i:i + step], n=m, axis=-1, **kw)
        out[i:i + step] = sfft.irfft(R * F, n=m, axis=-1, **kw)[:, :n]
    return out.reshape(rows.shape)


def filter_direct(rows, taps, double=True):
    rows = np.array(rows, dtype=np.float32, copy=True)
    n = rows.shape[-1]
    taps = np.ascontiguousarray(taps, dtype=np.float32)
    lib(double).orc_filter_direct(n, _addr(taps), rows.size // n, _addr(rows))
    return rows


def backproject(mr, rebin_block, conventions=BP_GPU, cols=None, double=True):
    """One 32-slice block; GPU layout [32][nx][ny] (CPU conventions: [32][ny][nx]).
    cols: optional (xi array, yi array) subset of voxel columns."""
    nx, ny = mr.rp.nx, mr.rp.ny
    img = np.zeros((mr.ri.n_slices_block, nx, ny), dtype=np.float32)
    reb = np.ascontiguousarray(rebin_block, dtype=np.float32)
    if cols is None:
        st = lib(double).orc_backproject(_mr(mr), _addr(reb), _addr(img), conventions, None, None, 0)
    else:
        cx = np.ascontiguousarray(cols[0], dtype=np.int32)
        cy = np.ascontiguousarray(cols[1], dtype=np.int32)
        st = lib(double).orc_backproject(_mr(mr), _addr(reb), _addr(img), conventions, _addr(cx), _addr(cy),
                                         len(cx))
    assert st == 0, st
    return img


def finalize(mr, image, mu_water=0.0, double=False):
    nxy = mr.rp.nx * mr.rp.ny
    n_final = int(np.floor(abs(np.float32(mr.rp.end_pos) - np.float32(mr.rp.start_pos)) /
                           np.float32(mr.rp.slice_thickness))) + 1
    out = np.zeros((n_final + 1) * nxy, dtype=np.float32)
    img = np.ascontiguousarray(image, dtype=np.float32)
    nf = lib(double).orc_finalize(_mr(mr), _addr(img), _addr(out), float(mu_water))
    assert nf > 0
    return out[:nf * nxy].reshape(nf, mr.rp.nx, mr.rp.ny)


def adaptive_filter(mr, raw):
    """FP32 restatement of the reference's adaptive filter; returns (out, thresholds)."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    out = np.empty_like(raw)
    thr = np.zeros(mr.ri.n_proj_pull, dtype=np.float32)
    lib(False).orc_adaptive_filter(_mr(mr), _addr(raw), _addr(out), _addr(thr))
    return out, thr


def adaptive_apply(mr, raw, thresholds):
    """The 5x5x5 smoothing step alone, with given per-view thresholds."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    thr = np.ascontiguousarray(thresholds, dtype=np.float32)
    out = np.empty_like(raw)
    l = lib(False)
    l.orc_adaptive_apply.argtypes = [C.c_void_p] * 4
    l.orc_adaptive_apply.restype = None
    l.orc_adaptive_apply(_mr(mr), _addr(raw), _addr(thr), _addr(out))
    return out


def rebin_filter(mr, raw, taps, double=True):
    """Rebin + FFT ramp filter (FP64 filter), float32 result."""
    r = rebin(mr, raw, double)
    return filter_fft(r, taps).astype(np.float32)


# ---------------------------------------------------------------------------
# the reference's own CPU twins (built from source into _ref/)

def ref_backproject_cpu(mr, rebin_block):
    """Run the reference backproject_cpu (src/backproject_cpu.cu:49) on a copy
    of mr; returns the block [32][ny][nx] it writes into ctd.image."""
    lib_ = ref_lib()
    assert lib_ is not None
    from ctbangbang_amd.structs import ReconMetadata
    m = ReconMetadata.from_buffer_copy(mr)
    reb = np.ascontiguousarray(rebin_block, dtype=np.float32)
    nblk = m.ri.n_slices_block * m.rp.nx * m.rp.ny
    image = np.zeros(max(m.ri.cb.block_idx, 1) * nblk, dtype=np.float32)
    m.ctd.rebin = reb.ctypes.data_as(C.POINTER(C.c_float))
    m.ctd.image = image.ctypes.data_as(C.POINTER(C.c_float))
    f = getattr(lib_, "_Z15backproject_cpuP14recon_metadata")
    f.argtypes = [C.c_void_p]
    f.restype = C.c_int
    assert f(C.addressof(m)) == 0
    b = m.ri.cb.block_idx - 1
    return image[b * nblk:(b + 1) * nblk].reshape(m.ri.n_slices_block, m.rp.ny, m.rp.nx)


def ref_finalize_cpu(mr, image):
    """Run the reference finalize_image_stack_cpu (src/finalize_image_stack_cpu.cu:24)."""
    lib_ = ref_lib()
    assert lib_ is not None
    from ctbangbang_amd.structs import ReconMetadata
    m = ReconMetadata.from_buffer_copy(mr)
    img = np.ascontiguousarray(image, dtype=np.float32)
    m.ctd.image = img.ctypes.data_as(C.POINTER(C.c_float))
    m.ctd.final_image_stack = None
    f = getattr(lib_, "_Z24finalize_image_stack_cpuP14recon_metadata")
    f.argtypes = [C.c_void_p]
    f.restype = C.c_int
    assert f(C.addressof(m)) == 0
    n_final = int(np.floor(abs(np.float32(m.rp.end_pos) - np.float32(m.rp.start_pos)) /
                           np.float32(m.rp.slice_thickness))) + 1
    n = n_final * m.rp.nx * m.rp.ny
    out = np.ctypeslib.as_array(m.ctd.final_image_stack, shape=(n,)).copy()
    C.CDLL(None).free(C.cast(m.ctd.final_image_stack, C.c_void_p))
    return out.reshape(n_final, m.rp.nx, m.rp.ny)


def ref_parse_config(path):
    """Run the reference parse_config (src/parse_config.cu:40) on a prm file."""
    lib_ = ref_lib()
    assert lib_ is not None
    from ctbangbang_amd.structs import ReconParams
    rp = ReconParams()
    f = getattr(lib_, "_Z12parse_configPcP12recon_params")
    f.argtypes = [C.c_char_p, C.c_void_p]
    f.restype = None
    f(str(path).encode(), C.addressof(rp))
    return rp
