"""ctypes wrapper over ``libksj_oracle.so`` (FP64 restatement) and, where it was
built, ``_ref/libksj_ref.so`` (the reference's own Eigen-free NUFFT sources).

TEST INFRASTRUCTURE ONLY (see the package docstring).  Every function takes and
returns numpy arrays; complex arrays are complex128 unless stated.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE_SO = os.path.join(_HERE, "libksj_oracle.so")
_REF_SO = os.path.join(_HERE, "_ref", "libksj_ref.so")
_ROUTE_SO = os.path.join(_HERE, "libksj_kb_route.so")

_lib = None
_ref = None
_route = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_sz = C.c_size_t


def build() -> None:
    """Compile the oracle (and the reference shim when the source tree is present)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "all"], check=True)
    if os.path.isdir("/root/reference/libs/numerics/kspacejet-nufft"):
        subprocess.run(["make", "-s", "-B", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_ORACLE_SO):
            subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
        L = C.CDLL(_ORACLE_SO)
        L.ora_nudft2_forward.argtypes = [_sz, _sz, C.c_double, C.c_double, _sz, _dp, _dp, _dp]
        L.ora_nudft2_adjoint.argtypes = [_sz, _sz, C.c_double, C.c_double, _sz, _dp, _dp, _dp]
        L.ora_ramp_dcf2.argtypes = [_sz, _dp, _dp]
        L.ora_radial_linear_gridding2.argtypes = [_sz, _sz, C.c_double, C.c_double, _sz, _dp, _dp, _dp, _dp]
        L.ora_fft1.argtypes = [_sz, _dp, C.c_int]
        L.ora_ifft2.argtypes = [_sz, _sz, _dp, C.c_int]
        L.ora_rss.argtypes = [_sz, _sz, _dp, _dp]
        L.ora_sense_combine.argtypes = [_sz, _sz, _dp, _dp, _dp]
        L.ora_i0.argtypes = [C.c_double]
        L.ora_i0.restype = C.c_double
        L.ora_kb_beta.argtypes = [C.c_int, C.c_double]
        L.ora_kb_beta.restype = C.c_double
        L.ora_kb_kernel.argtypes = [C.c_double, C.c_int, C.c_double]
        L.ora_kb_kernel.restype = C.c_double
        L.ora_kb_ft.argtypes = [C.c_double, _sz, C.c_int, C.c_double]
        L.ora_kb_ft.restype = C.c_double
        L.ora_grid_coordinate.argtypes = [C.c_float, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
        L.ora_bin2.argtypes = [_sz, _fp, C.c_uint32, C.c_uint32, _up, _fp, _up, _up]
        L.ora_kb_adjoint2.argtypes = [_sz, _sz, _sz, _sz, C.c_int, C.c_double, C.c_double, C.c_double, _sz,
                                      _dp, _dp, _dp, _dp]
        L.ora_golden_angle2.argtypes = [_sz, _sz, _dp]
        L.ora_nudft3_adjoint.argtypes = [_sz, _sz, _sz, C.c_double, C.c_double, C.c_double, _sz, _dp, _dp, _dp]
        L.ora_ramp_sq_dcf3.argtypes = [_sz, _dp, _dp]
        L.ora_bin3.argtypes = [_sz, _fp, C.c_uint32, C.c_uint32, C.c_uint32, _up, _up, _up]
        L.ora_kb_adjoint3.argtypes = [_sz, _sz, _sz, _sz, _sz, _sz, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _sz, _dp, _dp, _dp, _dp]
        L.ora_kb_adjoint3_voxels.argtypes = [_sz, _sz, _sz, _sz, _sz, _sz, C.c_int, C.c_double, C.c_double,
                                             C.c_double, C.c_double, _sz, _sz, _dp, _dp, _dp, _sz, _up, _sz, _up,
                                             _sz, _up, _dp]
        L.ora_golden_means3.argtypes = [_sz, _sz, _dp]
        L.ora_channel_covariance.argtypes = [_sz, _sz, _fp, C.c_int, _dp]
        L.ora_hermitian_eigen.argtypes = [_sz, _dp, _dp, _dp]
        L.ora_noise_model_estimate.argtypes = [_sz, _sz, _fp, C.c_double, _dp]
        L.ora_coil_compression_basis_estimate.argtypes = [_sz, _sz, _sz, _fp, _dp]
        L.ora_phase_correction_estimate.argtypes = [_sz, _sz, _sz, _fp, _fp]
        L.ora_walsh_maps.argtypes = [_sz, _sz, _sz, _sz, _sz, _dp, _dp]
        L.ora_nudft3_forward.argtypes = [_sz, _sz, _sz, C.c_double, C.c_double, C.c_double, _sz, _dp, _dp, _dp]
        L.ora_normalize_trajectory2.argtypes = [_sz, _fp, C.c_int, C.c_uint32, C.c_uint32, _fp]
        L.ora_channel_mix.argtypes = [_sz, _sz, _sz, _dp, _dp, _dp]
        L.ora_phase_correct.argtypes = [_sz, _sz, _sz, _fp, _fp, _fp]
        L.ora_readout_crop.argtypes = [_sz, _sz, _sz, _sz, _fp, _fp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF_SO)


def ref():
    """The reference's compiled NUFFT sources (None where oracle/_ref was not built)."""
    global _ref
    if _ref is None and ref_available():
        R = C.CDLL(_REF_SO)
        R.ref_ramp_dcf2_f64.argtypes = [_sz, _dp, _dp]
        R.ref_ramp_dcf2_f32.argtypes = [_sz, _fp, _fp]
        R.ref_radial_linear_gridding2_f64.argtypes = [_sz, _sz, C.c_double, C.c_double, _sz, _dp, _dp, _dp, _dp]
        R.ref_radial_linear_gridding2_f32.argtypes = [_sz, _sz, C.c_double, C.c_double, _sz, _fp, _fp, _fp, _fp]
        R.ref_direct_adjoint2_f64.argtypes = [_sz, _sz, C.c_double, C.c_double, _sz, _dp, _dp, _dp]
        R.ref_direct_forward2_f64.argtypes = [_sz, _sz, C.c_double, C.c_double, _sz, _dp, _dp, _dp]
        R.ref_radial_route_rss_batch.argtypes = [_sz, _sz, _sz, _sz, _sz, _fp, _fp, _fp, C.c_uint]
        R.ref_radial_route_rss_repeat.argtypes = [_sz, _sz, _sz, _sz, _sz, _fp, _fp, _fp, C.c_uint]
        _ref = R
    return _ref


def _c128(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128))


def _as_f64_pairs(x: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(_c128(x)).view(np.float64)


def _traj64(traj) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(traj, dtype=np.float64).reshape(-1, 2))


# --------------------------------------------------------------------------- direct NUDFT

def nudft2_forward(image, traj, row_origin=0.0, col_origin=0.0) -> np.ndarray:
    img = _c128(image)
    t = _traj64(traj)
    out = np.zeros(t.shape[0], np.complex128)
    lib().ora_nudft2_forward(img.shape[0], img.shape[1], row_origin, col_origin, t.shape[0], t.reshape(-1),
                             img.view(np.float64).reshape(-1), out.view(np.float64))
    return out


def nudft2_adjoint(values, traj, rows, cols, row_origin=0.0, col_origin=0.0) -> np.ndarray:
    v = _c128(values).reshape(-1)
    t = _traj64(traj)
    out = np.zeros((rows, cols), np.complex128)
    lib().ora_nudft2_adjoint(rows, cols, row_origin, col_origin, t.shape[0], t.reshape(-1), v.view(np.float64),
                             out.view(np.float64).reshape(-1))
    return out


def ramp_dcf2(traj) -> np.ndarray:
    t = _traj64(traj)
    out = np.zeros(t.shape[0], np.float64)
    lib().ora_ramp_dcf2(t.shape[0], t.reshape(-1), out)
    return out


def radial_linear_gridding2(values, traj, dcf, rows, cols, row_origin, col_origin) -> np.ndarray:
    v = _c128(values).reshape(-1)
    t = _traj64(traj)
    d = np.ascontiguousarray(np.asarray(dcf, np.float64))
    out = np.zeros((rows, cols), np.complex128)
    rc = lib().ora_radial_linear_gridding2(rows, cols, row_origin, col_origin, t.shape[0], v.view(np.float64),
                                           t.reshape(-1), d, out.view(np.float64).reshape(-1))
    if rc != 0:
        raise ValueError("trajectory outside [-pi, pi] or unmappable")
    return out


# --------------------------------------------------------------------------- FFT / combine

def fft1(x, sign: int) -> np.ndarray:
    y = _c128(x).copy()
    lib().ora_fft1(y.shape[0], y.view(np.float64), int(sign))
    return y


def ifft2(x, centred: bool = False) -> np.ndarray:
    y = _c128(x).copy()
    lib().ora_ifft2(y.shape[0], y.shape[1], y.view(np.float64).reshape(-1), 1 if centred else 0)
    return y


def ifft2_coil_images(kspace, centred: bool = False) -> np.ndarray:
    """[C, R, Cols] channel-major k-space -> [R, Cols, C] coil-last images."""
    k = _c128(kspace)
    return np.stack([ifft2(k[c], centred) for c in range(k.shape[0])], axis=-1)


def rss(coil_last) -> np.ndarray:
    x = _c128(coil_last)
    coils = x.shape[-1]
    pix = x.size // coils
    out = np.zeros(pix, np.float64)
    lib().ora_rss(pix, coils, x.view(np.float64).reshape(-1), out)
    return out.reshape(x.shape[:-1])


def sense_combine(coil_last, sens_last) -> np.ndarray:
    x = _c128(coil_last)
    s = _c128(sens_last)
    coils = x.shape[-1]
    pix = x.size // coils
    out = np.zeros(pix, np.complex128)
    lib().ora_sense_combine(pix, coils, x.view(np.float64).reshape(-1), s.view(np.float64).reshape(-1),
                            out.view(np.float64))
    return out.reshape(x.shape[:-1])


# --------------------------------------------------------------------------- Kaiser-Bessel

def i0(x: float) -> float:
    return lib().ora_i0(float(x))


def kb_beta(width: int, sigma: float) -> float:
    return lib().ora_kb_beta(int(width), float(sigma))


def kb_kernel(d: float, width: int, beta: float) -> float:
    return lib().ora_kb_kernel(float(d), int(width), float(beta))


def kb_ft(index: float, grid: int, width: int, beta: float) -> float:
    return lib().ora_kb_ft(float(index), int(grid), int(width), float(beta))


def grid_coordinate(w: float, extent: int):
    cell = C.c_uint32(0)
    frac = C.c_float(0)
    rc = lib().ora_grid_coordinate(C.c_float(w), extent, C.byref(cell), C.byref(frac))
    if rc != 0:
        raise ValueError("coordinate cannot be mapped")
    return cell.value, np.float32(frac.value)


def bin2(traj_f32, grid_rows: int, grid_cols: int):
    """Returns (bins[S] u32, frac[S,2] f32, cell_start[G_r*G_c+1] u32, perm[S] u32)."""
    t = np.ascontiguousarray(np.asarray(traj_f32, np.float32).reshape(-1, 2))
    S = t.shape[0]
    bins = np.zeros(S, np.uint32)
    frac = np.zeros((S, 2), np.float32)
    start = np.zeros(grid_rows * grid_cols + 1, np.uint32)
    perm = np.zeros(S, np.uint32)
    rc = lib().ora_bin2(S, t.reshape(-1), grid_rows, grid_cols, bins, frac.reshape(-1), start, perm)
    if rc != 0:
        raise ValueError("trajectory cannot be mapped to the grid")
    return bins, frac, start, perm


def kb_adjoint2(values, traj, dcf, n_rows, n_cols, oversampling=2.0, width=4, beta=None,
                row_origin=None, col_origin=None) -> np.ndarray:
    g_rows = int(round(n_rows * oversampling))
    g_cols = int(round(n_cols * oversampling))
    if beta is None:
        beta = kb_beta(width, oversampling)
    if row_origin is None:
        row_origin = (n_rows - 1) / 2.0
    if col_origin is None:
        col_origin = (n_cols - 1) / 2.0
    v = _c128(values).reshape(-1)
    t = _traj64(traj)
    d = np.ones(t.shape[0], np.float64) if dcf is None else np.ascontiguousarray(np.asarray(dcf, np.float64))
    out = np.zeros((n_rows, n_cols), np.complex128)
    rc = lib().ora_kb_adjoint2(n_rows, n_cols, g_rows, g_cols, int(width), float(beta), float(row_origin),
                               float(col_origin), t.shape[0], t.reshape(-1), v.view(np.float64), d,
                               out.view(np.float64).reshape(-1))
    if rc != 0:
        raise ValueError("trajectory cannot be mapped to the grid")
    return out


def kb_adjoint2_coils(kspace, traj, dcf, n, oversampling=2.0, width=4):
    """[C, S] channel-major samples -> [N, N, C] coil-last FP64 KB images."""
    k = _c128(kspace)
    return np.stack([kb_adjoint2(k[c], traj, dcf, n, n, oversampling, width) for c in range(k.shape[0])], axis=-1)


def golden_angle2(spokes: int, readout: int) -> np.ndarray:
    out = np.zeros((spokes * readout, 2), np.float64)
    lib().ora_golden_angle2(spokes, readout, out.reshape(-1))
    return out


def ref_radial_route_rss(values_f32, traj_f32, rows: int, cols: int, scans: int = 1, threads: int = 1) -> np.ndarray:
    """The reference radial route (ramp DCF + bilinear gridding + radix-2 iFFT per channel, then RSS)
    run `scans` times over one scan's data on `threads` host threads (compiled reference sources)."""
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref was not built")
    v = np.ascontiguousarray(values_f32, np.complex64)
    t = np.ascontiguousarray(traj_f32, np.float32).reshape(-1, 2)
    out = np.zeros((rows, cols), np.float32)
    rc = R.ref_radial_route_rss_repeat(scans, rows, cols, v.shape[0], v.shape[1], v.view(np.float32).reshape(-1),
                                       t.reshape(-1), out.reshape(-1), threads)
    if rc != 0:
        raise ValueError("reference route rejected its input")
    return out


def kb_route_rss_f32(values_f32, traj_f32, n: int, width: int = 4, scans: int = 1, threads: int = 1,
                     timed: bool = False):
    """FP32 CPU restatement of the GPU radial route (ramp DCF, KB width `width` with the
    Beatty beta on the 2n x 2n grid, radix-2 inverse FFT, deapodization, crop, RSS), the
    like-for-like CPU baseline; the scan is reconstructed `scans` times on `threads`
    threads.  values [C, S] complex64, traj [S, 2] float32 rad/px -> rss [n, n] float32
    (and, when `timed`, the wall seconds of the scans after the per-trajectory tables)."""
    v = np.ascontiguousarray(values_f32, np.complex64)
    t = np.ascontiguousarray(traj_f32, np.float32).reshape(-1, 2)
    out = np.zeros((n, n), np.float32)
    secs = C.c_double(0)
    rc = _route_lib().ora_kb_route_rss_f32(n, t.shape[0], v.shape[0], t.reshape(-1), v.view(np.float32).reshape(-1), width,
                                     scans, threads, out.reshape(-1), C.byref(secs))
    if rc != 0:
        raise ValueError("CPU route rejected its input")
    return (out, secs.value) if timed else out


def _route_lib():
    global _route
    if _route is None:
        if not os.path.exists(_ROUTE_SO):
            subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
        R = C.CDLL(_ROUTE_SO)
        R.ora_kb_route_rss_f32.argtypes = [_sz, _sz, _sz, _fp, _fp, C.c_int, _sz, C.c_int, _fp,
                                           C.POINTER(C.c_double)]
        R.ora_kb_route_rss_f32.restype = C.c_int
        R.ora_cart_route_rss_f32.argtypes = [_sz, _sz, _sz, _fp, C.c_int, _sz, C.c_int, _fp, C.POINTER(C.c_double)]
        R.ora_cart_route_rss_f32.restype = C.c_int
        R.ora_kb_route3_rss_f32.argtypes = [_sz, _sz, _sz, _fp, _fp, C.c_int, C.c_int, _fp, C.POINTER(C.c_double)]
        R.ora_kb_route3_rss_f32.restype = C.c_int
        R.ora_assemble_acquisitions_f32.argtypes = [_fp, _up, _sz, _sz, _sz, _fp, C.POINTER(C.c_double)]
        R.ora_assemble_acquisitions_f32.restype = C.c_int
        _route = R
    return _route


def assemble_acquisitions_c(packed, counts, channels: int, frames: int = 1):
    """The replay's copy loop in C (memcpy per acquisition and channel; the CPU baseline of the
    acquisition assembly): packed complex64 acquisitions -> (frame [channels][S] complex64, wall
    seconds of `frames` repeats on one thread), or (None, 0.0) on the route's rejections."""
    p = np.ascontiguousarray(np.asarray(packed, np.complex64).reshape(-1))
    cnt = np.ascontiguousarray(np.asarray(counts, np.uint32))
    out = np.zeros((channels, int(cnt.sum(dtype=np.uint64))), np.complex64)
    secs = C.c_double(0)
    rc = _route_lib().ora_assemble_acquisitions_f32(p.view(np.float32), cnt, cnt.size, channels, frames,
                                                    out.view(np.float32).reshape(-1), C.byref(secs))
    return (out, secs.value) if rc == 0 else (None, 0.0)


def cart_route_rss_f32(kspace_f32, centred: bool = True, scans: int = 1, threads: int = 1, timed: bool = False):
    """FP32 CPU restatement of the Cartesian route: per coil the inverse 2-D FFT with
    Normalization::inverse (centred or not), then RSS in coil order.  kspace [C, rows, cols]
    complex64 (powers of two) -> rss [rows, cols] float32 (and, when `timed`, the wall
    seconds of `scans` repeats on `threads` threads)."""
    k = np.ascontiguousarray(kspace_f32, np.complex64)
    c, r, w = k.shape
    out = np.zeros((r, w), np.float32)
    secs = C.c_double(0)
    rc = _route_lib().ora_cart_route_rss_f32(r, w, c, k.view(np.float32).reshape(-1), 1 if centred else 0, scans,
                                             threads, out.reshape(-1), C.byref(secs))
    if rc != 0:
        raise ValueError("Cartesian CPU route rejected its input")
    return (out, secs.value) if timed else out


def kb_route3_rss_f32(values_f32, traj_f32, n: int, width: int = 4, threads: int = 1, timed: bool = False):
    """FP32 CPU restatement of the 3-D kooshball route (|k|^2 DCF, KB `width` with the
    Beatty beta on the (2n)^3 grid, radix-2 inverse FFT pruned axis by axis,
    deapodization, RSS); threads take whole coils.  values [C, S] complex64, traj [S, 3]
    float32 (z, y, x) rad/px -> rss [n, n, n] float32 (and, when `timed`, the wall
    seconds of the coils after the per-trajectory tables)."""
    v = np.ascontiguousarray(values_f32, np.complex64)
    t = np.ascontiguousarray(traj_f32, np.float32).reshape(-1, 3)
    out = np.zeros((n, n, n), np.float32)
    secs = C.c_double(0)
    rc = _route_lib().ora_kb_route3_rss_f32(n, t.shape[0], v.shape[0], t.reshape(-1), v.view(np.float32).reshape(-1),
                                            width, threads, out.reshape(-1), C.byref(secs))
    if rc != 0:
        raise ValueError("3-D CPU route rejected its input")
    return (out, secs.value) if timed else out


# ---------------------------------------------------------------- 3-D (kooshball)
# Coordinates [S][3] in (z, y, x) order, radians per pixel; volumes [z][y][x].  The
# reference has no 3-D route: these extend its 2-D definitions axis by axis
# (eigen_nufft_direct_nudft.cpp:77-97, eigen_radial_gridding.cpp:262-277).

def _traj3(traj) -> np.ndarray:
    t = np.ascontiguousarray(np.asarray(traj, np.float64))
    if t.ndim != 2 or t.shape[1] != 3:
        raise ValueError("3-D trajectory must be [S, 3]")
    return t


def nudft3_adjoint(values, traj, shape, origin=None) -> np.ndarray:
    """Exact FP64 adjoint sum_s y_s exp(+i w . (r - origin)), no DCF, no normalisation."""
    nz, ny, nx = shape
    if origin is None:
        origin = ((nz - 1) / 2.0, (ny - 1) / 2.0, (nx - 1) / 2.0)
    v = _c128(values).reshape(-1)
    t = _traj3(traj)
    out = np.zeros(shape, np.complex128)
    lib().ora_nudft3_adjoint(nz, ny, nx, *map(float, origin), t.shape[0], t.reshape(-1), v.view(np.float64),
                             out.view(np.float64).reshape(-1))
    return out


def nudft3_forward(image, traj, origin=None) -> np.ndarray:
    """Exact FP64 forward model: [S] complex128 = sum_v x_v exp(-i w_s . (v - origin))."""
    x = _c128(image)
    nz, ny, nx = x.shape
    t = _traj3(traj)
    if origin is None:
        origin = ((nz - 1) / 2, (ny - 1) / 2, (nx - 1) / 2)
    out = np.zeros(t.shape[0], np.complex128)
    lib().ora_nudft3_forward(nz, ny, nx, float(origin[0]), float(origin[1]), float(origin[2]), t.shape[0],
                             t.reshape(-1), x.view(np.float64).reshape(-1), out.view(np.float64))
    return out


def ramp_sq_dcf3(traj) -> np.ndarray:
    t = _traj3(traj)
    out = np.zeros(t.shape[0], np.float64)
    lib().ora_ramp_sq_dcf3(t.shape[0], t.reshape(-1), out)
    return out


def bin3(traj_f32, grid):
    """FP64 binning of float32 (z, y, x) coordinates: (bins, cell_start, perm)."""
    t = np.ascontiguousarray(np.asarray(traj_f32, np.float32))
    S = t.shape[0]
    gz, gy, gx = grid
    bins = np.zeros(S, np.uint32)
    start = np.zeros(gz * gy * gx + 1, np.uint32)
    perm = np.zeros(S, np.uint32)
    if lib().ora_bin3(S, t.reshape(-1), gz, gy, gx, bins, start, perm) != 0:
        raise ValueError("trajectory cannot be mapped to the grid")
    return bins, start, perm


def kb_adjoint3(values, traj, dcf, shape, oversampling=2.0, width=4, beta=None, origin=None) -> np.ndarray:
    """FP64 KB gridding adjoint of one coil onto a sigma-oversampled grid, centred crop, deapodized."""
    nz, ny, nx = shape
    gz, gy, gx = (int(round(n * oversampling)) for n in shape)
    if beta is None:
        beta = kb_beta(width, oversampling)
    if origin is None:
        origin = ((nz - 1) / 2.0, (ny - 1) / 2.0, (nx - 1) / 2.0)
    v = _c128(values).reshape(-1)
    t = _traj3(traj)
    d = np.ones(t.shape[0], np.float64) if dcf is None else np.ascontiguousarray(np.asarray(dcf, np.float64))
    out = np.zeros(shape, np.complex128)
    rc = lib().ora_kb_adjoint3(nz, ny, nx, gz, gy, gx, int(width), float(beta), *map(float, origin), t.shape[0],
                               t.reshape(-1), v.view(np.float64), d, out.view(np.float64).reshape(-1))
    if rc != 0:
        raise ValueError("trajectory cannot be mapped to the grid")
    return out


def kb_adjoint3_voxels(values, traj, dcf, shape, zs, ys, xs, oversampling=2.0, width=4, beta=None, origin=None,
                       threads: int = 8) -> np.ndarray:
    """kb_adjoint3 at the voxel lattice zs x ys x xs only ([len(zs), len(ys), len(xs)] complex128),
    for volumes whose full FP64 oracle would take minutes.  The samples are split over
    `threads` host threads (the C call releases the GIL) and the partial sums added."""
    import concurrent.futures as cf
    nz, ny, nx = shape
    gz, gy, gx = (int(round(n * oversampling)) for n in shape)
    if beta is None:
        beta = kb_beta(width, oversampling)
    if origin is None:
        origin = ((nz - 1) / 2.0, (ny - 1) / 2.0, (nx - 1) / 2.0)
    v = _c128(values).reshape(-1)
    t = _traj3(traj)
    S = t.shape[0]
    d = np.ones(S, np.float64) if dcf is None else np.ascontiguousarray(np.asarray(dcf, np.float64))
    idx = [np.ascontiguousarray(np.asarray(a, np.uint32)) for a in (zs, ys, xs)]
    bounds = np.linspace(0, S, max(1, threads) + 1).astype(np.int64)

    def part(i):
        out = np.zeros((len(idx[0]), len(idx[1]), len(idx[2])), np.complex128)
        rc = lib().ora_kb_adjoint3_voxels(nz, ny, nx, gz, gy, gx, int(width), float(beta), *map(float, origin),
                                          int(bounds[i]), int(bounds[i + 1]), t.reshape(-1), v.view(np.float64), d,
                                          len(idx[0]), idx[0], len(idx[1]), idx[1], len(idx[2]), idx[2],
                                          out.view(np.float64).reshape(-1))
        if rc != 0:
            raise ValueError("trajectory cannot be mapped to the grid")
        return out

    with cf.ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return sum(ex.map(part, range(len(bounds) - 1)))


def golden_means3(spokes: int, readout: int) -> np.ndarray:
    out = np.zeros((spokes * readout, 3), np.float64)
    lib().ora_golden_means3(spokes, readout, out.reshape(-1))
    return out


# ---------------------------------------------------------------- trajectory units
TRAJ_UNITS = {"unspecified": 0, "cycles_per_fov": 1, "radians_per_pixel": 2, "encoded_matrix_index": 3}


def normalize_trajectory2(raw_kx_ky, units: str, encoded_rows: int, encoded_cols: int):
    """Route normalisation (noncartesian_rss_hdf5.cpp:776-834): ([S, 2] (ky, kx) float32 rad/px, 0)
    or (None, code) with the first failure's code (1 non-finite, 2 outside Nyquist, 3 unspecified,
    4 encoded dimension < 2)."""
    raw = np.ascontiguousarray(np.asarray(raw_kx_ky, np.float32))
    out = np.zeros_like(raw)
    rc = lib().ora_normalize_trajectory2(raw.shape[0], raw.reshape(-1), TRAJ_UNITS[units], encoded_rows,
                                         encoded_cols, out.reshape(-1))
    return (out, 0) if rc == 0 else (None, rc)


def assemble_acquisitions(acquisitions, channels: int):
    """The route's replay copy loop (noncartesian_rss_hdf5.cpp:1077-1083): acquisitions, each a flat
    complex64 array in the ISMRMRD layout samples[sample + channel * ns] (dataset_reader.hpp:57), ->
    (frame [channels][S] complex64, 0), acquisition a's channel-c row at columns [o_a, o_a + ns_a)
    with o_a the samples written before it (:1104); or (None, 1) on the route's rejections
    (:880-884): no acquisition, one without samples, channels outside [1, 64]."""
    if not acquisitions or not 1 <= channels <= 64:
        return None, 1
    rows = [np.asarray(a, np.complex64).reshape(-1) for a in acquisitions]
    if any(r.size == 0 or r.size % channels for r in rows):
        return None, 1
    frame = np.zeros((channels, sum(r.size // channels for r in rows)), np.complex64)
    written = 0
    for r in rows:
        n = r.size // channels
        for c in range(channels):
            frame[c, written:written + n] = r[c * n:(c + 1) * n]
        written += n
    return frame, 0


# ---------------------------------------------------------------- k-space conditioning
def channel_mix(w, x) -> np.ndarray:
    """FP64 W[out, in] @ X[in, samples] (complex128 out): noise prewhitening and coil
    compression of channel-major k-space (noncartesian_noise_prewhiten.cpp:120-136)."""
    w = _c128(w)
    x = _c128(x)
    out_ch, in_ch = w.shape
    assert x.shape[0] == in_ch
    xs = x.reshape(in_ch, -1)
    y = np.zeros((out_ch, xs.shape[1]), np.complex128)
    lib().ora_channel_mix(out_ch, in_ch, xs.shape[1], w.view(np.float64).reshape(-1), xs.view(np.float64).reshape(-1),
                          y.view(np.float64).reshape(-1))
    return y.reshape((out_ch,) + x.shape[1:])


def phase_correct(x, phase) -> np.ndarray:
    """complex64 x[C, R, K] * phase[C, K] with float products and sums (phase_correct.cpp:127-143)."""
    x = np.ascontiguousarray(np.asarray(x, np.complex64))
    p = np.ascontiguousarray(np.asarray(phase, np.complex64))
    C_, R, K = x.shape
    assert p.shape == (C_, K)
    y = np.zeros_like(x)
    lib().ora_phase_correct(C_, R, K, x.view(np.float32).reshape(-1), p.view(np.float32).reshape(-1),
                            y.view(np.float32).reshape(-1))
    return y


def readout_crop(x, out_cols: int, offset: int) -> np.ndarray:
    """complex64 x[..., in_cols] -> x[..., offset:offset + out_cols] (readout_oversampling_remove.cpp:165-181)."""
    x = np.ascontiguousarray(np.asarray(x, np.complex64))
    in_cols = x.shape[-1]
    lines = x.size // in_cols
    y = np.zeros(x.shape[:-1] + (out_cols,), np.complex64)
    lib().ora_readout_crop(lines, in_cols, out_cols, offset, x.view(np.float32).reshape(-1),
                           y.view(np.float32).reshape(-1))
    return y


# ---------------------------------------------------------------- calibration
# providers/kspacejet-calibration/src/operators/*.cpp restated: complex arrays in, complex128
# (or complex64 for the phase model) out.

def _c64_pairs(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, np.complex64)).view(np.float32).reshape(-1)


def channel_covariance(x, centred: bool) -> np.ndarray:
    """[C, S] complex -> [C, C] complex128 (centred: conj(x_r - m_r)(x_c - m_c); else x_r conj(x_c))."""
    x = np.ascontiguousarray(np.asarray(x, np.complex64))
    C, S = x.shape
    out = np.zeros((C, C), np.complex128)
    lib().ora_channel_covariance(C, S, _c64_pairs(x), int(centred), out.view(np.float64).reshape(-1))
    return out


def hermitian_eigen(a):
    """The reference's largest-pivot Jacobi in FP64: (eigenvalues ascending, vectors in columns)."""
    a = np.ascontiguousarray(np.asarray(a, np.complex128))
    n = a.shape[0]
    vals = np.zeros(n, np.float64)
    vecs = np.zeros((n, n), np.complex128)
    if lib().ora_hermitian_eigen(n, a.view(np.float64).reshape(-1), vals, vecs.view(np.float64).reshape(-1)) != 0:
        raise ValueError("eigen-decomposition failed")
    return vals, vecs


def noise_model_estimate(x, floor: float = float(np.float32(1e-6))) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, np.complex64))
    C, S = x.shape
    out = np.zeros((C, C), np.complex128)
    if lib().ora_noise_model_estimate(C, S, _c64_pairs(x), float(floor), out.view(np.float64).reshape(-1)) != 0:
        raise ValueError("noise model estimate failed")
    return out


def coil_compression_basis_estimate(x, virtual_channels: int) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, np.complex64))
    P = x.shape[0]
    S = x.size // P
    out = np.zeros((virtual_channels, P), np.complex128)
    if lib().ora_coil_compression_basis_estimate(P, virtual_channels, S, _c64_pairs(x),
                                                 out.view(np.float64).reshape(-1)) != 0:
        raise ValueError("coil compression basis estimate failed")
    return out


def phase_correction_estimate(ref) -> np.ndarray:
    """[C, lines, K] complex64 -> [C, K] complex64."""
    ref = np.ascontiguousarray(np.asarray(ref, np.complex64))
    C, L, K = ref.shape
    out = np.zeros((C, K), np.complex64)
    lib().ora_phase_correction_estimate(C, L, K, _c64_pairs(ref), out.view(np.float32).reshape(-1))
    return out


# ---------------------------------------------------------------- sensitivity maps

def walsh_maps(coil_last, window: int, iterations: int) -> np.ndarray:
    """[rows, cols, C] complex -> [rows, cols, C] complex128 maps (Walsh's adaptive method:
    dominant eigenvector of the windowed coil covariance, coil 0 real and >= 0)."""
    x = _c128(coil_last)
    rows, cols, coils = x.shape
    out = np.zeros_like(x)
    lib().ora_walsh_maps(rows, cols, coils, window, iterations, x.view(np.float64).reshape(-1),
                         out.view(np.float64).reshape(-1))
    return out
