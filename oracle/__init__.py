"""CPU oracle for the visibility -> dirty-image path.

TEST INFRASTRUCTURE ONLY.  This package wraps ``oracle/blink_oracle.c`` (a
plain-C restatement of the reference imager's CPU path) with numpy-facing
helpers.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` may import it, and only as the checker: the product library
(``imager_amd``) never imports, loads or calls anything in here.

Reference behaviour restated (file:line in the reference repository):
  * ``baseline_flags``        src/pacer_imager.cpp:60-86
  * ``uv_max``                src/pacer_imager.cpp:213-221
  * ``delta_uv``              src/pacer_imager.cpp:497-499
  * ``frequency_hz``          src/pacer_imager.cpp:528-555
  * ``zenith_uvw``            src/antenna_positions.cpp:52-90
  * ``geometric_corrections`` src/corrections.cpp:7-39 (w[bl], src/gpu/corrections_gpu.cpp:61)
  * ``cable_corrections``     src/corrections.cpp:41-74
  * ``cells``                 src/pacer_imager.cpp:289,332-369
  * ``gridding``              src/pacer_imager.cpp:250-432
  * ``fft2d``                 src/pacer_imager.cpp:150-152 (FFTW_BACKWARD, unnormalised)
  * ``fft_shift``             src/pacer_imager.cpp:30-35,110-128
  * ``image``                 src/pacer_imager.cpp:134-208
  * ``image_averaging_*``     src/utils.cpp:4-22, src/gpu/gpu_utils.cpp:143-161
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libblink_oracle.so")
_lib = None

POL = {"XX": 0, "YY": 1, "I": 2, "V": 3}

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_i64 = ctypes.c_int64
_c_int = ctypes.c_int
_c_dbl = ctypes.c_double


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, IEEE flags)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    L.oracle_n_baselines.restype = _c_i64
    L.oracle_n_baselines.argtypes = [_c_int]
    L.oracle_baseline_flags.argtypes = [_c_int, _i32p, _c_int, _i32p]
    L.oracle_uv_max.argtypes = [_f32p, _f32p, _i32p, _c_i64,
                                ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_dbl)]
    L.oracle_delta_uv.argtypes = [ctypes.c_float, _c_dbl, _c_dbl, _c_int,
                                  ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_dbl)]
    L.oracle_frequency_hz.restype = _c_dbl
    L.oracle_frequency_hz.argtypes = [_c_dbl, _c_int, _c_dbl, _c_int, _c_int, _c_int]
    L.oracle_zenith_uvw.argtypes = [_c_int, _f64p, _f64p, _f64p, _f32p, _f32p, _f32p]
    L.oracle_geometric_corrections.argtypes = [_f32p, _c_int, _c_int, _c_int, _f32p, _f64p]
    L.oracle_cable_corrections.argtypes = [_f32p, _c_int, _c_int, _c_int, _f64p, _f64p]
    L.oracle_cells.argtypes = [_f32p, _f32p, _i32p, _c_i64, _f64p, _c_int, _c_dbl, _c_dbl,
                               _c_dbl, _c_int, _i32p]
    L.oracle_cell.restype = _c_int
    L.oracle_cell.argtypes = [ctypes.c_float, ctypes.c_float, _c_dbl, _c_dbl, _c_dbl, _c_dbl,
                              _c_int] + [ctypes.POINTER(_c_int)] * 4
    for name, gp in (("oracle_gridding_f32", _f32p), ("oracle_gridding_f64", _f64p)):
        fn = getattr(L, name)
        fn.restype = _c_i64
        fn.argtypes = [_f32p, _c_int, _c_int, _c_int, _f32p, _f32p, _i32p, _f64p, _c_dbl,
                       _c_dbl, _c_dbl, _c_int, _c_int, _c_int, gp, gp]
    L.oracle_kb_kernels.argtypes = [_c_int, _c_int, _f64p]
    for name, gp in (("oracle_gridding_kb_f32", _f32p), ("oracle_gridding_kb_f64", _f64p)):
        fn = getattr(L, name)
        fn.restype = _c_i64
        fn.argtypes = [_f32p, _c_int, _c_int, _c_int, _f32p, _f32p, _i32p, _f64p, _c_dbl,
                       _c_dbl, _c_dbl, _c_int, _c_int, _c_int, _c_int, _c_int, gp, gp]
    L.oracle_fft2d_f32.restype = _c_int
    L.oracle_fft2d_f32.argtypes = [_f32p, _f32p, _c_int, _c_int, _c_int]
    L.oracle_fft2d_f64.restype = _c_int
    L.oracle_fft2d_f64.argtypes = [_f64p, _f64p, _c_int, _c_int, _c_int]
    L.oracle_fft_shift_f32.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_size_t]
    L.oracle_fft_shift_f64.argtypes = [_f64p, ctypes.c_size_t, ctypes.c_size_t]
    L.oracle_image_f32.argtypes = [_f32p, _f32p, _c_int, _c_int, _f32p, _c_int]
    L.oracle_image_f64.argtypes = [_f64p, _f64p, _c_int, _c_int, _f64p, _c_int]
    L.oracle_norm_shift_f32.argtypes = [_f32p, _f32p, _c_int, _c_int, _c_int]
    L.oracle_run_f32.argtypes = [_f32p, _c_int, _c_int, _c_int, _f32p, _f32p, _f32p, _i32p,
                                 _f64p, _f64p, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_int,
                                 _c_int, _f32p, _f32p, _f32p, _c_int]
    L.oracle_image_averaging_real.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_size_t, _f32p]
    L.oracle_image_averaging_complex.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_size_t, _f32p]
    _lib = L
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _as_float_view(c: np.ndarray) -> np.ndarray:
    """complex64/complex128 array -> contiguous float view (interleaved)."""
    c = np.ascontiguousarray(c)
    return c.view(np.float32 if c.dtype == np.complex64 else np.float64)


def n_baselines(n_ant: int) -> int:
    return n_ant * (n_ant + 1) // 2


def baseline_flags(n_ant: int, flagged_antennas=()) -> np.ndarray:
    out = np.zeros(n_baselines(n_ant), np.int32)
    fl = _i32(list(flagged_antennas) or [0])
    lib().oracle_baseline_flags(n_ant, fl, len(flagged_antennas), out)
    return out


def uv_max(u_m, v_m, flags):
    um, vm = _c_dbl(), _c_dbl()
    lib().oracle_uv_max(_f32(u_m), _f32(v_m), _i32(flags), len(u_m), ctypes.byref(um),
                        ctypes.byref(vm))
    return um.value, vm.value


def delta_uv(oversampling_factor: float, u_max: float, v_max: float, n_pixels: int):
    du, dv = _c_dbl(), _c_dbl()
    lib().oracle_delta_uv(oversampling_factor, u_max, v_max, n_pixels, ctypes.byref(du),
                          ctypes.byref(dv))
    return du.value, dv.value


def frequency_hz(frequency_resolution_mhz, n_averaged_channels, coarse_bw_mhz, coarse_channel,
                 fine_channel, cotter_compatible=True) -> float:
    return lib().oracle_frequency_hz(frequency_resolution_mhz, n_averaged_channels,
                                     coarse_bw_mhz, coarse_channel, fine_channel,
                                     1 if cotter_compatible else 0)


def zenith_uvw(x, y, z):
    n_ant = len(x)
    nb = n_baselines(n_ant)
    u, v, w = (np.zeros(nb, np.float32) for _ in range(3))
    lib().oracle_zenith_uvw(n_ant, _f64(x), _f64(y), _f64(z), u, v, w)
    return u, v, w


def geometric_corrections(vis: np.ndarray, n_ant: int, w_m, freq_hz) -> np.ndarray:
    """Returns a corrected copy of complex64 vis [t][ch][bl][4]."""
    out = np.array(vis, dtype=np.complex64, copy=True, order="C")
    n_t, n_ch = out.shape[0], out.shape[1]
    lib().oracle_geometric_corrections(_as_float_view(out), n_ant, n_ch, n_t, _f32(w_m),
                                       _f64(freq_hz))
    return out


def cable_corrections(vis: np.ndarray, n_ant: int, cable_m, freq_hz) -> np.ndarray:
    out = np.array(vis, dtype=np.complex64, copy=True, order="C")
    n_t, n_ch = out.shape[0], out.shape[1]
    lib().oracle_cable_corrections(_as_float_view(out), n_ant, n_ch, n_t, _f64(cable_m),
                                   _f64(freq_hz))
    return out


def cell(u_m, v_m, freq_hz, delta_u, delta_v, min_uv, n_pixels):
    r = [_c_int() for _ in range(4)]
    ok = lib().oracle_cell(float(u_m), float(v_m), freq_hz, delta_u, delta_v, min_uv, n_pixels,
                           *[ctypes.byref(x) for x in r])
    return (tuple(x.value for x in r) if ok else None)


def cells(u_m, v_m, flags, freq_hz, delta_u, delta_v, min_uv, n_pixels) -> np.ndarray:
    """int32 [n_ch][n_bl][4] = (u, v, u_conj, v_conj) cell indices, -1 if not gridded."""
    freq = _f64(np.atleast_1d(freq_hz))
    nb = len(u_m)
    out = np.zeros((len(freq), nb, 4), np.int32)
    lib().oracle_cells(_f32(u_m), _f32(v_m), _i32(flags), nb, freq, len(freq), delta_u, delta_v,
                       min_uv, n_pixels, out)
    return out


def gridding(vis, n_ant, u_m, v_m, flags, freq_hz, delta_u, delta_v, min_uv, n_pixels,
             pol="XX", uniform=False, precision="f32", kernel_size=0, with_sinc=True):
    """Grid complex64 vis [t][ch][bl][4].  Returns (grids, counts, n_gridded) with
    grids complex [t][ch][N][N] and counts [t][ch][N][N].  kernel_size > 0
    selects the convolutional (Kaiser-Bessel) mode of the reference."""
    vis = np.ascontiguousarray(vis, dtype=np.complex64)
    n_t, n_ch = vis.shape[0], vis.shape[1]
    shape = (n_t, n_ch, n_pixels, n_pixels)
    kb = kernel_size > 0
    if precision == "f32":
        g = np.zeros(shape, np.complex64)
        c = np.zeros(shape, np.float32)
        fn = lib().oracle_gridding_kb_f32 if kb else lib().oracle_gridding_f32
    else:
        g = np.zeros(shape, np.complex128)
        c = np.zeros(shape, np.float64)
        fn = lib().oracle_gridding_kb_f64 if kb else lib().oracle_gridding_f64
    args = [_as_float_view(vis), n_ant, n_ch, n_t, _f32(u_m), _f32(v_m), _i32(flags),
            _f64(freq_hz), delta_u, delta_v, min_uv, n_pixels,
            POL[pol] if isinstance(pol, str) else int(pol), 1 if uniform else 0]
    if kb:
        args += [int(kernel_size), 1 if with_sinc else 0]
    n = fn(*args, _as_float_view(g), c)
    return g, c, n


def kb_kernels(kernel_size: int, with_sinc: bool = True) -> np.ndarray:
    """The convolutional mode's kernel table [1023][K] (double)."""
    t = np.zeros((1023, kernel_size), np.float64)
    lib().oracle_kb_kernels(int(kernel_size), 1 if with_sinc else 0, t)
    return t


def fft2d(grids: np.ndarray, sign: int = +1) -> np.ndarray:
    """Unnormalised batched 2-D FFT over the last two axes (sign=+1: backward)."""
    g = np.ascontiguousarray(grids)
    N = g.shape[-1]
    batch = int(np.prod(g.shape[:-2])) if g.ndim > 2 else 1
    out = np.zeros_like(g)
    if g.dtype == np.complex64:
        rc = lib().oracle_fft2d_f32(_as_float_view(g), _as_float_view(out), N, batch, sign)
    else:
        g = g.astype(np.complex128)
        out = np.zeros_like(g)
        rc = lib().oracle_fft2d_f64(_as_float_view(g), _as_float_view(out), N, batch, sign)
    if rc != 0:
        raise ValueError(f"oracle FFT does not support N={N}")
    return out


def fft_shift(image: np.ndarray) -> np.ndarray:
    """Returns a shifted copy of one complex image [ny][nx]."""
    out = np.array(image, copy=True, order="C")
    ny, nx = out.shape
    if out.dtype == np.complex64:
        lib().oracle_fft_shift_f32(_as_float_view(out), nx, ny)
    else:
        out = out.astype(np.complex128)
        lib().oracle_fft_shift_f64(_as_float_view(out), nx, ny)
    return out


def image(grids: np.ndarray, counts: np.ndarray, n_threads: int = 0) -> np.ndarray:
    """FFT + 1/sum(counts) + fft_shift of grids [..., N, N] (dtype selects the variant)."""
    N = grids.shape[-1]
    n_img = int(np.prod(grids.shape[:-2])) if grids.ndim > 2 else 1
    if grids.dtype == np.complex64:
        g = np.ascontiguousarray(grids)
        c = _f32(counts)
        out = np.zeros_like(g)
        lib().oracle_image_f32(_as_float_view(g), c, N, n_img, _as_float_view(out), n_threads)
    else:
        g = np.ascontiguousarray(grids, dtype=np.complex128)
        c = _f64(counts)
        out = np.zeros_like(g)
        lib().oracle_image_f64(_as_float_view(g), c, N, n_img, _as_float_view(out), n_threads)
    return out


def norm_shift(images: np.ndarray, counts: np.ndarray, n_threads: int = 0) -> np.ndarray:
    """In place on transformed complex64 images [..., N, N]: multiply each by
    (float)(1/sum(counts)) and fft_shift (the steps of image() after the FFT)."""
    assert images.dtype == np.complex64 and images.flags.c_contiguous
    N = images.shape[-1]
    n_img = int(np.prod(images.shape[:-2])) if images.ndim > 2 else 1
    lib().oracle_norm_shift_f32(_as_float_view(images), _f32(counts), N, n_img, n_threads)
    return images


def run_f32(vis, n_ant, u_m, v_m, w_m, flags, freq_hz, cable_m, apply_geom, apply_cable,
            delta_u, delta_v, min_uv, n_pixels, pol="XX", n_threads: int = 0):
    """The reference CPU pipeline (corrections + gridding + image), FP32 semantics.
    Returns (images, corrected_vis)."""
    v = np.array(vis, dtype=np.complex64, copy=True, order="C")
    n_t, n_ch = v.shape[0], v.shape[1]
    shape = (n_t, n_ch, n_pixels, n_pixels)
    g = np.empty(shape, np.complex64)
    c = np.empty(shape, np.float32)
    out = np.empty(shape, np.complex64)
    lib().oracle_run_f32(_as_float_view(v), n_ant, n_ch, n_t, _f32(u_m), _f32(v_m), _f32(w_m),
                         _i32(flags), _f64(freq_hz), _f64(cable_m), int(apply_geom),
                         int(apply_cable), delta_u, delta_v, min_uv, n_pixels, POL[pol],
                         _as_float_view(g), c, _as_float_view(out), n_threads)
    return out, v


def run_f64(vis, n_ant, u_m, v_m, w_m, flags, freq_hz, cable_m, apply_geom, apply_cable,
            delta_u, delta_v, min_uv, n_pixels, pol="XX", n_threads: int = 0):
    """FP64 oracle image: corrections in double applied to a complex128 copy,
    gridding/FFT/normalisation in double.  Indices follow the same formula."""
    v = np.array(vis, dtype=np.complex128, copy=True)
    n_t, n_ch = v.shape[0], v.shape[1]
    nb = v.shape[2]
    freq = _f64(freq_hz)
    c = 299792458.0
    if apply_geom or apply_cable:
        phase = np.zeros((n_ch, nb))
        if apply_geom:
            phase += 2.0 * np.pi * np.asarray(w_m, np.float64)[None, :] * freq[:, None] / c
        if apply_cable:
            a1, a2 = antennas_of_baselines(nb)
            L = _f64(cable_m)
            phase += -2.0 * np.pi * (L[a2] - L[a1])[None, :] * freq[:, None] / c
        v = v * np.exp(1j * phase)[None, :, :, None]
    g, cnt, _ = _gridding_f64_from_c128(v, n_ant, u_m, v_m, flags, freq, delta_u, delta_v,
                                        min_uv, n_pixels, pol)
    return image(g, cnt, n_threads)


def antennas_of_baselines(n_bl: int):
    bl = np.arange(n_bl, dtype=np.float64)
    a1 = np.floor(-0.5 + np.sqrt(0.25 + 2.0 * bl)).astype(np.int64)
    a2 = (np.arange(n_bl) - (a1 + 1) * a1 // 2).astype(np.int64)
    return a1, a2


def _select_pol_c128(v, pol):
    if pol == "XX":
        return v[..., 0]
    if pol == "YY":
        return v[..., 3]
    if pol == "V":
        return 0.5j * (v[..., 1] - v[..., 2])
    return 0.5 * (v[..., 0] + v[..., 3])


def _gridding_f64_from_c128(v, n_ant, u_m, v_m, flags, freq, delta_u, delta_v, min_uv,
                            n_pixels, pol):
    """Double-precision gridding of complex128 visibilities with the reference's
    cells (the cells come from the C oracle, so indices are bit-identical)."""
    n_t, n_ch, nb = v.shape[:3]
    cl = cells(u_m, v_m, flags, freq, delta_u, delta_v, min_uv, n_pixels)
    N = n_pixels
    g = np.zeros((n_t, n_ch, N, N), np.complex128)
    cnt = np.zeros((n_t, n_ch, N, N), np.float64)
    val = _select_pol_c128(v, pol)
    for t in range(n_t):
        for ch in range(n_ch):
            x = val[t, ch]
            ok = (cl[ch, :, 0] >= 0) & ~np.isnan(x.real) & ~np.isnan(x.imag) \
                & (np.abs(x.real) < 1e12) & (np.abs(x.imag) < 1e12)
            c = cl[ch][ok]
            xs = x[ok]
            p = c[:, 1].astype(np.int64) * N + c[:, 0]
            q = c[:, 3].astype(np.int64) * N + c[:, 2]
            gf = g[t, ch].reshape(-1)
            cf = cnt[t, ch].reshape(-1)
            np.add.at(gf, p, xs)
            np.add.at(gf, q, np.conj(xs))
            np.add.at(cf, p, 1.0)
            np.add.at(cf, q, 1.0)
    return g, cnt, None


def image_averaging_real(images: np.ndarray) -> np.ndarray:
    im = np.ascontiguousarray(images, dtype=np.complex64)
    n_img = im.shape[0]
    size = int(np.prod(im.shape[1:]))
    out = np.zeros(im.shape[1:], np.complex64)
    lib().oracle_image_averaging_real(_as_float_view(im), size, n_img, _as_float_view(out))
    return out


def image_averaging_complex(images: np.ndarray) -> np.ndarray:
    im = np.ascontiguousarray(images, dtype=np.complex64)
    n_img = im.shape[0]
    size = int(np.prod(im.shape[1:]))
    out = np.zeros(im.shape[1:], np.complex64)
    lib().oracle_image_averaging_complex(_as_float_view(im), size, n_img, _as_float_view(out))
    return out


def peak_normalised_error(img, ref):
    """max|img - ref| / max|ref| over complex pixels, and the RMS version."""
    img = np.asarray(img, np.complex128)
    ref = np.asarray(ref, np.complex128)
    peak = np.max(np.abs(ref))
    d = np.abs(img - ref)
    return float(np.max(d) / peak), float(np.sqrt(np.mean(d * d)) / peak)
