"""ctypes binding of the C ABI in include/blink_mi355x.h.

This module is plumbing: it loads ``libblink_mi355x.so`` (built in-tree by
``imager_amd.build``) and turns its return codes into Python exceptions.
Device memory and streams come from PyTorch tensors; every compute call runs
the HIP kernels of the library -- there is no CPU fallback, and loading fails
loudly when the library has not been built.

torch must be imported before the library is loaded: both link the HIP
runtime under the soname ``libamdhip64.so.7`` and the process must hold one
copy of it (torch's).
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the process's HIP runtime first)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libblink_mi355x.so")

BLINK_OK = 0
BLINK_ERR_INVALID_ARGUMENT = 1
BLINK_ERR_RUNTIME = 2
BLINK_ERR_HIP = 3
BLINK_ERR_UNSUPPORTED = 4

POL = {"XX": 0, "YY": 1, "I": 2, "V": 3}
WEIGHTING_NATURAL, WEIGHTING_UNIFORM = 0, 1
WEIGHTING = {"N": WEIGHTING_NATURAL, "": WEIGHTING_NATURAL, "U": WEIGHTING_UNIFORM}
CORR_GEOMETRIC, CORR_CABLE = 1, 2
FFT_FORWARD, FFT_BACKWARD = -1, 1
GRID_TILED, GRID_ATOMIC = 0, 1
KERNEL_KAISER_BESSEL, KERNEL_KAISER_BESSEL_NO_SINC = 1, 2
TRANSFORM_FFT, TRANSFORM_DFT_BF16X3, TRANSFORM_DFT_FP16X3 = 0, 1, 2
MAX_PIXELS = 8192

# Every symbol include/blink_mi355x.h declares.
EXPORTED_SYMBOLS = (
    "blink_create", "blink_destroy", "blink_last_error", "blink_version", "blink_fft_supported",
    "blink_apply_geometric_corrections", "blink_apply_cable_corrections", "blink_gridding",
    "blink_gridding_kernel", "blink_zenith_uvw", "blink_baseline_flags", "blink_uv_max",
    "blink_uv_cells", "blink_fft2d_c2c", "blink_vector_sum", "blink_fft_shift_and_norm",
    "blink_image", "blink_image_averaging", "blink_image_batch", "blink_image_batch_noalloc",
    "blink_release_cache",
    "blink_profile_enable", "blink_profile_get", "blink_profile_kernels", "blink_profile_reset",
    "blink_batch_tiles", "blink_set_batch_tiles",
)


class BlinkError(RuntimeError):
    """A failed call; ``code`` is the blink_status value."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[blink {code}] {message}")
        self.code = code


class BlinkInvalidArgument(BlinkError, ValueError):
    """BLINK_ERR_INVALID_ARGUMENT (the reference's std::invalid_argument)."""


class BlinkUnsupported(BlinkError, ValueError):
    """BLINK_ERR_UNSUPPORTED."""


class ImageSet(ctypes.Structure):
    """struct blink_image_set."""
    _fields_ = [
        ("vis", ctypes.c_void_p),
        ("n_ant", ctypes.c_int), ("n_ch", ctypes.c_int), ("n_t", ctypes.c_int),
        ("u_m", ctypes.c_void_p), ("v_m", ctypes.c_void_p), ("w_m", ctypes.c_void_p),
        ("bl_flags", ctypes.c_void_p), ("freq_hz", ctypes.c_void_p), ("cable_m", ctypes.c_void_p),
        ("corrections", ctypes.c_uint),
        ("delta_u", ctypes.c_double), ("delta_v", ctypes.c_double), ("min_uv", ctypes.c_double),
        ("n_pixels", ctypes.c_int), ("pol", ctypes.c_int),
        ("images", ctypes.c_void_p), ("count_sums", ctypes.c_void_p),
        ("transform", ctypes.c_int), ("weighting", ctypes.c_int),
        ("average", ctypes.c_void_p), ("kernel_size", ctypes.c_int), ("kernel_mode", ctypes.c_int),
    ]


_lib = None


def load():
    """Loads the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m imager_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d, i64, u = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_uint
    sig = {
        "blink_create": (i, [i, ctypes.POINTER(vp)]),
        "blink_destroy": (None, [vp]),
        "blink_last_error": (ctypes.c_char_p, [vp]),
        "blink_version": (ctypes.c_char_p, []),
        "blink_fft_supported": (i, [i]),
        "blink_apply_geometric_corrections": (i, [vp, vp, i, i, i, vp, vp, vp]),
        "blink_apply_cable_corrections": (i, [vp, vp, i, i, i, vp, vp, vp]),
        "blink_gridding": (i, [vp, vp, i, i, i, vp, vp, vp, vp, d, d, d, i, i, i, i, vp, vp, vp]),
        "blink_gridding_kernel": (i, [vp, vp, i, i, i, vp, vp, vp, vp, d, d, d, i, i, i, i, i,
                                      vp, vp, vp]),
        "blink_zenith_uvw": (i, [vp, vp, vp, vp, i, vp, vp, vp, vp]),
        "blink_baseline_flags": (i, [vp, vp, i, vp, vp]),
        "blink_uv_max": (i, [vp, vp, vp, vp, i, vp, vp]),
        "blink_uv_cells": (i, [vp, vp, vp, vp, i, vp, i, d, d, d, i, vp, vp]),
        "blink_fft2d_c2c": (i, [vp, vp, vp, i, i, i, vp]),
        "blink_vector_sum": (i, [vp, vp, i64, i, vp, vp]),
        "blink_fft_shift_and_norm": (i, [vp, vp, i, i, i, vp, i, vp]),
        "blink_image": (i, [vp, vp, vp, i, i, vp, i, vp]),
        "blink_image_averaging": (i, [vp, vp, i64, i, i, vp, vp]),
        "blink_image_batch": (i, [vp, ctypes.POINTER(ImageSet), i, vp]),
        "blink_image_batch_noalloc": (i, [vp, ctypes.POINTER(ImageSet), i, vp]),
        "blink_release_cache": (i, [vp]),
        "blink_profile_enable": (i, [vp, i]),
        "blink_profile_get": (i, [vp, ctypes.c_char_p, ctypes.POINTER(i64),
                                  ctypes.POINTER(ctypes.c_double)]),
        "blink_profile_kernels": (i, [vp, ctypes.c_char_p, i]),
        "blink_profile_reset": (i, [vp]),
        "blink_batch_tiles": (i, [vp, i, ctypes.POINTER(i), ctypes.POINTER(i)]),
        "blink_set_batch_tiles": (i, [vp, i, i, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _ = u
    _lib = L
    return L


def version() -> str:
    return load().blink_version().decode()


def fft_supported(n: int) -> bool:
    return bool(load().blink_fft_supported(int(n)))


def _ptr(t) -> int | None:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    if not t.is_cuda:
        raise ValueError("expected a device (cuda) tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def _stream(stream) -> int | None:
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Context:
    """A blink_ctx on one device.  Methods take torch device tensors."""

    def __init__(self, device: int = 0):
        L = load()
        h = ctypes.c_void_p()
        rc = L.blink_create(int(device), ctypes.byref(h))
        if rc != BLINK_OK:
            raise BlinkError(rc, f"blink_create({device}) failed")
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            load().blink_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def _check(self, rc: int):
        if rc == BLINK_OK:
            return
        msg = load().blink_last_error(self._h).decode()
        if rc == BLINK_ERR_INVALID_ARGUMENT:
            raise BlinkInvalidArgument(rc, msg)
        if rc == BLINK_ERR_UNSUPPORTED:
            raise BlinkUnsupported(rc, msg)
        raise BlinkError(rc, msg)

    # (a)
    def apply_geometric_corrections(self, vis, n_ant, n_ch, n_t, w_m, freq_hz, stream=None):
        self._check(load().blink_apply_geometric_corrections(
            self._h, _ptr(vis), n_ant, n_ch, n_t, _ptr(w_m), _ptr(freq_hz), _stream(stream)))

    def apply_cable_corrections(self, vis, n_ant, n_ch, n_t, cable_m, freq_hz, stream=None):
        self._check(load().blink_apply_cable_corrections(
            self._h, _ptr(vis), n_ant, n_ch, n_t, _ptr(cable_m), _ptr(freq_hz), _stream(stream)))

    # (b)
    def gridding(self, vis, n_ant, n_ch, n_t, u_m, v_m, bl_flags, freq_hz, delta_u, delta_v,
                 min_uv, n_pixels, pol, weighting, grids, counts, algo=GRID_TILED, stream=None):
        self._check(load().blink_gridding(
            self._h, _ptr(vis), n_ant, n_ch, n_t, _ptr(u_m), _ptr(v_m), _ptr(bl_flags),
            _ptr(freq_hz), float(delta_u), float(delta_v), float(min_uv), int(n_pixels),
            POL[pol] if isinstance(pol, str) else int(pol), int(weighting), int(algo),
            _ptr(grids), _ptr(counts), _stream(stream)))

    def gridding_kernel(self, vis, n_ant, n_ch, n_t, u_m, v_m, bl_flags, freq_hz, delta_u,
                        delta_v, min_uv, n_pixels, pol, weighting, kernel_size, grids, counts,
                        kernel_mode=KERNEL_KAISER_BESSEL, stream=None):
        """Convolutional (Kaiser-Bessel) gridding of kernel_size^2 cells per visibility."""
        self._check(load().blink_gridding_kernel(
            self._h, _ptr(vis), n_ant, n_ch, n_t, _ptr(u_m), _ptr(v_m), _ptr(bl_flags),
            _ptr(freq_hz), float(delta_u), float(delta_v), float(min_uv), int(n_pixels),
            POL[pol] if isinstance(pol, str) else int(pol), int(weighting), int(kernel_size),
            int(kernel_mode), _ptr(grids), _ptr(counts), _stream(stream)))

    # array geometry
    def zenith_uvw(self, x, y, z, n_ant, u_m, v_m, w_m, stream=None):
        self._check(load().blink_zenith_uvw(self._h, _ptr(x), _ptr(y), _ptr(z), int(n_ant),
                                            _ptr(u_m), _ptr(v_m), _ptr(w_m), _stream(stream)))

    def baseline_flags(self, ant_flags, n_ant, bl_flags, stream=None):
        self._check(load().blink_baseline_flags(self._h, _ptr(ant_flags), int(n_ant),
                                                _ptr(bl_flags), _stream(stream)))

    def uv_max(self, u_m, v_m, bl_flags, n_bl, out, stream=None):
        self._check(load().blink_uv_max(self._h, _ptr(u_m), _ptr(v_m), _ptr(bl_flags), int(n_bl),
                                        _ptr(out), _stream(stream)))

    def uv_cells(self, u_m, v_m, bl_flags, n_bl, freq_hz, n_ch, delta_u, delta_v, min_uv,
                 n_pixels, cells, stream=None):
        self._check(load().blink_uv_cells(
            self._h, _ptr(u_m), _ptr(v_m), _ptr(bl_flags), int(n_bl), _ptr(freq_hz), int(n_ch),
            float(delta_u), float(delta_v), float(min_uv), int(n_pixels), _ptr(cells),
            _stream(stream)))

    # (c)
    def fft2d_c2c(self, inp, out, n_pixels, batch, direction=FFT_BACKWARD, stream=None):
        self._check(load().blink_fft2d_c2c(self._h, _ptr(inp), _ptr(out), int(n_pixels),
                                           int(batch), int(direction), _stream(stream)))

    # (d)
    def vector_sum(self, values, image_size, n_images, sums, stream=None):
        self._check(load().blink_vector_sum(self._h, _ptr(values), int(image_size), int(n_images),
                                            _ptr(sums), _stream(stream)))

    def fft_shift_and_norm(self, images, nx, ny, n_images, sums=None, n_sums=1, stream=None):
        self._check(load().blink_fft_shift_and_norm(self._h, _ptr(images), int(nx), int(ny),
                                                    int(n_images), _ptr(sums), int(n_sums),
                                                    _stream(stream)))

    def image(self, grids, counts, n_pixels, n_images, images, zero_grids=True, stream=None):
        self._check(load().blink_image(self._h, _ptr(grids), _ptr(counts), int(n_pixels),
                                       int(n_images), _ptr(images), 1 if zero_grids else 0,
                                       _stream(stream)))

    def image_averaging(self, images, image_size, n_images, avg, real_only=False, stream=None):
        self._check(load().blink_image_averaging(self._h, _ptr(images), int(image_size),
                                                 int(n_images), 1 if real_only else 0, _ptr(avg),
                                                 _stream(stream)))

    def image_batch(self, sets, stream=None):
        arr = (ImageSet * len(sets))(*sets)
        self._check(load().blink_image_batch(self._h, arr, len(sets), _stream(stream)))

    def image_batch_noalloc(self, sets, stream=None):
        """blink_image_batch_noalloc: the batched call with every scratch buffer
        in the library's fixed device region (no allocation)."""
        arr = (ImageSet * len(sets))(*sets)
        self._check(load().blink_image_batch_noalloc(self._h, arr, len(sets), _stream(stream)))

    def release_cache(self):
        self._check(load().blink_release_cache(self._h))

    def batch_tiles(self, n_pixels):
        """(lines per first-pass tile, output lines per second-pass block) of the
        batched kernels."""
        r, c = ctypes.c_int(), ctypes.c_int()
        self._check(load().blink_batch_tiles(self._h, int(n_pixels), ctypes.byref(r),
                                             ctypes.byref(c)))
        return r.value, c.value

    def set_batch_tiles(self, n_pixels, row_tile, col_tile):
        """Pins the batched kernels' tiles for one side (0, 0: back to the autotuner)."""
        self._check(load().blink_set_batch_tiles(self._h, int(n_pixels), int(row_tile),
                                                 int(col_tile)))

    # measurement hooks
    def profile_enable(self, on=True):
        self._check(load().blink_profile_enable(self._h, 1 if on else 0))

    def profile_reset(self):
        self._check(load().blink_profile_reset(self._h))

    def profile(self) -> dict:
        """{kernel: (launches, total_ms)} since the last reset (waits for the events)."""
        L = load()
        buf = ctypes.create_string_buffer(4096)
        self._check(L.blink_profile_kernels(self._h, buf, len(buf)))
        out = {}
        for name in filter(None, buf.value.decode().split(",")):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            self._check(L.blink_profile_get(self._h, name.encode(), ctypes.byref(n),
                                            ctypes.byref(ms)))
            out[name] = (n.value, ms.value)
        return out


def make_set(vis, n_ant, n_ch, n_t, u_m, v_m, w_m, bl_flags, freq_hz, cable_m, corrections,
             delta_u, delta_v, min_uv, n_pixels, pol, images, count_sums=None,
             transform=TRANSFORM_FFT, weighting=WEIGHTING_NATURAL, average=None,
             kernel_size=0, kernel_mode=KERNEL_KAISER_BESSEL) -> ImageSet:
    """One struct blink_image_set; `images` may be None when `average` is given."""
    return ImageSet(_ptr(vis), n_ant, n_ch, n_t, _ptr(u_m), _ptr(v_m), _ptr(w_m), _ptr(bl_flags),
                    _ptr(freq_hz), _ptr(cable_m), int(corrections), float(delta_u),
                    float(delta_v), float(min_uv), int(n_pixels),
                    POL[pol] if isinstance(pol, str) else int(pol), _ptr(images),
                    _ptr(count_sums), int(transform),
                    WEIGHTING[weighting] if isinstance(weighting, str) else int(weighting),
                    _ptr(average), int(kernel_size), int(kernel_mode) if kernel_size else 0)
