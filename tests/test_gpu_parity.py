"""GPU parity of every stage, called through the C ABI, against the CPU oracle.

Bars: cell indices and per-cell counts bit-exact; images within 1e-5
peak-normalised of the oracle's reference-precision (FP32) image, with the
error to the FP64 oracle image printed beside it; phase corrections within
1e-6 of |V| of the FP64-maths reference rotation.
"""
import numpy as np
import pytest

import oracle
from imager_amd import synth

pytestmark = pytest.mark.gpu

IMAGE_TOL = 1e-5          # peak-normalised, vs the FP32 oracle (reference precision)
CORR_TOL = 1e-6           # |dV| / |V| for the phase corrections


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _say(*a):
    print(*a, flush=True)


def _obs_dev(o):
    return dict(vis=_dev(o.vis), u=_dev(o.u), v=_dev(o.v), w=_dev(o.w), flags=_dev(o.flags),
                freq=_dev(o.freq), cable=_dev(o.cable))


# ---------------------------------------------------------------- cells

@pytest.mark.parametrize("case", ["eda2_180", "mwa_1024", "small_odd"])
def test_uv_cells_bit_exact(ctx, case):
    import torch
    if case == "eda2_180":
        o = synth.eda2_station(n_ch=4, freq=np.linspace(150e6, 190e6, 4))
    elif case == "mwa_1024":
        o = synth.mwa_array(n_ch=16, n_t=1)
    else:
        o = synth.small(n_ant=30, n_ch=5, n_t=1, n_pixels=45, seed=3)
    _say("case", case, o.n_ant, o.n_ch, o.n_pixels)
    ref = oracle.cells(o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv, o.n_pixels)
    d = _obs_dev(o)
    cells = torch.empty((o.n_ch, o.n_bl, 4), dtype=torch.int32, device="cuda")
    ctx.uv_cells(d["u"], d["v"], d["flags"], o.n_bl, d["freq"], o.n_ch, o.delta_u, o.delta_v,
                 o.min_uv, o.n_pixels, cells)
    assert np.array_equal(_host(cells), ref)


def test_uv_cells_exact_halves_and_min_uv(ctx):
    import torch
    c = 299792458.0
    u = np.array([2.5, -2.5, 0.0, 1.0, 3.0, -7.0, -8.0, 9.0], np.float32)
    v = np.array([0.0, 0.0, -0.5, 1.0, 0.0, 0.0, 0.0, 0.0], np.float32)
    flags = np.zeros(8, np.int32)
    freq = np.array([c])
    for min_uv in (-np.inf, 2.0):
        ref = oracle.cells(u, v, flags, freq, 1.0, 1.0, min_uv, 8)
        cells = torch.empty((1, 8, 4), dtype=torch.int32, device="cuda")
        ctx.uv_cells(_dev(u), _dev(v), _dev(flags), 8, _dev(freq), 1, 1.0, 1.0, min_uv, 8, cells)
        assert np.array_equal(_host(cells), ref)


# ---------------------------------------------------------- corrections

def test_phase_corrections(ctx):
    o = synth.small(n_ant=40, n_ch=6, n_t=3, n_pixels=64, seed=21, height_sigma_m=5.0)
    d = _obs_dev(o)
    ref = oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq)
    ctx.apply_geometric_corrections(d["vis"], o.n_ant, o.n_ch, o.n_t, d["w"], d["freq"])
    got = _host(d["vis"])
    err = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30))
    _say("geometric max rel err", err)
    assert err < CORR_TOL
    ref2 = oracle.cable_corrections(ref, o.n_ant, o.cable, o.freq)
    ctx.apply_cable_corrections(d["vis"], o.n_ant, o.n_ch, o.n_t, d["cable"], d["freq"])
    got2 = _host(d["vis"])
    err2 = np.max(np.abs(got2 - ref2) / np.maximum(np.abs(ref2), 1e-30))
    _say("cable max rel err", err2)
    assert err2 < 2 * CORR_TOL


def test_phase_corrections_remove_w_phase(ctx):
    """Closed form: a pure w-phase exp(-2 pi i w/lambda) corrects to 1."""
    rng = np.random.default_rng(9)
    n_ant = 64
    x, y, z = rng.normal(0, 300, n_ant), rng.normal(0, 300, n_ant), rng.normal(0, 10, n_ant)
    u, v, w = synth.zenith_uvw(x, y, z)
    freq = np.array([170e6, 200e6])
    lam = 299792458.0 / freq
    vis = np.exp(-2j * np.pi * w.astype(np.float64)[None, :] / lam[:, None])
    vis = np.repeat(vis[None, :, :, None], 4, axis=3).astype(np.complex64)
    dv = _dev(vis)
    ctx.apply_geometric_corrections(dv, n_ant, 2, 1, _dev(w), _dev(freq))
    assert np.max(np.abs(_host(dv) - 1.0)) < 1e-6


# ------------------------------------------------------------- gridding

@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("pol", ["XX", "YY", "I", "V"])
def test_gridding_counts_exact_grid_close(ctx, algo, pol):
    import torch
    o = synth.small(n_ant=48, n_ch=3, n_t=2, n_pixels=96, seed=4)
    _say("gridding algo", algo, "pol", pol)
    g_ref, c_ref, _ = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u,
                                      o.delta_v, o.min_uv, o.n_pixels, pol)
    d = _obs_dev(o)
    shape = (o.n_t, o.n_ch, o.n_pixels, o.n_pixels)
    grids = torch.zeros(shape, dtype=torch.complex64, device="cuda")
    counts = torch.zeros(shape, dtype=torch.float32, device="cuda")
    ctx.gridding(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["flags"], d["freq"],
                 o.delta_u, o.delta_v, o.min_uv, o.n_pixels, pol, 0, grids, counts, algo=algo)
    assert np.array_equal(_host(counts), c_ref)
    err = np.max(np.abs(_host(grids) - g_ref))
    assert err <= 1e-5 * np.max(np.abs(g_ref))


def test_gridding_accumulates_and_uniform(ctx):
    import torch
    o = synth.small(n_ant=32, n_ch=2, n_t=1, n_pixels=64, seed=8)
    d = _obs_dev(o)
    shape = (o.n_t, o.n_ch, o.n_pixels, o.n_pixels)
    grids = torch.zeros(shape, dtype=torch.complex64, device="cuda")
    counts = torch.zeros(shape, dtype=torch.float32, device="cuda")
    args = (d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["flags"], d["freq"], o.delta_u,
            o.delta_v, o.min_uv, o.n_pixels, "XX")
    ctx.gridding(*args, 0, grids, counts)
    ctx.gridding(*args, 0, grids, counts)   # grids accumulate across calls
    g1, c1, _ = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                                o.min_uv, o.n_pixels, "XX")
    assert np.array_equal(_host(counts), 2 * c1)
    assert np.max(np.abs(_host(grids) - 2 * g1)) <= 1e-5 * np.max(np.abs(2 * g1))
    grids.zero_()
    counts.zero_()
    ctx.gridding(*args, 1, grids, counts)
    gu, cu, _ = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                                o.min_uv, o.n_pixels, "XX", uniform=True)
    gg = _host(grids)
    assert np.array_equal(np.isnan(gg), np.isnan(gu))
    occ = ~np.isnan(gu)
    assert np.max(np.abs(gg[occ] - gu[occ])) <= 1e-5 * np.max(np.abs(gu[occ]))


def test_gridding_skips_nan_and_maxvis(ctx):
    import torch
    o = synth.small(n_ant=20, n_ch=1, n_t=1, n_pixels=40, seed=12)
    vis = o.vis.copy()
    good = np.where(o.flags == 0)[0]
    vis[0, 0, good[0], 0] = np.nan
    vis[0, 0, good[1], 0] = 3e12
    vis[0, 0, good[2], 0] = 1.00000007e12j
    vis[0, 0, good[3], 0] = 1e12j
    _, c_ref, _ = oracle.gridding(vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                                  o.min_uv, o.n_pixels, "XX")
    for algo in (0, 1):
        grids = torch.zeros((1, 1, 40, 40), dtype=torch.complex64, device="cuda")
        counts = torch.zeros((1, 1, 40, 40), dtype=torch.float32, device="cuda")
        ctx.gridding(_dev(vis), o.n_ant, 1, 1, _dev(o.u), _dev(o.v), _dev(o.flags), _dev(o.freq),
                     o.delta_u, o.delta_v, o.min_uv, 40, "XX", 0, grids, counts, algo=algo)
        assert np.array_equal(_host(counts), c_ref)


# ------------------------------------------------------------------ FFT

@pytest.mark.parametrize("n,batch", [(2, 3), (7, 2), (8, 2), (13, 2), (15, 2), (45, 1), (64, 3),
                                     (77, 2), (128, 2), (143, 1), (180, 3), (256, 2), (512, 1),
                                     (1024, 2), (2002, 1), (2048, 1), (4096, 1), (8192, 1)])
@pytest.mark.parametrize("direction", [1, -1])
def test_fft2d_matches_numpy(ctx, n, batch, direction):
    import torch
    _say("fft", n, batch, direction)
    rng = np.random.default_rng(n + batch)
    x = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n)))
    ref = np.fft.ifft2(x) * n * n if direction > 0 else np.fft.fft2(x)
    dx = _dev(x.astype(np.complex64))
    out = torch.empty_like(dx)
    ctx.fft2d_c2c(dx, out, n, batch, direction)
    got = _host(out)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    _say("  rel err", err)
    assert err < 3e-6
    # in place
    ctx.fft2d_c2c(dx, dx, n, batch, direction)
    assert np.array_equal(_host(dx), got)


# ---------------------------------------------- shift, norm, sums, averaging

def test_fft_shift_and_norm_kats(ctx):
    import torch
    # tests/gpu_utils_test.cpp:60-87: 3 images of 2x2, fnorm = 1
    a = np.tile(np.array([0, 1, 2, 3], np.complex64), 3)
    da = _dev(a)
    sums = _dev(np.ones(3))
    ctx.fft_shift_and_norm(da, 2, 2, 3, sums, 3)
    assert np.array_equal(_host(da), np.tile(np.array([3, 2, 1, 0], np.complex64), 3))
    # tests/imager_test.cpp:30-37 odd size 3x2 (the reference GPU throws here;
    # the CPU swap semantics are kept)
    b = _dev(np.arange(6).astype(np.complex64))
    ctx.fft_shift_and_norm(b, 3, 2, 1, None)
    assert np.array_equal(_host(b), np.arange(6)[::-1].astype(np.complex64))
    _ = torch


def test_fft_shift_and_norm_random(ctx):
    rng = np.random.default_rng(5)
    for n in (6, 7, 180):
        imgs = (rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))).astype(np.complex64)
        sums = np.array([12345.0, 7.0])
        d = _dev(imgs)
        ctx.fft_shift_and_norm(d, n, n, 2, _dev(sums), 2)
        got = _host(d)
        for k in range(2):
            f = np.float32(1.0 / sums[k])
            ref = oracle.fft_shift(imgs[k] * f)
            assert np.array_equal(got[k], ref)


def test_vector_sum_kat(ctx):
    import torch
    # tests/gpu_utils_test.cpp:20-57: 5 x 3000 values rand() % 50, exact
    rng = np.random.default_rng(50)
    vals = rng.integers(0, 50, (5, 3000)).astype(np.float32)
    sums = torch.empty(5, dtype=torch.float64, device="cuda")
    ctx.vector_sum(_dev(vals), 3000, 5, sums)
    assert np.array_equal(_host(sums), vals.astype(np.float64).sum(1))


def test_image_averaging_kat(ctx):
    import torch
    # tests/gpu_utils_test.cpp:91-124
    n_img, side = 4, 512
    vals = (np.arange(side * side) % 100).astype(np.float32)
    imgs = np.repeat(vals[None], n_img, 0).astype(np.complex64)
    out = torch.empty(side * side, dtype=torch.complex64, device="cuda")
    ctx.image_averaging(_dev(imgs), side * side, n_img, out)
    assert np.array_equal(_host(out), vals.astype(np.complex64))
    rng = np.random.default_rng(1)
    imgs = (rng.standard_normal((3, 100)) + 1j * rng.standard_normal((3, 100))).astype(np.complex64)
    out = torch.empty(100, dtype=torch.complex64, device="cuda")
    ctx.image_averaging(_dev(imgs), 100, 3, out)
    assert np.array_equal(_host(out), oracle.image_averaging_complex(imgs))
    ctx.image_averaging(_dev(imgs), 100, 3, out, real_only=True)
    assert np.array_equal(_host(out), oracle.image_averaging_real(imgs))


# ---------------------------------------------------------------- image

@pytest.mark.parametrize("n", [64, 180, 256, 45])
def test_image_from_grids(ctx, n):
    import torch
    o = synth.small(n_ant=40, n_ch=2, n_t=2, n_pixels=n, seed=n)
    _say("image", n)
    g, c, _ = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, n, "XX")
    ref = oracle.image(g, c)
    g64, c64, _ = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u,
                                  o.delta_v, o.min_uv, n, "XX", precision="f64")
    ref64 = oracle.image(g64, c64)
    dg, dc = _dev(g), _dev(c)
    imgs = torch.empty_like(dg)
    ctx.image(dg, dc, n, o.n_images, imgs, zero_grids=True)
    got = _host(imgs)
    e32 = oracle.peak_normalised_error(got, ref)
    e64 = oracle.peak_normalised_error(got, ref64)
    _say(f"  N={n} err vs FP32 oracle {e32[0]:.2e} (rms {e32[1]:.2e}), vs FP64 {e64[0]:.2e}")
    assert e32[0] < IMAGE_TOL
    assert not _host(dg).any() and not _host(dc).any()   # zeroed for the next round


def test_end_to_end_plain_path(ctx):
    """corrections -> gridding -> image through the stage entry points, as the
    reference's grid()/image() call them, vs the oracle's CPU pipeline."""
    import torch
    o = synth.small(n_ant=64, n_ch=4, n_t=2, n_pixels=128, seed=31, height_sigma_m=2.0)
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, o.n_pixels)
    ref64 = oracle.run_f64(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                           o.delta_u, o.delta_v, o.min_uv, o.n_pixels)
    d = _obs_dev(o)
    ctx.apply_geometric_corrections(d["vis"], o.n_ant, o.n_ch, o.n_t, d["w"], d["freq"])
    ctx.apply_cable_corrections(d["vis"], o.n_ant, o.n_ch, o.n_t, d["cable"], d["freq"])
    shape = (o.n_t, o.n_ch, o.n_pixels, o.n_pixels)
    grids = torch.zeros(shape, dtype=torch.complex64, device="cuda")
    counts = torch.zeros(shape, dtype=torch.float32, device="cuda")
    ctx.gridding(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["flags"], d["freq"],
                 o.delta_u, o.delta_v, o.min_uv, o.n_pixels, "XX", 0, grids, counts)
    imgs = torch.empty_like(grids)
    ctx.image(grids, counts, o.n_pixels, o.n_images, imgs)
    got = _host(imgs)
    e32 = oracle.peak_normalised_error(got, ref)
    e64 = oracle.peak_normalised_error(got, ref64)
    _say(f"plain path err vs FP32 oracle {e32[0]:.2e}, vs FP64 oracle {e64[0]:.2e}")
    assert e32[0] < IMAGE_TOL


@pytest.mark.parametrize("n,K,mode,pol,uniform", [(128, 7, 1, "XX", False),
                                                  (180, 10, 1, "I", False),
                                                  (64, 5, 2, "V", False),
                                                  (256, 1, 1, "YY", False),
                                                  (96, 7, 1, "XX", True)])
def test_gridding_kernel_matches_oracle(ctx, n, K, mode, pol, uniform):
    """Convolutional (Kaiser-Bessel) gridding vs the oracle's restatement of
    convolve_with_kernel: counts exact, grids within FP32 summation-order
    differences (1e-6 of the largest cell); uniform weighting keeps the
    reference's NaN / infinity in every cell without a nearest-cell count."""
    import torch
    from imager_amd import capi
    o = synth.small(n_ant=40, n_ch=2, n_t=2, n_pixels=n, seed=70 + K, height_sigma_m=1.0)
    d = _obs_dev(o)
    shape = (o.n_t, o.n_ch, n, n)
    grids = torch.zeros(shape, dtype=torch.complex64, device="cuda")
    counts = torch.zeros(shape, dtype=torch.float32, device="cuda")
    ctx.gridding_kernel(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["flags"], d["freq"],
                        o.delta_u, o.delta_v, o.min_uv, n, pol, 1 if uniform else 0, K, grids,
                        counts, kernel_mode=mode)
    g_ref, c_ref, _ = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u,
                                      o.delta_v, o.min_uv, n, pol, uniform=uniform,
                                      kernel_size=K, with_sinc=(mode == 1))
    g = _host(grids)
    assert np.array_equal(_host(counts), c_ref)
    if uniform:
        # x / 0 in every cell without a nearest-cell count (NaN or an
        # infinity, as the reference divides; which one depends on whether
        # mirrored contributions cancel exactly, i.e. on summation order)
        bad = ~np.isfinite(g_ref)
        assert np.array_equal(bad, c_ref == 0)
        assert np.array_equal(~np.isfinite(g), bad)
        g, g_ref = g[~bad], g_ref[~bad]
    err = np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref))
    _say(f"  kernel gridding N={n} K={K} mode={mode} {pol}: {err:.2e}")
    assert err < 1e-6


@pytest.mark.parametrize("n_ant,flagged", [(1, ()), (48, (0, 7, 47)), (256, (3,)),
                                            (512, ())])
def test_array_geometry_matches_oracle(ctx, n_ant, flagged):
    """Zenith UVW, baseline flags and u_max / v_max on the GPU, bit for bit
    with the oracle (CalculateUVW, UpdateFlags)."""
    import torch
    rng = np.random.default_rng(n_ant)
    x = rng.normal(0, 300.0, n_ant)
    y = rng.normal(0, 300.0, n_ant)
    z = rng.normal(0, 2.0, n_ant)
    n_bl = n_ant * (n_ant + 1) // 2
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    u = torch.empty(n_bl, dtype=torch.float32, device="cuda")
    v = torch.empty_like(u)
    w = torch.empty_like(u)
    ctx.zenith_uvw(dev(x), dev(y), dev(z), n_ant, u, v, w)
    ur, vr, wr = oracle.zenith_uvw(x, y, z)
    assert np.array_equal(_host(u), ur) and np.array_equal(_host(v), vr)
    assert np.array_equal(_host(w), wr)
    ant = np.zeros(n_ant, np.int32)
    ant[list(flagged)] = 1
    fl = torch.empty(n_bl, dtype=torch.int32, device="cuda")
    ctx.baseline_flags(dev(ant), n_ant, fl)
    assert np.array_equal(_host(fl), oracle.baseline_flags(n_ant, list(flagged)))
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    ctx.uv_max(u, v, fl, n_bl, out)
    um, vm = oracle.uv_max(ur, vr, oracle.baseline_flags(n_ant, list(flagged)))
    assert _host(out)[0] == um and _host(out)[1] == vm


# -------------------------------------------------------------- batched

def _batch(ctx, o, corrections=3, pol="XX"):
    import torch
    from imager_amd import capi
    d = _obs_dev(o)
    imgs = torch.empty((o.n_t, o.n_ch, o.n_pixels, o.n_pixels), dtype=torch.complex64,
                       device="cuda")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], corrections, o.delta_u, o.delta_v, o.min_uv,
                       o.n_pixels, pol, imgs, sums)
    return d, imgs, sums, st


# (128, 100): 5050 baselines over 65 plan rows, a dense grid whose row tiles
# gather the records; the sparser register-path cases gather compacted
# visibilities
@pytest.mark.parametrize("n,n_ant,pol", [(128, 48, "XX"), (180, 64, "I"), (256, 64, "V"),
                                         (1024, 128, "XX"), (2048, 64, "YY"), (4096, 32, "XX"),
                                         (90, 24, "XX"), (150, 32, "YY"), (128, 100, "I")])
def test_image_batch_matches_reference_pipeline(ctx, n, n_ant, pol):
    if (n, n_ant) == (1024, 128):
        # the 128-tile Gaussian array of the 1024^2 benchmark (1.5 km)
        o = synth.mwa_array(n_ch=3, n_t=2, n_pixels=1024, seed=1002, height_sigma_m=2.0)
    else:
        o = synth.small(n_ant=n_ant, n_ch=3, n_t=2, n_pixels=n, seed=n + n_ant,
                        height_sigma_m=2.0)
    _say("batch", n, n_ant, pol)
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, n, pol)
    ref64 = oracle.run_f64(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                           o.delta_u, o.delta_v, o.min_uv, n, pol)
    d, imgs, sums, st = _batch(ctx, o, 3, pol)
    ctx.image_batch([st])
    got = _host(imgs)
    e32 = oracle.peak_normalised_error(got, ref)
    e64 = oracle.peak_normalised_error(got, ref64)
    _say(f"  err vs FP32 oracle {e32[0]:.2e} (rms {e32[1]:.2e}), vs FP64 {e64[0]:.2e}")
    assert e32[0] < IMAGE_TOL
    _, c_ref, _ = oracle.gridding(oracle.cable_corrections(
        oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq), o.n_ant, o.cable, o.freq),
        o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv, n, pol)
    assert np.array_equal(_host(sums), c_ref.reshape(o.n_images, -1).sum(1).astype(np.float64))


@pytest.mark.parametrize("n", [128, 256, 512, 1024, 2048, 4096])
def test_image_batch_every_tile_variant(ctx, n):
    """Every built row-tile and column-block variant (pinned through
    blink_set_batch_tiles; unbuilt sizes are refused) against the oracle."""
    from imager_amd import capi
    o = synth.small(n_ant=48, n_ch=2, n_t=2, n_pixels=n, seed=3 * n, height_sigma_m=2.0)
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, n)
    d, imgs, sums, st = _batch(ctx, o)
    cands = (1, 2, 4, 8, 16, 32, 64, 128)
    built = {"row": [], "col": []}
    r_def, c_def = (abs(t) for t in ctx.batch_tiles(n))
    try:
        for kind in ("row", "col"):
            for t in cands:
                rt, ct = (t, c_def) if kind == "row" else (r_def, t)
                try:
                    ctx.set_batch_tiles(n, rt, ct)
                except capi.BlinkInvalidArgument:
                    continue
                built[kind].append(t)
                _say(f"batch N={n} tiles rows={rt} cols={ct}")
                imgs.fill_(float("nan"))
                ctx.image_batch([st])
                got = _host(imgs)
                assert oracle.peak_normalised_error(got, ref)[0] < IMAGE_TOL, (rt, ct)
                # the gridded plane is Hermitian: the images are real
                assert not np.any(got.imag), (rt, ct)
    finally:
        ctx.set_batch_tiles(n, 0, 0)
    assert len(built["row"]) >= 2 and len(built["col"]) >= 3, built


@pytest.mark.parametrize("n", [256, 1024, 180])
def test_image_batch_sparse_rows_nan_and_maxvis(ctx, n):
    """A 4-antenna array leaves almost every grid row empty (the register row
    tiles then hold only the few non-empty rows); NaN and |V| >= 1e12
    visibilities planted in the batched path are skipped as the reference
    does (src/pacer_imager.cpp:327-329), in the images and the count sums."""
    o = synth.small(n_ant=4, n_ch=3, n_t=2, n_pixels=n, seed=40 + n, height_sigma_m=1.0)
    good = np.flatnonzero(o.flags == 0)
    o.vis[0, 1, good[0], 0] = np.nan
    o.vis[1, 2, good[1], 0] = np.complex64(3e12)
    o.vis[1, 0, good[2], 0] = np.complex64(complex(0.5, -2e12))
    _say("sparse batch", n)
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, n)
    d, imgs, sums, st = _batch(ctx, o)
    ctx.image_batch([st])
    got = _host(imgs)
    assert oracle.peak_normalised_error(got, ref)[0] < IMAGE_TOL
    _, c_ref, _ = oracle.gridding(oracle.cable_corrections(
        oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq), o.n_ant, o.cable, o.freq),
        o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv, n, "XX")
    expect = c_ref.reshape(o.n_images, -1).sum(1).astype(np.float64)
    assert np.array_equal(_host(sums), expect)
    assert expect[1] == expect[0] - 2 and expect[5] == expect[0] - 2  # one visibility + conjugate


def test_image_batch_fully_flagged_image_set(ctx):
    """Every antenna flagged: no visibility is gridded, the count sums are 0 and
    the images are 0 * (float)(1/0) = NaN, the reference's arithmetic
    (src/pacer_imager.cpp:171-178) on an empty grid; a second set in the same
    call is unaffected."""
    import torch
    from imager_amd import capi
    o = synth.small(n_ant=16, n_ch=2, n_t=1, n_pixels=256, seed=61,
                    flagged_antennas=tuple(range(16)))
    o2 = synth.small(n_ant=16, n_ch=2, n_t=1, n_pixels=256, seed=62)
    d, imgs, sums, st = _batch(ctx, o)
    d2, imgs2, sums2, st2 = _batch(ctx, o2)
    ctx.image_batch([st, st2])
    assert np.all(_host(sums) == 0)
    assert np.all(np.isnan(_host(imgs).real))
    ref2, _ = oracle.run_f32(o2.vis, o2.n_ant, o2.u, o2.v, o2.w, o2.flags, o2.freq, o2.cable,
                             True, True, o2.delta_u, o2.delta_v, o2.min_uv, 256)
    assert oracle.peak_normalised_error(_host(imgs2), ref2)[0] < IMAGE_TOL


def test_image_batch_graph_replays_are_identical(ctx):
    import torch
    o = synth.small(n_ant=64, n_ch=4, n_t=2, n_pixels=256, seed=77)
    d, imgs, sums, st = _batch(ctx, o)
    ctx.image_batch([st])
    first = _host(imgs).copy()
    for k in range(3):
        imgs.fill_(float("nan"))
        scratch = torch.randn(1 << 20, device="cuda")   # eager work between replays
        scratch.mul_(2.0)
        ctx.image_batch([st])
        got = _host(imgs)
        assert not np.isnan(got).any()
        assert oracle.peak_normalised_error(got, first)[0] < 1e-6, k


@pytest.mark.parametrize("n,dft", [(1024, False), (180, False), (128, True)])
def test_image_batch_new_buffers_update_the_graph(ctx, n, dft):
    """A call repeated with other vis / images / count_sums buffers (the
    fields that do not key the graph cache): two buffer sets keep a graph
    each, a third re-points (updates) the least recently used one.  Every
    call's outputs land in that call's buffers and match the oracle, and
    going back to earlier buffers gives their images again."""
    import torch
    from imager_amd import capi
    o = synth.small(n_ant=48, n_ch=2, n_t=2, n_pixels=n, seed=300 + n, height_sigma_m=1.0)
    o2 = synth.small(n_ant=48, n_ch=2, n_t=2, n_pixels=n, seed=300 + n, height_sigma_m=1.0,
                     n_src=5)
    assert np.array_equal(o.u, o2.u) and not np.array_equal(o.vis, o2.vis)
    d = _obs_dev(o)
    vis2 = _dev(o2.vis)
    tr = capi.TRANSFORM_DFT_BF16X3 if dft else capi.TRANSFORM_FFT
    bufs = [(d["vis"], torch.empty((o.n_t, o.n_ch, n, n), dtype=torch.complex64, device="cuda"),
             torch.empty(o.n_images, dtype=torch.float64, device="cuda")) for _ in range(3)]
    bufs[1] = (vis2, bufs[1][1], bufs[1][2])
    refs = []
    for obs in (o, o2):
        r, _ = oracle.run_f32(obs.vis, obs.n_ant, obs.u, obs.v, obs.w, obs.flags, obs.freq,
                              obs.cable, True, True, obs.delta_u, obs.delta_v, obs.min_uv, n)
        refs.append(r)
    refs.append(refs[0])   # buffer set 2: the first visibilities, other outputs
    for k in (0, 1, 0, 1, 2, 1, 0, 2):
        vis, imgs, sums = bufs[k]
        imgs.fill_(float("nan"))
        st = capi.make_set(vis, o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                           d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, n, "XX",
                           imgs, sums, transform=tr)
        _say(f"buffers {k}, N={n}, dft={dft}")
        ctx.image_batch([st])
        err = oracle.peak_normalised_error(_host(imgs), refs[k])[0]
        assert err < IMAGE_TOL, (k, err)
    # the other buffer pair still holds its own images (not overwritten)
    assert oracle.peak_normalised_error(_host(bufs[0][1]), refs[0])[0] < IMAGE_TOL


def test_image_batch_two_sets(ctx):
    a = synth.small(n_ant=40, n_ch=2, n_t=1, n_pixels=128, seed=1)
    b = synth.small(n_ant=56, n_ch=3, n_t=2, n_pixels=180, seed=2)
    da, ia, sa, sta = _batch(ctx, a, 0)
    db, ib, sb, stb = _batch(ctx, b, 0)
    ctx.image_batch([sta, stb])
    for o, im in ((a, ia), (b, ib)):
        ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, False,
                                False, o.delta_u, o.delta_v, o.min_uv, o.n_pixels)
        assert oracle.peak_normalised_error(_host(im), ref)[0] < IMAGE_TOL


@pytest.mark.parametrize("n,dft", [(256, False), (1024, False), (180, False), (64, True)])
def test_image_batch_min_uv(ctx, n, dft):
    """A finite min_uv (the reference's sqrt(u^2 + v^2) <= min_uv skip) takes
    the plan's exact cell route: with a third of the baselines cut, images and
    exact count sums against the oracle on the register (256, 1024: transposed
    plan), generic (180) and DFT (64) paths."""
    import dataclasses
    import torch
    from imager_amd import capi
    o = synth.small(n_ant=40, n_ch=2, n_t=2, n_pixels=n, seed=900 + n, height_sigma_m=1.0)
    lam = 299792458.0 / o.freq[0]
    uvd = np.hypot(o.u.astype(np.float64), o.v.astype(np.float64)) / lam
    o = dataclasses.replace(o, min_uv=float(np.quantile(uvd[o.flags == 0], 0.33)))
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, n)
    _, c_ref, n_grid = oracle.gridding(oracle.cable_corrections(
        oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq), o.n_ant, o.cable, o.freq),
        o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv, n, "XX")
    _, c_all, n_all = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u,
                                      o.delta_v, -np.inf, n, "XX")
    assert 0 < n_grid < 0.8 * n_all   # the cut removes a good share of the visibilities
    d = _obs_dev(o)
    imgs = torch.empty((o.n_t, o.n_ch, n, n), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, n, "XX", imgs,
                       sums, transform=capi.TRANSFORM_DFT_BF16X3 if dft else capi.TRANSFORM_FFT)
    _say(f"batch min_uv={o.min_uv:.3g} N={n} dft={dft}")
    ctx.image_batch([st])
    got = _host(imgs)
    assert oracle.peak_normalised_error(got, ref)[0] < IMAGE_TOL
    assert np.array_equal(_host(sums), c_ref.reshape(o.n_images, -1).sum(1).astype(np.float64))


def test_image_batch_compacted_and_record_sets_in_one_group(ctx):
    """One 128^2 group holding a dense set (100 antennas, 5050 baselines over
    65 plan rows: its row tiles gather the records) and a sparse one (40
    antennas: compacted visibilities), so that one row launch takes both
    paths set by set: images and exact count sums against the oracle."""
    import torch
    from imager_amd import capi
    obs, keep, sets = [], [], []
    for k, n_ant in enumerate((100, 40)):
        o = synth.small(n_ant=n_ant, n_ch=2, n_t=2, n_pixels=128, seed=950 + k,
                        height_sigma_m=1.0, bl_flag_fraction=0.1)
        d = _obs_dev(o)
        imgs = torch.empty((o.n_t, o.n_ch, 128, 128), dtype=torch.complex64, device="cuda")
        sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
        sets.append(capi.make_set(d["vis"], n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"],
                                  d["flags"], d["freq"], d["cable"], 3, o.delta_u, o.delta_v,
                                  o.min_uv, 128, "XX", imgs, sums))
        obs.append(o)
        keep.append((d, imgs, sums))
    for rep in range(2):   # capture, then replay
        for _, imgs, _ in keep:
            imgs.fill_(float("nan"))
        _say(f"compacted + record sets, rep {rep}")
        ctx.image_batch(sets)
        for o, (_, imgs, sums) in zip(obs, keep):
            ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True,
                                    True, o.delta_u, o.delta_v, o.min_uv, 128)
            assert oracle.peak_normalised_error(_host(imgs), ref)[0] < IMAGE_TOL, (o.n_ant, rep)
            _, c_ref, _ = oracle.gridding(oracle.cable_corrections(
                oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq), o.n_ant, o.cable,
                o.freq), o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv,
                128, "XX")
            assert np.array_equal(_host(sums),
                                  c_ref.reshape(o.n_images, -1).sum(1).astype(np.float64))


def test_image_batch_grouped_sets(ctx):
    """Several sets per grid side in one call (one launch pair per side):
    different arrays, channel and timestep counts, flags and polarisations,
    register (256, 128) and generic (180) sides, plus a DFT set."""
    import torch
    from imager_amd import capi
    specs = [(256, 40, 2, 1, "XX", 0), (256, 64, 1, 3, "I", 0), (180, 48, 3, 1, "XX", 0),
             (256, 24, 2, 2, "YY", 0), (128, 32, 1, 1, "V", 0), (180, 40, 1, 2, "XX", 0),
             (64, 32, 2, 1, "XX", 1)]
    obs, keep, sets = [], [], []
    for k, (n, n_ant, n_ch, n_t, pol, dft) in enumerate(specs):
        o = synth.small(n_ant=n_ant, n_ch=n_ch, n_t=n_t, n_pixels=n, seed=500 + k,
                        height_sigma_m=1.0, flagged_antennas=(k % 5, 7), bl_flag_fraction=0.1)
        d = _obs_dev(o)
        imgs = torch.empty((n_t, n_ch, n, n), dtype=torch.complex64, device="cuda")
        sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
        st = capi.make_set(d["vis"], n_ant, n_ch, n_t, d["u"], d["v"], d["w"], d["flags"],
                           d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, n, pol, imgs,
                           sums, transform=capi.TRANSFORM_DFT_BF16X3 if dft else
                           capi.TRANSFORM_FFT)
        obs.append((o, pol))
        keep.append((d, imgs, sums))
        sets.append(st)
    for rep in range(2):   # capture, then replay
        for _, imgs, _ in keep:
            imgs.fill_(float("nan"))
        ctx.image_batch(sets)
        for (o, pol), (_, imgs, sums) in zip(obs, keep):
            ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True,
                                    True, o.delta_u, o.delta_v, o.min_uv, o.n_pixels, pol)
            err = oracle.peak_normalised_error(_host(imgs), ref)[0]
            _say(f"  grouped N={o.n_pixels} {o.n_ant} ant {o.n_ch}x{o.n_t} {pol}: {err:.2e}")
            assert err < IMAGE_TOL, (o.n_pixels, pol, rep)
            _, c_ref, _ = oracle.gridding(oracle.cable_corrections(
                oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq), o.n_ant, o.cable,
                o.freq), o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv,
                o.n_pixels, pol)
            assert np.array_equal(_host(sums),
                                  c_ref.reshape(o.n_images, -1).sum(1).astype(np.float64))


# ------------------------------------------------------------ error paths

def test_error_paths(ctx):
    import torch
    from imager_amd import capi
    x = torch.zeros((1, 17, 17), dtype=torch.complex64, device="cuda")
    with pytest.raises(capi.BlinkUnsupported):
        ctx.fft2d_c2c(x, x, 17, 1)
    g = torch.zeros((1, 1, 8, 8), dtype=torch.complex64, device="cuda")
    c = torch.zeros((1, 1, 8, 8), dtype=torch.float32, device="cuda")
    o = synth.small(n_ant=4, n_ch=1, n_t=1, n_pixels=8)
    d = _obs_dev(o)
    with pytest.raises(capi.BlinkInvalidArgument):
        ctx.gridding(d["vis"], 4, 1, 1, d["u"], d["v"], d["flags"], d["freq"], 1.0, 1.0, 0.0, 8,
                     7, 0, g, c)
    with pytest.raises(capi.BlinkInvalidArgument):
        ctx.image(g, c, 8, 1, g)   # images aliasing grids
    o2 = synth.small(n_ant=8, n_ch=1, n_t=1, n_pixels=45)
    d2 = _obs_dev(o2)
    avg = torch.zeros((45, 45), dtype=torch.complex64, device="cuda")
    st = capi.make_set(d2["vis"], 8, 1, 1, d2["u"], d2["v"], d2["w"], d2["flags"], d2["freq"],
                       d2["cable"], 3, o2.delta_u, o2.delta_v, o2.min_uv, 45, "XX", None,
                       average=avg)
    with pytest.raises(capi.BlinkInvalidArgument):
        ctx.image_batch([st])      # an odd side's mean needs the images


# ------------------------------------------ split-precision DFT on MFMA

DFT_TRANSFORMS = {"bf16x3": "TRANSFORM_DFT_BF16X3", "fp16x3": "TRANSFORM_DFT_FP16X3"}


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("n,n_ant,flux", [(64, 32, 1.0), (180, 128, 1.0), (256, 128, 1.0),
                                          (512, 64, 1.0), (90, 24, 1.0), (128, 48, 1e6),
                                          (128, 48, 1e-6)])
def test_image_batch_dft(ctx, prec, n, n_ant, flux):
    """image = D G D^T on MFMA over the Hermitian half plane with split
    bf16 or fp16 hi+lo operands: within 1e-5 (peak-normalised) of the FP32
    oracle and of the FP64 oracle, imaginary parts exactly 0, count sums
    exact.  Fluxes scaled by 1e6 and 1e-6 take fp16 outside its exponent
    range without the per-row power-of-two scale.  The error to the FP64
    oracle is printed next to the FP32 FFT path's."""
    import torch
    from imager_amd import capi
    o = synth.small(n_ant=n_ant, n_ch=3, n_t=2, n_pixels=n, seed=900 + n, height_sigma_m=1.0)
    o.vis *= np.float32(flux)
    _say("dft", prec, n, n_ant, flux)
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, n)
    ref64 = oracle.run_f64(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                           o.delta_u, o.delta_v, o.min_uv, n)
    d = _obs_dev(o)
    imgs = torch.empty((o.n_t, o.n_ch, n, n), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, n, "XX", imgs,
                       sums, transform=getattr(capi, DFT_TRANSFORMS[prec]))
    ctx.image_batch([st])
    got = _host(imgs)
    e32 = oracle.peak_normalised_error(got, ref)
    e64 = oracle.peak_normalised_error(got, ref64)
    line = (f"  DFT {prec} N={n} flux {flux:g}: vs FP32 oracle {e32[0]:.2e}, "
            f"vs FP64 max {e64[0]:.2e} rms {e64[1]:.2e}")
    imgs2 = torch.empty_like(imgs)
    st2 = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                        d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, n, "XX",
                        imgs2)
    ctx.image_batch([st2])
    f64 = oracle.peak_normalised_error(_host(imgs2), ref64)
    line += f"; FP32 FFT vs FP64 max {f64[0]:.2e} rms {f64[1]:.2e}"
    _say(line)
    assert e32[0] < IMAGE_TOL
    assert e64[0] < IMAGE_TOL
    assert np.all(got.imag == 0)
    _, c_ref, _ = oracle.gridding(oracle.cable_corrections(
        oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq), o.n_ant, o.cable, o.freq),
        o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv, n, "XX")
    assert np.array_equal(_host(sums), c_ref.reshape(o.n_images, -1).sum(1).astype(np.float64))


def test_image_batch_dft_rejects_large_grids(ctx):
    import torch
    from imager_amd import capi
    o = synth.small(n_ant=8, n_ch=1, n_t=1, n_pixels=1024)
    d = _obs_dev(o)
    imgs = torch.empty((1, 1, 1024, 1024), dtype=torch.complex64, device="cuda")
    st = capi.make_set(d["vis"], 8, 1, 1, d["u"], d["v"], d["w"], d["flags"], d["freq"],
                       d["cable"], 0, o.delta_u, o.delta_v, o.min_uv, 1024, "XX", imgs,
                       transform=capi.TRANSFORM_DFT_BF16X3)
    with pytest.raises(capi.BlinkUnsupported):
        ctx.image_batch([st])
