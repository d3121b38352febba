"""Grid sides beyond the powers of two and 2^a 3^b 5^c up to 4096: sides with
the prime factors 7, 11 and 13 (generic FFT passes of those radices), and
sides above 4096 up to 8192 (the reference's FFTW / hipFFT plans take any N,
src/pacer_imager.cpp:150-153, src/gpu/pacer_imager_hip.cpp:88; its
experiments image 8192^2, experiments/compare_blink_images.sh:8).

Each image of the batched call and of the stage entry points (blink_gridding
+ blink_image) against the oracle's reference-precision pipeline (peak-
normalised error below 1e-5) and every count exactly; the refusals that stay.
"""
import numpy as np
import pytest

import oracle
from imager_amd import synth

pytestmark = pytest.mark.gpu

IMAGE_TOL = 1e-5


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _say(*a):
    print(*a, flush=True)


def _obs(n, n_ch=2, n_t=1):
    if n >= 4096:
        # the 128-tile array (1.5 km) at the large sides
        return synth.mwa_array(n_ch=n_ch, n_t=n_t, n_pixels=n, seed=n, height_sigma_m=1.0)
    return synth.small(n_ant=48, n_ch=n_ch, n_t=n_t, n_pixels=n, seed=n, height_sigma_m=1.0)


def _reference(o):
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, o.n_pixels)
    vc = oracle.cable_corrections(oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq),
                                  o.n_ant, o.cable, o.freq)
    _, c, _ = oracle.gridding(vc, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, o.n_pixels, "XX")
    return ref, c


# 154 = 2 7 11, 208 = 2^4 13, 2002 = 2 7 11 13, 6000 = 2^4 3 5^3, 8192 = 2^13
@pytest.mark.parametrize("n,n_ch", [(154, 2), (208, 2), (2002, 2), (6000, 1), (8192, 1)])
def test_batch_other_sides(ctx, n, n_ch):
    import torch
    from imager_amd import capi
    o = _obs(n, n_ch=n_ch)
    ref, c = _reference(o)
    d = {k: _dev(getattr(o, k)) for k in ("vis", "u", "v", "w", "flags", "freq", "cable")}
    imgs = torch.full((o.n_t, o.n_ch, n, n), float("nan"), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], capi.CORR_GEOMETRIC | capi.CORR_CABLE, o.delta_u,
                       o.delta_v, o.min_uv, n, "XX", imgs, sums)
    for rep in range(2):   # capture, then replay
        _say(f"batch N={n}, {o.n_images} images, rep {rep}")
        ctx.image_batch([st])
        err = oracle.peak_normalised_error(_host(imgs), ref)[0]
        _say(f"  vs FP32 oracle {err:.2e}")
        assert err < IMAGE_TOL
        assert np.array_equal(_host(sums), c.reshape(o.n_images, -1).sum(1).astype(np.float64))


@pytest.mark.parametrize("n", [2002, 8192])
def test_stage_path_other_sides(ctx, n):
    """blink_apply_*_corrections + blink_gridding (tiled plan) + blink_image at
    a side with the factors 7, 11, 13 and at 8192."""
    import torch
    o = _obs(n, n_ch=1)
    ref, c = _reference(o)
    d = {k: _dev(getattr(o, k)) for k in ("vis", "u", "v", "w", "flags", "freq", "cable")}
    grids = torch.zeros((1, 1, n, n), dtype=torch.complex64, device="cuda")
    counts = torch.zeros((1, 1, n, n), dtype=torch.float32, device="cuda")
    images = torch.empty((1, 1, n, n), dtype=torch.complex64, device="cuda")
    _say(f"stage path N={n}")
    ctx.apply_geometric_corrections(d["vis"], o.n_ant, 1, 1, d["w"], d["freq"])
    ctx.apply_cable_corrections(d["vis"], o.n_ant, 1, 1, d["cable"], d["freq"])
    ctx.gridding(d["vis"], o.n_ant, 1, 1, d["u"], d["v"], d["flags"], d["freq"], o.delta_u,
                 o.delta_v, o.min_uv, n, "XX", 0, grids, counts)
    assert np.array_equal(_host(counts), c.reshape(1, 1, n, n))
    ctx.image(grids, counts, n, 1, images)
    err = oracle.peak_normalised_error(_host(images), ref)[0]
    _say(f"  vs FP32 oracle {err:.2e}")
    assert err < IMAGE_TOL


def test_side_refusals(ctx):
    """Still refused: a prime factor above 13 (17), sides above 8192, and
    uniform weighting on the batched path above 4096."""
    import torch
    from imager_amd import capi
    x = torch.zeros((1, 34, 34), dtype=torch.complex64, device="cuda")
    with pytest.raises(capi.BlinkUnsupported):
        ctx.fft2d_c2c(x, x, 34, 1)
    assert not capi.fft_supported(8194) and not capi.fft_supported(16384)
    o = synth.small(n_ant=8, n_ch=1, n_t=1, n_pixels=64)
    d = {k: _dev(getattr(o, k)) for k in ("vis", "u", "v", "w", "flags", "freq", "cable")}
    img = torch.empty((1, 1, 8, 8), dtype=torch.complex64, device="cuda")
    for n, weighting in ((34, "N"), (8192, "U")):
        st = capi.make_set(d["vis"], 8, 1, 1, d["u"], d["v"], d["w"], d["flags"], d["freq"],
                           d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, n, "XX", img,
                           weighting=weighting)
        with pytest.raises(capi.BlinkUnsupported):
            ctx.image_batch([st])


# odd sides on the batched path: 45 = 3^2 5, 77 = 7 11, 225 = 3^2 5^2, 1001 = 7 11 13
@pytest.mark.parametrize("n,pol", [(45, "XX"), (77, "I"), (225, "XX"), (1001, "YY")])
def test_batch_odd_sides(ctx, n, pol):
    """Odd N (the reference's CPU path images them, with its swap fftshift for
    odd sizes, src/pacer_imager.cpp:30-35,110-128): images, count sums and the
    mean (taken from the formed images) against the oracle, on capture and
    replay."""
    import torch
    from imager_amd import capi
    o = synth.small(n_ant=40, n_ch=2, n_t=2, n_pixels=n, seed=7 * n, height_sigma_m=1.0)
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, n, pol)
    vc = oracle.cable_corrections(oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq),
                                  o.n_ant, o.cable, o.freq)
    _, c, _ = oracle.gridding(vc, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, n, pol)
    d = {k: _dev(getattr(o, k)) for k in ("vis", "u", "v", "w", "flags", "freq", "cable")}
    imgs = torch.full((o.n_t, o.n_ch, n, n), float("nan"), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    avg = torch.full((n, n), float("nan"), dtype=torch.complex64, device="cuda")
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], capi.CORR_GEOMETRIC | capi.CORR_CABLE, o.delta_u,
                       o.delta_v, o.min_uv, n, pol, imgs, sums, average=avg)
    flat = np.ascontiguousarray(ref.reshape(-1, n, n))
    for rep in range(2):
        _say(f"odd side N={n} pol={pol}, rep {rep}")
        ctx.image_batch([st])
        err = oracle.peak_normalised_error(_host(imgs), ref)[0]
        _say(f"  vs FP32 oracle {err:.2e}")
        assert err < IMAGE_TOL
        assert np.array_equal(_host(sums), c.reshape(o.n_images, -1).sum(1).astype(np.float64))
        assert oracle.peak_normalised_error(_host(avg), oracle.image_averaging_complex(flat))[0] < IMAGE_TOL


def test_batch_odd_side_uniform_and_refusals(ctx):
    """Uniform weighting at an odd side (a grid whose every cell is covered, so
    the image is finite), and the odd-side refusals: the DFT transforms,
    convolutional gridding, an average without images."""
    import torch
    from imager_amd import capi
    from test_gpu_weighting import covering_array, _reference
    o = covering_array(45, n_ch=2, n_t=1, seed=9)
    ref, _ = _reference(o, uniform=True)
    assert np.isfinite(ref).all()
    d = {k: _dev(getattr(o, k)) for k in ("vis", "u", "v", "w", "flags", "freq", "cable")}
    imgs = torch.empty((o.n_t, o.n_ch, 45, 45), dtype=torch.complex64, device="cuda")
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, 45, "XX", imgs,
                       weighting="U")
    _say("odd side 45, uniform weighting")
    ctx.image_batch([st])
    err = oracle.peak_normalised_error(_host(imgs), ref)[0]
    _say(f"  vs FP32 oracle {err:.2e}")
    assert err < IMAGE_TOL
    avg = torch.empty((45, 45), dtype=torch.complex64, device="cuda")
    for kw, exc in ((dict(transform=capi.TRANSFORM_DFT_BF16X3), capi.BlinkUnsupported),
                    (dict(kernel_size=7), capi.BlinkUnsupported),
                    (dict(average=avg), capi.BlinkInvalidArgument)):
        bad = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                            d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, 45, "XX",
                            None if "average" in kw else imgs, **kw)
        with pytest.raises(exc):
            ctx.image_batch([bad])
