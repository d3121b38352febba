"""Convolutional (Kaiser-Bessel) gridding on the batched path, against the oracle.

Each visibility and its conjugate are spread over K x K cells with the
1023-times oversampled kernels of makeKernels / makeKaiserBesselKernel
(src/pacer_imager.cpp:565-651), placed as convolve_with_kernel and
calc_xy_grid2 place them (:654-750); counts stay at the nearest cells.  The
batched call grids the footprints into LDS row tiles fused with the row FFT
and must give the oracle's images (oracle_gridding_kb_f32 + image) within
1e-5 peak-normalised, with exact count sums, for odd and even K, both kernel
modes, and any grid side.
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


def _reference(o, K, with_sinc):
    v = oracle.cable_corrections(oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq),
                                 o.n_ant, o.cable, o.freq)
    g, c, _ = oracle.gridding(v, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, o.n_pixels, "XX", kernel_size=K, with_sinc=with_sinc)
    return oracle.image(g, c), c


def _set(o, images, sums, K, mode, average=None):
    from imager_amd import capi
    d = dict(vis=_dev(o.vis), u=_dev(o.u), v=_dev(o.v), w=_dev(o.w), flags=_dev(o.flags),
             freq=_dev(o.freq), cable=_dev(o.cable))
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, o.n_pixels,
                       "XX", images, sums, kernel_size=K, kernel_mode=mode, average=average)
    return st, d


@pytest.mark.parametrize("n,n_ant,K,mode", [(128, 40, 7, 1), (180, 64, 7, 1), (256, 48, 10, 1),
                                            (256, 48, 5, 2), (1024, 128, 7, 1),
                                            (64, 32, 31, 1)])
def test_kb_batch_matches_oracle(ctx, n, n_ant, K, mode):
    import torch
    if n == 1024:
        o = synth.mwa_array(n_ch=3, n_t=2, n_pixels=1024, seed=1010)
    else:
        o = synth.small(n_ant=n_ant, n_ch=3, n_t=2, n_pixels=n, seed=n + K, height_sigma_m=1.0)
    ref, c = _reference(o, K, mode == 1)
    images = torch.full((o.n_t, o.n_ch, n, n), float("nan"), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    st, keep = _set(o, images, sums, K, mode)
    for rep in range(2):   # capture, then replay
        print(f"KB batch N={n} K={K} mode={mode} rep {rep}", flush=True)
        images.fill_(float("nan"))
        ctx.image_batch([st])
        got = _host(images)
        err = oracle.peak_normalised_error(got, ref)[0]
        print(f"  vs FP32 oracle {err:.2e}", flush=True)
        assert err < IMAGE_TOL
        assert np.array_equal(_host(sums), c.reshape(o.n_images, -1).sum(1).astype(np.float64))


def test_kb_batch_group_average_and_nn_beside(ctx):
    """One call: two KB sets of one side (one imaged for its average only) and
    an NN set of the same side, which takes the register path."""
    import torch
    a = synth.small(n_ant=40, n_ch=2, n_t=2, n_pixels=256, seed=31, height_sigma_m=1.0)
    b = synth.small(n_ant=56, n_ch=3, n_t=1, n_pixels=256, seed=32, height_sigma_m=1.0)
    nn = synth.small(n_ant=48, n_ch=2, n_t=1, n_pixels=256, seed=33, height_sigma_m=1.0)
    ra, ca = _reference(a, 7, True)
    rb, cb = _reference(b, 7, True)
    rn, _ = oracle.run_f32(nn.vis, nn.n_ant, nn.u, nn.v, nn.w, nn.flags, nn.freq, nn.cable, True,
                           True, nn.delta_u, nn.delta_v, nn.min_uv, 256)
    avg_a = torch.empty((256, 256), dtype=torch.complex64, device="cuda")
    sums_a = torch.empty(a.n_images, dtype=torch.float64, device="cuda")
    img_b = torch.empty((b.n_t, b.n_ch, 256, 256), dtype=torch.complex64, device="cuda")
    img_n = torch.empty((nn.n_t, nn.n_ch, 256, 256), dtype=torch.complex64, device="cuda")
    sta, ka = _set(a, None, sums_a, 7, 1, average=avg_a)
    stb, kb = _set(b, img_b, None, 7, 1)
    stn, kn = _set(nn, img_n, None, 0, 1)
    ctx.image_batch([sta, stb, stn])
    flat = np.ascontiguousarray(ra.reshape(-1, 256, 256))
    e = oracle.peak_normalised_error(_host(avg_a), oracle.image_averaging_complex(flat))[0]
    print(f"KB average {e:.2e}", flush=True)
    assert e < IMAGE_TOL
    assert np.array_equal(_host(sums_a), ca.reshape(a.n_images, -1).sum(1).astype(np.float64))
    assert oracle.peak_normalised_error(_host(img_b), rb)[0] < IMAGE_TOL
    assert oracle.peak_normalised_error(_host(img_n), rn)[0] < IMAGE_TOL


def test_kb_batch_refusals(ctx):
    import torch
    from imager_amd import capi
    o = synth.small(n_ant=8, n_ch=1, n_t=1, n_pixels=64)
    img = torch.empty((1, 1, 64, 64), dtype=torch.complex64, device="cuda")
    d = [_dev(x) for x in (o.vis, o.u, o.v, o.w, o.flags, o.freq, o.cable)]
    mk = lambda **kw: capi.make_set(d[0], 8, 1, 1, d[1], d[2], d[3], d[4], d[5], d[6], 3,
                                    o.delta_u, o.delta_v, o.min_uv, 64, "XX", img, **kw)
    with pytest.raises(capi.BlinkInvalidArgument):
        ctx.image_batch([mk(kernel_size=32)])
    with pytest.raises(capi.BlinkUnsupported):
        ctx.image_batch([mk(kernel_size=7, weighting="U")])
    with pytest.raises(capi.BlinkUnsupported):
        ctx.image_batch([mk(kernel_size=7, transform=capi.TRANSFORM_DFT_BF16X3)])


class _env:
    """Sets environment variables for the calls inside the block."""

    def __init__(self, **kv):
        self.kv = kv
        self.old = {}

    def __enter__(self):
        import os
        for k, v in self.kv.items():
            self.old[k] = os.environ.get(k)
            os.environ[k] = str(v)

    def __exit__(self, *exc):
        import os
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("K,mode,freq", [(7, 1, (151.25e6, 159.375e6, 171.0e6)),
                                         (5, 2, (169.6e6, 185.0e6, 200.3e6)),
                                         (7, 1, (240e6, 300e6, 160e6))])
def test_kb_direct_window_256(ctx, K, mode, freq):
    """The 256-antenna station at 256^2 (dense, one timestep per channel) on
    the windowed direct row pass (BLINK_DIRECT=1: three images), the window
    widened by K/2 for the footprints; the 240 and 300 MHz channels need more
    than the LDS holds and take blocks of whole rows.  Images against the
    oracle's convolutional gridding within 1e-5, exact count sums, and the
    plan path's images (BLINK_DIRECT=0) within the same tolerance."""
    import torch
    o = synth.eda2_station(n_ch=3, n_t=1, n_pixels=256, seed=700 + K, freq=np.array(freq),
                           height_sigma_m=1.0, bl_flag_fraction=0.1)
    ref, c = _reference(o, K, mode == 1)
    images = torch.full((1, 3, 256, 256), float("nan"), dtype=torch.complex64, device="cuda")
    sums = torch.empty(3, dtype=torch.float64, device="cuda")
    st, keep = _set(o, images, sums, K, mode)
    out = {}
    for direct in ("1", "0"):
        with _env(BLINK_DIRECT=direct):
            for rep in range(2):
                ctx.profile_enable(rep == 1)
                ctx.profile_reset()
                images.fill_(float("nan"))
                print(f"KB direct={direct} 256^2 K={K} mode={mode} rep {rep}", flush=True)
                ctx.image_batch([st])
                got = _host(images)
                err = oracle.peak_normalised_error(got, ref)[0]
                print(f"  vs FP32 oracle {err:.2e}", flush=True)
                assert err < IMAGE_TOL
                assert np.array_equal(_host(sums), c.reshape(3, -1).sum(1).astype(np.float64))
                if rep == 1:
                    assert ("direct_prep" in ctx.profile()) == (direct == "1")
        out[direct] = _host(images).copy()
    ctx.profile_enable(False)
    assert oracle.peak_normalised_error(out["1"], out["0"])[0] < IMAGE_TOL


def test_kb_direct_chosen_for_many_images(ctx):
    """Without BLINK_DIRECT, 160 channels of the station at 256^2 with K = 7
    take the direct row pass by themselves."""
    import torch
    o = synth.eda2_station(n_ch=160, n_t=1, n_pixels=256, seed=720,
                           freq=synth.mwa_fine_channels(160), height_sigma_m=1.0)
    ref, c = _reference(o, 7, True)
    images = torch.empty((1, 160, 256, 256), dtype=torch.complex64, device="cuda")
    sums = torch.empty(160, dtype=torch.float64, device="cuda")
    st, keep = _set(o, images, sums, 7, 1)
    ctx.profile_enable(True)
    ctx.profile_reset()
    try:
        print("KB auto choice, 160 images of 256^2", flush=True)
        ctx.image_batch([st])
        assert "direct_prep" in ctx.profile()
    finally:
        ctx.profile_enable(False)
    assert oracle.peak_normalised_error(_host(images), ref)[0] < IMAGE_TOL
    assert np.array_equal(_host(sums), c.reshape(160, -1).sum(1).astype(np.float64))
