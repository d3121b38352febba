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
