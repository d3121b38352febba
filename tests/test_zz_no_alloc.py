"""blink_image_batch_noalloc (the batched call for callers that cannot allocate
device memory) against blink_image_batch, through the C ABI.

Two calls in a row, the second with a smaller input, so that the second finds
the fixed region as the first left it: images within 1e-5 peak-normalised of
blink_image_batch's (and of the oracle's) and count sums equal; then the
refusals (too large for the region, a side without a prepared FFT table, a DFT
transform) with no kernel launched.
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


def _both(ctx, o, transform=0, average=False):
    """Images (and average) of o from blink_image_batch and from the
    no-allocation entry point."""
    import torch
    from imager_amd import capi
    d = dict(vis=_dev(o.vis), u=_dev(o.u), v=_dev(o.v), w=_dev(o.w), flags=_dev(o.flags),
             freq=_dev(o.freq), cable=_dev(o.cable))
    out = []
    for fn in (ctx.image_batch, ctx.image_batch_noalloc):
        img = torch.full((o.n_t, o.n_ch, o.n_pixels, o.n_pixels), float("nan"),
                         dtype=torch.complex64, device="cuda")
        sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
        avg = (torch.full((o.n_pixels, o.n_pixels), float("nan"), dtype=torch.complex64,
                          device="cuda") if average else None)
        st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                           d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv,
                           o.n_pixels, "XX", img, sums, transform=transform, average=avg)
        print(f"{fn.__name__}: {o.n_ant} antennas, {o.n_images} images of {o.n_pixels}^2",
              flush=True)
        fn([st])
        out.append((_host(img), _host(sums), None if avg is None else _host(avg)))
    return out


def test_noalloc_matches_image_batch_on_two_calls(ctx):
    """Call 1: the 128-tile array at 1024^2, 12 images (several row-pass chunks
    in the region), with its average; call 2, smaller: the 256-antenna station
    at 180^2, 3 images.  Each against blink_image_batch and the oracle."""
    first = synth.mwa_array(n_ch=6, n_t=2, n_pixels=1024, seed=1002)
    second = synth.eda2_station(n_ch=3, n_t=1, n_pixels=180, seed=1004,
                                freq=synth.mwa_fine_channels(3))
    for o, average in ((first, True), (second, False)):
        (ref_img, ref_sums, ref_avg), (img, sums, avg) = _both(ctx, o, average=average)
        assert np.array_equal(sums, ref_sums)
        err = oracle.peak_normalised_error(img, ref_img)[0]
        print(f"  vs blink_image_batch {err:.2e}", flush=True)
        assert err < IMAGE_TOL
        r32, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True,
                                True, o.delta_u, o.delta_v, o.min_uv, o.n_pixels)
        assert oracle.peak_normalised_error(img, r32)[0] < IMAGE_TOL
        if average:
            assert oracle.peak_normalised_error(avg, ref_avg)[0] < IMAGE_TOL


@pytest.mark.parametrize("n", [128, 180, 256])
def test_noalloc_direct_row_pass(ctx, n):
    """The 256-antenna station's dense single-timestep sets on the direct row
    pass (BLINK_DIRECT=1: three images), whose set-up and row starts live in
    the fixed region (at 256^2 with the window's |u|, |v| maxima): the
    no-allocation call against blink_image_batch and the oracle, twice."""
    import os
    o = synth.eda2_station(n_ch=3, n_t=1, n_pixels=n, seed=1010 + n,
                           freq=synth.mwa_fine_channels(3))
    old = os.environ.get("BLINK_DIRECT")
    os.environ["BLINK_DIRECT"] = "1"
    try:
        for rep in range(2):
            (ref_img, ref_sums, _), (img, sums, _) = _both(ctx, o)
            assert np.array_equal(sums, ref_sums)
            assert oracle.peak_normalised_error(img, ref_img)[0] < IMAGE_TOL
    finally:
        if old is None:
            os.environ.pop("BLINK_DIRECT", None)
        else:
            os.environ["BLINK_DIRECT"] = old
    r32, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, n)
    assert oracle.peak_normalised_error(img, r32)[0] < IMAGE_TOL


def test_noalloc_refusals(ctx):
    import torch
    from imager_amd import capi
    # too large: 4096 channels of the 128-tile array need 270 MB of plan
    # entries alone (the visibilities are never read: the call is refused first)
    n_ant, n_ch = 128, 4096
    n_bl = n_ant * (n_ant + 1) // 2
    vis = torch.empty((1, n_ch, n_bl, 4), dtype=torch.complex64, device="cuda")
    o = synth.small(n_ant=n_ant, n_ch=1, n_t=1, n_pixels=1024)
    u, v, w, flags, cable = (_dev(o.u), _dev(o.v), _dev(o.w), _dev(o.flags), _dev(o.cable))
    freq = _dev(np.linspace(150e6, 200e6, n_ch))
    img = torch.empty((1, 1, 8, 8), dtype=torch.complex64, device="cuda")
    st = capi.make_set(vis, n_ant, n_ch, 1, u, v, w, flags, freq, cable, 3, o.delta_u, o.delta_v,
                       o.min_uv, 1024, "XX", img)
    with pytest.raises(capi.BlinkInvalidArgument):
        ctx.image_batch_noalloc([st])
    # a side without a prepared FFT table, and a DFT transform
    o = synth.small(n_ant=16, n_ch=1, n_t=1, n_pixels=90)
    d = [_dev(a) for a in (o.vis, o.u, o.v, o.w, o.flags, o.freq, o.cable)]
    img = torch.empty((1, 1, 90, 90), dtype=torch.complex64, device="cuda")
    for n, tr in ((90, 0), (128, capi.TRANSFORM_DFT_BF16X3)):
        st = capi.make_set(d[0], 16, 1, 1, d[1], d[2], d[3], d[4], d[5], d[6], 3, o.delta_u,
                           o.delta_v, o.min_uv, n, "XX", img, transform=tr)
        with pytest.raises(capi.BlinkUnsupported):
            ctx.image_batch_noalloc([st])
