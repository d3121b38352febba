"""Batched-path parity on the benchmark workloads' own arrays and shapes.

Each test images a slice of a benchmark configuration through
blink_image_batch and checks every image against the oracle's reference-
precision CPU pipeline (peak-normalised error below 1e-5) and every count sum
exactly:

  * the uneven batch: single-image sets drawn from synth.mixed_batch (both
    arrays, every grid side 128 .. 2048) in ONE call;
  * the large single-GPU load: 128 tiles at 4096^2 with 50 sources plus a
    diffuse field, two calls into one shared output buffer, with the row-pass
    budget cut so that every call runs several chunks (also 1024^2 and the
    generic 180^2 side);
  * the 256-antenna station at 180^2 and 256^2 over fine channels, with the
    FP32 FFT and both split-precision DFT transforms (those also within 1e-5
    of the FP64 oracle).
"""
import os

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


def _count_sums(o, vis=None, pol="XX"):
    v = o.vis if vis is None else vis
    vc = oracle.cable_corrections(oracle.geometric_corrections(v, o.n_ant, o.w, o.freq),
                                  o.n_ant, o.cable, o.freq)
    _, c, _ = oracle.gridding(vc, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, o.n_pixels, pol)
    return c.reshape(-1, o.n_pixels * o.n_pixels).sum(1).astype(np.float64)


def _pick_mixed(wl, per_key):
    """Indices of up to per_key images of every (grid side, array) pair."""
    picked = {}
    for i, o in enumerate(wl):
        lst = picked.setdefault((o.n_pixels, o.n_ant), [])
        if len(lst) < per_key:
            lst.append(i)
    return sorted(i for v in picked.values() for i in v)


@pytest.mark.parametrize("eager", [False, True])
def test_mixed_batch_in_one_call(ctx, eager):
    """The uneven batch: every grid side 128 .. 2048 from both arrays, each
    image with its own random baseline flags, imaged by one call (graph
    capture then replay, or the eager path with the compaction beside the
    plan on the second stream)."""
    import torch
    from imager_amd import capi
    wl = synth.mixed_batch(n_images=512, seed=1003)
    sel = _pick_mixed(wl, 2 if eager else 3)
    draw = [wl[i] for i in sel]
    sides = {o.n_pixels for o in draw}
    assert sides == {128, 180, 256, 512, 1024, 2048}
    assert {o.n_ant for o in draw} == {128, 256}
    assert 24 <= len(draw) <= 48
    corr = capi.CORR_GEOMETRIC | capi.CORR_CABLE
    shared, keep, sets = {}, [], []
    for o in draw:
        if id(o.u) not in shared:
            shared[id(o.u)] = dict(u=_dev(o.u), v=_dev(o.v), w=_dev(o.w), cable=_dev(o.cable))
        d = dict(shared[id(o.u)], vis=_dev(o.vis), flags=_dev(o.flags), freq=_dev(o.freq))
        img = torch.empty((1, 1, o.n_pixels, o.n_pixels), dtype=torch.complex64, device="cuda")
        sums = torch.empty(1, dtype=torch.float64, device="cuda")
        sets.append(capi.make_set(d["vis"], o.n_ant, 1, 1, d["u"], d["v"], d["w"], d["flags"],
                                  d["freq"], d["cable"], corr, o.delta_u, o.delta_v, o.min_uv,
                                  o.n_pixels, "XX", img, sums))
        keep.append((d, img, sums))
    refs = []
    for o in draw:
        ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True,
                                True, o.delta_u, o.delta_v, o.min_uv, o.n_pixels)
        refs.append((ref[0, 0], _count_sums(o)))
    ctx.profile_enable(eager)
    try:
        for rep in range(1 if eager else 2):     # capture, then replay
            for _, img, _ in keep:
                img.fill_(float("nan"))
            _say(f"mixed draw of {len(draw)} images, sides {sorted(sides)}, eager={eager}, "
                 f"rep {rep}")
            ctx.image_batch(sets)
            worst = 0.0
            for o, (_, img, sums), (ref, csum) in zip(draw, keep, refs):
                err = oracle.peak_normalised_error(_host(img)[0, 0], ref)[0]
                worst = max(worst, err)
                assert err < IMAGE_TOL, (o.name, o.n_pixels, o.n_ant, err)
                assert np.array_equal(_host(sums), csum), o.name
            _say(f"  worst error vs FP32 oracle {worst:.2e}")
    finally:
        ctx.profile_enable(False)


@pytest.mark.parametrize("n,inter_mib", [(4096, 64), (1024, 9), (180, 1)])
def test_streamed_calls_run_several_chunks(ctx, n, inter_mib):
    """Two calls (timesteps 0 and 1) into one shared output buffer, as the
    4096^2 workload streams its timesteps, with the row-pass budget cut to
    inter_mib MiB so that every call runs at least three chunks (launches
    with a first image index above 0).  4096^2: 128 tiles, 50 sources plus a
    diffuse Gaussian random field; 1024^2: the 128-tile array; 180^2: the
    256-antenna station on the generic kernels."""
    import torch
    from imager_amd import capi
    if n == 4096:
        o = synth.mwa_array(n_ch=4, n_t=2, n_pixels=4096, seed=1005, n_src=50,
                            freq=synth.mwa_fine_channels(768)[::191][:4], diffuse=2.0)
        per_image = 8 * (n // 2 + 1) * n
    elif n == 1024:
        o = synth.mwa_array(n_ch=6, n_t=2, n_pixels=1024, seed=1002)
        per_image = 8 * (n // 2 + 1) * n
    else:
        o = synth.eda2_station(n_ch=24, n_t=2, n_pixels=180, seed=1004,
                               freq=synth.mwa_fine_channels(24))
        per_image = 8 * (n // 2 + 1) * n
    chunk = max(1, (inter_mib << 20) // per_image)
    assert -(-o.n_ch // chunk) >= 3, "the budget must split each call into 3+ chunks"
    d = dict(u=_dev(o.u), v=_dev(o.v), w=_dev(o.w), flags=_dev(o.flags), freq=_dev(o.freq),
             cable=_dev(o.cable), vis=_dev(o.vis))
    images = torch.empty((1, o.n_ch, n, n), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_ch, dtype=torch.float64, device="cuda")
    corr = capi.CORR_GEOMETRIC | capi.CORR_CABLE
    calls = [capi.make_set(d["vis"][t:t + 1], o.n_ant, o.n_ch, 1, d["u"], d["v"], d["w"],
                           d["flags"], d["freq"], d["cable"], corr, o.delta_u, o.delta_v,
                           o.min_uv, n, "XX", images, sums) for t in range(o.n_t)]
    old = os.environ.get("BLINK_INTER_MIB")
    os.environ["BLINK_INTER_MIB"] = str(inter_mib)
    try:
        for t in range(o.n_t):
            images.fill_(float("nan"))
            _say(f"streamed N={n} t={t}: {o.n_ch} images in chunks of {chunk}")
            ctx.image_batch([calls[t]])
            got = _host(images)[0]
            ref, _ = oracle.run_f32(o.vis[t:t + 1], o.n_ant, o.u, o.v, o.w, o.flags, o.freq,
                                    o.cable, True, True, o.delta_u, o.delta_v, o.min_uv, n)
            errs = [oracle.peak_normalised_error(got[c], ref[0, c])[0] for c in range(o.n_ch)]
            _say(f"  worst error vs FP32 oracle {max(errs):.2e}")
            assert max(errs) < IMAGE_TOL, errs
            assert np.array_equal(_host(sums), _count_sums(o, o.vis[t:t + 1]))
        if n == 4096:
            ref64 = oracle.run_f64(o.vis[1:2, :1], o.n_ant, o.u, o.v, o.w, o.flags,
                                   o.freq[:1], o.cable, True, True, o.delta_u, o.delta_v,
                                   o.min_uv, n)
            e64 = oracle.peak_normalised_error(_host(images)[0, 0], ref64[0, 0])
            _say(f"  t=1 ch=0 vs FP64 oracle {e64[0]:.2e} (rms {e64[1]:.2e})")
            assert e64[0] < IMAGE_TOL
    finally:
        if old is None:
            os.environ.pop("BLINK_INTER_MIB", None)
        else:
            os.environ["BLINK_INTER_MIB"] = old


@pytest.mark.parametrize("transform", ["fft", "dft-bf16", "dft-fp16"])
@pytest.mark.parametrize("n", [180, 256])
def test_station_fine_channels(ctx, n, transform):
    """The 256-antenna station (32 896 baselines, a dense grid) over 6 fine
    channels from 169.6 MHz at its all-sky sides, through the FP32 FFT and
    the split-precision DFT transforms: within 1e-5 of the FP32 and of the
    FP64 oracle, count sums exact."""
    import torch
    from imager_amd import capi
    o = synth.eda2_station(n_ch=6, n_t=1, n_pixels=n, seed=1004,
                           freq=synth.mwa_fine_channels(6))
    tr = {"fft": capi.TRANSFORM_FFT, "dft-bf16": capi.TRANSFORM_DFT_BF16X3,
          "dft-fp16": capi.TRANSFORM_DFT_FP16X3}[transform]
    d = dict(u=_dev(o.u), v=_dev(o.v), w=_dev(o.w), flags=_dev(o.flags), freq=_dev(o.freq),
             cable=_dev(o.cable), vis=_dev(o.vis))
    images = torch.empty((1, o.n_ch, n, n), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_ch, dtype=torch.float64, device="cuda")
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, 1, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], 3, o.delta_u, o.delta_v, o.min_uv, n, "XX",
                       images, sums, transform=tr)
    _say(f"station {o.n_ant} ant N={n} {transform}")
    ctx.image_batch([st])
    got = _host(images)
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, n)
    ref64 = oracle.run_f64(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                           o.delta_u, o.delta_v, o.min_uv, n)
    e32 = oracle.peak_normalised_error(got, ref)
    e64 = oracle.peak_normalised_error(got, ref64)
    _say(f"  vs FP32 oracle {e32[0]:.2e}, vs FP64 max {e64[0]:.2e} rms {e64[1]:.2e}")
    assert e32[0] < IMAGE_TOL and e64[0] < IMAGE_TOL
    assert np.array_equal(_host(sums), _count_sums(o))
