"""The direct row pass: dense sets with one timestep per channel (every image
has a plan of its own) at sides whose half plane fits one workgroup's LDS
(128 on the register kernels, generic sides up to 186, odd ones included)
are gridded without a plan, one workgroup per image walking all baselines of
its image and computing their cells itself (k_grid_rows_direct,
k_grid_rows_direct_fft).  At 256^2 the workgroup holds the window of the
half plane that the image's cells fall in (k_grid_rows_direct_win), or,
when that window does not fit or a channel has several timesteps, blocks of
whole rows that each walk all baselines.

Against the oracle's reference-precision pipeline (peak-normalised error below
1e-5) and the oracle's counts (exact): the 256-antenna station at 128^2,
180^2, 175^2 and 90^2, the four polarisations, flagged baselines, NaN and
|V| >= 1e12 visibilities, a uv-distance cut, a group of two sets (one
averaged), on capture, replay and eagerly (where the launches show the direct
path); the same calls on the plan path (BLINK_DIRECT=0) for identical count
sums; and a sparse set with two timesteps forced onto the direct path.
"""
import os

import numpy as np
import pytest

import oracle
from imager_amd import synth

pytestmark = pytest.mark.gpu

IMAGE_TOL = 1e-5
FREQ = np.array([151.25e6, 159.375e6, 171.0e6])


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _say(*a):
    print(*a, flush=True)


class _env:
    """Sets environment variables for the calls inside the block."""

    def __init__(self, **kv):
        self.kv = kv
        self.old = {}

    def __enter__(self):
        for k, v in self.kv.items():
            self.old[k] = os.environ.get(k)
            os.environ[k] = str(v)

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _station(n, seed, **kw):
    o = synth.eda2_station(n_ch=3, n_t=1, n_pixels=n, seed=seed, freq=FREQ,
                           height_sigma_m=1.0, bl_flag_fraction=0.1, **kw)
    good = np.flatnonzero(o.flags == 0)
    o.vis[0, 1, good[5], 0] = np.nan
    o.vis[0, 2, good[9], 0] = np.complex64(3e12)
    o.vis[0, 0, good[11], 3] = np.complex64(complex(0.5, -2e12))
    return o


def _reference(o, pol):
    ref, _ = oracle.run_f32(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                            o.delta_u, o.delta_v, o.min_uv, o.n_pixels, pol)
    vc = oracle.cable_corrections(oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq),
                                  o.n_ant, o.cable, o.freq)
    _, c, _ = oracle.gridding(vc, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, o.n_pixels, pol)
    return ref, c.reshape(o.n_images, -1).sum(1).astype(np.float64)


def _set(o, pol, images, sums, average=None):
    from imager_amd import capi
    d = {k: _dev(getattr(o, k)) for k in ("vis", "u", "v", "w", "flags", "freq", "cable")}
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], capi.CORR_GEOMETRIC | capi.CORR_CABLE, o.delta_u,
                       o.delta_v, o.min_uv, o.n_pixels, pol, images, sums, average=average)
    return st, d


def _buffers(o):
    import torch
    n = o.n_pixels
    imgs = torch.empty((o.n_t, o.n_ch, n, n), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    return imgs, sums


@pytest.mark.parametrize("n,pol", [(128, "I"), (128, "YY"), (180, "XX"), (175, "V"),
                                   (90, "XX"), (256, "XX"), (256, "V"), (256, "I")])
def test_direct_sides(ctx, n, pol):
    o = _station(n, seed=500 + n)
    assert o.n_bl > 64 * (n // 2 + 1), "a dense grid"
    ref, c = _reference(o, pol)
    imgs, sums = _buffers(o)
    st, keep = _set(o, pol, imgs, sums)
    for rep in range(3):   # capture, replay, then eagerly with the launches counted
        eager = rep == 2
        ctx.profile_enable(eager)
        ctx.profile_reset()
        imgs.fill_(float("nan"))
        _say(f"direct N={n} pol={pol} rep {rep}")
        with _env(BLINK_DIRECT=1):   # 3 images: forced
            ctx.image_batch([st])
        got = _host(imgs)
        err = oracle.peak_normalised_error(got, ref)[0]
        _say(f"  vs FP32 oracle {err:.2e}")
        assert err < IMAGE_TOL
        assert not np.any(got.imag)
        assert np.array_equal(_host(sums), c)
        if eager:
            launches = ctx.profile()
            assert launches["direct_prep"][0] == 1, launches
            assert launches["grid_rows_fft"][0] == 1, launches
            assert "plan" not in launches, launches
    ctx.profile_enable(False)


@pytest.mark.parametrize("n,transform", [(180, "TRANSFORM_DFT_BF16X3"),
                                         (128, "TRANSFORM_DFT_FP16X3")])
def test_direct_dft_half_plane(ctx, n, transform):
    """The split-precision DFT transforms take their half plane from the
    direct walk too (k_grid_half_direct): images within 1e-5 of the FP32 and
    FP64 oracles, counts exact, and the plan path's images within the same
    tolerance."""
    import torch
    from imager_amd import capi
    o = _station(n, seed=900 + n)
    ref, c = _reference(o, "XX")
    ref64 = oracle.run_f64(o.vis, o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, True, True,
                           o.delta_u, o.delta_v, o.min_uv, n, "XX")
    d = {k: _dev(getattr(o, k)) for k in ("vis", "u", "v", "w", "flags", "freq", "cable")}
    imgs, sums = _buffers(o)
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], capi.CORR_GEOMETRIC | capi.CORR_CABLE, o.delta_u,
                       o.delta_v, o.min_uv, n, "XX", imgs, sums,
                       transform=getattr(capi, transform))
    out = {}
    for mode in ("1", "0"):
        with _env(BLINK_DIRECT=mode):
            for rep in range(3):
                eager = rep == 2
                ctx.profile_enable(eager)
                ctx.profile_reset()
                imgs.fill_(float("nan"))
                _say(f"DFT {transform} N={n}, BLINK_DIRECT={mode}, rep {rep}")
                ctx.image_batch([st])
                got = _host(imgs)
                e32 = oracle.peak_normalised_error(got, ref)[0]
                e64 = oracle.peak_normalised_error(got, ref64)[0]
                _say(f"  vs FP32 oracle {e32:.2e}, vs FP64 {e64:.2e}")
                assert e32 < IMAGE_TOL and e64 < IMAGE_TOL
                assert not np.any(got.imag)
                assert np.array_equal(_host(sums), c)
                if eager:
                    launches = ctx.profile()
                    assert ("direct_prep" in launches) == (mode == "1"), launches
                    assert ("plan" in launches) == (mode == "0"), launches
        out[mode] = _host(imgs).copy()
    ctx.profile_enable(False)
    # the gridded planes differ only in FP32 summation order, which the split
    # into 16-bit hi + lo operands can carry up to the transform's own error
    assert oracle.peak_normalised_error(out["1"], out["0"])[0] < IMAGE_TOL


@pytest.mark.parametrize("n", [128, 180, 256])
def test_direct_chosen_for_many_images(ctx, n):
    """Without BLINK_DIRECT a dense single-timestep set with images for at
    least half the CUs takes the direct path by itself, and a call of three
    such images does not (one workgroup per image would leave the GPU idle)."""
    import torch
    o = synth.eda2_station(n_ch=160, n_t=1, n_pixels=n, seed=970 + n,
                           freq=synth.mwa_fine_channels(160), height_sigma_m=1.0)
    ref, c = _reference(o, "XX")
    imgs, sums = _buffers(o)
    st, keep = _set(o, "XX", imgs, sums)
    small = _station(n, seed=980 + n)
    imgs2, sums2 = _buffers(small)
    st2, keep2 = _set(small, "XX", imgs2, sums2)
    ctx.profile_enable(True)
    try:
        for call, many in (([st], True), ([st2], False)):
            ctx.profile_reset()
            _say(f"auto choice N={n}, {'160' if many else '3'} images")
            ctx.image_batch(call)
            assert ("direct_prep" in ctx.profile()) == many
    finally:
        ctx.profile_enable(False)
    err = oracle.peak_normalised_error(_host(imgs), ref)[0]
    _say(f"  160 images vs FP32 oracle {err:.2e}")
    assert err < IMAGE_TOL
    assert np.array_equal(_host(sums), c)


@pytest.mark.parametrize("n", [128, 180, 256])
def test_direct_average_without_images(ctx, n):
    """A direct set imaged for its mean only (no image buffer): the mean from
    the row-pass buffers against the oracle's, and the count sums."""
    import torch
    o = _station(n, seed=950 + n)
    ref, c = _reference(o, "XX")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    avg = torch.full((n, n), float("nan"), dtype=torch.complex64, device="cuda")
    st, keep = _set(o, "XX", None, sums, average=avg)
    flat = np.ascontiguousarray(ref.reshape(-1, n, n))
    for rep in range(2):
        _say(f"direct average only N={n}, rep {rep}")
        with _env(BLINK_DIRECT=1):
            ctx.image_batch([st])
        err = oracle.peak_normalised_error(_host(avg), oracle.image_averaging_complex(flat))[0]
        _say(f"  mean vs oracle {err:.2e}")
        assert err < IMAGE_TOL
        assert np.array_equal(_host(sums), c)


@pytest.mark.parametrize("n", [128, 180, 256])
def test_direct_several_chunks(ctx, n):
    """A direct group cut into row-pass chunks (BLINK_INTER_MIB=1: 15 images
    per chunk at 128^2, 8 at 180^2, 3 at 256^2, of 20), so that the launches
    past the first chunk index their images from the chunk's offset."""
    o = synth.eda2_station(n_ch=20, n_t=1, n_pixels=n, seed=990 + n, height_sigma_m=1.0,
                           freq=synth.mwa_fine_channels(20))
    ref, c = _reference(o, "XX")
    imgs, sums = _buffers(o)
    st, keep = _set(o, "XX", imgs, sums)
    ctx.profile_enable(True)
    ctx.profile_reset()
    try:
        with _env(BLINK_DIRECT=1, BLINK_INTER_MIB=1):
            imgs.fill_(float("nan"))
            _say(f"direct in chunks N={n}")
            ctx.image_batch([st])
            launches = ctx.profile()
    finally:
        ctx.profile_enable(False)
    assert launches["grid_rows_fft"][0] == {128: 2, 180: 3, 256: 7}[n], launches
    assert launches["direct_prep"][0] == 1, launches
    err = oracle.peak_normalised_error(_host(imgs), ref)[0]
    _say(f"  vs FP32 oracle {err:.2e}")
    assert err < IMAGE_TOL
    assert np.array_equal(_host(sums), c)


@pytest.mark.parametrize("n", [128, 256])
def test_direct_forced_sparse_two_timesteps(ctx, n):
    """BLINK_DIRECT=1 puts a sparse set with two timesteps per channel on
    the direct path (each image walks the baselines itself; at 256^2 in
    blocks of whole rows, since the channel's images share their row
    starts)."""
    o = synth.small(n_ant=24, n_ch=3, n_t=2, n_pixels=n, seed=31, height_sigma_m=1.0)
    ref, c = _reference(o, "XX")
    imgs, sums = _buffers(o)
    st, keep = _set(o, "XX", imgs, sums)
    with _env(BLINK_DIRECT=1):
        for rep in range(3):
            eager = rep == 2
            ctx.profile_enable(eager)
            ctx.profile_reset()
            imgs.fill_(float("nan"))
            _say(f"forced direct, rep {rep}")
            ctx.image_batch([st])
            assert oracle.peak_normalised_error(_host(imgs), ref)[0] < IMAGE_TOL
            assert np.array_equal(_host(sums), c)
            if eager:
                assert "direct_prep" in ctx.profile()
    ctx.profile_enable(False)


@pytest.mark.parametrize("n", [128, 180, 256])
def test_direct_group_average_cut_and_plan_path(ctx, n):
    """Two sets in one group (the image -> set table) at 128^2, 180^2 and 256^2,
    the second with a uv-distance cut (the cell test's min_uv branch) and an
    average; then the same call on the plan path: identical count sums,
    images within 1e-6."""
    import torch
    a = _station(n, seed=601)
    b = _station(n, seed=602)
    b.min_uv = 3.0
    ra, ca = _reference(a, "XX")
    rb, cb = _reference(b, "XX")
    ia, sa = _buffers(a)
    ib, sb = _buffers(b)
    avg = torch.empty((n, n), dtype=torch.complex64, device="cuda")
    sta, ka = _set(a, "XX", ia, sa)
    stb, kb = _set(b, "XX", ib, sb, average=avg)
    out = {}
    for mode in ("1", "0"):
        with _env(BLINK_DIRECT=mode):
            for rep in range(3):
                eager = rep == 2
                ctx.profile_enable(eager)
                ctx.profile_reset()
                for t in (ia, ib, avg):
                    t.fill_(float("nan"))
                _say(f"group of two sets N={n}, BLINK_DIRECT={mode}, rep {rep}")
                ctx.image_batch([sta, stb])
                ea = oracle.peak_normalised_error(_host(ia), ra)[0]
                eb = oracle.peak_normalised_error(_host(ib), rb)[0]
                _say(f"  set a {ea:.2e}, set b {eb:.2e}")
                assert ea < IMAGE_TOL and eb < IMAGE_TOL
                assert np.array_equal(_host(sa), ca) and np.array_equal(_host(sb), cb)
                flat = np.ascontiguousarray(rb.reshape(-1, n, n))
                ev = oracle.peak_normalised_error(_host(avg), oracle.image_averaging_complex(flat))[0]
                assert ev < IMAGE_TOL
                if eager:
                    launches = ctx.profile()
                    assert ("direct_prep" in launches) == (mode == "1"), launches
        out[mode] = (_host(ia).copy(), _host(ib).copy(), _host(avg).copy())
    ctx.profile_enable(False)
    for x, y in zip(out["1"], out["0"]):
        assert oracle.peak_normalised_error(x, y)[0] < 1e-6


def test_direct_window_and_blocks(ctx):
    """256^2 images whose cells need more than the window LDS holds (channels
    at 240 and 300 MHz: |cell| up to 102 and 128, the latter wrapping at N/2)
    take blocks of whole rows, beside a set whose channels fit the window, in
    one call (the image -> set table): images within 1e-5 of the oracle,
    counts exact, and within 1e-6 of the plan path."""
    import torch
    a = _station(256, seed=611)
    b = synth.eda2_station(n_ch=2, n_t=1, n_pixels=256, seed=612,
                           freq=np.array([240e6, 300e6]), height_sigma_m=1.0,
                           bl_flag_fraction=0.1)
    ra, ca = _reference(a, "XX")
    rb, cb = _reference(b, "XX")
    ia, sa = _buffers(a)
    ib, sb = _buffers(b)
    sta, ka = _set(a, "XX", ia, sa)
    stb, kb = _set(b, "XX", ib, sb)
    out = {}
    for mode in ("1", "0"):
        with _env(BLINK_DIRECT=mode):
            for rep in range(2):
                ctx.profile_enable(rep == 1)
                ctx.profile_reset()
                for t in (ia, ib):
                    t.fill_(float("nan"))
                _say(f"window and blocks at 256^2, BLINK_DIRECT={mode}, rep {rep}")
                ctx.image_batch([sta, stb])
                ea = oracle.peak_normalised_error(_host(ia), ra)[0]
                eb = oracle.peak_normalised_error(_host(ib), rb)[0]
                _say(f"  window set {ea:.2e}, blocks set {eb:.2e}")
                assert ea < IMAGE_TOL and eb < IMAGE_TOL
                assert np.array_equal(_host(sa), ca) and np.array_equal(_host(sb), cb)
                if rep == 1:
                    assert ("direct_prep" in ctx.profile()) == (mode == "1")
        out[mode] = (_host(ia).copy(), _host(ib).copy())
    ctx.profile_enable(False)
    for x, y in zip(out["1"], out["0"]):
        assert oracle.peak_normalised_error(x, y)[0] < 1e-6


def test_direct_dft_256_keeps_the_plan(ctx):
    """The DFT transforms' direct half plane needs the whole half plane in
    LDS, which 256^2 exceeds: forced direct, such a set stays on the plan
    path, and its images match the oracle."""
    from imager_amd import capi
    o = _station(256, seed=620)
    ref, c = _reference(o, "XX")
    d = {k: _dev(getattr(o, k)) for k in ("vis", "u", "v", "w", "flags", "freq", "cable")}
    imgs, sums = _buffers(o)
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], capi.CORR_GEOMETRIC | capi.CORR_CABLE, o.delta_u,
                       o.delta_v, o.min_uv, 256, "XX", imgs, sums,
                       transform=capi.TRANSFORM_DFT_FP16X3)
    ctx.profile_enable(True)
    ctx.profile_reset()
    try:
        with _env(BLINK_DIRECT=1):
            _say("DFT fp16x3 at 256^2, forced direct")
            ctx.image_batch([st])
            launches = ctx.profile()
    finally:
        ctx.profile_enable(False)
    assert "direct_prep" not in launches and "plan" in launches, launches
    assert oracle.peak_normalised_error(_host(imgs), ref)[0] < IMAGE_TOL
    assert np.array_equal(_host(sums), c)


@pytest.mark.parametrize("n", [128, 180, 256])
def test_direct_all_baselines_flagged(ctx, n):
    """A direct set whose baselines are all flagged grids nothing: count sums
    0 and every pixel NaN (0 * (1/0)), as the oracle's images; a second set
    of the call, with its baselines, images normally (no rows of the empty
    one's leak into it)."""
    import dataclasses
    a = _station(n, seed=640 + n)
    b = dataclasses.replace(a, flags=np.ones_like(a.flags))
    ra, ca = _reference(a, "XX")
    rb, cb = _reference(b, "XX")
    assert not cb.any() and np.isnan(rb).all()
    ia, sa = _buffers(a)
    ib, sb = _buffers(b)
    sta, ka = _set(a, "XX", ia, sa)
    stb, kb = _set(b, "XX", ib, sb)
    for rep in range(2):
        for t in (ia, ib):
            t.fill_(0.0)
        _say(f"all flagged beside a normal set, N={n}, rep {rep}")
        with _env(BLINK_DIRECT=1):
            ctx.image_batch([stb, sta])
        assert np.array_equal(_host(sb), cb)
        assert np.isnan(_host(ib).real).all()
        assert oracle.peak_normalised_error(_host(ia), ra)[0] < IMAGE_TOL
        assert np.array_equal(_host(sa), ca)
