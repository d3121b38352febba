"""Uniform weighting and image averaging on the batched path, against the oracle.

Uniform ("U") weighting divides every cell by its count before the FFT
(src/pacer_imager.cpp:421-428), so an empty cell is 0/0 = NaN and the
reference's FFT makes the whole image NaN.  A meaningful (finite) uniformly
weighted image needs a grid with no empty cell: `covering_array` builds one,
with u, v chosen on integer cells (a channel at c Hz, so lambda = 1 m, and
delta_u = delta_v = 1 m) so that the baselines and their conjugates cover every
cell of the N x N grid.

Averaging: the mean of a set's images (image_averaging_gpu,
src/gpu/gpu_utils.cpp:143-161), formed on the FFT path from the row-pass
buffers without forming the images, checked against the oracle's complex and
real-part means of its reference-precision images (KAT of the averaging itself:
tests/gpu_utils_test.cpp:91-124, in test_gpu_parity.py).
"""
import os

import numpy as np
import pytest

import oracle
from imager_amd import synth

pytestmark = pytest.mark.gpu

IMAGE_TOL = 1e-5
C = 299792458.0


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _say(*a):
    print(*a, flush=True)


def covering_array(n, n_ch=2, n_t=2, seed=5):
    """An observation whose unflagged baselines and their conjugates cover every
    cell of the n x n grid (u, v integers in metres, lambda = 1 m, cell = 1 m)."""
    reps, seen = [], set()
    for v in range(n):
        for u in range(n):
            if (u, v) in seen:
                continue
            seen.add((u, v))
            seen.add(((-u) % n, (-v) % n))
            reps.append((u, v))
    k = len(reps)
    n_ant = 2
    while n_ant * (n_ant - 1) // 2 < k:
        n_ant += 1
    a1, a2 = synth.baseline_antennas(n_ant)
    flags = (a1 == a2).astype(np.int32)
    ok = np.flatnonzero(flags == 0)
    signed = lambda c: c if c <= n // 2 else c - n
    u = np.zeros(len(flags), np.float32)
    v = np.zeros(len(flags), np.float32)
    for j, bl in enumerate(ok):
        cu, cv = reps[j % k]
        u[bl], v[bl] = signed(cu), signed(cv)
    rng = np.random.default_rng(seed)
    vis = (rng.standard_normal((n_t, n_ch, len(flags), 4)) +
           1j * rng.standard_normal((n_t, n_ch, len(flags), 4))).astype(np.complex64)
    o = synth.Observation(name=f"covering_{n}", n_ant=n_ant, x=np.zeros(n_ant),
                          y=np.zeros(n_ant), z=np.zeros(n_ant), u=u, v=v,
                          w=rng.normal(0, 1, len(flags)).astype(np.float32), flags=flags,
                          cable=rng.uniform(50, 300, n_ant), freq=np.full(n_ch, C), vis=vis,
                          n_pixels=n, delta_u=1.0, delta_v=1.0)
    return o


def _reference(o, uniform, corrections=True):
    """The oracle's reference-precision images (FP32 semantics) and counts."""
    v = o.vis
    if corrections:
        v = oracle.cable_corrections(oracle.geometric_corrections(v, o.n_ant, o.w, o.freq),
                                     o.n_ant, o.cable, o.freq)
    g, c, _ = oracle.gridding(v, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, o.n_pixels, "XX", uniform=uniform)
    return oracle.image(g, c), c


def _set(o, images, sums=None, transform=0, weighting="N", average=None, corr=3):
    from imager_amd import capi
    d = dict(vis=_dev(o.vis), u=_dev(o.u), v=_dev(o.v), w=_dev(o.w), flags=_dev(o.flags),
             freq=_dev(o.freq), cable=_dev(o.cable))
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], corr, o.delta_u, o.delta_v, o.min_uv, o.n_pixels,
                       "XX", images, sums, transform=transform, weighting=weighting,
                       average=average)
    return st, d


@pytest.mark.parametrize("n,transform", [(128, 0), (256, 0), (180, 0), (128, 1), (128, 2)])
def test_uniform_weighting_covered_grid(ctx, n, transform):
    """Every cell gridded: finite uniformly weighted images within 1e-5 of the
    oracle (register, generic and both DFT transforms), count sums exact."""
    import torch
    o = covering_array(n)
    ref, c = _reference(o, uniform=True)
    assert np.all(c > 0), "the array must cover every cell"
    assert np.all(np.isfinite(ref))
    images = torch.empty((o.n_t, o.n_ch, n, n), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    st, keep = _set(o, images, sums, transform=transform, weighting="U")
    _say(f"uniform covered N={n} transform={transform}, {o.n_ant} antennas")
    ctx.image_batch([st])
    got = _host(images)
    err = oracle.peak_normalised_error(got, ref)[0]
    _say(f"  vs FP32 oracle {err:.2e}")
    assert err < IMAGE_TOL
    assert np.array_equal(_host(sums), c.reshape(o.n_images, -1).sum(1).astype(np.float64))


@pytest.mark.parametrize("n,transform", [(256, 0), (1024, 0), (180, 0), (4096, 0), (128, 1)])
def test_uniform_weighting_empty_cells_give_nan_images(ctx, n, transform):
    """An ordinary array leaves cells empty: 0/0 = NaN there, so the reference's
    FFT makes every pixel NaN; the batched path gives the same NaN images (real
    parts; the imaginary parts it writes are 0), and exact count sums."""
    import torch
    o = synth.small(n_ant=24, n_ch=2, n_t=1, n_pixels=n, seed=n, height_sigma_m=1.0)
    ref, c = _reference(o, uniform=True)
    assert np.any(c == 0) and np.all(np.isnan(ref.real))
    images = torch.zeros((o.n_t, o.n_ch, n, n), dtype=torch.complex64, device="cuda")
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    st, keep = _set(o, images, sums, transform=transform, weighting="U")
    _say(f"uniform with empty cells N={n} transform={transform}")
    ctx.image_batch([st])
    got = _host(images)
    assert np.all(np.isnan(got.real))
    assert np.array_equal(_host(sums), c.reshape(o.n_images, -1).sum(1).astype(np.float64))


def _check_average(got, ref_images):
    flat = np.ascontiguousarray(ref_images.reshape(-1, *ref_images.shape[-2:]))
    e_c = oracle.peak_normalised_error(got, oracle.image_averaging_complex(flat))[0]
    e_r = oracle.peak_normalised_error(got, oracle.image_averaging_real(flat))[0]
    _say(f"  average vs complex mean {e_c:.2e}, vs real-part mean {e_r:.2e}")
    assert e_c < IMAGE_TOL and e_r < IMAGE_TOL


@pytest.mark.parametrize("n,inter_mib,with_images", [(1024, 9, False), (1024, 0, True),
                                                     (180, 1, False), (4096, 0, False),
                                                     (512, 0, False)])
def test_average_without_forming_images(ctx, n, inter_mib, with_images):
    """The mean image of a set from the row-pass buffers, with and without the
    images (count sums then still written), over several chunks when the
    row-pass budget is cut (inter_mib), against the oracle's means."""
    import torch
    if n == 4096:
        o = synth.mwa_array(n_ch=2, n_t=1, n_pixels=4096, seed=77)
    elif n == 1024:
        o = synth.mwa_array(n_ch=6, n_t=2, n_pixels=1024, seed=78)
    else:
        o = synth.small(n_ant=64, n_ch=5, n_t=2, n_pixels=n, seed=n, height_sigma_m=1.0)
    ref, c = _reference(o, uniform=False)
    avg = torch.full((n, n), float("nan"), dtype=torch.complex64, device="cuda")
    images = (torch.empty((o.n_t, o.n_ch, n, n), dtype=torch.complex64, device="cuda")
              if with_images else None)
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    st, keep = _set(o, images, sums, average=avg)
    old = os.environ.get("BLINK_INTER_MIB")
    if inter_mib:
        os.environ["BLINK_INTER_MIB"] = str(inter_mib)
    try:
        for rep in range(2):   # capture, then replay
            avg.fill_(float("nan"))
            _say(f"average N={n} inter_mib={inter_mib} images={with_images} rep {rep}")
            ctx.image_batch([st])
            _check_average(_host(avg), ref)
            assert np.array_equal(_host(sums), c.reshape(o.n_images, -1).sum(1).astype(np.float64))
            if with_images:
                assert oracle.peak_normalised_error(_host(images), ref)[0] < IMAGE_TOL
    finally:
        if old is None:
            os.environ.pop("BLINK_INTER_MIB", None)
        else:
            os.environ["BLINK_INTER_MIB"] = old


def test_average_uniform_and_dft_and_mixed_group(ctx):
    """The mean of uniformly weighted images (covered 128^2 grid); a DFT set's
    mean (formed from its images); and one 256^2 group holding a set imaged for
    its mean only next to a set imaged normally."""
    import torch
    # uniform + average
    o = covering_array(128)
    ref, _ = _reference(o, uniform=True)
    avg = torch.empty((128, 128), dtype=torch.complex64, device="cuda")
    st, keep = _set(o, None, average=avg, weighting="U")
    _say("average of uniformly weighted images, 128^2")
    ctx.image_batch([st])
    _check_average(_host(avg), ref)
    # DFT + average (the images are formed, then averaged)
    o2 = synth.small(n_ant=48, n_ch=3, n_t=2, n_pixels=128, seed=12, height_sigma_m=1.0)
    ref2, _ = _reference(o2, uniform=False)
    img2 = torch.empty((o2.n_t, o2.n_ch, 128, 128), dtype=torch.complex64, device="cuda")
    avg2 = torch.empty((128, 128), dtype=torch.complex64, device="cuda")
    st2, keep2 = _set(o2, img2, average=avg2, transform=1)
    _say("average of DFT images, 128^2")
    ctx.image_batch([st2])
    _check_average(_host(avg2), ref2)
    assert oracle.peak_normalised_error(_host(img2), ref2)[0] < IMAGE_TOL
    # one group: mean only + images only
    a = synth.small(n_ant=40, n_ch=2, n_t=2, n_pixels=256, seed=21, height_sigma_m=1.0)
    b = synth.small(n_ant=56, n_ch=3, n_t=1, n_pixels=256, seed=22, height_sigma_m=1.0)
    ra, ca = _reference(a, uniform=False)
    rb, cb = _reference(b, uniform=False)
    avg_a = torch.empty((256, 256), dtype=torch.complex64, device="cuda")
    sums_a = torch.empty(a.n_images, dtype=torch.float64, device="cuda")
    img_b = torch.empty((b.n_t, b.n_ch, 256, 256), dtype=torch.complex64, device="cuda")
    sta, ka = _set(a, None, sums_a, average=avg_a)
    stb, kb = _set(b, img_b)
    for rep in range(2):
        _say(f"group of a mean-only set and an image set, rep {rep}")
        ctx.image_batch([sta, stb])
        _check_average(_host(avg_a), ra)
        assert np.array_equal(_host(sums_a), ca.reshape(a.n_images, -1).sum(1).astype(np.float64))
        assert oracle.peak_normalised_error(_host(img_b), rb)[0] < IMAGE_TOL


def test_batch_argument_checks(ctx):
    """Neither images nor average, an unknown weighting, and a DFT set without
    images are refused."""
    import torch
    from imager_amd import capi
    o = synth.small(n_ant=8, n_ch=1, n_t=1, n_pixels=64)
    img = torch.empty((1, 1, 64, 64), dtype=torch.complex64, device="cuda")
    avg = torch.empty((64, 64), dtype=torch.complex64, device="cuda")
    for kw in (dict(images=None), dict(images=img, weighting=7),
               dict(images=None, average=avg, transform=capi.TRANSFORM_DFT_BF16X3)):
        images = kw.pop("images")
        st, keep = _set(o, images, **kw)
        with pytest.raises(capi.BlinkInvalidArgument):
            ctx.image_batch([st])


@pytest.mark.parametrize("n,overlap_mib", [(1024, 10), (256, 1)])
def test_overlapped_chunks(ctx, n, overlap_mib):
    """Overlapped chunks (each chunk's second pass on the pass stream beside
    the next chunk's row pass, two row-pass buffers in turn, cached row-pass
    stores), forced by a small target chunk size on sets with several
    timesteps per channel: one group of two sets, one of them also averaged
    (its row sums accumulate across chunks on the pass stream), then a
    uniformly weighted set; every image, count sum and mean against the
    oracle, on capture, on replay and eagerly (where the launches are counted:
    several row / second-pass pairs)."""
    import torch
    if n == 1024:
        a = synth.mwa_array(n_ch=3, n_t=3, n_pixels=1024, seed=81)
        b = synth.mwa_array(n_ch=2, n_t=2, n_pixels=1024, seed=82)
    else:
        a = synth.small(n_ant=48, n_ch=4, n_t=3, n_pixels=n, seed=83, height_sigma_m=1.0)
        b = synth.small(n_ant=64, n_ch=2, n_t=2, n_pixels=n, seed=84, height_sigma_m=1.0)
    per_image = 8 * (n // 2 + 1) * n
    n_img = a.n_images + b.n_images
    assert n_img * per_image > 3 * (overlap_mib << 20), "3+ chunks wanted"
    assert n_img >= 2 * (a.n_ch + b.n_ch), "plans shared by 2+ images each"
    ra, ca = _reference(a, uniform=False)
    rb, cb = _reference(b, uniform=False)
    img_a = torch.empty((a.n_t, a.n_ch, n, n), dtype=torch.complex64, device="cuda")
    img_b = torch.empty((b.n_t, b.n_ch, n, n), dtype=torch.complex64, device="cuda")
    avg_a = torch.empty((n, n), dtype=torch.complex64, device="cuda")
    sums_a = torch.empty(a.n_images, dtype=torch.float64, device="cuda")
    sums_b = torch.empty(b.n_images, dtype=torch.float64, device="cuda")
    sta, ka = _set(a, img_a, sums_a, average=avg_a)
    stb, kb = _set(b, img_b, sums_b)
    u = covering_array(n if n <= 256 else 256, n_ch=3, n_t=2, seed=85)
    ru, _ = _reference(u, uniform=True)
    img_u = torch.empty((u.n_t, u.n_ch, u.n_pixels, u.n_pixels), dtype=torch.complex64,
                        device="cuda")
    stu, ku = _set(u, img_u, weighting="U")
    old = os.environ.get("BLINK_OVERLAP_MIB")
    os.environ["BLINK_OVERLAP_MIB"] = str(overlap_mib)
    try:
        for rep in range(3):   # capture, replay, then eagerly with the launches counted
            eager = rep == 2
            ctx.profile_enable(eager)
            ctx.profile_reset()
            for t in (img_a, img_b, avg_a, img_u):
                t.fill_(float("nan"))
            _say(f"overlapped chunks N={n} target {overlap_mib} MiB, rep {rep}")
            ctx.image_batch([sta, stb])
            ea = oracle.peak_normalised_error(_host(img_a), ra)[0]
            eb = oracle.peak_normalised_error(_host(img_b), rb)[0]
            _say(f"  set a {ea:.2e}, set b {eb:.2e}")
            assert ea < IMAGE_TOL and eb < IMAGE_TOL
            assert np.array_equal(_host(sums_a), ca.reshape(a.n_images, -1).sum(1).astype(np.float64))
            assert np.array_equal(_host(sums_b), cb.reshape(b.n_images, -1).sum(1).astype(np.float64))
            _check_average(_host(avg_a), ra)
            if eager:
                launches = ctx.profile()
                _say(f"  launches {launches}")
                assert launches["grid_rows_fft"][0] >= 3
                # the images' second passes and the average's
                assert launches["fft_cols"][0] == launches["grid_rows_fft"][0] + 1
            ctx.image_batch([stu])
            eu = oracle.peak_normalised_error(_host(img_u), ru)[0]
            _say(f"  uniform {eu:.2e}")
            assert eu < IMAGE_TOL
    finally:
        ctx.profile_enable(False)
        if old is None:
            os.environ.pop("BLINK_OVERLAP_MIB", None)
        else:
            os.environ["BLINK_OVERLAP_MIB"] = old
