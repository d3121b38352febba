"""Pins the CPU oracle (oracle/blink_oracle.c) before anything is compared to it.

Every check here is independent of the oracle's own formulas: the reference's
known-answer tests, a textbook DFT in extended precision, numpy's FFT
(pocketfft), brute-force nearest-cell search, closed-form images and
physical invariants.
"""
import numpy as np
import pytest

import oracle
from imager_amd import synth

C = 299792458.0


# ------------------------------------------------------------- fft_shift KATs

def test_fft_shift_kat_even():
    # tests/imager_test.cpp:21-28
    a = np.array([0, 1, 2, 3], np.complex64).reshape(2, 2)
    out = oracle.fft_shift(a)
    assert np.array_equal(out.reshape(-1), np.array([3, 2, 1, 0], np.complex64))


def test_fft_shift_kat_odd():
    # tests/imager_test.cpp:30-37: 3 (x) by 2 (y); swap semantics, not a roll
    a = np.arange(6).astype(np.complex64).reshape(2, 3)
    out = oracle.fft_shift(a)
    assert np.array_equal(out.reshape(-1), np.array([5, 4, 3, 2, 1, 0], np.complex64))


def test_fft_shift_and_norm_kat_three_images():
    # tests/gpu_utils_test.cpp:60-87 (fnorm = 1)
    for _ in range(3):
        a = np.array([0, 1, 2, 3], np.complex64).reshape(2, 2)
        assert np.array_equal(oracle.fft_shift(a).reshape(-1), np.array([3, 2, 1, 0], np.complex64))


@pytest.mark.parametrize("n", [2, 4, 6, 8, 180, 256])
def test_fft_shift_even_is_numpy_fftshift(n):
    rng = np.random.default_rng(n)
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    assert np.array_equal(oracle.fft_shift(a), np.fft.fftshift(a))


@pytest.mark.parametrize("n", [3, 5, 7, 9])
def test_fft_shift_odd_is_involution_fixing_centre(n):
    a = (np.arange(n * n) + 1j * np.arange(n * n)[::-1]).astype(np.complex64).reshape(n, n)
    once = oracle.fft_shift(a)
    assert np.array_equal(oracle.fft_shift(once), a)
    assert once[n // 2, n // 2] == a[n // 2, n // 2]


# -------------------------------------------------------------------- FFT

def _naive_dft2(x, sign):
    """Textbook 2-D DFT in extended precision."""
    n = x.shape[-1]
    k = np.arange(n)
    w = np.exp(sign * 2j * np.pi * np.outer(k, k).astype(np.longdouble) / n)
    xl = x.astype(np.clongdouble)
    return np.einsum("yv,...vu,xu->...yx", w, xl, w)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 20, 22, 25, 26,
                               30, 39, 45, 49, 77])
@pytest.mark.parametrize("sign", [+1, -1])
def test_fft_matches_textbook_dft(n, sign):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    ref = _naive_dft2(x, sign)
    got64 = oracle.fft2d(x.astype(np.complex128), sign)
    assert np.max(np.abs(got64 - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    got32 = oracle.fft2d(x.astype(np.complex64), sign)
    assert np.max(np.abs(got32 - ref)) <= 2e-6 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [64, 128, 154, 180, 208, 256])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    # backward, unnormalised == ifft2 * N^2 ; forward == fft2
    back = np.fft.ifft2(x) * n * n
    fwd = np.fft.fft2(x)
    assert np.max(np.abs(oracle.fft2d(x, +1) - back)) / np.max(np.abs(back)) < 1e-13
    assert np.max(np.abs(oracle.fft2d(x, -1) - fwd)) / np.max(np.abs(fwd)) < 1e-13
    e32 = np.max(np.abs(oracle.fft2d(x.astype(np.complex64), +1) - back)) / np.max(np.abs(back))
    assert e32 < 2e-6


def test_fft_parseval_and_roundtrip():
    rng = np.random.default_rng(3)
    n = 60
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    X = oracle.fft2d(x, +1)
    assert abs(np.sum(np.abs(X) ** 2) - n * n * np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(X) ** 2)
    back = oracle.fft2d(X, -1) / (n * n)
    assert np.max(np.abs(back - x)) < 1e-12


def test_fft_unsupported_size_raises():
    with pytest.raises(ValueError):
        oracle.fft2d(np.zeros((17, 17), np.complex128))


# ------------------------------------------------------------- cell indices

def _brute_cell(q):
    """Nearest integer by exhaustive search, ties away from zero."""
    lo = int(np.floor(q)) - 1
    cands = np.arange(lo, lo + 4)
    d = np.abs(cands - q)
    best = cands[d == d.min()]
    return int(best[np.argmax(np.abs(best))])


def test_cells_match_brute_force_nearest_search():
    obs = synth.small(n_ant=24, n_ch=3, n_t=1, n_pixels=48, seed=11)
    cl = oracle.cells(obs.u, obs.v, obs.flags, obs.freq, obs.delta_u, obs.delta_v, obs.min_uv,
                      obs.n_pixels)
    N = obs.n_pixels
    for ch, f in enumerate(obs.freq):
        lam = C / f  # IEEE double division, as src/pacer_imager.cpp:289
        for bl in range(obs.n_bl):
            if obs.flags[bl]:
                assert (cl[ch, bl] == -1).all()
                continue
            up = _brute_cell((float(obs.u[bl]) / lam) / obs.delta_u)
            vp = _brute_cell((float(obs.v[bl]) / lam) / obs.delta_v)
            assert tuple(cl[ch, bl]) == (up % N, vp % N, (-up) % N, (-vp) % N)


def test_cells_exact_half_rounds_away_from_zero():
    # u / lambda / delta_u = +-2.5 exactly: round() gives +-3 (rint would give +-2).
    f = C  # lambda = 1 m exactly
    N = 16
    assert oracle.cell(2.5, 0.0, f, 1.0, 1.0, -np.inf, N)[:2] == (3, 0)
    assert oracle.cell(-2.5, 0.0, f, 1.0, 1.0, -np.inf, N)[:2] == (N - 3, 0)
    assert oracle.cell(0.0, -0.5, f, 1.0, 1.0, -np.inf, N)[:2] == (0, N - 1)
    # conjugate cell of +2.5 is -3 -> N-3
    assert oracle.cell(2.5, 0.0, f, 1.0, 1.0, -np.inf, N)[2] == N - 3


def test_cells_min_uv_cut_and_wrap():
    f = C
    N = 8
    assert oracle.cell(1.0, 1.0, f, 1.0, 1.0, 2.0, N) is None        # sqrt(2) <= 2: cut
    assert oracle.cell(3.0, 0.0, f, 1.0, 1.0, 2.0, N) is not None    # 3 > 2
    # -N <= i < 0 wraps to N + i, exactly like wrap_index
    assert oracle.cell(-8.0, 0.0, f, 1.0, 1.0, -np.inf, N)[0] == 0
    assert oracle.cell(-7.0, 0.0, f, 1.0, 1.0, -np.inf, N)[0] == 1
    assert oracle.cell(9.0, 0.0, f, 1.0, 1.0, -np.inf, N)[0] == 1


def test_cell_extent_is_the_cell_of_the_largest_coordinate():
    """|cell| = |round((x / wavelength) / cell size)| never decreases with |x|,
    so every gridded cell of a channel lies within the cells of the largest
    unflagged |u| and |v| (the bound by which the direct row pass at 256^2
    sizes its window of the half plane).  Random layouts of 8-63 antennas,
    2-40 % of the baselines flagged, 50-300 MHz and oversampling 2.5-4 on a
    side large enough that no cell wraps."""
    rng = np.random.default_rng(12)
    n = 8192

    def unwrap(a):
        return np.where(a >= n // 2, a - n, a)

    for trial in range(16):
        n_ant = int(rng.integers(8, 64))
        scale = rng.uniform(5.0, 1500.0)
        x, y, z = (rng.normal(0.0, s, n_ant) for s in (scale, scale, 1.0))
        u, v, w = oracle.zenith_uvw(x, y, z)
        flags = (oracle.baseline_flags(n_ant).astype(bool)
                 | (rng.random(len(u)) < rng.uniform(0.02, 0.4))).astype(np.int32)
        u_max, v_max = oracle.uv_max(u, v, flags)
        du, dv = oracle.delta_uv(float(np.float32(rng.uniform(2.5, 4.0))), u_max, v_max, n)
        freq = rng.uniform(50e6, 300e6, 3)
        cl = oracle.cells(u, v, flags, freq, du, dv, -np.inf, n)
        good = flags == 0
        mu = float(np.abs(np.asarray(u, np.float32)[good]).max())
        mv = float(np.abs(np.asarray(v, np.float32)[good]).max())
        for ch, f in enumerate(freq):
            ok = cl[ch, :, 0] >= 0
            eu = abs(int(unwrap(np.array(oracle.cell(mu, 0.0, f, du, dv, -np.inf, n)[0]))))
            ev = abs(int(unwrap(np.array(oracle.cell(0.0, mv, f, du, dv, -np.inf, n)[1]))))
            assert eu < n // 2 and ev < n // 2
            assert np.abs(unwrap(cl[ch, ok, 0])).max() == eu, (trial, ch)
            assert np.abs(unwrap(cl[ch, ok, 1])).max() == ev, (trial, ch)


# ----------------------------------------------------------------- gridding

@pytest.fixture(scope="module")
def small_obs():
    return synth.small(n_ant=20, n_ch=2, n_t=2, n_pixels=40, seed=5)


def test_gridding_invariants(small_obs):
    o = small_obs
    g, c, n = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, o.n_pixels, "XX")
    assert n == o.n_images * int((o.flags == 0).sum())
    assert c.sum() == 2 * n
    # counts symmetric under (u, v) -> (-u, -v); grid Hermitian
    N = o.n_pixels
    idx = (-np.arange(N)) % N
    cm = c[..., idx, :][..., :, idx]
    gm = g[..., idx, :][..., :, idx]
    assert np.array_equal(c, cm)
    assert np.allclose(g, np.conj(gm), atol=1e-5)


def test_gridding_matches_brute_force_fp64(small_obs):
    o = small_obs
    g, c, _ = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, o.n_pixels, "XX")
    N = o.n_pixels
    ref = np.zeros((o.n_t, o.n_ch, N, N), np.complex128)
    cref = np.zeros((o.n_t, o.n_ch, N, N))
    for t in range(o.n_t):
        for ch, f in enumerate(o.freq):
            lam = C / f
            for bl in range(o.n_bl):
                if o.flags[bl]:
                    continue
                up = _brute_cell((float(o.u[bl]) / lam) / o.delta_u)
                vp = _brute_cell((float(o.v[bl]) / lam) / o.delta_v)
                val = complex(o.vis[t, ch, bl, 0])
                ref[t, ch, vp % N, up % N] += val
                ref[t, ch, (-vp) % N, (-up) % N] += val.conjugate()
                cref[t, ch, vp % N, up % N] += 1
                cref[t, ch, (-vp) % N, (-up) % N] += 1
    assert np.array_equal(c, cref)
    assert np.max(np.abs(g - ref)) < 1e-4


def test_gridding_skips_nan_maxvis_and_flags(small_obs):
    o = small_obs
    vis = o.vis.copy()
    good = np.where(o.flags == 0)[0]
    vis[0, 0, good[0], 0] = np.nan + 0j
    vis[0, 0, good[1], 0] = 2e12 + 0j
    # MAX_VIS = 1e12 is compared in double: float32(1e12) = 999999995904 is kept,
    # the next float above 1e12 (1000000061440) is skipped.
    vis[0, 0, good[2], 0] = 0 + 1.00000007e12j
    vis[0, 0, good[3], 0] = 0 + 1e12j
    _, c, n = oracle.gridding(vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, o.n_pixels, "XX")
    per_img = int((o.flags == 0).sum())
    assert n == o.n_images * per_img - 3
    assert c[0, 0].sum() == 2 * (per_img - 3)


def test_gridding_polarisations_closed_form(small_obs):
    o = small_obs
    args = (o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv, o.n_pixels)
    gxx, _, _ = oracle.gridding(o.vis, *args, pol="XX", precision="f64")
    gyy, _, _ = oracle.gridding(o.vis, *args, pol="YY", precision="f64")
    gi, _, _ = oracle.gridding(o.vis, *args, pol="I", precision="f64")
    assert np.allclose(gi, 0.5 * (gxx + gyy), atol=1e-5)
    # V = (i/2)(XY - YX): grid a copy whose XX holds that value
    v2 = o.vis.copy()
    v2[..., 0] = (0.5j * (o.vis[..., 1].astype(np.complex128) - o.vis[..., 2])).astype(np.complex64)
    gv, _, _ = oracle.gridding(o.vis, *args, pol="V", precision="f64")
    gv2, _, _ = oracle.gridding(v2, *args, pol="XX", precision="f64")
    assert np.allclose(gv, gv2, atol=1e-5)


def test_uniform_weighting_divides_by_count(small_obs):
    o = small_obs
    args = (o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv, o.n_pixels)
    g, c, _ = oracle.gridding(*args, pol="XX")
    gu, cu, _ = oracle.gridding(*args, pol="XX", uniform=True)
    assert np.array_equal(c, cu)
    occ = c > 0
    assert np.allclose(gu[occ], g[occ] / c[occ], rtol=1e-6)
    assert np.isnan(gu[~occ]).all()  # 0/0 in empty cells, as the reference


# -------------------------------------------------------------------- image

def test_image_single_visibility_closed_form():
    """One baseline at cell (up, vp): the dirty image is the cosine fringe
    Re(V exp(+2 pi i (up x' + vp y') / N)) with x' = x - N/2 (after the shift)."""
    N = 32
    up, vp = 3, -5
    V = 2.0 - 0.5j
    g = np.zeros((N, N), np.complex128)
    c = np.zeros((N, N))
    g[vp % N, up % N] += V
    g[(-vp) % N, (-up) % N] += np.conj(V)
    c[vp % N, up % N] += 1
    c[(-vp) % N, (-up) % N] += 1
    img = oracle.image(g[None], c[None])[0]
    xs = np.arange(N) - N // 2
    X, Y = np.meshgrid(xs, xs)
    ref = np.real(V * np.exp(2j * np.pi * (up * X + vp * Y) / N))
    assert np.max(np.abs(img - ref)) < 1e-12
    img32 = oracle.image(g[None].astype(np.complex64), c[None].astype(np.float32))[0]
    assert np.max(np.abs(img32 - ref)) < 1e-6


def test_image_dc_pixel_is_mean_real_part(small_obs):
    o = small_obs
    g, c, n = oracle.gridding(o.vis, o.n_ant, o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v,
                              o.min_uv, o.n_pixels, "XX", precision="f64")
    img = oracle.image(g, c)
    N = o.n_pixels
    ok = o.flags == 0
    for t in range(o.n_t):
        for ch in range(o.n_ch):
            mean_re = np.mean(o.vis[t, ch, ok, 0].real.astype(np.float64))
            assert abs(img[t, ch, N // 2, N // 2].real - mean_re) < 1e-9
            # Hermitian grid -> real image
            assert np.max(np.abs(img[t, ch].imag)) < 1e-9 * np.max(np.abs(img[t, ch]))


def test_image_f32_close_to_f64(small_obs):
    o = small_obs
    args = (o.n_ant, o.u, o.v, o.w, o.flags, o.freq, o.cable, False, False, o.delta_u,
            o.delta_v, o.min_uv, o.n_pixels)
    i32, _ = oracle.run_f32(o.vis, *args)
    i64 = oracle.run_f64(o.vis, *args)
    assert oracle.peak_normalised_error(i32, i64)[0] < 1e-6


# -------------------------------------------------------------- corrections

def test_geometric_correction_removes_w_phase():
    n_ant = 6
    rng = np.random.default_rng(1)
    x, y, z = rng.normal(0, 20, n_ant), rng.normal(0, 20, n_ant), rng.normal(0, 3, n_ant)
    u, v, w = synth.zenith_uvw(x, y, z)
    freq = np.array([150e6, 175e6, 199e6])
    lam = C / freq
    # a w-phase exp(-2 pi i w / lambda) on every pol product
    vis = np.exp(-2j * np.pi * w.astype(np.float64)[None, :] / lam[:, None])
    vis = np.repeat(vis[None, :, :, None], 4, axis=3).astype(np.complex64)
    out = oracle.geometric_corrections(vis, n_ant, w, freq)
    assert np.max(np.abs(out - 1.0)) < 3e-7


def test_cable_correction_removes_cable_phase():
    n_ant = 7
    rng = np.random.default_rng(2)
    L = rng.uniform(50, 300, n_ant)
    freq = np.array([160e6, 190e6])
    a1, a2 = synth.baseline_antennas(n_ant)
    lam = C / freq
    vis = np.exp(+2j * np.pi * (L[a2] - L[a1])[None, :] / lam[:, None])
    vis = np.repeat(vis[None, :, :, None], 4, axis=3).astype(np.complex64)
    out = oracle.cable_corrections(vis, n_ant, L, freq)
    assert np.max(np.abs(out - 1.0)) < 3e-7


def test_corrections_preserve_modulus_commute_and_identity(small_obs):
    o = small_obs
    g = oracle.geometric_corrections(o.vis, o.n_ant, o.w, o.freq)
    assert np.allclose(np.abs(g), np.abs(o.vis), rtol=1e-6, atol=1e-7)
    gc = oracle.cable_corrections(g, o.n_ant, o.cable, o.freq)
    cg = oracle.geometric_corrections(oracle.cable_corrections(o.vis, o.n_ant, o.cable, o.freq),
                                      o.n_ant, o.w, o.freq)
    assert np.max(np.abs(gc - cg)) < 1e-5
    assert np.array_equal(oracle.geometric_corrections(o.vis, o.n_ant, np.zeros_like(o.w), o.freq),
                          o.vis)
    same = np.full(o.n_ant, 123.0)
    assert np.array_equal(oracle.cable_corrections(o.vis, o.n_ant, same, o.freq), o.vis)


# --------------------------------------------------------------- misc KATs

def test_image_averaging_kat():
    # tests/gpu_utils_test.cpp:91-124: 4 images of 512^2, real values i % 100
    n_img, side = 4, 512
    vals = (np.arange(side * side) % 100).astype(np.float32)
    imgs = np.repeat(vals[None], n_img, 0).astype(np.complex64).reshape(n_img, side, side)
    ref = vals.reshape(side, side).astype(np.complex64)
    assert np.array_equal(oracle.image_averaging_complex(imgs), ref)
    assert np.array_equal(oracle.image_averaging_real(imgs), ref)


def test_baseline_flags_and_uv_max():
    f = oracle.baseline_flags(5, [2])
    a1, a2 = synth.baseline_antennas(5)
    assert np.array_equal(f, ((a1 == a2) | (a1 == 2) | (a2 == 2)).astype(np.int32))
    assert np.array_equal(f, synth.baseline_flags(5, [2]))
    u = np.array([0, 3, -7, 1, 2, 0, 9, 1, 1, 0, -4, 4, 4, 4, 0], np.float32)
    v = -u
    um, vm = oracle.uv_max(u, v, f)
    ok = f == 0
    assert um == np.max(np.abs(u[ok])) and vm == np.max(np.abs(v[ok]))


def test_frequency_cotter():
    # 768 x 40 kHz fine channels of coarse channel 145 x 1.28 MHz ... the
    # formula of src/pacer_imager.cpp:545-546 for channel 0 and 1
    f0 = oracle.frequency_hz(0.04, 1, 1.28, 133, 0, True)
    f1 = oracle.frequency_hz(0.04, 1, 1.28, 133, 1, True)
    assert abs(f0 - (133 * 1.28 - 0.64) * 1e6) < 1e-3
    assert abs((f1 - f0) - 40e3) < 1e-3
    g0 = oracle.frequency_hz(0.04, 1, 1.28, 133, 0, False)
    assert abs((g0 - f0) - 20e3) < 1e-3


def test_zenith_uvw_hand_kat():
    """Three antennas with answers worked out by hand.  Baselines are
    bl = a1 (a1 + 1) / 2 + a2, a2 <= a1: 0 = (0,0), 1 = (1,0), 2 = (1,1),
    3 = (2,0), 4 = (2,1), 5 = (2,2); u = -(x[a1] - x[a2]) formed in double
    and stored as FP32 (src/antenna_positions.cpp:81-89), v, w likewise."""
    x = np.array([1.5, -2.25, 0.1])
    y = np.array([16777217.0, 16777216.0, 0.0])   # 2^24 + 1 has no FP32 value
    z = np.array([0.0, 0.5, -0.25])
    u, v, w = oracle.zenith_uvw(x, y, z)
    assert u.dtype == np.float32 and len(u) == 6
    for a in (u, v, w):                       # autos
        assert a[0] == 0 and a[2] == 0 and a[5] == 0
    # -(x1 - x0) = 3.75 exactly; a swapped sign or operand order gives -3.75
    assert u[1] == 3.75
    # -(0.1 - 1.5) = 1.4 in double, stored as 11744051 / 2^23
    assert float(u[3]) == 11744051 / 2 ** 23
    # -(0.1 + 2.25) = -2.35, stored as -4928307 / 2^21
    assert float(u[4]) == -4928307 / 2 ** 21
    # the difference is formed in double: FP32 operands would both be 2^24
    # and give 0
    assert v[1] == 1.0
    # 2^24 + 1 rounds to even in FP32
    assert v[3] == 16777216.0 and v[4] == 16777216.0
    assert w[1] == -0.5 and w[3] == 0.25 and w[4] == 0.75


def test_delta_uv_closed_forms():
    """delta_u = oversampling * u_max / N with the oversampling factor a float
    member promoted to double (src/pacer_imager.cpp:497-499, src/pacer_imager.h:65)."""
    du, dv = oracle.delta_uv(2.0, 1024.0, 512.0, 1024)
    assert du == 2.0 and dv == 1.0            # u_max feeds delta_u, v_max delta_v
    du, dv = oracle.delta_uv(2.0, 90.0, 45.0, 180)
    assert du == 1.0 and dv == 0.5
    # 1.9 is not an FP32 value: the factor enters as 15938355 / 2^23
    du, dv = oracle.delta_uv(1.9, 1024.0, 2048.0, 1024)
    assert du == 15938355 / 2 ** 23 and dv == 2 * 15938355 / 2 ** 23
    assert du != 1.9
    # every baseline flagged: u_max = v_max = -inf
    du, dv = oracle.delta_uv(2.0, -np.inf, -np.inf, 180)
    assert du == -np.inf and dv == -np.inf


# ------------------------------------------ convolutional (Kaiser-Bessel) mode

def _kb_1d(table):
    """The 1-D kernel of K * 1023 samples the table rows were cut from
    (row r, column c = sample c*1023 + 1022 - r)."""
    K = table.shape[1]
    k1 = np.empty(K * 1023)
    for r in range(1023):
        k1[np.arange(K) * 1023 + (1022 - r)] = table[r]
    return k1


@pytest.mark.parametrize("K", [1, 5, 7, 11])
@pytest.mark.parametrize("with_sinc", [True, False])
def test_kb_kernel_is_kaiser_window_times_sinc(K, with_sinc):
    """Against scipy's Kaiser window (beta 8.6, an independent I0): sample m
    of the 1-D kernel is 1023 * kaiser[m] * sinc(m - mid) (odd K)."""
    from scipy.signal.windows import kaiser
    t = oracle.kb_kernels(K, with_sinc)
    k1 = _kb_1d(t)
    n = K * 1023
    mid = n // 2
    d = np.abs(np.arange(n) - mid).astype(np.float64)
    ratio = 1.0 / 1023
    sinc = np.where(d == 0, ratio, np.sin(np.pi * ratio * d) / (np.pi * np.maximum(d, 1)))
    if not with_sinc:
        sinc = np.full(n, ratio)
    expect = 1023 * kaiser(n, 8.6, sym=True) * sinc
    assert np.allclose(k1, expect, rtol=1e-7, atol=1e-12)


@pytest.mark.parametrize("K", [5, 7, 10])
def test_kb_kernel_rows(K):
    t = oracle.kb_kernels(K, True)
    # mirror: row r read backwards is row 1022 - r
    assert np.array_equal(t[::-1, ::-1], t)
    if K % 2:
        # zero fractional offset (row 511): the windowed sinc interpolates,
        # weight 1 at the centre cell and 0 at the others
        c = t[511]
        assert abs(c[K // 2] - 1.0) < 1e-12
        assert np.max(np.abs(np.delete(c, K // 2))) < 1e-15


def _kb_obs(u_cells, v_cells, vis_xx, n_pixels=64):
    """Single-channel observation at lambda = 1 m with delta_u = delta_v = 1,
    so u_m / lambda / delta_u is the given (possibly fractional) cell."""
    n_bl = len(u_cells)
    n_ant = int((np.sqrt(8 * n_bl + 1) - 1) / 2)
    assert n_ant * (n_ant + 1) // 2 == n_bl
    vis = np.zeros((1, 1, n_bl, 4), np.complex64)
    vis[0, 0, :, 0] = vis_xx
    freq = np.array([oracle.VEL_LIGHT if hasattr(oracle, "VEL_LIGHT") else 299792458.0])
    u = np.asarray(u_cells, np.float32)
    v = np.asarray(v_cells, np.float32)
    flags = np.zeros(n_bl, np.int32)
    return vis, n_ant, u, v, flags, freq, n_pixels


def test_kb_at_cell_centres_equals_nearest_neighbour():
    rng = np.random.default_rng(3)
    n_bl = 10
    u = rng.integers(-20, 20, n_bl)
    v = rng.integers(-20, 20, n_bl)
    val = (rng.normal(size=n_bl) + 1j * rng.normal(size=n_bl)).astype(np.complex64)
    vis, n_ant, u, v, flags, freq, N = _kb_obs(u, v, val)
    g_nn, c_nn, _ = oracle.gridding(vis, n_ant, u, v, flags, freq, 1.0, 1.0, -np.inf, N,
                                    precision="f64")
    g_kb, c_kb, _ = oracle.gridding(vis, n_ant, u, v, flags, freq, 1.0, 1.0, -np.inf, N,
                                    precision="f64", kernel_size=7)
    assert np.array_equal(c_nn, c_kb)
    assert np.max(np.abs(g_kb - g_nn)) < 1e-12 * np.max(np.abs(g_nn))


def test_kb_grid_hermitian_and_shift_equivariant():
    """One visibility at a fractional cell: the grid is Hermitian (the
    conjugate footprint mirrors the visibility's) and moving the visibility
    by whole cells moves its footprint by as many cells."""
    N = 64
    u0, v0 = 5.3, -7.42
    val = np.complex64(1.5 - 0.25j)
    vis, n_ant, u, v, flags, freq, _ = _kb_obs([u0], [v0], [val], N)
    g, c, n = oracle.gridding(vis, n_ant, u, v, flags, freq, 1.0, 1.0, -np.inf, N,
                              precision="f64", kernel_size=7)
    assert n == 1
    g = g[0, 0]
    mirrored = np.conj(np.roll(np.roll(g[::-1, ::-1], 1, axis=0), 1, axis=1))
    assert np.max(np.abs(g - mirrored)) < 1e-12
    # footprint: 7 x 7 cells around (5, -7) and their mirror
    assert np.count_nonzero(np.abs(g) > 0) == 2 * 49
    vis2, n_ant, u2, v2, flags, freq, _ = _kb_obs([u0 + 3], [v0 - 2], [val], N)
    g2, _, _ = oracle.gridding(vis2, n_ant, u2, v2, flags, freq, 1.0, 1.0, -np.inf, N,
                               precision="f64", kernel_size=7)
    # the visibility's own footprint moves by (+3, -2); compare away from the mirror
    sl = np.zeros((N, N), bool)
    for j in range(-3, 4):
        for i in range(-3, 4):
            sl[(-7 + j) % N, (5 + i) % N] = True
    moved = np.roll(np.roll(np.where(sl, g, 0), -2, axis=0), 3, axis=1)
    assert np.max(np.abs(np.where(np.roll(np.roll(sl, -2, 0), 3, 1), g2[0, 0], 0) - moved)) < 1e-6


def test_kb_counts_and_weight_sum():
    """Counts stay nearest-cell; every footprint carries V times the product
    of the two kernel rows' sums (the table row chosen by the fractional
    position, 1023 rows)."""
    N = 64
    u0, v0 = 2.25, 3.6
    val = np.complex64(2.0 + 1.0j)
    vis, n_ant, u, v, flags, freq, _ = _kb_obs([u0], [v0], [val], N)
    for K in (5, 7):
        g, c, _ = oracle.gridding(vis, n_ant, u, v, flags, freq, 1.0, 1.0, -np.inf, N,
                                  precision="f64", kernel_size=K)
        _, c_nn, _ = oracle.gridding(vis, n_ant, u, v, flags, freq, 1.0, 1.0, -np.inf, N,
                                     precision="f64")
        assert np.array_equal(c, c_nn)
        t = oracle.kb_kernels(K, True)
        rx = (int(np.round((u0 - round(u0)) * 1023)) + 1534) % 1023
        ry = (int(np.round((v0 - round(v0)) * 1023)) + 1534) % 1023
        s = t[rx].sum() * t[ry].sum()
        # total = V s + conj(V) s
        assert abs(g[0, 0].sum() - 2 * val.real * s) < 1e-9
