"""Seeded synthetic observations for the imaging path.

Each workload is an array layout, channel frequencies, timesteps and
visibilities made of point sources plus complex Gaussian noise:

  V_XX(bl, ch) = sum_s S_s exp(-2 pi i (u l_s + v m_s) / lambda_ch) + noise,
  V_YY = 0.9 V_XX + noise, V_XY = V_YX = noise,

with zenith u, v, w = -(x_a1 - x_a2) metres stored as FP32 (the convention
of CAntennaPositions::CalculateUVW, src/antenna_positions.cpp:81-89).
Randomness comes from numpy's PCG64 generator with a fixed seed per workload,
so every run, on a CPU or a GPU host, sees identical inputs.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0


@dataclass
class Observation:
    name: str
    n_ant: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    u: np.ndarray            # float32 [n_bl] metres
    v: np.ndarray
    w: np.ndarray
    flags: np.ndarray        # int32 [n_bl]
    cable: np.ndarray        # float64 [n_ant] metres
    freq: np.ndarray         # float64 [n_ch] Hz
    vis: np.ndarray          # complex64 [n_t][n_ch][n_bl][4]
    n_pixels: int
    oversampling: float = 2.0
    min_uv: float = -np.inf
    u_max: float = 0.0
    v_max: float = 0.0
    delta_u: float = 0.0
    delta_v: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def n_bl(self) -> int:
        return self.n_ant * (self.n_ant + 1) // 2

    @property
    def n_ch(self) -> int:
        return self.vis.shape[1]

    @property
    def n_t(self) -> int:
        return self.vis.shape[0]

    @property
    def n_images(self) -> int:
        return self.n_ch * self.n_t


def baseline_antennas(n_ant: int):
    """(a1, a2) of every baseline bl = a1(a1+1)/2 + a2, a2 <= a1."""
    a1 = np.repeat(np.arange(n_ant), np.arange(1, n_ant + 1))
    a2 = np.concatenate([np.arange(i + 1) for i in range(n_ant)])
    return a1, a2


def disk_layout(rng, n_ant: int, diameter_m: float):
    r = 0.5 * diameter_m * np.sqrt(rng.random(n_ant))
    th = 2 * np.pi * rng.random(n_ant)
    return r * np.cos(th), r * np.sin(th)


def gaussian_layout(rng, n_ant: int, sigma_m: float, clip_m: float):
    x = np.empty(0)
    y = np.empty(0)
    while len(x) < n_ant:
        px = rng.normal(0, sigma_m, 2 * n_ant)
        py = rng.normal(0, sigma_m, 2 * n_ant)
        keep = px * px + py * py <= clip_m * clip_m
        x = np.concatenate([x, px[keep]])
        y = np.concatenate([y, py[keep]])
    return x[:n_ant], y[:n_ant]


def zenith_uvw(x, y, z):
    a1, a2 = baseline_antennas(len(x))
    u = (-(x[a1] - x[a2])).astype(np.float32)
    v = (-(y[a1] - y[a2])).astype(np.float32)
    w = (-(z[a1] - z[a2])).astype(np.float32)
    return u, v, w


def baseline_flags(n_ant: int, flagged_antennas=()) -> np.ndarray:
    a1, a2 = baseline_antennas(n_ant)
    f = a1 == a2
    for a in flagged_antennas:
        f |= (a1 == a) | (a2 == a)
    return f.astype(np.int32)


def grid_scalars(u, v, flags, n_pixels, oversampling=2.0):
    """u_max, v_max over unflagged baselines and delta_u = osf*u_max/N (metres),
    as CPacerImager::CalculateUVW / grid() (src/pacer_imager.cpp:213-221,497-499)."""
    ok = flags == 0
    u_max = float(np.max(np.abs(u[ok].astype(np.float64)))) if ok.any() else -np.inf
    v_max = float(np.max(np.abs(v[ok].astype(np.float64)))) if ok.any() else -np.inf
    osf = float(np.float32(oversampling))
    return u_max, v_max, osf * u_max / n_pixels, osf * v_max / n_pixels


def point_sources(rng, n_src: int, l2m2_max: float = 0.8, flux=(1.0, 10.0)):
    ls, ms = [], []
    while len(ls) < n_src:
        l, m = rng.uniform(-1, 1, 2)
        if l * l + m * m < l2m2_max:
            ls.append(l)
            ms.append(m)
    s = rng.uniform(flux[0], flux[1], n_src)
    return np.array(ls), np.array(ms), s


def sky_visibilities(u, v, freq, l, m, s):
    """XX model visibilities [n_ch][n_bl] complex128 of point sources."""
    out = np.zeros((len(freq), len(u)), np.complex128)
    u64 = u.astype(np.float64)
    v64 = v.astype(np.float64)
    for k, f in enumerate(freq):
        lam = C_LIGHT / f
        ph = (-2j * np.pi / lam) * (np.outer(u64, l) + np.outer(v64, m))
        out[k] = np.exp(ph) @ s
    return out


def diffuse_visibilities(rng, u, v, freq, amplitude: float, index: float = 2.0,
                         uv_ref: float = 10.0):
    """XX visibilities [n_ch][n_bl] complex128 of a diffuse Gaussian random
    field with angular power spectrum P(q) ~ (q / uv_ref)^-index, q = |uv| in
    wavelengths: each (baseline, channel) draws a complex normal of variance
    amplitude^2 P(q), Hermitian by construction because only u, v of the
    baseline itself are used (the conjugate cell gets the conjugate value)."""
    u64 = u.astype(np.float64)
    v64 = v.astype(np.float64)
    out = np.empty((len(freq), len(u)), np.complex128)
    for k, f in enumerate(freq):
        lam = C_LIGHT / f
        q = np.hypot(u64, v64) / lam
        p = (np.maximum(q, 0.5) / uv_ref) ** (-index)
        z = rng.standard_normal((len(u), 2))
        out[k] = amplitude * np.sqrt(p / 2) * (z[:, 0] + 1j * z[:, 1])
    return out


def make_vis(rng, model_xx, n_t: int, sigma: float) -> np.ndarray:
    n_ch, n_bl = model_xx.shape
    vis = np.empty((n_t, n_ch, n_bl, 4), np.complex64)
    for t in range(n_t):
        noise = rng.standard_normal((n_ch, n_bl, 4, 2), dtype=np.float32) * np.float32(sigma)
        vis[t] = noise[..., 0] + 1j * noise[..., 1]
        vis[t, :, :, 0] += model_xx.astype(np.complex64)
        vis[t, :, :, 3] += (0.9 * model_xx).astype(np.complex64)
    return vis


def observation(name: str, n_ant: int, layout: str, n_ch: int, n_t: int, freq, n_pixels: int,
                n_src: int, sigma: float, seed: int, flagged_antennas=(), height_sigma_m=0.5,
                cable_range=(50.0, 300.0), diameter_m=35.0, gauss_sigma_m=300.0,
                clip_m=1500.0, bl_flag_fraction=0.0, diffuse=0.0) -> Observation:
    rng = np.random.Generator(np.random.PCG64(seed))
    if layout == "disk":
        x, y = disk_layout(rng, n_ant, diameter_m)
    else:
        x, y = gaussian_layout(rng, n_ant, gauss_sigma_m, clip_m)
    z = rng.normal(0, height_sigma_m, n_ant)
    u, v, w = zenith_uvw(x, y, z)
    flags = baseline_flags(n_ant, flagged_antennas)
    if bl_flag_fraction > 0:
        extra = rng.random(len(flags)) < bl_flag_fraction
        flags = (flags | extra).astype(np.int32)
    cable = rng.uniform(cable_range[0], cable_range[1], n_ant)
    freq = np.asarray(freq, np.float64)
    assert len(freq) == n_ch
    l, m, s = point_sources(rng, n_src)
    model = sky_visibilities(u, v, freq, l, m, s)
    if diffuse > 0:
        model += diffuse_visibilities(rng, u, v, freq, diffuse)
    vis = make_vis(rng, model, n_t, sigma)
    u_max, v_max, du, dv = grid_scalars(u, v, flags, n_pixels)
    return Observation(name=name, n_ant=n_ant, x=x, y=y, z=z, u=u, v=v, w=w, flags=flags,
                       cable=cable, freq=freq, vis=vis, n_pixels=n_pixels, u_max=u_max,
                       v_max=v_max, delta_u=du, delta_v=dv,
                       meta=dict(n_src=n_src, sigma=sigma, seed=seed, layout=layout))


def eda2_station(n_ch=1, n_t=1, n_pixels=180, seed=1001, freq=None, **kw) -> Observation:
    """256 antennas uniform in a 35 m disk, 159.375 MHz, 5 sources, sigma 0.1."""
    f = [159.375e6] * n_ch if freq is None else freq
    return observation("eda2_256ant", 256, "disk", n_ch, n_t, f, n_pixels, 5, 0.1, seed, **kw)


def mwa_array(n_ch=64, n_t=10, n_pixels=1024, seed=1002, freq=None, n_ant=128, n_src=20,
              **kw) -> Observation:
    """128 tiles, Gaussian layout (sigma 300 m, clipped at 1.5 km), channels
    170-200 MHz evenly spaced, 20 sources, sigma 0.5."""
    if freq is None:
        freq = (170.0 + 30.0 * np.arange(n_ch) / max(n_ch - 1, 1)) * 1e6
    return observation("mwa_128t", n_ant, "gauss", n_ch, n_t, freq, n_pixels, n_src, 0.5, seed,
                       **kw)


def mwa_fine_channels(n_ch=768, first_mhz=169.60, step_mhz=0.04):
    """768 x 40 kHz fine channels from 169.6 MHz (the MWA metafits used by the
    reference's tests, scripts/tests/phase_corrections/20200619163000.metafits)."""
    return (first_mhz + step_mhz * np.arange(n_ch)) * 1e6


def small(n_ant=16, n_ch=3, n_t=2, n_pixels=64, seed=7, layout="disk", diameter_m=35.0,
          n_src=3, sigma=0.1, freq=None, **kw) -> Observation:
    """A small random observation for unit tests."""
    if freq is None:
        freq = np.linspace(150e6, 200e6, n_ch)
    return observation(f"small_{n_ant}ant", n_ant, layout, n_ch, n_t, freq, n_pixels, n_src, sigma,
                       seed, diameter_m=diameter_m, **kw)


def mixed_batch(n_images=512, seed=1003, sizes=(128, 180, 256, 512, 1024, 2048),
                max_flag_fraction=0.3):
    """Uneven batch: every image from either the 128-tile Gaussian array or the
    256-antenna 35 m station, a grid side drawn from `sizes`, one fine channel
    at random, and its own random baseline flags (0 .. max_flag_fraction of the
    baselines on top of the autos), so the gridded visibility counts differ
    from image to image.  Returns one single-image Observation per image."""
    rng = np.random.Generator(np.random.PCG64(seed))
    arrays = []
    for n_ant, layout, n_src, sigma in ((128, "gauss", 20, 0.5), (256, "disk", 5, 0.1)):
        if layout == "disk":
            x, y = disk_layout(rng, n_ant, 35.0)
        else:
            x, y = gaussian_layout(rng, n_ant, 300.0, 1500.0)
        z = rng.normal(0, 0.5, n_ant)
        u, v, w = zenith_uvw(x, y, z)
        cable = rng.uniform(50.0, 300.0, n_ant)
        src = point_sources(rng, n_src)
        arrays.append((n_ant, layout, x, y, z, u, v, w, cable, src, sigma))
    chans = mwa_fine_channels(768)
    out = []
    for i in range(n_images):
        n_ant, layout, x, y, z, u, v, w, cable, (l, m, s), sigma = arrays[int(rng.integers(2))]
        n = int(sizes[int(rng.integers(len(sizes)))])
        frac = float(rng.uniform(0.0, max_flag_fraction))
        flags = (baseline_flags(n_ant) | (rng.random(len(u)) < frac)).astype(np.int32)
        freq = np.array([chans[int(rng.integers(len(chans)))]])
        model = sky_visibilities(u, v, freq, l, m, s)
        vis = make_vis(rng, model, 1, sigma)
        u_max, v_max, du, dv = grid_scalars(u, v, flags, n)
        out.append(Observation(name=f"mixed_{i}", n_ant=n_ant, x=x, y=y, z=z, u=u, v=v, w=w,
                               flags=flags, cable=cable, freq=freq, vis=vis, n_pixels=n,
                               u_max=u_max, v_max=v_max, delta_u=du, delta_v=dv,
                               meta=dict(layout=layout, flag_fraction=frac, seed=seed)))
    return out
