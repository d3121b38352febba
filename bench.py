#!/usr/bin/env python3
"""Benchmark of the visibility -> dirty-image path on one MI355X.

A step images one batch of synthetic visibilities, already resident in HBM:
phase corrections (geometric + cable), nearest-neighbour gridding, 2-D
backward FFT, fft_shift and normalisation, for every (channel, timestep) of
the workload, through blink_image_batch (replayed hipGraph).  Prints ONE JSON
line (the last line of stdout).

Default workload: 128-tile MWA-like array (8256 baselines), 64 channels
170-200 MHz x 10 timesteps = 640 images of 1024^2.  Other workloads:
  mwa128_768ch_1024    the same array, 768 fine channels x 1 timestep, 1024^2
  eda2_1ch_180         256-antenna station, 1 channel, one 180^2 image
  eda2_768ch_180/_256  the station over 768 fine channels (also --transform dft);
                       _128 and _512 the same at those sides
  mixed512             512 single-image sets: 128- or 256-antenna array, grid side
                       drawn from {128, 180, 256, 512, 1024, 2048}, own random flags
                       (0-30 %); one call per step, plus per-size throughput
  mwa128_768ch4t_4096  128 tiles, 768 channels x 4 timesteps of 4096^2 (50 sources +
                       diffuse field); a step streams the 4 timesteps through one
                       768-image (96 GiB) output buffer, one call each

Also measured, outside the timed region:
  * per-kernel GPU time (HIP events recorded by the library around each launch,
    on the stream it launches on) over K steps with the same kernels launched
    directly -> roofline of the dominant kernel;
  * the plain FP32 GPU path: one image at a time through the stage entry points
    (corrections, atomic gridding, FFT, shift+norm), on a sample of images;
  * parity of sample images against the CPU oracle (reference FP32 semantics)
    and the FP64 oracle;
  * the CPU baseline: the oracle's CPU/OpenMP restatement of the reference
    pipeline on a bounded sample of the same workload, on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "imaged channels/sec, 128-ant 1024² grid; % HBM roofline vs plain FP32 GPU path"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOADS = ("mwa128_64ch10t_1024", "mwa128_768ch_1024", "eda2_1ch_180", "eda2_768ch_128",
             "eda2_768ch_180", "eda2_768ch_256", "eda2_768ch_512", "mixed512",
             "mwa128_768ch4t_4096")


def make_workload(name: str):
    from imager_amd import synth
    if name == "mixed512":
        return synth.mixed_batch(n_images=512, seed=1003)
    if name == "mwa128_768ch4t_4096":
        return synth.mwa_array(n_ch=768, n_t=4, n_pixels=4096, seed=1005, n_src=50,
                               freq=synth.mwa_fine_channels(768), height_sigma_m=0.5,
                               diffuse=2.0)
    if name == "mwa128_64ch10t_1024":
        return synth.mwa_array(n_ch=64, n_t=10, n_pixels=1024, seed=1002, height_sigma_m=0.5)
    if name == "mwa128_768ch_1024":
        return synth.mwa_array(n_ch=768, n_t=1, n_pixels=1024, seed=1006,
                               freq=synth.mwa_fine_channels(768), height_sigma_m=0.5)
    if name == "eda2_1ch_180":
        return synth.eda2_station(n_ch=1, n_t=1, n_pixels=180, seed=1001)
    if name in ("eda2_768ch_128", "eda2_768ch_180", "eda2_768ch_256", "eda2_768ch_512"):
        n = int(name.rsplit("_", 1)[1])
        return synth.eda2_station(n_ch=768, n_t=1, n_pixels=n, seed=1004,
                                  freq=synth.mwa_fine_channels(768))
    raise SystemExit(f"unknown workload {name}")


def algorithmic_bytes(kernel: str, n_bl: int, n_pix: int, n_img: int) -> float | None:
    """Per-launch bytes of the unfused pipeline's stages (the figure the
    gridding + FFT roofline target is quoted in): gridding = 8 n_bl (one pol)
    + 8 n_bl (u, v) + 12 N^2 (grid + count write), 2-D FFT = 16 N^2 (read
    grid + write image), shift and normalisation fused = 0.  The fused
    kernels move fewer bytes than this (the grid never reaches HBM), so it is
    reported as an effective bandwidth, not as a roofline fraction."""
    g = 16.0 * n_bl + 12.0 * n_pix * n_pix
    f = 16.0 * n_pix * n_pix
    if kernel in ("grid_rows_fft", "grid_tiles"):
        return g * n_img
    if kernel == "grid_half_rows":   # DFT transform: half plane, complex only
        return (16.0 * n_bl + 8.0 * (n_pix // 2 + 1) * n_pix) * n_img
    if kernel == "fft_cols":
        return f * n_img
    return None


def plan_stats(o, transposed=True):
    """Per channel of observation o: the plan entries in plan rows 0 .. N/2
    (each gridded baseline's cell and conjugate cell that fall there) and the
    number of those rows that hold any entry.  The FFT sides' plan is
    transposed (plan rows = grid columns u), the DFT sides' is not."""
    import oracle
    n = o.n_pixels
    cells = oracle.cells(o.u, o.v, o.flags, o.freq, o.delta_u, o.delta_v, o.min_uv, n)
    k = 0 if transposed else 1
    rows = np.concatenate([cells[:, :, k], cells[:, :, k + 2]], axis=1)
    half = (rows >= 0) & (rows <= n // 2)
    entries = half.sum(1).astype(np.float64)
    nonempty = np.array([len(np.unique(rows[c][half[c]])) for c in range(rows.shape[0])],
                        np.float64)
    return entries, nonempty


def compacted(n_bl: int, n_pix: int) -> bool:
    """Whether the batched call gathers compacted visibilities for a set
    (register sides up to 2048 with at most 64 baselines per plan row)."""
    return (n_pix in (128, 256, 512, 1024, 2048) and n_bl <= 64 * (n_pix // 2 + 1)
            and os.environ.get("BLINK_COMPACT", "1") != "0")


# the grid side whose direct row pass holds a window of the half plane
DIRECT_WINDOW_SIDE = 256


def physical_bytes(kernel: str, o, stats, n_img: int, transform: str = "fft",
                   kernel_size: int = 0, direct: bool = False) -> float | None:
    """HBM bytes a fused kernel must move per launch over n_img images of
    observation o (images cycle over its channels), given the plan: what it
    has to read and write, with every byte counted once.

      grid_rows_fft  plan entries (4 B each) + the visibilities once (8 B per
                     baseline compacted, else the 32 B records) + the
                     transformed non-empty plan rows (8 N B each); on the
                     direct path (no plan): u, v, flags once (12 B per
                     baseline) + the records once + the plan rows up to the
                     last that holds cells (the rows the second pass reads)
      direct_prep    w and the phase coefficient per baseline, the count sums
                     (at 256^2 also u, v, flags for the |u|, |v| maxima)
      fft_cols       the non-empty plan rows (8 N B each) + the image (8 N^2 B)
      vis_compact    the records (32 B) + the compacted value (8 B) per baseline
      plan           per channel: u, v, flags (12 B) read, the packed cell (4 B)
                     written and read back per baseline, the entries written
      grid_half_rows plan entries + records + the half plane (8 (N/2+1) N B)
      dft_rows       the half plane + the stage-1 product (8 N pad16(N/2+1) B)
      dft_cols       the stage-1 product + the image"""
    n, n_bl = o.n_pixels, o.n_bl
    entries, nonempty = stats
    window = direct and n == DIRECT_WINDOW_SIDE
    if kernel_size and not direct:
        # convolutional gridding writes and reads every plan row 0 .. N/2
        nonempty = np.full_like(nonempty, n // 2 + 1)
    elif kernel_size:
        # on the direct row pass, the rows up to the last cell's plus K/2
        nonempty = np.minimum(nonempty + kernel_size // 2, n // 2 + 1)
    ch = np.arange(n_img) % o.n_ch
    e_sum = float(entries[ch].sum())
    r_sum = float(nonempty[ch].sum())
    h1 = n // 2 + 1
    pad = -(-h1 // 16) * 16
    if kernel == "grid_rows_fft" and direct:
        return 12.0 * n_bl + 32.0 * n_bl * n_img + 8.0 * n * r_sum
    if kernel == "direct_prep":
        return (24.0 if window else 12.0) * n_bl + 4.0 * n_img
    if kernel == "grid_rows_fft":
        vis_b = 8.0 if compacted(n_bl, n) else 32.0
        return 4.0 * e_sum + vis_b * n_bl * n_img + 8.0 * n * r_sum
    if kernel == "grid_rows_kb" and direct:   # u, v, flags and records once, the rows
        return 12.0 * n_bl + 32.0 * n_bl * n_img + 8.0 * n * r_sum
    if kernel == "grid_rows_kb":   # entries and records once (a lower bound), every row
        return 4.0 * e_sum + 32.0 * n_bl * n_img + 8.0 * n * r_sum
    if kernel == "fft_cols":
        return 8.0 * n * r_sum + 8.0 * n * n * n_img
    if kernel == "vis_compact":
        return 40.0 * n_bl * n_img
    if kernel == "plan":
        return 20.0 * n_bl * o.n_ch + 4.0 * float(entries.sum())
    if kernel == "grid_half_rows":
        return 4.0 * e_sum + 32.0 * n_bl * n_img + 8.0 * h1 * n * n_img
    if kernel == "dft_rows":
        return (8.0 * h1 * n + 8.0 * n * pad) * n_img
    if kernel == "dft_cols":
        return (8.0 * n * pad + 8.0 * n * n) * n_img
    return None


def dft_flops(kernel: str, n_pix: int, n_img: int):
    """(useful, executed) flops of ONE launch of a DFT-as-GEMM stage over the
    Hermitian half plane (H = N/2):
      dft_rows: complex (H+1) x N rows times the N x N matrix per image,
                useful 8 (H+1) N^2, executed 12 split real products;
      dft_cols: real part of N rows times the (H+1) x N matrix per image,
                useful 4 (H+1) N^2, executed 6 split real products;
    executed flops count every MFMA over the padded tile grid (m in tiles of
    256, n of 64, k of 16)."""
    h1 = n_pix // 2 + 1
    n_pad = -(-n_pix // 64) * 64
    if kernel == "dft_rows":
        m, k, terms, useful = n_img * h1, n_pix, 12, 8.0 * h1 * n_pix * n_pix * n_img
    else:
        m, k, terms, useful = n_img * n_pix, h1, 6, 4.0 * h1 * n_pix * n_pix * n_img
    m_pad = -(-m // 256) * 256
    k_pad = -(-k // 16) * 16
    executed = terms * 2.0 * m_pad * n_pad * k_pad
    return useful, executed


def read_traffic(kernel: str, workload: str):
    """HBM bytes per image measured with rocprofv3 PMC counters (the recorded
    launch's bytes over its images), when a summary for this kernel and
    workload has been recorded under profiles/."""
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        return float(d["hbm_bytes_per_launch"]) / float(d["images_per_launch"])
    except Exception:
        return None


def cpu_baseline(obs, budget_s: float, max_images: int = 128, reps: int = 5):
    """The reference CPU pipeline (the oracle's restatement, reference FP32
    semantics, OpenMP over images as src/pacer_imager.cpp:158,263) timed
    stage by stage on this host's cores, on the first images of the workload:
    corrections (geometric + cable), gridding, the 2-D FFT, and normalisation
    + fft_shift, all in C; the FFT stage also through torch's CPU FFT (MKL)
    with as many threads.  Each stage: one warm-up, then the median of `reps`
    runs.  The value charges the faster FFT, as an FFTW build of the
    reference would run (FFTW itself is not installed)."""
    import torch
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    torch.set_num_threads(threads)

    def median_s(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    def stages(n_t, n_ch):
        vis, freq = obs.vis[:n_t, :n_ch], obs.freq[:n_ch]
        k = n_t * n_ch

        def corrections():
            v = oracle.geometric_corrections(vis, obs.n_ant, obs.w, freq)
            return oracle.cable_corrections(v, obs.n_ant, obs.cable, freq)

        vis_c = corrections()
        grid = lambda: oracle.gridding(vis_c, obs.n_ant, obs.u, obs.v, obs.flags, freq,
                                       obs.delta_u, obs.delta_v, obs.min_uv, obs.n_pixels, "XX")
        g, c, _ = grid()
        t = {"corrections": median_s(corrections), "gridding": median_s(grid),
             "fft_c": median_s(lambda: oracle.fft2d(g, +1))}
        gt = torch.from_numpy(g)
        # unnormalised backward transform, as FFTW_BACKWARD
        t["fft_mkl"] = median_s(lambda: torch.fft.ifft2(gt, norm="forward"))
        img = oracle.fft2d(g, +1)
        t["shift_norm"] = median_s(lambda: oracle.norm_shift(img, c))
        return k, t

    # one image first, to size the sample to the budget
    _, t1 = stages(1, 1)
    per_img = (1 + reps) * sum(t1.values())
    want = int(min(max_images, max(1, budget_s / max(per_img, 1e-4))))
    n_ch = min(obs.n_ch, want)
    n_t = min(obs.n_t, max(1, want // n_ch))
    t0 = time.perf_counter()
    k, t = stages(n_t, n_ch)
    wall = time.perf_counter() - t0
    fft_backend = "MKL (torch CPU FFT)" if t["fft_mkl"] < t["fft_c"] else "oracle C FFT"
    total = t["corrections"] + t["gridding"] + min(t["fft_c"], t["fft_mkl"]) + t["shift_norm"]
    rate = {name: k / sec for name, sec in t.items()}
    return {"value": k / total, "unit": "images/s", "cores": threads, "kind": "port",
            "fft_backend": fft_backend,
            "sample": f"{n_t} timestep(s) x {n_ch} channel(s) = {k} images of the workload "
                      f"({obs.n_ant} antennas, {obs.n_pixels}^2), every stage 1 + {reps} runs "
                      f"on {threads} threads ({wall:.1f} s): corrections, gridding, norm+shift "
                      f"in oracle/blink_oracle.c; FFT charged with {fft_backend}",
            "stages_images_per_s": {
                "corrections": rate["corrections"], "gridding": rate["gridding"],
                "fft_oracle_c": rate["fft_c"], "fft_mkl": rate["fft_mkl"],
                "shift_norm": rate["shift_norm"]},
            "value_with_oracle_c_fft": k / (total - min(t["fft_c"], t["fft_mkl"]) + t["fft_c"])}


def time_run_batched(obs, steps: int, warmup: int):
    """CPacerImagerMI355X::run_batched() (the C++ class a reference caller
    binds) timed by tests/cpp/bin/bench_run_batched in a child process on an
    array of the same shape: host flags / UVW / frequencies and their upload,
    a new device image buffer per call, the captured graph updated in place."""
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "bench_run_batched")
    if not os.path.exists(exe):
        return {"skipped": "tests/cpp/bin/bench_run_batched is not built"}
    cmd = [exe, str(obs.n_ant), str(obs.n_ch), str(obs.n_t), str(obs.n_pixels), str(steps),
           str(max(warmup, 1)), "0"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=180)
    if r.returncode != 0:
        return {"failed": r.returncode, "stderr": r.stderr[-400:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    res["program"] = "tests/cpp/bin/bench_run_batched " + " ".join(cmd[1:])
    return res


def memory_reference(mib: int = 2048):
    """Write-only (fill) and copy bandwidth of this GPU with torch kernels, as a
    practical reference next to the 8 TB/s peak used for the roofline."""
    import torch
    n = (mib << 20) // 4
    x = torch.empty(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    y.copy_(x)
    torch.cuda.synchronize()
    res = {}
    for name, fn, nbytes in (("fill_gbs", lambda: x.fill_(2.0), 4 * n),
                             ("copy_gbs", lambda: y.copy_(x), 8 * n)):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(5):
            fn()
        b.record()
        torch.cuda.synchronize()
        res[name] = 5 * nbytes / (a.elapsed_time(b) * 1e-3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    return res


def build_job(name, wl, transform, kernel_size=0):
    """Device buffers and the list of blink_image_batch calls of one step."""
    import torch
    from imager_amd import capi
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    corr = capi.CORR_GEOMETRIC | capi.CORR_CABLE

    def upload(o):
        return dict(vis=dev(o.vis), u=dev(o.u), v=dev(o.v), w=dev(o.w), flags=dev(o.flags),
                    freq=dev(o.freq), cable=dev(o.cable))

    if name == "mixed512":
        shared = {}   # one device copy of u, v, w, cable per array
        sets, per_size, parity, keep = [], {}, [], []
        for o in wl:
            key = id(o.u)
            if key not in shared:
                shared[key] = dict(u=dev(o.u), v=dev(o.v), w=dev(o.w), cable=dev(o.cable))
            d = dict(shared[key], vis=dev(o.vis), flags=dev(o.flags), freq=dev(o.freq))
            img = torch.empty((1, 1, o.n_pixels, o.n_pixels), dtype=torch.complex64,
                              device="cuda")
            st = capi.make_set(d["vis"], o.n_ant, 1, 1, d["u"], d["v"], d["w"], d["flags"],
                               d["freq"], d["cable"], corr, o.delta_u, o.delta_v, o.min_uv,
                               o.n_pixels, "XX", img, transform=transform,
                               kernel_size=kernel_size)
            keep.append((d, img))
            sets.append(st)
            per_size.setdefault(o.n_pixels, []).append(st)
        seen = {}
        for i, o in enumerate(wl):     # two parity images per grid side
            if seen.get(o.n_pixels, 0) < 2:
                seen[o.n_pixels] = seen.get(o.n_pixels, 0) + 1
                parity.append((o, 0, 0, keep[i][1][0, 0], f"image{i}_{o.n_ant}ant_{o.n_pixels}"))
        counts = {n: len(v) for n, v in sorted(per_size.items())}
        return {"obs": wl[0], "n_pixels": None, "calls": [sets], "images_per_step": len(wl),
                "pipeline_obs": wl, "pipeline_repeat": 1, "per_size_calls": per_size,
                "parity": parity, "plain_ok": False, "keep": keep, "sizes": sorted(per_size),
                "dev": None,
                "data": "synthetic (seeded point sources + noise, per-image random flags)",
                "config": {"n_sets": len(wl), "images_per_side": counts,
                           "n_ant": "128 or 256", "flagging": "0-30 % per image"},
                "describe": f"mixed512: {len(wl)} single-image sets, per side {counts}"}

    o = wl
    d = upload(o)
    N = o.n_pixels
    sums = torch.empty(o.n_images, dtype=torch.float64, device="cuda")
    if name == "mwa128_768ch4t_4096":
        # 4 timesteps streamed through one 768-image output buffer, one call each
        images = torch.empty((1, o.n_ch, N, N), dtype=torch.complex64, device="cuda")
        n_bl = o.n_bl
        calls = []
        for t in range(o.n_t):
            vis_t = d["vis"][t:t + 1]
            calls.append([capi.make_set(vis_t, o.n_ant, o.n_ch, 1, d["u"], d["v"], d["w"],
                                        d["flags"], d["freq"], d["cable"], corr, o.delta_u,
                                        o.delta_v, o.min_uv, N, "XX", images,
                                        transform=transform, kernel_size=kernel_size)])
        t_last = o.n_t - 1
        # the buffer holds the last timestep's images after a step
        parity = [(o, t_last, int(c), images[0, int(c)], f"t{t_last}_ch{int(c)}")
                  for c in np.linspace(0, o.n_ch - 1, 8).round().astype(int)]
        return {"obs": o, "n_pixels": N, "calls": calls, "images_per_step": o.n_images,
                "pipeline_obs": [o], "pipeline_repeat": o.n_images, "parity": parity,
                "plain_ok": True, "keep": (d, images, sums), "sizes": [N], "dev": d,
                "data": "synthetic (seeded 50 point sources + diffuse Gaussian random field "
                        "+ noise)",
                "config": {"n_ant": o.n_ant, "n_baselines": n_bl, "n_ch": o.n_ch, "n_t": o.n_t,
                           "n_pixels": N, "calls_per_step": len(calls),
                           "output_buffer_images": o.n_ch},
                "describe": f"{name}: {o.n_ant} ant, {n_bl} baselines, {o.n_ch} ch x {o.n_t} t "
                            f"of {N}^2 in {len(calls)} calls per step"}
    images = torch.empty((o.n_t, o.n_ch, N, N), dtype=torch.complex64, device="cuda")
    st = capi.make_set(d["vis"], o.n_ant, o.n_ch, o.n_t, d["u"], d["v"], d["w"], d["flags"],
                       d["freq"], d["cable"], corr, o.delta_u, o.delta_v, o.min_uv, N, "XX",
                       images, sums, transform=transform, kernel_size=kernel_size)
    parity = []
    for i in np.unique(np.linspace(0, o.n_images - 1, 8).round().astype(int)):
        t, c = divmod(int(i), o.n_ch)
        parity.append((o, t, c, images[t, c], f"t{t}_ch{c}"))
    return {"obs": o, "n_pixels": N, "calls": [[st]], "images_per_step": o.n_images,
            "pipeline_obs": [o], "pipeline_repeat": o.n_images, "parity": parity,
            "plain_ok": True, "keep": (d, images, sums), "sizes": [N], "dev": d,
            "data": "synthetic (seeded point sources + noise)",
            "config": {"n_ant": o.n_ant, "n_baselines": o.n_bl, "n_ch": o.n_ch, "n_t": o.n_t,
                       "n_pixels": N},
            "describe": f"{name}: {o.n_ant} ant, {o.n_bl} baselines, {o.n_ch} ch x {o.n_t} t = "
                        f"{o.n_images} images of {N}^2"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mwa128_64ch10t_1024", choices=WORKLOADS)
    ap.add_argument("--transform", choices=["fft", "dft", "dft-fp16"], default="fft",
                    help="2-D transform of the batched path: in-house FP32 FFT, or the "
                         "split-precision DFT-as-GEMM on MFMA (N <= 512) with bf16 (dft) "
                         "or fp16 (dft-fp16) hi+lo operands")
    ap.add_argument("--kernel-size", type=int, default=0,
                    help="convolutional (Kaiser-Bessel with sinc) gridding with K x K footprints "
                         "instead of nearest-cell gridding (0)")
    ap.add_argument("--batch-tiles", action="append", default=[], metavar="N:ROWS:COLS",
                    help="pin the register kernels' row tile and column block for side N "
                         "(blink_set_batch_tiles) instead of autotuning; repeatable")
    ap.add_argument("--input-copies", type=int, default=1,
                    help="device copies of the visibilities used in turn by successive steps")
    ap.add_argument("--plain-images", type=int, default=64,
                    help="images run through the plain one-image-at-a-time GPU path")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-run-batched", action="store_true",
                    help="skip timing CPacerImagerMI355X::run_batched (C++ host class)")
    ap.add_argument("--profile-only", action="store_true",
                    help="only warmup + timed steps (for rocprofv3 runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: every image is independent, so images shard over "
              f"GPUs with no exchange, and only the one-GPU path is implemented; --gpus "
              f"{args.gpus} (WORLD_SIZE={world}) is not supported.",
              file=sys.stderr, flush=True)
        sys.exit(2)

    import torch
    import oracle
    from imager_amd import capi

    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "launches": 3}), flush=True)
        return

    t_gen = time.perf_counter()
    wl = make_workload(args.workload)
    log(f"workload {args.workload} generated in {time.perf_counter() - t_gen:.1f} s")
    ctx = capi.Context(0)
    for spec in args.batch_tiles:
        n_t, r_t, c_t = (int(x) for x in spec.split(":"))
        ctx.set_batch_tiles(n_t, r_t, c_t)
    transform = {"fft": capi.TRANSFORM_FFT, "dft": capi.TRANSFORM_DFT_BF16X3,
                 "dft-fp16": capi.TRANSFORM_DFT_FP16X3}[args.transform]
    job = build_job(args.workload, wl, transform, args.kernel_size)
    obs = job["obs"]            # the observation the plain path and the CPU baseline use
    N = job["n_pixels"]         # None for mixed sizes
    stream = torch.cuda.current_stream()
    log(job["describe"])

    # --input-copies K: K device copies of the visibilities, used in turn by
    # successive steps, so that a step cannot find its inputs in the 256 MiB
    # Infinity Cache from the step before (the 1024^2 workloads hold 169-203 MB)
    rotation = [job["calls"]]
    if args.input_copies > 1 and job.get("dev") is not None and len(job["calls"]) == 1:
        (st0,), = job["calls"]
        o = job["obs"]
        for _ in range(args.input_copies - 1):
            vis_k = job["dev"]["vis"].clone()
            job.setdefault("extra_vis", []).append(vis_k)
            st_k = capi.ImageSet.from_buffer_copy(st0)
            st_k.vis = vis_k.data_ptr()
            rotation.append([[st_k]])
    step_no = [0]

    def step():
        for call in rotation[step_no[0] % len(rotation)]:
            ctx.image_batch(call, stream)
        step_no[0] += 1

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = job["images_per_step"] * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.1f} images/s")
    # the parity sample is read back from what the last timed replay wrote
    parity_images = None if args.no_parity else [
        (o, t, ch, view.cpu().numpy(), label) for (o, t, ch, view, label) in job["parity"]]
    if args.profile_only:
        print(json.dumps({"metric": METRIC, "value": value, "unit": "images/s",
                          "ms_per_step": ms_per_step, "profile_only": True}), flush=True)
        ctx.close()
        return

    # per-kernel GPU time, HIP events on the launch stream, same K steps
    ctx.profile_reset()
    ctx.profile_enable(True)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    eager_ms = 1e3 * (time.perf_counter() - t1) / args.steps
    ctx.profile_enable(False)
    prof = ctx.profile()
    kernels = {}
    for name, (n, ms) in prof.items():
        kernels[name] = {"launches_per_step": n / args.steps, "avg_ms": ms / max(n, 1),
                         "ms_per_step": ms / args.steps}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    log("kernels: " + ", ".join(f"{k} {v['ms_per_step']:.3f} ms/step ({v['avg_ms']*1e3:.1f} us x "
                               f"{v['launches_per_step']:.0f})" for k, v in kernels.items()))
    roof = None
    stats = None
    direct = "direct_prep" in prof   # the direct row pass ran (no plan)
    if N is not None:
        stats = plan_stats(obs, transposed=args.transform == "fft")
        n_launch = max(prof.get("grid_rows_fft", (0, 0))[0], prof.get("grid_tiles", (0, 0))[0],
                       prof.get("grid_half_rows", (0, 0))[0],
                       prof.get("grid_rows_kb", (0, 0))[0]) // args.steps
        chunk = -(-job["images_per_step"] // max(n_launch, 1))
        imgs_per_launch = chunk if dom in ("grid_rows_fft", "grid_rows_kb", "fft_cols",
                                           "grid_tiles", "grid_half_rows", "dft_rows",
                                           "dft_cols") \
            else job["images_per_step"]
        avg_s = kernels[dom]["avg_ms"] * 1e-3
        if dom in ("dft_rows", "dft_cols"):
            useful, executed = dft_flops(dom, N, imgs_per_launch)
            ach = executed / avg_s / 1e12
            prec = "fp16" if args.transform == "dft-fp16" else "bf16"
            roof = {"bound": "mfma", "achieved": ach, "peak": 2500.0, "unit": "TFLOP/s",
                    "frac": ach / 2500.0, "traffic": None, "kernel": dom,
                    "images_per_launch": imgs_per_launch, "executed_flops_per_launch": executed,
                    "useful_flops_per_launch": useful,
                    "useful_tflops": useful / avg_s / 1e12,
                    "note": f"dense {prec} MFMA peak; executed = 3 split terms x 4 (dft_rows) "
                            "or 2 (dft_cols) real products over padded tiles"}
        else:
            phys = physical_bytes(dom, obs, stats, imgs_per_launch, args.transform,
                                  args.kernel_size, direct)
            if phys is not None:
                ach = phys / avg_s / 1e9
                per_image = read_traffic(dom, args.workload)
                traffic = per_image * imgs_per_launch if per_image else None
                roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                        "images_per_launch": imgs_per_launch, "bytes_per_launch": phys,
                        "bytes_model": "what the fused kernel must move (plan rows, image)"}
                alg = algorithmic_bytes(dom, obs.n_bl, N, imgs_per_launch)
                if alg is not None:
                    # the unfused stages' byte figure over the same time
                    roof["effective_gbs_unfused_figure"] = alg / avg_s / 1e9
                if traffic:
                    # the bytes the PMC counters saw for this launch, over the same time
                    roof["traffic_gbs"] = traffic / avg_s / 1e9
                    roof["traffic_frac"] = traffic / avg_s / 1e9 / HBM_PEAK_GBS
                    roof["traffic_over_model"] = traffic / phys
    # the DFT-as-GEMM stages against the dense MFMA peak, whichever kernel dominates
    dft_mfma = None
    if N is not None and any(k in kernels for k in ("dft_rows", "dft_cols")):
        dft_mfma = {}
        n_img = -(-job["images_per_step"] // max(kernels.get("dft_rows", {}).get(
            "launches_per_step", 1), 1))
        for k in ("dft_rows", "dft_cols"):
            if k not in kernels:
                continue
            useful, executed = dft_flops(k, N, n_img)
            avg_s = kernels[k]["avg_ms"] * 1e-3
            dft_mfma[k] = {"executed_tflops": executed / avg_s / 1e12,
                           "useful_tflops": useful / avg_s / 1e12,
                           "frac": executed / avg_s / 1e12 / 2500.0, "peak": 2500.0,
                           "images_per_launch": n_img, "avg_us": avg_s * 1e6}
    # whole step: every kernel's bytes (the same model) over the step time, and
    # the unfused gridding + FFT figure as an effective rate
    eff_bytes = sum(16.0 * o.n_bl + 28.0 * o.n_pixels ** 2 for o in job["pipeline_obs"])
    eff_bytes *= job["pipeline_repeat"]
    pipeline = {"effective_gbs_unfused_figure": eff_bytes / (ms_per_step * 1e-3) / 1e9}
    if stats is not None:
        step_bytes = 0.0
        for k, v in kernels.items():
            if k == "plan":   # no plan units on the direct path: an empty launch
                b = 0.0 if direct else \
                    v["launches_per_step"] * physical_bytes(k, obs, stats, obs.n_ch)
            else:
                b = physical_bytes(k, obs, stats, job["images_per_step"], args.transform,
                                   args.kernel_size, direct)
            step_bytes += b or 0.0
        pipeline.update({"bytes_per_step": step_bytes,
                         "achieved_gbs": step_bytes / (ms_per_step * 1e-3) / 1e9,
                         "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS})

    # per-size throughput of a mixed batch: one call per grid side
    per_size = None
    if job.get("per_size_calls"):
        per_size = {}
        for n_pix, call in sorted(job["per_size_calls"].items()):
            ctx.image_batch(call, stream)
            torch.cuda.synchronize()
            tq = time.perf_counter()
            for _ in range(args.steps):
                ctx.image_batch(call, stream)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - tq) / args.steps
            per_size[str(n_pix)] = {"images": len(call), "ms_per_call": 1e3 * dt,
                                    "images_per_s": len(call) / dt}
        log(f"per size: {per_size}")

    # plain path: one image at a time through the stage entry points
    plain = None
    if args.plain_images > 0 and job["plain_ok"] and not args.kernel_size:
        n_s = min(args.plain_images, obs.n_images)
        n_bl = obs.n_bl
        d = job["dev"]
        vis_s = d["vis"].reshape(obs.n_images, n_bl, 4)[:n_s].clone()
        grid = torch.zeros((N, N), dtype=torch.complex64, device="cuda")
        cnt = torch.zeros((N, N), dtype=torch.float32, device="cuda")
        img = torch.empty((N, N), dtype=torch.complex64, device="cuda")

        def plain_image(i):
            ch = i % obs.n_ch
            v1 = vis_s[i]
            f1 = d["freq"][ch:ch + 1]
            ctx.apply_geometric_corrections(v1, obs.n_ant, 1, 1, d["w"], f1, stream)
            ctx.apply_cable_corrections(v1, obs.n_ant, 1, 1, d["cable"], f1, stream)
            ctx.gridding(v1, obs.n_ant, 1, 1, d["u"], d["v"], d["flags"], f1, obs.delta_u,
                         obs.delta_v, obs.min_uv, N, "XX", 0, grid, cnt, algo=capi.GRID_ATOMIC,
                         stream=stream)
            ctx.image(grid, cnt, N, 1, img, zero_grids=True, stream=stream)

        for i in range(min(4, n_s)):
            plain_image(i)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for i in range(n_s):
            plain_image(i)
        torch.cuda.synchronize()
        dtp = time.perf_counter() - tp
        plain = {"value": n_s / dtp, "unit": "images/s", "images": n_s,
                 "path": "per image: geometric + cable corrections, atomic gridding, "
                         "row/column FFT, fused shift+norm (stage entry points)"}
        plain["batched_speedup"] = value / plain["value"]
        log(f"plain path: {plain['value']:.1f} images/s; batched is {plain['batched_speedup']:.1f}x")
        del vis_s, grid, cnt, img

    # the C++ class' run_batched() on the same shape, as a caller uses it
    run_batched = None
    if (not args.no_run_batched and not args.kernel_size and args.transform == "fft"
            and args.workload in ("mwa128_64ch10t_1024", "mwa128_768ch_1024")):
        run_batched = time_run_batched(obs, steps=max(args.steps, 5), warmup=args.warmup)
        log(f"run_batched: {run_batched}")

    # parity of the sample images of the last timed replay against the oracle
    parity = None
    if parity_images is not None:
        e32s, e64s, names = [], [], []
        for (o, t, ch, got, label) in parity_images:
            v1 = o.vis[t:t + 1, ch:ch + 1]
            if args.kernel_size:
                # the oracle's convolutional gridding, reference precision and FP64
                f1 = o.freq[ch:ch + 1]
                vc = oracle.cable_corrections(oracle.geometric_corrections(v1, o.n_ant, o.w, f1),
                                              o.n_ant, o.cable, f1)
                refs = []
                for prec in ("f32", "f64"):
                    g, c, _ = oracle.gridding(vc, o.n_ant, o.u, o.v, o.flags, f1, o.delta_u,
                                              o.delta_v, o.min_uv, o.n_pixels, "XX",
                                              kernel_size=args.kernel_size, precision=prec)
                    refs.append(oracle.image(g, c))
                r32, r64 = refs
            else:
                a = (o.n_ant, o.u, o.v, o.w, o.flags, o.freq[ch:ch + 1], o.cable, True,
                     True, o.delta_u, o.delta_v, o.min_uv, o.n_pixels, "XX")
                r32, _ = oracle.run_f32(v1, *a)
                r64 = oracle.run_f64(v1, *a)
            e32s.append(oracle.peak_normalised_error(got, r32[0, 0]))
            e64s.append(oracle.peak_normalised_error(got, r64[0, 0]))
            names.append(label)
        parity = {"images": names, "from": "the last timed graph replay",
                  "max_err_vs_fp32_oracle": max(e[0] for e in e32s),
                  "max_err_vs_fp64_oracle": max(e[0] for e in e64s),
                  "rms_err_vs_fp64_oracle": max(e[1] for e in e64s),
                  "tolerance": 1e-5}
        log(f"parity: {parity}")

    cpu = None
    if not args.no_cpu and not args.kernel_size:
        cpu = cpu_baseline(obs, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    config = dict(job["config"])
    config["input_copies"] = len(rotation)
    if args.kernel_size:
        config["gridding"] = f"Kaiser-Bessel {args.kernel_size}x{args.kernel_size}"
    config.update({"workload": args.workload, "images_per_step": job["images_per_step"],
                   "pol": "XX", "corrections": "geometric+cable",
                   "path": "blink_image_batch (hipGraph)",
                   "transform": {"fft": "fft fp32", "dft": "dft bf16 hi+lo on MFMA",
                                 "dft-fp16": "dft fp16 hi+lo on MFMA"}[args.transform]})
    out = {
        "metric": METRIC, "value": value, "unit": "images/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fft": "f32", "dft": "bf16 hi+lo, f32 accumulate",
                  "dft-fp16": "fp16 hi+lo, f32 accumulate"}[args.transform],
        "data": job["data"],
        "config": config,
        "roofline": roof,
        "dft_mfma": dft_mfma,
        "pipeline_roofline": pipeline,
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "run_batched": run_batched,
        "kernels": kernels,
        "eager_ms_per_step": eager_ms,
        "batch_tiles": {str(n): ctx.batch_tiles(n) for n in job["sizes"]},
        "hbm_reference": memory_reference(),
        "parity": parity,
    }
    if per_size is not None:
        out["per_size"] = per_size
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
