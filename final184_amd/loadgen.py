"""Drivers for the batched and the serving workloads (used by bench.py).

batched: 64 independent scenes of uneven size (triangle counts lognormal, median 20k,
clipped to [1k, 200k]; grids from {32, 64, 128, 256}^3; square G-buffers from {256, 512,
1024}; FP16 mips) built and traced by one batched call per stage, after the autotuner
has picked the bin size and screen tile for the batch. The same scenes run one call per
scene give the unbatched time beside it.

serving: a Poisson request stream into vgi.Server. Each request names a scene key; keys
follow a Zipf popularity over a population of scenes whose geometry comes from a pool of
seeded triangle soups (triangle counts lognormal, median 50k) placed in per-key grids
(128^3 or 256^3, per-key origin shift, so every key has its own volume). A request asks
for 1-8 frames at 512x512 .. 1920x1080, each a ray-cast G-buffer of its soup seen from a
seeded camera pose (9 diffuse cones + a roughness-driven specular cone); 10% of
requests re-voxelize (dynamic geometry). Latency is arrival (the scheduled Poisson time)
to completion of the request's last frame.
"""
from __future__ import annotations

import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import scenegen as sg


def _percentile(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


# ---------------------------------------------------------------------------
# batched
# ---------------------------------------------------------------------------

class _BatchScene:
    def __init__(self, ctx, vgi, w):
        self.w = w
        self.mesh = vgi.MeshData(ctx, w.positions.reshape(-1, 3), albedo=w.albedo)
        self.grid = vgi.make_grid(**w.grid)
        self.lights = vgi.make_lights(w.lights)
        self.gb = vgi.GBufferData(ctx, w.gb_pos, w.gb_nrm, w.width, w.height)
        self.cones = vgi.make_cones(**w.cones)
        self.od = vgi.DeviceArray.empty(ctx, (w.num_pixels, 4), np.float32)
        self.os = vgi.DeviceArray.empty(ctx, (w.num_pixels, 4), np.float32)
        self.scene = ctx.scene()


def run_batched(ctx, vgi, count=64, steps=5, reps=2, log=print, seed=sg.SEED_BATCH):
    t0 = time.perf_counter()
    ws = sg.batch_scenes(seed=seed, count=count)
    log(f"batched: {count} scenes generated in {time.perf_counter() - t0:.1f}s")
    sc = [_BatchScene(ctx, vgi, w) for w in ws]
    prec = vgi.PREC_FP16
    prev = ctx.get_tuning()
    vox = [(s.scene, [s.mesh], s.grid) for s in sc]
    mips = [(s.scene, s.lights) for s in sc]
    trace = [(s.scene, s.gb, s.cones, s.od, s.os) for s in sc]
    tuned = ctx.autotune_batch(vox, mips, trace, prec, reps=reps)
    log(f"batched: autotune {tuned.as_dict()}")

    def frame():
        ctx.voxelize_batch(vox)
        ctx.build_mips_batch(mips, prec)
        ctx.trace_batch(trace)

    frame()
    ctx.sync()
    ctx.timer_start()
    for _ in range(steps):
        frame()
    batch_ms = ctx.timer_stop() / steps
    # phase split of one batched frame
    ctx.set_profiling(True)
    frame()
    t = ctx.timings()
    ctx.set_profiling(False)
    # the same scenes one call per scene and stage
    ctx.sync()
    ctx.timer_start()
    for _ in range(max(1, steps // 2)):
        for s in sc:
            s.scene.voxelize([s.mesh], s.grid)
            s.scene.build_mips(s.lights, prec)
            s.scene.trace(s.gb, s.cones, s.od, s.os)
    alone_ms = ctx.timer_stop() / max(1, steps // 2)
    tris = sum(w.num_triangles for w in ws)
    px = sum(w.num_pixels for w in ws)
    out = dict(
        scenes=count, triangles=tris, pixels=px, mip_storage="fp16",
        grids={str(n): sum(1 for w in ws if w.grid["n"] == n) for n in (32, 64, 128, 256)},
        batch_frame_ms=batch_ms, one_call_per_scene_ms=alone_ms, batching_speedup=alone_ms / batch_ms,
        voxelize_ms=t.voxelize_total_ms, mips_ms=t.mips_ms, trace_ms=t.trace_ms,
        voxelize_mtris_s=tris / (t.voxelize_total_ms * 1e-3) / 1e6,
        trace_mpixels_s=px / (t.trace_ms * 1e-3) / 1e6,
        frame_mpixels_s=px / (batch_ms * 1e-3) / 1e6,
        autotune=tuned.as_dict())
    for s in sc:
        s.scene.release()
    ctx.put_tuning(prev)
    return out


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------

FRAME_SIZES = [(512, 512), (1024, 768), (1280, 720), (1920, 1080)]
POOL_PIXELS = 1920 * 1080


class _Soup:
    """A seeded triangle soup (per-triangle roughness 0.05-0.5) and, for one seeded pinhole
    camera inside it, a ray-cast G-buffer per frame size (centred crops of one 1920x1080
    image), all on the device."""

    def __init__(self, ctx, vgi, seed, ntri):
        pos, alb = sg.random_triangles(seed, ntri)
        rough = np.random.default_rng(seed).uniform(0.05, 0.5, ntri).astype(np.float32)
        cam = sg.camera_pose(seed, (0.2, 0.2, 0.2), (0.8, 0.8, 0.8))
        gp, gn, gr = sg.gbuffer_camera(pos, rough, cam, 1920, 1080)
        self.ntri = ntri
        self.eye = tuple(float(v) for v in cam[:3])
        self.mesh = vgi.MeshData(ctx, pos.reshape(-1, 3), albedo=alb)
        self.gbs = {}
        img_p, img_n, img_r = gp.reshape(1080, 1920, 3), gn.reshape(1080, 1920, 3), gr.reshape(1080, 1920)
        for (w, h) in FRAME_SIZES:
            y0, x0 = (1080 - h) // 2, (1920 - w) // 2
            crop = (slice(y0, y0 + h), slice(x0, x0 + w))
            self.gbs[(w, h)] = vgi.GBufferData(ctx, np.ascontiguousarray(img_p[crop]).reshape(-1, 3),
                                               np.ascontiguousarray(img_n[crop]).reshape(-1, 3), w, h,
                                               np.ascontiguousarray(img_r[crop]).reshape(-1))


class _OutPool:
    """Output buffers for frames in flight (largest frame size), handed out per request."""

    def __init__(self, ctx, vgi, count):
        self.free = [(vgi.DeviceArray(ctx, POOL_PIXELS * 16), vgi.DeviceArray(ctx, POOL_PIXELS * 16))
                     for _ in range(count)]

    def take(self, k):
        if len(self.free) < k:
            return None
        out, self.free = self.free[:k], self.free[k:]
        return out

    def give(self, bufs):
        self.free.extend(bufs)


def run_serving(ctx, vgi, seconds=60.0, qps=200.0, budget_gb=200.0, keys=1000, soups=24,
                zipf_s=1.0, revox=0.1, slots=4, max_batch=16, num_diffuse=9, out_buffers=192,
                log=print, seed=0xF1840005):
    rng = np.random.default_rng(seed)
    ntris = sg.lognormal_ints(seed, soups, 50_000.0, 1.0, 1_000, 200_000)
    t0 = time.perf_counter()
    src = [_Soup(ctx, vgi, seed + 7 * (i + 1), int(ntris[i])) for i in range(soups)]
    log(f"serving: {soups} soups ({int(ntris.sum())} triangles) + camera G-buffers ready in "
        f"{time.perf_counter() - t0:.1f}s")

    # per-key scene definition
    key_n = rng.choice([128, 256], size=keys)
    key_shift = rng.uniform(0.0, 0.05, size=(keys, 3))
    key_light = []
    for k in range(keys):
        d = rng.normal(size=3)
        d[1] = -abs(d[1]) - 0.5
        d /= np.linalg.norm(d)
        key_light.append(vgi.make_lights([dict(type="directional", vec=tuple(float(v) for v in d),
                                                color=(1.0, 1.0, 1.0))]))
    grids = [vgi.make_grid(tuple(float(-0.05 - v) for v in key_shift[k]), 1.1, int(key_n[k]),
                           int(math.log2(int(key_n[k])))) for k in range(keys)]
    zipf_w = 1.0 / np.arange(1, keys + 1) ** zipf_s
    zipf_w /= zipf_w.sum()
    perm = rng.permutation(keys)  # popularity rank -> key

    # request stream
    nreq = int(seconds * qps)
    gaps = rng.exponential(1.0 / qps, size=nreq)
    arrivals = np.cumsum(gaps)
    reqs = []
    for i in range(nreq):
        key = int(perm[rng.choice(keys, p=zipf_w)])
        nf = int(rng.integers(1, 9))
        frames = []
        for _ in range(nf):
            frames.append(FRAME_SIZES[int(rng.integers(len(FRAME_SIZES)))])
        reqs.append((key, frames, bool(rng.random() < revox)))
    cone_base = dict(num_diffuse=num_diffuse, specular=True, diffuse_aperture_deg=60.0,
                     diffuse_ring_deg=60.0, specular_aperture_deg=10.0, step_factor=0.5,
                     max_distance=1.0, alpha_cutoff=0.99, normal_offset=1.0, start_offset=1.0)
    outs = _OutPool(ctx, vgi, out_buffers)
    budget = int(budget_gb * 1e9)  # decimal gigabytes
    lat, builds, frames_done, px_done = [], 0, 0, 0
    held = {}
    with vgi.Server(ctx, budget, vgi.PREC_FP16, slots=slots, max_batch_requests=max_batch) as srv:
        t_start = vgi.now() + 0.05
        i = 0
        first_arrival = t_start + float(arrivals[0])
        last_done = first_arrival
        late = []
        while i < nreq or held:
            now = vgi.now()
            # submit everything that has arrived
            while i < nreq and t_start + arrivals[i] <= now:
                key, frames, rv = reqs[i]
                bufs = outs.take(len(frames))
                if bufs is None:
                    break  # out of output buffers: wait for completions
                s = src[key % soups]
                fl = [(s.gbs[(wdt, hgt)], vgi.make_cones(**cone_base, eye=s.eye), od, os_)
                      for (wdt, hgt), (od, os_) in zip(frames, bufs)]
                arr = t_start + float(arrivals[i])
                late.append(now - arr)
                srv.submit(key, [s.mesh], grids[key], key_light[key], fl, tag=i, revoxelize=rv,
                           arrival_s=arr)
                held[i] = (bufs, sum(w * h for w, h in frames), len(frames))
                i += 1
            if i < nreq:  # sleep in the poll until the next arrival (at most 2 ms)
                wait = min(2000, max(0, int((t_start + arrivals[i] - vgi.now()) * 1e6)))
            else:
                wait = 2000
            for c in srv.poll(wait_us=wait):
                bufs, px, nf = held.pop(c.tag)
                outs.give(bufs)
                if c.status != 0:
                    raise RuntimeError(f"request {c.tag} failed with status {c.status}")
                lat.append(c.done_s - c.arrival_s)
                builds += c.voxelized
                frames_done += nf
                px_done += px
                last_done = max(last_done, c.done_s)
        stats = srv.stats().as_dict()
    span = last_done - first_arrival
    return dict(
        requests=nreq, seconds=seconds, qps_offered=qps, qps_achieved=nreq / span,
        p50_ms=_percentile(lat, 50) * 1e3, p99_ms=_percentile(lat, 99) * 1e3,
        mean_ms=float(np.mean(lat)) * 1e3, max_ms=float(np.max(lat)) * 1e3,
        frames_per_s=frames_done / span, mpixels_per_s=px_done / span / 1e6,
        builds=builds, submit_lag_p99_ms=_percentile(late, 99) * 1e3,
        keys=keys, zipf_s=zipf_s, soups=soups, revoxelize_rate=revox, budget_gb=budget_gb,
        slots=slots, max_batch_requests=max_batch, mip_storage="fp16",
        cones=f"{num_diffuse} diffuse 60deg + 1 specular 10deg", server=stats)
