"""Traversal benchmark: incoherent closest-hit Mrays/s on one MI355X.

Workload: the 1,000,000-triangle fBm heightfield (1000x500 cells, 6 octaves, seed 3) with
4,194,304 incoherent cosine-distributed bounce rays (seed 4) and 4,194,304 coherent primary rays
(1024x512 pixels x 8 jittered samples). One step = one closest-hit pass over the incoherent batch,
rays and hits resident in HBM. Coherent closest-hit and any-hit shadow passes are timed separately
and reported as extra fields.

Prints one JSON line (rank 0). Only one GPU is supported.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
RAY_BYTES, HIT_BYTES, NODE_BYTES, TRI_BYTES, INST_BYTES = 32, 16, 64, 48, 64


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def timed_passes(fn, steps):
    """Runs fn() `steps` times between device synchronisations; returns (wall s, event ms per pass)."""
    import myrt_amd as m
    timer = m.Timer()
    m.synchronize()
    t0 = time.perf_counter()
    timer.start()
    for _ in range(steps):
        fn()
    timer.stop()
    m.synchronize()
    wall = time.perf_counter() - t0
    return wall, timer.elapsed_ms() / steps


def build_wavefront_scene(sc, with_sdf=True):
    import myrt_amd as m
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for mat in sc["materials"]:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    s.set_environment(1.0)
    for g, mt, mi in sc["instances"]:
        s.enqueue(gids[g], mt, mi)
    if with_sdf:
        sid = s.push_sdf(sc["sdf_tape"])
        s.enqueue_sdf(sid, sc["sdf_transform"])
    return s


def animated_frames(s, sc, d_r, n_rays, d_h, frames=5):
    """Frames that move every instance of the wavefront scene (a random translation of about 0.02
    scene units per axis and frame), prepared with a host rebuild of the instance BVH (the
    reference's behaviour on a changed drawable hash) and with the GPU refit; the device-matrix
    update alone; closest-hit throughput of the refitted TLAS after all frames."""
    import myrt_amd as m
    rng = np.random.default_rng(9)
    inst = [(g, np.asarray(mt, np.float32), mi) for g, mt, mi in sc["instances"]]
    out = {"instances": len(inst), "frames": frames}

    def step_frame():
        for k, (g, mt, mi) in enumerate(inst):
            mt = mt.copy()
            mt[:3, 3] += rng.normal(size=3).astype(np.float32) * 0.02
            inst[k] = (g, mt, mi)
            s.enqueue(g, mt, mi)

    step = lambda: s.trace_closest_device(d_r, n_rays, d_h)
    # Refits first; the closest-hit rate on the refitted TLAS is then compared with the rate on a
    # TLAS rebuilt one frame later (the same rays; the scene has moved by one more small step).
    for mode in ("refit", "rebuild"):
        s.set_tlas_refit(0 if mode == "rebuild" else 1 << 30)
        times, refitted = [], 0
        for _ in range(frames):
            step_frame()
            m.synchronize()
            t0 = time.perf_counter()
            p = s.prepare()
            m.synchronize()
            times.append(time.perf_counter() - t0)
            refitted += p["tlas_refitted"]
            if mode == "rebuild" or refitted < frames:
                continue
            mats = np.stack([np.ascontiguousarray(mt.T).ravel() for _, mt, _ in inst]).astype(np.float32)
            d_m = m.DeviceBuffer.from_array(mats)
            _, ms_dev = timed_passes(lambda: s.update_transforms_device(d_m, len(inst)), 5)
            out["device_update_ms"] = round(ms_dev, 4)
        out[mode + "_prepare_ms"] = round(1e3 * float(np.median(times)), 3)
        out[mode + "_frames_refitted"] = int(refitted)
        step()
        _, ms = timed_passes(step, 3)
        out["incoherent_closest_mrays_after_" + mode] = round(n_rays / (ms * 1e-3) / 1e6, 2)
    s.set_tlas_refit(0)
    return out


def wavefront_extras(args):
    """The 4M-triangle wavefront scene (2000 instanced icospheres, displaced ground, 64 SDF
    primitives): incoherent closest-hit rays/s and path-tracer samples/s at 1920x1080, 16 spp, depth 8."""
    import myrt_amd as m
    from myrt_amd import scenes
    log("wavefront scene: generating")
    sc = scenes.wavefront_scene()
    tris, refs = scenes.world_triangles(sc["geometries"], sc["instances"])
    rays = scenes.scene_bounce_rays(tris, refs, 1 << 22, seed=4)
    n_tris = int(tris.shape[0])
    del tris, refs
    d_r = m.DeviceBuffer.from_array(rays)
    d_h = m.DeviceBuffer(rays.size * HIT_BYTES)
    # Triangles only (the 4M instanced triangles), then the same scene with the SDF cluster.
    st_tri = build_wavefront_scene(sc, with_sdf=False)
    st_tri.prepare()
    step = lambda: st_tri.trace_closest_device(d_r, rays.size, d_h)
    step()
    _, ms_tri = timed_passes(step, max(3, args.steps // 2))
    st = st_tri.trace_stats(d_r, rays.size)
    st_tri.close()
    s = build_wavefront_scene(sc)
    t0 = time.perf_counter()
    prep = s.prepare()
    prep_s = time.perf_counter() - t0
    log("wavefront scene prepared in %.2f s: %s" % (prep_s, prep))
    step = lambda: s.trace_closest_device(d_r, rays.size, d_h)
    step()
    _, ms = timed_passes(step, max(3, args.steps // 2))
    W, H, spp, depth = 1920, 1080, 16, 8
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    r.render(s, W, H, 1)  # warm-up; allocates the queues
    st0 = r.stats()
    log("wavefront render %dx%d %d spp depth %d" % (W, H, spp, depth))
    r.render(s, W, H, spp)
    st1 = r.stats()
    dt = (st1["milliseconds"] - st0["milliseconds"]) * 1e-3
    n_rays = st1["rays"] - st0["rays"]
    # The same render with shading in material order instead of queue order (the default).
    r.set_material_sort(True)
    r.invalidate()
    r.render(s, W, H, 1)
    st0 = r.stats()
    log("wavefront render, shading in material order")
    r.render(s, W, H, spp)
    dt_sorted = (r.stats()["milliseconds"] - st0["milliseconds"]) * 1e-3
    log("wavefront scene: animated frames")
    anim = animated_frames(s, sc, d_r, rays.size, d_h)
    return {"triangles": n_tris, "sdf_primitives": 64, "prepare_s": round(prep_s, 3),
            "wide_nodes": prep["blas_wide_nodes"] + prep["tlas_wide_nodes"], "device_bytes": prep["device_bytes"],
            "incoherent_closest_mrays_triangles_only": round(rays.size / (ms_tri * 1e-3) / 1e6, 2),
            "incoherent_closest_mrays_with_sdf": round(rays.size / (ms * 1e-3) / 1e6, 2),
            "incoherent_node_fetches_per_ray": round(st["node_visits"] / rays.size, 3),
            "incoherent_tri_fetches_per_ray": round(st["tri_tests"] / rays.size, 3),
            "incoherent_instance_visits_per_ray": round(st["instance_visits"] / rays.size, 3),
            "render": {"width": W, "height": H, "spp": spp, "depth": depth, "seconds": round(dt, 4),
                       "samples_per_s": round(W * H * spp / dt, 1), "rays_traced": int(n_rays),
                       "mrays_per_s": round(n_rays / dt / 1e6, 2), "seconds_sorted_shading": round(dt_sorted, 4)},
            "animated_frames": anim}


def equal_time_study(args, W=480, H=270, depth=8, ref_spp=4096, budgets=(4, 16, 64), warm_frames=16):
    """Equal-time MSE against a converged render of the 4M-triangle wavefront scene: ref_spp
    samples of the plain tracer in 256-sample chunks over disjoint sample ranges, each chunk's mean
    accumulated in binary64 on the host. For each budget (the time the plain tracer takes for that
    many samples per pixel) every method gets the same time, its denoiser included:
    plain; plain + denoiser (as many samples as fit beside the denoiser); the radiance cache and
    cache + denoiser from a warm cache (warm_frames frames of 16 samples of the same view, an
    interactive session's history, not charged) and from a cold cache (a fresh network trained only
    inside the budget). The cache keeps training while it renders."""
    import myrt_amd as m
    from myrt_amd import scenes
    weights = os.path.join(ROOT, "myrt_amd", "data", "denoiser_unet.npy")
    sc = scenes.wavefront_scene()
    s = build_wavefront_scene(sc)
    s.prepare()
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)

    def renderer():
        r = m.Renderer()
        r.set_camera(view, proj)
        r.set_num_bounces(depth)
        return r

    log("equal-time study: reference %d spp at %dx%d (binary64 over 256-sample chunks)" % (ref_spp, W, H))
    ref_r = renderer()
    acc = np.zeros((H, W, 3), np.float64)
    chunks = ref_spp // 256
    for k in range(chunks):
        ref_r.set_sample_base(1 << 20 | (256 * k))  # sample indices no timed render uses
        ref_r.render(s, W, H, 256)
        acc += ref_r.image()[..., :3].astype(np.float64)
    ref = acc / chunks

    def mse(img):
        return float(((img[..., :3].astype(np.float64) - ref) ** 2).mean())

    def timed(fn):
        m.synchronize()
        t0 = time.perf_counter()
        out = fn()
        m.synchronize()
        return out, time.perf_counter() - t0

    def median_time(fn, k=3):
        """Median of k timed runs (a single render's time varies by several per cent); the last
        run's output is what the caller reads."""
        return float(np.median([timed(fn)[1] for _ in range(k)]))

    def fresh(r):
        r.invalidate()
        return r

    def fit(render_spp, T, t_extra, first):
        """The largest sample count whose render plus t_extra fits T (1 % slack), by rescaling
        each candidate's time (at most 6 candidates, each the median of 3 renders); returns
        (spp, seconds)."""
        spp, times, last = max(1, first), {}, None
        while spp not in times and len(times) < 6:
            times[spp] = render_spp(spp)
            last = spp
            lo = max([k for k, t in times.items() if t + t_extra <= T * 1.01], default=0)
            hi = min([k for k, t in times.items() if t + t_extra > T * 1.01], default=1 << 30)
            guess = int(spp * (T - t_extra) / times[spp])
            spp = min(max(guess, lo + 1), hi - 1) if hi - lo > 1 else max(lo, 1)
        best = max([k for k, t in times.items() if t + t_extra <= T * 1.01], default=1)
        if best != last:  # the renderer must hold the chosen render's image
            times[best] = render_spp(best)
        return best, times[best]

    plain = renderer()
    plain.render(s, W, H, 16)  # allocations and first launches stay out of the timings
    den = None
    if os.path.exists(weights):
        den = m.Denoiser()
        den.load(weights)
        plain.denoise(den)
    # Render paths end in the cache at vertex 2; training paths (one in 16, stratified) at vertex 4,
    # learning their tail from the cache (self-training); 8 Adam steps per 2^21 paths.
    cache_cfg = dict(mode=2, train_fraction=1.0 / 16, train_steps=8, min_bounce=2)
    train_cfg = dict(train_end=4, unbiased_fraction=0.0)

    def use_cache(r, cache):
        r.set_cache(cache, **cache_cfg)
        r.set_cache_training(**train_cfg)

    warm_cache = m.RadianceCache(seed=1)
    warm_cache.set_bounds((-12.5, -7.0, -9.5), (12.5, 6.0, 9.5))
    warm = renderer()
    use_cache(warm, warm_cache)
    _, warm_s = timed(lambda: [warm.render(s, W, H, 16) for _ in range(warm_frames)])
    cr = renderer()
    use_cache(cr, warm_cache)
    cr.render(s, W, H, 16)
    _, t16 = timed(lambda: fresh(cr).render(s, W, H, 16))
    res = {"width": W, "height": H, "depth": depth, "reference_spp": ref_spp, "reference_accumulation": "binary64 over "
           "256-sample fp32 chunks", "cache_warm_frames": warm_frames, "cache_warm_s": round(warm_s, 4),
           "cache_seconds_per_16spp": round(t16, 4), "cache_settings": dict(cache_cfg, **train_cfg), "budgets": []}
    cold_r = renderer()
    cold_r.render(s, W, H, 16)  # allocations stay out of the timings
    for b in budgets:
        T = median_time(lambda: fresh(plain).render(s, W, H, b))
        row = {"budget_spp": b, "budget_s": round(T, 4), "plain_mse": mse(plain.image())}
        # The reference's own post filter (bilateral, the app's sigmas) as the classical baseline.
        img, t_bl = timed(lambda: m.post_process(plain.image(), [m.POST_BILATERAL], 10.0, 0.05))
        row["plain_bilateral_mse"] = mse(img)
        row["bilateral_s"] = round(t_bl, 4)
        t_dn = 0.0
        if den is not None:
            _, t_dn = timed(lambda: plain.denoise(den))
            row["denoise_s"] = round(t_dn, 4)
            spp_pd, t_pd = fit(lambda k: median_time(lambda: fresh(plain).render(s, W, H, k)), T, t_dn,
                               int(b * (T - t_dn) / T))
            img, t_dn2 = timed(lambda: plain.denoise(den))
            row.update({"plain_denoised_spp": spp_pd, "plain_denoised_s": round(t_pd + t_dn2, 4),
                        "plain_denoised_mse": mse(img)})
        # Warm cache: as many samples as fit (with the denoiser when there is one).
        spp_c, tc = fit(lambda k: median_time(lambda: fresh(cr).render(s, W, H, k)), T, t_dn,
                        int((T - t_dn) / t16 * 16))
        row.update({"cache_spp": spp_c, "cache_s": round(tc, 4), "cache_mse": mse(cr.image())})
        if den is not None:
            img, t = timed(lambda: cr.denoise(den))
            row["cache_denoised_mse"] = mse(img)
            row["ratio_plain_over_cache_denoised"] = round(row["plain_mse"] / row["cache_denoised_mse"], 3)
            row["ratio_plain_over_plain_denoised"] = round(row["plain_mse"] / row["plain_denoised_mse"], 3)
            # Cold cache: a fresh network for every trial, trained only inside the budget.
            def cold_render(k):
                ts = []
                for _ in range(3):  # the median of 3, each from a fresh network
                    cold = m.RadianceCache(seed=2)
                    cold.set_bounds((-12.5, -7.0, -9.5), (12.5, 6.0, 9.5))
                    use_cache(cold_r, cold)
                    ts.append(timed(lambda: fresh(cold_r).render(s, W, H, k))[1])
                    cold_render.keep = cold
                return float(np.median(ts))
            spp_k, tk = fit(cold_render, T, t_dn, spp_c)
            img, t = timed(lambda: cold_r.denoise(den))
            row.update({"cold_cache_spp": spp_k, "cold_cache_s": round(tk + t, 4), "cold_cache_denoised_mse": mse(img),
                        "ratio_plain_over_cold_cache_denoised": round(row["plain_mse"] / mse(img), 3)})
        row["ratio_plain_over_cache"] = round(row["plain_mse"] / row["cache_mse"], 3)
        res["budgets"].append(row)
        log("budget %d spp: %s" % (b, row))
    fresh(cr).render(s, W, H, 256)
    res["cache_256spp_mse"] = mse(cr.image())  # bias floor of the trained cache
    fresh(plain).render(s, W, H, 256)
    res["plain_256spp_mse"] = mse(plain.image())
    s.close()
    return res


UNET_LEVELS = (32, 48, 64, 96, 128)


def unet_macs_per_pixel():
    """Multiply-adds per full-resolution pixel of the denoiser (logical channel counts)."""
    macs, c_in = 0.0, 10
    for l, c in enumerate(UNET_LEVELS):
        macs += (9 * c_in * c + 9 * c * c) / 4 ** l
        c_in = c
    for l in range(len(UNET_LEVELS) - 2, -1, -1):
        c = UNET_LEVELS[l]
        macs += (9 * (UNET_LEVELS[l + 1] + c) * c + 9 * c * c) / 4 ** l
    return macs + UNET_LEVELS[0] * 3


def c1_render(args, frames=20):
    """The plain renderer on the reference's small configuration (ground quad + 200 icospheres,
    64,002 triangles, 512x512, 1 spp, depth 4), one frame per call as an interactive loop renders
    it: the median frame time and samples/s, for comparison with the CPU baseline's restated
    reference renderer on the same scene."""
    import myrt_amd as m
    from myrt_amd import scenes
    sc = scenes.sphere_field_scene(200, 2, seed=1)
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for mat in sc["materials"]:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    s.set_environment(1.0)
    for g, mt, mi in sc["instances"]:
        s.enqueue(gids[g], mt, mi)
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], 1.0)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(4)
    r.render(s, 512, 512, 1)
    times = []
    for _ in range(frames):
        m.synchronize()
        t0 = time.perf_counter()
        r.render(s, 512, 512, 1)
        m.synchronize()
        times.append(time.perf_counter() - t0)
    st = r.stats()
    ms = float(np.median(times)) * 1e3
    out = {"width": 512, "height": 512, "spp": 1, "depth": 4, "frame_ms": round(ms, 3),
           "samples_per_s": round(512 * 512 / (ms * 1e-3), 1), "rays_per_sample": round(st["rays"] / st["samples"], 3)}
    r.close()
    s.close()
    return out


def learned_kernels(args):
    """Timings of the learned parts on the device: the denoiser at 1920x1080 (fp16 MFMA), the
    radiance-cache MLP inference on 2^20 queries and one training step on 2^16 samples (bf16 MFMA)."""
    import myrt_amd as m
    out = {}
    rng = np.random.default_rng(0)
    W, H = 1920, 1080
    raw = np.concatenate([rng.uniform(0, 2, (H, W, 3)), rng.uniform(0, 1, (H, W, 3)),
                          rng.normal(size=(H, W, 3)), rng.uniform(1, 20, (H, W, 1))], -1).astype(np.float32)
    d = m.Denoiser(seed=1)
    d_raw = m.DeviceBuffer.from_array(raw)
    d_out = m.DeviceBuffer(W * H * 16)
    step = lambda: m._check(m.lib().mrt_denoiser_denoise_device(d.h, d_raw.ptr, H, W, d_out.ptr, None))
    step()
    _, ms = timed_passes(step, 10)
    flops = 2.0 * unet_macs_per_pixel() * W * ((H + 15) // 16 * 16)
    out["denoiser_1080p"] = {"ms": round(ms, 3), "tflops": round(flops / (ms * 1e-3) / 1e12, 1),
                             "frac_of_fp16_mfma_peak": round(flops / (ms * 1e-3) / 2.5e15, 4),
                             "flop": flops}
    c = m.RadianceCache(seed=1)
    nq = 1 << 20
    q = np.zeros(nq, m.QUERY_DTYPE)
    q["pos"] = rng.uniform(-1, 1, (nq, 3))
    q["dir"] = rng.normal(size=(nq, 3))
    q["nrm"] = rng.normal(size=(nq, 3))
    q["roughness"] = rng.uniform(0, 1, nq)
    q["albedo"] = rng.uniform(0, 1, (nq, 3))
    d_q = m.DeviceBuffer.from_array(q)
    d_y = m.DeviceBuffer(nq * 12)
    inf = lambda: m._check(m.lib().mrt_cache_infer_device(c.h, d_q.ptr, nq, d_y.ptr, None))
    inf()
    _, ms_i = timed_passes(inf, 10)
    nt = 1 << 16
    tgt = m.DeviceBuffer.from_array(rng.uniform(0, 1, (nt, 3)).astype(np.float32))
    trn = lambda: m._check(m.lib().mrt_cache_train_device(c.h, d_q.ptr, tgt.ptr, nt, None))
    trn()
    _, ms_t = timed_passes(trn, 10)
    out["cache_infer_2^20"] = {"ms": round(ms_i, 3), "tflops": round(41344.0 * nq / (ms_i * 1e-3) / 1e12, 1),
                               "frac_of_bf16_mfma_peak": round(41344.0 * nq / (ms_i * 1e-3) / 2.5e15, 4)}
    out["cache_train_2^16"] = {"ms": round(ms_t, 3), "tflops": round(3 * 41344.0 * nt / (ms_t * 1e-3) / 1e12, 1)}
    return out


CPU_COMPILE = "g++ -O3 -march=native -ffp-contract=off -fno-fast-math (compiled on this host at bench time)"


def cpu_baseline_child(npz_path, threads_all):
    """The CPU baseline, run in a child process whose oracle library was built for this host.
    Times the reference's traversal (stackless bit-stack walk and Moller-Trumbore, restated) on
    bounded samples of the microbench's coherent, incoherent and shadow rays, and the reference's
    plain renderer (restated generate / trace / color / filter passes) on its small configuration
    (ground quad + 200 icospheres, 64,002 triangles, 512x512, 1 spp, depth 4), each at all host
    threads and at one thread, under the thread_pool clone; median of 5 after a warm-up."""
    import oracle
    from myrt_amd import scenes
    data = np.load(npz_path)
    osc = oracle.OracleScene([(data["idx"], data["pts"])], [(0, np.eye(4, dtype=np.float32))])
    out = {"traversal": {}, "render_c1": {}}
    for cls, any_hit in (("incoherent", False), ("coherent", False), ("shadow", True)):
        rays = data[cls]
        row = {}
        for label, threads, n in (("all_threads", 0, rays.size), ("one_thread", 1, rays.size // 8)):
            rate, used = osc.reference_time(rays[:n], any_hit=any_hit, threads=threads, chunk=65536, reps=5)
            row[label] = {"mrays_per_s": round(rate / 1e6, 4), "threads": used, "rays": int(n)}
        out["traversal"][cls] = row
    sc = scenes.sphere_field_scene(200, 2, seed=1)
    o = oracle.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"])
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], 1.0)
    for label, threads in (("all_threads", 0), ("one_thread", 1)):
        sps, rps, used, _ = o.reference_render_time(view, proj, 512, 512, 4, threads=threads, reps=5)
        out["render_c1"][label] = {"samples_per_s": round(sps, 1), "mrays_per_s": round(rps / 1e6, 4), "threads": used}
    print(json.dumps(out), flush=True)


def run_cpu_baseline(args, idx, pts, coh, inc, sh):
    """Builds the oracle for this host (falling back to the prebuilt -O2 library) and runs
    cpu_baseline_child in a subprocess; returns its result dict and the compile note."""
    import subprocess
    import tempfile
    tmp = tempfile.mkdtemp(prefix="myrt_cpu_baseline_")
    env = dict(os.environ)
    compile_note = CPU_COMPILE
    try:
        from oracle import build as ob
        env["MYRT_ORACLE_LIBRARY"] = ob.build_native(tmp)
    except Exception as ex:  # no compiler: the prebuilt library, which targets any x86-64
        log("native oracle build failed (%s); timing the prebuilt library" % str(ex).splitlines()[0])
        compile_note = "g++ -O2 -ffp-contract=off (prebuilt, generic x86-64)"
    n = args.cpu_sample
    npz = os.path.join(tmp, "rays.npz")
    np.savez(npz, idx=idx, pts=pts, incoherent=inc[:n], coherent=coh[:n], shadow=sh[:n])
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", npz], env=env,
                       capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline child failed:\n" + r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1]), compile_note


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only-incoherent", action="store_true",
                    help="time only the incoherent closest-hit passes (for profiler runs)")
    ap.add_argument("--no-wavefront", action="store_true", help="skip the 4M-triangle wavefront scene")
    ap.add_argument("--no-study", action="store_true", help="skip the equal-time MSE study")
    ap.add_argument("--only-study", action="store_true", help="run only the equal-time MSE study")
    ap.add_argument("--only-learned", action="store_true", help="time only the denoiser and cache kernels")
    ap.add_argument("--only-wavefront", action="store_true",
                    help="run only the 4M-triangle wavefront scene and print its fields (for profiler runs)")
    ap.add_argument("--cpu-sample", type=int, default=1 << 20, help="incoherent rays timed on the host")
    ap.add_argument("--cpu-baseline-child", metavar="NPZ", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel 3 times on BYTES (calibrates rocprof byte counters)")
    args = ap.parse_args()
    if args.cpu_baseline_child:
        cpu_baseline_child(args.cpu_baseline_child, 0)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only; multi-GPU runs are not part of this build "
              "(the path would shard rays across GPUs with a replicated scene).", file=sys.stderr)
        sys.exit(2)

    import myrt_amd as m
    from myrt_amd import scenes

    if args.copy_calibration:
        n = args.copy_calibration // 16 * 16
        a, b = m.DeviceBuffer(n), m.DeviceBuffer(n)
        for _ in range(3):
            m._check(m.lib().mrt_device_copy(b.ptr, a.ptr, n, None))
        m.synchronize()
        print(json.dumps({"copy_bytes": n, "launches": 3}), flush=True)
        return

    if args.only_learned:
        print(json.dumps(learned_kernels(args)), flush=True)
        return
    if args.only_study:
        print(json.dumps(equal_time_study(args)), flush=True)
        return
    if args.only_wavefront:
        print(json.dumps(wavefront_extras(args)), flush=True)
        return

    log("generating workload")
    (idx, pts, nrm, uv), coh, inc = scenes.traversal_workload()
    scene = m.Scene()
    t0 = time.perf_counter()
    g = scene.push_geometry(idx, pts, nrm, uv)
    scene.enqueue(g)
    prep = scene.prepare()
    build_s = time.perf_counter() - t0
    log("prepared: %s (%.2f s)" % (prep, build_s))
    name, cus = m.device_info()

    d_inc = m.DeviceBuffer.from_array(inc)
    d_coh = m.DeviceBuffer.from_array(coh)
    n_inc, n_coh = inc.size, coh.size
    d_hits = m.DeviceBuffer(max(n_inc, n_coh) * HIT_BYTES)
    d_occ = m.DeviceBuffer(max(n_inc, n_coh))

    step = lambda: scene.trace_closest_device(d_inc, n_inc, d_hits)
    for _ in range(args.warmup):
        step()
    m.synchronize()
    log("timing %d incoherent closest-hit passes" % args.steps)
    wall, ev_ms = timed_passes(step, args.steps)
    value = n_inc * args.steps / wall / 1e6
    kernel_rate = n_inc / (ev_ms * 1e-3)

    # Fetch counts per ray from the instrumented traversal (same kernel with counters).
    st_inc = scene.trace_stats(d_inc, n_inc)
    n_node = st_inc["node_visits"] / n_inc
    n_tri = st_inc["tri_tests"] / n_inc
    n_ins = st_inc["instance_visits"] / n_inc
    # Algorithmic bytes per ray: 32 (ray) + 16 (hit) + 64 per wide node + 48 per triangle record,
    # plus the 64-B instance record read when the ray enters the mesh's instance (once per ray
    # here), a named extra term.
    b_ray = RAY_BYTES + HIT_BYTES + NODE_BYTES * n_node + TRI_BYTES * n_tri + INST_BYTES * n_ins
    achieved = kernel_rate * b_ray / 1e9
    # HBM traffic per launch from the committed rocprofv3 counter passes over this command
    # (FETCH_SIZE and WRITE_SIZE, calibrated on the copy kernel).
    traffic, l2_hit = None, None
    pmc = os.path.join(ROOT, "profiles", "trace_kernel_heightfield_1M_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            rec = json.load(f)
        traffic, l2_hit = rec["traffic_bytes_per_launch"], rec["l2_hit_rate"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "l2_hit_rate": l2_hit,
                "bytes_per_ray": round(b_ray, 1),
                "bytes_per_ray_terms": {"ray_and_hit": RAY_BYTES + HIT_BYTES, "wide_nodes": round(NODE_BYTES * n_node, 1),
                                        "triangle_records": round(TRI_BYTES * n_tri, 1),
                                        "instance_record": round(INST_BYTES * n_ins, 1)},
                "node_fetches_per_ray": round(n_node, 3), "tri_fetches_per_ray": round(n_tri, 3),
                "instance_fetches_per_ray": round(n_ins, 3), "kernel_ms": round(ev_ms, 4),
                "kernel_ms_source": "HIP events on the launch stream around the timed passes"}
    if args.only_incoherent:
        print(json.dumps({"metric": "incoherent Mrays/s", "value": round(value, 2), "steps": args.steps,
                          "roofline": roofline}), flush=True)
        return

    # Coherent closest-hit and any-hit shadow rays (extra fields, same protocol).
    coh_step = lambda: scene.trace_closest_device(d_coh, n_coh, d_hits)
    coh_step()
    coh_wall, coh_ms = timed_passes(coh_step, max(3, args.steps // 2))
    scene.trace_closest_device(d_inc, n_inc, d_hits)
    hits_inc = d_hits.to_array(m.HIT_DTYPE, n_inc)
    sh = scenes.shadow_rays(inc, hits_inc)
    d_sh = m.DeviceBuffer.from_array(sh)
    sh_step = lambda: scene.trace_any_device(d_sh, sh.size, d_occ)
    sh_step()
    sh_wall, sh_ms = timed_passes(sh_step, max(3, args.steps // 2))

    st_coh = scene.trace_stats(d_coh, n_coh)

    wave = None if (args.only_incoherent or args.no_wavefront) else wavefront_extras(args)
    study = None if (args.only_incoherent or args.no_study) else equal_time_study(args)
    learned = None if args.only_incoherent else learned_kernels(args)
    c1 = None if args.only_incoherent else c1_render(args)

    cpu = None
    if not args.no_cpu_baseline and not args.only_incoherent:
        log("cpu baseline: reference traversal and renderer on the host (bounded samples)")
        res, note = run_cpu_baseline(args, idx, pts, coh, inc, sh)
        tr = res["traversal"]
        inc_all = tr["incoherent"]["all_threads"]
        cpu = {"value": inc_all["mrays_per_s"], "unit": "Mrays/s", "cores": inc_all["threads"], "kind": "port",
               "sample": "%d incoherent rays of this workload (1/8 of that at one thread), the reference's stackless "
                         "bit-stack traversal and Moller-Trumbore restated in C++ (oracle/), 65,536-ray jobs on a "
                         "thread_pool clone, median of 5 after one warm-up" % inc_all["rays"],
               "compile": note,
               "single_thread_mrays_per_s": tr["incoherent"]["one_thread"]["mrays_per_s"],
               "per_class": tr,
               "render_c1": dict(res["render_c1"], config="the reference's plain renderer (pt_generate / pt_trace / "
                                 "pt_color / pt_ray_filter and pbr.glsl restated, its two sphere lights) on ground "
                                 "quad + 200 icospheres, 512x512, 1 spp, depth 4, 64-row jobs")}

    out = {
        "metric": "incoherent Mrays/s",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded fBm heightfield and seeded ray sets)",
        "config": {"workload": "traversal microbench: 1,000,000-triangle fBm heightfield (1000x500, 6 octaves), "
                               "4,194,304 incoherent cosine bounce rays, closest hit",
                   "triangles": int(idx.shape[0]), "rays_per_step": int(n_inc), "device": name, "cus": cus},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": {
            "incoherent_kernel_mrays": round(kernel_rate / 1e6, 2),
            "coherent_closest_mrays": round(n_coh / (coh_ms * 1e-3) / 1e6, 2),
            "shadow_any_mrays": round(sh.size / (sh_ms * 1e-3) / 1e6, 2),
            "coherent_node_fetches_per_ray": round(st_coh["node_visits"] / n_coh, 3),
            "coherent_tri_fetches_per_ray": round(st_coh["tri_tests"] / n_coh, 3),
            "incoherent_hit_fraction": round(float(np.isfinite(hits_inc["t"]).mean()), 4),
            "bvh": {"wide_nodes": prep["blas_wide_nodes"], "tri_records": prep["blas_tri_records"],
                    "device_bytes": prep["device_bytes"], "host_build_s": round(build_s, 3)},
            "wavefront_4M": wave,
            "equal_time": study,
            "learned_kernels": learned,
            "c1_plain_render": c1,
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
