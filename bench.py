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
    return {"triangles": n_tris, "sdf_primitives": 64, "prepare_s": round(prep_s, 3),
            "wide_nodes": prep["blas_wide_nodes"] + prep["tlas_wide_nodes"], "device_bytes": prep["device_bytes"],
            "incoherent_closest_mrays_triangles_only": round(rays.size / (ms_tri * 1e-3) / 1e6, 2),
            "incoherent_closest_mrays_with_sdf": round(rays.size / (ms * 1e-3) / 1e6, 2),
            "incoherent_node_fetches_per_ray": round(st["node_visits"] / rays.size, 3),
            "incoherent_tri_fetches_per_ray": round(st["tri_tests"] / rays.size, 3),
            "incoherent_instance_visits_per_ray": round(st["instance_visits"] / rays.size, 3),
            "render": {"width": W, "height": H, "spp": spp, "depth": depth, "seconds": round(dt, 4),
                       "samples_per_s": round(W * H * spp / dt, 1), "rays_traced": int(n_rays),
                       "mrays_per_s": round(n_rays / dt / 1e6, 2)}}


def equal_time_study(args, W=480, H=270, depth=8, ref_spp=4096, budget_spp=64):
    """Equal-time MSE against a converged render of the 4M-triangle wavefront scene: the plain path
    tracer, plain + denoiser, radiance cache, and radiance cache + denoiser, each given the time the
    plain tracer takes for budget_spp samples per pixel."""
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

    log("equal-time study: reference %d spp at %dx%d" % (ref_spp, W, H))
    ref_r = renderer()
    done = 0
    while done < ref_spp:
        ref_r.render(s, W, H, 256)
        done += 256
    ref = ref_r.image()[..., :3].astype(np.float64)

    def mse(img):
        return float(((img[..., :3].astype(np.float64) - ref) ** 2).mean())

    def timed(fn):
        m.synchronize()
        t0 = time.perf_counter()
        out = fn()
        m.synchronize()
        return out, time.perf_counter() - t0

    warm = renderer()
    warm.render(s, W, H, 16)  # allocations and first-launch costs stay out of the timings
    plain = renderer()
    _, budget = timed(lambda: plain.render(s, W, H, budget_spp))
    res = {"width": W, "height": H, "depth": depth, "reference_spp": ref_spp, "budget_s": round(budget, 4),
           "plain": {"spp": budget_spp, "mse": mse(plain.image())}}
    den = None
    if os.path.exists(weights):
        den = m.Denoiser()
        den.load(weights)
        plain.denoise(den)  # warm-up
        img, t_dn = timed(lambda: plain.denoise(den))
        res["plain_denoised"] = {"spp": budget_spp, "mse": mse(img), "denoise_s": round(t_dn, 4)}
    # Radiance cache: the first pass only trains (unbiased samples), later passes end paths in it.
    lo, hi = np.array([-12.5, -7.0, -9.5], np.float32), np.array([12.5, 6.0, 9.5], np.float32)
    cache = m.RadianceCache(seed=1)
    cache.set_bounds(lo, hi)
    cr = renderer()
    cr.set_cache(cache, mode=1)
    spent, spp = 0.0, 0
    t_dn = 0.0 if den is None else res["plain_denoised"]["denoise_s"]
    while True:
        _, dt = timed(lambda: cr.render(s, W, H, 16))
        spent += dt
        spp += 16
        if spp == 16:
            cr.set_cache(cache, mode=2)
        if spent + dt * 0.5 + t_dn > budget:
            break
    res["cache"] = {"spp": spp, "seconds": round(spent, 4), "mse": mse(cr.image())}
    if den is not None:
        img, t = timed(lambda: cr.denoise(den))
        res["cache_denoised"] = {"spp": spp, "seconds": round(spent + t, 4), "mse": mse(img)}
        res["mse_ratio_plain_over_cache_denoised"] = round(res["plain"]["mse"] / res["cache_denoised"]["mse"], 3)
        res["mse_ratio_plain_over_plain_denoised"] = round(res["plain"]["mse"] / res["plain_denoised"]["mse"], 3)
    res["mse_ratio_plain_over_cache"] = round(res["plain"]["mse"] / res["cache"]["mse"], 3)
    s.close()
    return res


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
    ap.add_argument("--only-wavefront", action="store_true",
                    help="run only the 4M-triangle wavefront scene and print its fields (for profiler runs)")
    ap.add_argument("--cpu-sample", type=int, default=1 << 20, help="incoherent rays timed on the host")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel 3 times on BYTES (calibrates rocprof byte counters)")
    args = ap.parse_args()
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
                "bytes_per_ray": round(b_ray, 1), "node_fetches_per_ray": round(n_node, 3),
                "tri_fetches_per_ray": round(n_tri, 3), "kernel_ms": round(ev_ms, 4)}
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

    cpu = None
    if not args.no_cpu_baseline and not args.only_incoherent:
        import oracle
        sample = inc[: args.cpu_sample]
        log("cpu baseline: reference traversal on %d incoherent rays" % sample.size)
        osc = oracle.OracleScene([(idx, pts)], [(0, np.eye(4, dtype=np.float32))])
        rate, used = osc.reference_time(sample, threads=0, chunk=65536, reps=5)
        cpu = {"value": round(rate / 1e6, 4), "unit": "Mrays/s", "cores": used, "kind": "port",
               "sample": "%d incoherent rays of this workload; the reference's stackless bit-stack traversal "
                         "and Moller-Trumbore restated in C++ (oracle/), 65,536-ray jobs on a thread pool, "
                         "median of 5 after one warm-up" % sample.size}

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
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
