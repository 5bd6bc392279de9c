"""Trains the denoiser on (noisy, converged) image pairs rendered by this path tracer.

Each seeded training scene (myrt_amd.scenes.training_scene) is rendered at RES x RES, depth 8:
noisy inputs at 1, 2, 4, 8, 16, 32 and 64 samples per pixel (radiance plus primary-hit albedo, normal and
distance), and a converged target at TARGET_SPP. The network is trained with Adam on random
128 x 128 crops (flips and transposes as augmentation; input levels drawn in proportion to
sqrt(spp) by default) by the library's own fp32 forward and backward kernels, and the weights are
saved as a float32 .npy of the packed parameters.

Usage:
  python -m myrt_amd.train_denoiser [--scenes 48] [--val 4] [--res 256] [--target-spp 1024]
                                    [--steps 3000] [--batch 8] [--init weights.npy]
                                    [--out myrt_amd/data/denoiser_unet.npy]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import myrt_amd as m
from myrt_amd import scenes

INPUT_SPP = (1, 2, 4, 8, 16, 32, 64)
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "denoiser_unet.npy")


def log(msg):
    print("[train_denoiser] " + msg, flush=True)


def build_scene(sc):
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for mat in sc["materials"]:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    s.set_environment(sc.get("env", 1.0))
    for g, mt, mi in sc["instances"]:
        s.enqueue(gids[g], mt, mi)
    if "sdf_tape" in sc:
        s.enqueue_sdf(s.push_sdf(sc["sdf_tape"]), sc["sdf_transform"])
    return s


def render_pairs(sc, res, target_spp, depth=8):
    """Raw features at INPUT_SPP samples (len(INPUT_SPP), res, res, 10) and the target (res, res, 3)."""
    s = build_scene(sc)
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], 1.0)

    def renderer():
        r = m.Renderer()
        r.set_camera(view, proj)
        r.set_num_bounces(depth)
        return r

    r = renderer()
    inputs, done = [], 0
    for spp in INPUT_SPP:
        r.render(s, res, res, spp - done)
        done = spp
        inputs.append(r.raw_features())
    t = renderer()
    left = target_spp
    while left > 0:
        t.render(s, res, res, min(256, left))
        left -= min(256, left)
    target = t.image()[..., :3]
    s.close()
    return np.stack(inputs).astype(np.float32), target.astype(np.float32)


def augment(raw, tgt, rng):
    if rng.random() < 0.5:
        raw, tgt = raw[:, ::-1], tgt[:, ::-1]
    if rng.random() < 0.5:
        raw, tgt = raw[::-1], tgt[::-1]
    if rng.random() < 0.5:
        raw, tgt = raw.transpose(1, 0, 2), tgt.transpose(1, 0, 2)
    return raw, tgt


def mse(a, b):
    return float(((a.astype(np.float64) - b) ** 2).mean())


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=48)
    ap.add_argument("--val", type=int, default=4)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--level-weight", choices=("uniform", "sqrt"), default="sqrt",
                    help="how often each input spp level is drawn: uniformly, or in proportion to "
                         "sqrt(spp) so that the smaller errors of the cleaner inputs weigh as much "
                         "in the L1 loss as the noisy ones")
    ap.add_argument("--init", default="",
                    help="start from these packed parameters (.npy) instead of a seeded initialisation "
                         "(fine-tuning; Adam's moments start at zero)")
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--max-seconds", type=float, default=0.0,
                    help="wall-clock budget: rendering stops at 60 %% of it (fewer training scenes) and "
                         "training steps shrink to fit the rest; 0 = no limit")
    args = ap.parse_args(argv)

    t0 = time.time()
    data = []
    # Validation scenes first (seeds 0 .. val-1, never trained on), then training scenes.
    val = []
    for k in range(args.val):
        inputs, target = render_pairs(scenes.training_scene(k), args.res, args.target_spp)
        val.append((inputs.astype(np.float16), target))
    for k in range(args.scenes):
        if args.max_seconds and time.time() - t0 > 0.6 * args.max_seconds:
            log("render budget spent: %d training scenes" % k)
            break
        sc = scenes.training_scene(args.val + k)
        inputs, target = render_pairs(sc, args.res, args.target_spp)
        data.append((inputs.astype(np.float16), target))  # float16 inputs: half the host memory
        if k % 8 == 7 or k == args.scenes - 1:
            log("rendered %d / %d scenes (%.0f s)" % (k + 1, args.scenes, time.time() - t0))
    train = data

    d = m.Denoiser(seed=args.seed)
    if args.init:
        d.load(args.init)
        log("initialised from %s" % args.init)
    rng = np.random.default_rng(args.seed)
    level_p = np.sqrt(np.asarray(INPUT_SPP, np.float64)) if args.level_weight == "sqrt" else np.ones(len(INPUT_SPP))
    level_p /= level_p.sum()
    B, C, R = args.batch, args.crop, args.res
    losses = []
    t1 = time.time()
    steps = args.steps
    for step in range(args.steps):
        if step == 200 and args.max_seconds:
            per_step = (time.time() - t1) / 200
            left = args.max_seconds - (time.time() - t0) - 60.0  # validation and saving
            steps = max(400, min(args.steps, int(left / per_step)))
            log("%.1f ms per step: training %d steps" % (per_step * 1e3, steps))
        if step >= steps:
            break
        lr = args.lr * (0.02 + 0.98 * 0.5 * (1 + np.cos(np.pi * step / steps)))  # cosine decay to 2 %
        if step % 50 == 0:
            d.set_adam(lr=float(lr))
        raw = np.empty((B, C, C, 10), np.float32)
        tgt = np.empty((B, C, C, 3), np.float32)
        for b in range(B):
            inputs, target = train[int(rng.integers(len(train)))]
            lvl = int(rng.choice(len(INPUT_SPP), p=level_p))
            y, x = int(rng.integers(R - C + 1)), int(rng.integers(R - C + 1))
            rr, tt = augment(inputs[lvl, y:y + C, x:x + C].astype(np.float32), target[y:y + C, x:x + C], rng)
            raw[b], tgt[b] = rr, tt
        losses.append(d.train(raw, tgt))
        if step % 250 == 249 or step == steps - 1:
            log("step %d loss %.4f lr %.2e (%.0f s)" % (step + 1, float(np.mean(losses[-250:])), lr, time.time() - t1))

    # Validation: MSE of the noisy input and of the denoised image against the target, per spp.
    report = {}
    for i, spp in enumerate(INPUT_SPP):
        noisy = [mse(inp[i, ..., :3].astype(np.float32), tgt) for inp, tgt in val]
        den = [mse(d.denoise(inp[i].astype(np.float32))[..., :3], tgt) for inp, tgt in val]
        report[spp] = {"noisy_mse": float(np.mean(noisy)), "denoised_mse": float(np.mean(den)),
                       "gain": float(np.mean(noisy) / max(np.mean(den), 1e-12))}
        log("validation %2d spp: noisy MSE %.4e, denoised %.4e (x%.1f)" % (
            spp, report[spp]["noisy_mse"], report[spp]["denoised_mse"], report[spp]["gain"]))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    d.save(args.out)
    meta = {"args": vars(args), "train_scenes": len(train), "train_seeds": [args.val, args.val + len(train) - 1],
            "validation_seeds": [0, args.val - 1], "steps": steps, "input_spp": INPUT_SPP,
            "final_loss": float(np.mean(losses[-100:])),
            "validation": report, "render_s": round(t1 - t0, 1), "train_s": round(time.time() - t1, 1)}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log("saved %s" % args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
