"""Training and evaluation of the learned denoiser on pairs the path tracer itself renders.

Data: randomized worlds of the full-feature scene generator (mirror plane, 12 primitives, 4
convex-hull meshes, procedural textures, procedural HDR sky), each scene seed with its own
geometry, sky and textures, and per world random materials, geom positions, camera pose, sky
rotation / intensity and light intensities. Input: one path-traced sample per pixel with its
first-hit albedo, normal and depth; target: a high-sample render of the same world, in log space.
Training runs the U-Net's own matrix-core forward and backward kernels with Adam (L1 loss).

Evaluation compares, on held-out scene seeds, the display image (sRGB8) of
  - 1 sample + learned denoiser, and
  - k samples + a-trous (5 iterations), for the k whose GPU time first reaches the learned
    pipeline's,
against a high-sample reference render (PSNR over RGB, 8-bit values / 255).

  python -m mujoco_warp_pbr_amd.train_denoiser --scenes 32 --worlds 256 --steps 4000

prints one progress line per phase and a JSON summary, and writes the weights (FP32 master copy,
guide-statistics range, training and evaluation record) as safetensors.
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import learned
from mujoco_warp_pbr_amd import scenes

EV100_UNIT = math.log2(1.0 / 1.2)  # exposure 1


def log(msg: str) -> None:
  print(f"[train] {msg}", file=sys.stderr, flush=True)


# ----------------------------------------------------------------------------- scenes


SCENE_KINDS = ("full", "mesh")


def make_scene(kind: str, scene_seed: int, res: int):
  """The full-feature scene (mirror plane, 12 primitives, 4 convex hulls) or the mesh-heavy one
  (64 instances of 16 procedural meshes, ~200k triangles)."""
  if kind == "full":
    return scenes.full_scene(seed=scene_seed, res=res)
  if kind == "mesh":
    return scenes.mesh_scene(seed=scene_seed, res=res)
  raise ValueError(f"scene kind must be one of {SCENE_KINDS}, got {kind!r}")


def scene_context(scene_seed: int, nworld: int, res: int, device, max_depth: int = 4, rr_depth: int = 3, kind: str = "full"):
  """(host model, model, data, context) of `nworld` randomized worlds of one scene seed, path
  traced with first-hit AOVs (denoise=True)."""
  import mujoco_warp_pbr_amd as mj

  mjm = make_scene(kind, scene_seed, res)
  m = mj.put_model(mjm, device)
  d = mj.put_data(mjm, nworld, device)
  rc = mj.create_render_context(
    mjm,
    nworld=nworld,
    cam_res=(res, res),
    envmap=scenes.procedural_sky(scene_seed, 256, 512),
    tier=mj.Tier.PATH,
    path_samples=1,
    max_depth=max_depth,
    rr_depth=rr_depth,
    denoise=True,
    denoise_iterations=5,
    ev100=EV100_UNIT,
    device=device,
  )
  randomize(m, d, rc, nworld, seed=1000003 * (scene_seed + 1))
  mj.refit_bvh(m, d, rc, rebuild=True)
  return mjm, m, d, rc


def randomize(m, d, rc, nworld: int, seed: int) -> None:
  """Per-world materials, geom positions (xy offsets of the non-plane geoms), camera pose, sky
  rotation and intensity, light intensities."""
  rng = np.random.default_rng(seed)
  scenes.randomize_worlds(m, rc, nworld, seed=seed)
  scenes.jitter_cameras(d, seed=seed + 1, offset=0.25, angle_deg=8.0)
  dev = d.geom_xpos.device
  xpos = d.geom_xpos.cpu().numpy()
  xpos[:, 1:, :2] += rng.uniform(-0.35, 0.35, (nworld, xpos.shape[1] - 1, 2)).astype(np.float32)
  d.geom_xpos = torch.as_tensor(xpos, device=dev)
  rc.env_yaw = torch.as_tensor(rng.uniform(0.0, 2.0 * math.pi, nworld).astype(np.float32), device=dev)
  rc.env_intensity = torch.as_tensor(np.exp(rng.normal(0.0, 0.6, nworld)).astype(np.float32), device=dev)
  li = rc.light_intensity.cpu().numpy()
  li = np.broadcast_to(li[:1], (nworld, li.shape[1])) * np.exp(rng.normal(0.0, 0.6, (nworld, li.shape[1])))
  rc.light_intensity = torch.as_tensor(li.astype(np.float32), device=dev)


def render_pair(m, d, rc, nworld: int, target_spp: int, seed: int):
  """(input x BF16 (nworld, h, w, 32), log-space target (nworld * h * w, 3)) of one context."""
  from mujoco_warp_pbr_amd.render import render_path_samples

  rc.path_samples = target_spp
  render_path_samples(m, d, rc, seed=seed + 1)
  target = learned.log_target(rc, nworld)
  rc.path_samples = 1
  render_path_samples(m, d, rc, seed=seed)
  x = learned.pack_input(rc, 0, nworld)
  return x, target


def generate(jobs, res: int, target_spp: int, device):
  """Stacked training pairs (resident in HBM) of jobs [(scene kind, scene seed, worlds), ...]."""
  xs, ts = [], []
  for k, (kind, sd, nworld) in enumerate(jobs):
    t0 = time.perf_counter()
    _, m, d, rc = scene_context(sd, nworld, res, device, kind=kind)
    x, t = render_pair(m, d, rc, nworld, target_spp, seed=17 * sd + 5)
    torch.cuda.synchronize()
    xs.append(x.clone())
    ts.append(t.view(nworld, res * res, 3).clone())
    del rc, m, d
    log(f"scene {k + 1}/{len(jobs)} ({kind} seed {sd}): {nworld} worlds at {target_spp} spp in {time.perf_counter() - t0:.1f} s")
  return torch.cat(xs), torch.cat(ts)


def guide_stats(x: torch.Tensor) -> np.ndarray:
  """Per-image guide statistics (n, 4) of packed inputs."""
  n, h, w, c = x.shape
  stats = torch.zeros((n, learned.NSTATS), dtype=torch.float32, device=x.device)
  _abi.call("mjpbr_unet_guide_stats", _abi.ptr(x), n, h, w, c, 0, _abi.ptr(stats), 0, 0, torch.cuda.current_stream().cuda_stream)
  return stats.cpu().numpy()


# ----------------------------------------------------------------------------- training


def train(net: learned.UNet, X: torch.Tensor, T: torch.Tensor, steps: int, batch: int, lr: float, seed: int = 0, log_every: int = 200):
  """Adam on random batches (with random horizontal flips) of the stacked pairs; cosine
  learning-rate decay to 5% of `lr`. Returns the loss history (every log_every steps)."""
  n, h, w, _ = X.shape
  gen = torch.Generator(device=X.device).manual_seed(seed)
  T4 = T.view(n, h, w, 3)
  history = []
  t0 = time.perf_counter()
  running = None
  for step in range(steps):
    idx = torch.randint(0, n, (batch,), device=X.device, generator=gen)
    xb, tb = X[idx], T4[idx]
    if step % 2 == 1:
      xb, tb = torch.flip(xb, dims=(2,)), torch.flip(tb, dims=(2,))
    xb = xb.contiguous()
    tb = tb.reshape(-1, 3).contiguous()
    net.forward(xb, training=True)
    loss = net.backward(tb)
    frac = step / max(steps - 1, 1)
    net.adam_step(lr=lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * frac))))
    running = loss if running is None else running + loss
    if (step + 1) % log_every == 0 or step + 1 == steps:
      value = float(running) / ((step % log_every) + 1)
      running = None
      history.append((step + 1, value))
      log(f"step {step + 1}/{steps} loss {value:.5f} ({time.perf_counter() - t0:.1f} s)")
  return history


# ----------------------------------------------------------------------------- evaluation


def _rgb(t: torch.Tensor) -> torch.Tensor:
  t = t.to(torch.int64)
  return torch.stack([(t >> s) & 255 for s in (16, 8, 0)], -1).to(torch.float32) / 255.0


def psnr(a: torch.Tensor, b: torch.Tensor) -> float:
  """PSNR in dB of two packed sRGB8 images (int32 bits 0xAARRGGBB), over RGB."""
  mse = torch.mean((_rgb(a).double() - _rgb(b).double()) ** 2).item()
  return 10.0 * math.log10(1.0 / max(mse, 1e-12))


def ssim(a: torch.Tensor, b: torch.Tensor, width: int, height: int) -> float:
  """Mean SSIM (Wang et al. 2004: 11x11 Gaussian window, sigma 1.5, K1 0.01, K2 0.03, range 1)
  over the RGB channels of packed sRGB8 images [nworld, height * width]."""
  x = _rgb(a).reshape(-1, height, width, 3).permute(0, 3, 1, 2).reshape(-1, 1, height, width)
  y = _rgb(b).reshape(-1, height, width, 3).permute(0, 3, 1, 2).reshape(-1, 1, height, width)
  g = torch.exp(-((torch.arange(11, device=x.device, dtype=torch.float32) - 5.0) ** 2) / (2 * 1.5**2))
  g = g / g.sum()
  win = (g[:, None] * g[None, :])[None, None]

  def blur(t):
    return torch.nn.functional.conv2d(t, win)

  mx, my = blur(x), blur(y)
  sxx, syy, sxy = blur(x * x) - mx * mx, blur(y * y) - my * my, blur(x * y) - mx * my
  c1, c2 = 0.01**2, 0.03**2
  m = ((2 * mx * my + c1) * (2 * sxy + c2)) / ((mx * mx + my * my + c1) * (sxx + syy + c2))
  return float(m.mean().item())


def _time_render(m, d, rc, seed: int, reps: int = 3) -> float:
  import mujoco_warp_pbr_amd as mj

  mj.render(m, d, rc, seed=seed)  # warm (the first learned call loads the network)
  st = torch.cuda.current_stream()
  a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  torch.cuda.synchronize()
  a.record(st)
  for _ in range(reps):
    mj.render(m, d, rc, seed=seed)
  b.record(st)
  torch.cuda.synchronize()
  return a.elapsed_time(b) / reps


def timings(scene_seed: int, nworld: int, res: int, device, weights=None, atrous_spp=(1, 2, 3, 4, 6, 8), kind: str = "full") -> dict:
  """GPU milliseconds of each pipeline (path tracing + filter) on `nworld` worlds, and of the
  U-Net forward pass alone."""
  _, m, d, rc = scene_context(scene_seed, nworld, res, device, kind=kind)
  rc.denoiser_weights = str(weights) if weights else None
  out = {"worlds": nworld}
  rc.denoise, rc.denoiser, rc.path_samples = True, "learned", 1
  out["learned_ms"] = _time_render(m, d, rc, seed=7)
  den = learned.denoiser_for(rc)
  x = den.inputs[0]
  st = torch.cuda.current_stream()
  a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  den.net.forward(x)
  a.record(st)
  for _ in range(3):
    den.net.forward(x)
  b.record(st)
  b.synchronize()
  out["unet_ms"] = a.elapsed_time(b) / 3
  out["unet_pixels"] = int(x.shape[0] * x.shape[1] * x.shape[2])
  rc.denoiser = "atrous"
  out["atrous_ms"] = {}
  for k in atrous_spp:
    rc.path_samples = k
    out["atrous_ms"][k] = _time_render(m, d, rc, seed=7)
  return out


def quality(scene_seed: int, nworld: int, res: int, ref_spp: int, device, weights=None, atrous_spp=(1, 2, 3, 4, 6, 8),
            kind: str = "full") -> dict:
  """PSNR and SSIM against a `ref_spp` reference of: 1 sample unfiltered, 1 sample + learned
  denoiser, k samples + a-trous."""
  import mujoco_warp_pbr_amd as mj

  _, m, d, rc = scene_context(scene_seed, nworld, res, device, kind=kind)
  rc.denoiser_weights = str(weights) if weights else None
  rc.denoise, rc.path_samples = False, ref_spp
  mj.render(m, d, rc, seed=99991)
  ref = rc.base.rgb_data.clone()
  row = {"scene_seed": scene_seed, "kind": kind, "worlds": nworld, "ref_spp": ref_spp}

  def score():
    return psnr(rc.base.rgb_data, ref), ssim(rc.base.rgb_data, ref, res, res)

  rc.path_samples = 1
  mj.render(m, d, rc, seed=7)
  row["noisy_1spp"], row["noisy_1spp_ssim"] = score()
  rc.denoise, rc.denoiser = True, "learned"
  mj.render(m, d, rc, seed=7)
  row["learned"], row["learned_ssim"] = score()
  row["fallback_worlds"] = int(learned.denoiser_for(rc).fallback.sum().item())
  rc.denoiser = "atrous"
  row["atrous"], row["atrous_ssim"] = {}, {}
  for k in atrous_spp:
    rc.path_samples = k
    mj.render(m, d, rc, seed=7)
    row["atrous"][k], row["atrous_ssim"][k] = score()
  return row


def evaluate(scene_seeds, nworld: int, res: int, ref_spp: int, device, weights=None, atrous_spp=(1, 2, 3, 4, 6, 8),
             time_worlds: int | None = None, kind: str = "full") -> dict:
  """Quality on `nworld` worlds per held-out scene seed and GPU times on `time_worlds` worlds
  (default nworld) of the first seed; see summarize()."""
  rows = []
  for sd in scene_seeds:
    row = quality(sd, nworld, res, ref_spp, device, weights, atrous_spp, kind=kind)
    rows.append(row)
    log(f"eval seed {sd}: noisy {row['noisy_1spp']:.2f} dB, learned {row['learned']:.2f} dB, "
        + ", ".join(f"atrous@{k} {v:.2f} dB" for k, v in row["atrous"].items()))
  t = timings(scene_seeds[0], time_worlds or nworld, res, device, weights, atrous_spp, kind=kind)
  log(f"timing on {t['worlds']} worlds: learned {t['learned_ms']:.2f} ms (U-Net {t['unet_ms']:.2f} ms), "
      + ", ".join(f"atrous@{k} {v:.2f} ms" for k, v in t["atrous_ms"].items()))
  return summarize(rows, t)


def summarize(rows: list[dict], t: dict) -> dict:
  """Mean PSNRs over the evaluated scenes and the equal-time comparison: a-trous at the smallest
  sample count whose pipeline time reaches the learned pipeline's."""
  ks = sorted(rows[0]["atrous"].keys())
  learned_ms = float(t["learned_ms"])
  at = {k: (float(t["atrous_ms"][k]), float(np.mean([r["atrous"][k] for r in rows]))) for k in ks}
  equal = next((k for k in ks if at[k][0] >= learned_ms), ks[-1])
  flops = learned.flops_per_pixel() * t["unet_pixels"]
  flops_alg = learned.flops_per_pixel(padded=False) * t["unet_pixels"]
  ssim_at = {k: float(np.mean([r["atrous_ssim"][k] for r in rows])) for k in ks}
  out = {
    "kind": rows[0]["kind"],
    "scenes": len(rows),
    "quality_worlds": rows[0]["worlds"],
    "ref_spp": rows[0]["ref_spp"],
    "time_worlds": t["worlds"],
    "noisy_1spp_psnr": float(np.mean([r["noisy_1spp"] for r in rows])),
    "learned_psnr": float(np.mean([r["learned"] for r in rows])),
    "learned_ssim": float(np.mean([r["learned_ssim"] for r in rows])),
    "noisy_1spp_ssim": float(np.mean([r["noisy_1spp_ssim"] for r in rows])),
    "learned_ms": learned_ms,
    "unet_ms": float(t["unet_ms"]),
    "unet_tflops": flops / (t["unet_ms"] * 1e-3) / 1e12,
    "unet_tflops_algorithmic": flops_alg / (t["unet_ms"] * 1e-3) / 1e12,
    "fallback_worlds": int(sum(r["fallback_worlds"] for r in rows)),
    "atrous": {str(k): {"ms": at[k][0], "psnr": at[k][1], "ssim": ssim_at[k]} for k in ks},
    "equal_time_atrous_spp": equal,
    "equal_time_atrous_psnr": at[equal][1],
    "equal_time_atrous_ssim": ssim_at[equal],
    "equal_time_atrous_ms": at[equal][0],
  }
  out["gain_db"] = out["learned_psnr"] - out["equal_time_atrous_psnr"]
  return out


# ----------------------------------------------------------------------------- CLI


def main(argv=None) -> int:
  p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
  p.add_argument("--scenes", type=int, default=16, help="training seeds of the full-feature scene")
  p.add_argument("--worlds", type=int, default=256, help="randomized worlds per full-feature seed")
  p.add_argument("--mesh-scenes", type=int, default=0, help="training seeds of the mesh-heavy scene")
  p.add_argument("--mesh-worlds", type=int, default=128, help="randomized worlds per mesh-heavy seed")
  p.add_argument("--eval-mesh-res", type=int, default=0, help="also evaluate the mesh-heavy scene at this resolution")
  p.add_argument("--eval-mesh-time-worlds", type=int, default=1024)
  p.add_argument("--res", type=int, default=128)
  p.add_argument("--target-spp", type=int, default=256)
  p.add_argument("--steps", type=int, default=3000)
  p.add_argument("--batch", type=int, default=64)
  p.add_argument("--lr", type=float, default=1e-3)
  p.add_argument("--eval-scenes", type=int, default=2)
  p.add_argument("--eval-worlds", type=int, default=128)
  p.add_argument("--ref-spp", type=int, default=1024)
  p.add_argument("--out", default=str(learned.DEFAULT_WEIGHTS))
  p.add_argument("--seed", type=int, default=0)
  args = p.parse_args(argv)

  dev = torch.device("cuda:0")
  t0 = time.perf_counter()
  train_seeds = list(range(100, 100 + args.scenes))
  mesh_seeds = list(range(300, 300 + args.mesh_scenes))
  jobs = [("full", sd, args.worlds) for sd in train_seeds] + [("mesh", sd, args.mesh_worlds) for sd in mesh_seeds]
  # interleave the kinds so an early stop still trains on both
  jobs.sort(key=lambda j: (j[1] - (100 if j[0] == "full" else 300)) / max(args.scenes if j[0] == "full" else args.mesh_scenes, 1))
  log(f"generating {len(jobs)} scene seeds of pairs at {args.res}^2 ({args.target_spp}-spp targets)")
  X, T = generate(jobs, args.res, args.target_spp, dev)
  gen_s = time.perf_counter() - t0
  stats = guide_stats(X)
  net = learned.UNet(dev, seed=args.seed)
  log(f"training {net.parameter_count()} parameters for {args.steps} steps, batch {args.batch}")
  t1 = time.perf_counter()
  history = train(net, X, T, args.steps, args.batch, args.lr, seed=args.seed)
  torch.cuda.synchronize()
  train_s = time.perf_counter() - t1
  meta = {
    "guide_range": learned.guide_range_from_stats(stats),
    "train": {
      "full_scene_seeds": [train_seeds[0], train_seeds[-1]] if train_seeds else [],
      "worlds_per_scene": args.worlds,
      "mesh_scene_seeds": [mesh_seeds[0], mesh_seeds[-1]] if mesh_seeds else [],
      "worlds_per_mesh_scene": args.mesh_worlds,
      "images": int(X.shape[0]),
      "res": args.res,
      "target_spp": args.target_spp,
      "steps": args.steps,
      "batch": args.batch,
      "lr": args.lr,
      "loss": history,
      "generate_s": gen_s,
      "train_s": train_s,
    },
  }
  learned.save(net, args.out, meta)
  log(f"saved {args.out}")
  del X, T
  torch.cuda.empty_cache()
  result = {"train": meta["train"]}
  if args.eval_scenes > 0:
    eval_seeds = list(range(900, 900 + args.eval_scenes))
    log(f"evaluating on scene seeds {eval_seeds} ({args.eval_worlds} worlds, {args.ref_spp}-spp reference)")
    result["eval"] = evaluate(eval_seeds, args.eval_worlds, args.res, args.ref_spp, dev, weights=args.out)
    meta["eval"] = result["eval"]
    learned.save(net, args.out, meta)
  if args.eval_mesh_res > 0:
    log(f"evaluating the mesh-heavy scene at {args.eval_mesh_res}^2")
    result["eval_mesh"] = evaluate([950], 32, args.eval_mesh_res, args.ref_spp, dev, weights=args.out,
                                   time_worlds=args.eval_mesh_time_worlds, kind="mesh")
    meta["eval_mesh"] = result["eval_mesh"]
    learned.save(net, args.out, meta)
  print(json.dumps(result), flush=True)
  return 0


if __name__ == "__main__":
  sys.exit(main())
