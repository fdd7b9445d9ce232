"""Benchmark: 3D LIC volume + ray-cast frame on one MI355X.

Workload (the "plain method at full size" configuration):
  field   256^3 float32, 12 seeded Rankine vortices + uniform flow (0.1,0,0)
  noise   dense seeded uniform uint8, 256^3
  LIC     Gaussian filter (sigma = L/4), L = F = B = 40, step 0.005
  render  1024^2 ray cast, sample distance 1/256, vector-length TF
One step = one LIC-volume computation + one frame, inputs resident in HBM.

Prints one JSON line.  `value` is LIC voxel-samples per second over the
whole step (LIC + frame); the components, the FP16-storage variant and the
forward+adjoint time are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--image", type=int, default=1024)
    ap.add_argument("--L", type=int, default=40)
    ap.add_argument("--storage", choices=["fp32", "fp16", "unorm8"], default="fp32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip FP16 and adjoint legs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched timesteps x views leg")
    ap.add_argument("--no-target", action="store_true",
                    help="skip the FP64 check of the 256^3 LIC + 1024^2 image + adjoint")
    ap.add_argument("--no-precision", action="store_true", help="skip the 512^3 FP16-storage leg")
    ap.add_argument("--lic-brick", action="store_true",
                    help="FP32: the LIC volume kernel reads the 8^3-brick copy of the field (tuning lic.brick)")
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only 3 LIC launches and 3 copies of 256 MiB (rocprof counter runs)")
    ap.add_argument("--pmc-probe-fused", action="store_true",
                    help="only 3 fused on-the-fly LIC ray casts of the image (rocprof counter runs)")
    ap.add_argument("--pmc-probe-frame", action="store_true",
                    help="only 3 frames and 3 TF/opacity adjoints of the image with the automatic lanes per "
                         "ray, then the same with one lane per pixel (rocprof counter runs)")
    return ap.parse_args()


def lic_bytes_per_voxel(F: int, B: int, field_bytes: int = 12, noise_bytes: int = 1,
                        out_bytes: int = 4) -> int:
    """Algorithmic bytes of one LIC voxel: (1 + 2(F+B)) trilinear field fetches
    and (F+B+1) noise fetches of 8 texels each, plus the output."""
    return (1 + 2 * (F + B)) * 8 * field_bytes + (F + B + 1) * 8 * noise_bytes + out_bytes


def _profile_value(name: str, key: str):
    """A number from a committed rocprof summary line '<key>: <value>', if any."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1].split()[0].rstrip(","))
            except (ValueError, IndexError):
                return None
    return None


MI355X_CUS, MI355X_CLOCK_HZ = 256, 2.4e9
FP32_PEAK_TFLOPS = 157.3  # published MI355X FP32 figure (counts packed FMA)


def storage_kernel(vv, storage) -> str:
    return {vv.STORE_FP32: "k_lic_cached<float4>", vv.STORE_FP16: "k_lic_cached<half4_t>",
            vv.STORE_UNORM8: "k_lic_cached<u8x4_t>"}[storage]


def roofline(tag: str, kernel: str, ms: float, fetch_bytes: float) -> dict:
    """Roofline of the LIC kernel from its launch time `ms` (HIP events on the
    library's stream, in this run) and the per-launch counts of the committed
    rocprof counter summary of the same kernel and workload
    (profiles/<tag>_pmc.txt).

    The bound that holds is FP32 VALU issue: the fma/add/mul wave64
    instructions of a launch (SQ_INSTS_VALU_{FMA,ADD,MUL}_F32) issue at most
    once per clock per CU (the 157.3 TFLOP/s FP32 figure is 256 CUs x 2.4 GHz
    x 64 lanes x 2 flops x 2 for packed FP32, which the library's kernels do not
    emit), so `frac` = instructions / (launch clocks x 256 CUs).  Beside it:
    the FP32 flop rate against 157.3 TFLOP/s, HBM bytes (FETCH_SIZE +
    WRITE_SIZE, calibrated on a copy of known size) against 8 TB/s, L1
    cache-line accesses per clock per CU, L2 request bandwidth, and the
    texture-fetch bytes the reference's shader requests: those are served by
    the per-lane cell caches, L1 and L2, so they are a gather volume, not an
    HBM fraction."""
    name = f"{tag}_pmc.txt"
    sec = ms / 1000.0
    clocks = sec * MI355X_CUS * MI355X_CLOCK_HZ
    valu = _profile_value(name, "valu_f32_insts_per_launch")
    fma = _profile_value(name, "SQ_INSTS_VALU_FMA_F32 per launch")
    add = _profile_value(name, "SQ_INSTS_VALU_ADD_F32 per launch")
    mul = _profile_value(name, "SQ_INSTS_VALU_MUL_F32 per launch")
    traffic = _profile_value(name, "traffic_bytes_per_launch")
    l1 = _profile_value(name, "TCP_TOTAL_CACHE_ACCESSES per launch")
    l2 = _profile_value(name, "TCC requests per launch")
    rate = None if valu is None else valu / clocks
    r = {
        "bound": "valu_f32",
        "achieved": rate,
        "peak": 1.0,
        "unit": "wave64 FP32 fma/add/mul instructions per clock per CU",
        "frac": None if rate is None else min(rate, 1.0),
        "traffic": traffic,
        "kernel": kernel,
        "launch_ms": ms,
        "counts_from": f"profiles/{name}",
    }
    if None not in (fma, add, mul):
        tflops = (2.0 * fma + add + mul) * 64.0 / sec / 1e12
        r["flops"] = {"achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": tflops / FP32_PEAK_TFLOPS}
    if traffic is not None:
        gbs = traffic / sec / 1e9
        r["hbm"] = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                    "bytes_per_launch": traffic}
    if l1 is not None:
        r["l1_gather"] = {"accesses_per_launch": l1, "achieved": l1 / clocks, "peak": 1.0,
                          "unit": "cache-line accesses per clock per CU", "frac": min(l1 / clocks, 1.0)}
    if l2 is not None:
        l2_gbs = l2 * 128.0 / sec / 1e9
        l2_peak = 8 * 16 * 128.0 * MI355X_CLOCK_HZ / 1e9
        r["l2"] = {"requests_per_launch": l2, "achieved": l2_gbs, "peak": l2_peak, "unit": "GB/s",
                   "frac": l2_gbs / l2_peak, "peak_basis": "16 channels x 128 B/clk per XCD, 8 XCDs"}
    r["texture_fetch"] = {"bytes_per_launch": fetch_bytes, "rate_gbs": fetch_bytes / sec / 1e9,
                          "served_by": "per-lane cell caches, L1 and L2 (not HBM)"}
    return r



def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the engine runs on one GPU; --gpus must be 1 "
                          "(z-slab sharding of the LIC and image-tile sharding of the ray cast are not run)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import vectorvisualization_amd as vv
    from vectorvisualization_amd import fields

    n, L, W = args.n, args.L, args.image
    dims = (n, n, n)
    dev = torch.device("cuda:0")
    storage = {"fp32": vv.STORE_FP32, "fp16": vv.STORE_FP16, "unorm8": vv.STORE_UNORM8}[args.storage]

    print(f"[bench] generating {n}^3 Rankine field on {dev}", file=sys.stderr, flush=True)
    field = fields.rankine_vortices_torch(n, device=dev).contiguous()
    ctx = vv.Context(0)
    if args.lic_brick:
        ctx.set_tuning("lic.brick", 1)
    ctx.upload_field(field, dims, storage=storage)
    ctx.generate_noise(1, vv.NOISE_UNIFORM, 0.0, dims)
    lut, inv = vv.make_filter(vv.FILTER_GAUSSIAN, L)
    ctx.set_filter(lut, inv)
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = L
    p.step_size_lic = 0.005
    p.step_size_vol = 1.0 / 256.0
    p.gradient_scale = 3.0
    p.tf_index = vv.TF_INDEX_MAG
    ctx.set_lic_params(p)
    rgba, op = fields.tf_magnitude()
    ctx.set_tf(rgba, op)
    ext, sc, cen = ctx.field_geometry()
    cam = vv.default_camera(W, W, cen)
    image = torch.empty(W * W * 4, dtype=torch.float32, device=dev)

    stream = torch.cuda.ExternalStream(ctx.stream)
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    if args.pmc_probe:
        # counter runs: the LIC kernel, then a plain copy of known size whose
        # FETCH_SIZE / WRITE_SIZE calibrate the counters (256 MiB each way)
        for _ in range(3):
            ctx.compute_lic()
        ctx.synchronize()
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        ctx.close()
        print("[bench] pmc probe done", file=sys.stderr, flush=True)
        return

    if args.pmc_probe_fused:
        # counter runs of the fused caster: every sample integrates its own
        # streamline (no LIC volume), same field, filter and TF as the step
        for _ in range(3):
            ctx.render_fused(cam, out=image)
        ctx.synchronize()
        ctx.close()
        print("[bench] fused pmc probe done", file=sys.stderr, flush=True)
        return

    if args.pmc_probe_frame:
        # counter runs of the ray caster and its adjoint on this step's LIC
        # volume, with the adjoint's dL/dimage and opacity-scale gradient of
        # the pipeline leg
        ctx.compute_lic()
        dimg = torch.randn(W * W * 4, dtype=torch.float32, device=dev) * 1e-3
        dtf = torch.empty(1024, dtype=torch.float32, device=dev)
        dsig = torch.empty(n ** 3, dtype=torch.float32, device=dev)
        # the automatic lanes per ray, then one lane per pixel for comparison
        for G in (0, 1):
            ctx.set_tuning("render.volume_group", G)
            for _ in range(3):
                ctx.render(cam, out=image)
            for _ in range(3):
                ctx.render_adjoint(cam, dimg, dtf, dsig)
        ctx.set_tuning("render.volume_group", 0)
        ctx.synchronize()
        ctx.close()
        print("[bench] frame pmc probe done", file=sys.stderr, flush=True)
        return

    print("[bench] warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ctx.compute_lic()
        ctx.render(cam, out=image)
    _, ray_samples = ctx.render(cam, out=image, count_samples=True)
    torch.cuda.synchronize()

    # ---- timed region: K steps of (LIC volume + frame) ----
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    marks = [(ev(), ev(), ev()) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b, c in marks:
        a.record(stream)
        ctx.compute_lic()
        b.record(stream)
        ctx.render(cam, out=image)
        c.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lic_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in marks]))
    frame_ms = float(np.mean([b.elapsed_time(c) for _, b, c in marks]))
    step_ms = float(np.mean([a.elapsed_time(c) for a, _, c in marks]))
    wall_ms = wall * 1000.0 / args.steps

    vox = n ** 3
    samples = vox * (2 * L + 1)
    value = samples / (step_ms / 1000.0)
    # bytes of the three direction channels of one texel
    bpv = lic_bytes_per_voxel(L, L, {vv.STORE_FP32: 12, vv.STORE_FP16: 6, vv.STORE_UNORM8: 3}[storage])
    tag = f"lic_rankine{n}_gauss{L}_{args.storage}"

    out = {
        "metric": "LIC voxel-samples/s + ms/frame (256^3, 1024^2 view); fwd+adjoint vs plain GPU",
        "value": value,
        "unit": "voxel-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {vv.STORE_FP32: "f32", vv.STORE_FP16: "f16-storage/f32",
                  vv.STORE_UNORM8: "unorm8-storage/f32"}[storage],
        "data": "synthetic",
        "config": {
            "workload": f"LIC {n}^3 Rankine x12 + (0.1,0,0), dense uniform noise, Gaussian L={L} "
                        f"h=0.005; ray cast {W}^2 step 1/256 vector-length TF",
            "field": f"{n}^3 {args.storage}",
            "image": f"{W}x{W}",
            "steps_fwd_bwd": L,
        },
        "lic_ms": lic_ms,
        "frame_ms": frame_ms,
        "wall_ms_per_step": wall_ms,
        "lic_voxel_samples_per_s": samples / (lic_ms / 1000.0),
        "ray_samples_per_frame": int(ray_samples),
        "ray_samples_per_s": ray_samples / (frame_ms / 1000.0),
        "roofline": roofline(tag, storage_kernel(vv, storage), lic_ms, bpv * vox),
    }

    if not args.no_extras:
        def timed(fn, reps=3):
            fn()
            torch.cuda.synchronize()
            best = None
            for _ in range(reps):
                a, b = ev(), ev()
                a.record(stream)
                fn()
                b.record(stream)
                torch.cuda.synchronize()
                t = a.elapsed_time(b)
                best = t if best is None else min(best, t)
            return best

        dimg = torch.randn(W * W * 4, dtype=torch.float32, device=dev) * 1e-3
        dtf = torch.empty(1024, dtype=torch.float32, device=dev)
        dsig = torch.empty(vox, dtype=torch.float32, device=dev)
        legs = {}
        for name, st in (("fp32", vv.STORE_FP32), ("fp16", vv.STORE_FP16), ("unorm8", vv.STORE_UNORM8)):
            ctx.upload_field(field, dims, storage=st)
            l_ms = timed(ctx.compute_lic)
            f_ms = timed(lambda: ctx.render(cam, out=image))
            a_ms = timed(lambda: ctx.render_adjoint(cam, dimg, dtf, dsig))

            # a training step's render + backward: vv_render_record, then the
            # adjoint reusing its forward pass (what the autograd shim runs)
            def rec(ds):
                ctx.render_record(cam, out=image)
                ctx.render_adjoint(cam, dimg, dtf, ds)
            r_ms = timed(lambda: rec(dsig))
            r_tf_ms = timed(lambda: rec(None))
            legs[name] = dict(lic_ms=l_ms, frame_ms=f_ms, adjoint_ms=a_ms, fwd_adjoint_recorded_ms=r_ms,
                              fwd_adjoint_recorded_tf_only_ms=r_tf_ms, pipeline_ms=l_ms + r_ms)
        # the one-lane-per-voxel FP32 kernel without cell caches (information
        # only: the library runs the cached kernel for this workload)
        ctx.upload_field(field, dims, storage=vv.STORE_FP32)
        ctx.set_tuning("lic.cache", 0)
        uncached_lic = timed(ctx.compute_lic)
        ctx.set_tuning("lic.cache", 1)
        # the LIC-volume caster on the default view and three seeded
        # rotations: lanes per ray chosen automatically against one lane per
        # pixel (same images and sample counts)
        ctx.compute_lic()
        rng = np.random.default_rng(7)
        quats = [(0.0, 0.0, 0.0, 1.0)]
        for _ in range(3):
            q = rng.normal(size=4)
            quats.append(tuple(float(x) for x in q / np.linalg.norm(q)))
        views = []
        for q in quats:
            vcam = vv.default_camera(W, W, cen)
            for k in range(4):
                vcam.quat[k] = q[k]
            _, vns = ctx.render(vcam, out=image, count_samples=True)

            def rec_adj(ds):
                ctx.render_record(vcam, out=image)
                ctx.render_adjoint(vcam, dimg, dtf, ds)
            row = dict(quat=[round(x, 4) for x in q], ray_samples=int(vns))
            for tag, G in (("", 0), ("one_lane_per_pixel_", 1)):
                ctx.set_tuning("render.volume_group", G)
                row[f"{tag}ms"] = timed(lambda: ctx.render(vcam, out=image))
                row[f"{tag}record_ms"] = timed(lambda: ctx.render_record(vcam, out=image))
                row[f"{tag}record_tf_adjoint_ms"] = timed(lambda: rec_adj(None))
                row[f"{tag}record_adjoint_ms"] = timed(lambda: rec_adj(dsig))
            ctx.set_tuning("render.volume_group", 0)
            views.append(row)
        out["frame_views"] = {
            "what": f"LIC-volume ray cast at {W}^2 (FP32 field): the default view, then three seeded "
                    "rotations; frame, recorded frame (vv_render_record), recorded frame + TF adjoint and "
                    "recorded frame + TF and opacity-scale adjoint in ms, with the automatic lanes per ray "
                    "and with one lane per pixel",
            "views": views,
        }
        out["pipeline"] = {
            "what": "LIC volume + 1024^2 forward frame + TF/opacity adjoint (ms); adjoint_ms is a stand-alone "
                    "adjoint (its own forward walk), fwd_adjoint_recorded_ms a vv_render_record frame followed by "
                    "the adjoint that reuses it (pipeline_ms = lic_ms + that); speedups against the FP32-stored "
                    "path the library runs",
            "fp32": legs["fp32"],
            "fp16": legs["fp16"],
            "unorm8": legs["unorm8"],
            "fp16_speedup_vs_fp32": legs["fp32"]["pipeline_ms"] / legs["fp16"]["pipeline_ms"],
            "fp16_lic_speedup_vs_fp32": legs["fp32"]["lic_ms"] / legs["fp16"]["lic_ms"],
            "unorm8_speedup_vs_fp32": legs["fp32"]["pipeline_ms"] / legs["unorm8"]["pipeline_ms"],
            "fwd_plus_adjoint_over_frame": {k: legs[k]["fwd_adjoint_recorded_ms"] / legs[k]["frame_ms"]
                                            for k in legs},
            "fwd_plus_tf_adjoint_over_frame": {k: legs[k]["fwd_adjoint_recorded_tf_only_ms"] / legs[k]["frame_ms"]
                                               for k in legs},
            "fp32_uncached_lic_ms": uncached_lic,
        }
        out["fwd_plus_adjoint_ms"] = legs[args.storage]["fwd_adjoint_recorded_ms"]
        # the fused caster (lic3d_fragment.glsl): LIC of every ray sample on the fly
        fused = {}
        for name, st in (("fp32", vv.STORE_FP32), ("fp16", vv.STORE_FP16)):
            ctx.upload_field(field, dims, storage=st)
            for w in (256, W):
                fcam = vv.default_camera(w, w, cen)
                fimg = torch.empty(w * w * 4, dtype=torch.float32, device=dev)
                _, fns = ctx.render_fused(fcam, out=fimg, count_samples=True)
                f_ms = timed(lambda: ctx.render_fused(fcam, out=fimg))
                fused[f"{name}_{w}"] = dict(ms=f_ms, ray_samples=int(fns),
                                            lic_samples_per_s=fns * (2 * L + 1) / (f_ms / 1000.0))
                _, sns = ctx.render_slices(fcam, out=fimg, count_samples=True)
                s_ms = timed(lambda: ctx.render_slices(fcam, out=fimg))
                fused[f"slices_{name}_{w}"] = dict(ms=s_ms, samples=int(sns))
        out["fused"] = {
            "what": "on-the-fly LIC casters (fused ray caster and view-aligned slicing): every "
                    "sample integrates its own streamline (no LIC volume), same field/filter/TF, "
                    "square views; lanes per ray chosen automatically",
            **fused,
            "lic_volume_plus_frame_ms_fp32": legs["fp32"]["lic_ms"] + legs["fp32"]["frame_ms"],
        }
        ctx.upload_field(field, dims, storage=storage)
        ctx.compute_lic()

    if not args.no_batch:
        out["batch"] = batch_leg(args, dev)
        if args.storage == "fp32":
            fp16_args = argparse.Namespace(**{**vars(args), "storage": "fp16"})
            b16 = batch_leg(fp16_args, dev)
            out["batch"]["fp16"] = {k: b16[k] for k in ("batch_ms", "view_latency_ms_p50", "view_latency_ms_p99",
                                                        "aggregate_samples_per_s")}

    if not args.no_precision:
        out["precision"] = precision_leg(dev)

    if not args.no_target and n == 256 and W == 1024:
        out["fp64_check"] = fp64_check_leg(dev, field, L)
        out["tf_fit"] = tf_fit_leg(dev, field, L)

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(field.cpu().numpy(), dims, L, lut, inv, p,
                                           args.cpu_seconds, samples / vox)
    ctx.close()
    print(json.dumps(out), flush=True)


def batch_leg(args, dev, seed=7, nsteps=8, nviews=64, L=30):
    """8 timesteps of the rotating 3D double gyre with sizes drawn from
    {128,192,256,384}^3, 64 views with square resolutions uniform in
    256..1024 on seeded orbit poses, triangle filter L = 30: every LIC volume
    and every view in one captured launch (autotuned).  Compared with the
    same work issued call by call through the single-volume API."""
    import numpy as np
    import torch

    import vectorvisualization_amd as vv
    from vectorvisualization_amd import fields

    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.choice([128, 192, 256, 384], size=nsteps)]
    views = []
    for _ in range(nviews):
        w = int(rng.integers(256, 1025))
        views.append((int(rng.integers(nsteps)), w, float(rng.uniform(0, 2 * np.pi)),
                      float(rng.uniform(-0.6, 0.6))))
    print(f"[bench] batch: timesteps {sizes}, {nviews} views", file=sys.stderr, flush=True)
    storage = {"fp32": vv.STORE_FP32, "fp16": vv.STORE_FP16, "unorm8": vv.STORE_UNORM8}[args.storage]
    ctx = vv.Context(0)
    lut, inv = vv.make_filter(vv.FILTER_TRIANGLE, L)
    ctx.set_filter(lut, inv)
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = L
    p.step_size_lic = 0.005
    p.step_size_vol = 1.0 / 256.0
    p.gradient_scale = 3.0
    p.tf_index = vv.TF_INDEX_MAG
    ctx.set_lic_params(p)
    ctx.set_tf(*fields.tf_magnitude())
    b = vv.Batch(ctx)
    for i, n in enumerate(sizes):
        f = fields.double_gyre_3d_torch(n, 1.25 * i, device=dev)
        b.add_timestep(f, (n, n, n), storage=storage, noise_seed=100 + i, noise_kind=vv.NOISE_UNIFORM)
        del f
    images = []
    for t, w, yaw, pitch in views:
        img = torch.empty(w * w * 4, dtype=torch.float32, device=dev)
        b.add_view(t, fields.orbit_camera(w, w, (0.5, 0.5, 0.5), yaw, pitch), img)
        images.append(img)
    t = time.perf_counter()
    b.autotune()
    tune_s = time.perf_counter() - t
    b.run()
    ctx.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream)
    run_ms, lat, span = [], [], []
    for _ in range(max(1, args.batch_reps)):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        b.run()
        e1.record(stream)
        torch.cuda.synchronize()
        run_ms.append(e0.elapsed_time(e1))
        lat.append(b.view_latency_ms())
        span.append(b.view_span_ms())
    lat = np.concatenate(lat)
    span = np.concatenate(span)
    ray = b.samples()
    lic = sum(n ** 3 * (2 * L + 1) for n in sizes)
    batch_ms = float(np.median(run_ms))
    tuning = {f"{n}^3": b.tuning(i)[:6] for i, n in enumerate(sizes)}
    render_tile = b.tuning(0)[6:]
    b.close()

    # the same work call by call: per timestep one vv_compute_lic and its views' vv_render
    seq_ms = 0.0
    for i, n in enumerate(sizes):
        f = fields.double_gyre_3d_torch(n, 1.25 * i, device=dev)
        ctx.upload_field(f, (n, n, n), storage=storage)
        ctx.generate_noise(100 + i, vv.NOISE_UNIFORM, 0.0, (n, n, n))
        del f
        mine = [(k, v) for k, v in enumerate(views) if v[0] == i]

        def work():
            ctx.compute_lic()
            for k, (_, w, yaw, pitch) in mine:
                ctx.render(fields.orbit_camera(w, w, (0.5, 0.5, 0.5), yaw, pitch), out=images[k])

        work()
        ctx.synchronize()
        t = time.perf_counter()
        work()
        ctx.synchronize()
        seq_ms += (time.perf_counter() - t) * 1000.0
    ctx.close()
    return {
        "workload": f"{nsteps} double-gyre timesteps {sizes}, {nviews} views 256^2..1024^2, "
                    f"triangle L={L}, {args.storage}",
        "batch_ms": batch_ms,
        "view_latency_ms_p50": float(np.percentile(lat, 50)),
        "view_latency_ms_p99": float(np.percentile(lat, 99)),
        "view_latency_note": "from the start of the batch to the view's last tile (includes the LIC volumes and "
                             "views queued before it)",
        "view_span_ms_p50": float(np.percentile(span, 50)),
        "view_span_ms_p99": float(np.percentile(span, 99)),
        "view_span_note": "the view's own render: its first tile's start to its last tile's end",
        "lic_voxel_samples": int(lic),
        "ray_samples": int(ray),
        "aggregate_samples_per_s": (lic + ray) / (batch_ms / 1000.0),
        "call_by_call_ms": seq_ms,
        "speedup_vs_call_by_call": seq_ms / batch_ms,
        "autotune_s": tune_s,
        "lic_tile_and_brick": tuning,
        "render_tile": render_tile,
    }


def precision_leg(dev, n=512, L=60, box=64, W=1024):
    """512^3 vortex field stored as 3xFP16 (noise uint8, 4% sparse), cos^2
    filter L = 60, FP32 integration: LIC volume and 1024^2 image against the
    FP32-stored run (whole volume / image), and the LIC of a central 64^3
    subvolume against the FP64 oracle."""
    import numpy as np
    import torch

    import vectorvisualization_amd as vv
    from vectorvisualization_amd import fields
    from oracle import oracle as orc

    print(f"[bench] precision leg: {n}^3 fp16 vs fp32 vs FP64 oracle ({box}^3)", file=sys.stderr, flush=True)
    dims = (n, n, n)
    field = fields.rankine_vortices_torch(n, device=dev).contiguous()
    ctx = vv.Context(0)
    lut, inv = vv.make_filter(vv.FILTER_COS2, L)
    ctx.set_filter(lut, inv)
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = L
    p.step_size_lic = 0.005
    p.step_size_vol = 1.0 / 256.0
    p.gradient_scale = 3.0
    p.tf_index = vv.TF_INDEX_MAG
    ctx.set_lic_params(p)
    ctx.set_tf(*fields.tf_magnitude())
    ctx.generate_noise(1, vv.NOISE_BINARY, 0.04, dims)
    stream = torch.cuda.ExternalStream(ctx.stream)
    res = {}
    for name, st in (("fp32", vv.STORE_FP32), ("fp16", vv.STORE_FP16), ("unorm8", vv.STORE_UNORM8)):
        ctx.upload_field(field, dims, storage=st)
        ctx.compute_lic()
        ctx.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.compute_lic()
        e1.record(stream)
        lic = torch.empty(n ** 3, dtype=torch.float32, device=dev)
        ctx.download_lic(lic)
        cen = ctx.field_geometry()[2]
        img = torch.empty(W * W * 4, dtype=torch.float32, device=dev)
        ctx.render(vv.default_camera(W, W, cen), out=img)
        torch.cuda.synchronize()
        res[name] = dict(lic=lic, img=img, lic_ms=e0.elapsed_time(e1))
    ctx.close()
    b = res["fp32"]
    scale = float(b["lic"].abs().max())
    out = {
        "workload": f"{n}^3 Rankine x12, field 3xFP16 ({n ** 3 * 6 / 1e6:.0f} MB) vs FP32, 4% binary noise, "
                    f"cos^2 L={L} h=0.005; {W}^2 frame; also the UNORM8 field (4 B per voxel)",
        "lic_ms_fp32": b["lic_ms"],
        "lic_ms_fp16": res["fp16"]["lic_ms"],
        "lic_ms_unorm8": res["unorm8"]["lic_ms"],
    }
    for name in ("fp16", "unorm8"):
        a = res[name]
        dl = (a["lic"] - b["lic"]).double().abs()
        di = (a["img"] - b["img"]).double().abs()
        out[f"{name}_vs_fp32"] = {
            "lic_max_rel": float(dl.max()) / scale,
            "lic_rms_rel": float(dl.pow(2).mean().sqrt()) / scale,
            "image_max_abs": float(di.max()),
            "image_rms": float(di.pow(2).mean().sqrt()),
        }
    # FP64 oracle on a central box: the whole field is encoded in double
    host = field.cpu().numpy()
    del field
    f4 = orc.encode_field(host, dims, f64=True)
    del host
    noise = orc.gen_noise(1, vv.NOISE_BINARY, 0.04, dims)
    op = orc.params(steps_forward=L, steps_backward=L, step_size_lic=0.005, inv_filter_area=inv,
                    gradient_scale=3.0)
    b0 = n // 2 - box // 2
    try:
        cores = max(1, min(len(os.sched_getaffinity(0)), 16))
    except AttributeError:
        cores = 1
    t = time.perf_counter()
    ref = orc.lic_box(f4, dims, noise, 1, dims, lut, op, (b0, b0 + box) * 3, nthreads=cores)
    oracle_s = time.perf_counter() - t
    del f4
    ref = ref.reshape(box, box, box)
    rscale = float(np.max(np.abs(ref)))
    for name in ("fp32", "fp16", "unorm8"):
        sub = res[name]["lic"].view(n, n, n)[b0:b0 + box, b0:b0 + box, b0:b0 + box].cpu().numpy()
        d = np.abs(sub.astype(np.float64) - ref)
        out[f"{name}_vs_fp64_{box}^3"] = {"lic_max_rel": float(d.max()) / rscale,
                                         "lic_rms_rel": float(np.sqrt(np.mean(d ** 2))) / rscale}
    out["oracle_box_s"] = oracle_s
    return out


def fp64_check_leg(dev, field, L, W=1024, box_threads=16):
    """The benchmark workload itself (256^3 LIC, 1024^2 frame, TF/opacity
    adjoint) against the FP64 oracle, for FP32, FP16 and UNORM8 field storage:
    LIC, image and gradient errors.

    Early ray termination (dest.a > 0.95) is a discontinuity: a ray whose
    opacity ends within rounding of 0.95 takes one sample more or less than
    in FP64, whatever the storage.  The termination index of a GPU ray is that
    of the oracle's FP32 ray cast of the GPU's own LIC volume on the same
    stored field (the GPU image equals it to 1e-5).  Image errors are given
    over the pixels whose index is the FP64 one and over all pixels.  The
    adjoint is run twice, for a seeded dL/dimage and for the same with the
    pixels whose index differs in any storage set to zero."""
    import numpy as np

    import vectorvisualization_amd as vv
    from vectorvisualization_amd import fields
    from oracle import oracle as orc

    n = field.shape[0]
    dims = (n, n, n)
    try:
        cores = max(1, min(len(os.sched_getaffinity(0)), box_threads))
    except AttributeError:
        cores = 1
    print(f"[bench] fp64 check: {n}^3 LIC + {W}^2 frame + adjoint vs the FP64 oracle ({cores} threads)",
          file=sys.stderr, flush=True)
    host = field.cpu().numpy()
    lut, inv = vv.make_filter(vv.FILTER_GAUSSIAN, L)
    rgba, op = fields.tf_magnitude()
    dimg = (np.random.default_rng(17).normal(size=W * W * 4) * 1e-3).astype(np.float32)
    storages = (("fp32", vv.STORE_FP32), ("fp16", vv.STORE_FP16), ("unorm8", vv.STORE_UNORM8))
    ctx = vv.Context(0)
    ctx.generate_noise(1, vv.NOISE_UNIFORM, 0.0, dims)
    ctx.set_filter(lut, inv)
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = L
    p.step_size_lic = 0.005
    p.step_size_vol = 1.0 / 256.0
    p.gradient_scale = 3.0
    p.tf_index = vv.TF_INDEX_MAG
    ctx.set_lic_params(p)
    ctx.set_tf(rgba, op)
    cam = None
    gpu = {}
    for name, st in storages:
        ctx.upload_field(field, dims, storage=st)
        ctx.compute_lic()
        cen = ctx.field_geometry()[2]
        cam = vv.default_camera(W, W, cen)
        gpu[name] = dict(lic=ctx.download_lic(), img=ctx.render(cam))
    noise = orc.gen_noise(1, vv.NOISE_UNIFORM, 0.0, dims)
    op64 = orc.params(steps_forward=L, steps_backward=L, step_size_lic=0.005, inv_filter_area=inv,
                      gradient_scale=3.0, step_size_vol=1.0 / 256.0, tf_index=1)
    view = orc.view_setup(W, W)
    t = time.perf_counter()
    f4 = orc.encode_field(host, dims, f64=True)
    lic64 = orc.lic(f4, dims, noise, 1, dims, lut, op64, nthreads=cores)
    img64, ns64, _ = orc.raycast(f4, dims, lic64, 1, dims, rgba, op, op64, view, nthreads=cores, f64=True)
    oracle_s = time.perf_counter() - t
    flips = np.zeros(W * W, bool)
    for name, st in storages:
        if st == vv.STORE_UNORM8:
            f4s = orc.field_bytes_as_float(orc.encode_field_u8(host, dims))
        else:
            f4s = orc.encode_field(host, dims)
            if st == vv.STORE_FP16:
                f4s = orc.f16_to_f32(orc.f32_to_f16(f4s))
        ps = orc.Params.from_buffer_copy(op64)
        ps.field_unorm8 = 1 if st == vv.STORE_UNORM8 else 0
        _, ns_s, _ = orc.raycast(f4s, dims, gpu[name]["lic"], 1, dims, rgba, op, ps, view, nthreads=cores)
        gpu[name]["same"] = ns_s == ns64
        flips |= ~gpu[name]["same"]
    dimg_m = dimg.reshape(-1, 4).copy()
    dimg_m[flips] = 0.0
    dimg_m = dimg_m.ravel()
    t = time.perf_counter()
    ref = {k: orc.raycast_adjoint(f4, dims, lic64, 1, dims, rgba, op, op64, view, d)
           for k, d in (("all", dimg), ("masked", dimg_m))}
    oracle_s += time.perf_counter() - t
    del f4
    for name, st in storages:
        ctx.upload_field(field, dims, storage=st)
        ctx.compute_lic()
        gpu[name]["adj"] = {k: ctx.render_adjoint(cam, d) for k, d in (("all", dimg), ("masked", dimg_m))}
    ctx.close()

    def errs(got, want):
        d = np.abs(got.astype(np.float64) - want)
        return {"max_rel": float(d.max() / np.max(np.abs(want))),
                "rel_l2": float(np.linalg.norm(d) / np.linalg.norm(want))}

    lscale = float(np.max(np.abs(lic64)))
    out = {"workload": f"{n}^3 Rankine x12, dense uniform noise, Gaussian L={L} h=0.005; {W}^2 frame step "
                       f"1/256, vector-length TF; adjoint for a seeded dL/dimage (N(0, 1e-3))",
           "oracle_s": oracle_s, "oracle_threads": cores,
           "masked_pixels_frac": float(flips.mean())}
    for name, _ in storages:
        g = gpu[name]
        same = g["same"]
        di = np.abs(g["img"].astype(np.float64) - img64).reshape(-1, 4)
        dl = np.abs(g["lic"].astype(np.float64) - lic64)
        out[name] = {
            "lic_max_rel": float(dl.max()) / lscale,
            "image_max_abs_same_termination": float(di[same].max()),
            "image_max_abs_all_pixels": float(di.max()),
            "image_rms": float(np.sqrt(np.mean(di ** 2))),
            "termination_index_differs_frac": float(1.0 - same.mean()),
        }
        for k in ("all", "masked"):
            (gt, gs), (rt, rs) = g["adj"][k], ref[k]
            out[name][f"dtf_{k}_rays"] = errs(gt, rt)
            out[name][f"dsigma_{k}_rays"] = errs(gs, rs)
    out["fp16_meets_1e-3_image_bound"] = out["fp16"]["image_max_abs_same_termination"] <= 1e-3
    return out


def tf_fit_leg(dev, field, L, W=1024, steps=200, seed_target=3, seed_start=5):
    """Transfer-function optimisation at full size: the 256^3 LIC volume, a
    1024^2 view, L2 image loss against a target rendered from a seeded random
    256x4 TF.  The gradient of torch.autograd through the hand-written adjoint
    is checked against central differences (h = 1e-3) on 32 RGB entries and
    32 alpha entries, then Adam (lr 0.02, TF clamped to [0,1]) runs 200
    steps.  An alpha entry moves the early-termination index (dest.a > 0.95)
    of some of the million rays when it changes by h, a jump that the
    difference quotient sees and the adjoint (which holds the index fixed)
    does not; RGB entries leave every alpha unchanged.  The alpha entries are
    checked again on the loss restricted to the rays whose termination index
    stays the same (vv_record_ray_samples of the +h, 0 and -h passes), and a
    third time with the opacity table scaled by 0.01, where no ray terminates
    early and the loss is smooth."""
    import numpy as np
    import torch

    import vectorvisualization_amd as vv
    from vectorvisualization_amd import fields
    from vectorvisualization_amd.autograd import render_tf

    n = field.shape[0]
    dims = (n, n, n)
    print(f"[bench] tf fit: {n}^3, {W}^2, central differences on 32 entries + {steps} Adam steps",
          file=sys.stderr, flush=True)
    ctx = vv.Context(0)
    ctx.upload_field(field, dims)
    ctx.generate_noise(1, vv.NOISE_UNIFORM, 0.0, dims)
    lut, inv = vv.make_filter(vv.FILTER_GAUSSIAN, L)
    ctx.set_filter(lut, inv)
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = L
    p.step_size_lic = 0.005
    p.step_size_vol = 1.0 / 256.0
    p.gradient_scale = 3.0
    p.tf_index = vv.TF_INDEX_MAG
    ctx.set_lic_params(p)
    ctx.compute_lic()
    cen = ctx.field_geometry()[2]
    cam = vv.default_camera(W, W, cen)
    op = fields.tf_magnitude()[1]
    with torch.no_grad():
        target = render_tf(ctx, cam, torch.tensor(fields.tf_random(seed_target)[0].reshape(256, 4),
                                                  device=dev), op).clone()

    def loss_of(tf, op_=op, target_=target):
        img = render_tf(ctx, cam, tf, op_)
        return ((img.double() - target_.double()) ** 2).sum()

    tf0 = torch.tensor(fields.tf_random(seed_start)[0].reshape(256, 4), device=dev, requires_grad=True)
    loss_of(tf0).backward()
    g = tf0.grad.detach().cpu().numpy().astype(np.float64).ravel()
    rng = np.random.default_rng(0)
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    ch = np.arange(1024) % 4
    h = 1e-3

    def fd_worst(cand, g, op_=op, target_=target):
        idx = rng.choice(cand, min(32, cand.size), replace=False)
        worst = 0.0
        with torch.no_grad():
            for k in idx:
                e = torch.zeros(1024, device=dev)
                e[k] = h
                fd = (loss_of(tf0.detach() + e.view(256, 4), op_, target_).item() -
                      loss_of(tf0.detach() - e.view(256, 4), op_, target_).item()) / (2 * h)
                worst = max(worst, abs(fd - g[k]) / max(abs(fd), 1e-3 * np.abs(g).max()))
        return len(idx), float(worst)

    n_rgb, worst_rgb = fd_worst(np.flatnonzero(big & (ch < 3)), g)
    n_a, worst_a = fd_worst(np.flatnonzero(big & (ch == 3)), g)

    # alpha entries at the configuration's own settings, on the rays whose
    # early-termination index is the same at tf0 - h, tf0 and tf0 + h (the
    # per-ray sample counts of the recorded forward passes): central
    # differences of that masked loss against its adjoint
    H_, W_ = cam.height, cam.width

    def counts_of(tf_):
        with torch.no_grad():
            img_ = render_tf(ctx, cam, tf_, op)
        return img_, ctx.record_ray_samples(cam).copy()

    _, c0 = counts_of(tf0.detach())
    idx_m = rng.choice(np.flatnonzero(big & (ch == 3)), min(32, int((big & (ch == 3)).sum())), replace=False)
    worst_m, excluded = 0.0, []
    for k in idx_m:
        e = torch.zeros(1024, device=dev)
        e[k] = h
        ip, cp = counts_of(tf0.detach() + e.view(256, 4))
        im, cm = counts_of(tf0.detach() - e.view(256, 4))
        keep = (cp == c0) & (cm == c0)
        excluded.append(int((~keep).sum()))
        mt = torch.from_numpy(keep).to(dev).view(H_, W_, 1).double()
        fd = float(((((ip.double() - target.double()) ** 2) - ((im.double() - target.double()) ** 2)) * mt).sum()
                   / (2 * h))
        tfk = tf0.detach().clone().requires_grad_(True)
        (((render_tf(ctx, cam, tfk, op).double() - target.double()) ** 2) * mt).sum().backward()
        gk = float(tfk.grad.view(-1)[k])
        worst_m = max(worst_m, abs(fd - gk) / max(abs(fd), 1e-3 * np.abs(g).max()))

    # alpha entries again with no early termination: the opacity table scaled
    # by 0.01 keeps every sample's alpha <= 0.01 (<= 0.005 after the 1/256
    # step correction), so no ray of <= 443 samples reaches dest.a > 0.95 and
    # the loss is smooth in every TF entry; the largest alpha seen is reported
    op_lo = (np.asarray(op, np.float32) * np.float32(0.01)).astype(np.float32)
    with torch.no_grad():
        target_lo = render_tf(ctx, cam, torch.tensor(fields.tf_random(seed_target)[0].reshape(256, 4),
                                                     device=dev), op_lo).clone()
        alpha_max = float(target_lo.view(-1, 4)[:, 3].max().item())
    tfa = tf0.detach().clone().requires_grad_(True)
    loss_lo = loss_of(tfa, op_lo, target_lo)
    loss_lo.backward()
    g_lo = tfa.grad.detach().cpu().numpy().astype(np.float64).ravel()
    with torch.no_grad():
        alpha_max = max(alpha_max, float(render_tf(ctx, cam, tf0.detach(), op_lo).view(-1, 4)[:, 3].max().item()))
    big_lo = np.abs(g_lo) > 1e-3 * np.abs(g_lo).max()
    n_a_lo, worst_a_lo = fd_worst(np.flatnonzero(big_lo & (ch == 3)), g_lo, op_lo, target_lo)
    tf = tf0.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([tf], lr=0.02)
    losses = []
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        opt.zero_grad()
        loss = ((render_tf(ctx, cam, tf, op) - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            tf.clamp_(0.0, 1.0)
        losses.append(loss.item())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    ctx.close()
    return {
        "workload": f"{n}^3 LIC volume (Rankine x12, Gaussian L={L}), {W}^2 view, L2 image loss against a target "
                    f"from tf_random({seed_target}), start tf_random({seed_start})",
        "fd_rgb_entries": n_rgb,
        "fd_rgb_worst_rel_err": worst_rgb,
        "fd_alpha_entries": n_a,
        "fd_alpha_worst_rel_err": worst_a,
        "fd_alpha_masked": {"entries": len(idx_m), "worst_rel_err": worst_m,
                            "rays_excluded_mean": float(np.mean(excluded)), "rays_excluded_max": max(excluded),
                            "rule": "loss over the rays whose termination index is equal at tf0-h, tf0, tf0+h"},
        "fd_alpha_no_termination": {"opacity_table_scale": 0.01, "entries": n_a_lo,
                                    "worst_rel_err": worst_a_lo, "image_alpha_max": alpha_max},
        "adam_steps": steps,
        "loss_first": float(losses[0]),
        "loss_last": float(losses[-1]),
        "ms_per_adam_step": dt * 1000.0 / steps,
    }


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(field_np, dims, L, lut, inv, p, seconds, samples_per_voxel):
    """The oracle's FP32 LIC (its -O3 AVX2/FMA build, bit-identical to the
    checker) on a central z-slab of the same workload: one warm-up run, then
    the median of 5 timed runs, on all the threads this process may use
    (the CPU share of one GPU: the affinity mask, limited by OMP_NUM_THREADS)
    and on 1 thread.  The slab is sized so that the 12 runs take about
    `seconds` in all."""
    import numpy as np

    from oracle import oracle as orc

    n = dims[0]
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    cores = max(1, min(avail, int(omp))) if omp.isdigit() else avail
    f4 = orc.encode_field(field_np, dims)
    noise = orc.gen_noise(1, 1, 0.0, dims)
    op = orc.params(steps_forward=L, steps_backward=L, step_size_lic=0.005, inv_filter_area=inv,
                    gradient_scale=3.0)
    per_run = seconds / 12.0

    def run(threads, k):
        z0 = max(0, n // 2 - k // 2)
        t = time.perf_counter()
        orc.lic(f4, dims, noise, 1, dims, lut, op, z0=z0, z1=z0 + k, nthreads=threads, fast=True)
        return time.perf_counter() - t, z0

    res = {}
    for threads in (cores, 1):
        print(f"[bench] cpu baseline on {threads} thread(s)", file=sys.stderr, flush=True)
        one, _ = run(threads, 1)  # warm-up, and the slab size
        k = int(max(1, min(n, round(per_run / max(one, 1e-4)))))
        times = []
        for _ in range(5):
            dt, z0 = run(threads, k)
            times.append(dt)
        med = float(np.median(times))
        vox = k * n * n
        res[threads] = {"value": vox * samples_per_voxel / med, "median_s": med, "runs_s": times,
                        "slab": f"z {z0}..{z0 + k - 1} ({vox} voxels)"}
    allc = res[cores]
    return {
        "value": allc["value"],
        "unit": "voxel-samples/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle FP32 LIC (-O3 x86-64-v3 build) of the central z-slab {allc['slab']} of the same "
                  f"{n}^3 workload; median of 5 runs after a warm-up",
        "cpu_model": _cpu_model(),
        "host_threads_available": avail,
        "one_core": {"value": res[1]["value"], "median_s": res[1]["median_s"], "slab": res[1]["slab"]},
        "all_cores_runs_s": allc["runs_s"],
    }


if __name__ == "__main__":
    main()
