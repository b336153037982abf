#!/usr/bin/env python
"""Benchmark of the beta-surfel adjoint renderer on one MI355X.

One step = one forward pass (photon tracing, photon grid, camera gather,
post-processing) plus one adjoint pass of the L2 loss against a target image
of the same scene with jittered centers (the reference's training-step pair,
python/training.py:621-662), returning the image and the gradients of every
surfel parameter (--gradients opacity: only the opacity gradient, as the
reference's adjoint), on the plain single-view workload:

  1,048,576 clustered surfels (64 Gaussian clusters, sigma 0.1), 800x800,
  16 gather passes and 16 adjoint samples per pixel, depth 6, 2M photons,
  gather radius 0.02, fp32.

Metric: Mrays/s = closest-hit scene queries (photon bounces + camera-gather
surfel layers + adjoint bounces) per second of device time of the step.  The
step's inputs are resident in device memory when the timed region starts;
the wall-clock rate, which includes the host<->device image copies and the
loss on the host, is reported beside it (mrays_per_s_wall).

Independent-samples form ("independent_samples"): the same step with every
gather pass and adjoint sample seeded on its own (spp_decorrelate and
adjoint_decorrelate).  The reference seeds all gather passes of a pixel
alike and all its adjoint samples with 42, so its 16 spp are one sample
repeated; this form is 16 independent samples per pixel.

Also reported: the roofline of the dominant kernel by device time, priced
against HBM from its counted traffic (calibrated rocprofv3 FETCH_SIZE +
WRITE_SIZE of the same workload, summarised in profiles/) over the
launch time measured here, with what the counters say bounds it; the
reference's algorithmic byte model per work unit (counted in a separate,
untimed counting step) beside it as information; the same figures for the
other large kernels; and the CPU oracle timed on a bounded sample of the
same workload.

Served form ("served"): four resident clustered scenes (100k, 500k, 1M and
3M surfels with photon maps of 1, 2, 3 and 4M photons, depth 6, radius
0.02, two replicas each) take a Poisson stream of view requests (seed 1):
scene uniform, forward only with p = 0.7 else forward + adjoint against a
target image of the scene with jittered centers, resolution log-uniform in
area between 256x256 and 1920x1080, spp uniform in {1, 4, 16, 64}.  A bounded
prefix of the stream is run (--served requests; the full stream is 10,000),
first one request at a time (the plain single-view path on the same
requests), then all at once (the served form's capacity: views overlap on
the device) and then as they arrive; reported are p50/p99 latency from
arrival to completion at the arrival rate, throughput of the three runs and
whether every result is the same alone and in both mixes.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s fwd+adjoint at 1M surfels 800² 16spp vs plain; served p50/p99 ms"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--surfels", type=int, default=None, help="override the surfel count (tests)")
    ap.add_argument("--width", type=int, default=None)
    ap.add_argument("--height", type=int, default=None)
    ap.add_argument("--photons", type=int, default=None)
    ap.add_argument("--spp", type=int, default=None)
    ap.add_argument("--decorrelate", action="store_true", help="independent seed per gather pass")
    ap.add_argument("--gradients", choices=("full", "opacity"), default="full",
                    help="full: all parameter gradients (the benchmark's workload); opacity: the reference's adjoint")
    ap.add_argument("--cpu-photon-stride", type=int, default=8)
    ap.add_argument("--cpu-pixel-stride", type=int, default=80)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="concurrent oracle instances of the CPU baseline (capped at the host's CPU count)")
    ap.add_argument("--no-count", action="store_true", help="skip the work-counting step (roofline)")
    ap.add_argument("--decorrelated", type=int, default=3,
                    help="timed steps of the independent-samples form (every gather pass and adjoint sample "
                         "seeded on its own; 0: skip)")
    ap.add_argument("--served", type=int, default=400, help="requests of the served stream (0: skip)")
    ap.add_argument("--served-rate", type=float, default=2.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--served-alone", type=int, default=40,
                    help="head of the stream also run one request at a time and all at once (same results check)")
    ap.add_argument("--served-replicas", type=int, default=2)
    ap.add_argument("--served-batch-slots", type=int, default=0,
                    help="pixel slots per layer-walk launch of the served replicas (0: the renderer's default)")
    ap.add_argument("--served-grid-fraction", type=float, default=0.0,
                    help="fraction of the GPU a served replica's layer walk occupies (0: all)")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (0: leave the environment's).  The served form's "
                         "replicas each own a stream; with fewer hardware queues than streams, a small request's "
                         "kernels queue behind a large request's in the queue they share")
    ap.add_argument("--served-only", action="store_true", help="run only the served form")
    ap.add_argument("--batched", type=int, default=32, help="views of the batched form (0: skip)")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched form")
    ap.add_argument("--differentiation", type=int, default=2, help="autograd steps of the differentiation form (0: skip)")
    ap.add_argument("--differentiation-only", action="store_true", help="run only the differentiation form")
    ap.add_argument("--training", type=int, default=2,
                    help="timed iterations of the training-loop form (render, backward, optimizer step, apply, "
                         "rebuild_bvh, refetch; 0: skip)")
    ap.add_argument("--no-tune", action="store_true", help="batched form with the default tuning")
    ap.add_argument("--batched-tuning", type=str, default="",
                    help="batched form with this tuning (slots,entries,graph_rounds,grid_fraction) instead of the "
                         "autotuner's")
    ap.add_argument("--batch-passes", type=int, choices=(0, 1), default=1,
                    help="plain form: 1 = all gather passes / adjoint samples of the view as batches")
    ap.add_argument("--tuning", type=str, default="12582912,24,0,1",
                    help="plain form batch tuning as slots,entries,graph_rounds,grid_fraction (the default holds "
                         "all 16 passes of 800x800 in one batch, eager launches so each stage is timed)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only an elementwise kernel reading 1 GiB and writing 1 GiB: the known-size kernel "
                         "that scales rocprof FETCH_SIZE / WRITE_SIZE to bytes")
    ap.add_argument("--adjoint-decorrelate", action="store_true",
                    help="plain form: adjoint sample k seeds from 42 + k (with --decorrelate: every sample independent)")
    ap.add_argument("--form", choices=("plain", "independent", "batched", "served", "training", "bvh"), default=None,
                    help="run only this form (plain: the headline step; independent: the same step with every "
                         "gather pass and adjoint sample independently seeded)")
    ap.add_argument("--marker", action="store_true",
                    help="launch a marker kernel (torch's spin_kernel) right before the timed steps, so that a "
                         "profile can tell the timed steps' dispatches from set-up")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR", default=None,
                    help="summarise rocprofv3 CSV output (kernel trace and/or counter collections) under the given "
                         "directories: per kernel, the dispatches after the last marker kernel")
    ap.add_argument("--calibration", nargs="+", metavar="DIR", default=None,
                    help="with --pmc-summary: rocprofv3 FETCH_SIZE / WRITE_SIZE collections of --copy-calibration")
    return ap.parse_args()


def copy_calibration():
    import torch

    n = 1 << 28  # 2^28 fp32 = 1 GiB
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    marker()
    # An elementwise kernel (not the runtime's copy engine or blit path), so
    # the counters see an ordinary streaming read and write.
    torch.add(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))
    return 0


MARKER_KERNEL = "spin_kernel"


def marker():
    """A short kernel of a name no step launches (torch's spin_kernel): in a
    profile, the dispatches after it are the timed region's."""
    import torch

    torch.cuda.synchronize()
    torch.cuda._sleep(1000)
    torch.cuda.synchronize()


def _read_csvs(dirs, suffix):
    import csv
    import glob

    rows = []
    for d in dirs:
        for path in sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)):
            with open(path, newline="") as f:
                rows.extend(csv.DictReader(f))
    return rows


def _short(name: str) -> str:
    """Kernel name without namespaces, template arguments and parameter list."""
    name = name.replace("(anonymous namespace)::", "").split("(")[0].split("<")[0].strip()
    return name.split()[-1].split("::")[-1] if name else name


def _after_marker(rows, key_start):
    if rows and key_start not in rows[0]:
        key_start = "Dispatch_Id"  # dispatch ids grow with submission order
    marks = [int(r[key_start]) for r in rows if MARKER_KERNEL in r.get("Kernel_Name", "")]
    t0 = max(marks) if marks else -1
    return [r for r in rows if int(r[key_start]) > t0 and MARKER_KERNEL not in r.get("Kernel_Name", "")]


def pmc_summary(dirs, cal_dirs=None):
    """Per-kernel figures of the dispatches after the last marker kernel in
    rocprofv3 CSV output (--output-format csv): from a kernel trace the
    launches, durations and the resources of the code object (VGPRs, AGPRs,
    SGPRs, scratch and LDS bytes); from counter collections every counter's
    sum over those dispatches.  With calibration directories (the same
    FETCH_SIZE / WRITE_SIZE collections over --copy-calibration, 1 GiB read
    and 1 GiB written by one elementwise kernel) FETCH_SIZE and WRITE_SIZE are
    turned into bytes."""
    out = {"kernels": {}}
    kt = _read_csvs(dirs, "kernel_trace.csv")
    if kt:
        rows = _after_marker(kt, "Start_Timestamp")
        # Device busy time: the union of the dispatches' intervals (kernels of
        # several streams overlap, so their durations add up to more).
        busy, end = 0, None
        for s0, s1 in sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows):
            if end is None or s0 > end:
                busy += s1 - s0
                end = s1
            elif s1 > end:
                busy += s1 - end
                end = s1
        if rows:
            out["totals"] = {"device_busy_ms": round(busy * 1e-6, 3), "dispatches": len(rows)}
        for r in rows:
            k = out["kernels"].setdefault(_short(r["Kernel_Name"]), {"launches": 0, "total_ms": 0.0})
            k["launches"] += 1
            k["total_ms"] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
            # VGPR_Count is the profiler's field as it reads it (on gfx950 half
            # the compiler's VGPR count).
            for col, key in (("Arch_VGPR_Count", "vgpr_field"), ("VGPR_Count", "vgpr_field"),
                             ("Accum_VGPR_Count", "agpr"),
                             ("SGPR_Count", "sgpr"), ("Scratch_Size", "scratch_bytes_per_lane"),
                             ("Private_Segment_Size", "scratch_bytes_per_lane"), ("LDS_Block_Size", "lds_bytes"),
                             ("Group_Segment_Size", "lds_bytes")):
                if col in r and r[col] not in ("", None):
                    k[key] = int(float(r[col]))
        for k in out["kernels"].values():
            k["avg_ms"] = k["total_ms"] / k["launches"]
    # Counter collections: one directory per rocprofv3 --pmc pass (separate
    # processes running the same workload); per kernel and counter, the mean
    # per dispatch over that pass's dispatches after the marker.
    for d in dirs:
        cc = _read_csvs([d], "counter_collection.csv")
        if not cc:
            continue
        acc = {}
        for r in _after_marker(cc, "Start_Timestamp"):
            a = acc.setdefault((_short(r["Kernel_Name"]), r["Counter_Name"]), [0.0, set()])
            a[0] += float(r["Counter_Value"])
            a[1].add(r["Dispatch_Id"])
        for (name, counter), (total, ids) in acc.items():
            k = out["kernels"].setdefault(name, {})
            k.setdefault("counters_per_launch", {})[counter] = total / len(ids)
            k.setdefault("counter_launches", {})[counter] = len(ids)
    if cal_dirs:
        cal = {}
        for d in cal_dirs:  # one run per directory, each with its own marker
            for r in _after_marker(_read_csvs([d], "counter_collection.csv"), "Start_Timestamp"):
                if r["Counter_Name"] in ("FETCH_SIZE", "WRITE_SIZE"):
                    cal[r["Counter_Name"]] = max(cal.get(r["Counter_Name"], 0.0), float(r["Counter_Value"]))
        gib = float(1 << 30)
        f_fac = gib / (cal["FETCH_SIZE"] * 1024.0) if cal.get("FETCH_SIZE") else None
        w_fac = gib / (cal["WRITE_SIZE"] * 1024.0) if cal.get("WRITE_SIZE") else None
        out["calibration"] = {"fetch_kib_for_1GiB_read": cal.get("FETCH_SIZE"),
                              "write_kib_for_1GiB_written": cal.get("WRITE_SIZE"),
                              "fetch_factor": f_fac, "write_factor": w_fac}
        for k in out["kernels"].values():
            c = k.get("counters_per_launch", {})
            if "FETCH_SIZE" in c and "WRITE_SIZE" in c and f_fac and w_fac:
                k["hbm_bytes_per_launch"] = c["FETCH_SIZE"] * 1024.0 * f_fac + c["WRITE_SIZE"] * 1024.0 * w_fac
        if "totals" in out:
            out["totals"]["hbm_bytes"] = sum(k["hbm_bytes_per_launch"] * k["launches"] for k in out["kernels"].values()
                                             if "hbm_bytes_per_launch" in k and k.get("launches"))
            out["totals"]["kernel_ms"] = round(sum(k.get("total_ms", 0.0) for k in out["kernels"].values()), 3)
    # What the counters say bounds each kernel, as fractions of the kernel's
    # cycles on every CU (GRBM_GUI_ACTIVE sums the 8 XCDs' GPU-busy cycles;
    # the _sum counters sum 256 CUs).
    for k in out["kernels"].values():
        c = k.get("counters_per_launch", {})
        if c.get("GRBM_GUI_ACTIVE"):
            cyc = c["GRBM_GUI_ACTIVE"] / 8.0
            u = {}
            for name, key in (("TD_TD_BUSY_sum", "td_busy"), ("TA_TA_BUSY_sum", "ta_busy"),
                              ("TCP_PENDING_STALL_CYCLES_sum", "l1_stalled_on_l2")):
                if name in c:
                    u[key] = c[name] / 256.0 / cyc
            if "TCP_TOTAL_CACHE_ACCESSES_sum" in c and "TCP_TCC_READ_REQ_sum" in c:
                u["l1_hit"] = 1.0 - c["TCP_TCC_READ_REQ_sum"] / max(c["TCP_TOTAL_CACHE_ACCESSES_sum"], 1.0)
            k["utilisation"] = u
        if c.get("SQ_WAVE_CYCLES"):
            u = k.setdefault("utilisation", {})
            for name, key in (("SQ_WAIT_ANY", "wave_waiting"), ("SQ_ACTIVE_INST_VALU", "wave_valu_issue")):
                if name in c:
                    u[key] = c[name] / c["SQ_WAVE_CYCLES"]
    # Keep the kernels that take at least 1 % of the profiled device time
    # (all of them when there is no kernel trace).
    tot = sum(k.get("total_ms", 0.0) for k in out["kernels"].values())
    if tot > 0:
        out["kernels"] = {n: k for n, k in out["kernels"].items() if k.get("total_ms", 0.0) >= 0.01 * tot}
    for k in out["kernels"].values():
        k.pop("counter_launches", None)
        for c in list(k.get("counters_per_launch", {})):
            k["counters_per_launch"][c] = round(k["counters_per_launch"][c], 1)
        if "total_ms" in k:
            k["total_ms"] = round(k["total_ms"], 4)
            k["avg_ms"] = round(k["avg_ms"], 4)
        for key in ("utilisation",):
            if key in k:
                k[key] = {u: round(v, 4) for u, v in k[key].items()}
    print(json.dumps(out, indent=1, sort_keys=True))
    return 0


# Algorithmic bytes of the reference layouts per work unit.  Per closest-hit
# query: ray in + hit/next ray out 48 + 48 B, 24 B per AABB test, 8 B per
# popped node, 52 B per surfel test.  Per irradiance gather: 4 B per cell
# start, 4 B per non-empty cell end, 40 B per photon scanned.
def query_bytes(queries: int, w: dict) -> float:
    return 96.0 * queries + 24.0 * w["aabb"] + 8.0 * w["pops"] + 52.0 * w["prims"]


def gather_bytes(w: dict) -> float:
    return 4.0 * w["cells"] + 4.0 * w["cells_nonempty"] + 40.0 * w["photons"]


def load_traffic(path):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def served_requests(count: int, rate: float, seed: int = 1):
    from differentiablepointrendering_amd import scene

    return scene.served_requests(count, rate, seed)


def batched_cameras(count: int, seed: int = 2):
    from differentiablepointrendering_amd import scene

    return scene.batched_cameras(count, seed)


def batched_leg(a):
    from differentiablepointrendering_amd import autotune, scene
    from differentiablepointrendering_amd.pale import Renderer

    t0 = time.time()
    n = 500_000
    pts, tgt = scene.clustered_surfels(n, seed=0)
    settings = dict(photons=2_000_000, bounces=6, forward_passes=1, gather_passes=16, adjoint_bounces=6,
                    adjoint_passes=16, seed=0, gather_radius=0.02,
                    full_gradients=1 if a.gradients == "full" else 0)
    cams, spp = batched_cameras(a.batched)
    seeds = [100 + i for i in range(len(cams))]
    slots = sum(c.width * c.height * s for c, s in zip(cams, spp))
    meshes = [scene.light_quad()]
    t = Renderer.from_scene(meshes, [], tgt, settings)
    t.build_photon_map()
    targets = [x["raw"][..., :3].copy() for x in t.render_views(cams, spp, seeds)]
    t.close()
    r = Renderer.from_scene(meshes, [], pts, settings)
    r.build_photon_map()
    print(f"[bench] batched: {len(cams)} views, {slots / 1e6:.1f} M pixel samples, setup {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    def dloss(raw, tg):
        d = raw[..., :3] - tg
        return (d / d.size).astype(np.float32)

    keys = ("opacity", "position", "scale", "albedo", "beta") if a.gradients == "full" else ("opacity",)
    # Plain: one view after another through the single-view path.
    tw = time.time()
    p_ms, p_q, p_img, p_grad = 0.0, 0, [], {k: 0.0 for k in keys}
    for c, s, sd, tg in zip(cams, spp, seeds, targets):
        img = r.render_view(c, s, sd)["raw"]
        st = r.stats()
        p_ms += st["forward_ms"]
        p_q += st["gather_queries"]
        g = r.render_view_backward(c, s, dloss(img, tg))
        st = r.stats()
        p_ms += st["backward_ms"]
        p_q += st["adjoint_queries"]
        p_img.append(img)
        for k in keys:
            p_grad[k] = p_grad[k] + g[k]
    plain_wall = time.time() - tw
    # Autotuner on the workload itself (how many passes of which views share
    # a batch is what the batch size decides, so a reduced sample misleads it).
    tune = None
    if a.batched_tuning:
        sl, en, gr, fr = a.batched_tuning.split(",")
        r.set_batch_tuning(int(sl), int(en), int(gr), float(fr))
    elif not a.no_tune:
        ns = len(cams)
        sc, ss = cams, spp
        sd = [dloss(np.zeros((c.height, c.width, 4), np.float32), tg) for c, tg in zip(sc, targets[:ns])]
        tune = autotune.tune(r, sc, ss, seeds[:ns], sd, total_slots=slots)
        print(f"[bench] batched tuning {tune['tuning']} ({tune['seconds']}s)", file=sys.stderr, flush=True)
    else:
        r.set_batch_tuning(**autotune.DEFAULT)
    # Batched: all views in one call each way (one warm-up first).
    r.render_views(cams[:2], spp[:2], seeds[:2])
    if a.marker:
        marker()
    tw = time.time()
    out = r.render_views(cams, spp, seeds)
    st = r.stats()
    b_ms, b_q, batches = st["forward_ms"], st["gather_queries"], r.last_batches
    walk_avg = st["camera_trace_ms"] / max(1, st["camera_trace_launches"])
    g = r.render_views_backward(cams, spp, [dloss(o["raw"], tg) for o, tg in zip(out, targets)])
    st = r.stats()
    b_ms += st["backward_ms"]
    b_q += st["adjoint_queries"]
    batched_wall = time.time() - tw
    same_images = all(np.array_equal(o["raw"].view(np.uint32), p.view(np.uint32)) for o, p in zip(out, p_img))
    grad_err = {}
    for k in keys:
        scale = max(float(np.abs(p_grad[k]).max()), 1e-30)
        grad_err[k] = float(np.abs(g[k] - p_grad[k]).max() / scale)
    used_tuning = r.batch_tuning()
    r.close()
    res = {
        "workload": f"{len(cams)} views (eyes on the radius-4 sphere, 128x96..1024x768 log-uniform area, spp "
                    f"{{4,8,16}}) of {n} clustered surfels, 2M photons, depth 6, forward + adjoint ({a.gradients} "
                    f"gradients, summed over views)",
        "pixel_samples": slots,
        "mrays_per_s": round(b_q / (b_ms * 1e-3) / 1e6, 2),
        "ms": round(b_ms, 2),
        "batches": batches,
        "plain_one_view_at_a_time": {"mrays_per_s": round(p_q / (p_ms * 1e-3) / 1e6, 2), "ms": round(p_ms, 2)},
        "speedup_vs_plain": round((p_ms / b_ms) * (b_q / max(p_q, 1)), 3),
        "wall_s": {"batched": round(batched_wall, 2), "plain": round(plain_wall, 2)},
        "same_queries": b_q == p_q,
        "same_images_as_plain": bool(same_images),
        "gradient_max_err_over_max": grad_err,
        "roofline": counter_roofline("batched", "k_cam_trace", walk_avg),
        "tuning": tune["tuning"] if tune else used_tuning,
        "autotune": {"sample": f"the {len(cams)} views at their spp", "seconds": tune["seconds"], "table": tune["table"]}
        if tune else None,
    }
    print(f"[bench] batched {res['mrays_per_s']} Mrays/s vs one view at a time "
          f"{res['plain_one_view_at_a_time']['mrays_per_s']}; same images {same_images}; grad err {grad_err}",
          file=sys.stderr, flush=True)
    return res


def differentiation_leg(a):
    """2M clustered surfels, 512x512, 32 spp, depth 8: forward + backward of
    an L2 loss through render_surfels (torch.autograd over every surfel
    parameter) on a scene camera; the adjoint paths are regenerated from
    their seeds, so device memory does not grow with spp or depth."""
    import torch

    from differentiablepointrendering_amd import scene
    from differentiablepointrendering_amd.pale import Renderer
    from differentiablepointrendering_amd.torch_ops import PARAMS, render_surfels

    t0 = time.time()
    n, W = 2_000_000, 512
    pts, tgt = scene.clustered_surfels(n, seed=0)
    cam = scene.default_camera(W, W)
    settings = dict(photons=2_000_000, bounces=8, forward_passes=1, gather_passes=32, adjoint_bounces=8,
                    adjoint_passes=32, seed=0, gather_radius=0.02, full_gradients=1)
    meshes = [scene.light_quad()]
    t = Renderer.from_scene(meshes, [cam], tgt, settings)
    target = torch.from_numpy(t.render_forward()[cam.name + "_raw"][..., :3].copy()).cuda()
    t.close()
    r = Renderer.from_scene(meshes, [cam], pts, settings)
    r.set_batch_tuning(16 << 20, 24, 0, 1.0)
    p = r.get_point_parameters()
    params = {k: torch.from_numpy(p[k]).cuda().requires_grad_(True) for k in PARAMS}
    mem0 = Renderer.device_bytes()

    def step():
        for v in params.values():
            v.grad = None
        img = render_surfels(r, params, camera=cam.name)
        q = r.stats()
        loss = 0.5 * ((img[..., :3] - target) ** 2).mean()
        loss.backward()
        torch.cuda.synchronize()
        b = r.stats()
        return float(loss.detach()), q["forward_ms"] + b["backward_ms"], q["photon_queries"] + q["gather_queries"] + b[
            "adjoint_queries"]

    step()
    ms, qs, tw = [], [], time.time()
    for _ in range(a.differentiation):
        loss, m, q = step()
        ms.append(m)
        qs.append(q)
    wall = (time.time() - tw) / max(1, a.differentiation)
    gnorm = {k: float(v.grad.norm()) for k, v in params.items()}
    res = {
        "workload": f"{n} clustered surfels, {W}x{W}, 32 spp, depth 8, 2M photons: forward + backward of an L2 "
                    f"loss through render_surfels (torch.autograd, all surfel parameters)",
        "mrays_per_s": round(float(np.mean(qs)) / (float(np.mean(ms)) * 1e-3) / 1e6, 2),
        "device_ms_per_step": round(float(np.mean(ms)), 2),
        "wall_ms_per_step": round(wall * 1e3, 1),
        "loss": loss,
        "gradient_norms": gnorm,
        "renderer_device_bytes": Renderer.device_bytes(),
        "renderer_device_bytes_before_first_step": mem0,
        "setup_s": round(time.time() - t0, 1),
    }
    r.close()
    print(f"[bench] differentiation {res['mrays_per_s']} Mrays/s, {res['device_ms_per_step']} ms per step",
          file=sys.stderr, flush=True)
    return res


def training_leg(a):
    """One iteration of the reference's training loop (training.py:621-750)
    at the plain workload's scene: render_forward, the L2 gradient of the
    raw image, render_backward with every gradient, an Adam step on the
    seven parameter arrays (numpy here, torch there), apply_point_
    optimization, rebuild_bvh and the parameters fetched again.  Wall time
    of each phase; rebuild_bvh is the whole scene build (the surfel BVH on
    the GPU, in std::nth_element's element order, with the meshes, TLAS and
    host mirrors around it) and its upload."""
    from differentiablepointrendering_amd import scene
    from differentiablepointrendering_amd.pale import Renderer

    spec = scene.config_scene(1)
    cam = spec.cameras[0]
    settings = dict(spec.settings, full_gradients=1)
    tgt = Renderer.from_scene(spec.meshes, spec.cameras, spec.target_points, settings)
    target = tgt.render_forward()[cam.name + "_raw"][..., :3].copy()
    tgt.close()
    r = Renderer.from_spec(spec, settings)
    keys = ("position", "tangent_u", "tangent_v", "scale", "albedo", "opacity", "beta")
    params = r.get_point_parameters()
    m = {k: np.zeros_like(params[k]) for k in keys}
    v = {k: np.zeros_like(params[k]) for k in keys}
    phases = []
    for it in range(1, a.training + 2):
        ph = {}
        t = time.perf_counter()
        raw = r.render_forward()[cam.name + "_raw"][..., :3]
        ph["render_forward"] = time.perf_counter() - t
        t = time.perf_counter()
        grads, _ = r.render_backward({cam.name: ((raw - target) / raw.size).astype(np.float32)})
        ph["render_backward"] = time.perf_counter() - t
        t = time.perf_counter()
        for k in keys:
            g = grads[k]
            m[k] = 0.9 * m[k] + 0.1 * g
            v[k] = 0.999 * v[k] + 0.001 * g * g
            params[k] = (params[k] - 1e-3 * (m[k] / (1 - 0.9 ** it)) /
                         (np.sqrt(v[k] / (1 - 0.999 ** it)) + 1e-8)).astype(np.float32)
        params["opacity"] = np.clip(params["opacity"], 0.0, 1.0)
        ph["optimizer_step_host"] = time.perf_counter() - t
        t = time.perf_counter()
        r.apply_point_optimization({k: params[k] for k in keys})
        ph["apply_point_optimization"] = time.perf_counter() - t
        t = time.perf_counter()
        r.rebuild_bvh()
        ph["rebuild_bvh"] = time.perf_counter() - t
        on_device = bool(r.stats()["bvh_on_device"])
        t = time.perf_counter()
        params = r.get_point_parameters()
        ph["refetch"] = time.perf_counter() - t
        m = {k: np.zeros_like(params[k]) for k in keys}  # a fresh optimizer over the new order
        v = {k: np.zeros_like(params[k]) for k in keys}
        if it > 1:  # the first iteration warms up
            phases.append(ph)
        print(f"[bench] training iteration {it}: {sum(ph.values()) * 1e3:.0f} ms", file=sys.stderr, flush=True)
    r.close()
    out = {k: round(float(np.mean([p[k] for p in phases])) * 1e3, 2) for k in phases[0]}
    return {"workload": f"{len(spec.points['position'])} clustered surfels, 800x800, 16 spp, depth 6: one "
                        f"training.py iteration (host arrays, as the reference's drivers)",
            "iterations": len(phases), "ms": out, "total_ms": round(sum(out.values()), 2),
            "bvh_rebuilt_on": "GPU" if on_device else "host", "host_threads": os.cpu_count()}


def bvh_leg(a):
    """rebuild_bvh of the plain workload's 1,048,576 surfels after
    apply_point_optimization of the current parameters (SceneBuild.cpp:642-
    772 each time), `warmup` untimed then `steps` timed: wall time of the
    whole scene build and upload, and where the surfel BVH was built."""
    from differentiablepointrendering_amd import scene
    from differentiablepointrendering_amd.pale import Renderer

    spec = scene.config_scene(1)
    r = Renderer.from_spec(spec, dict(spec.settings, photons=20_000))
    step = {k: v for k, v in r.get_point_parameters().items() if k != "shape"}
    ms = []
    for i in range(a.warmup + a.steps):
        if i == a.warmup and a.marker:
            marker()
        r.apply_point_optimization(step)
        t = time.perf_counter()
        r.rebuild_bvh()
        dt = (time.perf_counter() - t) * 1e3
        if i >= a.warmup:
            ms.append(dt)
        step = {k: v for k, v in r.get_point_parameters().items() if k != "shape"}
        print(f"[bench] rebuild {i}: {dt:.1f} ms", file=sys.stderr, flush=True)
    on_device = bool(r.stats()["bvh_on_device"])
    r.close()
    return {"workload": f"{len(spec.points['position'])} clustered surfels: rebuild_bvh after "
                        f"apply_point_optimization", "rebuilds": len(ms), "ms": round(float(np.mean(ms)), 2),
            "ms_min": round(float(np.min(ms)), 2), "bvh_built_on": "GPU" if on_device else "host"}


def served_hbm():
    """HBM rate of the served form while the device is busy, from the counted
    traffic of every kernel of a served run (its summary in profiles/:
    calibrated FETCH_SIZE + WRITE_SIZE per launch times launches) over the
    union of the dispatches' intervals in the same run's kernel trace (the
    replicas' kernels overlap, so their summed durations would undercount
    the rate)."""
    prof = load_traffic(profile_path("served"))
    tot = (prof or {}).get("totals", {})
    if not tot.get("hbm_bytes") or not tot.get("device_busy_ms"):
        return None
    b, t = float(tot["hbm_bytes"]), float(tot["device_busy_ms"])
    gbs = b / (t * 1e-3) / 1e9
    return {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "bytes": b, "device_busy_ms": round(t, 2), "kernel_ms": tot.get("kernel_ms"),
            "source": "profiles/" + PROFILES["served"]}


def served_leg(a):
    import threading

    import torch

    from differentiablepointrendering_amd import scene
    from differentiablepointrendering_amd.pale import Renderer
    from differentiablepointrendering_amd.serve import SceneServer, ViewRequest, latency_summary

    free, total = torch.cuda.mem_get_info()
    budget = int(0.9 * total)
    stream = served_requests(a.served, a.served_rate)
    t0 = time.time()
    tuning = {}
    if a.served_batch_slots:
        tuning["slots_per_batch"] = a.served_batch_slots
    if a.served_grid_fraction:
        tuning["grid_fraction"] = a.served_grid_fraction
    srv = SceneServer(hbm_budget_bytes=budget, tuning=tuning or None)
    targets = SceneServer(hbm_budget_bytes=budget)
    names = []
    resident = {}
    for k, (n, photons) in enumerate(scene.SERVED_SCENES):
        name = f"clustered_{n // 1000}k"
        names.append(name)
        pts, tgt = scene.clustered_surfels(n, seed=k)
        settings = dict(photons=photons, bounces=6, forward_passes=1, gather_passes=16, adjoint_bounces=6,
                        adjoint_passes=16, seed=0, gather_radius=0.02)
        meshes = [scene.light_quad()]
        resident[name] = srv.add_scene(name, meshes, pts, settings, replicas=a.served_replicas)
        if any(r[1] == k and r[4] for r in stream):
            targets.add_scene(name, meshes, tgt, settings, replicas=1)
        print(f"[bench] served scene {name}: {resident[name] / 2**30:.2f} GiB resident "
              f"({a.served_replicas} replicas)", file=sys.stderr, flush=True)
    # Target images of the forward + adjoint requests (untimed; one sample
    # per pixel of the scene with jittered centers).
    reqs = []
    for i, (t, sc, cam, spp, adjoint) in enumerate(stream):
        target = None
        if adjoint:
            target = targets.render(ViewRequest(names[sc], cam, 1, 10_000 + i)).raw[..., :3].copy()
        reqs.append((t, ViewRequest(names[sc], cam, spp, 1 + i, target=target)))
    targets.close()
    setup_s = time.time() - t0
    print(f"[bench] served setup {setup_s:.1f}s; {len(reqs)} requests, "
          f"{sum(r.adjoint for _, r in reqs)} with adjoint", file=sys.stderr, flush=True)
    head = reqs[:max(1, min(a.served_alone, len(reqs)))]

    # One request at a time: the plain single-view path on the head of the
    # stream.
    tp = time.perf_counter()
    alone = []
    for i, (_, r) in enumerate(head):
        alone.append(srv.render(r))
        if (i + 1) % 10 == 0:
            print(f"[bench] served alone {i + 1}/{len(head)}", file=sys.stderr, flush=True)
    plain_s = time.perf_counter() - tp
    plain = latency_summary(alone, plain_s)
    small_alone = [m.latency_ms for m in alone
                   if m.request.camera.width * m.request.camera.height * m.request.spp <= (1 << 20)]

    # All at once: the capacity of the served form (views overlap on the
    # device), on the same requests.
    if a.marker:
        marker()
    tb = time.perf_counter()
    burst = [f.result() for f in [srv.submit(r) for _, r in head]]
    burst_s = max(m.finished for m in burst) - tb
    cap = latency_summary(burst, burst_s)
    print(f"[bench] served burst: {cap['mrays_per_s']:.1f} Mrays/s in {burst_s:.1f}s", file=sys.stderr, flush=True)

    # As they arrive (Poisson), continuously scheduled: the latency.
    futs = []
    ts = time.perf_counter()
    for i, (t, r) in enumerate(reqs):
        dt = ts + t - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        futs.append(srv.submit(r))
        if (i + 1) % 50 == 0:
            done = sum(f.done() for f in futs)
            print(f"[bench] served arrivals {i + 1}/{len(reqs)} submitted, {done} done", file=sys.stderr,
                  flush=True)
    mixed = [f.result() for f in futs]
    makespan = max(m.finished for m in mixed) - ts
    served = latency_summary(mixed, makespan)
    # Latency by request size (pixel samples = width x height x spp).
    small = [m for m in mixed if m.request.camera.width * m.request.camera.height * m.request.spp <= (1 << 20)]
    by_size = {"small_max_samples": 1 << 20, "small_requests": len(small),
               "small_p50_ms": round(float(np.percentile([m.latency_ms for m in small], 50)), 2) if small else None,
               "small_p99_ms": round(float(np.percentile([m.latency_ms for m in small], 99)), 2) if small else None,
               "small_alone_p50_ms": round(float(np.percentile(small_alone, 50)), 2) if small_alone else None}
    same_images = same_grads = True
    mismatches = []
    for i, (x, y) in [(i, p) for i, p in enumerate(zip(alone, mixed))] + [(i, p) for i, p in enumerate(zip(alone, burst))]:
        img_ok = np.array_equal(x.raw.view(np.uint32), y.raw.view(np.uint32))
        grad_ok = True
        gerr = None
        if x.grads is not None:
            gx, gy = x.grads["opacity"], y.grads["opacity"]
            scale = max(float(np.abs(gx).max()), 1e-30)
            grad_ok = bool(np.allclose(gy, gx, rtol=1e-4, atol=1e-4 * scale))
            gerr = float(np.abs(gy - gx).max() / scale)
        same_images &= img_ok
        same_grads &= grad_ok
        if not (img_ok and grad_ok):
            mismatches.append({"request": i, "scene": x.request.scene, "replicas": [x.replica, y.replica],
                               "size": [x.request.camera.width, x.request.camera.height], "spp": x.request.spp,
                               "pixels_differing": int(np.sum(np.any(x.raw != y.raw, axis=-1))),
                               "max_abs_pixel_diff": float(np.abs(x.raw - y.raw).max()),
                               "grad_max_err_over_max": gerr})
    groups = srv.groups_served
    srv.close()
    dev_bytes_after = Renderer.device_bytes()
    out = {
        "workload": f"{len(reqs)} requests (prefix of the seed-1 stream; the first {len(head)} also alone and all "
                    f"at once) over {len(names)} resident scenes "
                    f"{'/'.join(names)}, Poisson {a.served_rate}/s, 30% forward+adjoint, 256x256..1920x1080, "
                    f"spp {{1,4,16,64}}, {a.served_replicas} replicas per scene",
        "arrival_rate_per_s": a.served_rate,
        "p50_ms": round(served["p50_ms"], 2),
        "p99_ms": round(served["p99_ms"], 2),
        "mean_wait_ms": round(served["mean_wait_ms"], 2),
        "mrays_per_s": round(served["mrays_per_s"], 2),
        "requests_per_s": round(served["requests_per_s"], 3),
        "makespan_s": round(makespan, 3),
        "latency_by_size": by_size,
        "head_requests_alone_and_all_at_once": len(head),
        "tuning": tuning or "renderer default",
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "capacity_all_at_once": {"mrays_per_s": round(cap["mrays_per_s"], 2),
                                 "requests_per_s": round(cap["requests_per_s"], 3),
                                 "p50_ms": round(cap["p50_ms"], 2), "p99_ms": round(cap["p99_ms"], 2),
                                 "total_s": round(burst_s, 3)},
        "plain_one_at_a_time": {"mrays_per_s": round(plain["mrays_per_s"], 2),
                                "requests_per_s": round(plain["requests_per_s"], 3),
                                "p50_ms": round(plain["p50_ms"], 2), "p99_ms": round(plain["p99_ms"], 2),
                                "total_s": round(plain_s, 3)},
        "queries": served["queries"],
        "same_results_alone_and_mixed": bool(same_images and same_grads),
        "groups_served": groups,
        "same_images": bool(same_images),
        "same_gradients": bool(same_grads),
        "mismatches": mismatches[:8],
        "hbm_while_busy": served_hbm(),
        "hbm_budget_bytes": budget,
        "resident_bytes": resident,
        "setup_s": round(setup_s, 1),
        "device_bytes_after_close": dev_bytes_after,
    }
    print(f"[bench] served p50 {out['p50_ms']} ms p99 {out['p99_ms']} ms at {a.served_rate}/s; capacity "
          f"{out['capacity_all_at_once']['mrays_per_s']} Mrays/s all at once vs "
          f"{out['plain_one_at_a_time']['mrays_per_s']} one at a time; same results: "
          f"{out['same_results_alone_and_mixed']}", file=sys.stderr, flush=True)
    return out


def cpu_leg(a, spec, settings, target, cam, step_queries=None):
    """The oracle (the sequential CPU restatement) on a bounded sample of the
    same workload: photons/stride photons, every stride-th pixel, all passes
    and samples, forward + adjoint.  `threads` instances run the same sample
    at once, one per host thread (ctypes releases the GIL inside the oracle),
    and the rate is their queries over the wall time from the common start to
    the last finish: the throughput of the host's cores on this work."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    from oracle.oracle import Oracle

    threads = max(1, min(a.cpu_threads, os.cpu_count() or 1))
    start = threading.Barrier(threads)

    def one(_):
        o = Oracle.from_spec(spec, settings)
        o.set_sample(a.cpu_photon_stride, a.cpu_pixel_stride)
        start.wait()
        t0 = time.time()
        raw = o.render_forward()[cam.name + "_raw"][..., :3]
        o.render_backward({cam.name: ((raw - target) / raw.size).astype(np.float32)})
        t1 = time.time()
        cs = o.stats()
        o.close()
        return cs["photon_queries"] + cs["gather_queries"] + cs["adjoint_queries"], t0, t1

    with ThreadPoolExecutor(threads) as ex:
        runs = list(ex.map(one, range(threads)))
    q = sum(r[0] for r in runs)
    wall = max(r[2] for r in runs) - min(r[1] for r in runs)
    per = [r[0] / (r[2] - r[1]) / 1e6 for r in runs]
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    out = {"value": round(q / wall / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "per_thread_value": round(float(np.mean(per)), 4),
           "host": {"cpu": model, "cpu_count": os.cpu_count()},
           "sample": f"{threads} concurrent oracle instances (one per host thread), each on the same scene with "
                     f"photons/{a.cpu_photon_stride} and every {a.cpu_pixel_stride}th pixel, "
                     f"{settings['gather_passes']} spp, depth {settings['bounces']}: {runs[0][0]} queries each, "
                     f"{wall:.1f}s wall",
           "rate_is": "queries of the sample over its wall time (measured on the sample, not extrapolated)",
           "sample_fraction_of_step": {"photons": 1.0 / a.cpu_photon_stride, "pixels": 1.0 / a.cpu_pixel_stride}}
    if step_queries:
        # What one full step would take on these cores at the sample's rate.
        out["full_step_s_at_this_rate"] = round(step_queries / (out["value"] * 1e6), 1)
    print(f"[bench] cpu baseline {out['value']} Mrays/s on {threads} threads ({wall:.1f}s)", file=sys.stderr,
          flush=True)
    return out


# Counter summaries in profiles/ per form (written by --pmc-summary).
PROFILES = {
    "plain": "step_kernels_1m_surfels_800px_16spp_reference_seeds.json",
    "independent": "step_kernels_1m_surfels_800px_16spp_independent_samples.json",
    "batched": "batched_32_views_500k_surfels.json",
    "served": "served_4_scenes_24_requests.json",
}


def profile_path(form):
    return os.path.join(ROOT, "profiles", PROFILES.get(form, f"{form}.json"))


def counter_roofline(form, kernel, avg_launch_ms, algorithmic_bytes=None):
    """Roofline of `kernel` priced against HBM from its counted traffic: the
    calibrated FETCH_SIZE + WRITE_SIZE bytes per launch of the same workload
    (its summary in profiles/, rocprofv3 --pmc) over the launch duration
    measured here with HIP events.  What bounds the kernel comes from the
    same profile's unit counters (TD / TA busy, L1 hit rate, issue), with
    HBM's own fraction beside them.  The reference byte model per work unit
    (every lane charged for its own node and surfel loads, most of which
    the L1 serves) is kept as an informational rate."""
    prof = load_traffic(profile_path(form))
    out = {"kernel": kernel, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
           "traffic": None, "avg_launch_ms": round(avg_launch_ms, 3)}
    k = (prof or {}).get("kernels", {}).get(kernel, {})
    if "hbm_bytes_per_launch" in k and avg_launch_ms > 0:
        b = float(k["hbm_bytes_per_launch"])
        ach = b / (avg_launch_ms * 1e-3) / 1e9
        out.update(achieved=round(ach, 1), frac=round(ach / HBM_PEAK_GBS, 4), traffic=b,
                   traffic_source="profiles/" + PROFILES.get(form, ""),
                   traffic_profile_avg_launch_ms=round(k.get("avg_ms", 0.0), 3))
        if ach > HBM_PEAK_GBS:
            out["invalid"] = "counted HBM bytes over the measured time exceed the HBM peak"
    else:
        out["note"] = f"no counter profile of {kernel} for this form (profiles/{PROFILES.get(form)})"
    u = k.get("utilisation")
    if u:
        lim = {kk: round(v, 3) for kk, v in u.items()}
        if out["frac"] is not None:
            lim["hbm"] = out["frac"]
        busy = {kk: v for kk, v in lim.items() if kk in ("td_busy", "ta_busy", "hbm")}
        top = max(busy, key=busy.get) if busy else None
        names = {"td_busy": "vector-memory data return (TD)", "ta_busy": "vector-memory addressing (TA)",
                 "hbm": "HBM bandwidth"}
        out["limiter"] = {"unit": names.get(top, top), "fractions": lim}
    for key in ("vgpr_field", "agpr", "sgpr", "scratch_bytes_per_lane", "lds_bytes"):
        if key in k:
            out.setdefault("resources", {})[key] = k[key]
    if algorithmic_bytes:
        out["algorithmic"] = {"bytes_per_launch": algorithmic_bytes,
                              "gbs": round(algorithmic_bytes / (avg_launch_ms * 1e-3) / 1e9, 1),
                              "model": "reference layouts per work unit: 96 B per query, 24 B per slab test, 8 B "
                                       "per popped node, 52 B per surfel test (per lane, cache hits included)"}
    return out


def run_plain(a, spec, settings, target, steps, warmup, form, count=True):
    """Timed forward + adjoint steps of one renderer on the plain workload.
    Device time per step comes from the renderer's HIP events (its own
    stream); wall time brackets each step from the host."""
    from differentiablepointrendering_amd.pale import Renderer

    cam = spec.cameras[0]
    t_build = time.time()
    r = Renderer.from_spec(spec, settings)
    if a.tuning:
        sl, en, gr, fr = a.tuning.split(",")
        r.set_batch_tuning(int(sl), int(en), int(gr), float(fr))
    build_s = time.time() - t_build
    bvh = r.stats()
    print(f"[bench] {form}: renderer ready in {build_s:.2f}s (BVH build {bvh['bvh_build_ms']:.0f} ms, "
          f"{'GPU' if bvh['bvh_on_device'] else 'host'})", file=sys.stderr, flush=True)

    def step():
        img = r.render_forward()
        raw = img[cam.name + "_raw"][..., :3]
        dl = ((raw - target) / raw.size).astype(np.float32)
        r.render_backward({cam.name: dl})
        return r.stats()

    for i in range(warmup):
        step()
        print(f"[bench] {form}: warmup {i + 1}/{warmup} done", file=sys.stderr, flush=True)
    if a.marker:
        marker()
    dev_ms, wall_ms, queries = [], [], []
    acc = {}
    tw0 = time.time()
    for i in range(steps):
        ts = time.time()
        st = step()
        wall_ms.append((time.time() - ts) * 1e3)
        dev_ms.append(st["forward_ms"] + st["backward_ms"])
        queries.append(st["photon_queries"] + st["gather_queries"] + st["adjoint_queries"])
        for k in ("gather_ms", "gather_launches", "camera_trace_ms", "camera_trace_launches", "photon_trace_ms",
                  "photon_trace_launches", "adjoint_trace_ms", "adjoint_trace_launches", "adjoint_transport_ms",
                  "adjoint_transport_launches"):
            acc[k] = acc.get(k, 0) + st[k]
        print(f"[bench] {form}: step {i + 1}/{steps}: device {dev_ms[-1]:.1f} ms, wall {wall_ms[-1]:.1f} ms, "
              f"{queries[-1] / 1e6:.1f} M queries", file=sys.stderr, flush=True)
    total_wall = time.time() - tw0
    last = st
    ms_step = float(np.mean(dev_ms))
    q_step = float(np.mean(queries))
    res = {
        "value": round(q_step / (ms_step * 1e-3) / 1e6, 2),
        "ms_per_step": round(ms_step, 3),
        "queries_per_step": int(q_step),
        "ms_per_step_wall": round(float(np.mean(wall_ms)), 3),
        "mrays_per_s_wall": round(q_step / (np.mean(wall_ms) * 1e-3) / 1e6, 2),
        "camera_samples_per_s": round(cam.width * cam.height * settings["gather_passes"] / (ms_step * 1e-3), 1),
        "phase_ms": {k: round(last[k], 3) for k in ("photon_trace_ms", "grid_ms", "camera_trace_ms", "camera_sort_ms",
                                                    "gather_ms", "adjoint_trace_ms", "adjoint_transport_ms",
                                                    "forward_ms", "backward_ms")},
        "camera_rounds": last["camera_rounds"],
        "renderer_setup_s": round(build_s, 3),
        "scene_build_ms": round(bvh["bvh_build_ms"], 1),
        "bvh_on_device": bool(bvh["bvh_on_device"]),
        "timed_wall_s": round(total_wall, 3),
        "batch_tuning": r.batch_tuning(),
    }
    # Average launch of each large kernel over the timed steps (HIP events on
    # the renderer's stream).
    avg = {}
    for name, key in (("k_cam_trace", "camera_trace"), ("k_entry_gather", "gather"), ("k_photon_trace", "photon_trace"),
                      ("k_adj_trace", "adjoint_trace"), ("k_adj_transport", "adjoint_transport")):
        n_l = acc.get(f"{key}_launches", 0)
        if n_l:
            avg[name] = {"avg_launch_ms": acc[f"{key}_ms"] / n_l, "ms_per_step": acc[f"{key}_ms"] / steps,
                         "launches_per_step": n_l / steps}
    res["kernels"] = avg
    algorithmic = {}
    if count:
        r.set_count_work(True)
        cst = step()
        r.set_count_work(False)
        work = {k: cst[f"work_{k}"] for k in ("photon", "gather", "adjoint")}
        res["work_per_step"] = work
        res["counted_step_matches_timed"] = bool(cst["gather_queries"] == last["gather_queries"])
        algorithmic["k_entry_gather"] = gather_bytes(work["gather"]) / max(1, cst["gather_launches"])
        algorithmic["k_cam_trace"] = query_bytes(cst["gather_queries"], work["gather"]) / max(
            1, cst["camera_trace_launches"])
        # Whole step in the same byte model, for the cross-check against the
        # step's device time.
        step_bytes = (query_bytes(cst["photon_queries"], work["photon"]) +
                      query_bytes(cst["gather_queries"], work["gather"]) + gather_bytes(work["gather"]) +
                      query_bytes(cst["adjoint_queries"], work["adjoint"]) + gather_bytes(work["adjoint"]))
        res["algorithmic_step"] = {"bytes": step_bytes, "gbs_over_step": round(step_bytes / (ms_step * 1e-3) / 1e9, 1),
                                   "note": "per-lane byte model; its rate above the HBM peak counts L1/L2 hits"}
    dom = max(avg, key=lambda k: avg[k]["ms_per_step"]) if avg else None
    if dom:
        res["roofline"] = counter_roofline(form, dom, avg[dom]["avg_launch_ms"], algorithmic.get(dom))
        res["kernel_rooflines"] = {k: counter_roofline(form, k, v["avg_launch_ms"], algorithmic.get(k))
                                   for k, v in avg.items() if k != dom}
    r.close()
    del r
    return res


def independent_leg(a, spec, settings, target):
    """The plain workload with every sample independently seeded: gather pass
    k draws seed + k (spp_decorrelate) and adjoint sample k seeds from 42 + k
    (adjoint_decorrelate).  With the reference's seeding the 16 gather passes
    of a pixel trace one ray and its 16 adjoint samples one path (the layer
    walk places a pixel's passes side by side in a wave, so their loads
    coalesce); here every sample is a distinct ray and path, so "16 spp" is
    16 independent samples.  Same surfels, target and tuning."""
    s2 = dict(settings, spp_decorrelate=1, adjoint_decorrelate=1)
    res = run_plain(a, spec, s2, target, a.decorrelated, 1, "independent", count=not a.no_count)
    print(f"[bench] independent samples {res['value']} Mrays/s, {res['ms_per_step']} ms per step",
          file=sys.stderr, flush=True)
    return res


def main():
    a = parse()
    if a.hw_queues:
        # Read by the HIP runtime when it starts (nothing has touched the GPU
        # yet): one hardware queue per replica stream of the served form.
        os.environ["GPU_MAX_HW_QUEUES"] = str(a.hw_queues)
    if a.copy_calibration:
        return copy_calibration()
    if a.pmc_summary:
        return pmc_summary(a.pmc_summary, a.calibration)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        return 2
    if a.served_only or a.form == "served":
        print(json.dumps({"served": served_leg(a)}))
        return 0
    if a.batched_only or a.form == "batched":
        print(json.dumps({"batched": batched_leg(a)}))
        return 0
    if a.differentiation_only:
        print(json.dumps({"differentiation": differentiation_leg(a)}))
        return 0
    if a.form == "training":
        print(json.dumps({"training_iteration": training_leg(a)}))
        return 0
    if a.form == "bvh":
        print(json.dumps({"bvh_rebuild": bvh_leg(a)}))
        return 0

    from differentiablepointrendering_amd import scene
    from differentiablepointrendering_amd.pale import Renderer

    spec = scene.config_scene(1, n=a.surfels, width=a.width, height=a.height)
    settings = dict(spec.settings)
    if a.photons:
        settings["photons"] = a.photons
    if a.spp:
        settings["gather_passes"] = settings["adjoint_passes"] = a.spp
    settings["spp_decorrelate"] = 1 if a.decorrelate else 0
    settings["adjoint_decorrelate"] = 1 if a.adjoint_decorrelate else 0
    settings["full_gradients"] = 1 if a.gradients == "full" else 0
    settings["batch_passes"] = a.batch_passes
    cam = spec.cameras[0]
    W, H = cam.width, cam.height
    n = len(spec.points["position"])
    print(f"[bench] scene {n} surfels {W}x{H} spp {settings['gather_passes']} photons {settings['photons']}",
          file=sys.stderr, flush=True)

    # Target image: the same scene with jittered centers (untimed).
    tgt = Renderer.from_scene(spec.meshes, spec.cameras, spec.target_points, settings)
    target = tgt.render_forward()[cam.name + "_raw"][..., :3].copy()
    tgt.close()
    del tgt

    if a.form == "independent":
        print(json.dumps({"independent_samples": independent_leg(a, spec, settings, target)}))
        return 0
    independent = bool(a.decorrelate and a.adjoint_decorrelate)
    form = "independent" if independent else ("plain" if not (a.decorrelate or a.adjoint_decorrelate) else "custom")
    p = run_plain(a, spec, settings, target, a.steps, a.warmup, form, count=not a.no_count)

    cpu = None
    if not a.no_cpu and a.form is None:
        cpu = cpu_leg(a, spec, settings, target, cam, p["queries_per_step"])

    out = {
        "metric": METRIC,
        "value": p["value"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": p["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator of the benchmark configuration)",
        "config": {
            "workload": f"plain single view: {n} clustered surfels, {W}x{H}, {settings['gather_passes']} spp, "
                        f"depth {settings['bounces']}, {settings['photons']} photons, radius "
                        f"{settings['gather_radius']}, forward + adjoint ({a.gradients} gradients)",
            "surfels": n, "width": W, "height": H, "spp": settings["gather_passes"],
            "photons": settings["photons"], "depth": settings["bounces"],
            "gather_passes_share_seed": not a.decorrelate,
            "adjoint_samples_share_seed": not a.adjoint_decorrelate,
            "batch_passes": a.batch_passes,
            "batch_tuning": p["batch_tuning"],
            "gradients": a.gradients,
        },
    }
    for k in ("queries_per_step", "ms_per_step_wall", "mrays_per_s_wall", "camera_samples_per_s", "phase_ms",
              "camera_rounds", "renderer_setup_s", "scene_build_ms", "bvh_on_device", "timed_wall_s", "roofline", "kernel_rooflines", "work_per_step",
              "counted_step_matches_timed", "algorithmic_step"):
        if k in p:
            out[k] = p[k]
    out["cpu_baseline"] = cpu
    if a.decorrelated > 0 and form == "plain" and a.form is None:
        out["independent_samples"] = independent_leg(a, spec, settings, target)
    if a.form is None:
        if a.batched > 0:
            out["batched"] = batched_leg(a)
        if a.differentiation > 0:
            out["differentiation"] = differentiation_leg(a)
        if a.training > 0:
            out["training_iteration"] = training_leg(a)
        if a.served > 0:
            out["served"] = served_leg(a)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
