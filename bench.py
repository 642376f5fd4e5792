#!/usr/bin/env python3
"""Benchmark of the distance-field path on one MI355X.

Headline workload: exact squared EDT of a 512^3 bitmask volume (a seeded
union of 300 rotated ellipsoids with semi-axes 4-64 voxels and 100 boxes),
giving int32 squared distances and float32 distances.  One step = one full
EDT of the volume, inputs and outputs resident in HBM.  value = voxels per
second in Gvoxels/s.

Also reported on the same line:
  roofline      dominant kernel of the step: algorithmic bytes per launch /
                its mean duration (HIP events on the launch stream, recorded
                during the timed steps) against 8 TB/s
  cpu_baseline  the oracle's exact separable EDT (oracle/dol_oracle.c) on
                the same volume, one thread, on this host; its output is
                also compared with the GPU's, voxel for voxel
  extras        the signed 3D JumpFlood pipeline (border init + JumpFlood in
                half-cell units + float distances) on the same volume, and
                batched glyph SDFs against per-image launches

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "distance-occlusion_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "Gvoxels/s + % HBM roofline, 512³ bitmask EDT; batched glyphs/s vs per-image"

# algorithmic HBM bytes per voxel of each kernel of the 3D EDT: what it must
# read and write at least once
EDT3D_BYTES_PER_VOXEL = {
    "edt_rows": 0.125 + 4.0,          # mask bits in, row field out
    "edt_cols": 4.0 + 4.0,            # row field in, y field out
    "edt_slabs": 4.0 + 4.0 + 4.0,     # y field in, int32 + float32 out
    "edt3d_xy": 0.125 + 4.0,          # fused x+y pass: mask in, y field out
    "edt3d_z": 4.0 + 4.0 + 4.0,       # z pass: y field in, int32 + float32 out
}
EDT3D_PIPELINE_BYTES_PER_VOXEL = 24.125   # 0.125 + 4 + 8 + 4 + 4 + 4 (three passes)


def timed(fn, steps, torch):
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(steps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def load_traffic(kernel):
    """HBM bytes per launch measured with rocprofv3 counters (profiles/), if any."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get(kernel)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path runs on one MI355X; see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    import torch
    import dolhip

    L = dolhip.load()
    if L.device_count() < 1:
        print("no gfx950 device: " + L.last_error(), file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda")
    N = args.size
    w = h = d = N
    cells = w * h * d

    print(f"# generating {N}^3 scene", file=sys.stderr, flush=True)
    mask = L.scene_shapes_3d(w, h, d, 300, 100, seed=3)
    mask_t = torch.from_numpy(mask.view(np.int32)).to(dev)
    d2 = torch.empty(cells, dtype=torch.int32, device=dev)
    fl = torch.empty(cells, dtype=torch.float32, device=dev)
    work = torch.empty(L.workspace("edt_3d", w, h, d), dtype=torch.uint8, device=dev)

    def step():
        L.call_dev("edt_3d", w, h, d, mask_t, 0, d2, fl, work)

    print("# warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    L.profile_enable(True)
    ms = timed(step, args.steps, torch)
    kernels = {}
    for k in EDT3D_BYTES_PER_VOXEL:
        avg, n = L.profile_query(k)
        if n:
            kernels[k] = {"avg_ms": avg, "launches": n}
    L.profile_enable(False)
    print(f"# edt3d {ms:.3f} ms/step, kernels {kernels}", file=sys.stderr, flush=True)

    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    dom_bytes = EDT3D_BYTES_PER_VOXEL[dom] * cells
    achieved = dom_bytes / (kernels[dom]["avg_ms"] * 1e-3) / 1e9
    roofline = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": load_traffic(dom),
        "algorithmic_bytes": int(dom_bytes),
        "avg_ms": round(kernels[dom]["avg_ms"], 4),
    }
    gvox = cells / (ms * 1e-3) / 1e9
    pipeline_pct = EDT3D_PIPELINE_BYTES_PER_VOXEL * cells / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS

    result = {
        "metric": METRIC,
        "value": round(gvox, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded union of 300 rotated ellipsoids and 100 boxes)",
        "config": {"workload": f"exact squared EDT of a {N}^3 bitmask volume, int32 + float32 out",
                   "voxels": cells, "fill": round(float(np.unpackbits(mask.view(np.uint8)).mean()), 4)},
        "hbm_roofline_pct_pipeline": round(100 * pipeline_pct, 2),
        "kernels": {k: round(v["avg_ms"], 4) for k, v in kernels.items()},
        "roofline": roofline,
    }

    # ---- CPU baseline (the oracle, one thread) and full-size parity
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bind import load_oracle

        O = load_oracle()
        print("# cpu baseline", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        ref = O.edt_3d(w, h, d, mask)
        cpu_s = time.perf_counter() - t0
        got = d2.cpu().numpy().reshape(d, h, w)
        mism = int((got != ref).sum())
        result["cpu_baseline"] = {
            "value": round(cells / cpu_s / 1e9, 5),
            "unit": "Gvoxels/s",
            "cores": 1,
            "kind": "port",
            "sample": f"the full {N}^3 volume once (exact separable EDT of oracle/dol_oracle.c, "
                      f"{cpu_s:.1f} s on {os.cpu_count()} host cores, one used)",
        }
        result["parity"] = {"edt3d_int32_mismatches_vs_oracle": mism}
        if mism:
            print(f"# PARITY FAILURE: {mism} voxels differ from the oracle", file=sys.stderr)
    else:
        result["cpu_baseline"] = None

    # ---- extras: signed 3D JumpFlood pipeline on the same volume
    if not args.no_extras:
        del work
        delta = torch.empty(cells * 3, dtype=torch.int16, device=dev)
        jwork = torch.empty(L.workspace("jump_flood_3d", w, h, d), dtype=torch.uint8, device=dev)

        def jstep():
            L.call_dev("signed_border_3d", w, h, d, mask_t, delta, jwork)
            L.call_dev("signed_distance_3d", w, h, d, mask_t, delta, fl)

        jstep()
        L.profile_enable(True)
        jms = timed(jstep, 2, torch)
        jk = {}
        for k in ("border_3d", "jfa3d_pass", "signed_dist_3d"):
            avg, n = L.profile_query(k)
            jk[k] = round(avg, 4)
        L.profile_enable(False)
        result["jfa3d_signed"] = {
            "ms": round(jms, 3),
            "gvoxels_per_s": round(cells / (jms * 1e-3) / 1e9, 3),
            "kernels_avg_ms": jk,
            "algorithmic_bytes_per_voxel": 0.125 + 6 + 10 * 12 + 4,
        }
        del delta, jwork

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
