#!/usr/bin/env python3
"""el_iso RK4 time-stepping throughput on one MI355X.

Workload (default): the 400x400x200 curvilinear topography model (3 layers
with correlated perturbations, double couple, 400 receivers) stepped in FP32
through the C ABI of gpu-cgfd3d-elastic_amd/libcgfd_el_iso.so.  A "step" is
one full RK4 step (4 stages) over the whole grid with receivers recorded;
W warmup steps run first, then exactly K steps are timed.  Inputs are
resident in HBM before the timed region.

Prints one JSON line with the throughput in Gcell-RK-stages/s
(= ni*nj*nk*4*K / t), the HBM roofline of the dominant (interior) kernel
measured with HIP events around its launches, and a CPU baseline: the
oracle (a C++/OpenMP restatement of the reference loop) timed on a few
steps of the same grid on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gpu-cgfd3d-elastic_amd")
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

import cgfd  # noqa: E402
import workloads  # noqa: E402

METRIC = "Gcell-RK-stages/s el_iso FP32 vs FP16-store; served shots/hour, p50/p99"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per interior cell and RK stage of the fused dataflow:
# mean of 39/57/57/39 words (wavefield in, coefficients, RK levels out)
ALGO_BYTES = {"f32": 192.0, "f16": 120.0, "bf16": 120.0, "f64": 384.0}
STORE = {"f32": cgfd.STORE_F32, "f16": cgfd.STORE_F16, "bf16": cgfd.STORE_BF16, "f64": cgfd.STORE_F64}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(m, steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as C

    import orc

    L = orc.lib()
    P = orc.problem_struct(m)
    nr = len(m["recv_indx"])
    seis = np.zeros((nr, 9, steps))
    rc = L.orc_run(C.byref(P), 32, m["dt"], steps, None, None, orc.p(seis), None, None)
    if rc != 0:
        raise RuntimeError("oracle run failed")
    sec = L.orc_last_loop_seconds()
    cells = workloads.cells(m)
    return dict(value=round(cells * 4 * steps / sec / 1e9, 5), unit="Gcell-RK-stages/s",
                cores=int(L.orc_max_threads()), kind="port",
                sample=f"{steps} full RK4 steps of the same {m['ni']}x{m['nj']}x{m['nk']} grid, FP32, "
                       f"unfused reference stage order, {sec:.1f} s loop")


def measured_traffic(workload, storage):
    """HBM bytes per interior launch from the committed counter summary, if any.

    The summary holds FETCH_SIZE + WRITE_SIZE per dispatch of the interior
    kernel on this workload, scaled by the same counters on a copy kernel of
    known size."""
    path = os.path.join(ROOT, "profiles", f"k_march_{workload}_{storage}_hbm_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--workload", default="topo", choices=["topo", "small"])
    ap.add_argument("--storage", default="f32", choices=sorted(STORE))
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain", action="store_true", help="per-point kernel everywhere")
    ap.add_argument("--tile", default="", metavar="SHAPE,ZCHUNK",
                    help="pin the interior tile instead of autotuning (profiling runs)")
    ap.add_argument("--timed-launch-steps", type=int, default=8,
                    help="steps of the per-launch event timing pass (0: skip)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (counter calibration)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch  # plumbing only: a plain copy of known size for HBM counter scaling

        n = args.copy_calib * 1024 * 1024 // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_each_way": n * 4, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(the decomposition is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    t0 = time.time()
    m = workloads.topo() if args.workload == "topo" else workloads.small()
    log(f"workload {m['name']} {m['ni']}x{m['nj']}x{m['nk']} prepared in {time.time() - t0:.1f}s; "
        f"dt {m['dt']} (dtmax {m['dtmax']:.5f})")
    if m["dt"] > m["dtmax"]:
        raise SystemExit(f"dt {m['dt']} exceeds the stability limit {m['dtmax']}")

    nt_total = args.warmup + args.steps
    sess = cgfd.Session(m["ni"], m["nj"], m["nk"], storage=STORE[args.storage],
                        flags=cgfd.FLAG_PLAIN if args.plain else 0)
    nt_total += args.timed_launch_steps
    workloads.upload(sess, m, nt_total=nt_total)
    if args.tile:
        sh, zc = (int(v) for v in args.tile.split(","))
        sess.set_tile(sh, zc)
    log("uploaded; warmup")
    sess.run(0, args.warmup, m["dt"])
    log(f"timing {args.steps} steps")
    # the timed region: K steps, 8-step periods replayed as captured graphs
    t1 = time.perf_counter()
    sess.run(args.warmup, args.steps, m["dt"])
    t2 = time.perf_counter()
    run_ms = sess.last_run_timing()[0]
    # then a short pass with HIP events around every interior launch (on the
    # session stream the kernels run on) for the kernel's average duration
    int_avg_ms, n_int = 0.0, 0
    if args.timed_launch_steps > 0:
        sess.set_timing(True)
        sess.run(args.warmup + args.steps, args.timed_launch_steps, m["dt"])
        _, int_avg_ms, n_int = sess.last_run_timing()
        step_ms_timed = sess.last_run_timing()[0] / args.timed_launch_steps
    seis = sess.read_seismo(0)
    finite = bool(np.isfinite(seis).all())
    tile = sess.get_tile()
    sess.close()
    wall = t2 - t1
    cells = workloads.cells(m)
    value = cells * 4 * args.steps / wall / 1e9
    log(f"{value:.3f} Gcell-RK-stages/s wall, device {run_ms:.1f} ms, interior avg {int_avg_ms:.4f} ms x {n_int}")

    roof = None
    if n_int > 0 and int_avg_ms > 0:
        icells = workloads.interior_cells(m)
        ach = ALGO_BYTES[args.storage] * icells / (int_avg_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None,
                    kernel="k_march (interior z-marching sweep)", interior_cells=icells, avg_launch_ms=round(int_avg_ms, 5),
                    tile=f"{tile['tx']}x{tile['ty']} threads, {tile['zchunk']} planes",
                    interior_share_of_step=round(4 * int_avg_ms / step_ms_timed, 4))
        traffic = measured_traffic(m["name"], args.storage)
        if traffic:
            roof["traffic"] = traffic

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline: {args.cpu_steps} steps")
        cpu = cpu_baseline(m, args.cpu_steps)

    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gcell-RK-stages/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.storage,
        "data": "synthetic",
        "config": {"workload": f"{m['name']}: {m['ni']}x{m['nj']}x{m['nk']} el_iso RK4, "
                               f"{'curvilinear topography, 3 perturbed layers, double couple, 400 receivers' if m['name'] == 'topo' else 'Cartesian homogeneous, explosion, 20 receivers'}, "
                               f"free surface + 10-layer CFS-PML, dt {m['dt']}",
                   "cells": cells, "storage": args.storage, "kernel_path": "plain" if args.plain else "zmarch",
                   "seismograms_finite": finite, "device_ms": round(run_ms, 2)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
