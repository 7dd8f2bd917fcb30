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
# per (point, PML face) pair in the sweep: nine memory variables, mean of
# 27/45/45/27 words (memory in, RK levels in and out)
AUX_BYTES = {"f32": 144.0, "f16": 72.0, "bf16": 72.0, "f64": 288.0}
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


def measured_traffic(workload, storage, split, tile):
    """HBM bytes per sweep launch from the committed counter summary, if any.

    The summary holds FETCH_SIZE + WRITE_SIZE per dispatch of the z-marching
    kernel on this workload and work split, scaled by the same counters on a
    copy kernel of known size; a line keyed by the tile ("[32x8/16]") wins
    over an unkeyed one."""
    name = "planes" if split == cgfd.SPLIT_PLANES else "shell"
    path = os.path.join(ROOT, "profiles", f"k_march_{workload}_{storage}_{name}_hbm_counters.txt")
    if not os.path.exists(path):
        return None
    key = f"[{tile['tx']}x{tile['ty']}/{tile['zchunk']}]"
    plain = None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch" + key):
            return float(line.split("=")[1])
        if line.startswith("traffic_bytes_per_launch ") or line.startswith("traffic_bytes_per_launch="):
            plain = float(line.split("=")[1])
    return plain


STREAM_SIZES = ((64, 64, 64), (128, 128, 96), (256, 256, 128), (400, 400, 200), (512, 512, 256))


def served_stream_sample(n, nt_range, load, seed, nseq_per_size, nrec=20):
    """Served stream: n seeded shot requests on five resident topography
    models (64^3, 128x128x96, 256x256x128, 400x400x200, 512x512x256; seeded
    hills, 3 perturbed layers, 10-layer CFS-PML, free surface), each request
    on a uniformly drawn model, FP16 storage with probability 1/2, random
    source and steps in nt_range, arriving as a Poisson stream at `load`
    times the rate of sequential plain runs.  Sequential plain run = the
    request alone in an FP32 single-shot session of its model (graphs on),
    timed on nseq_per_size requests per model and scaled by steps.  Returns
    shots/hour both ways and the served latency percentiles."""
    rng = np.random.default_rng(seed)
    models, reqs = [], []
    for q, (ni, nj, nk) in enumerate(STREAM_SIZES):
        m = workloads.topo(seed=seed + 10 * q, nt=nt_range[1], ni=ni, nj=nj, nk=nk)
        m["recv_indx"], m["recv_frac"] = m["recv_indx"][:nrec], m["recv_frac"][:nrec]
        models.append(m)
    pick = rng.integers(0, len(models), n)
    fp16 = rng.random(n) < 0.5
    per_model = [workloads.shot_requests(m, n, seed + 1000 + q, nt_range=nt_range) for q, m in enumerate(models)]
    for r in range(n):
        reqs.append(dict(model=int(pick[r]), fp16=bool(fp16[r]), **per_model[pick[r]][r]))
    # sequential plain runs: per model, a few requests alone in an FP32 single-shot session
    step_s, alone = {}, {}
    for q, m in enumerate(models):
        sess = cgfd.Session(m["ni"], m["nj"], m["nk"])
        workloads.upload(sess, m, nt_total=8)
        sess.run(0, 8, m["dt"])  # autotune and first captures outside the clock
        mine = [r for r in range(n) if reqs[r]["model"] == q][:nseq_per_size]
        t_used, steps = 0.0, 0
        for r in mine:
            sess.reset_shot(0)
            sess.upload_source(0, reqs[r]["src"])
            sess.upload_receivers(0, m["recv_indx"], np.asarray(m["recv_frac"], np.float32), reqs[r]["nt"])
            t0 = time.perf_counter()
            sess.run(0, reqs[r]["nt"], m["dt"])
            sess.read_seismo(0)
            t_used += time.perf_counter() - t0
            steps += reqs[r]["nt"]
            alone[r] = sess.read_seismo(0)
        sess.close()
        step_s[q] = t_used / steps if steps else 0.0
    t_seq = sum(reqs[r]["nt"] * step_s[reqs[r]["model"]] for r in range(n))
    seq_rate = n / t_seq
    # served: every model resident in FP32 and FP16 form, continuous batching
    srv = cgfd.Server(max_batch=16, chunk_steps=16)
    mid = {}
    for q, m in enumerate(models):
        arrs = workloads.model_arrays(m)
        ex = np.max([cgfd.storage_scales(cgfd.STORE_F16, m["ni"], m["nj"], m["nk"], m["metric"], m["lam"], m["mu"],
                                         m["slw"], reqs[r]["src"], m["dt"]) for r in range(n)
                     if reqs[r]["model"] == q and reqs[r]["fp16"]] or [[0] * 9], axis=0)
        for h in (False, True):
            mid[(q, h)] = srv.add_model(m["ni"], m["nj"], m["nk"], dt=m["dt"],
                                        storage=cgfd.STORE_F16 if h else cgfd.STORE_F32,
                                        scale_exp2=ex if h else None, max_nrecv=nrec, max_nt=nt_range[1], **arrs)
        del arrs
    # warm every group (autotune, graph capture) with one short request each
    for key in mid:
        m = models[key[0]]
        t = srv.submit(mid[key], 16, per_model[key[0]][0]["src"], m["recv_indx"], m["recv_frac"])
        srv.wait(t)
        srv.result(t)
    gaps = rng.exponential(1.0 / (load * seq_rate), n)
    tickets = []
    start = time.perf_counter()
    t_next = start
    for r in range(n):
        t_next += gaps[r]
        while time.perf_counter() < t_next:
            time.sleep(0.0005)
        m = models[reqs[r]["model"]]
        tickets.append(srv.submit(mid[(reqs[r]["model"], reqs[r]["fp16"])], reqs[r]["nt"], reqs[r]["src"],
                                  m["recv_indx"], m["recv_frac"]))
    lat, errs = [], []
    for r, t in enumerate(tickets):
        srv.wait(t)
        seis, _, t_done = srv.result(t)
        lat.append(t_done)
        if r in alone and not reqs[r]["fp16"]:
            d = np.linalg.norm(alone[r])
            errs.append(float(np.linalg.norm(seis - alone[r]) / d) if d > 0 else 0.0)
    end = time.perf_counter()
    stats = srv.stats()
    srv.close()
    served_rate = n / (end - start)
    lat = np.array(lat)
    sizes = "/".join(f"{a}x{b}x{c}" for a, b, c in STREAM_SIZES)
    return dict(
        workload=f"{n} shot requests on 5 resident topography models ({sizes}), model uniform, steps "
                 f"U{list(nt_range)}, f0 U[1,4] Hz, {int(fp16.sum())} FP16-storage + {int((~fp16).sum())} FP32, "
                 f"Poisson arrivals at {load}x the sequential rate",
        served_shots_per_hour=round(served_rate * 3600, 1),
        sequential_plain_shots_per_hour=round(seq_rate * 3600, 1),
        speedup=round(served_rate / seq_rate, 3),
        p50_latency_s=round(float(np.percentile(lat, 50)), 4),
        p99_latency_s=round(float(np.percentile(lat, 99)), 4),
        sequential_sample=f"{nseq_per_size} requests per model timed alone in an FP32 single-shot session, "
                          f"per-step time scaled to all {sum(r['nt'] for r in reqs)} steps",
        sequential_ms_per_step={f"{a}x{b}x{c}": round(step_s[q] * 1e3, 4) for q, (a, b, c) in enumerate(STREAM_SIZES)},
        served_vs_alone_max_rel_l2=max(errs) if errs else None,
        groups=stats["groups"], rejected=stats["rejected"])


def batch_sample(n, nt_range, seed, nseq, storage):
    """n shot requests on the 100x100x60 medium submitted at once to the
    server (one batched group, launches per RK stage shared by all its
    shots, 8-step periods replayed as graphs); shots/hour against the same
    shots run one after another in an FP32 single-shot session."""
    m = workloads.small()
    reqs = workloads.shot_requests(m, n, seed, nt_range=nt_range)
    ni, nj, nk, dt = m["ni"], m["nj"], m["nk"], m["dt"]
    nrec = len(m["recv_indx"])
    sess = cgfd.Session(ni, nj, nk)
    workloads.upload(sess, m, nt_total=8)
    sess.run(0, 8, dt)
    t0 = time.perf_counter()
    steps_seq = 0
    for q in range(nseq):
        r = reqs[q]
        sess.reset_shot(0)
        sess.upload_source(0, r["src"])
        sess.upload_receivers(0, m["recv_indx"], np.asarray(m["recv_frac"], np.float32), r["nt"])
        sess.run(0, r["nt"], dt)
        sess.read_seismo(0)
        steps_seq += r["nt"]
    t_seq = (time.perf_counter() - t0) * sum(r["nt"] for r in reqs) / steps_seq
    sess.close()
    st = STORE[storage]
    ex = None
    if st == cgfd.STORE_F16:
        ex = np.max([cgfd.storage_scales(st, ni, nj, nk, m["metric"], m["lam"], m["mu"], m["slw"], r["src"], dt)
                     for r in reqs], axis=0)
    srv = cgfd.Server(max_batch=n, chunk_steps=64)
    mid = srv.add_model(ni, nj, nk, dt=dt, storage=st, scale_exp2=ex, max_nrecv=nrec, max_nt=nt_range[1],
                        **workloads.model_arrays(m))
    t = srv.submit(mid, 16, reqs[0]["src"], m["recv_indx"], m["recv_frac"])  # warm: autotune, capture
    srv.wait(t)
    srv.result(t)
    start = time.perf_counter()
    tickets = [srv.submit(mid, r["nt"], r["src"], m["recv_indx"], m["recv_frac"]) for r in reqs]
    lat = []
    for t in tickets:
        srv.wait(t)
        lat.append(srv.result(t)[2])
    end = time.perf_counter()
    srv.close()
    rate, seq = n / (end - start), n / t_seq
    return dict(workload=f"{n} shot requests submitted at once on the 100x100x60 medium, steps U{list(nt_range)}, "
                         f"f0 U[1,4] Hz, {storage} storage, one batched group",
                shots_per_hour=round(rate * 3600, 1), sequential_plain_shots_per_hour=round(seq * 3600, 1),
                speedup=round(rate / seq, 3), p50_latency_s=round(float(np.percentile(lat, 50)), 3),
                p99_latency_s=round(float(np.percentile(lat, 99)), 3))


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
    ap.add_argument("--pml-point", action="store_true",
                    help="PML slabs by the per-point kernel (the sweep keeps to the box inside them)")
    ap.add_argument("--tile", default="", metavar="SHAPE,ZCHUNK",
                    help="pin the interior tile instead of autotuning (profiling runs)")
    ap.add_argument("--timed-launch-steps", type=int, default=8,
                    help="steps of the per-launch event timing pass (0: skip)")
    ap.add_argument("--reduced-steps", type=int, default=200,
                    help="timed steps of the FP16/BF16-storage runs of the same workload (0: skip)")
    ap.add_argument("--serve-shots", type=int, default=40, help="served-stream sample size (0: skip)")
    ap.add_argument("--batch-shots", type=int, default=64, help="batched-shots sample size (0: skip)")
    ap.add_argument("--batch-steps", default="500,2000", help="step range of the batched shots")
    ap.add_argument("--serve-steps", default="40,160", help="step range of the served requests")
    ap.add_argument("--serve-load", type=float, default=4.0, help="arrival rate / sequential plain rate")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (counter calibration)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch  # plumbing only: a plain copy of known size for HBM counter scaling

        n = args.copy_calib * 1024 * 1024 // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        y.copy_(x)
        e0.record()
        for _ in range(5):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        print(json.dumps({"copy_calib_bytes_each_way": n * 4, "copies": 6, "ms_per_copy": round(ms, 4),
                          "copy_GBps_read_plus_write": round(2 * n * 4 / ms / 1e6, 1)}), flush=True)
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
    flags = (cgfd.FLAG_PLAIN if args.plain else 0) | (cgfd.FLAG_PML_POINT if args.pml_point else 0)
    sess = cgfd.Session(m["ni"], m["nj"], m["nk"], storage=STORE[args.storage], flags=flags)
    nt_total += args.timed_launch_steps
    workloads.upload(sess, m, nt_total=nt_total)
    if args.storage == "f16":
        # FP16 needs power-of-two scales to keep stresses in range (BF16 has FP32's exponent range)
        sess.set_storage_scale(cgfd.storage_scales(cgfd.STORE_F16, m["ni"], m["nj"], m["nk"], m["metric"], m["lam"],
                                                   m["mu"], m["slw"], m["src"], m["dt"]))
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
    split = sess.get_split()
    sess.close()
    wall = t2 - t1
    cells = workloads.cells(m)
    value = cells * 4 * args.steps / wall / 1e9
    log(f"{value:.3f} Gcell-RK-stages/s wall, device {run_ms:.1f} ms, interior avg {int_avg_ms:.4f} ms x {n_int}")

    roof = None
    if n_int > 0 and int_avg_ms > 0:
        if split == cgfd.SPLIT_SHELL:
            icells, pairs = workloads.interior_cells(m), 0
        else:
            icells, pairs = workloads.sweep_cells(m), workloads.sweep_pml_pairs(m)
        algo = ALGO_BYTES[args.storage] * icells + AUX_BYTES[args.storage] * pairs
        ach = algo / (int_avg_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None,
                    kernel="k_march (z-marching sweep" + (", inside the PML slabs)" if split == cgfd.SPLIT_SHELL
                                                          else ", whole planes with the PML slabs)"),
                    sweep_cells=icells, pml_face_pairs=pairs, algorithmic_bytes_per_launch=algo,
                    avg_launch_ms=round(int_avg_ms, 5),
                    tile=f"{tile['tx']}x{tile['ty']} threads, {tile['zchunk']} planes",
                    sweep_share_of_step=round(4 * int_avg_ms / step_ms_timed, 4))
        traffic = None if args.plain else measured_traffic(m["name"], args.storage, split, tile)
        if traffic:
            roof["traffic"] = traffic

    reduced = None
    if args.reduced_steps > 0 and args.storage == "f32":
        reduced = {}
        ex = cgfd.storage_scales(cgfd.STORE_F16, m["ni"], m["nj"], m["nk"], m["metric"], m["lam"], m["mu"], m["slw"],
                                 m["src"], m["dt"])
        # FP16 / BF16 storage, and the FP64 path (the GPU oracle of the same scheme) on fewer steps
        for name, st in (("f16", cgfd.STORE_F16), ("bf16", cgfd.STORE_BF16), ("f64", cgfd.STORE_F64)):
            nst = args.reduced_steps if st != cgfd.STORE_F64 else max(8, args.reduced_steps // 4)
            rs = cgfd.Session(m["ni"], m["nj"], m["nk"], storage=st)
            workloads.upload(rs, m, nt_total=args.warmup + nst)
            if st == cgfd.STORE_F16:
                rs.set_storage_scale(ex)
            rs.run(0, args.warmup, m["dt"])
            ta = time.perf_counter()
            rs.run(args.warmup, nst, m["dt"])
            tb = time.perf_counter()
            fin = bool(np.isfinite(rs.read_seismo(0)).all())
            rs.close()
            v = cells * 4 * nst / (tb - ta) / 1e9
            reduced[name] = dict(value=round(v, 4), unit="Gcell-RK-stages/s",
                                 ms_per_step=round((tb - ta) * 1e3 / nst, 4),
                                 speedup_vs_f32=round(v / value, 3), steps=nst,
                                 scales_exp2=ex if st == cgfd.STORE_F16 else [0] * 9, seismograms_finite=fin)
            log(f"{name} storage: {v:.3f} Gcell-RK-stages/s ({v / value:.2f}x FP32)")

    serving = None
    if args.serve_shots > 0:
        lo, hi = (int(v) for v in args.serve_steps.split(","))
        log(f"serving sample: {args.serve_shots} shots")
        serving = served_stream_sample(args.serve_shots, (lo, hi), args.serve_load, 7, 2)
        log(f"served {serving['served_shots_per_hour']} shots/h vs sequential "
            f"{serving['sequential_plain_shots_per_hour']}; p50 {serving['p50_latency_s']} s p99 "
            f"{serving['p99_latency_s']} s")

    batched = None
    if args.batch_shots > 0:
        lo, hi = (int(v) for v in args.batch_steps.split(","))
        batched = {}
        for st in ("f32", "f16"):
            log(f"batched shots: {args.batch_shots} x {st}")
            batched[st] = batch_sample(args.batch_shots, (lo, hi), 11, 4, st)
            log(f"  {batched[st]['shots_per_hour']} shots/h vs sequential "
                f"{batched[st]['sequential_plain_shots_per_hour']} ({batched[st]['speedup']}x)")

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
                   "cells": cells, "storage": args.storage,
                   "kernel_path": "plain" if args.plain else ("zmarch planes" if split == cgfd.SPLIT_PLANES
                                                                 else "zmarch + point shell"),
                   "seismograms_finite": finite, "device_ms": round(run_ms, 2)},
        "roofline": roof,
        "reduced_storage": reduced,
        "serving": serving,
        "batched_shots": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
