#!/usr/bin/env python3
"""el_iso RK4 time-stepping throughput on one MI355X.

Headline workload: the 800x800x400 explosion model -- the 100x100x60
configuration's physics (Cartesian 100 m grid, homogeneous Vp 6000 / Vs
3464 / rho 2670, flat free surface, 10-layer CFS-PML on the other five
faces, explosive moment tensor Mxx=Myy=Mzz=1e16 N m 3 km deep, Ricker 2 Hz,
dt 0.005 s, 20 surface receivers) at the size of the FP32-vs-FP16-store
comparison -- stepped in FP32 through the C ABI of
gpu-cgfd3d-elastic_amd/libcgfd_el_iso.so.  A "step" is one full RK4 step (4
stages) over the whole grid with receivers recorded; W warmup steps run
first, then exactly K steps are timed.  Inputs are resident in HBM before
the timed region.

The JSON line also carries: the HBM roofline of the dominant kernel (the
z-marching sweep, average launch time from HIP events on its stream, 192
algorithmic bytes per swept point and stage) beside the streaming read/copy
ceilings measured in the same run; FP16 / BF16 storage and the FP64 path on
the same model with their relative L2 errors against FP64; the 400x400x200
topography model (curvilinear metrics, all 12 coefficient streams); 64
batched shots on the small medium; a served stream over five resident
models; and the CPU baseline (the oracle, a C++/OpenMP restatement of the
reference loop) on a bounded sample of the headline model.
"""
import argparse
import json
import os
import platform
import re
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
# work splits of a stage (cgfd_get_split): profile names, the swept kernel(s) and the path
SPLIT_NAMES = {cgfd.SPLIT_PLANES: "planes", cgfd.SPLIT_SHELL: "shell", cgfd.SPLIT_CORE: "core"}
SWEEP_KERNEL = {
    cgfd.SPLIT_PLANES: "k_march (z-marching sweep, whole planes with the PML slabs)",
    cgfd.SPLIT_SHELL: "k_march (z-marching sweep, inside the PML slabs)",
    cgfd.SPLIT_CORE: "k_march (z-marching sweep, whole planes: PML-free core launch + PML-form rim launch, "
                     "timed together per stage)",
}
KERNEL_PATH = {cgfd.SPLIT_PLANES: "zmarch planes", cgfd.SPLIT_SHELL: "zmarch + point shell",
               cgfd.SPLIT_CORE: "zmarch core + rim"}
# The roofline unit: algorithmic HBM bytes per interior cell and RK stage of
# the reference's dataflow with the RK update fused, mean of 39/57/57/39
# words (wavefield in, 12 coefficients, RK levels in and out).
ALGO_BYTES = {"f32": 192.0, "f16": 120.0, "bf16": 120.0, "f64": 384.0}
# This build's own dataflow (stage values): wavefield words per swept point
# and stage, S0 9 in + 9 out, S1/S2 9 in + 9 W^n + 9 out, S3 9 in + 27 (W^n,
# U1, U2) + 9 out, mean 29.25; plus 12 coefficient words (3, the medium,
# when the metric is constant and passed as kernel arguments).  The nine
# PML memory variables of a (point, face) pair follow the same pattern.
WAVE_WORDS = 29.25
ES_WAVE = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}
ES_MODEL = {"f32": 4, "f16": 4, "bf16": 4, "f64": 8}


def path_bytes(storage, const_metric):
    return WAVE_WORDS * ES_WAVE[storage] + (3 if const_metric else 12) * ES_MODEL[storage]


AUX_BYTES = {k: WAVE_WORDS * v for k, v in ES_WAVE.items()}
STORE = {"f32": cgfd.STORE_F32, "f16": cgfd.STORE_F16, "bf16": cgfd.STORE_BF16, "f64": cgfd.STORE_F64}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """(model name, logical CPUs, sockets, physical cores) from /proc/cpuinfo."""
    model, n, phys, cores = "", 0, set(), set()
    pid = None
    try:
        for line in open("/proc/cpuinfo"):
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "model name":
                model = model or val
                n += 1
            elif key == "physical id":
                pid = val
                phys.add(val)
            elif key == "core id":
                cores.add((pid, val))
    except OSError:
        pass
    return model or platform.processor(), n, len(phys), len(cores)


def cpu_baseline(m, steps):
    """The oracle (C++/OpenMP restatement of the reference's unfused per-point
    loop, FP32) on `steps` full RK4 steps of workload m."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
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
    model, ncpu, nsock, ncore = host_cpu()
    return dict(value=round(cells * 4 * steps / sec / 1e9, 5), unit="Gcell-RK-stages/s",
                cores=int(L.orc_max_threads()), kind="port",
                sample=f"{steps} full RK4 steps of the same {m['ni']}x{m['nj']}x{m['nk']} model, FP32, unfused "
                       f"reference stage order (RHS then axpys), {sec:.1f} s loop",
                host=f"{model}; {ncpu} logical CPUs visible, {nsock} socket(s), {ncore} physical cores; "
                     f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} OMP_PLACES={os.environ.get('OMP_PLACES')}")


def measured_traffic(workload, storage, split, tile):
    """(HBM bytes per sweep launch, profiled tile) from the committed counter
    summary, or (None, None).

    The summary holds FETCH_SIZE + WRITE_SIZE per dispatch of the z-marching
    kernel on this workload and work split, scaled by the same counters on
    streaming kernels of known size, one line per profiled tile
    ("traffic_bytes_per_launch[32x8/20] = ...").  The line of the tile this
    run uses wins; otherwise the line of the same thread shape with the
    nearest z chunk (the chunk changes only the priming planes), and the
    profiled tile is reported beside the figure."""
    name = SPLIT_NAMES[split]
    path = os.path.join(ROOT, "profiles", f"k_march_{workload}_{storage}_{name}_hbm_counters.txt")
    if not os.path.exists(path):
        return None, None
    best = None
    for line in open(path):
        mt = re.match(r"traffic_bytes_per_launch\[(\d+)x(\d+)/(\d+)\]\s*=\s*([0-9.eE+]+)", line)
        if not mt:
            continue
        tx, ty, zc, val = int(mt.group(1)), int(mt.group(2)), int(mt.group(3)), float(mt.group(4))
        if (tx, ty) != (tile["tx"], tile["ty"]):
            continue
        d = abs(zc - tile["zchunk"])
        if best is None or d < best[0]:
            best = (d, val, f"{tx}x{ty}/{zc}")
    return (best[1], best[2]) if best else (None, None)


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


STORE_NAMES = {"f32": "FP32", "f16": "FP16 storage, FP32 arithmetic", "bf16": "BF16 storage, FP32 arithmetic",
               "f64": "FP64 (the GPU oracle of the same scheme)"}


def scales_for(m, st, src=None):
    return cgfd.storage_scales(st, m["ni"], m["nj"], m["nk"], m["metric"], m["lam"], m["mu"], m["slw"],
                               m["src"] if src is None else src, m["dt"])


def timed_session(m, storage, warmup, steps, launch_steps, flags=0, tile="", split=""):
    """Upload m into a single-shot session, run warmup steps, time `steps`
    steps (8-step periods replayed as captured graphs), then a short pass with
    HIP events around every sweep launch.  Returns a dict of the measurements."""
    st = STORE[storage]
    nt_total = warmup + steps + launch_steps
    sess = cgfd.Session(m["ni"], m["nj"], m["nk"], storage=st, flags=flags)
    workloads.upload(sess, m, nt_total=nt_total)
    if st == cgfd.STORE_F16:
        sess.set_storage_scale(scales_for(m, st))
    if split:
        sess.set_split({v: k for k, v in SPLIT_NAMES.items()}[split])
    if tile:
        sh, zc = (int(v) for v in tile.split(","))
        sess.set_tile(sh, zc)
    sess.run(0, warmup, m["dt"])
    t1 = time.perf_counter()
    sess.run(warmup, steps, m["dt"])
    t2 = time.perf_counter()
    out = dict(wall=t2 - t1, device_ms=sess.last_run_timing()[0], int_avg_ms=0.0, n_int=0, step_ms_timed=0.0)
    if launch_steps > 0:
        sess.set_timing(True)
        sess.run(warmup + steps, launch_steps, m["dt"])
        run_ms, out["int_avg_ms"], out["n_int"] = sess.last_run_timing()
        out["step_ms_timed"] = run_ms / launch_steps
    out["finite"] = bool(np.isfinite(sess.read_seismo(0)).all())
    out["tile"] = sess.get_tile()
    out["split"] = sess.get_split()
    out["const_metric"] = sess.const_metric()
    sess.close()
    return out


def roofline(m, storage, r, stream):
    """Dominant kernel: the z-marching sweep.  achieved = 192 B (FP32 unit of
    the fused RK4 dataflow, interior point and stage) x swept points / the
    sweep's average launch time; the PML memory traffic and the path's own
    needs are reported beside it."""
    if r["n_int"] == 0 or r["int_avg_ms"] <= 0:
        return None
    split = r["split"]
    if split == cgfd.SPLIT_SHELL:
        cells, pairs = workloads.interior_cells(m), 0
    else:
        cells, pairs = workloads.sweep_cells(m), workloads.sweep_pml_pairs(m)
    unit = ALGO_BYTES[storage]
    sec = r["int_avg_ms"] * 1e-3
    ach = unit * cells / sec / 1e9
    path_unit = path_bytes(storage, r["const_metric"])
    tile = r["tile"]
    roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4),
                traffic=None,
                kernel=SWEEP_KERNEL[split],
                bytes_per_point_stage=unit, swept_points=cells, avg_launch_ms=round(r["int_avg_ms"], 5),
                algorithmic_bytes_per_launch=unit * cells,
                tile=f"{tile['tx']}x{tile['ty']} threads, {tile['zchunk']} planes",
                tile_arg=f"{tile['shape']},{tile['zchunk']}",
                sweep_share_of_step=round(4 * r["int_avg_ms"] / r["step_ms_timed"], 4) if r["step_ms_timed"] else None,
                const_metric=r["const_metric"],
                path_bytes_per_point_stage=path_unit,
                path_GBps=round(path_unit * cells / sec / 1e9, 1),
                pml_face_pairs=pairs,
                path_with_pml_GBps=round((path_unit * cells + AUX_BYTES[storage] * pairs) / sec / 1e9, 1))
    if stream:
        roof["stream_read_GBps"] = round(stream[0], 1)
        roof["stream_copy_GBps"] = round(stream[1], 1)
        roof["stream_read_dword_GBps"] = round(stream[2], 1)
        roof["path_frac_of_stream_read"] = round(roof["path_GBps"] / stream[0], 4)
    traffic, ttile = measured_traffic(m["name"], storage, split, tile)
    if traffic:
        roof["traffic"] = traffic
        roof["traffic_tile"] = ttile
        roof["traffic_GBps"] = round(traffic / sec / 1e9, 1)
    return roof


def _rel(a, b):
    return float(f"{np.linalg.norm(a - b) / np.linalg.norm(b):.4g}")


def precision_leg(m, steps):
    """FP64, FP32, FP16 and BF16 storage from the same start for `steps`
    steps: throughput of each (steps 8.. timed) and relative L2 errors
    against the FP64 run (and against FP32 for the 16-bit forms): the
    velocity seismograms (Vx, Vy, Vz at every receiver, what the stations
    record), all nine components, a stride-8 sample of the final wavefield,
    and the velocity-seismogram error over the first n steps for a growing n
    (the error's growth with run length)."""
    out, ref, ref32 = {}, None, None
    cells = workloads.cells(m)
    ni, nj, nk = m["ni"], m["nj"], m["nk"]
    cnt = ((ni - 4) // 8 + 1, (nj - 4) // 8 + 1, (nk - 4) // 8 + 1)
    marks = [n for n in (80, 160, 320, 640, 1000, 1500, 2000) if n < steps] + [steps]
    for name in ("f64", "f32", "f16", "bf16"):
        st = STORE[name]
        s = cgfd.Session(ni, nj, nk, storage=st)
        workloads.upload(s, m, nt_total=steps)
        ex = None
        if st == cgfd.STORE_F16:
            ex = scales_for(m, st)
            s.set_storage_scale(ex)
        s.run(0, 8, m["dt"])   # autotune and first graph captures outside the clock
        ta = time.perf_counter()
        it = 8
        while it < steps:   # chunks of at most 400 steps, each logged
            n = min(400, steps - it)
            s.run(it, n, m["dt"])
            it += n
            if it < steps:
                log(f"{name}: step {it} of {steps}")
        tb = time.perf_counter()
        seis = s.read_seismo(0)
        wave = s.fetch_subvolume(0, cgfd.LEVEL_PRE, list(range(9)), (2, 2, 2), cnt, (8, 8, 8), dtype=np.float64)
        s.close()
        v = cells * 4 * (steps - 8) / (tb - ta) / 1e9
        rec = dict(value=round(v, 4), unit="Gcell-RK-stages/s", ms_per_step=round((tb - ta) * 1e3 / (steps - 8), 4),
                   steps=steps, scales_exp2=ex if ex is not None else [0] * 9,
                   seismograms_finite=bool(np.isfinite(seis).all()))
        if name == "f64":
            ref = (seis, wave)
        else:
            rec["rel_l2_velocity_seismograms_vs_f64"] = _rel(seis[:, :3], ref[0][:, :3])
            rec["rel_l2_seismograms_vs_f64"] = _rel(seis, ref[0])
            rec["rel_l2_wavefield_vs_f64"] = _rel(wave, ref[1])
            rec["rel_l2_velocity_seismograms_vs_f64_by_steps"] = {
                str(n): _rel(seis[:, :3, :n], ref[0][:, :3, :n]) for n in marks}
            if name == "f32":
                ref32 = (seis, wave)
            else:
                rec["rel_l2_velocity_seismograms_vs_f32"] = _rel(seis[:, :3], ref32[0][:, :3])
                rec["rel_l2_wavefield_vs_f32"] = _rel(wave, ref32[1])
        out[name] = rec
        log(f"{name}: {v:.3f} Gcell-RK-stages/s over {steps - 8} steps" +
            (f", rel L2 vs FP64 velocity seismograms {rec['rel_l2_velocity_seismograms_vs_f64']:.3e}, all nine "
             f"{rec['rel_l2_seismograms_vs_f64']:.3e}, wavefield {rec['rel_l2_wavefield_vs_f64']:.3e}"
             if name != "f64" else ""))
    for name in ("f16", "bf16", "f64"):
        out[name]["speedup_vs_f32"] = round(out[name]["value"] / out["f32"]["value"], 3)
    return out


def topo_leg(steps, warmup, stream):
    """The 400x400x200 topography model (curvilinear metric, all coefficient
    streams): FP32 with its sweep roofline, FP16 and BF16 storage."""
    m = workloads.topo(nt=warmup + steps + 8)
    cells = workloads.cells(m)
    out = {}
    for name in ("f32", "f16", "bf16"):
        r = timed_session(m, name, warmup, steps, 8 if name == "f32" else 0)
        v = cells * 4 * steps / r["wall"] / 1e9
        out[name] = dict(value=round(v, 4), unit="Gcell-RK-stages/s", ms_per_step=round(r["wall"] * 1e3 / steps, 4),
                         steps=steps, seismograms_finite=r["finite"],
                         split=SPLIT_NAMES[r["split"]])
        if name == "f32":
            out[name]["roofline"] = roofline(m, name, r, stream)
        log(f"topography {name}: {v:.3f} Gcell-RK-stages/s")
    for name in ("f16", "bf16"):
        out[name]["speedup_vs_f32"] = round(out[name]["value"] / out["f32"]["value"], 3)
    out["workload"] = ("topo: 400x400x200 el_iso RK4, curvilinear topography, 3 perturbed layers, double couple, "
                       "400 receivers, free surface + 10-layer CFS-PML, dt 0.005")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--workload", default="explosion800", choices=["explosion800", "topo", "small"])
    ap.add_argument("--storage", default="f32", choices=sorted(STORE))
    ap.add_argument("--cpu-steps", type=int, default=4, help="steps of the CPU baseline sample (0: skip)")
    ap.add_argument("--plain", action="store_true", help="per-point kernel everywhere")
    ap.add_argument("--pml-point", action="store_true",
                    help="PML slabs by the per-point kernel (the sweep keeps to the box inside them)")
    ap.add_argument("--full-metric", action="store_true", help="keep the per-point metric streams")
    ap.add_argument("--tile", default="", metavar="SHAPE,ZCHUNK",
                    help="pin the interior tile instead of autotuning (profiling runs)")
    ap.add_argument("--split", default="", choices=["", "planes", "core", "shell"],
                    help="pin the work split (default: the autotuner chooses)")
    ap.add_argument("--timed-launch-steps", type=int, default=8,
                    help="steps of the per-launch event timing pass (0: skip)")
    ap.add_argument("--error-steps", type=int, default=320,
                    help="steps of the FP64/FP32/FP16/BF16 runs compared against FP64 (0: skip)")
    ap.add_argument("--topo-steps", type=int, default=48, help="timed steps of the topography legs (0: skip)")
    ap.add_argument("--stream-mib", type=int, default=2048, help="buffer size of the streaming ceilings (0: skip)")
    ap.add_argument("--serve-shots", type=int, default=40, help="served-stream sample size (0: skip)")
    ap.add_argument("--batch-shots", type=int, default=64, help="batched-shots sample size (0: skip)")
    ap.add_argument("--batch-steps", default="500,2000", help="step range of the batched shots")
    ap.add_argument("--serve-steps", default="40,160", help="step range of the served requests")
    ap.add_argument("--serve-load", type=float, default=4.0, help="arrival rate / sequential plain rate")
    ap.add_argument("--only-headline", action="store_true", help="skip every leg after the headline (profiling)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="run the streaming read and copy kernels once over MIB MiB and exit (the known-size "
                         "dispatches that scale the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(the decomposition is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        r, c = cgfd.hbm_stream(args.copy_calib << 20, 1)
        rd = cgfd.hbm_stream_dword(args.copy_calib << 20, 1)
        print(json.dumps({"copy_calib_mib": args.copy_calib, "read_GBps": round(r, 1), "copy_GBps": round(c, 1),
                          "read_dword_GBps": round(rd, 1)}))
        return
    if args.only_headline:
        args.error_steps = args.topo_steps = args.stream_mib = args.serve_shots = args.batch_shots = 0
        args.cpu_steps = 0

    t0 = time.time()
    nt = args.warmup + args.steps + args.timed_launch_steps
    if args.workload == "explosion800":
        m = workloads.small(ni=800, nj=800, nk=400, nt=max(nt, args.error_steps))
    elif args.workload == "topo":
        m = workloads.topo(nt=nt)
    else:
        m = workloads.small(nt=max(nt, args.error_steps))
    log(f"workload {m['name']} {m['ni']}x{m['nj']}x{m['nk']} prepared in {time.time() - t0:.1f}s; "
        f"dt {m['dt']} (dtmax {m['dtmax']:.5f})")
    if m["dt"] > m["dtmax"]:
        raise SystemExit(f"dt {m['dt']} exceeds the stability limit {m['dtmax']}")

    flags = (cgfd.FLAG_PLAIN if args.plain else 0) | (cgfd.FLAG_PML_POINT if args.pml_point else 0) | \
        (cgfd.FLAG_FULL_METRIC if args.full_metric else 0)
    log(f"timing {args.steps} steps ({args.warmup} warmup)")
    r = timed_session(m, args.storage, args.warmup, args.steps, args.timed_launch_steps, flags, args.tile, args.split)
    cells = workloads.cells(m)
    value = cells * 4 * args.steps / r["wall"] / 1e9
    log(f"{value:.3f} Gcell-RK-stages/s wall, device {r['device_ms']:.1f} ms, sweep avg {r['int_avg_ms']:.4f} ms "
        f"x {r['n_int']}, tile {r['tile']}, constant metric {r['const_metric']}")

    stream = None
    if args.stream_mib > 0:
        stream = cgfd.hbm_stream(args.stream_mib << 20, 10) + (cgfd.hbm_stream_dword(args.stream_mib << 20, 10),)
        log(f"streaming ceilings: read {stream[0]:.0f} GB/s, copy {stream[1]:.0f} GB/s (read + write), "
            f"4-byte-load read {stream[2]:.0f} GB/s")
    roof = roofline(m, args.storage, r, stream) if not args.plain else None

    precision = None
    if args.error_steps > 0 and args.storage == "f32":
        log(f"precision leg: FP64 / FP32 / FP16 / BF16, {args.error_steps} steps each")
        precision = precision_leg(m, args.error_steps)

    topo = None
    if args.topo_steps > 0 and args.workload != "topo":
        log(f"topography legs: {args.topo_steps} steps")
        topo = topo_leg(args.topo_steps, args.warmup, stream)

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
    if args.cpu_steps > 0:
        log(f"cpu baseline: {args.cpu_steps} steps")
        cpu = cpu_baseline(m, args.cpu_steps)
        # the 100x100x60 configuration in full (501 executed steps), as the CPU baseline plan has it
        m1 = workloads.small()
        c1 = cpu_baseline(m1, m1["nt_total"])
        cpu["config_100x100x60_full"] = dict(value=c1["value"], unit=c1["unit"], sample=c1["sample"])
        log(f"cpu baseline {cpu['value']} Gcell-RK-stages/s ({cpu['cores']} threads); 100x100x60 in full "
            f"{c1['value']}")

    desc = {"explosion800": "explosion model at 800x800x400: Cartesian 100 m, homogeneous Vp 6000 / Vs 3464 / "
                            "rho 2670, explosive source 3 km deep, Ricker 2 Hz, 20 surface receivers",
            "small": "explosion model at 100x100x60: Cartesian 100 m, homogeneous, explosive source 3 km deep, "
                     "Ricker 2 Hz, 20 surface receivers",
            "topo": "curvilinear topography, 3 perturbed layers, double couple, 400 receivers"}[args.workload]
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gcell-RK-stages/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["wall"] * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.storage,
        "data": "synthetic",
        "config": {"workload": f"{m['name']}: {m['ni']}x{m['nj']}x{m['nk']} el_iso RK4, {desc}, free surface + "
                               f"10-layer CFS-PML, dt {m['dt']}",
                   "cells": cells, "storage": args.storage,
                   "kernel_path": "plain" if args.plain else KERNEL_PATH[r["split"]],
                   "seismograms_finite": r["finite"], "device_ms": round(r["device_ms"], 2)},
        "roofline": roof,
        "reduced_storage": precision,
        "topography": topo,
        "serving": serving,
        "batched_shots": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
