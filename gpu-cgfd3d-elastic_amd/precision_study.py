#!/usr/bin/env python3
"""Accuracy of reduced-precision storage on the explosion model.

Runs the 100 m Cartesian explosion model (homogeneous Vp 6000 / Vs 3464 /
rho 2670, free surface, 10-layer CFS-PML, Mxx=Myy=Mzz=1e16 N m Ricker 2 Hz at
3 km depth, 20 seeded surface receivers, dt 0.005 s) at a chosen size in one
storage type, and saves the seismograms and a strided sample of the final
wavefield.  With --ref, also prints the relative L2 error of both against
the reference files (e.g. the FP64 run).

  python precision_study.py --storage f64 --size 800,800,400 --steps 2000 --out f64.npz
  python precision_study.py --storage f16 --size 800,800,400 --steps 2000 --out f16.npz --ref f64.npz f32.npz
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import numpy as np  # noqa: E402

import cgfd  # noqa: E402
import workloads  # noqa: E402

STORE = {"f32": cgfd.STORE_F32, "f16": cgfd.STORE_F16, "bf16": cgfd.STORE_BF16, "f64": cgfd.STORE_F64}


def rel_l2(a, b):
    d = np.linalg.norm(np.asarray(b, np.float64))
    return float(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / d) if d > 0 else 0.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--storage", default="f32", choices=sorted(STORE))
    ap.add_argument("--size", default="800,800,400")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--stride", type=int, default=8, help="sampling stride of the final wavefield")
    ap.add_argument("--out", required=True)
    ap.add_argument("--ref", nargs="*", default=[])
    args = ap.parse_args()
    ni, nj, nk = (int(v) for v in args.size.split(","))
    t0 = time.time()
    m = workloads.small(nt=args.steps, ni=ni, nj=nj, nk=nk)
    print(f"[prec] model {ni}x{nj}x{nk} ready in {time.time() - t0:.0f} s", flush=True)
    st = STORE[args.storage]
    s = cgfd.Session(ni, nj, nk, storage=st)
    workloads.upload(s, m)
    ex = [0] * 9
    if st == cgfd.STORE_F16:
        ex = cgfd.storage_scales(st, ni, nj, nk, m["metric"], m["lam"], m["mu"], m["slw"], m["src"], m["dt"])
        s.set_storage_scale(ex)
    del m["metric"], m["x"], m["y"], m["z"]
    print(f"[prec] uploaded in {time.time() - t0:.0f} s; running {m['nt_total']} steps", flush=True)
    t1 = time.time()
    done = 0
    while done < m["nt_total"]:
        n = min(200, m["nt_total"] - done)
        s.run(done, n, m["dt"])
        done += n
        print(f"[prec] step {done}/{m['nt_total']} ({time.time() - t1:.0f} s)", flush=True)
    secs = time.time() - t1
    seis = s.read_seismo(0, dtype=np.float32)
    k = args.stride
    snap = s.fetch_subvolume(0, cgfd.LEVEL_PRE, list(range(9)), (0, 0, 0),
                             ((ni - 1) // k + 1, (nj - 1) // k + 1, (nk - 1) // k + 1), (k, k, k))
    s.close()
    np.savez(args.out, seismo=seis, snapshot=snap, storage=args.storage, scales=np.array(ex),
             seconds=secs, steps=m["nt_total"], size=np.array([ni, nj, nk]))
    res = dict(storage=args.storage, size=[ni, nj, nk], steps=m["nt_total"], run_seconds=round(secs, 1),
               gcell_rk_stages_per_s=round(ni * nj * nk * 4 * m["nt_total"] / secs / 1e9, 3), scales_exp2=ex,
               finite=bool(np.isfinite(seis).all() and np.isfinite(snap).all()))
    for r in args.ref:
        ref = np.load(r, allow_pickle=False)
        name = str(ref["storage"])
        res[f"seismo_rel_l2_vs_{name}"] = rel_l2(seis, ref["seismo"])
        res[f"snapshot_rel_l2_vs_{name}"] = rel_l2(snap, ref["snapshot"])
        # velocity only (what the receivers of the reference record)
        res[f"seismo_velocity_rel_l2_vs_{name}"] = rel_l2(seis[:, :3], ref["seismo"][:, :3])
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
