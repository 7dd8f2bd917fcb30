"""Mutation fuzz of the file readers through the C ABI: truncated, spliced and
perturbed copies of valid station lists, .src source files and md3lay /
md3grd media files must each give a result or a CgfdError, never a crash or
a hang.  (A longer run of the same mutations under AddressSanitizer and
UBSan found the out-of-range conversions and the read before a buffer that
the readers now reject.)  Seeded, so every run tries the same inputs."""
import os
import random

import numpy as np
import pytest

import cgfd

HERE = os.path.dirname(os.path.abspath(__file__))
NI, NJ, NK = 20, 18, 16
TOKENS = [" ", "\n", "#", "-1", "0", "1", "2", "3", "1e30", "nan", "inf", "ricker", "gaussian", "klauder",
          "999999", "-999999", "0.5", "abc", "\t", "2147483647", "20001e30", "1e-30"]


def _grid():
    return cgfd.grid_cartesian(NI, NJ, NK, 100.0, 0.0, 100.0, 0.0, 100.0, -(NK - 1) * 100.0)


def _mutate(rng, s):
    for _ in range(1 + rng.randrange(2)):
        if not s:
            break
        pos = rng.randrange(len(s) + 1)
        op = rng.randrange(4)
        if op == 0:
            s = s[:pos] + s[pos + 1 + rng.randrange(8):]
        elif op == 1:
            s = s[:pos] + rng.choice(TOKENS) + s[pos:]
        elif op == 2 and pos < len(s):
            s = s[:pos] + rng.choice(" \n0123456789.-e#a") + s[pos + 1:]
        else:
            s = s[:pos]
    return s


def _seeds_text():
    ts = np.arange(11) * 0.02
    rows = "\n".join(f"{1 + 2 * t:.6g} {-5 * t ** 3 + t:.6g} 0.5" for t in ts)
    return {
        "src": [open(os.path.join(HERE, "golden", "test_source.src")).read(),
                "# comment\nev\n3\n0 1.0\n2 0\n0 1\n10.2 11.7 5.4\n55 3 3\n12 13 2.0\n0.05 gaussian 0.2 0.3\n"
                "1 2 3 4 5 6\n0.0 ricker 9 9\n7 7 7 7 7 7\n0.10 ricker 2.5\n10 20 30 40\n",
                "ev\n2\n0 0.3\n2 1\n0 0\n6 7 8\n9 9 9\n0 gaussian 0.05 0.1\n30 60 90 3.2e10 1.5 2.0e6\n"
                "0 gaussian 0.05 0.1\n45 30 -60 -1 0.5 1.0e6\n",
                f"ev\n2\n1 0.02 11\n3 0\n0 0\n4 4 4\n5 5 5\n0.0\n{rows}\n0.01\n{rows}\n",
                "ev\n2\n0 0.2\n1 0\n1 1\n1234 876 420\n500 600 100\n0 ricker 3 0.1\n1 1 1\n0 klauder 3 0.1\n1 1 1\n"],
        "list": [open(os.path.join(HERE, "golden", "station.list")).read(),
                 "# a comment\n\n4\na 1 1 1234.0 1050.0 350.0\nb 1 0 777.0 888.0 -1020.0\nc 0 0 5.5 7.25 3.0\n"
                 "far 0 1 500.0 2.0 0.0\n"],
    }


def _media_seeds():
    lay = ["elastic_isotropic", "2", "3 3 -1000 -1000 2000 2000"]
    for z, r, vp, vs in ((100, 2000, 3000, 1700), (-800, 2500, 5000, 2900)):
        lay += [f"{z} {r} 0 0 {vp} 0 0 {vs} 0 0"] * 9
    grd = ["elastic_isotropic", "2", "2 2", "3 3 -1000 -1000 2000 2000"]
    for v in ((100, 2000, 3000, 1700), (-500, 2100, 3200, 1800), (-500, 2500, 5000, 2900), (-2000, 2600, 5200, 3000)):
        grd += [" ".join(str(a) for a in v)] * 9
    return "\n".join(lay) + "\n", "\n".join(grd) + "\n"


@pytest.mark.timeout(300)
def test_source_and_station_readers_survive_mutations(tmp_path):
    x, y, z = _grid()
    mu = np.full((NK + 6, NJ + 6, NI + 6), 3.0e10)
    rng = random.Random(20261203)
    seeds = _seeds_text()
    p = tmp_path / "m.txt"
    outcomes = {"ok": 0, "error": 0}
    for it in range(1500):
        p.write_text(_mutate(rng, rng.choice(seeds["src" if it % 2 == 0 else "list"])))
        try:
            if it % 2 == 0:
                info, src, _, _ = cgfd.source_file(p, NI, NJ, NK, 0.0, 0.01, x, y, z, mu=mu,
                                                   max_table_bytes=1 << 26)
                assert info["nsrc"] == src["nsrc"] >= 0
            else:
                cgfd.station_file(p, NI, NJ, NK, x, y, z, max_n=1000)
            outcomes["ok"] += 1
        except cgfd.CgfdError:
            outcomes["error"] += 1
    # both outcomes occur: the mutations reach past the first line
    assert outcomes["ok"] > 20 and outcomes["error"] > 20, outcomes


@pytest.mark.timeout(300)
def test_media_readers_survive_mutations(tmp_path):
    x, y, z = _grid()
    rng = random.Random(11)
    lay, grd = _media_seeds()
    p = tmp_path / "m.txt"
    outcomes = {"ok": 0, "error": 0}
    for it in range(400):
        p.write_text(_mutate(rng, grd if it % 2 else lay))
        method = rng.choice(["loc", "har", "ari"])
        try:
            if it % 2:
                lam, mu, rho = cgfd.media_grid_file(NI, NJ, NK, x, y, z, p, method)
            else:
                lam, mu, rho = cgfd.media_layer_file(NI, NJ, NK, x, y, z, p, method)
            assert lam.shape == mu.shape == rho.shape
            outcomes["ok"] += 1
        except cgfd.CgfdError:
            outcomes["error"] += 1
    assert outcomes["ok"] > 10 and outcomes["error"] > 10, outcomes
