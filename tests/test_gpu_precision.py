"""GPU: reduced-precision storage (FP16 / BF16 wavefield and PML memory with
power-of-two scales, FP32 arithmetic) against the FP64 path of the same scheme.

The FP32 path must agree with FP64 to FP32 rounding; the 16-bit formats are
held to the bounds below (relative L2 over all seismogram components), and
the measured errors are printed for the record.
"""
import numpy as np
import pytest

import cgfd
from problem import Problem, rel_l2

pytestmark = pytest.mark.gpu

BOUND = {cgfd.STORE_F32: 1e-4, cgfd.STORE_F16: 3e-2, cgfd.STORE_BF16: 2e-1}
NAME = {cgfd.STORE_F32: "f32", cgfd.STORE_F16: "f16", cgfd.STORE_BF16: "bf16"}


@pytest.fixture(scope="module")
def fp64_run():
    P = Problem(ni=64, nj=60, nk=48, curvilinear=True, hetero=0.03, nt_total=320, seed=31,
                pml=(8, 8, 8, 8, 8, 0), f0=3.0, t0=0.3, nrecv=12)
    s = cgfd.Session(P.ni, P.nj, P.nk, storage=cgfd.STORE_F64)
    P.upload(s, dtype=np.float64)
    s.run(0, P.nt_total, P.dt)
    ref = s.read_seismo(0)
    w = s.fetch_level(0, cgfd.LEVEL_PRE)
    s.close()
    return P, ref, w


@pytest.mark.parametrize("storage", [cgfd.STORE_F32, cgfd.STORE_F16, cgfd.STORE_BF16], ids=["f32", "f16", "bf16"])
def test_storage_error_vs_fp64(fp64_run, storage):
    P, ref, wref = fp64_run
    s = cgfd.Session(P.ni, P.nj, P.nk, storage=storage)
    P.upload(s, dtype=np.float32)
    ex = cgfd.storage_scales(storage, P.ni, P.nj, P.nk, P.metric, P.lam, P.mu, P.slw, P.src, P.dt)
    if storage in (cgfd.STORE_F16, cgfd.STORE_BF16):
        s.set_storage_scale(ex)
    s.run(0, P.nt_total, P.dt)
    seis = s.read_seismo(0)
    w = s.fetch_level(0, cgfd.LEVEL_PRE)
    s.close()
    assert np.isfinite(seis).all() and np.isfinite(w).all()
    e_seis = rel_l2(seis, ref)
    e_snap = rel_l2(w, wref)
    print(f"\n[precision] {NAME[storage]} scales {ex}: seismogram rel L2 {e_seis:.3e}, "
          f"final wavefield rel L2 {e_snap:.3e}")
    assert e_seis <= BOUND[storage]
    assert e_snap <= BOUND[storage]


@pytest.mark.parametrize("curv", [True, False], ids=["curv", "cart"])
@pytest.mark.parametrize("storage", [cgfd.STORE_F16, cgfd.STORE_BF16], ids=["f16", "bf16"])
def test_two_column_sweep_matches_one_column(storage, curv):
    """The 16-bit interior sweep that handles two columns per lane (tile shapes
    7 and 8) computes what the one-column sweep (shape 2) computes: same FP32
    arithmetic, so the two differ only where a contraction or summation order
    flips a 16-bit rounding.  An odd ni and x-PML widths 3 and 4 put box edges
    at both column parities."""
    P = Problem(ni=75, nj=41, nk=37, curvilinear=curv, pml=(3, 4, 5, 2, 4, 0), nt_total=48, seed=12,
                hetero=0.04, f0=4.0, t0=0.25)
    ex = cgfd.storage_scales(storage, P.ni, P.nj, P.nk, P.metric, P.lam, P.mu, P.slw, P.src, P.dt)
    res = {}
    for shape, zc in ((2, 16), (7, 11), (8, 24)):
        # the two-column sweep runs inside the PML slabs (per-point shell split)
        s = cgfd.Session(P.ni, P.nj, P.nk, storage=storage, flags=cgfd.FLAG_PML_POINT)
        P.upload(s, dtype=np.float32)
        s.set_storage_scale(ex)
        s.set_tile(shape, zc)
        s.run(0, P.nt_total, P.dt)
        assert s.get_tile()["shape"] == shape
        res[shape] = (s.read_seismo(0), s.fetch_level(0, cgfd.LEVEL_PRE))
        s.close()
    tol = 2e-3 if storage == cgfd.STORE_F16 else 2e-2
    for shape in (7, 8):
        e_seis = rel_l2(res[shape][0], res[2][0])
        e_w = rel_l2(res[shape][1], res[2][1])
        print(f"\n[two-column] {NAME[storage]} shape {shape}: seismogram {e_seis:.2e}, wavefield {e_w:.2e}")
        assert np.isfinite(res[shape][1]).all()
        assert e_seis <= tol and e_w <= tol
