"""GPU parity: the HIP el_iso path against the CPU oracle on identical inputs.

Tolerances: FP32 relative L2 error <= 1e-5 per component (the two differ only
in FMA contraction and summation order); FP64 <= 1e-11.
"""
import numpy as np
import pytest

import cgfd
from problem import Problem, rel_l2

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-11


def _sess(P, storage=cgfd.STORE_F32, flags=0, nshots=1):
    s = cgfd.Session(P.ni, P.nj, P.nk, storage=storage, nshots=nshots, flags=flags)
    P.upload(s, dtype=np.float64 if storage == cgfd.STORE_F64 else np.float32)
    return s


def _cmp_levels(a, b, tol, what):
    for c in range(9):
        e = rel_l2(a[c], b[c])
        assert e <= tol, f"{what}: component {c} rel L2 {e:.3e} > {tol}"


def _aux_concat(sess, P, level):
    sizes = P.face_sizes()
    parts = []
    for f in range(6):
        if sizes[f] == 0:
            continue
        parts.append(sess.fetch_pml_aux(0, f, level, int(sizes[f])).ravel())
    return np.concatenate(parts) if parts else np.zeros(0)


@pytest.mark.parametrize("flags", [cgfd.FLAG_PLAIN, 0, cgfd.FLAG_PML_POINT], ids=["plain", "zmarch", "pmlpoint"])
def test_onestage_rhs_all_pairs(flags):
    """Every (pair, stage) RHS incl. traction image, vlow rows, PML and source."""
    P = Problem(ni=40, nj=34, nk=30, curvilinear=True, pml=(4, 4, 4, 4, 4, 0), hetero=0.05,
                init_random=True, nt_total=6, seed=3)
    s = _sess(P, flags=flags)
    s.run(0, 3, P.dt)  # non-zero PML memory
    w = s.fetch_level(0, cgfd.LEVEL_PRE)
    aux = _aux_concat(s, P, cgfd.LEVEL_PRE)
    assert np.abs(aux).max() > 0
    sizes = P.face_sizes()
    for ipair in range(8):
        for istage in range(4):
            it = 3 + ipair
            s.onestage(0, cgfd.LEVEL_PRE, ipair, istage, it)
            rg = s.fetch_level(0, cgfd.LEVEL_RHS)
            ro, ao = P.oracle_stage(w, ipair, istage, it, aux=aux)
            _cmp_levels(rg, ro, TOL32, f"rhs pair {ipair} stage {istage}")
            ag = _aux_concat(s, P, cgfd.LEVEL_RHS)
            off = 0
            for f in range(6):
                n = int(sizes[f]) * 9
                if n == 0:
                    continue
                e = rel_l2(ag[off:off + n], ao[off:off + n])
                assert e <= TOL32, f"aux rhs face {f} pair {ipair} stage {istage}: {e:.3e}"
                off += n
    s.close()


@pytest.mark.parametrize("flags", [cgfd.FLAG_PLAIN, 0, cgfd.FLAG_PML_POINT], ids=["plain", "zmarch", "pmlpoint"])
@pytest.mark.parametrize("curv", [False, True], ids=["cart", "curv"])
def test_run_matches_oracle_fp32(flags, curv):
    P = Problem(ni=40, nj=36, nk=32, curvilinear=curv, pml=(5, 5, 5, 5, 5, 0), nt_total=40, seed=1,
                hetero=0.03 if curv else 0.0)
    s = _sess(P, flags=flags)
    s.run(0, P.nt_total, P.dt)
    sg = s.read_seismo(0)
    wg = s.fetch_level(0, cgfd.LEVEL_PRE)
    so, wo, _ = P.oracle_run(32)
    assert np.abs(so).max() > 0
    e = rel_l2(sg, so)
    assert e <= TOL32, f"seismograms rel L2 {e:.3e}"
    _cmp_levels(wg, wo, TOL32, "final wavefield")
    s.close()


def test_run_matches_oracle_fp64():
    P = Problem(ni=36, nj=32, nk=28, curvilinear=True, pml=(4, 4, 4, 4, 4, 0), nt_total=30, seed=2, hetero=0.02)
    s = _sess(P, storage=cgfd.STORE_F64)
    s.run(0, P.nt_total, P.dt)
    sg = s.read_seismo(0)
    wg = s.fetch_level(0, cgfd.LEVEL_PRE)
    so, wo, _ = P.oracle_run(64)
    assert rel_l2(sg, so) <= TOL64
    _cmp_levels(wg, wo, TOL64, "final wavefield fp64")
    s.close()


def test_force_source_and_no_free_surface():
    P = Problem(ni=32, nj=30, nk=28, free_top=False, pml=(4, 4, 4, 4, 4, 4), source="force", nt_total=30, seed=5)
    s = _sess(P)
    s.run(0, P.nt_total, P.dt)
    so, wo, _ = P.oracle_run(32)
    assert rel_l2(s.read_seismo(0), so) <= TOL32
    _cmp_levels(s.fetch_level(0, cgfd.LEVEL_PRE), wo, TOL32, "wavefield")
    s.close()


def test_chunked_run_equals_single_run():
    """cgfd_run(it0, n) in pieces gives the same result as one call (pairs follow it % 8)."""
    P = Problem(ni=32, nj=28, nk=26, nt_total=21, seed=4)
    a = _sess(P)
    a.run(0, 21, P.dt)
    b = _sess(P)
    b.run(0, 5, P.dt)
    b.run(5, 9, P.dt)
    b.run(14, 7, P.dt)
    assert np.array_equal(a.read_seismo(0), b.read_seismo(0))
    assert np.array_equal(a.fetch_level(0, cgfd.LEVEL_PRE), b.fetch_level(0, cgfd.LEVEL_PRE))
    a.close()
    b.close()


def test_multishot_equals_single_shot():
    P = Problem(ni=36, nj=30, nk=28, nt_total=16, seed=6, curvilinear=True)
    one = _sess(P)
    one.run(0, P.nt_total, P.dt)
    many = _sess(P, nshots=3)
    many.run(0, P.nt_total, P.dt)
    ref = one.read_seismo(0)
    for shot in range(3):
        assert np.array_equal(many.read_seismo(shot), ref)
        assert np.array_equal(many.fetch_level(shot, cgfd.LEVEL_PRE), one.fetch_level(0, cgfd.LEVEL_PRE))
    one.close()
    many.close()


@pytest.mark.parametrize("split", ["planes", "shell", "core"])
@pytest.mark.parametrize("shape,zchunk", [(0, 7), (1, 48), (2, 5), (3, 24), (0, 96), (1, 33)])
def test_every_interior_tile_shape(shape, zchunk, split):
    """Each z-marching tile shape and a ragged z chunk reproduce the oracle,
    under every work split (a pinned tile runs the whole-plane split unless
    the per-point shell or the core split is asked for)."""
    P = Problem(ni=75, nj=41, nk=37, curvilinear=True, pml=(4, 3, 5, 2, 4, 0), nt_total=12, seed=12, hetero=0.04)
    s = _sess(P, flags=cgfd.FLAG_PML_POINT if split == "shell" else 0)
    if split == "core":
        s.set_split(cgfd.SPLIT_CORE)
    s.set_tile(shape, zchunk)
    s.run(0, P.nt_total, P.dt)
    t = s.get_tile()
    assert t["shape"] == shape and t["zchunk"] == zchunk
    assert s.get_split() == {"planes": cgfd.SPLIT_PLANES, "shell": cgfd.SPLIT_SHELL, "core": cgfd.SPLIT_CORE}[split]
    so, wo, _ = P.oracle_run(32)
    assert rel_l2(s.read_seismo(0), so) <= TOL32
    _cmp_levels(s.fetch_level(0, cgfd.LEVEL_PRE), wo, TOL32, f"shape {shape}")
    s.close()


@pytest.mark.parametrize("storage", ["f32", "f64"])
@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_core_split_wide_grid(shape, storage):
    """The core split on a grid wide enough for a core under every tile
    shape (64-wide tiles too), with a free surface and unequal faces: the
    PML-free core and the PML-form rim boxes around it reproduce the oracle
    (FP32 and the FP64 path) and the whole-plane split to rounding."""
    P = Problem(ni=150, nj=44, nk=30, pml=(6, 9, 5, 7, 3, 0), nt_total=40, seed=15, hetero=0.03,
                src_ijk=(40.3, 20.6, 14.2))
    st, prec, tol = (cgfd.STORE_F32, 32, TOL32) if storage == "f32" else (cgfd.STORE_F64, 64, TOL64)
    runs = {}
    for split in (cgfd.SPLIT_CORE, cgfd.SPLIT_PLANES):
        s = _sess(P, storage=st)
        s.set_split(split)
        s.set_tile(shape, 9)
        s.run(0, P.nt_total, P.dt)
        assert s.get_split() == split
        runs[split] = (s.read_seismo(0), s.fetch_level(0, cgfd.LEVEL_PRE))
        s.close()
    so, wo, _ = P.oracle_run(prec)
    errs = {split: rel_l2(seis, so) for split, (seis, _) in runs.items()}
    for split, (seis, w) in runs.items():
        assert errs[split] <= tol, errs
        _cmp_levels(w, wo, tol, f"split {split}")
    _cmp_levels(runs[cgfd.SPLIT_CORE][1], runs[cgfd.SPLIT_PLANES][1], tol / 10, "core vs planes")


def test_pinned_split_survives_the_tile_cache():
    """A split pinned with cgfd_set_split is kept (and its tile tuned under
    it) even when an unpinned session of the same model filled the process's
    tile cache first; every pinned split reproduces the oracle."""
    P = Problem(ni=96, nj=40, nk=30, pml=(5, 5, 5, 5, 5, 0), nt_total=12, seed=16, hetero=0.02)
    a = _sess(P)
    a.run(0, 2, P.dt)
    a.close()
    so, wo, _ = P.oracle_run(32)
    for split in (cgfd.SPLIT_CORE, cgfd.SPLIT_PLANES, cgfd.SPLIT_SHELL):
        s = _sess(P)
        s.set_split(split)
        s.run(0, P.nt_total, P.dt)
        assert s.get_split() == split
        assert rel_l2(s.read_seismo(0), so) <= TOL32
        _cmp_levels(s.fetch_level(0, cgfd.LEVEL_PRE), wo, TOL32, f"pinned split {split}")
        s.set_split(-1)
        s.close()


def test_autotune_picks_a_valid_tile():
    P = Problem(ni=48, nj=40, nk=36, nt_total=4, seed=13, pml=(5, 5, 5, 5, 5, 0))
    s = _sess(P)
    s.run(0, 2, P.dt)
    t = s.get_tile()
    assert 0 <= t["shape"] < 4 and t["zchunk"] >= 1
    assert s.get_split() in (cgfd.SPLIT_SHELL, cgfd.SPLIT_PLANES, cgfd.SPLIT_CORE)
    s.close()


def test_autotuned_split_matches_oracle():
    """Whatever split and tile the autotuner keeps, the run reproduces the oracle."""
    P = Problem(ni=64, nj=50, nk=40, curvilinear=True, pml=(6, 6, 6, 6, 6, 0), nt_total=10, seed=14, hetero=0.03)
    s = _sess(P)
    s.run(0, P.nt_total, P.dt)
    so, wo, _ = P.oracle_run(32)
    assert rel_l2(s.read_seismo(0), so) <= TOL32
    _cmp_levels(s.fetch_level(0, cgfd.LEVEL_PRE), wo, TOL32, f"autotuned split {s.get_split()}")
    s.close()
