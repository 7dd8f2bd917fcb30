"""Power-of-two storage scales: the physical bound holds on an FP64 oracle run."""
import numpy as np
import pytest

import cgfd
from problem import Problem


@pytest.mark.parametrize("source", ["moment", "force"])
def test_scaled_fp16_values_stay_in_range(source):
    P = Problem(ni=40, nj=36, nk=32, nt_total=120, seed=3, source=source, pml=(5, 5, 5, 5, 5, 0), t0=0.15, f0=4.0)
    ex = cgfd.storage_scales(cgfd.STORE_F16, P.ni, P.nj, P.nk, P.metric, P.lam, P.mu, P.slw, P.src, P.dt)
    seis, wave, _ = P.oracle_run(64)
    peak = np.maximum(np.abs(wave).reshape(9, -1).max(1), np.abs(seis).max(axis=(0, 2)))
    stored = peak * np.exp2(-np.asarray(ex, dtype=np.float64))
    # stays below the 2^12 target (16x under the FP16 maximum) and far above
    # the FP16 normal minimum 2^-14
    assert np.all(stored <= 4096.0)
    assert np.all(stored >= 2.0 ** -6)
    # velocity components share one exponent, stress components another
    assert len(set(ex[:3])) == 1 and len(set(ex[3:])) == 1


def test_other_storages_keep_unit_scale():
    P = Problem(ni=20, nj=18, nk=16, nt_total=10, seed=1)
    for st in (cgfd.STORE_F32, cgfd.STORE_BF16, cgfd.STORE_F64):
        assert cgfd.storage_scales(st, P.ni, P.nj, P.nk, P.metric, P.lam, P.mu, P.slw, P.src, P.dt) == [0] * 9
