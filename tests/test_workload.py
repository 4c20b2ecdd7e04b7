"""Seeded synthetic fields: determinism and the config recipes."""
import numpy as np

from salsa_amd import capi, workload as W


def test_uniform_is_a_pure_function_of_seed_stream_index():
    a = W.uniform(3, 7, np.arange(1000, dtype=np.uint64))
    b = W.uniform(3, 7, np.arange(1000, dtype=np.uint64))
    c = W.uniform(4, 7, np.arange(1000, dtype=np.uint64))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert 0.0 <= a.min() and a.max() < 1.0 and abs(a.mean() - 0.5) < 0.05


def test_config1_fields():
    d = capi.bin_desc()
    f = W.config_fields(d, 1)
    g = f.grid
    assert (g.nzp, g.nxp, g.nyp, g.ncells) == (42, 12, 12, 2560)
    idx = g.interior_index()
    T = f.temp[idx]
    assert 275. <= T.min() and T.max() <= 295.
    rh = f.rv[idx] / f.rsl[idx]
    assert 0.95 <= rh.min() and rh.max() <= 1.01
    assert 0.25 < f.cloudy.mean() < 0.35
    assert f.n[0].shape == (11, g.ncell3d) and f.m[0].shape == (22, g.ncell3d)
    assert np.all(f.n[2] == 0.0)                       # no precipitation initially
    f2 = W.config_fields(d, 1)
    assert np.array_equal(f.m[1], f2.m[1])


def test_salsa_units_pack_matches_driver_arithmetic():
    d = capi.bin_desc()
    f = W.config_fields(d, 1)
    T, p, rs, rsi, rv, numc, volc = W.to_salsa_units(d, f)
    idx = f.grid.interior_index()
    b = 5
    assert np.array_equal(numc[:, b], f.n[0][b, idx] * f.dn[idx])
    assert np.array_equal(volc[:, b, 0], f.m[0][b, idx] * f.dn[idx] / 1830.)
