"""The oracle's level-4 sedimentation (so_sedim: sedim_SALSA,
DepositionSlow, DepositionFast, terminal_vel; mcrp.f90:745-1275,
mo_particle_external_properties.f90:43-100), pinned by:

* fall speeds of water drops against the Gunn & Kinzer (1949) measurements
  (20 C, 1013 hPa) within 8 %, the Stokes law at 10 um, and continuity of
  the three regimes at 19 um and 1.07 mm;
* the precipitation scheme conserving sum_k N/dzt and sum_k M/dzt per column
  on a uniform grid (particles only move between levels, the level below
  the surface included; on a stretched grid the reference's split between
  two levels is not conservative, as written);
* the slow scheme's flux form: per column sum_k dM_w/dt adn/dzt equals the
  surface water flux that the surface rain rate reports;
* the fall distance on a uniform grid worked by hand.
Runs the FP64 oracle only."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, les_case

ALVL = 2.5e6


def sed_fields(desc, grid, seed):
    shape, dzt, f, n_p, n_t, m_p, m_t = les_case(desc, 4, grid, seed)
    from salsa_amd import workload as W
    rng = np.random.default_rng(seed + 7)
    n3 = int(np.prod(shape))
    for a in n_p + m_p:
        np.abs(a, out=a)
    # the state SALSA_diagnostics leaves: no particles without mass
    for c in range(3):
        nb = n_p[c].shape[0]
        tot = sum(m_p[c][s * nb:(s + 1) * nb] for s in range(desc.iwa + 1))
        n_p[c][tot <= 1e-25 * np.maximum(n_p[c], 1.)] = 0.
    tk = f["temp"]
    fields = dict(tk=tk, th=tk * 1.02, adn=np.full(n3, 1.1) + 0.1 * rng.random(n3),
                  ustar=0.1 + 0.4 * rng.random(shape[1] * shape[2]))
    n_t = [rng.normal(size=a.shape) * 1e-3 * a for a in n_p]
    m_t = [rng.normal(size=a.shape) * 1e-3 * a for a in m_p]
    return shape, dzt, fields, n_p, m_p, n_t, m_t


def run(O, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, tlt, sed=(1, 1, 1), dtlt=1.0):
    L = O.make_layout(species=(0,))
    return O.sedim(L, shape, dtlt, dzt, fld["tk"], fld["th"], fld["adn"], fld["ustar"], n_p, m_p, n_t, m_t, tlt,
                   sed)


def test_fall_speed_against_measurements(oracle_mod):
    O = oracle_mod
    mu, rho_a = 1.81e-5, 1.204                    # air at 20 C, 1013 hPa
    for d, v in ((0.1e-3, 0.27), (0.2e-3, 0.72), (0.5e-3, 2.06), (1.0e-3, 4.03), (2.0e-3, 6.49), (4.0e-3, 8.83)):
        assert O.terminal_vel(d, 998.2, rho_a, mu, 1.0) == pytest.approx(v, rel=0.08), d
    d = 10e-6
    assert O.terminal_vel(d, 998.2, rho_a, mu, 1.0) == pytest.approx(d * d * (998.2 - rho_a) * 9.81 / (18 * mu),
                                                                      rel=1e-15)
    for edge in (19e-6, 1.07e-3):
        lo, hi = O.terminal_vel(edge * (1 - 1e-9), 998.2, rho_a, mu, 1.0), O.terminal_vel(edge * (1 + 1e-9), 998.2,
                                                                                          rho_a, mu, 1.0)
        assert abs(hi - lo) / lo < 0.12, edge


def test_precipitation_conserved_per_column(oracle_mod):
    from salsa_amd import capi
    desc = capi.bin_desc()
    shape, dzt, fld, n_p, m_p, n_t, m_t = sed_fields(desc, (20, 4, 4), 5)
    # a uniform grid: where the drops land between two levels, the reference
    # moves the lower share without the dzt(kf)/dzt(k) factor (mcrp.f90:1246),
    # which conserves sum_k N/dzt only when the levels are equally thick
    dzt = np.full_like(dzt, 1. / 25.)
    n_t = [np.zeros_like(a) for a in n_p]
    m_t = [np.zeros_like(a) for a in m_p]
    tlt = np.zeros(int(np.prod(shape)))
    rr, sr, vt = run(oracle_mod, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, tlt, sed=(0, 0, 1))
    nzp, nxp, nyp = shape
    assert vt.max() > 0.5, "no fast drops: the case does not exercise the fall distance"
    for a, src in ((n_t[2], n_p[2]), (m_t[2], m_p[2])):
        s = a.reshape(a.shape[0], nyp, nxp, nzp)[:, 2:nyp - 2, 2:nxp - 2, :]
        s0 = src.reshape(s.shape[0], nyp, nxp, nzp)[:, 2:nyp - 2, 2:nxp - 2, :]
        col = (s / dzt[None, None, None, :]).sum(axis=-1)
        # rounding: drops that stay mostly in their level leave N - (1 - f) N
        scale = (np.abs(s0) / dzt[None, None, None, :]).sum(axis=-1) + 1e-300
        assert np.all(np.abs(col) <= 1e-14 * scale)
        assert (np.abs(s) > 0).sum() > 0


def test_slow_flux_matches_surface_rate(oracle_mod):
    from salsa_amd import capi
    desc = capi.bin_desc()
    shape, dzt, fld, n_p, m_p, n_t, m_t = sed_fields(desc, (12, 4, 4), 6)
    nzp, nxp, nyp = shape
    nw = desc.iwa
    for c, nb in ((0, desc.nbins), (1, desc.ncld)):
        n_p[c][nb - 1] = 0.                       # the rate sums start one entry early (as written): keep it empty
    for c in (0, 1):
        n_t[c][:] = 0.
        m_t[c][:] = 0.
    tlt = np.zeros(int(np.prod(shape)))
    rr, sr, vt = run(oracle_mod, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, tlt, sed=(1, 1, 0))
    adn = fld["adn"].reshape(nyp, nxp, nzp)
    tot = 0.
    for c, nb in ((0, desc.nbins), (1, desc.ncld)):
        w = m_t[c][nw * nb:(nw + 1) * nb].reshape(nb, nyp, nxp, nzp).sum(axis=0)
        tot = tot + (w * adn / dzt[None, None, :]).sum(axis=-1)
    surf = sr.reshape(nyp, nxp)
    inner = (slice(2, nyp - 2), slice(2, nxp - 2))
    # sum_k dM/dt adn / dzt = rfl(2) adn(2) = -(surface rate) / alvl; the sum
    # telescopes level fluxes up to 1e4 times the surface term, hence 1e-9
    np.testing.assert_allclose(tot[inner], -surf[inner] / ALVL, rtol=1e-9, atol=1e-30)
    assert np.all(surf[inner] > 0)


def test_fall_distance_uniform_grid(oracle_mod):
    """One precipitation bin at level k of a uniform grid (thickness D): the
    drops fall fdmax = vc dt; they land fi = floor(fdmax/D) levels down, and
    the overshoot fraction f = fdmax/D - fi goes one level further (all of
    it when f > 0.5)."""
    from salsa_amd import capi
    desc = capi.bin_desc()
    nzp = 30
    shape = (nzp, 5, 5)
    n3 = nzp * 25
    D = 20.
    dzt = np.full(nzp, 1. / D)
    fld = dict(tk=np.full(n3, 285.), th=np.full(n3, 290.), adn=np.full(n3, 1.15), ustar=np.full(25, 0.3))
    for dtlt in (3.0, 7.0, 11.0):
        n_p = [np.zeros((desc.nbins, n3)), np.zeros((desc.ncld, n3)), np.zeros((desc.nprc, n3))]
        m_p = [np.zeros((desc.nbins * 2, n3)), np.zeros((desc.ncld * 2, n3)), np.zeros((desc.nprc * 2, n3))]
        n_t = [np.zeros_like(a) for a in n_p]
        m_t = [np.zeros_like(a) for a in m_p]
        k = 25
        x = k + nzp * (2 + 5 * 2)
        b = 3
        N, d = 100., 1.0e-3
        n_p[2][b, x] = N
        m_p[2][1 * desc.nprc + b, x] = N * 1000. * 0.5235988 * d ** 3
        rr, sr, vt = run(oracle_mod, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, np.zeros(n3), sed=(0, 0, 1),
                         dtlt=dtlt)
        vc = vt[b, x]
        fdmax = vc * dtlt
        fi = int(np.floor(fdmax / D))
        f = fdmax / D - fi
        got = n_t[2][b, x - k:x - k + nzp] * dtlt
        expect = np.zeros(nzp)
        expect[k] -= N
        if f > 0.5:
            expect[k - fi - 1] += N
        else:
            expect[k - fi - 1] += f * N
            expect[k - fi] += (1 - f) * N
        np.testing.assert_allclose(got, expect, rtol=1e-12, atol=1e-9 * N)


# ---- level 5: ice (DepositionFast flag 4, mcrp.f90:880-905, 1073-1275) ----

ALVI = 2.834e6


def sed_fields5(desc, grid, seed):
    """Level-5 LES fields (SO4 + dust, ice with rime) in the state
    SALSA_diagnostics leaves, with the fields sedimentation reads."""
    shape, dzt, f, n_p, n_t, m_p, m_t = les_case(desc, 5, grid, seed)
    rng = np.random.default_rng(seed + 7)
    n3 = int(np.prod(shape))
    for a in n_p + m_p:
        np.abs(a, out=a)
    for c in range(4):
        nb = n_p[c].shape[0]
        nsw = desc.iwa + (2 if c == 3 else 1)
        tot = sum(m_p[c][s * nb:(s + 1) * nb] for s in range(nsw))
        n_p[c][tot <= 1e-25 * np.maximum(n_p[c], 1.)] = 0.
    tk = f["temp"]
    fields = dict(tk=tk, th=tk * 1.02, adn=np.full(n3, 1.1) + 0.1 * rng.random(n3),
                  ustar=0.1 + 0.4 * rng.random(shape[1] * shape[2]))
    n_t = [rng.normal(size=a.shape) * 1e-3 * a for a in n_p]
    m_t = [rng.normal(size=a.shape) * 1e-3 * a for a in m_p]
    return shape, dzt, fields, n_p, m_p, n_t, m_t


def run5(O, shape, dzt, fld, n_p, m_p, n_t, m_t, tlt, sed=(1, 1, 1, 1), dtlt=1.0):
    L = O.make_layout(species=(0, 3))
    P = O.make_params(level=5)
    return O.sedim5(L, P, shape, dtlt, dzt, fld["tk"], fld["th"], fld["adn"], fld["ustar"], n_p, m_p, n_t, m_t, tlt,
                    sed)


def test_ice_fall_speed_chain_against_reference(oracle_mod):
    """The per-bin fall speed of ice against the compiled reference (calcDiamLES
    both ways, getDiameter, getShapeCoefficients, terminal_vel flag 4;
    tests/golden/ice_fall_cfg5.npz): the diameters, shape coefficients,
    density and Cunningham factor to the last bit, the fall speed to 1e-12
    (pow and exp of the two runtimes).  The cases span the four mass-diameter
    regimes of mo_ice_shape.f90."""
    O = oracle_mod
    d = np.load(os.path.join(GOLDEN, "ice_fall_cfg5.npz"))
    L = O.make_layout(species=(0, 3))
    P = O.make_params(level=5)
    keys = ("dwet", "dnsp", "alpha", "beta", "gamma", "sigma", "zdn", "gg", "vc")
    n, ni = d["numc"].shape
    got = np.array([[O.sed_ice_vc(L, P, d["T"][c], d["adn"][c], d["numc"][c, b], d["mass"][c, b]) for b in range(ni)]
                    for c in range(n)])
    for q, k in enumerate(keys):
        rtol = 1e-12 if k == "vc" else 0.
        np.testing.assert_allclose(got[:, :, q], d[k], rtol=rtol, atol=0, err_msg=k)
    # regimes: spherical (small or graupel), dense non-spherical, partially rimed
    sph = np.isclose(d["alpha"], 0.5235988 * 917.)
    dense = d["alpha"] == P.ice_alpha
    rimed = ~sph & ~dense
    mtot = d["mass"].sum(axis=-1) / d["numc"]
    assert sph.sum() > 20 and dense.sum() > 5 and rimed.sum() > 20
    assert mtot[sph].min() < 1e-12 < mtot[sph].max()          # below Mth and in the graupel band


def test_ice_conserved_per_column(oracle_mod):
    """Ice moves only between levels (the level below the surface included):
    on a uniform grid sum_k N/dzt and sum_k M_s/dzt of every species, rime
    included, are conserved per column."""
    from salsa_amd import capi
    desc = capi.bin_desc(species=("SO4", "DU"))
    shape, dzt, fld, n_p, m_p, n_t, m_t = sed_fields5(desc, (20, 4, 4), 5)
    dzt = np.full_like(dzt, 1. / 25.)
    n_t = [np.zeros_like(a) for a in n_p]
    m_t = [np.zeros_like(a) for a in m_p]
    tlt = np.zeros(int(np.prod(shape)))
    rr, sr, ir, si, vt, vi = run5(oracle_mod, shape, dzt, fld, n_p, m_p, n_t, m_t, tlt, sed=(0, 0, 0, 1),
                                  dtlt=20.)
    nzp, nxp, nyp = shape
    assert vi.max() > 0.3, "no fast ice: the case does not exercise the fall distance"
    assert not any(a.any() for a in n_t[:3] + m_t[:3]) and not rr.any() and not vt.any()
    for a, src in ((n_t[3], n_p[3]), (m_t[3], m_p[3])):
        s = a.reshape(a.shape[0], nyp, nxp, nzp)[:, 2:nyp - 2, 2:nxp - 2, :]
        s0 = src.reshape(s.shape[0], nyp, nxp, nzp)[:, 2:nyp - 2, 2:nxp - 2, :]
        col = (s / dzt[None, None, None, :]).sum(axis=-1)
        scale = (np.abs(s0) / dzt[None, None, None, :]).sum(axis=-1) + 1e-300
        assert np.all(np.abs(col) <= 1e-14 * scale)
        assert (np.abs(s) > 0).sum() > 0


def test_ice_surface_rate_from_lowest_level(oracle_mod):
    """Ice only at the level above the surface: every bin there deposits, so
    the surface rate (the deposited water and rime, times alvi) equals that
    level's ice rate (water plus rime of every bin, times alvi), and the rime
    alone already gives a positive rate; the latent heat is alvi, so tlt
    changes by -(deposited water + rime)/dtlt alvi/cp th/tk summed with the
    landing level below the surface."""
    from salsa_amd import capi
    desc = capi.bin_desc(species=("SO4", "DU"))
    shape, dzt, fld, n_p, m_p, n_t, m_t = sed_fields5(desc, (12, 4, 4), 8)
    nzp, nxp, nyp = shape
    n3 = nzp * nxp * nyp
    nb = desc.nice
    k = np.arange(n3) % nzp
    n_p[3][:, k != 1] = 0.
    m_p[3][:, k != 1] = 0.
    m_p[3][(desc.iwa - 1) * nb + nb - 1] = 0.     # the rate sum starts one entry early (as written): keep it empty
    for rime_only in (False, True):
        mp = [a.copy() for a in m_p]
        if rime_only:
            mp[3][desc.iwa * nb:(desc.iwa + 1) * nb] = 0.
        n_t = [np.zeros_like(a) for a in n_p]
        m_t = [np.zeros_like(a) for a in mp]
        tlt = np.zeros(n3)
        rr, sr, ir, si, vt, vi = run5(oracle_mod, shape, dzt, fld, n_p, mp, n_t, m_t, tlt, sed=(0, 0, 0, 1))
        irk = ir.reshape(nyp, nxp, nzp)[2:nyp - 2, 2:nxp - 2, 1]
        sik = si.reshape(nyp, nxp)[2:nyp - 2, 2:nxp - 2]
        assert np.all(irk > 0)
        np.testing.assert_allclose(sik, irk, rtol=1e-13)
        # tlt: levels 0 and 1 of the interior columns (the ice leaves level 1)
        w = sum(m_t[3][q] for q in range(desc.iwa * nb, (desc.irim + 1) * nb))
        exp_tlt = -w * (ALVI / 1005.) * fld["th"] / fld["tk"]
        inner = np.zeros((nyp, nxp, nzp), bool)
        inner[2:nyp - 2, 2:nxp - 2, :nzp - 1] = True
        np.testing.assert_allclose(tlt[inner.ravel()], exp_tlt[inner.ravel()], rtol=1e-12, atol=1e-300)
        assert np.abs(tlt).max() > 0


def test_level5_liquid_categories_as_level4(oracle_mod):
    """At level 5 the aerosol, cloud and precipitation passes are those of
    level 4 on the same arrays (the wet species; mcrp.f90:797-878)."""
    from salsa_amd import capi
    desc = capi.bin_desc(species=("SO4", "DU"))
    shape, dzt, fld, n_p, m_p, n_t, m_t = sed_fields5(desc, (12, 6, 5), 9)
    L = oracle_mod.make_layout(species=(0, 3))
    tl4, tl5 = np.zeros(int(np.prod(shape))), np.zeros(int(np.prod(shape)))
    nt4, mt4 = [a.copy() for a in n_t[:3]], [a.copy() for a in m_t[:3]]
    r4 = oracle_mod.sedim(L, shape, 1.0, dzt, fld["tk"], fld["th"], fld["adn"], fld["ustar"], n_p[:3], m_p[:3],
                          nt4, mt4, tl4)
    nt5, mt5 = [a.copy() for a in n_t], [a.copy() for a in m_t]
    r5 = run5(oracle_mod, shape, dzt, fld, n_p, m_p, nt5, mt5, tl5, sed=(1, 1, 1, 0))
    for a, b in zip(r4, (r5[0], r5[1], r5[4])):
        np.testing.assert_array_equal(a, b)
    for c in range(3):
        np.testing.assert_array_equal(nt4[c], nt5[c])
        np.testing.assert_array_equal(mt4[c], mt5[c])
    np.testing.assert_array_equal(tl4, tl5)
    np.testing.assert_array_equal(nt5[3], n_t[3])
    assert not r5[2].any() and not r5[3].any() and not r5[5].any()
