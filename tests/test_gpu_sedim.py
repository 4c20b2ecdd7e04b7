"""salsa_sedim_device (sedimentation: DepositionSlow for aerosols and cloud
droplets, DepositionFast for precipitation and, at level 5, ice; mcrp.f90:
745-1275) against the oracle's so_sedim / so_sedim5 on the same fields.  The fall speeds go through pow,
exp, log and cube roots, whose device and host results may differ in the
last bit (the device forms some powers as products and a Halley-refined
cube root, and the ice fall speed's powers as one exponential of a sum of
logarithms, a few 1e-14 from the power form), and the tendencies are
differences of level fluxes: the bar is
1e-10 of each quantity's scale in its column (the fields themselves for the
tendencies).
Fall-distance decisions compare such speeds with level depths; none of the
cases here sits within rounding of a decision."""
import numpy as np
import pytest

from test_oracle_sedim import sed_fields, run

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    from salsa_amd import capi as c
    c.lib()
    return c


def _compare(capi, oracle_mod, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, seed, sed, dtlt=1.0):
    """One call of salsa_sedim_device against so_sedim on the same inputs."""
    import torch
    nzp, nxp, nyp = shape
    n3 = nzp * nxp * nyp
    tlt = np.random.default_rng(seed).normal(size=n3) * 1e-4
    ont, omt, otlt = [a.copy() for a in n_t], [a.copy() for a in m_t], tlt.copy()
    rr, sr, vt = run(oracle_mod, desc, shape, dzt, fld, n_p, m_p, ont, omt, otlt, sed=sed, dtlt=dtlt)
    dev = torch.device("cuda")
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    dn, dm, dnt, dmt = ([D(a) for a in x] for x in (n_p, m_p, n_t, m_t))
    df = {k: D(v) for k, v in fld.items()}
    dtl, drr, dsr, dvt = D(tlt), torch.zeros(n3, dtype=torch.float64, device=dev), \
        torch.zeros(nxp * nyp, dtype=torch.float64, device=dev), torch.zeros((desc.nprc, n3), dtype=torch.float64,
                                                                             device=dev)
    ctx = capi.Context(desc, capi.params(level=4), nzp, nxp, nyp)
    ctx.sedim_device(dtlt, dzt, df, dn, dm, dnt, dmt, dtl, drr, dsr, dvt, sed=sed)
    ctx.sync()
    ctx.close()

    def close(name, got, want, scale):
        err = np.abs(got - want)
        bad = err > 1e-10 * scale + 1e-300
        print("  %s: max err/scale %.2e" % (name, float((err / np.maximum(scale, 1e-300)).max())), flush=True)
        assert not bad.any(), (name, np.argwhere(bad)[:5])

    colmax = lambda a: np.broadcast_to(np.abs(a).reshape(-1, nyp * nxp, nzp).max(axis=2, keepdims=True),
                                       (a.reshape(-1, nyp * nxp, nzp).shape)).reshape(a.shape)
    for c in range(3):
        close("n_t[%d]" % c, dnt[c].cpu().numpy(), ont[c], colmax(n_p[c]) / dtlt + np.abs(n_t[c]))
        close("m_t[%d]" % c, dmt[c].cpu().numpy(), omt[c], colmax(m_p[c]) / dtlt + np.abs(m_t[c]))
    close("tlt", dtl.cpu().numpy(), otlt, np.abs(otlt) + np.abs(otlt - tlt).max())
    close("rrate", drr.cpu().numpy(), rr, np.abs(rr).max() + 1e-300)
    close("sfcrrate", dsr.cpu().numpy(), sr, np.abs(sr).max() + 1e-300)
    close("VtPrc", dvt.cpu().numpy(), vt, np.abs(vt) + 1e-300)
    return ont, vt


# (the 130- and 250-level grids run columns of three and four wavefronts:
# the level-above fluxes and the fall-distance gathers cross wavefronts)
@pytest.mark.parametrize("seed,grid,sed", [(1, (20, 8, 8), (1, 1, 1)), (2, (40, 12, 10), (1, 1, 1)),
                                           (3, (20, 8, 8), (0, 0, 1)), (4, (20, 8, 8), (1, 0, 0)),
                                           (5, (130, 8, 8), (1, 1, 1)), (6, (250, 6, 6), (1, 1, 1))])
def test_sedim_matches_oracle(capi, oracle_mod, seed, grid, sed):
    desc = capi.bin_desc()
    shape, dzt, fld, n_p, m_p, n_t, m_t = sed_fields(desc, grid, seed)
    print("case sedimentation, seed %d, grid %s, switches %s" % (seed, shape, sed), flush=True)
    ont, vt = _compare(capi, oracle_mod, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, seed, sed)
    if sed[2]:
        assert vt.max() > 0.5 and np.abs(ont[2] - n_t[2]).max() > 0


@pytest.mark.parametrize("dz,dtlt,nzp", [(20., 3.0, 30), (20., 7.0, 30), (20., 11.0, 30), (2., 5.0, 40)])
def test_sedim_multi_level_falls(capi, oracle_mod, dz, dtlt, nzp):
    """Precipitation falling several levels in one call: a uniform grid of
    20 m levels with dtlt 3, 7 and 11 s, and 2 m levels with dtlt 5 s, where
    the drops cross four levels and more (the landing window of the fall
    kernel and its spill path beyond it).  All three categories sediment."""
    desc = capi.bin_desc()
    shape, dzt, fld, n_p, m_p, n_t, m_t = sed_fields(desc, (nzp, 8, 8), 11)
    dzt = np.full_like(dzt, 1. / dz)
    print("case sedimentation multi-level, dz %.0f m, dtlt %.0f s, grid %s" % (dz, dtlt, shape), flush=True)
    ont, vt = _compare(capi, oracle_mod, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, 11, (1, 1, 1), dtlt)
    levels = vt.max() * dtlt / dz
    print("  largest fall %.1f levels" % levels, flush=True)
    assert levels > (4 if dz < 10 else 1)


def test_fresh_context_reports_no_faults(capi):
    """A context that has not stepped has no cell flags: sync after a
    sedimentation-only call reports success and the flags read zero, whatever
    memory the context's flag words were carved from."""
    desc = capi.bin_desc()
    for _ in range(3):
        ctx = capi.Context(desc, capi.params(level=4), 12, 8, 8)
        ctx.sync()
        assert not ctx.cell_flags().any()
        ctx.close()


def _compare5(capi, oracle_mod, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, seed, sed, dtlt=1.0):
    """One level-5 call of salsa_sedim_device against so_sedim5."""
    import torch
    from test_oracle_sedim import run5
    nzp, nxp, nyp = shape
    n3 = nzp * nxp * nyp
    tlt = np.random.default_rng(seed).normal(size=n3) * 1e-4
    ont, omt, otlt = [a.copy() for a in n_t], [a.copy() for a in m_t], tlt.copy()
    rr, sr, ir, si, vt, vi = run5(oracle_mod, shape, dzt, fld, n_p, m_p, ont, omt, otlt, sed=sed, dtlt=dtlt)
    dev = torch.device("cuda")
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    dn, dm, dnt, dmt = ([D(a) for a in x] for x in (n_p, m_p, n_t, m_t))
    df = {k: D(v) for k, v in fld.items()}
    Z = lambda *sh: torch.full(sh, 7.0, dtype=torch.float64, device=dev)    # outputs are overwritten
    dtl, drr, dsr, dvt = D(tlt), Z(n3), Z(nxp * nyp), Z(desc.nprc, n3)
    dir_, dsi, dvi = Z(n3), Z(nxp * nyp), Z(desc.nice, n3)
    ctx = capi.Context(desc, capi.params(level=5), nzp, nxp, nyp)
    ctx.sedim_device(dtlt, dzt, df, dn, dm, dnt, dmt, dtl, drr, dsr, dvt, sed=sed, irate=dir_, sfcirate=dsi,
                     vt_ice=dvi)
    ctx.sync()
    ctx.close()

    def close(name, got, want, scale):
        err = np.abs(got - want)
        bad = err > 1e-10 * scale + 1e-300
        print("  %s: max err/scale %.2e" % (name, float((err / np.maximum(scale, 1e-300)).max())), flush=True)
        assert not bad.any(), (name, np.argwhere(bad)[:5])

    colmax = lambda a: np.broadcast_to(np.abs(a).reshape(-1, nyp * nxp, nzp).max(axis=2, keepdims=True),
                                       (a.reshape(-1, nyp * nxp, nzp).shape)).reshape(a.shape)
    for c in range(4):
        close("n_t[%d]" % c, dnt[c].cpu().numpy(), ont[c], colmax(n_p[c]) / dtlt + np.abs(n_t[c]))
        close("m_t[%d]" % c, dmt[c].cpu().numpy(), omt[c], colmax(m_p[c]) / dtlt + np.abs(m_t[c]))
    close("tlt", dtl.cpu().numpy(), otlt, np.abs(otlt) + np.abs(otlt - tlt).max())
    for name, got, want in (("rrate", drr, rr), ("sfcrrate", dsr, sr), ("irate", dir_, ir), ("sfcirate", dsi, si)):
        close(name, got.cpu().numpy(), want, np.abs(want).max() + 1e-300)
    close("VtPrc", dvt.cpu().numpy(), vt, np.abs(vt) + 1e-300)
    close("VtIce", dvi.cpu().numpy(), vi, np.abs(vi) + 1e-300)
    return ont, omt, vi, ir


@pytest.mark.parametrize("seed,grid,sed", [(21, (20, 8, 8), (1, 1, 1, 1)), (22, (40, 12, 10), (1, 1, 1, 1)),
                                           (23, (20, 8, 8), (0, 0, 0, 1)), (25, (130, 8, 8), (1, 1, 1, 1)),
                                           (26, (250, 6, 6), (1, 1, 1, 1))])
def test_sedim_level5_matches_oracle(capi, oracle_mod, seed, grid, sed):
    """Level 5 (SO4 + dust, ice with rime): the liquid categories and ice
    sedimentation (DepositionFast flag 4) against so_sedim5."""
    from test_oracle_sedim import sed_fields5
    desc = capi.bin_desc(species=("SO4", "DU"))
    shape, dzt, fld, n_p, m_p, n_t, m_t = sed_fields5(desc, grid, seed)
    print("case sedimentation level 5, seed %d, grid %s, switches %s" % (seed, shape, sed), flush=True)
    ont, omt, vi, ir = _compare5(capi, oracle_mod, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, seed, sed)
    assert vi.max() > 0.1 and np.abs(ont[3] - n_t[3]).max() > 0 and ir.max() > 0


def test_sedim_level5_ice_multi_level(capi, oracle_mod):
    """Ice falling several levels in one call (2 m levels, dtlt 20 s)."""
    from test_oracle_sedim import sed_fields5
    desc = capi.bin_desc(species=("SO4", "DU"))
    shape, dzt, fld, n_p, m_p, n_t, m_t = sed_fields5(desc, (40, 8, 8), 24)
    dzt = np.full_like(dzt, 1. / 2.)
    print("case sedimentation level 5 multi-level, grid %s" % (shape,), flush=True)
    ont, omt, vi, ir = _compare5(capi, oracle_mod, desc, shape, dzt, fld, n_p, m_p, n_t, m_t, 24, (0, 0, 0, 1),
                                 dtlt=20.)
    levels = vi.max() * 20. / 2.
    print("  largest fall %.1f levels" % levels, flush=True)
    assert levels > 4


def test_sedim_refuses_too_many_ice_bins(capi):
    import torch
    desc = capi.bin_desc(nbin=(1, 4), nprc=1, species=("SO4", "DU"), nice=17)   # 35 bins, within the context capacity
    ctx = capi.Context(desc, capi.params(level=5), 8, 6, 6)
    z = lambda: torch.zeros(8 * 36 * 40, dtype=torch.float64, device="cuda")
    f = {k: z() for k in ("tk", "th", "adn", "ustar")}
    with pytest.raises(RuntimeError, match="16 ice bins"):
        ctx.sedim_device(1.0, np.ones(8), f, [z()] * 4, [z()] * 4, [z()] * 4, [z()] * 4, z(), z(), z(), z(),
                         irate=z(), sfcirate=z(), vt_ice=z())
    ctx.close()
