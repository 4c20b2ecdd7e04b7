"""salsa_les_update_device (tend_constrain2, update_sclrs, tend0 and
SALSA_diagnostics on device, constrain_SALSA.f90:18-402, step.f90:330-333)
against the oracle's so_les_update on the same fields: bit-exact for the
updated numbers and masses, the zeroed tendencies, the zeroed charge times
and the diagnostic sums (every operation is a comparison, a move or a sum
in the reference's order; contraction is off on both sides)."""
import numpy as np
import pytest

from helpers import les_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    from salsa_amd import capi as c
    c.lib()
    return c


# Level 4 with SO4 (78 tracer rows) and level 5 with SO4 (106 rows) run the
# diagnostics pass with the rows staged in LDS; level 5 with SO4 and DU (139
# rows) runs it on the LES arrays directly.
@pytest.mark.parametrize("level,seed,grid,dry", [(4, 1, (12, 10, 9), 1), (4, 2, (30, 16, 16), 1),
                                                 (5, 3, (12, 10, 9), 2), (5, 4, (30, 16, 16), 2),
                                                 (5, 5, (30, 16, 16), 1)])
def test_les_update_matches_oracle(capi, oracle_mod, level, seed, grid, dry):
    import torch
    species = ("SO4", "DU")[:dry]
    desc = capi.bin_desc(species=species)
    shape, dzt, f, n_p, n_t, m_p, m_t = les_case(desc, level, grid, seed)
    nzp, nxp, nyp = shape
    n3 = nzp * nxp * nyp
    print("case LES update, level %d, seed %d, grid %s, species %s" % (level, seed, shape, species), flush=True)
    nrow_ct = desc.nbins + desc.ncld + desc.nprc
    ct = np.ones((nrow_ct, n3))
    # oracle
    L = oracle_mod.make_layout(species=(0, 3)[:dry])
    on, ont, om, omt = ([a.copy() for a in x] for x in (n_p, n_t, m_p, m_t))
    oct_ = ct.copy()
    four = lambda x: (list(x) + [None] * 4)[:4]
    dzf2 = dzt[nzp - 3] / dzt[nzp - 2]
    odiag = oracle_mod.les_update(L, level, shape, 1.0, dzf2, f["temp"], f["rh"], f["rhi"], f["rp"], four(on),
                                  four(ont), four(om), four(omt), oct_)
    # device
    dev = torch.device("cuda")
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    dn, dnt, dm, dmt = ([D(a) for a in x] for x in (n_p, n_t, m_p, m_t))
    df = {k: D(v) for k, v in f.items()}
    dct = D(ct)
    ddiag = torch.zeros((6, n3), dtype=torch.float64, device=dev)
    ctx = capi.Context(desc, capi.params(level=level), nzp, nxp, nyp)
    ctx.les_update_device(1.0, dzt, df, dn, dnt, dm, dmt, ddiag, charge_time=dct)
    ctx.sync()
    ctx.close()
    moved = sum(int((a != b).sum()) for a, b in zip(on, n_p))
    print("  oracle changed %d numbers; charge times zeroed: %d" % (moved, int((oct_ == 0).sum())), flush=True)
    assert moved > 0 and (oct_ == 0).any()
    for c in range(len(n_p)):
        np.testing.assert_array_equal(dn[c].cpu().numpy(), on[c], err_msg="numbers, category %d" % c)
        np.testing.assert_array_equal(dm[c].cpu().numpy(), om[c], err_msg="masses, category %d" % c)
        assert not dnt[c].any() and not dmt[c].any()
    np.testing.assert_array_equal(dct.cpu().numpy(), oct_)
    rows = 6 if level == 5 else 4
    np.testing.assert_array_equal(ddiag.cpu().numpy()[:rows], odiag[:rows])


# The optional fields: a_gaerop (tiny aerosols to the gas phase, lscndgas)
# on the tiny-aerosol case with OC + SO4 and with NO + NH at level 5 (every
# gas row the reference names; 139 rows: removal from the LES arrays), on the
# level-4 SO4 case (78 rows: removal staged in LDS), and with a_indefp
# (clamped at level 5) on the level-5 SO4+DU case.
@pytest.mark.parametrize("case", ["tiny_oc_so4", "tiny_no_nh", "les4", "les5"])
def test_les_update_extra_fields_match_oracle(capi, oracle_mod, case):
    import torch
    from test_oracle_les import tiny_aerosol_case
    if case.startswith("tiny"):
        level = 5
        species, ids = (("OC", "SO4"), (1, 0)) if case == "tiny_oc_so4" else (("NO", "NH"), (5, 6))
        desc = capi.bin_desc(species=species)
        shape, dzt, f, n_p, n_t, m_p, m_t = tiny_aerosol_case(desc, 4, level)
        dzt = np.ones(shape[0])                       # dzf2 = 1
    else:
        level = 4 if case == "les4" else 5
        species, ids = (("SO4",), (0,)) if level == 4 else (("SO4", "DU"), (0, 3))
        desc = capi.bin_desc(species=species)
        shape, dzt, f, n_p, n_t, m_p, m_t = les_case(desc, level, (12, 10, 9), 3)
    nzp, nxp, nyp = shape
    n3 = nzp * nxp * nyp
    print("case LES update with a_gaerop%s: %s, level %d, grid %s" % (" and a_indefp" if level == 5 else "", case,
                                                                      level, shape), flush=True)
    rng = np.random.default_rng(9)
    gas = rng.uniform(1e10, 1e12, (5, n3))
    ind = rng.uniform(-1., 2., (3, n3)) if level == 5 else None
    # oracle
    L = oracle_mod.make_layout(species=ids)
    on, ont, om, omt = ([a.copy() for a in x] for x in (n_p, n_t, m_p, m_t))
    ogas, oind = gas.copy(), None if ind is None else ind.copy()
    four = lambda x: (list(x) + [None] * 4)[:4]
    dzf2 = dzt[nzp - 3] / dzt[nzp - 2]
    odiag = oracle_mod.les_update(L, level, shape, 1.0, dzf2, f["temp"], f["rh"], f["rhi"], f["rp"], four(on),
                                  four(ont), four(om), four(omt), gaerop=ogas, indefp=oind)
    # device
    dev = torch.device("cuda")
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    dn, dnt, dm, dmt = ([D(a) for a in x] for x in (n_p, n_t, m_p, m_t))
    df = {k: D(v) for k, v in f.items()}
    dgas, dind = D(gas), None if ind is None else D(ind)
    ddiag = torch.zeros((6, n3), dtype=torch.float64, device=dev)
    ctx = capi.Context(desc, capi.params(level=level), nzp, nxp, nyp)
    ctx.set_les_extra_fields(dgas, dind)
    ctx.les_update_device(1.0, dzt, df, dn, dnt, dm, dmt, ddiag)
    ctx.sync()
    ctx.close()
    moved = int((ogas != gas).sum())
    print("  oracle changed %d gas entries" % moved, flush=True)
    assert moved > 0
    np.testing.assert_array_equal(dgas.cpu().numpy(), ogas)
    if ind is not None:
        np.testing.assert_array_equal(dind.cpu().numpy(), oind)
        assert (oind == 0).any() and (oind == 1).any()
    for c in range(len(n_p)):
        np.testing.assert_array_equal(dn[c].cpu().numpy(), on[c], err_msg="numbers, category %d" % c)
        np.testing.assert_array_equal(dm[c].cpu().numpy(), om[c], err_msg="masses, category %d" % c)
    rows = 6 if level == 5 else 4
    np.testing.assert_array_equal(ddiag.cpu().numpy()[:rows], odiag[:rows])


def test_les_update_rejects_other_halos(capi):
    import torch
    desc = capi.bin_desc()
    ctx = capi.Context(desc, capi.params(level=4, hz=0, hxy=0), 8, 6, 6)
    n3 = 8 * 6 * 6
    z = lambda: torch.zeros(n3 * 40, dtype=torch.float64, device="cuda")
    f = {k: z() for k in ("temp", "rh", "rp")}
    with pytest.raises(RuntimeError, match="halo"):
        ctx.les_update_device(1.0, np.ones(8), f, [z()] * 3, [z()] * 3, [z()] * 3, [z()] * 3, z())
    ctx.close()
