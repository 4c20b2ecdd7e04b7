"""Cloud-base activation (params.lsactiv = 2: cloud_activation with
lsactiv%mode 2 after condensation, mo_salsa.f90:115-119; ActCloudBase and
activate3, mo_salsa_cloud.f90:161-375, 1199-1378) on the GPU against the
oracle's act_cloud_base (itself bit-identical to the compiled reference on
tests/golden/act_cloudbase.npz):

* the step's state is that of the same step without activation (the mode
  changes nothing in the cell);
* the activated droplets written to pa_nactd / pa_vactd equal the oracle's
  activation of the state the step produced (the state it saw: no process
  follows it here), to 1e-9 of the cell's activated total;
* against the oracle's own step with activation on the same inputs, to
  1e-6 of the cell's activated total on at least 95 % of the cells (the
  critical-bin search is a discrete decision on the post-condensation
  state);
* with autoconv2 before it and distr_update after it, the activation sees
  the state between the two, as the oracle's step does."""
import numpy as np
import pytest

from helpers import FlatCells, full_volc, golden, oracle_layout_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    from salsa_amd import capi as c
    c.lib()
    return c


def run(capi, desc, level, prec, cells, w, lsactiv):
    import torch
    fc = FlatCells(desc, level, *cells)
    prm = capi.params(level=level, precision=prec, hz=0, hxy=0, lsactiv=lsactiv)
    ctx = capi.Context(desc, prm, fc.n, 1, 1)
    n_p = [x.copy() for x in fc.n_p]
    m_p = [x.copy() for x in fc.m_p]
    f = dict(fc.f)
    f["rv"] = f["rv"].copy()
    nwet = desc.nspec - 1
    dev = torch.device("cuda")
    dw = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    na = torch.full((desc.ncld, fc.n), -1.0, dtype=torch.float64, device=dev)
    va = torch.full((nwet * desc.ncld, fc.n), -1.0, dtype=torch.float64, device=dev)
    if lsactiv == 2:
        ctx.set_cloud_base_fields(dw, na, va)
    ctx.step_host(f, n_p, m_p, flags=1)
    ctx.sync()
    out = fc.unpack(n_p, m_p, f["rv"])
    inp = fc.unpack(fc.n_p, fc.m_p, fc.f["rv"])
    ctx.close()
    return inp, out, na.cpu().numpy(), va.cpu().numpy(), fc


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("name", ["step_cfg1.npz", "step_cfg4.npz", "step_cfg5.npz"])
def test_gpu_cloud_base_activation(capi, oracle_mod, name, prec):
    O = oracle_mod
    g = golden(name)
    level = int(g["level"])
    species = tuple(int(s) for s in g["species"])
    desc = capi.bin_desc(species=species)
    nt = desc.nbins + desc.ncld + desc.nprc + (desc.nice if level == 5 else 0)
    cells = (g["T"], g["p"], g["rs"], g["rsi"], g["rv"], g["numc"][:, :nt], full_volc(g["volc"])[:, :nt])
    n = len(g["T"])
    rng = np.random.default_rng(7)
    w = rng.uniform(-0.5, 6., n)
    w[:4] = [-0.5, 0., 0.3, 5.]
    print("case %s prec %d, %d cells" % (name, prec, n), flush=True)
    inp0, out0, _, _, _ = run(capi, desc, level, prec, cells, w, 0)
    inp, out, na, va, fc = run(capi, desc, level, prec, cells, w, 2)
    # the mode leaves the cell as it is
    for a, b in zip(out0, out):
        np.testing.assert_array_equal(a, b)
    L = oracle_layout_for(g)
    nwet = desc.nspec - 1
    dn = fc.f["dn"]
    rho = np.array([desc.rholiq[s] for s in range(nwet)])
    scale_of = lambda nact, vact: (np.abs(nact).sum(axis=1) + 1e-300, np.abs(vact).sum(axis=(1, 2)) + 1e-300)
    # 1. against the oracle's activation of the state the step produced
    nco = np.zeros((n, L.ntotal))
    vco = np.zeros((n, L.ntotal, 9))
    nco[:, :nt] = out[1]
    vco[:, :nt] = out[2]
    on = np.zeros((n, desc.ncld))
    ov = np.zeros((n, desc.ncld, nwet))
    for c in range(n):
        a, v = O.act_cloud_base(L, g["T"][c], g["p"][c], w[c], nco[c], vco[c])
        on[c] = a / dn[c]
        ov[c] = v[:, :nwet] * rho[None, :] / dn[c]
    gn = na.T
    gv = va.reshape(nwet, desc.ncld, n).transpose(2, 1, 0)
    sn, sv = scale_of(on, ov)
    en = np.abs(gn - on).max(axis=1) / sn
    ev = np.abs(gv - ov).max(axis=(1, 2)) / sv
    print("  vs oracle on the step's own state: N %.2e V %.2e, cells activating %d / %d"
          % (en.max(), ev.max(), int((on.sum(axis=1) > 0).sum()), n), flush=True)
    assert (w <= 0).sum() >= 2 and np.all(gn[w <= 0] == 0) and np.all(gv[w <= 0] == 0)
    assert (on.sum(axis=1) > 0).sum() > n // 2
    assert en.max() <= 1e-9 and ev.max() <= 1e-9
    # 2. against the oracle's own step with activation (FP64 path)
    if prec == 0:
        nci = np.zeros((n, L.ntotal))
        vci = np.zeros((n, L.ntotal, 9))
        nci[:, :nt] = inp[1]
        vci[:, :nt] = inp[2]
        P = O.make_params(level=level, lsactiv=2)
        _, _, _, _, sn_, sv_ = O.step_cells_activ(L, P, g["T"], g["p"], g["rs"], g["rsi"], inp[0], nci, vci, w)
        on2 = sn_ / dn[:, None]
        ov2 = sv_[:, :, :nwet] * rho[None, None, :] / dn[:, None, None]
        s2n, s2v = scale_of(on2, ov2)
        e2 = np.maximum(np.abs(gn - on2).max(axis=1) / s2n, np.abs(gv - ov2).max(axis=(1, 2)) / s2v)
        print("  vs the oracle's step: median %.2e, cells above 1e-6: %d" % (np.median(e2), int((e2 > 1e-6).sum())))
        assert (e2 <= 1e-6).mean() >= 0.95


@pytest.mark.parametrize("name", ["step_cfg4.npz", "step_cfg5.npz"])
def test_gpu_cloud_base_activation_with_later_processes(capi, oracle_mod, name):
    """autoconv2 before and distr_update after the activation (mo_salsa.f90:
    110-149): the activation sees the state after autoconv2 and before the
    re-binning, so the result differs from an activation of the step's final
    state; against the oracle's step with both processes (FP64)."""
    O = oracle_mod
    g = golden(name)
    level = int(g["level"])
    species = tuple(int(s) for s in g["species"])
    desc = capi.bin_desc(species=species)
    nt = desc.nbins + desc.ncld + desc.nprc + (desc.nice if level == 5 else 0)
    cells = (g["T"], g["p"], g["rs"], g["rsi"], g["rv"], g["numc"][:, :nt], full_volc(g["volc"])[:, :nt])
    n = len(g["T"])
    w = np.random.default_rng(8).uniform(0.1, 6., n)
    print("case %s with autoconv2 and distr_update, %d cells" % (name, n), flush=True)
    import torch
    fc = FlatCells(desc, level, *cells)
    prm = capi.params(level=level, precision=0, hz=0, hxy=0, lsactiv=2, lsauto_mode=2, lsdistupdate=1)
    ctx = capi.Context(desc, prm, fc.n, 1, 1)
    n_p = [x.copy() for x in fc.n_p]
    m_p = [x.copy() for x in fc.m_p]
    f = dict(fc.f)
    f["rv"] = f["rv"].copy()
    nwet = desc.nspec - 1
    dev = torch.device("cuda")
    na = torch.zeros((desc.ncld, n), dtype=torch.float64, device=dev)
    va = torch.zeros((nwet * desc.ncld, n), dtype=torch.float64, device=dev)
    ctx.set_cloud_base_fields(torch.from_numpy(w).to(dev), na, va)
    ctx.step_host(f, n_p, m_p, flags=1)
    ctx.sync()
    inp = fc.unpack(fc.n_p, fc.m_p, fc.f["rv"])
    ctx.close()
    L = oracle_layout_for(g)
    P = O.make_params(level=level, lsactiv=2, lsauto_mode=2, lsdistupdate=1)
    nci = np.zeros((n, L.ntotal))
    vci = np.zeros((n, L.ntotal, 9))
    nci[:, :nt] = inp[1]
    vci[:, :nt] = inp[2]
    _, _, _, _, on, ov = O.step_cells_activ(L, P, g["T"], g["p"], g["rs"], g["rsi"], inp[0], nci, vci, w)
    dn = fc.f["dn"]
    rho = np.array([desc.rholiq[s] for s in range(nwet)])
    on = on / dn[:, None]
    ov = ov[:, :, :nwet] * rho[None, None, :] / dn[:, None, None]
    gn = na.cpu().numpy().T
    gv = va.cpu().numpy().reshape(nwet, desc.ncld, n).transpose(2, 1, 0)
    e = np.maximum(np.abs(gn - on).max(axis=1) / (np.abs(on).sum(axis=1) + 1e-300),
                   np.abs(gv - ov).max(axis=(1, 2)) / (np.abs(ov).sum(axis=(1, 2)) + 1e-300))
    print("  vs the oracle's step: median %.2e, cells above 1e-6: %d, activating %d"
          % (np.median(e), int((e > 1e-6).sum()), int((on.sum(axis=1) > 0).sum())), flush=True)
    assert (on.sum(axis=1) > 0).sum() > n // 2
    assert (e <= 1e-6).mean() >= 0.95
