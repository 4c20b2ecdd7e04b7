"""The Fortran drop-in run_SALSA (uclales-salsa_amd/fortran/mo_salsa_amd_driver.f90)
executed on the GPU.

oracle/_ref/dropin_drive (built by `make -C oracle ref` against the reference's
module files) builds the LES Diag and Prog FieldArrays of a config-1 grid and
calls the drop-in run_SALSA twice with the reference's dummy list: as the
initialization call of init.f90:101-107 (initialize = .TRUE.) and as the
time-step call of step.f90:317-328, each on zeroed tendencies.  Both calls'
tendencies and rate fields must be bitwise those of salsa_step_host called
from Python with the same flags, and their new states must match the oracle
(with and without the initialization equilibration) within the FP64 bar.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, check_condition_aware, oracle_sensitivity

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_drive")
MAGIC = 370020


@pytest.fixture(scope="module")
def capi():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    from salsa_amd import capi as c
    c.lib()
    return c


def _write_input(path, desc, f, tstep, time, istp, wp=None):
    g = f.grid
    hdr = np.array([MAGIC, g.nzp, g.nxp, g.nyp, f.level, desc.nbin[0], desc.nbin[1], desc.nprc, desc.nice,
                    desc.nspec_dry, istp, 0 if wp is None else 2], np.int32)
    ids = np.zeros(7, np.int32)
    ids[:desc.nspec_dry] = list(desc.dry_species)[:desc.nspec_dry]
    dh = np.array(list(desc.reglim) + [desc.prc_dlo, desc.prc_vratio, desc.ice_dlo, desc.ice_vratio])
    ncat = 4 if f.level == 5 else 3
    with open(path, "wb") as fh:
        hdr.tofile(fh)
        ids.tofile(fh)
        dh.astype(np.float64).tofile(fh)
        np.array([tstep, time], np.float64).tofile(fh)
        for a in (f.press, f.temp, f.rsl, f.rsi, f.dn, f.rv):
            np.ascontiguousarray(a, dtype=np.float64).tofile(fh)
        for c in range(ncat):
            np.ascontiguousarray(f.n[c], dtype=np.float64).tofile(fh)
            np.ascontiguousarray(f.m[c], dtype=np.float64).tofile(fh)
        if wp is not None:
            np.ascontiguousarray(wp, dtype=np.float64).tofile(fh)


def _read_output(path, desc, f, act=False):
    g = f.grid
    n3 = g.nzp * g.nxp * g.nyp
    ncat = 4 if f.level == 5 else 3
    nrf = 17 if f.level == 5 else 12
    raw = np.fromfile(path, dtype=np.float64)
    off = 0
    calls = []
    for _ in range(2):
        rv_t = raw[off:off + n3]
        off += n3
        n_t, m_t = [], []
        for c in range(ncat):
            a = f.n[c].size
            n_t.append(raw[off:off + a].reshape(f.n[c].shape))
            off += a
            a = f.m[c].size
            m_t.append(raw[off:off + a].reshape(f.m[c].shape))
            off += a
        rates = raw[off:off + nrf * n3].reshape(nrf, n3)
        off += nrf * n3
        if act:
            nwet = desc.nspec - 1
            na = raw[off:off + desc.ncld * n3].reshape(desc.ncld, n3)
            off += desc.ncld * n3
            va = raw[off:off + nwet * desc.ncld * n3].reshape(nwet * desc.ncld, n3)
            off += nwet * desc.ncld * n3
            calls.append((rv_t, n_t, m_t, rates, na, va))
        else:
            calls.append((rv_t, n_t, m_t, rates))
    assert off == raw.size
    return calls


@pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/dropin_drive is built by `make -C oracle ref`")
def test_fortran_dropin_run_salsa(capi, oracle_mod, tmp_path):
    from salsa_amd import workload as W
    desc = capi.bin_desc()
    f = W.config_fields(desc, 1)
    g = f.grid
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, desc, f, 1.0, 10.0, 1)
    print("dropin_drive on the config-1 grid %dx%dx%d" % (g.nzp, g.nxp, g.nyp), flush=True)
    env = dict(os.environ, SALSA_AMD_PRECISION="fp64")
    r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0 and "dropin_drive done" in r.stdout, (r.stdout[-2000:], r.stderr[-2000:])
    calls = _read_output(fout, desc, f)

    # the same calls through the C ABI from Python: bitwise the same results
    n3 = g.nzp * g.nxp * g.nyp
    for k, flags in enumerate((capi.STEP_INIT, 0)):
        ctx = capi.Context(desc, capi.params(level=4, precision=capi.SALSA_FP64), g.nzp, g.nxp, g.nyp)
        fd = dict(press=f.press, temp=f.temp, rsl=f.rsl, rsi=f.rsi, dn=f.dn, rv=f.rv)
        n_t = [np.zeros_like(x) for x in f.n[:3]]
        m_t = [np.zeros_like(x) for x in f.m[:3]]
        rv_t = np.zeros_like(f.rv)
        rates = np.zeros((12, n3))
        ctx.step_host(fd, f.n[:3], f.m[:3], n_t, m_t, rv_t, flags=flags, rates=rates, time=10.0, istp=1)
        ctx.close()
        frv, fn, fm, frates = calls[k]
        assert np.array_equal(frv, rv_t), k
        for c in range(3):
            assert np.array_equal(fn[c], n_t[c]), (k, c)
            assert np.array_equal(fm[c], m_t[c]), (k, c)
        assert np.array_equal(frates, rates), k

    # the new states against the oracle, with and without the initialization
    T, p, rs, rsi, rv, numc, volc = W.to_salsa_units(desc, f)
    L = oracle_mod.make_layout()
    P = oracle_mod.make_params(level=4)
    nt = numc.shape[1]
    nco = np.zeros((len(T), L.ntotal))
    vco = np.zeros((len(T), L.ntotal, 9))
    nco[:, :nt] = numc
    vco[:, :nt] = volc
    outs = []
    for k, init in enumerate((True, False)):
        frv, fn, fm, _ = calls[k]
        fo = W.Fields(grid=g, level=4, press=f.press, temp=f.temp, rsl=f.rsl, rsi=f.rsi, dn=f.dn,
                      rv=f.rv + frv, n=[f.n[c] + fn[c] for c in range(3)] + [None],
                      m=[f.m[c] + fm[c] for c in range(3)] + [None])
        _, _, _, _, grv, gnc, gvc = W.to_salsa_units(desc, fo)
        orv, onc, ovc, st = oracle_mod.step_cells(L, P, T, p, rs, rsi, rv, nco, vco, init=init)
        sens = oracle_sensitivity(oracle_mod, L, P, T, p, rs, rsi, rv, nco, vco, init=init)
        # the tendency round trip m_p + dt*(m_new - m_p)/dt costs a few ulp of m_p
        e, nflag = check_condition_aware((grv, gnc, gvc), (orv, onc[:, :nt], ovc[:, :nt]),
                                         (sens[0], sens[1][:, :nt], sens[2][:, :nt]), 1e-9, 2,
                                         "init" if init else "step")
        print("  %s call: max rel rv %.2e N %.2e V %.2e %.2e; inside the allowance %d"
              % ((("init" if init else "step"),) + tuple(e) + (nflag,)))
        assert nflag <= 10
        outs.append(gvc)
    # the initialization call changed the aerosol water
    iwa = desc.iwa
    d = np.abs(outs[0][:, :desc.nbins, iwa] - outs[1][:, :desc.nbins, iwa])
    assert np.count_nonzero(d > 1e-6 * np.abs(outs[1][:, :desc.nbins, iwa])) > 0.05 * d.size


@pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/dropin_drive is built by `make -C oracle ref`")
def test_fortran_dropin_cloud_base_activation(capi, tmp_path):
    """lsactiv mode 2 through the drop-in: run_SALSA copies wp to the
    device, the step activates at cloud base, and pa_nactd / pa_vactd come
    back in the reference's layout (mo_salsa_driver.f90:360-366).  Both
    calls (initialization and time step) bitwise those of salsa_step_host
    with the same fields registered from Python, halos zero, and droplets
    activated where the updraft is positive."""
    import torch
    from salsa_amd import workload as W
    desc = capi.bin_desc()
    f = W.config_fields(desc, 1)
    g = f.grid
    n3 = g.nzp * g.nxp * g.nyp
    wp = np.random.default_rng(4).uniform(-0.5, 4., n3)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, desc, f, 1.0, 10.0, 1, wp=wp)
    print("dropin_drive with cloud-base activation on the config-1 grid %dx%dx%d" % (g.nzp, g.nxp, g.nyp),
          flush=True)
    env = dict(os.environ, SALSA_AMD_PRECISION="fp64")
    r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0 and "dropin_drive done" in r.stdout, (r.stdout[-2000:], r.stderr[-2000:])
    calls = _read_output(fout, desc, f, act=True)
    nwet = desc.nspec - 1
    dev = torch.device("cuda")
    for k, flags in enumerate((capi.STEP_INIT, 0)):
        ctx = capi.Context(desc, capi.params(level=4, precision=capi.SALSA_FP64, lsactiv=2), g.nzp, g.nxp, g.nyp)
        dw = torch.from_numpy(wp).to(dev)
        na = torch.zeros((desc.ncld, n3), dtype=torch.float64, device=dev)
        va = torch.zeros((nwet * desc.ncld, n3), dtype=torch.float64, device=dev)
        ctx.set_cloud_base_fields(dw, na, va)
        fd = dict(press=f.press, temp=f.temp, rsl=f.rsl, rsi=f.rsi, dn=f.dn, rv=f.rv)
        n_t = [np.zeros_like(x) for x in f.n[:3]]
        m_t = [np.zeros_like(x) for x in f.m[:3]]
        rv_t = np.zeros_like(f.rv)
        rates = np.zeros((12, n3))
        ctx.step_host(fd, f.n[:3], f.m[:3], n_t, m_t, rv_t, flags=flags, rates=rates, time=10.0, istp=1)
        ctx.sync()
        pna, pva = na.cpu().numpy(), va.cpu().numpy()
        ctx.close()
        frv, fn, fm, frates, fna, fva = calls[k]
        assert np.array_equal(frv, rv_t), k
        for c in range(3):
            assert np.array_equal(fn[c], n_t[c]), (k, c)
            assert np.array_equal(fm[c], m_t[c]), (k, c)
        assert np.array_equal(frates, rates), k
        assert np.array_equal(fna, pna), k
        assert np.array_equal(fva, pva), k
        # interior cells with an updraft activate; halos and w <= 0 stay zero
        interior = np.zeros((g.nyp, g.nxp, g.nzp), bool)
        interior[2:-2, 2:-2, 1:-1] = True
        interior = interior.reshape(-1)
        act = fna.sum(axis=0) > 0
        print("  call %d: %d cells activate, largest pa_nactd %.3e /kg" % (k, int(act.sum()), fna.max()), flush=True)
        assert act.sum() > 0 and not (act & ~interior).any() and not (act & (wp <= 0)).any()
