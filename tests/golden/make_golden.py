"""Write the golden fixtures of tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_harness (the UCLALES-SALSA Fortran compiled from its own
sources by `make -C oracle ref`) on seeded cells and stores inputs and outputs
as small compressed .npz files (data only).  Needs the reference tree, so it
runs only in the build container:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "uclales-salsa_amd"))

import oracle as O  # noqa: E402
import refio  # noqa: E402
from salsa_amd import workload as W  # noqa: E402


def pad(L, numc, volc):
    n, nt0 = numc.shape
    nc = np.zeros((n, L.ntotal))
    vc = np.zeros((n, L.ntotal, 9))
    nc[:, :nt0] = numc
    vc[:, :nt0] = volc
    return nc, vc


def pick(f, n, seed):
    """Sample cells so that cloudy, humid and dry cells are all present."""
    rng = np.random.default_rng(seed)
    T, p, rs, rsi, rv, numc, volc = f
    rh = rv / rs
    cloudy = numc[:, 11:19].max(axis=1) > 1.0
    groups = [np.where(cloudy)[0], np.where(~cloudy & (rh > 0.98))[0], np.where(~cloudy & (rh <= 0.98))[0]]
    out = []
    for g in groups:
        if len(g):
            out.append(rng.choice(g, size=min(len(g), n // 3 + 1), replace=False))
    return np.sort(np.concatenate(out))[:n]


def save_step(name, L, level, P, cells, nsteps=1, lcharge=0, enh=0.0, kstore=False, distr=False, activ=False,
              init=False):
    T, p, rs, rsi, rv, nc, vc = cells
    R = refio.run(L, P, level, T, p, rs, rsi, rv, nc, vc, nsteps=nsteps, lcharge=lcharge, charge_enh=enh,
                  kstore=kstore, distr=distr, activ=activ, init=init)
    ns = L.nspec
    np.savez_compressed(
        os.path.join(HERE, name),
        level=level, nsteps=nsteps, lcharge=lcharge, charge_enh=enh, kstore=int(kstore), distr=int(distr),
        lsactiv=int(activ), init=int(init),
        lsauto=P.lsauto_state, lsauto_mode=P.lsauto_mode, nbin=np.array(list(L.nbin)), nprc=L.nprc, nice=L.nice,
        species=np.array(list(L.dry_species)[:L.nspec_dry]),
        T=T, p=p, rs=rs, rsi=rsi, rv=rv, numc=nc, volc=vc[:, :, :ns],
        out_rv=R["rv"], out_numc=R["numc"], out_volc=R["volc"][:, :, :ns], out_rates=R["rates"])
    print("wrote", name, len(T), "cells")


def kstore_case():
    """lscoag%mode = 2: one update step then two steps on the stored kernels,
    cloudy cells with precipitation (written separately: `--only kstore`)."""
    L = O.make_layout()
    P4 = O.make_params(level=4)
    f4 = W.to_salsa_units(L, W.config_fields(L, 4, grid_override=(20, 8, 8)), level=4)
    T, p, rs, rsi, rv, numc, volc = f4
    nc, vc = pad(L, numc, volc)
    cells = (T, p, rs, rsi, rv, nc, vc)
    sel = pick(cells, 120, 4)[:60]
    save_step("step_cfg4_kstore.npz", L, 4, P4, tuple(x[sel] for x in cells), nsteps=3, kstore=True)


def distr_cases():
    """Steps ending with distr_update (mo_salsa.f90:148-149): config-1 cells
    over three steps, cloudy cells with precipitation over two, mixed-phase
    cells over one (written separately: `--only distr`)."""
    L = O.make_layout()
    P4 = O.make_params(level=4, lsdistupdate=1)
    f1 = W.to_salsa_units(L, W.config_fields(L, 1), level=4)
    nc, vc = pad(L, f1[5], f1[6])
    cells = tuple(f1[:5]) + (nc, vc)
    sel = pick(cells, 120, 11)
    save_step("step_cfg1_distr.npz", L, 4, P4, tuple(x[sel] for x in cells), nsteps=3, distr=True)
    f4 = W.to_salsa_units(L, W.config_fields(L, 4, grid_override=(20, 8, 8)), level=4)
    nc, vc = pad(L, f4[5], f4[6])
    cells = tuple(f4[:5]) + (nc, vc)
    sel = pick(cells, 120, 14)
    save_step("step_cfg4_distr.npz", L, 4, P4, tuple(x[sel] for x in cells), nsteps=2, distr=True)
    L5 = O.make_layout(species=(0, 3))
    P5 = O.make_params(level=5, lsdistupdate=1)
    f5 = W.to_salsa_units(L5, W.config_fields(L5, 5, grid_override=(20, 8, 8)), level=5)
    sel = np.sort(np.random.default_rng(15).choice(len(f5[0]), 80, replace=False))
    save_step("step_cfg5_distr.npz", L5, 5, P5, tuple(x[sel] for x in f5), distr=True)


def autoconv2_cases():
    """lsauto%mode = 2 (autoconv2 after condensation, mo_salsa.f90:110-112, and
    no precipitation formation inside coagulation): config-1 cells over three
    steps, cloudy cells with precipitation over two with distr_update, and
    mixed-phase cells over one (written separately: `--only autoconv2`)."""
    L = O.make_layout()
    P4 = O.make_params(level=4, lsauto_mode=2)
    f1 = W.to_salsa_units(L, W.config_fields(L, 1), level=4)
    nc, vc = pad(L, f1[5], f1[6])
    cells = tuple(f1[:5]) + (nc, vc)
    sel = pick(cells, 120, 21)
    save_step("step_cfg1_auto2.npz", L, 4, P4, tuple(x[sel] for x in cells), nsteps=3)
    f4 = W.to_salsa_units(L, W.config_fields(L, 4, grid_override=(20, 8, 8)), level=4)
    nc, vc = pad(L, f4[5], f4[6])
    cells = tuple(f4[:5]) + (nc, vc)
    sel = pick(cells, 120, 24)
    save_step("step_cfg4_auto2_distr.npz", L, 4, O.make_params(level=4, lsauto_mode=2, lsdistupdate=1),
              tuple(x[sel] for x in cells), nsteps=2, distr=True)
    L5 = O.make_layout(species=(0, 3))
    P5 = O.make_params(level=5, lsauto_mode=2)
    f5 = W.to_salsa_units(L5, W.config_fields(L5, 5, grid_override=(20, 8, 8)), level=5)
    sel = np.sort(np.random.default_rng(25).choice(len(f5[0]), 80, replace=False))
    save_step("step_cfg5_auto2.npz", L5, 5, P5, tuple(x[sel] for x in f5))


def activation_cases():
    """Interstitial cloud activation (lsactiv%mode 1, mo_salsa.f90:115-119):
    config-1 cells over three steps, cloudy cells with precipitation over two
    with autoconv2 and distr_update as well, and mixed-phase cells over one
    (written separately: `--only activation`)."""
    L = O.make_layout()
    f1 = W.to_salsa_units(L, W.config_fields(L, 1), level=4)
    nc, vc = pad(L, f1[5], f1[6])
    cells = tuple(f1[:5]) + (nc, vc)
    sel = pick(cells, 120, 31)
    save_step("step_cfg1_activ.npz", L, 4, O.make_params(level=4, lsactiv=1), tuple(x[sel] for x in cells),
              nsteps=3, activ=True)
    f4 = W.to_salsa_units(L, W.config_fields(L, 4, grid_override=(20, 8, 8)), level=4)
    nc, vc = pad(L, f4[5], f4[6])
    cells = tuple(f4[:5]) + (nc, vc)
    sel = pick(cells, 120, 34)
    save_step("step_cfg4_activ_auto2_distr.npz", L, 4,
              O.make_params(level=4, lsactiv=1, lsauto_mode=2, lsdistupdate=1), tuple(x[sel] for x in cells),
              nsteps=2, distr=True, activ=True)
    L5 = O.make_layout(species=(0, 3))
    f5 = W.to_salsa_units(L5, W.config_fields(L5, 5, grid_override=(20, 8, 8)), level=5)
    sel = np.sort(np.random.default_rng(35).choice(len(f5[0]), 80, replace=False))
    save_step("step_cfg5_activ.npz", L5, 5, O.make_params(level=5, lsactiv=1), tuple(x[sel] for x in f5),
              activ=True)


def init_cases():
    """run_SALSA's initialization call (initialize = .TRUE., mo_salsa_driver.f90:
    282-283): equilibration at RH 0.3 before the step, then a second, normal
    step.  Config-1 cells (water drawn far from the RH-0.3 equilibrium, so the
    initialization moves it) and mixed-phase cells over one step (written
    separately: `--only init`)."""
    L = O.make_layout()
    f1 = W.to_salsa_units(L, W.config_fields(L, 1), level=4)
    nc, vc = pad(L, f1[5], f1[6])
    cells = tuple(f1[:5]) + (nc, vc)
    sel = pick(cells, 120, 41)
    save_step("step_cfg1_init.npz", L, 4, O.make_params(level=4), tuple(x[sel] for x in cells), nsteps=2,
              init=True)
    L5 = O.make_layout(species=(0, 3))
    f5 = W.to_salsa_units(L5, W.config_fields(L5, 5, grid_override=(20, 8, 8)), level=5)
    sel = np.sort(np.random.default_rng(45).choice(len(f5[0]), 80, replace=False))
    save_step("step_cfg5_init.npz", L5, 5, O.make_params(level=5), tuple(x[sel] for x in f5), init=True)


def ice_fall_cases():
    """DepositionFast's per-bin fall-speed chain for ice (flag 4, mcrp.f90:
    1146-1176) through the reference's calcDiamLES, getDiameter,
    getShapeCoefficients and terminal_vel (harness mode 3), on SO4 + dust
    ice bins whose particle masses (1e-14 to 1e-7 kg) and rime fractions
    (0 to 0.95) cover the four mass-diameter regimes of mo_ice_shape.f90
    (written separately: `--only icefall`)."""
    L5 = O.make_layout(species=(0, 3))
    P5 = O.make_params(level=5)
    rng = np.random.default_rng(51)
    n, nt, ni, i0 = 60, L5.ntotal, L5.nice, L5.ntotal - L5.nice
    T = rng.uniform(240., 275., n)
    adn = rng.uniform(0.8, 1.3, n)
    numc = np.zeros((n, nt))
    volc = np.zeros((n, nt, 9))
    num = 10 ** rng.uniform(-5, 5, (n, ni))
    mp = 10 ** rng.uniform(-14, -7, (n, ni))
    fr = rng.uniform(0., 0.95, (n, ni))
    fr[rng.random((n, ni)) < 0.2] = 0.
    tot = mp * num
    dry = tot * 10 ** rng.uniform(-4, -2, (n, ni))
    numc[:, i0:] = num
    volc[:, i0:, 0] = 0.3 * dry                     # SO4
    volc[:, i0:, 1] = 0.7 * dry                     # dust
    volc[:, i0:, 2] = tot * (1. - fr) - dry         # pristine ice (water)
    volc[:, i0:, 3] = tot * fr                      # rime
    z = np.zeros(n)
    R = refio.run(L5, P5, 5, T, adn, z, z, z, numc, volc, mode=3)
    np.savez_compressed(os.path.join(HERE, "ice_fall_cfg5.npz"), T=T, adn=adn, numc=num,
                        mass=volc[:, i0:, :L5.nspec],
                        **{k: R[k] for k in ("dwet", "dnsp", "alpha", "beta", "gamma", "sigma", "zdn", "gg", "vc")})
    print("wrote ice_fall_cfg5.npz", n * ni, "bins")


def cloud_base_cases():
    """Cloud-base activation (cloud_activation with lsactiv%mode 2, harness
    mode 4) of config-1 and config-4 cells with updrafts w from -0.5 to 8
    m/s (w <= 0 activates nothing), and of SO4 + dust cells at level 5
    (written separately: `--only cloudbase`)."""
    out = {}
    for name, L, cfg, level, seed in (("l4", O.make_layout(), 1, 4, 61), ("l4b", O.make_layout(), 4, 4, 62),
                                      ("l5", O.make_layout(species=(0, 3)), 5, 5, 63)):
        P = O.make_params(level=level)
        f = W.to_salsa_units(L, W.config_fields(L, cfg, grid_override=(20, 8, 8) if cfg != 1 else None),
                             level=level)
        nc, vc = pad(L, f[5], f[6])
        rng = np.random.default_rng(seed)
        sel = np.sort(rng.choice(len(f[0]), 100, replace=False))
        w = rng.uniform(-0.5, 8., len(sel))
        w[:8] = [-0.5, 0., 0.05, 0.2, 1., 3., 6., 8.]
        R = refio.run(L, P, level, f[0][sel], f[1][sel], f[2][sel], w, f[4][sel], nc[sel], vc[sel], mode=4)
        out.update({name + "_T": f[0][sel], name + "_p": f[1][sel], name + "_w": w, name + "_numc": nc[sel],
                    name + "_volc": vc[sel][:, :, :L.nspec], name + "_nact": R["nact"],
                    name + "_vact": R["vact"][:, :, :L.nspec]})
        print(name, "activated cells:", int((R["nact"].sum(axis=1) > 0).sum()), "of", len(sel))
    np.savez_compressed(os.path.join(HERE, "act_cloudbase.npz"), **out)
    print("wrote act_cloudbase.npz")


def main():
    if len(sys.argv) > 2 and sys.argv[1] == "--only" and sys.argv[2] == "cloudbase":
        return cloud_base_cases()
    if len(sys.argv) > 2 and sys.argv[1] == "--only" and sys.argv[2] == "icefall":
        return ice_fall_cases()
    if len(sys.argv) > 2 and sys.argv[1] == "--only" and sys.argv[2] == "init":
        return init_cases()
    if len(sys.argv) > 2 and sys.argv[1] == "--only" and sys.argv[2] == "kstore":
        return kstore_case()
    if len(sys.argv) > 2 and sys.argv[1] == "--only" and sys.argv[2] == "distr":
        return distr_cases()
    if len(sys.argv) > 2 and sys.argv[1] == "--only" and sys.argv[2] == "autoconv2":
        return autoconv2_cases()
    if len(sys.argv) > 2 and sys.argv[1] == "--only" and sys.argv[2] == "activation":
        return activation_cases()
    if not refio.available():
        sys.exit("oracle/_ref/ref_harness missing: run `make -C oracle ref` first")
    # warm layout nbin=(3,4): aerosol 11, cloud 8, precip 7 (+ 7 inert ice bins)
    L = O.make_layout()
    P4 = O.make_params(level=4)
    R = refio.run(L, P4, 4, *[np.zeros(1)] * 5, np.zeros((1, L.ntotal)), np.zeros((1, L.ntotal, 9)))
    np.savez_compressed(os.path.join(HERE, "bins_legal.npz"), vlolim=R["vlolim"], vhilim=R["vhilim"],
                        dmid=R["dmid"], precpbins=R["precpbins"])
    f1 = W.to_salsa_units(L, W.config_fields(L, 1), level=4)
    T, p, rs, rsi, rv, numc, volc = f1
    nc, vc = pad(L, numc, volc)
    cells = (T, p, rs, rsi, rv, nc, vc)
    sel = pick(cells, 240, 1)
    save_step("step_cfg1.npz", L, 4, P4, tuple(x[sel] for x in cells))
    save_step("step_cfg1_charge.npz", L, 4, O.make_params(level=4, lcharge=1, charge_coll_enh=0.3),
              tuple(x[sel[:60]] for x in cells), lcharge=1, enh=0.3)
    # kernel matrices of cloudy cells
    cl = sel[nc[sel, 11:19].max(axis=1) > 1.0][:12]
    Rk = refio.run(L, P4, 4, T[cl], p[cl], rs[cl], rsi[cl], rv[cl], nc[cl], vc[cl], mode=2)
    np.savez_compressed(os.path.join(HERE, "kmat_cfg1.npz"), T=T[cl], p=p[cl], numc=nc[cl],
                        volc=vc[cl][:, :, :L.nspec], kmat=Rk["kmat"])
    print("wrote kmat_cfg1.npz", len(cl), "cells")
    # cloudy cells with precipitation present, with and without precip formation
    f4 = W.to_salsa_units(L, W.config_fields(L, 4, grid_override=(20, 8, 8)), level=4)
    T, p, rs, rsi, rv, numc, volc = f4
    nc, vc = pad(L, numc, volc)
    cells = (T, p, rs, rsi, rv, nc, vc)
    sel = pick(cells, 120, 4)
    save_step("step_cfg4.npz", L, 4, P4, tuple(x[sel] for x in cells))
    save_step("step_cfg4_noauto.npz", L, 4, O.make_params(level=4, lsauto_state=0),
              tuple(x[sel[:60]] for x in cells))
    # mixed phase: SO4 + dust, ice bins, T 240-275 K
    L5 = O.make_layout(species=(0, 3))
    P5 = O.make_params(level=5)
    R5 = refio.run(L5, P5, 5, *[np.zeros(1)] * 5, np.zeros((1, L5.ntotal)), np.zeros((1, L5.ntotal, 9)))
    np.savez_compressed(os.path.join(HERE, "bins_ice.npz"), vlolim=R5["vlolim"], vhilim=R5["vhilim"],
                        dmid=R5["dmid"], precpbins=R5["precpbins"])
    f5 = W.to_salsa_units(L5, W.config_fields(L5, 5, grid_override=(20, 8, 8)), level=5)
    sel = np.random.default_rng(5).choice(len(f5[0]), 120, replace=False)
    save_step("step_cfg5.npz", L5, 5, P5, tuple(x[np.sort(sel)] for x in f5))
    kstore_case()
    distr_cases()


if __name__ == "__main__":
    main()
