"""Run the compiled reference harness (oracle/_ref/ref_harness) on a batch of
cells.  TEST INFRASTRUCTURE ONLY; available only where the reference tree was
compiled (the build container), used to write tests/golden fixtures and to
pin the C restatement."""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "ref_harness")
MAGIC = 370010


def available():
    return os.path.exists(HARNESS)


def run(L, P, level, T, p, rs, rsi, rv, numc, volc, dt=1.0, nsteps=1, mode=1, lcharge=0,
        charge_enh=0.0, kstore=False, distr=False, activ=False, init=False):
    """L, P: oracle Layout/Params (ctypes); numc [n, nt], volc [n, nt, 9] in
    SALSA units with nt = L.ntotal.  Returns dict with rv, numc, volc, rates
    (mode 1) or kmat [n, nt, nt] (mode 2), plus the reference bin table."""
    n = len(T)
    nt = L.ntotal
    hdr = np.zeros(16, np.int32)
    hdr[:] = [MAGIC, 1, L.nbin[0], L.nbin[1], L.nprc, L.nice, L.nspec_dry, level,
              P.lscoag, P.lscnd, P.lsauto_state, P.lsauto_mode, lcharge, nsteps, n, nt]
    ids = np.zeros(7, np.int32)
    ids[:L.nspec_dry] = list(L.dry_species)[:L.nspec_dry]
    dh = np.zeros(16)
    dh[:11] = [L.reglim[0], L.reglim[1], L.reglim[2], L.reglim[3], L.prc_dlo, L.prc_vratio,
               L.ice_dlo, L.ice_vratio, dt, charge_enh, P.eddy_dis]
    dh[11] = 1.0 if activ else 0.0
    dh[12] = 1.0 if init else 0.0          # first step is run_SALSA(initialize=.TRUE.)
    sw = np.zeros(16, np.int32)
    sw[:13] = [P.lscgaa, P.lscgcc, P.lscgpp, P.lscgii, P.lscgca, P.lscgpa, P.lscgia, P.lscgpc,
               P.lscgic, P.lscgip, P.lscndh2oae, P.lscndh2ocl, P.lscndh2oic]
    sw[13] = 1 if kstore else 0
    sw[14] = 1 if distr else 0
    sw[15] = mode
    cells = np.zeros((n, 5 + nt + 9 * nt))
    cells[:, 0] = T
    cells[:, 1] = p
    cells[:, 2] = rs
    cells[:, 3] = rsi
    cells[:, 4] = rv
    cells[:, 5:5 + nt] = numc
    cells[:, 5 + nt:] = np.asarray(volc).reshape(n, nt * 9)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            hdr.tofile(fh)
            ids.tofile(fh)
            dh.tofile(fh)
            sw.tofile(fh)
            cells.tofile(fh)
        pr = subprocess.run([HARNESS, fin, fout], check=True, cwd=td, capture_output=True, text=True)
        raw = open(fout, "rb").read()
    compute_s = None
    for line in pr.stdout.splitlines():
        if line.startswith("compute_s"):
            compute_s = float(line.split()[1])
    off = 0
    ntr, nprc = np.frombuffer(raw, np.int32, 2, off)
    off += 8
    tab = np.frombuffer(raw, np.float64, 3 * ntr + nprc, off).copy()
    off += 8 * (3 * ntr + nprc)
    res = {"vlolim": tab[:ntr], "vhilim": tab[ntr:2 * ntr], "dmid": tab[2 * ntr:3 * ntr],
           "precpbins": tab[3 * ntr:]}
    if mode == 4:
        out = np.frombuffer(raw, np.float64, n * 10 * L.ncld, off).reshape(n, L.ncld, 10)
        res["nact"] = out[:, :, 0].copy()
        res["vact"] = out[:, :, 1:].copy()
    elif mode == 3:
        out = np.frombuffer(raw, np.float64, n * 10 * L.nice, off).reshape(n, L.nice, 10)
        for q, key in enumerate(("dwet", "dnsp", "alpha", "beta", "gamma", "sigma", "zdn", "gg", "vc", "avis")):
            res[key] = out[:, :, q].copy()
    elif mode == 1:
        w = 1 + nt + 9 * nt + 200
        out = np.frombuffer(raw, np.float64, n * w, off).reshape(n, w)
        res["rv"] = out[:, 0].copy()
        res["numc"] = out[:, 1:1 + nt].copy()
        res["volc"] = out[:, 1 + nt:1 + nt + 9 * nt].reshape(n, nt, 9).copy()
        res["rates"] = out[:, 1 + nt + 9 * nt:].reshape(n, 20, 10).copy()
    else:
        out = np.frombuffer(raw, np.float64, n * nt * nt, off).reshape(n, nt, nt)
        res["kmat"] = np.transpose(out, (0, 2, 1)).copy()
    res["compute_s"] = compute_s
    return res
