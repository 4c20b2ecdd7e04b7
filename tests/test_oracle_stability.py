"""The reference algorithm's own sensitivity in a cell of the 32x32x80
workload (interior cell 3568, 20 steps of 1 s): a relative change of 1e-11 in
the collision-kernel enhancement grows by about two orders of magnitude per
step once the cell's smallest cloud bin starts to oscillate in the
condensation substeps, and saturates near 1e-3.  The derivative of such a
trajectory is finite but carries no information about any perturbation that
FP64 can resolve; salsa_gradient_device flags these cells
(CELL_GRAD_UNSTABLE) and leaves them out of its dJ/dE total.  Runs the FP64
oracle only."""
import numpy as np


def test_oracle_trajectory_unstable_cell(oracle_mod):
    from salsa_amd import capi, workload as W
    O = oracle_mod
    desc = capi.bin_desc()
    T, p, rs, rsi, rv, numc, volc = W.to_salsa_units(desc, W.config_fields(desc, 3))
    i = 3568
    L = O.make_layout()
    nt = numc.shape[1]
    cloud0 = desc.nbins                      # first cloud bin, master index

    def trajectory(enh):
        P = O.make_params(level=4, lcharge=1, charge_coll_enh=enh)
        nc = np.zeros((1, L.ntotal))
        nc[:, :nt] = numc[i:i + 1]
        vc = np.zeros((1, L.ntotal, 9))
        vc[:, :nt] = volc[i:i + 1]
        x = rv[i:i + 1].copy()
        out = []
        for _ in range(20):
            st = O.step_batch(L, P, T[i:i + 1].copy(), p[i:i + 1].copy(), rs[i:i + 1].copy(), rsi[i:i + 1].copy(),
                              x, nc, vc, 1.0, 1)
            assert st == 0
            out.append(vc[0, cloud0, desc.nspec - 2])    # water of the first cloud bin
        return np.array(out)

    a = trajectory(0.0)
    b = trajectory(1e-11)
    d = np.abs(b - a) / np.abs(a)
    print("relative divergence per step:", " ".join("%.1e" % x for x in d))
    grow = d[6:10] / np.maximum(d[5:9], 1e-300)
    assert np.all(grow > 30.)                # two orders of magnitude per step
    assert d.max() > 1e-5 and d.max() < 1e-1  # then bounded: the clamp keeps the values finite
