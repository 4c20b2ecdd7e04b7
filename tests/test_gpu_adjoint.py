"""Reverse-mode gradient of the multi-step trajectory (salsa_adjoint_device):
the objective of salsa_gradient_device (J = precipitation volume
concentration summed over cells after n in-place steps), with dJ/dN0 of
every bin and dJ/dE from one adjoint sweep over checkpointed steps.

Checks:
  * against the forward-mode tangents of salsa_gradient_device, which are
    themselves checked against central differences of the oracle
    (test_gpu_gradient.py): per element, as a dot product with random
    directions, and per cell for dJ/dE, both precision paths, checkpoint
    intervals 1, 3, 5 and 10 (the checkpoint interval must not change a bit
    of the result);
  * against central differences of the FP64 oracle over 10 steps with the
    forward-mode test's bars (check_against_oracle), both precision paths;
  * at the 32x32x80 workload over 20 steps with checkpoints every 5 steps
    against the forward tangents, with the run time of both;
  * the torch.autograd.Function wrapper (salsa_amd.autograd);
  * the configurations the adjoint does not cover are refused.

The two derivative codes differentiate the same discrete step, but their
trajectories differ in the order of floating-point sums; on cells whose
trajectory amplifies rounding (flagged by either code, or whose derivative
contribution |g x0| exceeds 100 J) the comparison is left out.  Bar per
element: |g_adj - g_fwd| x0 <= tol (J + max_b |g_fwd,b x0,b|) of the cell,
i.e. relative to the cell's largest derivative contribution, with tol 1e-9
on the FP64 path (J per cell within 1e-10), for all but 0.1 % of the
elements, and every element within the finite-difference bar 1e-4.  On the FP32-mixed path the two
codes evaluate the collision kernels in different FP32 forms (the forward
tangents in dual arithmetic, the adjoint with the step kernel's FP64 fall
speeds), which moves J by up to a few 1e-7; there the bars are the
precision path's own: J within 1e-5, derivatives within 1e-4 (the finite
difference bar).
"""
import time

import numpy as np
import pytest

from helpers import FlatCells

pytestmark = pytest.mark.gpu

TOL = {0: (1e-10, 1e-9), 1: (1e-5, 1e-4)}   # precision: (J per cell, derivatives)


@pytest.fixture(scope="module")
def capi():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    from salsa_amd import capi as c
    c.lib()
    return c


def _flat_run(capi, desc, prec, cells, nsteps, every=None, **kw):
    """Forward-mode (every None) or reverse-mode gradient on flat cells."""
    import torch
    T, p, rs, rsi, rv, numc, volc = cells
    fc = FlatCells(desc, 4, T, p, rs, rsi, rv, numc, volc)
    ctx = capi.Context(desc, capi.params(level=4, precision=prec, hz=0, hxy=0, **kw), fc.n, 1, 1)
    dev = torch.device("cuda")
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    f = {k: D(v) for k, v in fc.f.items()}
    n_p = [D(x) for x in fc.n_p]
    m_p = [D(x) for x in fc.m_p]
    n0 = [x.clone() for x in n_p]
    djdn = [torch.zeros_like(x) for x in n_p]
    jc = torch.zeros(fc.n, dtype=torch.float64, device=dev)
    je = torch.zeros_like(jc)
    if every is None:
        J, dJdE = ctx.gradient_device(f, n_p, m_p, djdn, tstep=1.0, nsteps=nsteps, j_cell=jc, djde_cell=je)
    else:
        J, dJdE = ctx.adjoint_device(f, n_p, m_p, djdn, tstep=1.0, nsteps=nsteps, checkpoint_every=every,
                                     j_cell=jc, djde_cell=je)
    flags = ctx.cell_flags()
    ctx.close()
    for a, b in zip(n_p, n0):            # inputs are left untouched
        assert torch.equal(a, b)
    g = np.concatenate([x.cpu().numpy() for x in djdn], axis=0).T      # [cell, bin]
    return J, dJdE, jc.cpu().numpy(), je.cpu().numpy(), g, flags


def compare(name, capi, fwd, adj, numc, prec):
    """Element, dot-product and per-cell dJ/dE agreement of the adjoint with
    the forward tangents on the cells neither code flags."""
    Jf, dEf, jcf, jef, gf, flf = fwd
    Ja, dEa, jca, jea, ga, fla = adj
    jtol, rel = TOL[prec]
    ok = ((flf | fla) & capi.CELL_GRAD_UNSTABLE) == 0
    contrib = np.abs(gf) * numc
    scale = np.abs(jcf) + contrib.max(axis=1)
    ok &= contrib.max(axis=1) <= 100. * np.abs(jcf) + 1e-300
    print("%s: %d cells, flagged fwd %d adj %d, compared %d" % (
        name, len(ok), int((flf & capi.CELL_GRAD_UNSTABLE != 0).sum()),
        int((fla & capi.CELL_GRAD_UNSTABLE != 0).sum()), int(ok.sum())), flush=True)
    assert ok.mean() >= 0.9
    # the value part: the same trajectory up to summation order
    dj = np.abs(jca - jcf) / np.maximum(np.abs(jcf), 1e-300)
    print("  J per cell: max rel diff %.2e" % dj[ok].max())
    assert np.all(dj[ok] <= jtol)
    # a cell without precipitation and without any derivative must give exact zeros
    d = np.abs(ga - gf) * numc
    err = np.where(scale[:, None] > 0, d / np.maximum(scale, 1e-300)[:, None], np.where(d == 0, 0., np.inf))
    e = err[ok]
    print("  dJ/dN0: max |diff| x0 / cell scale %.2e (%d elements; %d beyond %.0e, %d beyond 1e-6)"
          % (e.max(), e.size, int((e > rel).sum()), rel, int((e > 1e-6).sum())))
    assert np.all(np.isfinite(ga[ok]))
    # rounding differences of the two trajectories, amplified where a cell
    # passes close to a threshold of the algorithm, may move a few elements
    # beyond the bar; every element stays within the finite-difference bar
    assert (e > rel).mean() <= 1e-3 and e.max() <= 1e-4
    ee = np.abs(jea - jef) / np.maximum(np.abs(jcf) + np.abs(jef), 1e-300)
    print("  dJ/dE per cell: max rel diff %.2e (%d cells beyond %.0e)" % (ee[ok].max(), int((ee[ok] > rel).sum()), rel))
    assert (ee[ok] > rel).mean() <= 1e-3 and ee[ok].max() <= 1e-4
    # dot products with random directions: u_b relative changes of N0_b and v of E
    rng = np.random.default_rng(11)
    for _ in range(3):
        u = rng.standard_normal(gf.shape) * numc
        v = rng.standard_normal(len(jef))
        df = (gf * u).sum(axis=1)[ok].sum() + (jef * v)[ok].sum()
        da = (ga * u).sum(axis=1)[ok].sum() + (jea * v)[ok].sum()
        s = (np.abs(gf * u).sum(axis=1) + np.abs(jef * v))[ok].sum()
        print("  dot product: fwd %.12e adj %.12e (rel to sum of |terms| %.2e)" % (df, da, abs(da - df) / s))
        assert abs(da - df) <= rel * s
    # the totals: J over every cell, dJ/dE over the cells the adjoint does not flag
    assert abs(Ja - jca.sum()) <= 1e-12 * abs(Ja)
    okA = (fla & capi.CELL_GRAD_UNSTABLE) == 0
    assert abs(dEa - jea[okA].sum()) <= 1e-12 * np.abs(jea[okA]).sum() + 1e-300


def test_adjoint_matches_forward_tangents(capi):
    """Workload-3 cells on a 20x8x8 grid, 10 steps: every bin's dJ/dN0 and
    dJ/dE per cell, both precision paths, four checkpoint intervals."""
    from salsa_amd import workload as W
    desc = capi.bin_desc()
    f = W.config_fields(desc, 3, grid_override=(20, 8, 8))
    cells = W.to_salsa_units(desc, f)
    nsteps = 10
    numc = cells[5]
    for prec in (capi.SALSA_FP64, capi.SALSA_FP32_MIXED):
        print("case adjoint vs forward tangents, precision %d, %d cells, %d steps" % (prec, len(numc), nsteps),
              flush=True)
        fwd = _flat_run(capi, desc, prec, cells, nsteps)
        assert fwd[2].max() > 0, "no precipitation formed"
        ref = None
        for every in (5, 1, 3, 10):
            adj = _flat_run(capi, desc, prec, cells, nsteps, every=every)
            assert not np.any(adj[5] & capi.CELL_NONFINITE)
            if ref is None:
                compare("precision %d, checkpoints every %d" % (prec, every), capi, fwd, adj, numc, prec)
                ref = adj
            else:
                # recomputation from checkpoints repeats the forward pass bit for bit
                for x, y in zip(adj[2:5], ref[2:5]):
                    assert np.array_equal(x, y), every
                assert adj[0] == ref[0] and adj[1] == ref[1]
                assert np.array_equal(adj[5], ref[5])


def test_adjoint_against_oracle_fd(capi, oracle_mod):
    """dJ/dN0 for every bin of every cell and dJ/dE against central
    differences of the FP64 oracle over a 10-step trajectory on workload-3
    cells, both precision paths, with the bars of the forward-mode test."""
    from salsa_amd import workload as W
    from test_gpu_gradient import OracleFD, check_against_oracle, small_cells
    desc = capi.bin_desc()
    cells = small_cells(desc, W, cfg=3)
    nsteps = 10
    print("case adjoint vs oracle FD, %d cells, %d steps" % (len(cells[0]), nsteps), flush=True)
    ref = OracleFD(oracle_mod, 4, 0, desc, cells, nsteps)
    assert ref.Jo.max() > 0
    for prec in (0, 1):
        J, dJdE, jc, je, g, flags = _flat_run(capi, desc, prec, cells, nsteps, every=5)
        assert not np.any(flags & capi.CELL_NONFINITE)
        check_against_oracle(ref, cells, prec, J, dJdE, jc, je, g, flags, capi.CELL_GRAD_UNSTABLE)


def test_adjoint_workload3(capi):
    """32x32x80 cells, 20 steps, checkpoints every 5 steps, both precision
    paths, against the forward tangents; the run times of both are printed."""
    import torch
    from salsa_amd import workload as W
    desc = capi.bin_desc()
    f = W.config_fields(desc, 3)
    g = f.grid
    nsteps = 20
    dev = torch.device("cuda")
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fd = {k: D(getattr(f, k)) for k in ("press", "temp", "rsl", "rsi", "dn", "rv")}
    n_p = [D(x) for x in f.n[:3]]
    m_p = [D(x) for x in f.m[:3]]
    idx = g.interior_index()
    numc = np.concatenate([f.n[c][:, idx] for c in range(3)], axis=0).T
    print("case workload 3 adjoint, %d cells, %d steps" % (len(idx), nsteps), flush=True)
    for prec in (capi.SALSA_FP64, capi.SALSA_FP32_MIXED):
        out = []
        for every in (None, 5):
            ctx = capi.Context(desc, capi.params(level=4, precision=prec), g.nzp, g.nxp, g.nyp)
            djdn = [torch.zeros_like(x) for x in n_p]
            jc = torch.zeros(len(idx), dtype=torch.float64, device=dev)
            je = torch.zeros_like(jc)
            for rep in range(2):          # the second call runs with its buffers in place
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                if every is None:
                    J, dJdE = ctx.gradient_device(fd, n_p, m_p, djdn, tstep=1.0, nsteps=nsteps, j_cell=jc,
                                                  djde_cell=je)
                else:
                    J, dJdE = ctx.adjoint_device(fd, n_p, m_p, djdn, tstep=1.0, nsteps=nsteps,
                                                 checkpoint_every=every, j_cell=jc, djde_cell=je)
                dt = time.perf_counter() - t0
            print("precision %d %s: %.3f s, %.3e cell-steps/s" % (
                prec, "forward tangents" if every is None else "adjoint", dt, len(idx) * nsteps / dt), flush=True)
            fl = ctx.cell_flags()
            ctx.close()
            gg = np.concatenate([x.cpu().numpy()[:, idx] for x in djdn], axis=0).T
            out.append((J, dJdE, jc.cpu().numpy(), je.cpu().numpy(), gg, fl))
        compare("workload 3 precision %d" % prec, capi, out[0], out[1], numc, prec)


def test_adjoint_autograd_function(capi):
    """salsa_amd.autograd.precip_volume: J.backward() fills n_p[c].grad and
    E.grad with the adjoint's dJ/dN0 and dJ/dE, scaled by the upstream
    gradient."""
    import torch
    from salsa_amd import workload as W
    from salsa_amd.autograd import precip_volume
    desc = capi.bin_desc()
    f = W.config_fields(desc, 3, grid_override=(12, 8, 8))
    g = f.grid
    dev = torch.device("cuda")
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fd = {k: D(getattr(f, k)) for k in ("press", "temp", "rsl", "rsi", "dn", "rv")}
    m_p = [D(x) for x in f.m[:3]]
    ctx = capi.Context(desc, capi.params(level=4, precision=capi.SALSA_FP64, lcharge=1, charge_coll_enh=0.25),
                       g.nzp, g.nxp, g.nyp)
    n_p = [D(x).requires_grad_(True) for x in f.n[:3]]
    E = torch.tensor(1.25, dtype=torch.float64, device=dev, requires_grad=True)
    J = precip_volume(ctx, fd, n_p, m_p, E, nsteps=6, checkpoint_every=2)
    (3.0 * J).backward()
    djdn = [torch.zeros_like(x) for x in n_p]
    J2, dJdE = ctx.adjoint_device(fd, [x.detach() for x in n_p], m_p, djdn, nsteps=6, checkpoint_every=2)
    ctx.close()
    assert float(J.detach()) == J2
    assert float(E.grad) == 3.0 * dJdE
    for a, b in zip(n_p, djdn):
        assert torch.equal(a.grad, 3.0 * b)
    assert any(float(b.abs().max()) > 0 for b in djdn)
    with pytest.raises(ValueError):
        precip_volume(ctx, fd, n_p, m_p, torch.tensor(1.0, dtype=torch.float64, device=dev), nsteps=1)


@pytest.mark.parametrize("kw", [dict(level=5), dict(lsauto_mode=2), dict(lsactiv=1), dict(lsdistupdate=1),
                                dict(every=0)])
def test_adjoint_refuses_uncovered_configurations(capi, kw):
    import torch
    from salsa_amd import workload as W
    kw = dict(kw)
    every = kw.pop("every", 5)
    level = kw.pop("level", 4)
    desc = capi.bin_desc(species=("SO4", "DU") if level == 5 else ("SO4",))
    f = W.config_fields(desc, 5 if level == 5 else 3, grid_override=(8, 8, 8))
    g = f.grid
    dev = torch.device("cuda")
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fd = {k: D(getattr(f, k)) for k in ("press", "temp", "rsl", "rsi", "dn", "rv")}
    ncat = 4 if level == 5 else 3
    n_p = [D(x) for x in f.n[:ncat]]
    m_p = [D(x) for x in f.m[:ncat]]
    ctx = capi.Context(desc, capi.params(level=level, **kw), g.nzp, g.nxp, g.nyp)
    djdn = [torch.zeros_like(x) for x in n_p]
    with pytest.raises(capi.SalsaError) as e:
        ctx.adjoint_device(fd, n_p, m_p, djdn, nsteps=2, checkpoint_every=every)
    ctx.close()
    assert e.value.code == capi.SALSA_ERR_ARG
