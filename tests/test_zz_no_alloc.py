"""salsa_gradient_device_noalloc (no device allocation; static work space)
against salsa_gradient_device through the C ABI: two calls in a row, the
second on a smaller grid, each bitwise equal to the allocating entry point on
the same inputs (same kernels, fixed-order sums); and a context above
SALSA_NOALLOC_MAX_CELLS is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    from salsa_amd import capi as c
    c.lib()
    return c


def _case(capi, grid):
    import torch
    from salsa_amd import workload as W
    desc = capi.bin_desc()
    f = W.config_fields(desc, 3, grid_override=grid)
    dev = torch.device("cuda")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fd = dict(press=T(f.press), temp=T(f.temp), rsl=T(f.rsl), rsi=T(f.rsi), dn=T(f.dn), rv=T(f.rv))
    return desc, f, fd, [T(x) for x in f.n[:3]], [T(x) for x in f.m[:3]]


def _run(capi, case, prec, noalloc):
    import torch
    desc, f, fd, n_p, m_p = case
    g = f.grid
    ctx = capi.Context(desc, capi.params(level=4, precision=prec), g.nzp, g.nxp, g.nyp)
    djdn = [torch.zeros_like(x) for x in n_p]
    jc = torch.zeros(g.ncells, dtype=torch.float64, device=n_p[0].device)
    dc = torch.zeros_like(jc)
    J, dJdE = ctx.gradient_device(fd, n_p, m_p, djdn, nsteps=3, j_cell=jc, djde_cell=dc, noalloc=noalloc)
    flags = ctx.cell_flags()
    ctx.close()
    return J, dJdE, [x.cpu().numpy() for x in djdn], jc.cpu().numpy(), dc.cpu().numpy(), flags


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_gradient_matches_allocating_entry(capi, prec):
    big = _case(capi, (20, 8, 8))
    small = _case(capi, (12, 4, 4))
    print("no-alloc gradient, precision %d: 1280 then 192 cells" % prec, flush=True)
    got = [_run(capi, big, prec, True), _run(capi, small, prec, True)]
    ref = [_run(capi, big, prec, False), _run(capi, small, prec, False)]
    for k, (a, b) in enumerate(zip(got, ref)):
        assert a[0] == b[0] and a[1] == b[1], (k, a[:2], b[:2])
        for x, y in zip(a[2], b[2]):
            assert np.array_equal(x, y), k
        assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4]), k
        assert np.array_equal(a[5], b[5]), k
        assert a[0] > 0. and np.count_nonzero(a[3]) > 0


def test_noalloc_gradient_refuses_oversize_grid(capi):
    import torch
    desc = capi.bin_desc()
    # 1032 x 32 x 32 interior cells > 2^20
    g = (1034, 36, 36)
    ctx = capi.Context(desc, capi.params(level=4, precision=0), *g)
    assert ctx.ncells > capi.NOALLOC_MAX_CELLS
    # the size check comes before any field is read: one-element stand-ins
    x = lambda: torch.zeros(1, dtype=torch.float64, device="cuda")
    fd = dict(press=x(), temp=x(), rsl=x(), rsi=x(), dn=x(), rv=x())
    with pytest.raises(capi.SalsaError) as ei:
        ctx.gradient_device(fd, [x(), x(), x()], [x(), x(), x()], [x(), x(), x()], noalloc=True)
    ctx.close()
    assert ei.value.code == capi.SALSA_ERR_ARG
    assert "SALSA_NOALLOC_MAX_CELLS" in str(ei.value)
