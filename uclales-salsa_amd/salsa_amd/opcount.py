"""Algorithmic work of one SALSA coagulation + condensation cell-step.

The contract figures used for roofline numbers (not re-derived from the
implementation):

  flops  F = 100 P + 60 B + 2 P (1 + n_s) + 10 B + 45 B_c + S (16 B_c + 4 B_c) + F_eq
         P   unique kernel pairs, B bins, n_s wet species, B_c bins in the
             condensation loop (2a..2b + cloud + precip + ice),
         S   substeps (50 at dt = 1 s), F_eq = 40 x 3 bins x 3 iterations
  bytes  read 77 FP64 (24 N + 48 m + 5 thermo at the 10/7/7 layout) and
         read-modify-write the tendencies, = 1,784 B/cell-step for the
         host-layout drop-in (config 2); generalised below to any layout.
"""


def layout_counts(desc, level=4):
    nb, nc, npr = desc.nbins, desc.ncld, desc.nprc
    ni = desc.nice if level == 5 else 0
    B = nb + nc + npr + ni
    # unique pairs of the enabled kernel blocks (aa, cc, pp, ii self; ca, pa, ia, pc, ic, ip cross)
    self_pairs = sum(n * (n + 1) // 2 for n in (nb, nc, npr, ni))
    cross = nb * nc + nb * npr + nb * ni + nc * npr + nc * ni + npr * ni
    P = self_pairs + cross
    ns = desc.nspec - 1 if level == 4 else desc.nspec - 1
    Bc = (desc.fn2b - desc.in2a + 1) + nc + npr + ni
    return P, B, ns, Bc


def flops_per_cell_step(desc, level=4, substeps=50, f_eq=40 * 3 * 3):
    P, B, ns, Bc = layout_counts(desc, level)
    return 100 * P + 60 * B + 2 * P * (1 + ns) + 10 * B + 45 * Bc + substeps * (16 * Bc + 4 * Bc) + f_eq


def bytes_per_cell_step(desc, level=4, tendency=True):
    """HBM bytes of the FP64 LES-layout boundary: every N and m field plus the
    5 thermodynamic fields (+ rsi at level 5) read once; with tendency=True the
    N, m and rv tendencies are read and written (+=), otherwise the new state
    is written."""
    ncat = 4 if level == 5 else 3
    nbs = [desc.nbins, desc.ncld, desc.nprc, desc.nice][:ncat]
    nwet = desc.nspec - 1
    fields = sum(nb * (1 + (nwet if c < 3 else desc.nspec)) for c, nb in enumerate(nbs))
    thermo = 5 + (1 if level == 5 else 0)
    if tendency:
        return 8 * (fields + thermo) + 16 * (fields + 1)
    return 8 * (fields + thermo) + 8 * (fields + 1)
