"""Shared test helpers: move an oracle two-site step into a chemps2_amd context."""
import numpy as np

from plancheck import oracle_tensors


def load_site(ctx, o, site, econst=0.0):
    """Problem, sectors of every boundary and the operators Heff reads at `site`."""
    ctx.set_problem(o.L, int(o.group), o.irreps(), o.N, o.TwoS, o.Irrep, o.mx(), econst)
    for b in range(o.L + 1):
        ctx.set_sectors(b, o.sectors(b))
    for cnt, mr in ((site - 1, 1), (site + 1, 0)):
        if 0 <= cnt <= o.L - 2:
            for kind, k, m, meta, data in oracle_tensors(o, cnt):
                ctx.upload(cnt, mr, kind, k, m, int(meta[1]), int(meta[2]), int(meta[3]), data)
    return ctx.prepare(site)


def symm_to_prog(o, x):
    """Divide block kappa by sqrt(2SR+1) (Sobject.cpp:638-650)."""
    lab, k2i = o.kappa()
    y = x.copy()
    for k in range(len(lab)):
        y[k2i[k]:k2i[k + 1]] /= np.sqrt(lab[k, 7] + 1.0)
    return y
