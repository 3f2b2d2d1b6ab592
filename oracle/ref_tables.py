"""The reference's own module arrays (module/matmod.f90 layout, 1-based values)
built from an oracle model, as the Fortran preparation phase hands them to
the drop-in loader (heomq_model_from_reference) -- TEST INFRASTRUCTURE ONLY.
Also the input of the C port of the apply (oracle/heom_apply_port.c)."""
import numpy as np


def reference_arrays(o):
    M, L, d = o.M, o.L, o.d
    ntier = L + 1
    nfirst = np.array([o.first[n] + 1 for n in range(ntier)], np.int64)
    nlast = np.array([o.first[n + 1] for n in range(ntier)], np.int64)
    # indextable(1) = 0 is tier 0's empty record (main/ball2drawer.f90:40-44);
    # tier 1 starts at position 2
    ifirst, table = [1], [0]
    pos = 2
    for n in range(1, ntier):
        ifirst.append(pos)
        flat = (o.tiers[n] + 1).ravel()
        table.extend(flat.tolist())
        pos += flat.size
    indexcoef, itype, ff, fr, ref = [], [], [], [], np.zeros(o.N, np.int64)
    for n in range(1, L + 1):
        lo, lof = o.lower[n]
        up = o.upper[n] if n < L else None
        for a in range(len(o.tiers[n])):
            g = o.first[n] + a
            ref[g] = len(indexcoef) + 1
            for r in range(n):
                indexcoef.append(lo[a, r, 0] + 1); itype.append(lo[a, r, 1])
                ff.append(lof[a, r, 0]); fr.append(lof[a, r, 1])
            if up is not None:
                for e in range(M):
                    indexcoef.append(up[0][a, e, 0] + 1); itype.append(max(up[0][a, e, 1], 0))
                    ff.append(up[1][a, e, 0]); fr.append(up[1][a, e, 1])
    half = M // 2
    indexdraw = np.zeros(half, np.int32)
    for k in range(half):
        nb, it, fact = o.lookup((k,))
        assert it == 0
        indexdraw[k] = nb - o.first[1] + 1
    norbs, nspin, ncor, nalf = o.norbs, 2, o.ncor, o.nalf
    cb = np.zeros((2, nalf, ncor, nspin, norbs), complex)
    cd = np.zeros_like(cb)
    dbs = np.zeros((2, nalf, ncor, nspin, norbs))
    for k in range(M):
        cb[o.sgn[k], o.alf[k], o.cor[k], o.spin[k], o.orb[k]] = o.cb[k]
        cd[o.sgn[k], o.alf[k], o.cor[k], o.spin[k], o.orb[k]] = o.cd[k]
        dbs[o.sgn[k], o.alf[k], o.cor[k], o.spin[k], o.orb[k]] = o.dbsqrt[k]
    ams = np.zeros((nspin, norbs, d, d))
    for m in range(norbs):
        for s in range(nspin):
            ams[s, m] = o.ams[m][s].T          # C order [s][m][j][i] == Fortran (i,j,m,s)
    return dict(
        nrho=d, nunk=o.N, ntier=ntier, norbs=norbs, nspin=nspin, nalf=nalf, ncor=ncor,
        igroundsteady=1, indextable=np.array(table, np.int32), nfirst=nfirst, nlast=nlast,
        ifirst=np.array(ifirst, np.int64), indexcoef=np.array(indexcoef, np.int64),
        itypecoef=np.array(itype, np.int8), ifactfcoef=np.array(ff, np.int8),
        ifactrcoef=np.array(fr, np.int8), index_coef_ref=ref, indexdraw=indexdraw,
        cb=cb.ravel(), cd=cd.ravel(), cgamma=o.cgamma, dipm=o.dipm,
        ilead=(o.alf + 1).astype(np.int32), mspin=(o.spin + 1).astype(np.int32),
        morbs=(o.orb + 1).astype(np.int32), mpm=(o.sgn + 1).astype(np.int32),
        eleadinfty=o.eleadinfty.T.ravel(), hs=o.H.T.ravel(), amsall=ams.ravel(),
        dbsqrt=dbs.ravel() if o.p.lscale else None)
