"""Test helper: profile sets written in the COMER text format ("COMER
profile v2.0", the condensed layout of TextWriteProfileCondensed,
libpro/srcpro/PMProfileModel.cpp:2295-2510), so that the library's reader
can be checked on files of known content.

Values are written as the reference writes them: probabilities times the
scale, rounded (SS probabilities truncated); transitions as probabilities
(our sets hold their logs: MM, MI, MD as log p, IM, II, DM, DD as 0.4 log p;
ID and DI, which the search path does not use, are written as 0); observed
counts, HDP posteriors and the statistical parameters as placeholders.
"""
import numpy as np

ALPHA = "ARNDCQEGHILKMFPSTWYV"


def _probs_from_logs(t7):
    """the 9 transition probabilities of a row of 7 stored logs"""
    t7 = np.asarray(t7, np.float64)
    p = np.where(t7 <= -32768.0, 0.0, np.exp(np.where(np.arange(7) >= 3, t7 / 0.4, t7)))
    mm, mi, md, im, ii, dm, dd = p
    return [mm, mi, md, im, ii, 0.0, dm, 0.0, dd]


def write_profile(fh, view, p, desc, name, scale=10000, with_cv=None, with_ss=True):
    L = int(view.len[p])
    d0 = int(view.dist[p])
    ns, npro = int(view.npos), int(view.nprof)
    ntr = ns + npro
    with_cv = bool(view.cv) if with_cv is None else with_cv
    w = fh.write
    w("COMER profile v2.0\n")
    w("DESC: %s\n" % desc)
    w("FILE: %s\n" % name)
    w("CMD: \n")
    w("LEN: %d\n" % L)
    w("NOS: %d\n" % 100)
    w("EFFNOS: %.1f\n" % float(view.eno[p]))
    w("SCALE: %d\n" % scale)
    w("".join(" %c" % c for c in ALPHA))
    w("\nNULL" + "".join(" %d" % int(np.rint(scale * float(view.bkg[a * npro + p]))) for a in range(20)))
    w("\nPOST" + " 0" * 20)

    def trans_row(row):
        t7 = [view.trans[k * ntr + d0 + p + row] for k in range(7)]
        return "".join(" %d" % int(np.rint(scale * x)) for x in _probs_from_logs(t7))

    w("\n" + trans_row(0))
    w("\n 1000 0 0")
    for m in range(L):
        pos = d0 + m
        res = view.aa[pos]
        res = res.decode() if isinstance(res, bytes) else chr(res) if isinstance(res, int) else res
        w("\n%d %s" % (m + 1, res))
        w("".join(" %d" % int(np.rint(scale * float(view.trgfrq[a * ns + pos]))) for a in range(20)))
        w("\n" + trans_row(m + 1))
        w("\n 1000 0 0")
        w("\n %d %d" % (int(np.rint(scale * 0.05)), 1))
        w("\n %d" % int(np.rint(scale * 0.9)))
        w("\n %d" % (m % 7))
        if with_cv:
            w("\n CV: %d %d %d" % (int(np.rint(scale * float(view.cvprior[pos]))),
                                   int(np.rint(scale * float(view.cvnorm2[pos]))), 19))
            w("".join(" %d" % int(np.rint(scale * float(view.cv[i * ns + pos]))) for i in range(19)))
        if with_ss:
            pr = [float(view.ssprb[j * ns + pos]) for j in range(3)]
            w("\n SS:%s" % "CEH"[int(np.argmax(pr))])
            w("".join(" %d" % int(scale * x) for x in pr))
    w("\n")
    w("K, Lambda:\n")
    w("Computed  ungapped, 0.1000 0.3000\n")
    w("Reference ungapped, 0.1000 0.3000\n")
    w("Entropy, 0.5000; Expected, -1.0000\n")
    w("*\n")


def write_set(path, view, profiles=None, scale=10000, with_cv=None, with_ss=True):
    profiles = range(int(view.nprof)) if profiles is None else profiles
    with open(path, "w") as fh:
        for p in profiles:
            write_profile(fh, view, p, "synthetic profile %d" % p, "pro%04d.pro" % p, scale, with_cv, with_ss)
