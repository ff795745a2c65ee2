"""Fits the table of c2_lse1p (comer2_amd/csrc/c2_fmath.h): log(1 + e^d) for
d in (-16, 0] as 64 degree-4 polynomials in u = d - (-16 + i/4) on intervals
of width 1/4, least squares on 200 Chebyshev nodes per interval in double,
coefficients rounded to float.  Prints the C initializer (hex floats, 8 per
interval: c0..c4 and three zeros, so an interval is 32 bytes) and the
largest error of the FP32 evaluation against log1p(exp(d)) in double.

    python comer2_amd/fit_lse_table.py
"""
import numpy as np

NINT, LO, W, DEG = 64, -16.0, 0.25, 4


def fit():
    k = np.arange(200)
    x = np.cos(np.pi * (k + 0.5) / 200)
    u = (x + 1) / 2 * W
    V = np.vander(u, DEG + 1, increasing=True)
    out = []
    for i in range(NINT):
        y = np.log1p(np.exp(LO + W * i + u))
        c, *_ = np.linalg.lstsq(V, y, rcond=None)
        out.append(c.astype(np.float32))
    return np.array(out)


def evaluate(C, d):
    """the FP32 operation sequence of c2_lse1p (fma emulated in double)"""
    d = d.astype(np.float32)
    t = ((d + np.float32(16.0)) * np.float32(4.0)).astype(np.float32)
    i = np.minimum(t.astype(np.int64), NINT - 1)
    c = (np.float32(0.25) * i.astype(np.float32) + np.float32(-16.0)).astype(np.float32)
    u = (d - c).astype(np.float32)
    p = C[i, DEG]
    for k in range(DEG - 1, -1, -1):
        p = (p.astype(np.float64) * u.astype(np.float64) + C[i, k].astype(np.float64)).astype(np.float32)
    return np.where(d > np.float32(-16.0), p, np.float32(0.0))


def main():
    C = fit()
    d = np.linspace(-16, 0, 4000001).astype(np.float32)
    err = np.abs(evaluate(C, d).astype(np.float64) - np.log1p(np.exp(d.astype(np.float64))))
    print("/* max abs error %.3g */" % err.max())
    rows = []
    for i in range(NINT):
        vals = [float(v).hex() + "f" for v in C[i]] + ["0x0p+0f"] * 3
        rows.append("    " + ", ".join(vals) + ",")
    print("#define C2_LSE1P_TABLE { \\")
    print(" \\\n".join(rows) + " \\")
    print("}")


if __name__ == "__main__":
    main()
