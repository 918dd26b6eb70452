"""CPU oracle of the per-aggregate features read by the parameter network
(TEST INFRASTRUCTURE ONLY).  Features (float64) of each aggregate c of a level
with matrix A (scipy CSR) and aggregate map agg, averaged over the members i of
c, with a_ii the first stored diagonal entry, s = sum_{j!=i}|a_ij|,
q = sum_{j!=i} a_ij, mx / mn the largest / smallest nonzero |a_ij| (j != i),
pos = sum of positive off-diagonal entries, in = sum_{j!=i, agg j = agg i}|a_ij|:
 0 log2(size)  1 mean s/a_ii  2 mean q/a_ii  3 mean mx/s  4 mean mn/mx  5 mean in/s
 6 mean pos/s  7 mean row length/16  8 level/10  9 std of log10|a_ii|
 10 max mx/s  11 min s/a_ii   (ratios with a zero denominator count as 0).
"""
from __future__ import annotations

import numpy as np


def aggregate_features(ia, ja, va, agg, nagg, level):
    n = ia.size - 1
    rows = np.zeros((n, 8))
    for i in range(n):
        di, s, q, mx, mn, pos, inn = 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0
        found = False
        for k in range(ia[i], ia[i + 1]):
            j, v = ja[k], va[k]
            if j == i:
                if not found:
                    di, found = v, True
                continue
            av = abs(v)
            s += av
            q += v
            if v > 0:
                pos += v
            if av > 0:
                mx = max(mx, av)
                mn = av if mn == 0 else min(mn, av)
            if agg[i] > -1 and agg[j] == agg[i]:
                inn += av
        ad = abs(di)
        rows[i] = [s / ad if ad > 0 else 0, q / ad if ad > 0 else 0, mx / s if s > 0 else 0,
                   mn / mx if mx > 0 else 0, inn / s if s > 0 else 0, pos / s if s > 0 else 0,
                   (ia[i + 1] - ia[i]) / 16.0, np.log10(ad) if ad > 0 else 0]
    F = np.zeros((nagg, 12))
    for c in range(nagg):
        mem = np.nonzero(agg == c)[0]
        m = mem.size
        if m == 0:
            continue
        R = rows[mem]
        F[c, 0] = np.log2(m)
        F[c, 1:8] = R[:, :7].mean(0)
        F[c, 8] = level / 10.0
        mu = R[:, 7].mean()
        F[c, 9] = np.sqrt(max((R[:, 7] ** 2).mean() - mu * mu, 0.0))
        F[c, 10] = R[:, 2].max()
        F[c, 11] = R[:, 0].min()
    return F
