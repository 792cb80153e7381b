"""Minimal FCIDUMP reader for the oracle tests (test infrastructure only).

Follows the reference reader's semantics (squantop.cpp:79-161): NORB from the
header; a line `x i j k l` with all indices nonzero stores the chemist integral
(ij|kl) at physicist position two_mo[i,k,j,l] together with its 7 symmetric
copies; `x i j 0 0` fills one_mo[i,j] and one_mo[j,i]; `x 0 0 0 0` is ecore.
"""
import re

import numpy as np


def read(path):
    with open(path) as f:
        text = f.read()
    m = re.search(r"(&END|/)", text)
    header, body = text[: m.start()], text[m.end():]
    norb = int(re.search(r"NORB\s*=\s*(\d+)", header).group(1))
    n = norb
    one = np.zeros((n, n))
    two = np.zeros((n, n, n, n))
    ecore = 0.0
    vals = body.split()
    for t in range(0, len(vals) - 4, 5):
        x = float(vals[t])
        i, j, k, l = (int(v) for v in vals[t + 1:t + 5])
        if i and j and k and l:
            i, j, k, l = i - 1, j - 1, k - 1, l - 1
            for (a, b, c, d) in ((i, k, j, l), (k, i, l, j), (j, k, i, l), (i, l, j, k),
                                 (j, l, i, k), (l, j, k, i), (k, j, l, i), (l, i, k, j)):
                two[a, b, c, d] = x
        elif i and j:
            one[i - 1, j - 1] = x
            one[j - 1, i - 1] = x
        else:
            ecore = x
    return ecore, one, two
