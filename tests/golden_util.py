"""Load the reference-written fixtures of tests/golden/*.npz as
{case: {field: array}}."""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(group):
    z = np.load(os.path.join(HERE, f"{group}.npz"), allow_pickle=False)
    out = {}
    for key in z.files:
        case, field = key.split("/", 1)
        out.setdefault(case, {})[field] = z[key]
    return out


def spacing(case):
    e, n, ndim = case["ext"], case["n"], int(case["ndim"])
    return tuple((e[2 * a + 1] - e[2 * a]) / n[a] if a < ndim else 1.0 for a in range(3))
