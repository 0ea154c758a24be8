"""Loader for the golden vectors written by tests/golden/make_golden.py."""
import glob
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(prefix):
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, prefix + "*.npz")))


def load(name):
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["W"] = int(d["W"])
    d["H"] = int(d["H"])
    if "P" in d:
        p = int(d["P"])
        d["P"] = None if p == -1000 else p
    return d
