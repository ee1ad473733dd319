import glob
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


def golden_snapshots(max_n=None, pattern="*_s*.npz", exclude_big=True):
    out = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, pattern))):
        if "_expect_" in p or os.path.basename(p).startswith(("ed_", "qn_")):
            continue
        # the production-size superblocks (n from 1.4e5 to 2e6) have tests of their own
        if exclude_big and ("L64" in p or "_m800_" in p):
            continue
        out.append(p)
    return out


def golden_ids(paths):
    return [os.path.basename(p)[:-4] for p in paths]
