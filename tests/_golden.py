"""Helpers to read the golden vectors under tests/golden (written once from
the reference by tests/golden/make_golden.py)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def meta():
    with open(os.path.join(GOLDEN, "golden_meta.json")) as f:
        return json.load(f)


def rel_l1(a, b):
    """Relative L1 error per variable (last axis) of a against b; variables
    that are identically zero in b report the absolute L1."""
    ax = tuple(range(a.ndim - 1))
    num = np.sum(np.abs(a - b), axis=ax)
    den = np.sum(np.abs(b), axis=ax)
    return np.where(den > 0, num / np.where(den > 0, den, 1), num)
