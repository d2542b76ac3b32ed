"""Shared helpers for the parity tests."""
import numpy as np


def rel_l2(a, b):
    """relative L2 error of a against b (absolute when b is zero)"""
    nb = float(np.linalg.norm(b))
    d = float(np.linalg.norm(np.asarray(a) - np.asarray(b)))
    return d / nb if nb > 0 else d


def field_errors(a, b):
    return [rel_l2(a[m], b[m]) for m in range(a.shape[0])]
