"""Loaders for the golden vectors in tests/golden (written by make_golden.py
from the compiled reference)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def tiny(dtype=np.float64):
    suffix = "d" if np.dtype(dtype) == np.float64 else "f"
    meta = json.load(open(os.path.join(GOLDEN, f"tiny_{suffix}.json")))
    data = np.load(os.path.join(GOLDEN, f"tiny_{suffix}.npz"))
    return meta["cases"], data


def config0():
    meta = json.load(open(os.path.join(GOLDEN, "config0.json")))
    samples = np.load(os.path.join(GOLDEN, "config0_samples_d.npz"))
    return meta, samples


def materials():
    return json.load(open(os.path.join(GOLDEN, "materials.json")))


def balls():
    return json.load(open(os.path.join(GOLDEN, "ball.json")))


def cli():
    return json.load(open(os.path.join(GOLDEN, "cli.json")))


def fromhex(x):
    return float.fromhex(x)


def sha256(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def same_bits(a, b):
    """Equal shapes, dtypes and bit patterns, signed zeros included
    (np.array_equal treats -0 and +0 as equal)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    u = {8: np.uint64, 4: np.uint32}[a.dtype.itemsize]
    return bool(np.array_equal(a.view(u), b.view(u)))
