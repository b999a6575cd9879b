"""Readers for the golden vectors in tests/golden/ (written by make_golden.py from the reference)."""
import base64
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def hexv(lst):
    return np.array([float.fromhex(s) for s in lst])


def b64(s):
    return np.frombuffer(base64.b64decode(s), dtype="<f8").copy()


def cfg_of(d):
    """JSON lists back to the tuple shapes mcitest.chain_setup expects."""
    c = dict(d)
    if "move" in c:
        c["move"] = tuple(c["move"])
    if "irange" in c:
        c["irange"] = tuple(c["irange"])
    c["pdfs"] = [tuple(p) for p in c.get("pdfs", [])]
    c["obs"] = [tuple(o) for o in c.get("obs", [])]
    return c
