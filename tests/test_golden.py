"""Golden fixtures (tests/golden, written by make_golden.py from the oracle):
the oracle must reproduce them bit for bit (CPU), and the HIP kernels must
match them within the parity tolerances (GPU)."""
import json
import os

import numpy as np
import pytest

import gargantua_amd as gg
from tests.parity import compare

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(HERE, "cases.json")))


def _load(name):
    d = np.load(os.path.join(HERE, f"{name}.npz"))
    return gg.Scene.from_dict(CASES[name]), {k: d[k] for k in d.files}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(oracle, name):
    scene, want = _load(name)
    got = oracle.render(scene, rgb=True)
    for k, v in want.items():
        np.testing.assert_array_equal(got[k], v, err_msg=f"{name}:{k}")


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_matches_golden(renderer, name, precision):
    scene, want = _load(name)
    got = renderer.render(scene, precision, rgb=True)
    c = compare(got, want)
    print(name, precision, c)
    assert c["class_agree"] == 1.0
    assert c["err_le_1e-6"] >= 0.99
    assert np.mean(np.abs(got["rgb"].astype(int) - want["rgb"].astype(int)) <= 1) >= 0.99
    disk = want["cls"] == gg.GG_CLASS_DISK
    if disk.any():
        rel = np.abs(got["intensity"][disk] - want["intensity"][disk]) / np.abs(want["intensity"][disk])
        assert np.mean(rel <= 1e-6) >= 0.99
