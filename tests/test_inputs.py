"""The synthetic benchmark inputs match their recorded SHA-256
(tests/golden/inputs.sha256), so the generators are pinned."""
import os

import pytest

from tests.golden import make_input_hashes as mih

HERE = os.path.dirname(os.path.abspath(__file__))


def _recorded():
    out = {}
    for line in open(os.path.join(HERE, "golden", "inputs.sha256")):
        if line.strip():
            h, k = line.split()
            out[k] = h
    return out


def test_generator_hashes():
    rec = _recorded()
    got = mih.entries(with_jgrid=False)
    for k, v in got.items():
        assert rec[k] == v, k


@pytest.mark.slow
def test_jgrid512_hash():
    rec = _recorded()
    p, w, _, _ = mih.workloads.standard("jgrid512")
    assert mih.sha(p, w) == rec["jgrid512"]
