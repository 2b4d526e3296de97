"""Known-answer tests of the oracle's integer functions (random.glsl tea / lcg /
rnd, the PCG32 ray streams, Morton codes) against tests/golden/kat.json,
whose values come from independent Python implementations pinned to
published vectors (see tests/golden/make_golden.py)."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat.json")))


def test_golden_generator_matches_published_vectors():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    mg.check_published()
    assert mg.make() == KAT


def test_tea16(oracle_lib):
    for a, b, want in KAT["tea16"]:
        assert oracle_lib.tea(a, b) == want, (a, b)


def test_lcg_and_rnd(oracle_lib):
    for s, vals in KAT["lcg"]:
        assert oracle_lib.lcg_seq(s, len(vals)) == vals
    for s, vals in KAT["rnd"]:
        got = oracle_lib.rnd_seq(s, len(vals))
        assert [np.float32(v) for v in got] == [np.float32(v) for v in vals]


def test_pcg32(oracle_lib):
    for st, sq, vals in KAT["pcg32"]:
        assert list(oracle_lib.pcg32(st, sq, len(vals))) == vals, (st, sq)


def test_ray_stream_u01(oracle_lib):
    for seed, purpose, index, vals in KAT["u01"]:
        for k, v in enumerate(vals):
            assert np.float32(oracle_lib.lib().or_u01_stream(seed, purpose, index, k)) == np.float32(v)


def test_morton(oracle_lib):
    for x, y, z, want in KAT["morton3"]:
        assert oracle_lib.morton3(x, y, z) == want


@pytest.mark.parametrize("bad", [1, 2])
def test_kat_catches_planted_error(bad):
    """A plausible mistake in the routine (wrong shift, wrong key word) must
    change the known answers."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    key = list(mg.REF_KEY)
    if bad == 1:
        key[0], key[1] = key[1], key[0]
    else:
        key[3] ^= 1
    assert mg.tea_rounds(1, 0, tuple(key), 16)[0] != KAT["tea16"][1][2]
