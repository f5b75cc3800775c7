"""Pins the oracle's Morton encode/decode (morton3D.h:36-47,163-178).

Known answers: samples of the reference's own lookup tables (tests/golden/
morton_lut_kat.json, written by tests/golden/make_morton_kat.py). Closed form: the bit
interleave x -> bit 3i, y -> 3i+1, z -> 3i+2, checked by brute force.
"""
import json
import os

import numpy as np

import oracle

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "morton_lut_kat.json")))


def interleave(x, y, z):
    m = 0
    for i in range(21):
        m |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
    return m


def test_encode_matches_reference_luts():
    t = KAT["tables"]
    for axis, name in enumerate(["Morton3D_encode_x_256", "Morton3D_encode_y_256",
                                 "Morton3D_encode_z_256"]):
        assert t[name]["size"] == 256
        for idx, val in t[name]["samples"]:
            c = [0, 0, 0]
            c[axis] = idx
            got = int(oracle.morton_encode([c[0]], [c[1]], [c[2]])[0])
            assert got == val, (name, idx)


def test_decode_matches_reference_luts():
    t = KAT["tables"]
    for axis, name in enumerate(["Morton3D_decode_x_512", "Morton3D_decode_y_512",
                                 "Morton3D_decode_z_512"]):
        assert t[name]["size"] == 512
        for idx, val in t[name]["samples"]:
            xyz = oracle.morton_decode([idx])
            assert int(xyz[axis][0]) == val, (name, idx)


def test_encode_brute_force():
    rng = np.random.default_rng(0)
    x, y, z = (rng.integers(0, 1 << 21, 3000, dtype=np.uint64) for _ in range(3))
    got = oracle.morton_encode(x, y, z)
    want = [interleave(int(a), int(b), int(c)) for a, b, c in zip(x, y, z)]
    assert [int(g) for g in got] == want


def test_decode_inverts_encode():
    rng = np.random.default_rng(1)
    x, y, z = (rng.integers(0, 1 << 21, 5000, dtype=np.uint64) for _ in range(3))
    dx, dy, dz = oracle.morton_decode(oracle.morton_encode(x, y, z))
    assert np.array_equal(dx, x.astype(np.uint32))
    assert np.array_equal(dy, y.astype(np.uint32))
    assert np.array_equal(dz, z.astype(np.uint32))


def test_exhaustive_small_grid_order():
    """On an 8^3 grid, codes are a permutation of 0..511 with Morton order."""
    g = np.arange(8, dtype=np.uint64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    codes = oracle.morton_encode(X.ravel(), Y.ravel(), Z.ravel())
    assert sorted(int(c) for c in codes) == list(range(512))
    assert int(oracle.morton_encode([7], [7], [7])[0]) == 511
    assert int(oracle.morton_encode([1], [0], [0])[0]) == 1
    assert int(oracle.morton_encode([0], [1], [0])[0]) == 2
    assert int(oracle.morton_encode([0], [0], [1])[0]) == 4
