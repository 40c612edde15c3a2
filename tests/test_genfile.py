"""Generator interchange files (include/hss_hip.h, hss_gen_header): the library's host-only
reader and writer (hss_gen_read / hss_gen_write) against the numpy ones
(hssmatrix.save_generators / load_generators), round trips on every golden fixture,
and malformed files.  No device is used."""
import pathlib

import numpy as np
import pytest

from conftest import H, rect_hss

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLD = np.load(pathlib.Path(__file__).with_name("golden") / "hss_golden.npz")
NAMES = [str(s) for s in GOLD["names"]]


def fixture(name):
    return GOLD[name + "/nodes"], GOLD[name + "/blob"]


def test_header_layout():
    from hssmatrices_jl_amd.hssmatrix import GEN_HEADER_DTYPE
    assert GEN_HEADER_DTYPE.itemsize == 40
    assert "} hss_gen_header; /* 40 bytes */" in (ROOT / "include" / "hss_hip.h").read_text()


@pytest.mark.parametrize("name", NAMES)
def test_roundtrip_both_ways(tmp_path, hip, name):
    nodes, blob = fixture(name)
    p_py, p_c = tmp_path / "py.hssg", tmp_path / "c.hssg"
    H.save_generators(p_py, nodes, blob)
    hip.gen_write(p_c, nodes, blob)
    # the two writers agree byte for byte
    assert p_py.read_bytes() == p_c.read_bytes()
    for reader in (H.load_generators, hip.gen_read):
        for p in (p_py, p_c):
            n2, b2 = reader(p)
            assert n2.tobytes() == np.ascontiguousarray(nodes).tobytes()
            assert np.array_equal(b2, blob)


@pytest.mark.parametrize("name", NAMES)
def test_loaded_generators_reproduce_the_products(tmp_path, oracle, name):
    """Everything a product needs travels in the file: the oracle's matvec on the loaded
    generators reproduces the extended-precision fixture Y = A X."""
    nodes, blob = fixture(name)
    p = tmp_path / "a.hssg"
    H.save_generators(p, nodes, blob)
    n2, b2 = H.load_generators(p)
    X = GOLD[name + "/X"]
    Y = oracle.matmul(n2, b2, X)
    assert np.linalg.norm(Y - GOLD[name + "/Y"]) <= 1e-14 * np.linalg.norm(GOLD[name + "/Y"])


def test_rectangular_and_host_tree_roundtrip(tmp_path, hip):
    rng = np.random.default_rng(3)
    h = rect_hss(rng, [7, 12, 9, 15], [10, 6, 14, 8], 3, 5)
    nodes, blob = H.pack(h)
    p = tmp_path / "r.hssg"
    hip.gen_write(p, nodes, blob)
    n2, b2 = hip.gen_read(p)
    assert np.array_equal(H.full(H.unpack(n2, b2)), H.full(h))


def _corrupt(tmp_path, data: bytes, name="bad.hssg"):
    p = tmp_path / name
    p.write_bytes(data)
    return p


def test_malformed_files_are_refused(tmp_path, hip):
    nodes, blob = fixture("uneven_n97_leaf10")
    good = tmp_path / "good.hssg"
    H.save_generators(good, nodes, blob)
    raw = good.read_bytes()
    cases = {
        "magic": b"HSSX" + raw[4:],
        "version": raw[:4] + (2).to_bytes(4, "little") + raw[8:],
        "elem": raw[:8] + (4).to_bytes(4, "little") + raw[12:],
        "truncated": raw[:-8],
        "trailing": raw + b"\0" * 8,
        "header_only": raw[:40],
        "short_header": raw[:20],
    }
    for what, data in cases.items():
        p = _corrupt(tmp_path, data, what + ".hssg")
        with pytest.raises(ValueError):
            hip.gen_read(p)
        with pytest.raises(ValueError):
            H.load_generators(p)
    # a well-formed file whose table hss_mat_create would refuse (children's rows do not add up)
    bad = nodes.copy()
    bad["m"][0] += 1
    p = tmp_path / "table.hssg"
    H.save_generators(p, bad, blob)
    with pytest.raises(ValueError):
        hip.gen_read(p)
    # and the library's writer refuses to write it
    with pytest.raises((ValueError, H.DimensionMismatch)):
        hip.gen_write(tmp_path / "w.hssg", bad, blob)
    with pytest.raises(OSError):
        hip.gen_read(tmp_path / "missing.hssg")
