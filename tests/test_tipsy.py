"""Tipsy files and ChaNGa's ASCII array outputs (changa_amd/tipsy.py):
the reader against bytes of the reference's own teststep/king_soft.bin
(the fixture king_soft_head.bin, decoded independently into
king_soft.npz), round trips in both byte orders, and the array layout
(count, then every x, every y, every z)."""
import os

import numpy as np
import pytest

from changa_amd import tipsy

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_read_king_soft_head():
    t = tipsy.read(os.path.join(GOLDEN, "king_soft_head.bin"))
    assert t["big_endian"] and len(t["gas"]) == 0 and len(t["star"]) == 0
    d = t["dark"]
    ref = np.load(os.path.join(GOLDEN, "king_soft.npz"))
    assert len(d) == 64
    np.testing.assert_array_equal(d["mass"], ref["mass"][:64])
    np.testing.assert_array_equal(d["pos"], ref["pos"][:64])
    np.testing.assert_array_equal(d["vel"], ref["vel"][:64])
    np.testing.assert_array_equal(d["eps"], ref["eps"][:64])
    pos, mass, soft = tipsy.particles(t)
    assert pos.dtype == np.float64 and pos.shape == (64, 3)
    np.testing.assert_array_equal(soft, ref["eps"][:64].astype(np.float64))


@pytest.mark.parametrize("big_endian", [True, False])
def test_round_trip(tmp_path, big_endian):
    rng = np.random.default_rng(3)
    gas = {"mass": rng.uniform(size=5), "pos": rng.normal(size=(5, 3)), "hsmooth": np.full(5, 0.1),
           "rho": rng.uniform(size=5)}
    dark = {"mass": rng.uniform(size=7), "pos": rng.normal(size=(7, 3)),
            "vel": rng.normal(size=(7, 3)), "eps": np.full(7, 0.02)}
    star = {"mass": rng.uniform(size=3), "pos": rng.normal(size=(3, 3)), "eps": np.full(3, 0.03),
            "tform": np.arange(3.0)}
    path = tmp_path / "t.bin"
    tipsy.write(path, 0.25, gas=gas, dark=dark, star=star, big_endian=big_endian)
    assert os.path.getsize(path) == 32 + 5 * 48 + 7 * 36 + 3 * 44
    t = tipsy.read(path)
    assert t["big_endian"] == big_endian and t["time"] == 0.25
    np.testing.assert_array_equal(t["gas"]["pos"], gas["pos"].astype(np.float32))
    np.testing.assert_array_equal(t["dark"]["vel"], dark["vel"].astype(np.float32))
    np.testing.assert_array_equal(t["star"]["tform"], star["tform"].astype(np.float32))
    pos, mass, soft = tipsy.particles(t)
    assert pos.shape == (15, 3)
    np.testing.assert_array_equal(soft[:5], np.float32(0.1))
    np.testing.assert_array_equal(soft[5:12], np.float32(0.02))


def test_not_tipsy(tmp_path):
    path = tmp_path / "x.bin"
    path.write_bytes(bytes(50))
    with pytest.raises(ValueError):
        tipsy.read(path)


def test_array_layout(tmp_path):
    acc = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]]) / 3.0
    pot = np.array([-1.5, 2.25e-12])
    tipsy.write_outputs(str(tmp_path / "out"), acc, pot)
    lines = (tmp_path / "out.acc2").read_text().split()
    assert lines[0] == "2"
    # every x, then every y, then every z, "%.14g"
    assert lines[1:] == [f"{v:.14g}" for v in (acc[0, 0], acc[1, 0], acc[0, 1], acc[1, 1],
                                                acc[0, 2], acc[1, 2])]
    np.testing.assert_allclose(tipsy.read_array(str(tmp_path / "out.acc2"), vector=True), acc,
                               rtol=1e-13)
    np.testing.assert_array_equal(tipsy.read_array(str(tmp_path / "out.pot")), pot)


def test_periodic_workload(tmp_path):
    dark = {"mass": np.full(4, 0.25), "pos": np.array([[-0.5, -0.25, 0.0], [0.1, 0.2, 0.3],
                                                       [0.49, -0.49, 0.0], [0.0, 0.0, 0.0]]),
            "eps": np.full(4, 0.01)}
    path = tmp_path / "box.bin"
    tipsy.write(path, 0.0, dark=dark)
    w = tipsy.workload(path, period=1.0)
    assert w["periodic"] and w["L"] == 1.0
    assert (w["pos"] >= 0).all() and (w["pos"] < 1).all()
    np.testing.assert_allclose(w["pos"][1], np.array([0.6, 0.7, 0.8]), atol=1e-6)
