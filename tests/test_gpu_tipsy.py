"""teststep/king_soft.bin through the tipsy path on the GPU, against the one
gravity number the reference's own tests hold for it: the potential energy
of the first data row of teststep/pkdtest.log (PKDGRAV, theta 0.7,
hexadecapole, nBucket 8), 1/2 sum m phi = -64.84776
(TreePiece.cpp:1420,1431 form the same sum). The particles are the fixture
tests/golden/king_soft.npz (the file's FP32 records), written back as a
tipsy file and evaluated by changa_amd.tipsy.evaluate (GPU tree, GPU lists,
precision form) with the log's parameters.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_gpu_king_soft_pkdgrav_energy(device, tmp_path):
    from changa_amd import tipsy
    ref = json.load(open(os.path.join(GOLD, "pkdtest_t0.json")))
    k = np.load(os.path.join(GOLD, "king_soft.npz"))
    path = tmp_path / "king_soft.bin"
    tipsy.write(path, 0.0, dark={"mass": k["mass"], "pos": k["pos"], "vel": k["vel"],
                                 "eps": k["eps"]})
    t = tipsy.read(path)
    assert len(t["dark"]) == 36000 and t["big_endian"]
    # the first 64 records are the bytes of the reference file (the fixture
    # king_soft_head.bin holds them after a 32-byte header)
    head = open(os.path.join(GOLD, "king_soft_head.bin"), "rb").read()
    assert open(path, "rb").read(len(head))[32:] == head[32:]
    acc, pot = tipsy.evaluate(path, str(tmp_path / "king"), theta=ref["theta"],
                              bucket=ref["nBucket"], device=device)
    m = k["mass"].astype(np.float64)
    pe = 0.5 * float((m * pot).sum())
    print(f"king_soft GPU PE {pe:.7f} vs PKDGRAV {ref['potential']}", flush=True)
    assert pe == pytest.approx(ref["potential"], rel=2e-5)
    # Newton's third law on the whole isolated system: sum m a ~ 0 relative to sum m |a|
    ma = (m[:, None] * acc).sum(axis=0)
    assert np.linalg.norm(ma) <= 1e-4 * float((m * np.linalg.norm(acc, axis=1)).sum())
