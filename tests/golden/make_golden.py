"""Writes the committed golden fixtures (run in the container that holds the
reference tree; the outputs are data, the reference stays where it is).

  moments_golden.npz  inputs and outputs of the reference's moments.c
                      (momMakeFmomr, momShiftFmomr, momRescaleFmomr,
                      momScaledAddFmomr, momFmomr2Momc, momEvalFmomrcm),
                      compiled from its own source into oracle/_ref by
                      oracle/Makefile.
  king_soft.npz       the particles of teststep/king_soft.bin (tipsy,
                      big-endian: 36,000 dark particles: mass, pos, vel, eps)
  pkdtest_t0.json     the t = 0 energies of teststep/pkdtest.log (PKDGRAV,
                      theta = 0.7, hexadecapole, nBucket 8)
  king_soft_head.bin  the first 64 particles of teststep/king_soft.bin as a
                      tipsy file (header counts set to 64, records byte for
                      byte)

Usage: python tests/golden/make_golden.py [/root/reference]
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def moments_vectors():
    from oracle import oracle as O
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libmoments_ref.so is not built (make -C oracle)")
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rng = np.random.default_rng(20261202)
    n = 64
    out = {}
    # momMakeFmomr: (m, u, x, y, z) -> FMOMR
    mk_in = np.column_stack([rng.uniform(0.1, 2, n), rng.uniform(0.5, 3, n), rng.normal(size=(n, 3))])
    mk_out = np.zeros((n, 22))
    for i in range(n):
        R.momMakeFmomr(P(mk_out[i]), *mk_in[i])
    out["make_in"], out["make_out"] = mk_in, mk_out
    # momShiftFmomr: (q, u, d) -> q'
    sh_q = rng.normal(size=(n, 22))
    sh_q[:, 0] = np.abs(sh_q[:, 0]) + 1
    sh_ud = np.column_stack([rng.uniform(0.5, 2, n), 0.5 * rng.normal(size=(n, 3))])
    sh_out = sh_q.copy()
    for i in range(n):
        R.momShiftFmomr(P(sh_out[i]), *sh_ud[i])
    out["shift_q"], out["shift_ud"], out["shift_out"] = sh_q, sh_ud, sh_out
    # momRescaleFmomr: (q, unew, uold)
    rs_u = rng.uniform(0.3, 3, size=(n, 2))
    rs_out = sh_q.copy()
    for i in range(n):
        R.momRescaleFmomr(P(rs_out[i]), *rs_u[i])
    out["rescale_u"], out["rescale_out"] = rs_u, rs_out
    # momScaledAddFmomr: (qr, ur, qa, ua)
    sa_qa = rng.normal(size=(n, 22))
    sa_u = rng.uniform(0.3, 3, size=(n, 2))
    sa_out = sh_q.copy()
    for i in range(n):
        qa = sa_qa[i].copy()
        R.momScaledAddFmomr(P(sa_out[i]), sa_u[i, 0], P(qa), sa_u[i, 1])
    out["sadd_qa"], out["sadd_u"], out["sadd_out"] = sa_qa, sa_u, sa_out
    # momFmomr2Momc
    mc_out = np.zeros((n, 32))
    for i in range(n):
        q = sh_q[i].copy()
        R.momFmomr2Momc(P(q), P(mc_out[i]))
    out["momc_out"] = mc_out
    # momEvalFmomrcm: (q, u, r) -> pot, a
    ev_u = rng.uniform(0.5, 2, n)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ev_r = dirs * rng.uniform(2.0, 10.0, size=(n, 1)) * ev_u[:, None]
    ev_out = np.zeros((n, 4))
    for i in range(n):
        q = sh_q[i].copy()
        pot, ax, ay, az, mg = (np.zeros(1) for _ in range(5))
        x, y, z = ev_r[i]
        R.momEvalFmomrcm(P(q), ev_u[i], 1.0 / np.linalg.norm(ev_r[i]), x, y, z, P(pot), P(ax),
                         P(ay), P(az), P(mg))
        ev_out[i] = [pot[0], ax[0], ay[0], az[0]]
    out["eval_u"], out["eval_r"], out["eval_out"] = ev_u, ev_r, ev_out
    np.savez_compressed(os.path.join(HERE, "moments_golden.npz"), **out)


def king_soft(ref):
    d = open(os.path.join(ref, "teststep", "king_soft.bin"), "rb").read()
    n = int(np.frombuffer(d[8:12], dtype=">i4")[0])
    ndark = int(np.frombuffer(d[20:24], dtype=">i4")[0])
    assert n == ndark == 36000
    a = np.frombuffer(d[32:32 + 36 * n], dtype=">f4").reshape(n, 9)
    np.savez_compressed(os.path.join(HERE, "king_soft.npz"), mass=a[:, 0].astype(np.float32),
                        pos=a[:, 1:4].astype(np.float32), vel=a[:, 4:7].astype(np.float32),
                        eps=a[:, 7].astype(np.float32))
    for line in open(os.path.join(ref, "teststep", "pkdtest.log")):
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        # columns: time redshift TotalE Kinetic Potential ...
        json.dump({"time": float(f[0]), "total_energy": float(f[2]), "kinetic": float(f[3]),
                   "potential": float(f[4]), "theta": 0.7, "nBucket": 8, "iOrder": 4,
                   "source": "teststep/pkdtest.log first data row"},
                  open(os.path.join(HERE, "pkdtest_t0.json"), "w"), indent=1)
        break


def king_soft_head(ref, n=64):
    """The first n dark particles of teststep/king_soft.bin as a tipsy file
    of their own: the file's big-endian header with nbodies = ndark = n, then
    the n records byte for byte (a reader fixture)."""
    d = open(os.path.join(ref, "teststep", "king_soft.bin"), "rb").read()
    head = bytearray(d[:32])
    head[8:12] = np.array([n], dtype=">i4").tobytes()
    head[20:24] = np.array([n], dtype=">i4").tobytes()
    with open(os.path.join(HERE, "king_soft_head.bin"), "wb") as f:
        f.write(bytes(head) + d[32:32 + 36 * n])


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    moments_vectors()
    king_soft(ref)
    king_soft_head(ref)
    print("fixtures written to", HERE)
