"""Seeded synthetic inputs for the five gravity workloads (numpy PCG64,
seed = 1000 + workload number). Positions are FP64; every particle has mass
1/N.

  plummer      isolated Plummer sphere, N = 65,536 (and N = 2^20), a = 1, M = 1,
               eps = 0.01, theta = 0.7, bucket 16
  zeldovich    unit periodic box, ng^3 lattice displaced by a Zel'dovich field
               with P(k) ~ k^-2.5 (ng = 128, rms 0.02; ng = 256, rms 0.15),
               eps = 1/(20 ng), Ewald on (nReplicas = 1, fEwCut = 2.6,
               |h|^2 <= 10), theta = 0.7, bucket 16
  halos        2^22 particles: 80% in 512 Hernquist halos (dN/dM ~ M^-1.9 over
               3 decades, a ~ M^(1/3), a_max = 0.02, truncated at 20 a), 20%
               uniform background; theta = 0.6, bucket 32
"""
import numpy as np


def _rng(cfg):
    return np.random.default_rng(1000 + cfg)


def _isotropic(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def plummer(n=65536, cfg=1, a=1.0, eps=0.01, theta=0.7, bucket=16):
    rng = _rng(cfg)
    X = rng.uniform(0.0, 0.999, size=n)
    r = a / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    pos = r[:, None] * _isotropic(rng, n)
    return dict(name=f"plummer_N{n}", pos=np.ascontiguousarray(pos),
                mass=np.full(n, 1.0 / n), soft=np.full(n, eps), theta=theta, bucket=bucket,
                periodic=False, L=0.0, n_replicas=0, ew_cut=0.0, ewh_cut=0.0, active=None)


def zeldovich(ng=128, rms=0.02, cfg=2, theta=0.7, bucket=16, active_frac=None):
    rng = _rng(cfg)
    n = ng ** 3
    white = rng.standard_normal((ng, ng, ng))
    wk = np.fft.rfftn(white)
    del white
    kx = 2 * np.pi * np.fft.fftfreq(ng, d=1.0 / ng)
    kz = 2 * np.pi * np.fft.rfftfreq(ng, d=1.0 / ng)
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[0, 0, 0] = 1.0
    amp = k2 ** (-2.5 / 4.0)  # sqrt(P(k)), P ~ k^-2.5
    amp[0, 0, 0] = 0.0
    dk = wk * amp
    del wk, amp
    disp = []
    for K in (KX, KY, KZ):
        psi_k = 1j * K * dk / k2
        disp.append(np.fft.irfftn(psi_k, s=(ng, ng, ng), axes=(0, 1, 2)).astype(np.float64))
    del dk, KX, KY, KZ, k2
    psi = np.stack([d.reshape(-1) for d in disp], axis=1)
    del disp
    cur = np.sqrt(np.mean(np.sum(psi ** 2, axis=1)))
    psi *= rms / cur
    g = (np.arange(ng) + 0.5) / ng
    GX, GY, GZ = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([GX.reshape(-1), GY.reshape(-1), GZ.reshape(-1)], axis=1) + psi
    pos -= np.floor(pos)
    pos[pos >= 1.0] = 0.0
    active = None
    if active_frac is not None:
        active = rng.uniform(size=n) < active_frac
    return dict(name=f"zeldovich_ng{ng}_rms{rms}", pos=np.ascontiguousarray(pos),
                mass=np.full(n, 1.0 / n), soft=np.full(n, 1.0 / (20.0 * ng)), theta=theta,
                bucket=bucket, periodic=True, L=1.0, n_replicas=1, ew_cut=2.6,
                ewh_cut=np.sqrt(10.0), active=active)


def halos(n=1 << 22, cfg=3, n_halos=512, bg_frac=0.2, a_max=0.02, theta=0.6, bucket=32,
          eps=None):
    rng = _rng(cfg)
    n_bg = int(round(bg_frac * n))
    n_h = n - n_bg
    u = rng.uniform(size=n_halos)
    # dN/dM ~ M^-1.9 on [1, 1000]: inverse CDF of M^-0.9
    lo = 1000.0 ** -0.9
    M = (1.0 - u * (1.0 - lo)) ** (-1.0 / 0.9)
    frac = M / M.sum()
    counts = np.floor(frac * n_h).astype(np.int64)
    rest = n_h - counts.sum()
    counts[np.argsort(-(frac * n_h - counts))[:rest]] += 1
    a_h = a_max * (M / M.max()) ** (1.0 / 3.0)
    centres = rng.uniform(size=(n_halos, 3))
    parts = []
    xmax = (20.0 / 21.0) ** 2
    for h in range(n_halos):
        c = int(counts[h])
        X = rng.uniform(0.0, xmax, size=c)
        s = np.sqrt(X)
        r = a_h[h] * s / (1.0 - s)
        parts.append(centres[h] + r[:, None] * _isotropic(rng, c))
    parts.append(rng.uniform(size=(n_bg, 3)))
    pos = np.ascontiguousarray(np.concatenate(parts, axis=0))
    if eps is None:
        eps = 1.0 / (20.0 * n ** (1.0 / 3.0))
    return dict(name=f"halos_N{n}", pos=pos, mass=np.full(n, 1.0 / n), soft=np.full(n, eps),
                theta=theta, bucket=bucket, periodic=False, L=0.0, n_replicas=0, ew_cut=0.0,
                ewh_cut=0.0, active=None)


def workload(cfg, scale=None):
    """The five workloads by number; `scale` shrinks them for tests."""
    if cfg == 1:
        return plummer(n=scale or 65536, cfg=1)
    if cfg == 2:
        return zeldovich(ng=scale or 128, rms=0.02, cfg=2)
    if cfg == 3:
        return halos(n=scale or (1 << 22), cfg=3)
    if cfg == 4:
        return plummer(n=scale or (1 << 20), cfg=4)
    if cfg == 5:
        return zeldovich(ng=scale or 256, rms=0.15, cfg=5, active_frac=0.1)
    raise ValueError(cfg)
