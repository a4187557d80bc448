"""Write the golden fixtures under tests/golden/ from the CPU oracle.

Run in the development container (python tests/golden/make_golden.py).
Both sides regenerate the a_lm from wavemoth_amd.alm.power_law_alm with
the stored seed; each fixture stores a checksum of that input so a drift
of the generator is caught before any comparison.

Fixtures:
  golden_n64_l128.npz   Nside=64, lmax=mmax=128, nmaps=1: full map, and the
                        Legendre-stage work (m, odd, ring) for 8 m values
  golden_small.npz      Nside in {1,2,4,8,16,32}, lmax=2 Nside, nmaps in
                        {1,3}: full maps
  golden_n1024_l2048.npz  Nside=1024, lmax=mmax=2048, nmaps=1: every pixel
                        of 24 rings (poles, cap/belt boundaries, equator,
                        random), every 251st pixel of the map, per-ring rms
  golden_n2048_l4096.npz  Nside=2048, lmax=mmax=4096, nmaps=3 (map k of the
                        K=1 case is column 0 of the same a_lm): map 0 of
                        14 whole rings (both poles, cap rings whose split
                        Bluestein length is 4096, the cap/belt boundaries,
                        belt rings of both phi0 parities, the equator),
                        every 2003rd pixel of all 3 maps, per-ring and
                        per-map rms
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from wavemoth_amd.alm import power_law_alm  # noqa: E402


def alm_checksum(alm):
    return np.array([np.sum(np.abs(alm) ** 2), np.sum(alm.real), np.sum(alm.imag)])


def small():
    d = {}
    for N in [1, 2, 4, 8, 16, 32]:
        for K in [1, 3]:
            L = 2 * N
            seed = 100 * N + K
            alm = power_law_alm(L, L, K, seed)
            mp = O.OraclePlan(N, L, L, K).synthesize(alm)
            key = "n%d_k%d" % (N, K)
            d[key + "_map"] = mp
            d[key + "_seed"] = np.array([seed])
            d[key + "_alm_checksum"] = alm_checksum(alm)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **d)


def config1():
    N, L, K, seed = 64, 128, 1, 64128
    alm = power_law_alm(L, L, K, seed)
    op = O.OraclePlan(N, L, L, K)
    work = op.legendre_stage(alm)
    mp = op.c2r(op.ring_spectra(work))
    ms = np.array([0, 1, 2, 17, 63, 64, 127, 128])
    np.savez_compressed(os.path.join(HERE, "golden_n64_l128.npz"), map=mp, seed=np.array([seed]),
                        alm_checksum=alm_checksum(alm), work_m=ms, work=work[ms])


def config2():
    N, L, K, seed = 1024, 2048, 1, 10242048
    alm = power_law_alm(L, L, K, seed)
    t0 = time.time()
    op = O.OraclePlan(N, L, L, K)
    t1 = time.time()
    mp = op.synthesize(alm)
    t2 = time.time()
    print("config2 precompute %.1fs execute %.1fs" % (t1 - t0, t2 - t1), flush=True)
    offs = op.ring_offsets
    nrings = 4 * N - 1
    rng = np.random.default_rng(5)
    fixed = [0, 1, 2, N - 2, N - 1, N, N + 1, 2 * N - 2, 2 * N - 1, 2 * N, 3 * N - 1, 3 * N,
             nrings - 2, nrings - 1]
    extra = sorted(set(rng.integers(0, nrings, 40).tolist()) - set(fixed))[:10]
    rings = np.array(sorted(set(fixed + extra)))
    ring_pix = np.concatenate([mp[offs[i]:offs[i + 1], 0] for i in rings])
    ring_rms = np.array([np.sqrt(np.mean(mp[offs[i]:offs[i + 1], 0] ** 2))
                         for i in range(nrings)])
    stride_idx = np.arange(0, mp.shape[0], 251)
    np.savez_compressed(os.path.join(HERE, "golden_n1024_l2048.npz"), seed=np.array([seed]),
                        alm_checksum=alm_checksum(alm), rings=rings, ring_pix=ring_pix,
                        ring_rms=ring_rms, stride_idx=stride_idx, stride_pix=mp[stride_idx, 0],
                        map_norm=np.array([np.linalg.norm(mp)]))


def metric_size():
    N, L, K, seed = 2048, 4096, 3, 20484096
    alm = power_law_alm(L, L, K, seed)
    t0 = time.time()
    op = O.OraclePlan(N, L, L, K)
    t1 = time.time()
    mp = op.synthesize(alm)
    t2 = time.time()
    print("metric size precompute %.1fs execute %.1fs" % (t1 - t0, t2 - t1), flush=True)
    offs = op.ring_offsets
    nrings = 4 * N - 1
    # ring index iring (0 = north pole): cap rings iring < N hold 4 (iring + 1)
    # pixels; their split Bluestein length is next_pow2(2 (iring + 1) - 1),
    # 4096 from iring = 1024 up.  Belt rings alternate phi0 = 0 / pi / (4N).
    rings = np.array([0, 1, 511, 1024, 1500, 2046, N - 1, N, N + 1, 2 * N - 2, 2 * N - 1,
                      3 * N - 1, 3 * N, nrings - 1])
    ring_pix = np.concatenate([mp[offs[i]:offs[i + 1], 0] for i in rings])
    ring_rms = np.array([np.sqrt(np.mean(mp[offs[i]:offs[i + 1]] ** 2, axis=0))
                         for i in range(nrings)])
    stride_idx = np.arange(0, mp.shape[0], 2003)
    np.savez_compressed(os.path.join(HERE, "golden_n2048_l4096.npz"), seed=np.array([seed]),
                        alm_checksum=alm_checksum(alm), alm1_checksum=alm_checksum(alm[:, :1]),
                        rings=rings, ring_pix=ring_pix, ring_rms=ring_rms,
                        stride_idx=stride_idx, stride_pix=mp[stride_idx],
                        map_norm=np.linalg.norm(mp, axis=0))


if __name__ == "__main__":
    which = sys.argv[1:] or ["small", "config1", "config2"]
    for w in which:
        t = time.time()
        globals()[w]()
        print(w, "%.1fs" % (time.time() - t), flush=True)
