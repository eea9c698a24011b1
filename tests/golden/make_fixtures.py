"""Generates the committed fixtures under tests/golden/.

    python tests/golden/make_fixtures.py

bessel_roots.json  roots j_{m,k} of J_m, m = 0..4, k = 1..511, from
                   mpmath.besseljzero at 30 digits (independent of scipy's
                   jn_zeros, which the reference calls: transformer.py:45).
bessel_values.json J_m(x) at 40 points per order from mpmath.besselj.
oracle_tiny.npz    the CPU oracle's state after 3 PIC steps of a 64 x 24,
                   M=1 laser-plasma set-up (regression fixture of the oracle
                   itself; regenerate after an intended oracle change).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def bessel():
    import mpmath
    mpmath.mp.dps = 30
    roots = {str(m): [float(mpmath.besseljzero(m, k)) for k in range(1, 512)] for m in range(5)}
    with open(os.path.join(HERE, "bessel_roots.json"), "w") as f:
        json.dump(roots, f)
    xs = [0.003, 0.05, 0.4, 1.0, 1.9, 2.1, 3.7, 6.0, 11.0, 25.0, 77.7, 160.0, 400.0, 1600.0]
    vals = {str(m): [[x, float(mpmath.besselj(m, x))] for x in xs] for m in range(5)}
    with open(os.path.join(HERE, "bessel_values.json"), "w") as f:
        json.dump(vals, f)


def tiny_trajectory():
    from oracle import oracle as O
    from tests import simpair as S
    from chimeracl_amd import config as C
    g = S.tiny_grid(Nx=64, Nr=24, M=1)
    s = O.OracleSolver(dict(g))
    O.add_gausian_pulse(s, S.tiny_laser(g))
    eo, io = C.species_config("C1", s.Args)
    arr = C.load_plasma(g, (2, 2, 4), 0.002, seed=0)
    e, i = O.OracleParticles(eo), O.OracleParticles(io)
    e.add_particles(arr)
    i.add_particles(arr)
    loop = O.OracleLoop(s, [e, i])
    for _ in range(3):
        loop.step()
    out = {}
    rng = np.random.default_rng(7)
    for k in S.real_names(1):
        a = s.f[k][1:]
        out["norm_" + k] = np.array([np.linalg.norm(a)])
        idx = rng.integers(0, a.size, 16)
        out["samp_" + k] = a.ravel()[idx]
        out["idx_" + k] = idx
    for k in ("px", "py", "pz", "x"):
        out["p_" + k] = e.data[k][::97]
    out["sort_head"] = e.data["sort_indx"][:512]
    out["np_stay"] = np.array([e.Args["Np_stay"]])
    np.savez_compressed(os.path.join(HERE, "oracle_tiny.npz"), **out)


if __name__ == "__main__":
    bessel()
    tiny_trajectory()
    print("fixtures written to", HERE)
