"""Writes tests/golden/boys_mp.npz: Boys function golden vectors at 40
significant digits.

F_m(T) = 1F1(m + 1/2; m + 3/2; -T) / (2m + 1), evaluated with mpmath's hyp1f1
at 40 digits (an independent arbitrary-precision routine), m = 0..8, at T on
and between the points of a 0.125-spaced grid (offsets 0, +-1/16, +-1/32),
around 50 and 60, at tiny and at large T.  Run once; the committed .npz is
the fixture.
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
rng = np.random.default_rng(7)
base = np.arange(0.0, 64.0, 0.125)
T = np.concatenate([
    [0.0, 1e-300, 1e-30, 1e-16, 1e-12, 1e-8, 1e-4, 1e-2],
    base[rng.choice(base.size, 80, replace=False)],
    base[1:][rng.choice(base.size - 1, 80, replace=False)] - 1.0 / 16,
    base[rng.choice(base.size, 80, replace=False)] + 1.0 / 16,
    base[rng.choice(base.size, 40, replace=False)] + 1.0 / 32,
    rng.uniform(0.0, 64.0, 120),
    [49.9375, 49.99999, 50.0, 50.00001, 50.0625, 59.9375, 59.99999, 60.0, 60.00001, 60.0625],
    np.geomspace(64.0, 1e6, 30),
])
m = np.arange(9)
F = np.array([[float(mp.hyp1f1(mm + mp.mpf(1) / 2, mm + mp.mpf(3) / 2, -mp.mpf(float(t))) / (2 * mm + 1))
               for mm in m] for t in T])
np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "boys_mp.npz"), T=T, m=m, F=F)
print(F.shape)
