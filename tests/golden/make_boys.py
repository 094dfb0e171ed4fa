"""Writes tests/golden/boys.npz: Boys function golden vectors.

F_m(T) = 1F1(m + 1/2; m + 3/2; -T) / (2m + 1), evaluated with scipy 1.15's
hyp1f1 (an independent library routine), m = 0..16, on T from 0 to 200.
Run once; the committed .npz is the fixture.
"""
import os

import numpy as np
from scipy.special import hyp1f1

T = np.concatenate([[0.0, 1e-12, 1e-8, 1e-4, 1e-2], np.linspace(0.05, 50.0, 400),
                    np.geomspace(50.0, 200.0, 40)])
m = np.arange(17)
F = np.array([[hyp1f1(mm + 0.5, mm + 1.5, -t) / (2 * mm + 1) for mm in m] for t in T])
np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "boys.npz"), T=T, m=m, F=F)
print(F.shape)
