"""Writes the small regression fixtures under tests/golden/ from the float64 oracle.

Run from the repository root: python tests/golden/make_golden.py
The fixtures pin the oracle against accidental change; the oracle itself is pinned by the
closed-form, invariant and brute-force tests in tests/test_oracle_*.py.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from manifoldem_python_amd import synth  # noqa: E402
from oracle import distance as od, dm  # noqa: E402


def main():
    s = synth.make_pd(synth.seeds(1)[0], 12, 24, conj_frac=0.25)
    p = dict(nPix=24, pix_size=1.5, Cs=2.7, EkV=300.0, gaussEnv=np.inf, AmpContrast=0.1)
    out = od.op([s.ind, s.q, s.df, None, 0], dict(type="Butter", Qc=0.5, N=8), s.raw, s.nStot,
                dict(avgOnly=False), p)
    np.savez_compressed(os.path.join(HERE, "stage1_n12_N24.npz"), raw=s.raw, q=s.q, df=s.df,
                        ind=s.ind, nStot=s.nStot, N=24, D=out["D"], imgAll=out["imgAll"],
                        CTF=out["CTF"])
    rng = np.random.default_rng(5)
    x = np.sort(rng.uniform(0, 1, 40))
    D = (x[:, None] - x[None, :]) ** 2 * 1e4 + rng.uniform(0, 1, (40, 40))
    D = (D + D.T) / 2
    np.fill_diagonal(D, 0.0)
    le = np.linspace(-150, 150, 301)
    lamb, psi, sigma, mu, _, lsw, popt, R2 = dm.op(D, 25, 3, 60000, num_eigs=8,
                                                  rng=np.random.default_rng(6), dense=True,
                                                  logEps=le)
    np.savez_compressed(os.path.join(HERE, "dm_n40_k25.npz"), D=D, logEps=le, lamb=lamb,
                        sigma=sigma, mu=mu, logSumWij=lsw, popt=popt)
    print("wrote fixtures")


if __name__ == "__main__":
    main()
