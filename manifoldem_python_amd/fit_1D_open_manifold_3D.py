"""Drop-in for ``fit_1D_open_manifold_3D`` (``fit_1D_open_manifold_3D.py:163-259``):
op(psi) -> (a, b, tau), vectorised over points (see fit1d)."""
from . import fit1d


def op(psi):
    return fit1d.fit(psi)
