"""The occupied/virtual split of an orbital spectrum (src/Base/orbital_spaces.py).

Pair layout everywhere in this package: z[i, a], occupied index slow and
virtual index fast (src/SingleReference/LinearResponse/block_actions.py:99-115).
"""
import numpy as np


def get_occ_virt_indices(eps, nocc):
    """occ = [0, nocc), virt = [nocc, norb) (orbital_spaces.py:6-9)."""
    norb = len(eps)
    return np.arange(nocc), np.arange(nocc, norb)


def particle_hole_window(eps, nocc):
    """(gap, span, mu): smallest and largest eps_a - eps_i and the mid-gap
    chemical potential (orbital_spaces.py:12-20)."""
    eps = np.asarray(eps, float)
    homo, lumo = eps[:nocc].max(), eps[nocc:].min()
    return lumo - homo, eps[nocc:].max() - eps[:nocc].min(), 0.5 * (homo + lumo)


def ov_energies(eps, nocc):
    """d_(i,a) = eps_a - eps_i, flattened occupied-slow (orbital_spaces.py:31-34)."""
    eps = np.asarray(eps, float)
    return (eps[nocc:][None, :] - eps[:nocc][:, None]).ravel()


def bse_pair_diagonal(eps, nocc):
    """eps_a - eps_i as (n_occ, n_vir) (block_actions.py:99-115)."""
    return ov_energies(eps, nocc).reshape(nocc, -1)


def pair_symmetry(orbsym, occ, virt):
    """The irrep label of every occupied-virtual pair, occupied-slow: the XOR
    of the two orbitals' D2h labels (pyscf's irrep IDs modulo 10); None
    without orbsym (LR/davidson.py:623-628)."""
    if orbsym is None:
        return None
    d2h = np.asarray(orbsym) % 10
    return (d2h[np.asarray(occ)][:, None] ^ d2h[np.asarray(virt)][None, :]).ravel()
