"""Write oracle fixtures that take too long to recompute inside a GPU test.

fock_c2h6_631g.npz: seeded symmetric P (rng seed 1, 0.5*(A+A^T), A ~ N(0,1))
and G = J - K/2 from the oracle's full nbas^4 loop (refcpu.fock2e) for ethane,
6-31G (librint_amd.mol.alkane_631g(2)).

Run: python tests/golden/make_oracle_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from librint_amd import mol  # noqa: E402
from oracle import refcpu  # noqa: E402


def main():
    m = mol.alkane_631g(2)
    n = m.nao()
    rng = np.random.default_rng(1)
    P = rng.normal(size=(n, n))
    P = 0.5 * (P + P.T)
    G = refcpu.fock2e(m, P)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fock_c2h6_631g.npz"), P=P, G=G)


if __name__ == "__main__":
    main()
