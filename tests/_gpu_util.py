"""Shared helpers of the GPU parity tests (HIP library vs the CPU oracle)."""
import numpy as np

import _oracle as O
from mantiis_amd.hip import Lattice

TRANSPORT = dict(dx=1.0e-9, Fbody=1.0e-8, T=300.0, P=2.0e6)      # test/transport_test.cpp:9-16
ADSORPTION = dict(dx=12.0e-9, Fbody=1.0e-8, T=300.0, P=2.0e6)    # test/adsorption_test.cpp:10-17


def pair(pore, kn, lp, precision="fp64", collision=2, wall=2, **phys):
    o = O.Oracle(pore, kn, lp, **phys)
    o.set_operators(collision, wall)
    g = Lattice(pore, kn, lp, precision=precision, collision=collision, wall=wall, **phys)
    return o, g


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def banner(msg):
    print(f"[gpu-case] {msg}", flush=True)
