"""Brute-force statevector simulator used to pin the oracle (test infrastructure).

Gate matrices are row-major U[out][in]; for a two-qubit gate on (a, b) the
first listed qubit ``a`` is the least-significant bit of the 4x4 index.
"""
from __future__ import annotations

import numpy as np


class StateVector:
    def __init__(self, n):
        self.n = n
        self.psi = np.zeros((2,) * n, dtype=np.complex128)
        self.psi[(0,) * n] = 1.0

    def apply1(self, u, q):
        self.psi = np.moveaxis(np.tensordot(u, self.psi, axes=([1], [q])), 0, q)

    def apply2(self, u, a, b):
        u4 = u.reshape(2, 2, 2, 2)  # [b_out, a_out, b_in, a_in]
        psi = np.tensordot(u4, self.psi, axes=([2, 3], [b, a]))  # -> b_out, a_out, rest
        self.psi = np.moveaxis(psi, [0, 1], [b, a])

    def expect1(self, o, q):
        phi = np.moveaxis(np.tensordot(o, self.psi, axes=([1], [q])), 0, q)
        return complex(np.vdot(self.psi, phi))

    def expect2(self, o, a, b):
        u4 = o.reshape(2, 2, 2, 2)
        phi = np.moveaxis(np.tensordot(u4, self.psi, axes=([2, 3], [b, a])), [0, 1], [b, a])
        return complex(np.vdot(self.psi, phi))


def kicked_layer(sv, edges, a, b, g):
    """One kicked-Ising Floquet step: exp(-i a ZZ) on every edge, then
    exp(-i b Z), then exp(-i g X) on every site (circuit.h:44-101)."""
    zz = np.diag([np.exp(-1j * a), np.exp(1j * a), np.exp(1j * a), np.exp(-1j * a)])
    uz = np.diag([np.exp(-1j * b), np.exp(1j * b)])
    ux = np.array([[np.cos(g), -1j * np.sin(g)], [-1j * np.sin(g), np.cos(g)]])
    for (i, j) in edges:
        sv.apply2(zz, i, j)
    for q in range(sv.n):
        sv.apply1(uz, q)
    for q in range(sv.n):
        sv.apply1(ux, q)


X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
ZZ = np.diag([1, -1, -1, 1]).astype(np.complex128)
