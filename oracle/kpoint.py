"""CPU FP64 oracle for the k-point H|psi> (numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the library's
k-point path (qx_khpsi_slot); the product package never imports it.

The reference has no k-point band algebra (Sources/kpoints.f90:33-39 fixes
k = 0), so this restates its Gamma Hamiltonian apply -- hcnonloc (local
potential through the FFT grid, Sources/engrad.f90:411-606), ekinetic
(|G|^2 c, Sources/engrad.f90:66-127) and hckbpp_g (Sources/nlkbpp.f90:
1284-1393) -- for complex Bloch coefficients c(G) at k + G, directly from the
definitions with numpy's FFT:

    u(r)  = sum_G c(G) e^{iGr}                 (numpy ifftn x N)
    H c   = |k+G|^2 c + FFT[V u] / N + sum_p beta_p D_p <beta_p|c>

hk_dense builds the same operator as a matrix from the convolution theorem
(no band transforms): the two must agree, which pins the FFT index and sign
conventions of hk_apply.

Pinning (tests/test_kpoints_cpu.py): at k = 0 a coefficient set with
c(-G) = conj c(G) reproduces the C oracle's Gamma H (pinned to the reference
build); the labels of a k-point shifted by a reciprocal vector give the same
H (the operator depends on k + G only); hermiticity <a|H b> = conj <b|H a>.
"""
from __future__ import annotations

import numpy as np


def hk_apply(c, hkl, kin, V, beta=None, D=None):
    """H_k c for bands c (nb, npw) complex on the sphere hkl (npw, 3) with
    kin = |k+G|^2 (npw,), V the local potential on the grid (n3, n2, n1)
    (x fastest in memory), beta (npj, npw) complex projector coefficients,
    D (npj,)."""
    c = np.atleast_2d(np.asarray(c, np.complex128))
    n3, n2, n1 = V.shape
    N = n1 * n2 * n3
    h = np.asarray(hkl)
    ix, iy, iz = np.mod(h[:, 0], n1), np.mod(h[:, 1], n2), np.mod(h[:, 2], n3)
    out = np.empty_like(c)
    for b in range(c.shape[0]):
        F = np.zeros((n3, n2, n1), np.complex128)
        F[iz, iy, ix] = c[b]
        u = np.fft.ifftn(F) * N
        w = np.fft.fftn(V * u) / N
        out[b] = w[iz, iy, ix] + kin * c[b]
    if beta is not None and len(beta):
        P = np.conj(beta) @ c.T  # (npj, nb)
        out += ((np.asarray(D)[:, None] * P).T) @ beta
    return out


def hk_dense(hkl, kin, V, beta=None, D=None):
    """H_k as a dense matrix from its definition, without transforming bands:
    |k+G|^2 on the diagonal, V^(G - G') (the grid's discrete transform of V,
    indices modulo the grid) and sum_p beta_p(G) D_p conj(beta_p(G'))."""
    n3, n2, n1 = V.shape
    vh = np.fft.fftn(V) / V.size
    h = np.asarray(hkl)
    dz = np.mod(h[:, None, 2] - h[None, :, 2], n3)
    dy = np.mod(h[:, None, 1] - h[None, :, 1], n2)
    dx = np.mod(h[:, None, 0] - h[None, :, 0], n1)
    H = vh[dz, dy, dx]
    H[np.diag_indices_from(H)] += kin
    if beta is not None and len(beta):
        H = H + (np.asarray(beta).T * np.asarray(D)[None, :]) @ np.conj(beta)
    return H
