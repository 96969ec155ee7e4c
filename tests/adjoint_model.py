"""Dense numpy model of the discrete adjoint used by liblmhip (test helper).

Forward step (what the device computes for s = 1):
    side S(X) = (U X)^H with U X by Horner, y_{K+1} = x, y_m = x + c_m H y_{m+1},
    c_m = -i dt / m;   P_{k+1} = S(S(P_k)) = U P_k U^H.

Reverse step k (Lambda = dL/dP as a Hermitian cotangent, dL = Re tr(Lambda^H dP)):
    A:  x_A = U^H P_{k+1}           (= (U P_k)^H up to the unitarity defect of U)
        reverse Horner on x_A with cotangent Lambda_{k+1}:
            g_1 = Lambda_{k+1};  F += c_m conj(g_m) y_{m+1}^T;  xbar += g_m;  g_{m+1} = conj(c_m) H g_m
        Ybar^H = xbar^H
    B:  x_B = P_k = U^H (x_A)^H, or the stored checkpoint when k % C == 0
        the same reverse Horner on x_B with cotangent Ybar^H -> Lambda_k
    dL/dH_ij (on the pattern of H) = F_ij summed over both sides; the on-site
    gradient is Re F_ii and a field strength a gets Re sum_ij F_ij dH_ij/da.

Everything here is dense and O(N^3); it is only meant for lattices of a few
dozen sites, to pin the algorithm against finite differences on the CPU and
to give the device gradients a same-math reference.
"""
import numpy as np


def coeffs(dt, K):
    return [-1j * dt / m for m in range(1, K + 1)]


def apply_U(H, X, dt, K):
    c = coeffs(dt, K)
    Y = X.copy()
    for m in range(K, 0, -1):
        Y = X + c[m - 1] * (H @ Y)
    return Y


def apply_Udag(H, X, dt, K):
    return apply_U(H, X, -dt, K)


def step(H, P, dt, K):
    Y = apply_U(H, P, dt, K).conj().T
    return apply_U(H, Y, dt, K).conj().T


def forward(Hs, dts, Ks, P0):
    Ps = [np.array(P0, dtype=np.complex128)]
    for H, dt, K in zip(Hs, dts, Ks):
        Ps.append(step(H, Ps[-1], dt, K))
    return Ps


def side_adjoint(H, x, G, dt, K):
    """x: input columns of the side, G: cotangent of U x.  Returns (U^H G, F)."""
    c = coeffs(dt, K)
    ys = {K + 1: x}
    y = x
    for m in range(K, 1, -1):
        y = x + c[m - 1] * (H @ y)
        ys[m] = y
    g = G
    xbar = np.zeros_like(G)
    F = np.zeros(H.shape, dtype=np.complex128)
    for m in range(1, K + 1):
        F += c[m - 1] * (g.conj() @ ys[m + 1].T)
        xbar += g
        g = np.conj(c[m - 1]) * (H @ g)
    xbar += g
    return xbar, F


def backward(Hs, dts, Ks, P_M, Lam_M, checkpoints, C):
    """Returns (Lambda_0, [F_k]) with F_k the dL/dH_k matrix of step k."""
    Lam = 0.5 * (Lam_M + Lam_M.conj().T)
    Pn = np.array(P_M, dtype=np.complex128)
    Fs = [None] * len(Hs)
    for k in range(len(Hs) - 1, -1, -1):
        H, dt, K = Hs[k], dts[k], Ks[k]
        xA = apply_Udag(H, Pn, dt, K)
        xbarA, FA = side_adjoint(H, xA, Lam, dt, K)
        if k % C == 0:
            xB = np.array(checkpoints[k], dtype=np.complex128)
        else:
            xB = apply_Udag(H, xA.conj().T, dt, K)
        xbarB, FB = side_adjoint(H, xB, xbarA.conj().T, dt, K)
        Fs[k] = FA + FB
        Lam = xbarB
        Pn = xB
    return Lam, Fs
