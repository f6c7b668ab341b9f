"""DG operators on [0,1] derived from first principles (numpy, float64).

Independent of the literal tables: used by the tests to pin what each table
*means* (orientation, scaling) to ~1e-12, so a transposed or mis-scaled table
cannot go unnoticed.  Basis: Lagrange polynomials phi_k through the N+1
Gauss-Legendre nodes xi_k of [0,1].

  FLCoeff[k] = phi_k(0)          FRCoeff[k] = phi_k(1)          F0 = FLCoeff
  Kxi[k][m]  = w_m phi_k'(xi_m)   (= int_0^1 phi_k' phi_m, exact by GL)
  K1[k][m]   = phi_k(1) phi_m(1) - Kxi[m][k]   (time matrix), iK1 = inv(K1)^T
  dudx[l][n] = phi_n'(xi_l)
  equidistantGridProjector1d[k][j] = phi_k(j/N)
  fineGridProjector1d[s][k][j]     = phi_k((s + xi_j)/3)
"""
import numpy as np


def gl_nodes_weights(n):
    x, w = np.polynomial.legendre.leggauss(n + 1)
    return 0.5 * (x + 1.0), 0.5 * w


def lagrange_eval(nodes, k, x):
    """phi_k(x) for the Lagrange basis through ``nodes``."""
    x = np.asarray(x, dtype=np.float64)
    r = np.ones_like(x)
    for j, xj in enumerate(nodes):
        if j != k:
            r = r * (x - xj) / (nodes[k] - xj)
    return r


def lagrange_deriv(nodes, k, x):
    x = np.asarray(x, dtype=np.float64)
    tot = np.zeros_like(x)
    for i, xi in enumerate(nodes):
        if i == k:
            continue
        term = np.ones_like(x) / (nodes[k] - xi)
        for j, xj in enumerate(nodes):
            if j != k and j != i:
                term = term * (x - xj) / (nodes[k] - xj)
        tot = tot + term
    return tot


def operators(n):
    xi, w = gl_nodes_weights(n)
    B = n + 1
    FL = np.array([lagrange_eval(xi, k, 0.0) for k in range(B)])
    FR = np.array([lagrange_eval(xi, k, 1.0) for k in range(B)])
    D = np.array([[lagrange_deriv(xi, k, xi[m]) for m in range(B)] for k in range(B)])  # D[k][m] = phi_k'(xi_m)
    Kxi = D * w[None, :]
    K1 = np.outer(FR, FR) - Kxi.T
    iK1 = np.linalg.inv(K1).T
    dudx = D.T
    eq = np.array([[lagrange_eval(xi, k, (j / n) if n > 0 else 0.5) for j in range(B)] for k in range(B)])
    fine = np.array([[[lagrange_eval(xi, k, (s + xi[j]) / 3.0) for j in range(B)]
                      for k in range(B)] for s in range(3)])
    return dict(gl_nodes=xi, gl_weights=w, FLCoeff=FL, FRCoeff=FR, F0=FL.copy(),
                Kxi=Kxi, iK1=iK1, dudx=dudx, equidistantGridProjector1d=eq,
                fineGridProjector1d=fine)
