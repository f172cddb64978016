"""Analytic AO moment integrals <mu| x^a y^b z^c |nu> about the origin, for
the electrostatic-potential pins (the role of BasisEvaluator::
cartesian_moments in tests/electrostatic_potential_test.cpp:98-138).

A path of its own: each one-dimensional factor of a product of two Cartesian
Gaussians is expanded about the product centre P by the binomial theorem,
    (x - A)^i (x - B)^j x^e = sum C(i, r) C(j, s) C(e, m) PA^(i-r) PB^(j-s) P^(e-m) y^(r+s+m),
with y = x - P and the Gaussian moments
    int y^k exp(-p y^2) dy = (k - 1)!! / (2p)^(k/2) sqrt(pi / p)  (k even, 0 odd).
Contraction weights and the Cartesian-to-pure transform come from the
synthetic-basis description (kumo_hip.synthetic.Basis), whose overlap is the
e = 0 case."""
import math

import numpy as np

from kumo_hip.synthetic import _CART, Basis


def _dfact(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2, dtype=float)))


def _basis(d):
    pc = d.p_coefficients if d.p_coefficients is not None else np.zeros_like(d.exponents)
    return Basis(d.shell_types, d.nprim, d.shell_xyz, d.exponents, d.coefficients, pc)


def _primitives(basis):
    cen, ex, pw, wt, ao = [], [], [], [], []
    ncart = 0
    starts = []
    for (o, l, e, w, T) in basis.blocks:
        starts.append(ncart)
        for k, q in enumerate(_CART[l]):
            for p in range(len(e)):
                cen.append(o)
                ex.append(e[p])
                pw.append(q)
                wt.append(w[k, p])
                ao.append(ncart + k)
        ncart += len(_CART[l])
    return (np.asarray(cen, float), np.asarray(ex, float), np.asarray(pw, int), np.asarray(wt, float),
            np.asarray(ao, int), ncart, starts)


def _one_dim(na, nb, e, a, b, xa, xb):
    """int (x-A)^na (x-B)^nb x^e exp(-a(x-A)^2 - b(x-B)^2) dx for arrays of pairs."""
    p = a + b
    P = (a * xa + b * xb) / p
    pa, pb = P - xa, P - xb
    out = np.zeros(a.shape)
    for r in range(int(na.max()) + 1):
        for s in range(int(nb.max()) + 1):
            for m in range(e + 1):
                k = r + s + m
                if k % 2:
                    continue
                sel = (na >= r) & (nb >= s)
                if not sel.any():
                    continue
                gk = _dfact(k - 1) / (2.0 * p[sel]) ** (k // 2) * np.sqrt(math.pi / p[sel])
                cr = np.array([math.comb(int(n), r) for n in na[sel]], float)
                cs = np.array([math.comb(int(n), s) for n in nb[sel]], float)
                out[sel] += (cr * cs * math.comb(e, m) * pa[sel] ** (na[sel] - r) * pb[sel] ** (nb[sel] - s)
                             * P[sel] ** (e - m) * gk)
    return out * np.exp(-a * b / p * (xa - xb) ** 2)


def ao_moments(d, powers):
    """AO matrices (FCHK AO order) of x^a y^b z^c for each (a, b, c) in powers."""
    basis = _basis(d)
    cen, ex, pw, wt, ao, ncart, starts = _primitives(basis)
    A, B = np.meshgrid(np.arange(len(ex)), np.arange(len(ex)), indexing="ij")
    A, B = A.ravel(), B.ravel()
    W = np.zeros((ncart, len(ex)))
    W[ao, np.arange(len(ex))] = wt
    rows = []
    for (o, l, e, w, T), c0 in zip(basis.blocks, starts):
        nc = len(_CART[l])
        r = np.zeros((nc if T is None else T.shape[0], ncart))
        r[:, c0:c0 + nc] = np.eye(nc) if T is None else T
        rows.append(r)
    M = np.vstack(rows)
    out = []
    for power in powers:
        s = np.ones(A.shape)
        for axis in range(3):
            s *= _one_dim(pw[A, axis], pw[B, axis], power[axis], ex[A], ex[B], cen[A, axis], cen[B, axis])
        prim = s.reshape(len(ex), len(ex))
        out.append(M @ (W @ prim @ W.T) @ M.T)
    return out


def electron_moments(d):
    """Charge, dipole (3) and second moments (xx, xy, xz, yy, yz, zz) of the
    total electron density about the origin: sums of P * <mu|x^k|nu>."""
    powers = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1),
              (0, 0, 2)]
    mats = ao_moments(d, powers)
    vals = [float(np.sum(d.p_total * m)) for m in mats]
    return vals[0], np.array(vals[1:4]), np.array(vals[4:10])
