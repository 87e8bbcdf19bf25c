"""Oracle: the method of moving asymptotes (test infrastructure only).

topoptlab calls ``mmapy.mmasub`` (``topology_optimization.py:591-607``) with
the defaults of ``optimizer/mma_utils.py:8-60``.  mmapy is a third-party
package (pyproject.toml:24, no version pin) that is not vendored in the
reference and not installed here, so this file restates the published
algorithm it implements:

* K. Svanberg, "The method of moving asymptotes - a new method for
  structural optimization", IJNME 24 (1987) 359-373: moving asymptotes,
  move limits, the convex separable approximation;
* K. Svanberg, "MMA and GCMMA - two methods for nonlinear optimization"
  (2007 note): the subproblem with the artificial variables y, z and its
  primal-dual interior-point solution (``subsolv``), epsimin = 1e-7.

Parity is anchored on the reference's own call site and golden: the MBB
60x20 run of ``tests/test_compliance.py:62-86`` (233.709917, 5 decimals).
The call site passes ``iter=i`` with ``i`` left over from the load-case loop
(``topology_optimization.py:463``), i.e. 0 for one load case, so the
asymptotes are re-initialised to x -/+ asyinit*(xmax-xmin) every iteration;
the loop here reproduces that.
"""
from __future__ import annotations

import numpy as np


def mmasub(m, n, iter, xval, xmin, xmax, xold1, xold2, f0val, df0dx, fval, dfdx, low, upp,
           a0, a, c, d, move=0.5, asyinit=0.5, asydecr=0.7, asyincr=1.2, asymin=0.01,
           asymax=10.0, raa0=1e-5, albefa=0.1, **kwargs):
    """One MMA step.  Column vectors (n,1)/(m,1), dfdx (m,n).  Returns
    (xmma, ymma, zmma, lam, xsi, eta, mu, zet, s, low, upp)."""
    epsimin = 1e-7
    een = np.ones((n, 1))
    eem = np.ones((m, 1))
    # asymptotes
    if iter <= 2:
        low = xval - asyinit * (xmax - xmin)
        upp = xval + asyinit * (xmax - xmin)
    else:
        zzz = (xval - xold1) * (xold1 - xold2)
        factor = een.copy()
        factor[zzz > 0] = asyincr
        factor[zzz < 0] = asydecr
        low = xval - factor * (xold1 - low)
        upp = xval + factor * (upp - xold1)
        low = np.minimum(np.maximum(low, xval - asymax * (xmax - xmin)), xval - asymin * (xmax - xmin))
        upp = np.maximum(np.minimum(upp, xval + asymax * (xmax - xmin)), xval + asymin * (xmax - xmin))
    # bounds of the subproblem: asymptote distance and move limit
    alfa = np.maximum(np.maximum(low + albefa * (xval - low), xval - move * (xmax - xmin)), xmin)
    beta = np.minimum(np.minimum(upp - albefa * (upp - xval), xval + move * (xmax - xmin)), xmax)
    # convex separable approximation: p0, q0, P, Q, b
    xmami = np.maximum(xmax - xmin, 1e-5 * een)
    xmamiinv = een / xmami
    ux1 = upp - xval
    ux2 = ux1 * ux1
    xl1 = xval - low
    xl2 = xl1 * xl1
    uxinv = een / ux1
    xlinv = een / xl1
    p0 = np.maximum(df0dx, 0)
    q0 = np.maximum(-df0dx, 0)
    pq0 = 0.001 * (p0 + q0) + raa0 * xmamiinv
    p0 = (p0 + pq0) * ux2
    q0 = (q0 + pq0) * xl2
    P = np.maximum(dfdx, 0)
    Q = np.maximum(-dfdx, 0)
    PQ = 0.001 * (P + Q) + raa0 * (eem @ xmamiinv.T)
    P = (P + PQ) * ux2.T
    Q = (Q + PQ) * xl2.T
    b = P @ uxinv + Q @ xlinv - fval
    xmma, ymma, zmma, lam, xsi, eta, mu, zet, s = subsolv(m, n, epsimin, low, upp, alfa, beta,
                                                          p0, q0, P, Q, a0, a, b, c, d)
    return xmma, ymma, zmma, lam, xsi, eta, mu, zet, s, low, upp


def _residual(m, n, epsi, x, y, z, lam, xsi, eta, mu, zet, s, low, upp, alfa, beta, p0, q0, P,
              Q, a0, a, b, c, d):
    ux1 = upp - x
    xl1 = x - low
    plam = p0 + P.T @ lam
    qlam = q0 + Q.T @ lam
    gvec = P @ (1.0 / ux1) + Q @ (1.0 / xl1)
    dpsidx = plam / (ux1 * ux1) - qlam / (xl1 * xl1)
    rex = dpsidx - xsi + eta
    rey = c + d * y - mu - lam
    rez = a0 - zet - a.T @ lam
    relam = gvec - a * z - y + s - b
    rexsi = xsi * (x - alfa) - epsi
    reeta = eta * (beta - x) - epsi
    remu = mu * y - epsi
    rezet = zet * z - epsi
    res = lam * s - epsi
    return np.concatenate([rex, rey, rez, relam, rexsi, reeta, remu, rezet, res])


def subsolv(m, n, epsimin, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d):
    """Primal-dual Newton method for the MMA subproblem (Svanberg 2007,
    section 5): relaxed KKT conditions with epsi -> 0 (factor 0.1 per outer
    step), Newton steps on the reduced (lam, z) system for m < n, step
    length limited to keep all slack variables positive (factor 1.01) and
    halved until the residual norm decreases."""
    een = np.ones((n, 1))
    eem = np.ones((m, 1))
    epsi = 1.0
    x = 0.5 * (alfa + beta)
    y = eem.copy()
    z = np.array([[1.0]])
    lam = eem.copy()
    xsi = np.maximum(een / (x - alfa), een)
    eta = np.maximum(een / (beta - x), een)
    mu = np.maximum(eem, 0.5 * c)
    zet = np.array([[1.0]])
    s = eem.copy()
    args = (low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d)
    while epsi > epsimin:
        residu = _residual(m, n, epsi, x, y, z, lam, xsi, eta, mu, zet, s, *args)
        residunorm = np.sqrt((residu.T @ residu).item())
        residumax = np.max(np.abs(residu))
        ittt = 0
        while residumax > 0.9 * epsi and ittt < 200:
            ittt += 1
            ux1 = upp - x
            xl1 = x - low
            ux2, xl2 = ux1 * ux1, xl1 * xl1
            ux3, xl3 = ux1 * ux2, xl1 * xl2
            plam = p0 + P.T @ lam
            qlam = q0 + Q.T @ lam
            gvec = P @ (1.0 / ux1) + Q @ (1.0 / xl1)
            GG = P * (1.0 / ux2).T - Q * (1.0 / xl2).T
            dpsidx = plam / ux2 - qlam / xl2
            delx = dpsidx - epsi / (x - alfa) + epsi / (beta - x)
            dely = c + d * y - lam - epsi / y
            delz = a0 - a.T @ lam - epsi / z
            dellam = gvec - a * z - y - b + epsi / lam
            diagx = 2.0 * (plam / ux3 + qlam / xl3) + xsi / (x - alfa) + eta / (beta - x)
            diagxinv = 1.0 / diagx
            diagy = d + mu / y
            diagyinv = 1.0 / diagy
            diaglamyi = s / lam + diagyinv
            if m < n:
                blam = dellam + dely / diagy - GG @ (delx / diagx)
                bb = np.concatenate([blam, delz])
                Alam = np.diag(diaglamyi[:, 0]) + (GG * diagxinv.T) @ GG.T
                AA = np.block([[Alam, a], [a.T, -zet / z]])
                solut = np.linalg.solve(AA, bb)
                dlam = solut[:m]
                dz = solut[m:m + 1]
                dx = -delx / diagx - (GG.T @ dlam) / diagx
            else:
                diaglamyiinv = 1.0 / diaglamyi
                dellamyi = dellam + dely / diagy
                Axx = np.diag(diagx[:, 0]) + (GG.T * diaglamyiinv.T) @ GG
                azz = zet / z + a.T @ (a / diaglamyi)
                axz = -GG.T @ (a / diaglamyi)
                bx = delx + GG.T @ (dellamyi / diaglamyi)
                bz = delz - a.T @ (dellamyi / diaglamyi)
                AA = np.block([[Axx, axz], [axz.T, azz]])
                solut = np.linalg.solve(AA, -np.concatenate([bx, bz]))
                dx = solut[:n]
                dz = solut[n:n + 1]
                dlam = (GG @ dx) / diaglamyi - dz * (a / diaglamyi) + dellamyi / diaglamyi
            dy = -dely / diagy + dlam / diagy
            dxsi = -xsi + epsi / (x - alfa) - (xsi * dx) / (x - alfa)
            deta = -eta + epsi / (beta - x) + (eta * dx) / (beta - x)
            dmu = -mu + epsi / y - (mu * dy) / y
            dzet = -zet + epsi / z - zet * dz / z
            ds = -s + epsi / lam - (s * dlam) / lam
            xx = np.concatenate([y, z, lam, xsi, eta, mu, zet, s])
            dxx = np.concatenate([dy, dz, dlam, dxsi, deta, dmu, dzet, ds])
            stmxx = np.max(-1.01 * dxx / xx)
            stmalfa = np.max(-1.01 * dx / (x - alfa))
            stmbeta = np.max(1.01 * dx / (beta - x))
            steg = 1.0 / max(stmalfa, stmbeta, stmxx, 1.0)
            old = (x, y, z, lam, xsi, eta, mu, zet, s)
            itto = 0
            resinew = 2.0 * residunorm
            while resinew > residunorm and itto < 50:
                itto += 1
                x, y, z, lam, xsi, eta, mu, zet, s = (
                    o + steg * dv for o, dv in zip(old, (dx, dy, dz, dlam, dxsi, deta, dmu, dzet, ds)))
                residu = _residual(m, n, epsi, x, y, z, lam, xsi, eta, mu, zet, s, *args)
                resinew = np.sqrt((residu.T @ residu).item())
                steg = steg / 2.0
            residunorm = resinew
            residumax = np.max(np.abs(residu))
        epsi = 0.1 * epsi
    return x, y, z, lam, xsi, eta, mu, zet, s


def mma_defaultkws(n, n_constr, ft=None):
    """optimizer/mma_utils.py:8-60."""
    return {"nconstr": n_constr, "xmin": np.zeros((n, 1)), "xmax": np.ones((n, 1)),
            "low": np.ones((n, 1)), "upp": np.ones((n, 1)), "a0": 1.0,
            "a": np.zeros((n_constr, 1)), "c": 10000 * np.ones((n_constr, 1)),
            "d": np.zeros((n_constr, 1)),
            "move": 0.1 if ft in (5, 6) else (0.05 if ft == 7 else 0.2),
            "asyinit": 0.5, "asydecr": 0.7, "asyincr": 1.2, "asymin": 0.01, "asymax": 10,
            "raa0": 1e-5, "albefa": 0.1}
