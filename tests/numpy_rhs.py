"""An independent, vectorised numpy restatement of the air RHS
(WENO-Z5 component-wise + Lax-Friedrichs splitting, second-order viscous
fluxes) for checking the CPU oracle on tiny grids.

It is written from the equations rather than from the oracle's loops: the
split fluxes are formed from the physical Euler flux, F+- = (F +- J lambda Q)/2,
WENO-Z weights follow Borges et al. (2008) with p = 2, and the viscous stress
uses whole-array central differences.  Agreement with the oracle is to
round-off, not bitwise.
"""
import numpy as np

R_U, MW_AIR, GAMMA = 8314.462618, 28.9647, 1.4
R_AIR = R_U / MW_AIR
CP = GAMMA * R_AIR / (GAMMA - 1)


def _shift(a, axis, s):
    """a[..., t + s, ...] along `axis` on the padded index (no wrap needed:
    callers slice away the invalid rim)."""
    return np.roll(a, -s, axis=axis)


def weno5z_vec(vp, vm, eps):
    """vp[r], vm[r] stacks of the 5 stencil values (upwind / downwind)."""
    def side(v, d):
        q = [(2 * v[2] + 5 * v[3] - v[4]) / 6, (-v[1] + 5 * v[2] + 2 * v[3]) / 6,
             (2 * v[0] - 7 * v[1] + 11 * v[2]) / 6]
        b = [13 / 12 * (v[2] - 2 * v[3] + v[4]) ** 2 + 0.25 * (3 * v[2] - 4 * v[3] + v[4]) ** 2,
             13 / 12 * (v[1] - 2 * v[2] + v[3]) ** 2 + 0.25 * (v[1] - v[3]) ** 2,
             13 / 12 * (v[0] - 2 * v[1] + v[2]) ** 2 + 0.25 * (v[0] - 4 * v[1] + 3 * v[2]) ** 2]
        tau = np.abs(b[0] - b[2])
        a = [d[k] * (1 + (tau / (eps + b[k])) ** 2) for k in range(3)]
        return (a[0] * q[0] + a[1] * q[1] + a[2] * q[2]) / (a[0] + a[1] + a[2])
    return side(vp, (0.3, 0.6, 0.1)) + side(vm[::-1], (0.3, 0.6, 0.1))


def rhs(cv, bv, mul, jac, met, g, eps_scale, v_ref, prandtl, viscous=True):
    """dq (5, mz, my, mx) from padded arrays shaped (planes, Z, Y, X)."""
    nz, ny, nx = cv.shape[1:]
    inner = (slice(g, nz - g), slice(g, ny - g), slice(g, nx - g))
    p, T = bv[4], bv[5]
    c = np.sqrt(GAMMA * R_AIR * T)
    dq = np.zeros((5,) + tuple(n - 2 * g for n in (nz, ny, nx)))
    axes = {0: 3, 1: 2, 2: 1}  # direction -> numpy axis of the padded (plane, z, y, x) arrays
    for d in range(3):
        ax = axes[d] - 1          # axis within a single plane (z, y, x)
        k = met[3 * d:3 * d + 3]
        U = (cv[1] * k[0] + cv[2] * k[1] + cv[3] * k[2]) / cv[0]
        lam = np.abs(U) + c * np.sqrt(k[0] ** 2 + k[1] ** 2 + k[2] ** 2)
        F = np.stack([cv[0] * U, cv[1] * U + p * k[0], cv[2] * U + p * k[1], cv[3] * U + p * k[2],
                      (cv[4] + p) * U]) * jac
        Fp = 0.5 * (F + jac * lam * cv)
        Fm = 0.5 * (F - jac * lam * cv)
        mj = k * jac
        s = np.sum((mj + np.stack([_shift(m, ax, 1) for m in mj])) ** 2, axis=0)
        e0 = 1e-20 * eps_scale * 0.25 * s
        eps = [e0, e0 * v_ref ** 2, e0 * v_ref ** 2, e0 * v_ref ** 2, e0 * v_ref ** 4]
        # face t+1/2 stored at index t
        fc = np.stack([weno5z_vec(np.stack([_shift(Fp[l], ax, r) for r in range(-2, 3)]),
                                  np.stack([_shift(Fm[l], ax, r) for r in range(-1, 4)]), eps[l])
                       for l in range(5)])
        diff = fc - np.stack([_shift(fc[l], ax, -1) for l in range(5)])
        dq -= diff[(slice(None),) + inner]
    if not viscous:
        return dq
    u = bv[1:4]
    T = bv[5]
    for d in range(3):
        ax = axes[d] - 1
        sh = lambda a, s: _shift(a, ax, s)
        # the two other directions
        others = [e for e in range(3) if e != d]
        ma = 0.5 * (met + np.stack([sh(m, 1) for m in met]))          # face-averaged metric (9)
        var = [u[0], u[1], u[2], T]
        grads = []
        for f in var:
            gcomp = [None, None, None]
            gcomp[d] = sh(f, 1) - f
            for e in others:
                axe = axes[e] - 1
                sp = lambda a, s: np.roll(a, -s, axis=axe)
                gcomp[e] = 0.25 * (sp(f, 1) - sp(f, -1) + sh(sp(f, 1), 1) - sh(sp(f, -1), 1))
            # chain rule: d/dx_c = sum_r g_r * m(r, c)
            grads.append([sum(gcomp[r] * ma[3 * r + cc] for r in range(3)) for cc in range(3)])
        (ux, uy, uz), (vx, vy, vz), (wx, wy, wz), (tx, ty, tz) = grads
        mu = 0.5 * (mul + sh(mul, 1))
        txx = mu * (4 * ux - 2 * vy - 2 * wz) / 3
        tyy = mu * (4 * vy - 2 * ux - 2 * wz) / 3
        tzz = mu * (4 * wz - 2 * ux - 2 * vy) / 3
        txy, txz, tyz = mu * (uy + vx), mu * (uz + wx), mu * (vz + wy)
        kj = [0.5 * (met[3 * d + cc] * jac + sh(met[3 * d + cc] * jac, 1)) for cc in range(3)]
        f1 = kj[0] * txx + kj[1] * txy + kj[2] * txz
        f2 = kj[0] * txy + kj[1] * tyy + kj[2] * tyz
        f3 = kj[0] * txz + kj[1] * tyz + kj[2] * tzz
        um, vm, wm = (0.5 * (q + sh(q, 1)) for q in u)
        f4 = um * f1 + vm * f2 + wm * f3 + mu / prandtl * CP * (kj[0] * tx + kj[1] * ty + kj[2] * tz)
        fv = np.stack([f1, f2, f3, f4])
        diff = fv - np.stack([sh(f, -1) for f in fv])
        dq[1:] += diff[(slice(None),) + inner]
    return dq
