"""Float64 torch restatement of the oracle's discretisation, differentiated
by torch.autograd.  Test infrastructure only: it provides the reference
gradients the hand-written adjoint is checked against (same formulas as
oracle/cdd_oracle.py, which cites the reference file:line of each)."""
import numpy as np
import torch

from oracle import cdd_oracle as O

ME = O.ME


def _k(shape):
    K = O.ktools(shape)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))
    return [t(K.kx), t(K.ky), t(K.kz)], t(K.kSq), t(K.kSqSq), t(K.kmask)


def sigma(beta, mu, nu, load_shift=None):
    if callable(load_shift):  # a load that depends on the state (uniaxial control)
        load_shift = load_shift(beta)
    shape = tuple(beta.shape[1:])
    d = len(shape)
    k, kSq0, kSqSq0, kmask = _k(shape)
    lamb = O.lame_lambda(mu, nu)
    dims = tuple(range(1, d + 1))
    kb = torch.fft.rfftn(beta, dim=dims)
    kSq = kSq0 + ME
    kSqSq = kSqSq0 + ME
    kbe = []
    for i, j in O.COMPS:
        acc = torch.zeros(kb.shape[1:], dtype=kb.dtype)
        for m, n in O.COMPS:
            T = (lamb / (lamb + 2.0 * mu) * k[i] * k[j] / kSq * (m == n)
                 + k[i] / kSq * ((k[m] * (j == n)) + (k[n] * (j == m)))
                 - 2.0 * (lamb + mu) / (lamb + 2.0 * mu) * (k[i] * k[j] * k[m] * k[n]) / kSqSq)
            acc = acc + kb[3 * m + n] * T
        kbe.append(acc * kmask)
    be = torch.fft.irfftn(torch.stack(kbe), s=shape, dim=dims) - beta
    tr = be[0] + be[4] + be[8]
    out = []
    for i, j in O.COMPS:
        s = lamb * tr * (i == j) + mu * (be[3 * i + j] + be[3 * j + i])
        if load_shift is not None:
            s = s + load_shift[3 * i + j]
        out.append(s)
    return torch.stack(out)


def _cond(x):
    return x.to(torch.float64)


def minmod3(a, b, c):
    ca, cb, cc = _cond(a > 0), _cond(b > 0), _cond(c > 0)
    cab, cbc, cac = _cond(a < b), _cond(b < c), _cond(a < c)
    cabc = ca * cb * cc
    nabc = (1 - ca) * (1 - cb) * (1 - cc)
    return (cabc * (cab * cac * a + (1 - cab) * cbc * b + (1 - cbc) * (1 - cac) * c)
            + nabc * ((1 - cab) * (1 - cac) * a + cab * (1 - cbc) * b + cbc * cac * c))


def max3(a, b, c):
    cab, cbc, cac = _cond(a > b), _cond(b > c), _cond(a > c)
    return cab * cac * a + (1 - cab) * cbc * b + (1 - cbc) * (1 - cac) * c


def min3(a, b, c):
    cab, cbc, cac = _cond(a < b), _cond(b < c), _cond(a < c)
    return cab * cac * a + (1 - cab) * cbc * b + (1 - cbc) * (1 - cac) * c


def maxn(vals):
    z = torch.zeros_like(vals[0])
    if len(vals) == 4:
        return max3(vals[0], vals[1], max3(vals[2], vals[3], z))
    return max3(vals[0], vals[1], max3(vals[2], vals[3], max3(vals[4], vals[5], max3(vals[6], vals[7], z))))


def minn(vals):
    z = torch.zeros_like(vals[0])
    if len(vals) == 4:
        return min3(vals[0], vals[1], min3(vals[2], vals[3], z))
    return min3(vals[0], vals[1], min3(vals[2], vals[3], min3(vals[4], vals[5], min3(vals[6], vals[7], z))))


def find_derivatives(f, axis):
    Dx = 1.0 / f.shape[-1]
    ax = axis + 1
    roll = torch.roll
    diff = roll(f, -1, ax) - f
    d21 = (roll(diff, -1, ax) - diff) / Dx / Dx
    d22 = (roll(diff, -1, ax) - roll(diff, 1, ax)) / Dx / Dx * 0.5
    d23 = roll(d21, 1, ax)
    d2 = minmod3(d21, d22, d23)
    return diff / Dx - 0.5 * Dx * d2, roll(diff, 1, ax) / Dx + 0.5 * Dx * roll(d2, 1, ax)


def _rhomod(us):
    d = len(us)
    r2 = ME + sum((us[a] * us[a]).sum(0) for a in range(d))
    for j in range(3):
        r2 = r2 - us[0][j] ** 2 - us[1][3 + j] ** 2 - 2 * us[0][3 + j] * us[1][j]
        if d == 3:
            r2 = r2 - us[2][6 + j] ** 2 - 2 * us[1][6 + j] * us[2][3 + j] - 2 * us[2][j] * us[0][6 + j]
    return torch.sqrt(r2)


def velocity(us, sig, Lambda, lvp):
    d = len(us)
    v = [torch.zeros_like(sig[0]) for _ in range(3)]
    for n in range(3):
        for l in range(3):
            for a in range(d):
                v[l] = v[l] + us[a][3 * l + n] * sig[3 * a + n]
            for a in range(d):
                v[a] = v[a] - us[a][3 * l + n] * sig[3 * l + n]
    if not lvp and Lambda is not None:
        st = (sig[0] + sig[4] + sig[8]) / 3.0 * Lambda
        for a in range(d):
            v[a] = v[a] + st * (us[a][0] + us[a][4] + us[a][8])
        for l in range(3):
            acc = sum(us[a][3 * l + a] for a in range(d))
            v[l] = v[l] - st * acc
    if lvp:
        tr = [us[a][0] + us[a][4] + us[a][8] for a in range(d)] + [torch.zeros_like(sig[0])] * (3 - d)
        R = [sum(us[a][3 * n + a] for a in range(d)) for n in range(3)]
        rr = ME + sum(tr[a] * tr[a] - 2 * tr[a] * R[a] for a in range(3)) + sum(R[n] * R[n] for n in range(3))
        q = [tr[a] - R[a] for a in range(3)]
        rv = ME + sum(v[a] * q[a] for a in range(3))
        v = [v[a] + q[a] * (rv / rr) for a in range(3)]
    rm = _rhomod(us)
    return [vl / rm for vl in v]


def hamiltonian(us, V, Lambda, lvp):
    d = len(us)
    H = [sum(us[a][c] * V[a] for a in range(d)) for c in range(9)]
    for j in range(3):
        for l in range(3):
            for a in range(d):
                H[3 * a + j] = H[3 * a + j] - V[l] * us[a][3 * l + j]
    if not lvp and Lambda is not None:
        q = []
        for a in range(3):
            qa = sum(us[a][4 * b] for b in range(3) if b != a) if a < d else torch.zeros_like(V[0])
            qa = qa - sum(us[b][3 * a + b] for b in range(d) if b != a)
            q.append(qa)
        J = -Lambda / 3.0 * sum(V[a] * q[a] for a in range(3))
        for i in range(3):
            H[4 * i] = H[4 * i] + J
    return torch.stack(H)


def flux(beta, sig, Lambda, lvp):
    d = beta.dim() - 1
    dv = [find_derivatives(beta, a) for a in range(d)]
    corners = [tuple((c >> (d - 1 - a)) & 1 for a in range(d)) for c in range(1 << d)]
    us = {c: [dv[a][c[a]] for a in range(d)] for c in corners}  # 0 = '+', 1 = '-'
    V = {c: velocity(us[c], sig, Lambda, lvp) for c in corners}
    sp = [maxn([V[c][a] for c in corners]) for a in range(d)]
    sm = [-minn([V[c][a] for c in corners]) for a in range(d)]
    at = [sp[a] + sm[a] + ME for a in range(d)]
    deriv = 0
    for c in corners:
        w = 1
        for a in range(d):
            w = w * (sm[a] if c[a] == 0 else sp[a])
        deriv = deriv - hamiltonian(us[c], V[c], Lambda, lvp) * w
    for a in range(d):
        deriv = deriv / at[a]
    for a in range(d):
        deriv = deriv + (dv[a][0] - dv[a][1]) * sm[a] * sp[a] / at[a]
    return deriv


def uniaxial_shift(t, rate, init_strain, mu, nu):
    """RunLoadingMPI.py:68-83: sigma += (lam d_ij + 2 mu d_iz d_jz) betaE_zz with
    betaE_zz = rate t - (<beta_zz> - init), as a function of the state."""
    lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
    w = torch.tensor([lam, 0, 0, 0, lam, 0, 0, 0, lam + 2.0 * mu], dtype=torch.float64)

    def shift(beta):
        bzz = rate * t - (beta[8].mean() - init_strain)
        return w * bzz
    return shift


def mean_stress(beta, mu, nu, load_shift=None):
    """Mean of the loaded stress, components xx yy zz xy xz yz."""
    s = sigma(beta, mu, nu, load_shift)
    return torch.stack([s[c].mean() for c in (0, 4, 8, 1, 2, 5)])


def energy(beta, mu, nu, load_shift=None):
    s = sigma(beta, mu, nu, load_shift)
    e = 0
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for m in range(3):
                    S = (2.0 * (i == k) * (j == m) - (i == m) * (j == k) - nu * (i == j) * (k == m) / (1.0 + nu)) / (2.0 * mu)
                    if S != 0.0:
                        e = e + 0.5 * S * s[3 * i + j] * s[3 * k + m]
    return e.sum()


def rk_step(beta, dt, Lambda, lvp, mu, nu, order, load_shift=None):
    L = lambda b: flux(b, sigma(b, mu, nu, load_shift), Lambda, lvp)
    F0 = beta + dt * L(beta)
    if order == 2:
        return 0.5 * beta + 0.5 * (F0 + dt * L(F0))
    F1 = 0.75 * beta + 0.25 * (F0 + dt * L(F0))
    return beta / 3.0 + 2.0 / 3.0 * (F1 + dt * L(F1))
