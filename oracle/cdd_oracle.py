"""CPU oracle for the continuum-dislocation-dynamics (CDD) stepper.

TEST INFRASTRUCTURE ONLY.  This module is a NumPy float64 restatement of the
reference Python path of mattbierbaum/cuda-plasticity, written to check the
HIP implementation.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path never
does, and never falls back to it.

Every function cites the reference file:line it restates (paths relative to
the reference repository root).  Fields are NumPy arrays of shape
``(9, N0, N1[, N2])``; component ``c = 3*i + j`` stands for (i, j) with
x, y, z = 0, 1, 2 -- the sorted component order of
``Plasticity/Fields/Fields.py:21,247-249``.

Deliberate deviations (none change the arithmetic of the restated path):
  * the reference is Python 2; integer ``/`` in shapes is written ``//``;
  * ``Field`` dictionaries are replaced by the (9, ...) array;
  * the reference mask forms of MinMod3/Max/Min are kept literally.

Pinning status (tests/test_oracle_*.py): every function below is pinned by
closed forms, identities or brute force, listed in each docstring.  The
TVD-RK2 stepper has no reference counterpart (the reference only has RK3)
and is pinned by its closed-form amplification factor only.
"""
import numpy as np

# Plasticity/NumericalMethods.py:4
ME = np.finfo(float).eps

X, Y, Z = 0, 1, 2
COMPS = [(i, j) for i in range(3) for j in range(3)]


def cidx(i, j):
    return 3 * i + j


# Plasticity/Constants.py:16-22 (Levi-Civita symbol)
PERM = np.zeros((3, 3, 3))
PERM[0, 1, 2] = PERM[1, 2, 0] = PERM[2, 0, 1] = 1.0
PERM[0, 2, 1] = PERM[2, 1, 0] = PERM[1, 0, 2] = -1.0


def lame_lambda(mu, nu):
    """lambda = 2 mu nu / (1 - 2 nu)  (Plasticity/PlasticityStates/PlasticityState.py:123)."""
    return 2.0 * mu * nu / (1.0 - 2.0 * nu)


# ---------------------------------------------------------------------------
# k-space tools: Plasticity/GridArray/FourierSpaceTools.py:5-61
# ---------------------------------------------------------------------------
def _identity(a):
    # FourierSpaceTools.py:5-7 -- theta - 2 pi [theta > pi]; Nyquist stays +pi
    greater = (a > np.pi).astype(float)
    return a - 2 * np.pi * greater


class KTools:
    """k-vector grids on the half spectrum (FourierSpaceTools.py:13-61).

    2D uses ``gridShape[1]`` as the scale of both axes (:29-30); 3D scales
    each axis by its own size (:47-49).  ``kz`` is zero in 2D (:31).
    ``func``: the wrap ``_identity`` (the state's ``ktools``) or ``np.sin``
    (its ``sinktools``, PlasticityState.py:35-36).  ``kmask`` does not depend
    on it.  Pinned by: closed forms k = 2 pi n_wrapped and N sin(2 pi n / N)
    (tests/test_oracle_fft_sigma.py, tests/test_oracle_sink.py).
    """

    def __init__(self, shape, func=None):
        func = _identity if func is None else func
        shape = tuple(shape)
        self.shape = shape
        d = len(shape)
        if d == 2:
            n0, n1 = shape
            ks = (n0, n1 // 2 + 1)
            ix, iy = np.meshgrid(np.arange(ks[0]), np.arange(ks[1]), indexing="ij")
            self.kx = n1 * func(2.0 * np.pi * ix / float(n0))
            self.ky = n1 * func(2.0 * np.pi * iy / float(n1))
            self.kz = np.zeros(ks)
            self.kSq = self.kx * self.kx + self.ky * self.ky
            self.kmask = 1.0 - ((ix == 0) * (iy == 0)).astype(float)
        elif d == 3:
            n0, n1, n2 = shape
            ks = (n0, n1, n2 // 2 + 1)
            ix, iy, iz = np.meshgrid(np.arange(ks[0]), np.arange(ks[1]),
                                     np.arange(ks[2]), indexing="ij")
            self.kx = n0 * func(2.0 * np.pi * ix / float(n0))
            self.ky = n1 * func(2.0 * np.pi * iy / float(n1))
            self.kz = n2 * func(2.0 * np.pi * iz / float(n2))
            self.kSq = self.kx * self.kx + self.ky * self.ky + self.kz * self.kz
            self.kmask = 1.0 - ((ix == 0) * (iy == 0) * (iz == 0)).astype(float)
        else:
            raise ValueError("only 2D and 3D grids are on the stepper path")
        self.kSqSq = self.kSq * self.kSq
        self.k = [self.kx, self.ky, self.kz]


_KT = {}


def ktools(shape, type="k"):
    """Per-shape cache (PlasticityState.py:15-16,35-39): type 'k' the state's
    ktools, 'sinK' its sinktools (FourierSpaceTools(gridShape, func=sin))."""
    shape = tuple(shape)
    if type not in ("k", "sinK"):
        raise ValueError("type must be 'k' or 'sinK'")
    key = (shape, type)
    if key not in _KT:
        _KT[key] = KTools(shape, None if type == "k" else np.sin)
    return _KT[key]


# GridArray.py:40-44 (numpy.fft; irfftn with s=gridShape)
def rfftn(a):
    return np.fft.rfftn(a)


def irfftn(a, shape):
    return np.fft.irfftn(a, s=shape, axes=tuple(range(len(shape))))


def fft_field(beta):
    """Field.FFT: one rfftn per component (Fields.py:52-62)."""
    return np.stack([rfftn(beta[c]) for c in range(beta.shape[0])])


def ifft_field(kf, shape):
    """Field.IFFT (Fields.py:64-74)."""
    return np.stack([irfftn(kf[c], shape) for c in range(kf.shape[0])])


# ---------------------------------------------------------------------------
# Stress: PlasticityState.py:85-230, 357-390
# ---------------------------------------------------------------------------
def beta_elastic(beta, mu, nu):
    """CalculateBetaElastic(type='k') (PlasticityState.py:357-390).

    T_ijmn = lam/(lam+2mu) k_i k_j/K2 d_mn + k_i/K2 (k_m d_jn + k_n d_jm)
             - 2(lam+mu)/(lam+2mu) k_i k_j k_m k_n / K4,
    K2 = kSq + ME, K4 = kSqSq + ME; betaE = IFFT(kmask * T.betaP^) - betaP.
    Pinned by: constant / compatible / antisymmetric / single-mode closed
    forms and agreement with the 'betaP' (M-tensor) source.
    """
    shape = beta.shape[1:]
    K = ktools(shape)
    k = K.k
    lamb = lame_lambda(mu, nu)
    kb = fft_field(beta)
    kSq = K.kSq + ME
    kSqSq = K.kSqSq + ME
    kbe = np.zeros_like(kb)
    for i, j in COMPS:
        acc = np.zeros(kb.shape[1:], dtype=complex)
        for m, n in COMPS:
            T = (lamb / (lamb + 2.0 * mu) * k[i] * k[j] / kSq * (m == n)
                 + k[i] / kSq * ((k[m] * (j == n)) + (k[n] * (j == m)))
                 - 2.0 * (lamb + mu) / (lamb + 2.0 * mu) * (k[i] * k[j] * k[m] * k[n]) / kSqSq)
            acc = acc + kb[cidx(m, n)] * T
        kbe[cidx(i, j)] = acc * K.kmask
    return ifft_field(kbe, shape) - beta


def sigma(beta, mu=1.0, nu=0.3):
    """CalculateSigma(type='k', source='betaE') (PlasticityState.py:118-127).

    sigma_ij = lam tr(betaE) d_ij + mu (betaE_ij + betaE_ji).
    """
    be = beta_elastic(beta, mu, nu)
    lamb = lame_lambda(mu, nu)
    tr = be[0] + be[4] + be[8]
    out = np.zeros_like(beta)
    for i, j in COMPS:
        out[cidx(i, j)] = out[cidx(i, j)] + lamb * tr * (i == j) + mu * (be[cidx(i, j)] + be[cidx(j, i)])
    return out


def sigma_betaP_source(beta, mu=1.0, nu=0.3, type="k"):
    """CalculateSigma(type, source='betaP') (PlasticityState.py:151-182).

    Independent formula (the M_ijmn tensor, also written out in
    plasticity_kernel.cu:891-908) used to cross-check ``sigma``; type 'sinK'
    takes the k vectors of the sinktools (the mask stays the ktools' one).
    Pinned by: agreement with ``sigma`` (type 'k') and with the 'rho' source
    plus the mean terms it removes (both types).
    """
    shape = beta.shape[1:]
    K = ktools(shape, type)
    k = K.k
    kb = fft_field(beta)
    kSq = K.kSq + ME
    kSqSq = K.kSqSq + ME
    out_k = np.zeros_like(kb)
    for i, j in COMPS:
        acc = np.zeros(kb.shape[1:], dtype=complex)
        for m, n in COMPS:
            M = ((2.0 * mu * nu / (1.0 - nu)) * ((k[m] * k[n] * (i == j) + k[i] * k[j] * (m == n)) / kSq - (i == j) * (m == n))
                 - mu * ((i == m) * (j == n) + (i == n) * (j == m))
                 - (2.0 * mu / (1.0 - nu)) * k[i] * k[j] * k[m] * k[n] / kSqSq
                 + mu * (k[i] * k[n] * (j == m) + k[i] * k[m] * (j == n) + k[j] * k[n] * (i == m) + k[j] * k[m] * (i == n)) / kSq)
            acc = acc + M * kb[cidx(m, n)]
        out_k[cidx(i, j)] = acc * K.kmask
    s = ifft_field(out_k, shape)
    lamb = lame_lambda(mu, nu)
    tr = beta[0].mean() + beta[4].mean() + beta[8].mean()
    for i, j in COMPS:
        s[cidx(i, j)] -= lamb * tr * (i == j) + mu * (beta[cidx(i, j)].mean() + beta[cidx(j, i)].mean())
    return s


def rho_fourier(beta, type="k"):
    """CalculateRhoFourier(type) (PlasticityState.py:264-304).

    rho_ij = -eps_ilm d_l beta_mj via i k (type 'k') or i N sin(2 pi n / N)
    (type 'sinK').  Pinned by rho(grad u) = 0, div rho = 0 and the
    single-mode closed form (type 'k'); for 'sinK', by its equality with
    ``rho_symmetric`` (central differences) and the single-mode factor
    sin(theta) / theta.
    """
    shape = beta.shape[1:]
    d = len(shape)
    K = ktools(shape, type)
    kb = fft_field(beta)
    D = [kb * 1.0j * K.kx, kb * 1.0j * K.ky]
    if d == 3:
        D.append(kb * 1.0j * K.kz)
    kr = np.zeros_like(kb)
    for i in range(3):
        if d == 2:
            kr[cidx(X, i)] = -D[1][cidx(Z, i)]
            kr[cidx(Y, i)] = D[0][cidx(Z, i)]
            kr[cidx(Z, i)] = D[1][cidx(X, i)] - D[0][cidx(Y, i)]
        else:
            kr[cidx(X, i)] = D[2][cidx(Y, i)] - D[1][cidx(Z, i)]
            kr[cidx(Y, i)] = D[0][cidx(Z, i)] - D[2][cidx(X, i)]
            kr[cidx(Z, i)] = D[1][cidx(X, i)] - D[0][cidx(Y, i)]
    return ifft_field(kr, shape)


def sigma_rho_source(beta, mu=1.0, nu=0.3, type="k"):
    """CalculateSigma(type, source='rho') (PlasticityState.py:128-150): the
    free-boundary stress from the Nye tensor,
      sigma_ij(k) = sum_mn K_ijmn(k) rho_mn(k),
      K_ijmn = sum_s (-i mu k_s / K2) (eps_ins d_jm + eps_jns d_im
               + 2 eps_mns / (1 - nu) (k_i k_j / K2 - d_ij)),  K2 = kSq + ME,
    times the ktools' kmask, with rho = CalculateRhoFourier(type) in real
    space transformed again.  Pinned by: equal to ``sigma`` plus the mean
    terms lam tr<beta> d_ij + mu (<beta_ij> + <beta_ji>) that the fixed-volume
    sources keep (type 'k'), and to ``sigma_betaP_source`` without its mean
    correction (type 'sinK'), on band-limited fields.
    """
    shape = beta.shape[1:]
    K = ktools(shape, type)
    k = K.k
    kSq = K.kSq + ME
    krho = fft_field(rho_fourier(beta, type))
    ks = np.zeros_like(krho)
    for i, j in COMPS:
        if j < i:
            continue
        acc = np.zeros(krho.shape[1:], dtype=complex)
        for m, n in COMPS:
            for s_ in range(3):
                Kijmn = ((-1.0j) * mu * k[s_] / kSq) * (PERM[i, n, s_] * (j == m) + PERM[j, n, s_] * (i == m)
                                                        + (2.0 * PERM[m, n, s_] / (1.0 - nu)) * (k[i] * k[j] / kSq - (i == j)))
                acc = acc + Kijmn * krho[cidx(m, n)]
        ks[cidx(i, j)] = acc * ktools(shape).kmask
        ks[cidx(j, i)] = ks[cidx(i, j)]
    return ifft_field(ks, shape)


def rho_symmetric(beta):
    """CalculateRhoSymmetric (PlasticityState.py:232-262): the Nye tensor from
    SymmetricDerivative (NumericalMethods.py:6-13, order 1),
    D_a f = 0.5 N (f(i+1) - f(i-1)) along axis a with N = gridShape[-1].
    Pinned by: the Levi-Civita form -eps_ikl D_k beta_lj on the same
    differences, and equality with rho_fourier(type='sinK') (the central
    difference of a Fourier mode is i N sin(2 pi n / N) times it)."""
    shape = beta.shape[1:]
    d = len(shape)
    N = shape[-1]
    D = [0.5 * N * (np.roll(beta, -1, axis=1 + a) - np.roll(beta, 1, axis=1 + a)) for a in range(d)]
    r = np.zeros_like(beta)
    for i in range(3):
        if d == 2:
            r[cidx(X, i)] = -D[1][cidx(Z, i)]
            r[cidx(Y, i)] = D[0][cidx(Z, i)]
            r[cidx(Z, i)] = D[1][cidx(X, i)] - D[0][cidx(Y, i)]
        else:
            r[cidx(X, i)] = D[2][cidx(Y, i)] - D[1][cidx(Z, i)]
            r[cidx(Y, i)] = D[0][cidx(Z, i)] - D[2][cidx(X, i)]
            r[cidx(Z, i)] = D[1][cidx(X, i)] - D[0][cidx(Y, i)]
    return r


def elastic_energy_density(sig, mu=1.0, nu=0.3):
    """CalculateElasticEnergy (PlasticityState.py:317-330).

    e = 1/2 S_ijkm sigma_ij sigma_km,
    S = (2 d_ik d_jm - d_im d_jk - nu/(1+nu) d_ij d_km) / (2 mu).
    """
    e = np.zeros(sig.shape[1:])
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for m in range(3):
                    S = (2.0 * (i == k) * (j == m) - (i == m) * (j == k) - nu * (i == j) * (k == m) / (1.0 + nu)) / (2.0 * mu)
                    if S != 0.0:
                        e = e + 0.5 * S * sig[cidx(i, j)] * sig[cidx(k, m)]
    return e


def elastic_energy(beta, mu=1.0, nu=0.3):
    """Total energy = plain sum of the density (Observers/Observer.py:86,99)."""
    return elastic_energy_density(sigma(beta, mu, nu), mu, nu).sum()


# ---------------------------------------------------------------------------
# Limiter and selections: NumericalMethods.py:352-497
# ---------------------------------------------------------------------------
def minmod3(a, b, c):
    """MinMod3 arithmetic-mask form (NumericalMethods.py:367-378)."""
    cond_a = (a > 0)
    cond_b = (b > 0)
    cond_c = (c > 0)
    cond_ab = (a < b)
    cond_bc = (b < c)
    cond_ac = (a < c)
    cond_abc = cond_a * cond_b * cond_c
    cond_not_abc = (1 - cond_a) * (1 - cond_b) * (1 - cond_c)
    return (cond_abc * (cond_ab * cond_ac * a + (1 - cond_ab) * cond_bc * b + (1 - cond_bc) * (1 - cond_ac) * c)
            + cond_not_abc * ((1 - cond_ab) * (1 - cond_ac) * a + cond_ab * (1 - cond_bc) * b + cond_bc * cond_ac * c))


def max3(a, b, c):
    """Max3 (NumericalMethods.py:425-430)."""
    cond_ab = (a > b)
    cond_bc = (b > c)
    cond_ac = (a > c)
    return cond_ab * cond_ac * a + (1 - cond_ab) * cond_bc * b + (1 - cond_bc) * (1 - cond_ac) * c


def min3(a, b, c):
    """Min3 (NumericalMethods.py:462-467)."""
    cond_ab = (a < b)
    cond_bc = (b < c)
    cond_ac = (a < c)
    return cond_ab * cond_ac * a + (1 - cond_ab) * cond_bc * b + (1 - cond_bc) * (1 - cond_ac) * c


def max5(a, b, c, d, e):
    return max3(a, b, max3(c, d, e))  # NumericalMethods.py:432-433


def min5(a, b, c, d, e):
    return min3(a, b, min3(c, d, e))  # NumericalMethods.py:469-470


def max9(a, b, c, d, e, f, g, h, i):
    return max5(a, b, c, d, max5(e, f, g, h, i))  # :435-436


def min9(a, b, c, d, e, f, g, h, i):
    return min5(a, b, c, d, min5(e, f, g, h, i))  # :472-473


THETA = 1.0  # FieldDynamics/CentralUpwindHJ3D.py:9


def find_derivatives(field, axis, Dx=None):
    """FindDerivatives (FieldDynamics/CentralUpwindHJ3D.py:11-34).

    field: (ncomp, *grid); returns (deriv_p, deriv_m) with the same shape.
    ``rollfield(f, n, dir)`` is ``numpy.roll(f, n, dir)`` per component
    (NumericalMethods.py:352-356); the grid axis is ``axis + 1`` here.
    Pinned by: exactness on quadratics/linears, extremum -> 0, and the
    minmod table.
    """
    if Dx is None:
        Dx = 1.0 / field.shape[-1]
    ax = axis + 1
    roll = np.roll
    diff = roll(field, -1, ax) - field
    diff2_1 = THETA * (roll(diff, -1, ax) - diff) / Dx / Dx
    diff2_2 = (roll(diff, -1, ax) - roll(diff, 1, ax)) / Dx / Dx * 0.5
    diff2_3 = roll(diff2_1, 1, ax)
    diff2 = minmod3(diff2_1, diff2_2, diff2_3)
    deriv_p = diff / Dx - 0.5 * Dx * diff2
    deriv_m = roll(diff, 1, ax) / Dx + 0.5 * Dx * roll(diff2, 1, ax)
    return deriv_p, deriv_m


# ---------------------------------------------------------------------------
# Velocities and Hamiltonians
# ---------------------------------------------------------------------------
def _rhomod(us):
    """|rho| of one corner: BetaP{2,3}D_RhoMod
    (FieldDynamics/CentralUpwindHJBetaPDynamics3D.py:34-58)."""
    d = len(us)
    rho2 = ME
    for c in range(9):
        for a in range(d):
            rho2 = rho2 + us[a][c] * us[a][c]
    for j in range(3):
        rho2 = rho2 - us[0][cidx(X, j)] * us[0][cidx(X, j)]
        rho2 = rho2 - us[1][cidx(Y, j)] * us[1][cidx(Y, j)]
        if d == 3:
            rho2 = rho2 - us[2][cidx(Z, j)] * us[2][cidx(Z, j)]
        rho2 = rho2 - 2 * us[0][cidx(Y, j)] * us[1][cidx(X, j)]
        if d == 3:
            rho2 = rho2 - 2 * us[1][cidx(Z, j)] * us[2][cidx(Y, j)]
            rho2 = rho2 - 2 * us[2][cidx(X, j)] * us[0][cidx(Z, j)]
    return np.sqrt(rho2)


def velocity_glide_climb(us, sig, Lambda):
    """BetaP2D_Velocity / BetaP3D_Velocity with climb suppression Lambda
    (CentralUpwindHJBetaPDynamics3D.py:77-142; 2D loops :83-100).

    us: list over derivative axes of (9, *grid) one-sided derivatives.
    Returns [v_x, v_y, v_z].  Pinned by the Levi-Civita form
    v_l = sigma'_st (d_s beta_lt - d_l beta_st)/|rho| with
    sigma' = sigma - Lambda tr(sigma)/3 I.
    """
    d = len(us)
    v = [np.zeros(sig.shape[1:]) for _ in range(3)]
    for n in range(3):
        for l in range(3):
            for a in range(d):
                v[l] = v[l] + us[a][cidx(l, n)] * sig[cidx(a, n)]
            for a in range(d):
                v[a] = v[a] - us[a][cidx(l, n)] * sig[cidx(l, n)]
    if Lambda != 0:
        sigma_tr = (sig[0] + sig[4] + sig[8]) / 3.0 * Lambda
        for a in range(d):
            v[a] = v[a] + sigma_tr * (us[a][0] + us[a][4] + us[a][8])
        for l in range(3):
            acc = us[0][cidx(l, X)]
            for a in range(1, d):
                acc = acc + us[a][cidx(l, a)]
            v[l] = v[l] - sigma_tr * acc
    rm = _rhomod(us)
    return [vl / rm for vl in v]


def velocity_lvp(us, sig):
    """Glide-only ('lvp') velocity (CentralUpwindHJBetaPGlideOnlyDynamics.py:
    2D :69-128, 3D :131-202).  The code ADDS q (rhov/rhorho), the docstring
    writes a minus; this restates the code.  Pinned by rhorho - ME = |q|^2
    and the closed form of v'.q.
    """
    d = len(us)
    ux, uy = us[0], us[1]
    uz = us[2] if d == 3 else None
    # climb velocity (no division): :78-91 / :140-155
    v = [np.zeros(sig.shape[1:]) for _ in range(3)]
    for n in range(3):
        for l in range(3):
            for a in range(d):
                v[l] = v[l] + us[a][cidx(l, n)] * sig[cidx(a, n)]
            for a in range(d):
                v[a] = v[a] - us[a][cidx(l, n)] * sig[cidx(l, n)]
    if d == 2:
        rhorho = ME
        uxTr = ux[0] + ux[4] + ux[8]
        uyTr = uy[0] + uy[4] + uy[8]
        rhorho = rhorho + uxTr * uxTr + uyTr * uyTr
        rhorho = rhorho - 2 * uxTr * (ux[cidx(X, X)] + uy[cidx(X, Y)])
        rhorho = rhorho - 2 * uyTr * (ux[cidx(Y, X)] + uy[cidx(Y, Y)])
        for n in range(3):
            t = ux[cidx(n, X)] + uy[cidx(n, Y)]
            rhorho = rhorho + t * t
        rhov = ME
        rhov = rhov + v[0] * (ux[cidx(Z, Z)] + ux[cidx(Y, Y)] - uy[cidx(X, Y)])
        rhov = rhov + v[1] * (uy[cidx(X, X)] + uy[cidx(Z, Z)] - ux[cidx(Y, X)])
        rhov = rhov + -v[2] * (ux[cidx(Z, X)] + uy[cidx(Z, Y)])
        v[0] = v[0] + (ux[cidx(Y, Y)] + ux[cidx(Z, Z)] - uy[cidx(X, Y)]) * rhov / rhorho
        v[1] = v[1] + (uy[cidx(X, X)] + uy[cidx(Z, Z)] - ux[cidx(Y, X)]) * rhov / rhorho
        v[2] = v[2] - (ux[cidx(Z, X)] + uy[cidx(Z, Y)]) * rhov / rhorho
    else:
        rhorho = ME
        uxTr = ux[0] + ux[4] + ux[8]
        uyTr = uy[0] + uy[4] + uy[8]
        uzTr = uz[0] + uz[4] + uz[8]
        rhorho = rhorho + uxTr * uxTr + uyTr * uyTr + uzTr * uzTr
        rhorho = rhorho - 2 * uxTr * (ux[cidx(X, X)] + uy[cidx(X, Y)] + uz[cidx(X, Z)])
        rhorho = rhorho - 2 * uyTr * (ux[cidx(Y, X)] + uy[cidx(Y, Y)] + uz[cidx(Y, Z)])
        rhorho = rhorho - 2 * uzTr * (ux[cidx(Z, X)] + uy[cidx(Z, Y)] + uz[cidx(Z, Z)])
        for n in range(3):
            t = ux[cidx(n, X)] + uy[cidx(n, Y)] + uz[cidx(n, Z)]
            rhorho = rhorho + t * t
        rhov = ME
        rhov = rhov + v[0] * (ux[cidx(Z, Z)] + ux[cidx(Y, Y)] - uy[cidx(X, Y)] - uz[cidx(X, Z)])
        rhov = rhov + v[1] * (uy[cidx(X, X)] + uy[cidx(Z, Z)] - ux[cidx(Y, X)] - uz[cidx(Y, Z)])
        rhov = rhov + v[2] * (uz[cidx(X, X)] + uz[cidx(Y, Y)] - ux[cidx(Z, X)] - uy[cidx(Z, Y)])
        factor = rhov / rhorho
        v[0] = v[0] + (ux[cidx(Y, Y)] + ux[cidx(Z, Z)] - uy[cidx(X, Y)] - uz[cidx(X, Z)]) * factor
        v[1] = v[1] + (uy[cidx(X, X)] + uy[cidx(Z, Z)] - ux[cidx(Y, X)] - uz[cidx(Y, Z)]) * factor
        v[2] = v[2] + (uz[cidx(X, X)] + uz[cidx(Y, Y)] - ux[cidx(Z, X)] - uy[cidx(Z, Y)]) * factor
    rm = _rhomod(us)
    return [vl / rm for vl in v]


def hamiltonian(us, V, Lambda):
    """H_2D / H_3D (CentralUpwindHJBetaPDynamics3D.py:161-201; glide-only
    module :216-242 is the Lambda = 0 case).  Pinned by
    H_ij = eps_ilm v_l rho_mj and tr(H') = (1 - Lambda) tr(H).
    """
    d = len(us)
    H = us[0] * V[0]
    for a in range(1, d):
        H = H + us[a] * V[a]
    H = H.copy()
    for j in range(3):
        for l in range(3):
            for a in range(d):
                H[cidx(a, j)] = H[cidx(a, j)] - V[l] * us[a][cidx(l, j)]
    if Lambda != 0:
        if d == 2:
            ux, uy = us
            J = -Lambda / 3.0 * (V[0] * (ux[cidx(Y, Y)] + ux[cidx(Z, Z)] - uy[cidx(X, Y)])
                                 + V[1] * (uy[cidx(X, X)] + uy[cidx(Z, Z)] - ux[cidx(Y, X)])
                                 - V[2] * (ux[cidx(Z, X)] + uy[cidx(Z, Y)]))
        else:
            ux, uy, uz = us
            J = -Lambda / 3.0 * (V[0] * (ux[cidx(Y, Y)] + ux[cidx(Z, Z)] - uy[cidx(X, Y)] - uz[cidx(X, Z)])
                                 + V[1] * (uy[cidx(X, X)] + uy[cidx(Z, Z)] - ux[cidx(Y, X)] - uz[cidx(Y, Z)])
                                 + V[2] * (uz[cidx(X, X)] + uz[cidx(Y, Y)] - ux[cidx(Z, X)] - uy[cidx(Z, Y)]))
        for i in range(3):
            H[cidx(i, i)] = H[cidx(i, i)] + J
    return H


def velocity_wojzz(us, sig, Lambda):
    """BetaP2D_Velocity of the woJzz variant
    (FieldDynamics/CentralUpwindHJBetaPDynamics_woJzz.py:54-83, 2D only):
    the climb velocity plus, for Lambda != 0,
    sigma_tr = tr(sigma)/2 Lambda times (ux_yy - uy_xy, uy_xx - ux_yx,
    -(ux_zx + uy_zy)), divided by |rho|.  Pinned by: Lambda = 0 reduces to
    the glide + climb velocity; rotation equivariance of the flux."""
    ux, uy = us
    v = [np.zeros(sig.shape[1:]) for _ in range(3)]
    for n in range(3):
        for l in range(3):
            v[l] = v[l] + ux[cidx(l, n)] * sig[cidx(X, n)]
            v[l] = v[l] + uy[cidx(l, n)] * sig[cidx(Y, n)]
            v[X] = v[X] - ux[cidx(l, n)] * sig[cidx(l, n)]
            v[Y] = v[Y] - uy[cidx(l, n)] * sig[cidx(l, n)]
    if Lambda != 0:
        sigma_tr = (sig[0] + sig[4] + sig[8]) / 2. * Lambda
        v[X] = v[X] + sigma_tr * (ux[cidx(Y, Y)] - uy[cidx(X, Y)])
        v[Y] = v[Y] + sigma_tr * (uy[cidx(X, X)] - ux[cidx(Y, X)])
        v[Z] = v[Z] - sigma_tr * (ux[cidx(Z, X)] + uy[cidx(Z, Y)])
    rm = _rhomod(us)
    return [vl / rm for vl in v]


def hamiltonian_wojzz(us, V, Lambda):
    """H_2D of the woJzz variant (CentralUpwindHJBetaPDynamics_woJzz.py:95-106):
    the glide + climb Hamiltonian without its Lambda term, then
    J = Lambda/2 (V_x (ux_yy - uy_xy) + V_y (uy_xx - ux_yx) - V_z (ux_zx + uy_zy))
    subtracted from H_xx and H_yy only."""
    ux, uy = us
    H = hamiltonian(us, V, 0.0)
    if Lambda != 0:
        J = Lambda / 2. * (V[0] * (ux[cidx(Y, Y)] - uy[cidx(X, Y)]) + V[1] * (uy[cidx(X, X)] - ux[cidx(Y, X)])
                           - V[2] * (ux[cidx(Z, X)] + uy[cidx(Z, Y)]))
        for i in (X, Y):
            H[cidx(i, i)] = H[cidx(i, i)] - J
    return H


# ---------------------------------------------------------------------------
# Central-upwind flux: CentralUpwindHJ3D.py:63-271
# ---------------------------------------------------------------------------
GLIDE_CLIMB = "glide_climb"
GLIDE_ONLY = "lvp"
GLIDE_CLIMB_WOJZZ = "wojzz"  # 2D only; rhs_zz is held at 0


def _velocity(dynamics, us, sig, Lambda):
    if dynamics == GLIDE_ONLY:
        return velocity_lvp(us, sig)
    if dynamics == GLIDE_CLIMB_WOJZZ:
        return velocity_wojzz(us, sig, Lambda)
    return velocity_glide_climb(us, sig, Lambda)


def _hamiltonian(dynamics, us, V, Lambda):
    if dynamics == GLIDE_CLIMB_WOJZZ:
        return hamiltonian_wojzz(us, V, Lambda)
    return hamiltonian(us, V, Lambda)


def central_upwind_flux(beta, sig, dynamics=GLIDE_CLIMB, Lambda=0.0, Dx=None):
    """CentralUpwindHJDynamics.CalculateFlux for the BetaP dynamics
    (CentralUpwindHJ3D.py:99-188 (2D), :189-271 (3D)).  The glide-only
    dynamics has no climb-suppression term (Lambda is ignored).

    Returns (rhs, speed) with speed = a_tot + b_tot (+ c_tot).
    Pinned by: constant field -> rhs = 0 and speed = d ME; translation
    and 90-degree rotation equivariance; 3D z-invariant == 2D.
    """
    d = beta.ndim - 1
    lam = 0.0 if dynamics == GLIDE_ONLY else Lambda
    if dynamics == GLIDE_CLIMB_WOJZZ and d != 2:
        raise ValueError("the woJzz dynamics is 2D only")
    if d == 2:
        dxp, dxm = find_derivatives(beta, 0, Dx)
        dyp, dym = find_derivatives(beta, 1, Dx)
        v_mm = _velocity(dynamics, [dxm, dym], sig, lam)
        v_mp = _velocity(dynamics, [dxm, dyp], sig, lam)
        v_pm = _velocity(dynamics, [dxp, dym], sig, lam)
        v_pp = _velocity(dynamics, [dxp, dyp], sig, lam)
        a_p = max5(v_mm[0], v_mp[0], v_pm[0], v_pp[0], 0.)
        a_m = np.fabs(min5(v_mm[0], v_mp[0], v_pm[0], v_pp[0], 0.))
        b_p = max5(v_mm[1], v_mp[1], v_pm[1], v_pp[1], 0.)
        b_m = np.fabs(min5(v_mm[1], v_mp[1], v_pm[1], v_pp[1], 0.))
        a_tot = a_p + a_m + ME
        b_tot = b_p + b_m + ME
        H_pp = _hamiltonian(dynamics, [dxp, dyp], v_pp, lam)
        H_pm = _hamiltonian(dynamics, [dxp, dym], v_pm, lam)
        H_mp = _hamiltonian(dynamics, [dxm, dyp], v_mp, lam)
        H_mm = _hamiltonian(dynamics, [dxm, dym], v_mm, lam)
        rhs = (-(H_pp * a_m * b_m + H_pm * a_m * b_p + H_mp * a_p * b_m + H_mm * a_p * b_p) / a_tot / b_tot
               + ((dxp - dxm) * a_m * a_p / a_tot) + ((dyp - dym) * b_m * b_p / b_tot))
        if dynamics == GLIDE_CLIMB_WOJZZ:
            rhs[cidx(Z, Z)] = 0.  # CentralUpwindHJBetaPDynamics_woJzz.py:128-135
        return rhs, a_tot + b_tot
    dxp, dxm = find_derivatives(beta, 0, Dx)
    dyp, dym = find_derivatives(beta, 1, Dx)
    dzp, dzm = find_derivatives(beta, 2, Dx)
    dx = {'+': dxp, '-': dxm}
    dy = {'+': dyp, '-': dym}
    dz = {'+': dzp, '-': dzm}
    opts = ['---', '--+', '-+-', '-++', '+--', '+-+', '++-', '+++']
    V = {o: _velocity(dynamics, [dx[o[0]], dy[o[1]], dz[o[2]]], sig, lam) for o in opts}
    a_p = max9(*[V[o][0] for o in opts], 0.)
    a_m = np.fabs(min9(*[V[o][0] for o in opts], 0.))
    b_p = max9(*[V[o][1] for o in opts], 0.)
    b_m = np.fabs(min9(*[V[o][1] for o in opts], 0.))
    c_p = max9(*[V[o][2] for o in opts], 0.)
    c_m = np.fabs(min9(*[V[o][2] for o in opts], 0.))
    a_tot = a_p + a_m + ME
    b_tot = b_p + b_m + ME
    c_tot = c_p + c_m + ME
    wa = {'+': a_m, '-': a_p}
    wb = {'+': b_m, '-': b_p}
    wc = {'+': c_m, '-': c_p}
    # summation order of CentralUpwindHJ3D.py:243-258
    deriv = None
    for o in ['+++', '++-', '+-+', '+--', '-++', '-+-', '--+', '---']:
        H = hamiltonian([dx[o[0]], dy[o[1]], dz[o[2]]], V[o], lam)
        term = H * wa[o[0]] * wb[o[1]] * wc[o[2]]
        deriv = -term if deriv is None else deriv - term
    tempa = (dxp - dxm) * a_m * a_p / a_tot
    tempb = (dyp - dym) * b_m * b_p / b_tot
    tempc = (dzp - dzm) * c_m * c_p / c_tot
    rhs = deriv / a_tot / b_tot / c_tot + tempa + tempb + tempc
    return rhs, a_tot + b_tot + c_tot


# ---------------------------------------------------------------------------
# External loads (uniform stress shifts)
# ---------------------------------------------------------------------------
class Load:
    """Uniform stress added by the dynamics' GetSigma override.

    kind 'none'; 'strain' -- RunPython.py:21-26,44-49 (ExternalStrain with
    strain(t) = initial + rate (t - t0)); 'stress' -- RunPython.py:28-32;
    'uniaxial' -- RunLoadingMPI.py:68-83 (betaE_zz = rate t - (<beta_zz> -
    init)); 'tensor' -- sigma_ij += s0_ij + rate_ij (t - t0) (the shear of
    the adjoint workload, no reference counterpart).
    elastic: (mu, nu) of the 'strain' load's Hooke's law; RunPython's
    ExternalStrain uses its module constants (0.5, 0.3) whatever the state's
    (RunPython.py:19-26); None: the dynamics' own mu, nu.  The uniaxial
    control uses the state's (RunLoadingMPI.py:76-77).
    """

    def __init__(self, kind="none", initial=(0., 0., 0.), rate=(0., 0., 0.), t0=0.0,
                 init_strain=0.0, uniaxial_rate=0.0, tensor0=None, tensor_rate=None, elastic=None):
        self.kind = kind
        self.elastic = None if elastic is None else (float(elastic[0]), float(elastic[1]))
        self.initial = np.asarray(initial, dtype=float)
        self.rate = np.asarray(rate, dtype=float)
        self.t0 = t0
        self.init_strain = init_strain
        self.uniaxial_rate = uniaxial_rate
        self.tensor0 = np.zeros(9) if tensor0 is None else np.asarray(tensor0, float).reshape(9)
        self.tensor_rate = np.zeros(9) if tensor_rate is None else np.asarray(tensor_rate, float).reshape(9)

    def apply(self, sig, beta, t, mu, nu):
        lamb = lame_lambda(mu, nu)
        if self.kind == "none":
            return sig
        sig = sig.copy()
        if self.kind == "strain":
            smu, snu = self.elastic if self.elastic is not None else (mu, nu)
            slamb = lame_lambda(smu, snu)
            e = self.initial + self.rate * (t - self.t0)
            tr = e[0] + e[1] + e[2]
            for i in range(3):
                sig[cidx(i, i)] += slamb * tr + 2.0 * smu * e[i]
        elif self.kind == "stress":
            s = self.initial + self.rate * (t - self.t0)
            for i in range(3):
                sig[cidx(i, i)] += s[i]
        elif self.kind == "uniaxial":
            bzz = self.uniaxial_rate * t - (np.average(beta[8]) - self.init_strain)
            for i, j in COMPS:
                sig[cidx(i, j)] += (lamb * (i == j) + 2.0 * mu * (i == Z) * (j == Z)) * bzz
        elif self.kind == "tensor":
            s = self.tensor0 + self.tensor_rate * (t - self.t0)
            for c in range(9):
                sig[c] += s[c]
        else:
            raise ValueError(self.kind)
        return sig


class Dynamics:
    """BetaPDynamics.CalculateFlux = GetSigma + central-upwind flux
    (CentralUpwindHJBetaPDynamics3D.py:214-221; glide-only :255-262)."""

    def __init__(self, dynamics=GLIDE_CLIMB, Lambda=0.0, mu=1.0, nu=0.3, load=None):
        self.dynamics = dynamics
        self.Lambda = Lambda
        self.mu = mu
        self.nu = nu
        self.load = load if load is not None else Load()

    def get_sigma(self, beta, t):
        s = sigma(beta, self.mu, self.nu)
        return self.load.apply(s, beta, t, self.mu, self.nu)

    def flux(self, t, beta):
        return central_upwind_flux(beta, self.get_sigma(beta, t), self.dynamics, self.Lambda)


def instability(beta):
    """LargeKInstabilityChecking_FieldMover.CheckInstability's measure
    (Plasticity/FieldMovers/LargeKInstabilityChecking_FieldMover.py:73-89):
    rhoM = |rho| per cell (CalculateRhoFourier().modulus()), KrhoM =
    fftn(rhoM), value = |KrhoM[N/2, ..., N/2]|.  Pinned by the closed form
    sum over cells of (-1)^(i+j[+k]) rhoM."""
    r = rho_fourier(beta)
    rhoM = np.sqrt((r * r).sum(axis=0))
    K = np.fft.fftn(rhoM)
    mid = tuple(n // 2 for n in rhoM.shape)
    return float(np.sqrt(K[mid].real ** 2 + K[mid].imag ** 2))


def mean_stress(beta, t, mu=1.0, nu=0.3, load=None):
    """Mean stress of the state at time t, external load included: the
    stress-strain point of PlottingTools.StrainStress (PlottingTools.py:69-100:
    sigma(source='betaE') averaged, plus lam tr(e) d + 2 mu e of the imposed
    strain) and, for the uniaxial strain-rate control, the mean of the
    loaded sigma of RunLoadingMPI.py:68-83.  Returns (xx, yy, zz, xy, xz, yz)."""
    dyn = Dynamics(mu=mu, nu=nu, load=load)
    s = dyn.get_sigma(beta, t)
    return np.array([s[cidx(i, j)].mean() for i, j in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))])


# ---------------------------------------------------------------------------
# Time stepping: FieldMovers/FieldMover.py:14-141
# ---------------------------------------------------------------------------
def cfl_dt(speed, shape, cfl):
    """dt = CFL / (max(gridShape) * max speed) (FieldMover.py:94-96)."""
    return cfl / (max(shape) * speed.max())


def tvd_rk3_step(flux, t, u, t_end, cfl, fixed_dt=None, dt_bound=None):
    """CalculateOneAdaptiveTimeStep (FieldMover.py:61-141) or, with fixed_dt,
    CalculateOneFixedTimeStep (:38-58).  flux(t, u) -> (rhs, speed).
    Every stage is evaluated at the step's start time (:123,133).
    dt_bound caps the adaptive step (the operator-splitting mover,
    FieldMover.py:359-360; the plain mover has it commented out, :107-108).
    Returns (u_new, dt).  Pinned by the RK3 amplification factor.
    """
    rhs, speed = flux(t, u)
    dt = fixed_dt if fixed_dt is not None else cfl_dt(speed, u.shape[1:], cfl)
    if fixed_dt is None and dt_bound is not None and dt > dt_bound:
        dt = dt_bound
    if t + dt > t_end:
        dt = t_end - t
    L0 = dt * rhs
    F0 = 1.0 * u + 1.0 * L0
    F2 = 1.0 / 3.0 * u + 0. * L0 + 0. * F0
    rhs, _ = flux(t, F0)
    L1 = dt * rhs
    F1 = 3.0 / 4.0 * u + 0. * L0 + 1.0 / 4.0 * F0 + 1.0 / 4.0 * L1
    F2 = F2 + 0. * L1 + 2.0 / 3.0 * F1
    rhs, _ = flux(t, F1)
    L2 = dt * rhs
    F2 = F2 + 2.0 / 3.0 * L2
    return F2, dt


def tvd_rk2_step(flux, t, u, t_end, cfl, fixed_dt=None, dt_bound=None):
    """SSP (TVD) RK2 -- no reference counterpart; same CFL rule and stage-time
    convention as the reference RK3 (FieldMover.py:94-111,123):
    u1 = u + dt L(u);  u' = u/2 + (u1 + dt L(u1))/2.
    Pinned by the closed-form amplification 1 + z + z^2/2.
    """
    rhs, speed = flux(t, u)
    dt = fixed_dt if fixed_dt is not None else cfl_dt(speed, u.shape[1:], cfl)
    if fixed_dt is None and dt_bound is not None and dt > dt_bound:
        dt = dt_bound
    if t + dt > t_end:
        dt = t_end - t
    u1 = u + dt * rhs
    rhs, _ = flux(t, u1)
    return 0.5 * u + 0.5 * (u1 + dt * rhs), dt


def run(dyn, beta, t0, nsteps, cfl, order=3, t_end=np.inf, fixed_dt=None, observer=None):
    """Mover.Run restricted to a fixed number of steps (FieldMover.py:20-36):
    t += dt after each step, observers called after every step."""
    step = tvd_rk3_step if order == 3 else tvd_rk2_step
    t = t0
    u = beta
    dts = []

    def f(tt, uu):
        return dyn.flux(tt, uu)

    for _ in range(nsteps):
        u, dt = step(f, t, u, t_end, cfl, fixed_dt)
        t += dt
        dts.append(dt)
        if observer is not None:
            observer(t, u)
        if t >= t_end:
            break
    return u, t, np.array(dts)


# ---------------------------------------------------------------------------
# Vacancy coupling: BetaP_VacancyDynamics (FieldDynamics/VacancyDynamics.py:
# 16-70) advanced by OperatorSplittingTVDRK_FieldMover (FieldMovers/
# FieldMover.py:318-380)
# ---------------------------------------------------------------------------
class Vacancy:
    """alpha: chemical stress sigma_ii -= alpha c (GetSigma, :24-28); beta:
    c += beta (tr beta_new - tr beta_old) after a step (CalculateVFlux,
    :37-39); gamma: diffusion constant (DiffusionEvolution, :40-70);
    dt_bound: the mover's dtBound (FieldMover.py:359-360), None for none."""

    def __init__(self, alpha=1.0, beta=1.0, gamma=1.0, dt_bound=None):
        self.alpha = alpha
        self.beta = beta
        self.gamma = gamma
        self.dt_bound = dt_bound


def vacancy_diffuse(c, dt, gamma):
    """DiffusionEvolution (VacancyDynamics.py:40-70):
    irfftn(rfftn(c) exp(kSq * -gamma * dt)).  Pinned by: a single Fourier
    mode decays by exp(-gamma dt k^2) exactly, the mean is kept
    (tests/test_oracle_vacancy.py)."""
    kt = ktools(c.shape)
    kc = rfftn(c)
    k = kt.kSq * -gamma * dt
    return irfftn(kc * np.exp(k), c.shape)


def vacancy_sigma(dyn, vac, beta, c, t):
    """GetSigma of the vacancy dynamics (VacancyDynamics.py:24-28; the loaded
    variant RunPython.py:93-101 adds the load afterwards): sigma of the
    state, sigma_ii -= alpha c, then the load."""
    s = sigma(beta, dyn.mu, dyn.nu).copy()
    for i in range(3):
        s[cidx(i, i)] = s[cidx(i, i)] - vac.alpha * c
    return dyn.load.apply(s, beta, t, dyn.mu, dyn.nu)


def run_vacancy(dyn, vac, beta, c, t0, nsteps, cfl, order=3, t_end=np.inf, fixed_dt=None):
    """OperatorSplittingTVDRK_FieldMover.Run restricted to nsteps steps
    (FieldMover.py:20-36 loop, CalculateOneAdaptiveTimeStep :349-380): every
    stage's flux sees the stress of vacancy_sigma with the concentration of
    the step start; dt from the first stage's speeds capped at dtBound and
    clamped to the end time; after the step c += beta (tr F2 - tr u), then c
    diffuses over dt.  The fixed-step variant (:322-348) recouples the
    step-start beta^P (apparently unintended: its F_2 is commented out);
    here both variants keep the advanced state.  Returns (u, c, t, dts).
    Pinned by tests/test_oracle_vacancy.py (diffusion closed form, stress
    shift, trace balance, Lambda = 1 trace conservation, dt cap)."""
    if np.shape(c) != np.shape(beta)[1:]:
        raise ValueError("c must have the grid shape")
    step = tvd_rk3_step if order == 3 else tvd_rk2_step
    t, u, dts = t0, beta, []
    for _ in range(nsteps):
        cs = c

        def f(tt, uu):
            return central_upwind_flux(uu, vacancy_sigma(dyn, vac, uu, cs, tt), dyn.dynamics, dyn.Lambda)

        prev = u[cidx(X, X)] + u[cidx(Y, Y)] + u[cidx(Z, Z)]
        u, dt = step(f, t, u, t_end, cfl, fixed_dt, dt_bound=vac.dt_bound)
        tr = u[cidx(X, X)] + u[cidx(Y, Y)] + u[cidx(Z, Z)]
        c = c + vac.beta * (tr - prev)
        c = vacancy_diffuse(c, dt, vac.gamma)
        t += dt
        dts.append(dt)
        if t >= t_end:
            break
    return u, c, t, np.array(dts)


# ---------------------------------------------------------------------------
# Disorder: the stress of a disorder field d coupled to the dislocation
# density (BetaP_DisorderDynamics.GetSigma, FieldDynamics/DisorderDynamics.py:
# 22-39, with d held by the dynamics; DisorderState.CalculateSigma,
# PlasticityStates/DisorderState.py:15-30, with d the state's ('s', 's')
# component, the RunPython.py:196-214 'x' runs).
# ---------------------------------------------------------------------------
def disorder_weight(rho, d):
    """rho'_ij = d rho_ij / rhomod, rhomod = ME + sum_mn 2 rho_mn rho_mn
    accumulated in component order (DisorderState.py:19-25)."""
    rhomod = ME
    for c in range(9):
        rhomod = rhomod + 2 * rho[c] * rho[c]
    return np.stack([d * rho[c] / rhomod for c in range(9)])


def disorder_curl(rhop, sig=None):
    """sigma_ij += sum_{l,t} perm[l,i,t] irfftn(k_l rfftn(rho'_tj)) in the
    loop order of DisorderState.py:26-29 (l = z skipped in 2D).  The k-space
    factor is the real k_l, not i k_l, and irfftn keeps only the real part of
    the kz = 0 and Nyquist planes, as written: along the last axis the
    operator is |d/dz|, and along another axis a mode contributes only at
    its Nyquist wavenumber.  Pinned by those single-mode closed forms
    (tests/test_oracle_disorder.py)."""
    shape = rhop.shape[1:]
    d = len(shape)
    kt = ktools(shape)
    out = np.zeros((9,) + shape) if sig is None else sig.copy()
    F = [rfftn(rhop[c]) for c in range(9)]
    for i in range(3):
        for j in range(3):
            for l in range(3):
                for t in range(3):
                    if PERM[l, i, t] != 0 and (d == 3 or l != Z):
                        out[cidx(i, j)] = out[cidx(i, j)] + PERM[l, i, t] * irfftn(kt.k[l] * F[cidx(t, j)], shape)
    return out


def disorder_sigma(beta, d, mu, nu):
    """DisorderState.CalculateSigma (DisorderState.py:15-30): the state's
    stress (source 'betaE') plus the disorder term of disorder_weight(rho, d)
    with rho the spectral Nye tensor.  beta: the nine tensor components."""
    return disorder_curl(disorder_weight(rho_fourier(beta[:9]), d), sigma(beta[:9], mu, nu))


class DisorderDynamics(Dynamics):
    """The BetaP dynamics with the disorder stress.  disorder=None: d is the
    state's tenth component u[9], advected by the flux with the Hamiltonian
    H_s = sum_a v_a d_a (H_2D/H_3D apply deriv_a * V[a] to every component of
    the state's field, CentralUpwindHJBetaPDynamics3D.py:186-188, and the
    RK combinations to all ten); else d is the given fixed field
    (BetaP_DisorderDynamics) and the state has nine components.  A load,
    which the reference's disorder classes do not apply, is added after the
    disorder term as ExternalStrain/ExternalStress would be."""

    def __init__(self, dynamics=GLIDE_CLIMB, Lambda=0.0, mu=1.0, nu=0.3, load=None, disorder=None):
        Dynamics.__init__(self, dynamics, Lambda, mu, nu, load)
        self.disorder = disorder

    def field_of(self, u):
        return self.disorder if self.disorder is not None else u[9]

    def get_sigma(self, u, t):
        s = disorder_sigma(u[:9], self.field_of(u), self.mu, self.nu)
        return self.load.apply(s, u[:9], t, self.mu, self.nu)

    def flux(self, t, u):
        return central_upwind_flux(u, self.get_sigma(u, t), self.dynamics, self.Lambda)


# ---------------------------------------------------------------------------
# The CUDA executable's numerics (plasticity_kernel.cu centralHJ :118-697,
# findDerivatives :53-90, simpleTVD :1112-1137, TVD3rd :1140-1206, with the
# defaults of plasticity.h: theta = 1, LLF, PYTHON_COMPATIBILITY_DIVIDE, DOUBLE;
# RunCUDA.py:54-81 adds lambda (gcd 0, mdp 1, lvp 0 + NEWGLIDEONLY) and
# CFLsafeFactor 0.5).  Same physics as the Python path, different numerics:
# one local Lax-Friedrichs speed per cell, |rho| + sqrt(ME) in place of
# sqrt(|rho|^2 + ME), its own second-difference form, forward Euler.
# ---------------------------------------------------------------------------
def cuda_minmod3(a, b, c):
    """minmod3 of plasticity_kernel.cu:41-51 (min if all > 0, max if all < 0,
    else 0)."""
    mn = np.minimum(np.minimum(a, b), c)
    mx = np.maximum(np.maximum(a, b), c)
    return np.where((a > 0) & (b > 0) & (c > 0), mn, np.where((a < 0) & (b < 0) & (c < 0), mx, 0.0))


def cuda_derivatives(field, axis, N):
    """findDerivatives with theta (plasticity_kernel.cu:53-90): returns the
    (+, -) one-sided derivatives in units of the grid (times N).  Pinned
    against find_derivatives (the same limiter written with differences of
    differences) to rounding."""
    ax = axis + 1
    r = np.roll
    val, val_r, val_l = field, r(field, -1, ax), r(field, 1, ax)
    val_rr, val_ll = r(field, -2, ax), r(field, 2, ax)
    diff_p = val_r - val
    diff_m = val - val_l
    d2 = THETA * (val_r + val_l - 2 * val)
    dp = (diff_p - 0.5 * cuda_minmod3(THETA * (val_rr + val - 2 * val_r), d2, 0.5 * (val_rr - val_r - val + val_l))) * N
    dm = (diff_m + 0.5 * cuda_minmod3(THETA * (val_ll + val - 2 * val_l), d2, 0.5 * (val_r - val - val_l + val_ll))) * N
    return dp, dm


def cuda_velocity_numerator(us, sig, lam, glide_only):
    """The corner velocity before the |rho| division (plasticity_kernel.cu:
    289-456): the glide + climb velocity with climb suppression lambda or,
    glide_only (NEWGLIDEONLY, :389-452), the lvp projection.  The 3D v_z
    subtracts uz[y,n] sigma[z,n] where its comment (and the Python path)
    have uz[y,n] sigma[y,n] (:364-369); restated as written."""
    d = len(us)

    def U(a, i, j):
        return us[a][cidx(i, j)]

    def S(i, j):
        return sig[cidx(i, j)]
    vx = U(1, 0, 0) * S(1, 0)
    vx = vx - U(0, 1, 0) * S(1, 0)
    vx = vx - U(0, 2, 0) * S(2, 0)
    vx = vx + U(1, 0, 1) * S(1, 1)
    vx = vx - U(0, 1, 1) * S(1, 1)
    vx = vx - U(0, 2, 1) * S(2, 1)
    vx = vx + U(1, 0, 2) * S(1, 2)
    vx = vx - U(0, 1, 2) * S(1, 2)
    vx = vx - U(0, 2, 2) * S(2, 2)
    if d == 3:
        vx = vx + U(2, 0, 0) * S(2, 0)
        vx = vx + U(2, 0, 1) * S(2, 1)
        vx = vx + U(2, 0, 2) * S(2, 2)
    vy = U(0, 1, 0) * S(0, 0)
    vy = vy - U(1, 0, 0) * S(0, 0)
    vy = vy - U(1, 2, 0) * S(2, 0)
    vy = vy + U(0, 1, 1) * S(0, 1)
    vy = vy - U(1, 0, 1) * S(0, 1)
    vy = vy - U(1, 2, 1) * S(2, 1)
    vy = vy + U(0, 1, 2) * S(0, 2)
    vy = vy - U(1, 0, 2) * S(0, 2)
    vy = vy - U(1, 2, 2) * S(2, 2)
    if d == 3:
        vy = vy + U(2, 1, 0) * S(2, 0)
        vy = vy + U(2, 1, 1) * S(2, 1)
        vy = vy + U(2, 1, 2) * S(2, 2)
    vz = U(0, 2, 0) * S(0, 0)
    vz = vz + U(1, 2, 0) * S(1, 0)
    vz = vz + U(0, 2, 1) * S(0, 1)
    vz = vz + U(1, 2, 1) * S(1, 1)
    vz = vz + U(0, 2, 2) * S(0, 2)
    vz = vz + U(1, 2, 2) * S(1, 2)
    if d == 3:
        vz = vz - U(2, 0, 0) * S(0, 0)
        vz = vz - U(2, 0, 1) * S(0, 1)
        vz = vz - U(2, 0, 2) * S(0, 2)
        vz = vz - U(2, 1, 0) * S(2, 0)
        vz = vz - U(2, 1, 1) * S(2, 1)
        vz = vz - U(2, 1, 2) * S(2, 2)
    st = (S(0, 0) + S(1, 1) + S(2, 2)) / 3. * lam
    vx = vx + st * (U(0, 0, 0) + U(0, 1, 1) + U(0, 2, 2))
    vy = vy + st * (U(1, 0, 0) + U(1, 1, 1) + U(1, 2, 2))
    vx = vx - st * (U(0, 0, 0) + U(1, 0, 1))
    vy = vy - st * (U(0, 1, 0) + U(1, 1, 1))
    vz = vz - st * (U(0, 2, 0) + U(1, 2, 1))
    if d == 3:
        vz = vz + st * (U(2, 0, 0) + U(2, 1, 1) + U(2, 2, 2))
        vx = vx - st * U(2, 0, 2)
        vy = vy - st * U(2, 1, 2)
        vz = vz - st * U(2, 2, 2)
    if glide_only:
        txr = U(0, 0, 0) + U(0, 1, 1) + U(0, 2, 2)
        tyr = U(1, 0, 0) + U(1, 1, 1) + U(1, 2, 2)
        if d == 2:
            rr = ME + (txr * txr + tyr * tyr)
            rr = rr + (U(0, 0, 0) + U(1, 0, 1)) * (U(0, 0, 0) + U(1, 0, 1))
            rr = rr + (U(0, 1, 0) + U(1, 1, 1)) * (U(0, 1, 0) + U(1, 1, 1))
            rr = rr + (U(0, 2, 0) + U(1, 2, 1)) * (U(0, 2, 0) + U(1, 2, 1))
            rr = rr - 2 * txr * (U(0, 0, 0) + U(1, 0, 1))
            rr = rr - 2 * tyr * (U(0, 1, 0) + U(1, 1, 1))
            rv = ME + vx * (U(0, 2, 2) + U(0, 1, 1) - U(1, 0, 1))
            rv = rv + vy * (U(1, 0, 0) + U(1, 2, 2) - U(0, 1, 0))
            rv = rv - vz * (U(0, 2, 0) + U(1, 2, 1))
            rv = rv / rr
            vx = vx + (U(0, 1, 1) + U(0, 2, 2) - U(1, 0, 1)) * rv
            vy = vy + (U(1, 0, 0) + U(1, 2, 2) - U(0, 1, 0)) * rv
            vz = vz - (U(0, 2, 0) + U(1, 2, 1)) * rv
        else:
            tzr = U(2, 0, 0) + U(2, 1, 1) + U(2, 2, 2)
            rx = U(0, 0, 0) + U(1, 0, 1) + U(2, 0, 2)
            ry = U(0, 1, 0) + U(1, 1, 1) + U(2, 1, 2)
            rz = U(0, 2, 0) + U(1, 2, 1) + U(2, 2, 2)
            rr = ME + (txr * txr + tyr * tyr + tzr * tzr)
            rr = rr + rx * rx
            rr = rr + ry * ry
            rr = rr + rz * rz
            rr = rr - 2 * txr * rx
            rr = rr - 2 * tyr * ry
            rr = rr - 2 * tzr * rz
            qx = U(0, 2, 2) + U(0, 1, 1) - U(1, 0, 1) - U(2, 0, 2)
            qy = U(1, 0, 0) + U(1, 2, 2) - U(0, 1, 0) - U(2, 1, 2)
            qz = U(2, 0, 0) + U(2, 1, 1) - U(0, 2, 0) - U(1, 2, 1)
            rv = ME + vx * qx
            rv = rv + vy * qy
            rv = rv + vz * qz
            rv = rv / rr
            vx = vx + (U(0, 1, 1) + U(0, 2, 2) - U(1, 0, 1) - U(2, 0, 2)) * rv
            vy = vy + (U(1, 0, 0) + U(1, 2, 2) - U(0, 1, 0) - U(2, 1, 2)) * rv
            vz = vz + (U(2, 0, 0) + U(2, 1, 1) - U(0, 2, 0) - U(1, 2, 1)) * rv
    return [vx, vy, vz]


def cuda_rho_norm(us):
    """|rho| of a corner without ME, negative values clamped (:232-287)."""
    d = len(us)
    r = 0.0
    for a in range(d):
        for i in range(3):
            if i == a:
                continue
            for j in range(3):
                r = r + us[a][cidx(i, j)] * us[a][cidx(i, j)]
    for j in range(3):
        r = r - 2 * us[0][cidx(1, j)] * us[1][cidx(0, j)]
    if d == 3:
        for j in range(3):
            r = r - 2 * us[0][cidx(2, j)] * us[2][cidx(0, j)]
        for j in range(3):
            r = r - 2 * us[1][cidx(2, j)] * us[2][cidx(1, j)]
    return np.sqrt(np.maximum(r, 0.0))


def cuda_llf_flux(beta, sig, lam, glide_only=False):
    """centralHJ with LLF (plasticity_kernel.cu:118-697): corners k (x sign
    k % 2, y sign (k % 4) / 2, z sign k / 4; 0 is '+'), velocities divided by
    |rho| + sqrt(ME), one speed a = max over corners and axes of |v|, then
    rhs = sum_k -H_k a^d / prod(2a + ME) + sum_axes (u+ - u-) a^2 / (2a + ME)
    and speed = sum_axes (2a + ME).  Pinned by tests/test_oracle_cuda.py."""
    d = beta.ndim - 1
    N = beta.shape[-1]
    du = []
    for a in range(d):
        p, m = cuda_derivatives(beta, a, N)
        du += [p, m]
    ncor = 1 << d
    V, US = [], []
    for k in range(ncor):
        s = [k % 2, (k % 4) // 2, k // 4]
        us = [du[2 * a + s[a]] for a in range(d)]
        v = cuda_velocity_numerator(us, sig, lam, glide_only)
        irho = cuda_rho_norm(us) + np.sqrt(ME)
        V.append([vv / irho for vv in v])
        US.append(us)
    amax = np.zeros(beta.shape[1:])
    for k in range(ncor):
        for l in range(3):
            amax = np.maximum(amax, np.abs(V[k][l]))
    at = amax + amax + ME
    deriv = np.zeros_like(beta)
    for k in range(ncor):
        us, v = US[k], V[k]
        for i in range(3):
            for j in range(3):
                c = cidx(i, j)
                h = us[0][c] * v[0] + us[1][c] * v[1]
                if d == 3:
                    h = h + us[2][c] * v[2]
                if i < d:
                    h = h - v[0] * us[i][cidx(0, j)]
                    h = h - v[1] * us[i][cidx(1, j)]
                    h = h - v[2] * us[i][cidx(2, j)]
                if i == j:
                    q = (v[0] * (us[0][cidx(1, 1)] + us[0][cidx(2, 2)] - us[1][cidx(0, 1)])
                         + v[1] * (us[1][cidx(0, 0)] + us[1][cidx(2, 2)] - us[0][cidx(1, 0)])
                         - v[2] * (us[0][cidx(2, 0)] + us[1][cidx(2, 1)]))
                    if d == 3:
                        q = q - v[0] * us[2][cidx(0, 2)] - v[1] * us[2][cidx(1, 2)] + v[2] * (
                            us[2][cidx(0, 0)] + us[2][cidx(1, 1)])
                    if not glide_only:
                        h = h - q / 3. * lam
                w = -h * amax * amax
                if d == 3:
                    w = w * amax
                for a in range(d):
                    w = w / at
                deriv[c] = deriv[c] + w
    for a in range(d):
        deriv = deriv + (du[2 * a] - du[2 * a + 1]) * amax * amax / at
    return deriv, at * d


def cuda_euler_step(flux, t, u, t_end, cfl, N, dt_cap=None):
    """simpleTVD (plasticity_kernel.cu:1112-1137): dt = CFL / max speed / N,
    capped at maxNucleationTimestep under DYNAMIC_NUCLEATION (:1122-1125),
    past the end time dt = t_end - t + ME; u += dt rhs."""
    rhs, speed = flux(t, u)
    dt = cfl / speed.max() / N
    if dt_cap is not None and dt > dt_cap:
        dt = dt_cap
    if t + dt > t_end:
        dt = t_end - t + ME
    return u + dt * rhs, dt


def cuda_tvd3_step(flux, t, u, t_end, cfl, N, dt_cap=None):
    """TVD3rd (plasticity_kernel.cu:1140-1206): the Shu-Osher stages of the
    Python mover (every stage at the step's start time) with simpleTVD's dt
    rule and end clamp.  Pinned against tvd_rk3_step, whose stages it shares,
    by tests/test_oracle_cuda.py."""
    rhs, speed = flux(t, u)
    dt = cfl / speed.max() / N
    if dt_cap is not None and dt > dt_cap:
        dt = dt_cap
    if t + dt > t_end:
        dt = t_end - t + ME
    L0 = dt * rhs
    F0 = u + L0
    rhs, _ = flux(t, F0)
    F1 = 3. / 4. * u + 1. / 4. * F0 + 1. / 4. * (dt * rhs)
    rhs, _ = flux(t, F1)
    return 1. / 3. * u + 2. / 3. * F1 + 2. / 3. * (dt * rhs), dt


def nucleation_cap(source, mu, nu):
    """maxNucleationTimestep = 1 / sqrt(|sigma|) of the sigma entry of largest
    magnitude over all nine components of the source's stress
    (plasticity.cu:190-199, reduceMax = Iamax, plasticity_kernel.cu:1096-1103)."""
    s = sigma(source, mu, nu)
    return 1.0 / np.sqrt(np.abs(s).max())


def run_cuda(dyn, beta, t0, nsteps, cfl=0.5, t_end=np.inf, order=1, source=None, rate=0.0, dt_cap=None):
    """The CUDA executable's run loop (plasticity.cu:242-271) for a fixed
    number of steps with the LLF flux and forward Euler (TVDstep = simpleTVD,
    plasticity.h:9) or, order 3, TVD3rd; dyn gives Lambda (its 'lambda'),
    glide-only (NEWGLIDEONLY), elastic constants and the load.  source,
    rate, dt_cap: DYNAMIC_NUCLEATION, every flux gains rate * source
    (calculateFlux's updateField(rhs, 1/TIME_FINAL, beta0dot),
    plasticity_kernel.cu:1089-1091) and dt is capped at dt_cap.
    Returns (u, t, dts)."""
    glide_only = dyn.dynamics == GLIDE_ONLY
    lam = 0.0 if glide_only else dyn.Lambda
    N = beta.shape[-1]
    step = cuda_euler_step if order == 1 else cuda_tvd3_step
    t, u, dts = t0, beta, []

    def flux(tt, uu):
        rhs, speed = cuda_llf_flux(uu, dyn.get_sigma(uu, tt), lam, glide_only)
        if source is not None:
            rhs = rhs + rate * source
        return rhs, speed
    for _ in range(nsteps):
        u, dt = step(flux, t, u, t_end, cfl, N, dt_cap)
        t += dt
        dts.append(dt)
        if t >= t_end:
            break
    return u, t, np.array(dts)


def run_cuda_vacancy(dyn, vac, beta, c, t0, nsteps, cfl=0.5, t_end=np.inf):
    """The CUDA executable's VACANCIES variant under its simpleTVD step
    (plasticity_kernel.cu:176-196,284-290,676-699,700-775,1112-1137): the
    stress at every cell loses vacancycost c on the diagonal (after the
    load), the tensor rhs is the LLF flux, the vacancy component's rhs is
    the trace of the tensor rhs, u and c advance by dt times their rhs, and
    c then diffuses by exp(-vacancydiffusion dt k^2) (calculateKDiffusion).
    vac.alpha is vacancycost, vac.gamma vacancydiffusion.  Returns
    (u, c, t, dts)."""
    glide_only = dyn.dynamics == GLIDE_ONLY
    lam = 0.0 if glide_only else dyn.Lambda
    N = beta.shape[-1]
    t, u, dts = t0, beta, []
    for _ in range(nsteps):
        sig = dyn.get_sigma(u, t).copy()
        for i in range(3):
            sig[cidx(i, i)] = sig[cidx(i, i)] - vac.alpha * c
        rhs, speed = cuda_llf_flux(u, sig, lam, glide_only)
        dt = cfl / speed.max() / N
        if t + dt > t_end:
            dt = t_end - t + ME
        u = u + dt * rhs
        c = c + dt * ((rhs[0] + rhs[4]) + rhs[8])
        c = vacancy_diffuse(c, dt, vac.gamma)
        t += dt
        dts.append(dt)
        if t >= t_end:
            break
    return u, c, t, np.array(dts)
