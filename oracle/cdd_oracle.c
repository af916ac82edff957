/* CPU oracle for the CDD stepper, C/OpenMP restatement.
 *
 * TEST INFRASTRUCTURE ONLY.  A float64 restatement, function by function, of
 * oracle/cdd_oracle.py (itself a restatement of the reference Python path of
 * mattbierbaum/cuda-plasticity), written so the GPU path can be checked at
 * grid sizes and step counts where the NumPy oracle is too slow (3D 128^3,
 * 200 RK3 steps).  Only tests/ and the cpu_baseline leg of bench.py load
 * it; the product path never does.
 *
 * It follows the NumPy oracle's arithmetic operation by operation (mask-form
 * minmod/max/min, per-corner velocities and Hamiltonians in the reference's
 * summation order, Shu-Osher stages with their literal zero terms).  The
 * FFT is a radix-2 complex transform of its own in numpy's conventions:
 * rfftn as the full transform of the real field (last-axis indices
 * 0..n/2 used), irfftn as inverse transforms along the leading axes followed
 * by pocketfft's c2r along the last axis (Hermitian extension, imaginary
 * parts of its DC and Nyquist bins dropped).  The two oracles agree to
 * rounding (tests/test_oracle_c.py).
 *
 * Reference files restated (via oracle/cdd_oracle.py): PlasticityState.py:
 * 85-127,357-390 (stress), FourierSpaceTools.py:5-61 (k grids),
 * NumericalMethods.py:352-497 (limiter, Max/Min), CentralUpwindHJ3D.py:11-271
 * (derivatives, flux), CentralUpwindHJBetaPDynamics3D.py:34-201 and
 * CentralUpwindHJBetaPGlideOnlyDynamics.py:69-242 (velocities,
 * Hamiltonians), FieldMover.py:61-141 (RK3, CFL), RunPython.py:21-66 and
 * RunLoadingMPI.py:68-83 (loads).  Grids are square with side n a power of
 * two. */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ME 2.220446049250313e-16 /* NumericalMethods.py:4 */
#define CI(i, j) (3 * (i) + (j))
#define XX 0
#define YY 1
#define ZZ 2

typedef struct {
  int d, n, dyn; /* dyn: 0 glide+climb, 1 glide-only (lvp) */
  int order;     /* 2 (SSP-RK2) or 3 (TVD-RK3) */
  double Lambda, mu, nu, cfl;
  int load_kind; /* 0 none, 1 strain, 2 stress, 3 uniaxial, 4 tensor */
  double init[9], rate[9], t0, init_strain, uniaxial_rate;
  double load_mu, load_nu; /* the strain load's elastic constants (0: mu, nu) */
  double fixed_dt;          /* > 0: CalculateOneFixedTimeStep (FieldMover.py:38-58) */
} oc_cfg;

/* ------------------------------------------------------------ FFT */
/* twiddles exp(-2 pi i k / n), k < n/2, one table per n (long double) */
static double complex* tw_tab[32];
static const double complex* twiddles(int n) {
  int lg = 0;
  while ((1 << lg) < n) ++lg;
  double complex* t;
#pragma omp critical(cdd_oracle_tw)
  {
    if (!tw_tab[lg]) {
      t = (double complex*)malloc(sizeof(double complex) * (n / 2 > 0 ? n / 2 : 1));
      for (int k = 0; k < n / 2; ++k) {
        long double th = -2.0L * 3.14159265358979323846264338327950288L * k / n;
        t[k] = (double)cosl(th) + I * (double)sinl(th);
      }
      tw_tab[lg] = t;
    }
    t = tw_tab[lg];
  }
  return t;
}

static void fft1d(double complex* a, int n, int sign, const double complex* tw) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double complex t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int step = n / len;
    for (int k = 0; k < len / 2; ++k) {
      const double complex w = sign < 0 ? tw[k * step] : conj(tw[k * step]);
      for (int i = 0; i < n; i += len) {
        double complex u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

/* unnormalised transform of an n^d array along axis ax (sign -1 forward);
 * half = 1 skips the lines whose last-axis index exceeds n/2 (ax < d - 1),
 * which the half-spectrum inverse never reads */
static void fft_axis(double complex* a, int d, int n, int ax, int sign, int half) {
  const long long cells = d == 3 ? (long long)n * n * n : (long long)n * n;
  long long stride = 1;
  for (int b = ax + 1; b < d; ++b) stride *= n;
  const long long lines = cells / n;
  const double complex* tw = twiddles(n);
#pragma omp parallel
  {
    double complex* buf = (double complex*)malloc(sizeof(double complex) * n);
#pragma omp for schedule(static)
    for (long long l = 0; l < lines; ++l) {
      const long long hi = l / stride, lo = l % stride;
      if (half && lo % n > n / 2) continue;
      const long long base = hi * stride * n + lo;
      for (int i = 0; i < n; ++i) buf[i] = a[base + i * stride];
      fft1d(buf, n, sign, tw);
      for (int i = 0; i < n; ++i) a[base + i * stride] = buf[i];
    }
    free(buf);
  }
}

/* numpy.fft.rfftn of a real field (kept as the full spectrum; only the
 * last-axis indices 0..n/2 are used afterwards) */
static void rfftn_full(double complex* a, int d, int n) {
  for (int ax = 0; ax < d; ++ax) fft_axis(a, d, n, ax, -1, 0);
}

/* numpy.fft.irfftn(s = grid) of the half spectrum stored in the last-axis
 * slots 0..n/2 of a: inverse complex transforms along the other axes, then
 * the c2r along the last axis, which takes the Hermitian extension of the
 * half line and ignores the imaginary parts of its DC and Nyquist bins
 * (pocketfft's c2r).  The real result is left in the real parts. */
static void irfftn_half(double complex* a, int d, int n) {
  const long long cells = d == 3 ? (long long)n * n * n : (long long)n * n;
  for (int ax = 0; ax < d - 1; ++ax) fft_axis(a, d, n, ax, +1, 1);
  const long long lines = cells / n;
  const double s = 1.0 / (double)cells;
  const double complex* tw = twiddles(n);
#pragma omp parallel
  {
    double complex* buf = (double complex*)malloc(sizeof(double complex) * n);
#pragma omp for schedule(static)
    for (long long l = 0; l < lines; ++l) {
      double complex* row = a + l * n;
      buf[0] = creal(row[0]);
      buf[n / 2] = creal(row[n / 2]);
      for (int k = 1; k < n / 2; ++k) {
        buf[k] = row[k];
        buf[n - k] = conj(row[k]);
      }
      fft1d(buf, n, +1, tw);
      for (int i = 0; i < n; ++i) row[i] = creal(buf[i]) * s;
    }
    free(buf);
  }
}

/* FourierSpaceTools.py:5-7: theta - 2 pi [theta > pi] */
static double ident(double th) { return th - 2.0 * M_PI * (double)(th > M_PI); }

/* ------------------------------------------------------------ stress */
/* sigma of beta at time t, external load included (CalculateSigma 'betaE' +
 * GetSigma load).  beta, sig: 9 planes of n^d. */
void oc_sigma(const oc_cfg* c, const double* beta, double t, double* sig) {
  const int d = c->d, n = c->n;
  const long long cells = d == 3 ? (long long)n * n * n : (long long)n * n;
  const double mu = c->mu, lamb = 2.0 * c->mu * c->nu / (1.0 - 2.0 * c->nu);
  /* spectra kept between calls (the oracle is driven from one host thread) */
  static double complex *kb = NULL, *ke = NULL;
  static long long cap = 0;
  if (cap < cells) {
    free(kb);
    free(ke);
    kb = (double complex*)malloc(sizeof(double complex) * 9 * cells);
    ke = (double complex*)malloc(sizeof(double complex) * 9 * cells);
    cap = cells;
  }
  for (int q = 0; q < 9; ++q) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < cells; ++i) kb[q * cells + i] = beta[q * cells + i];
    rfftn_full(kb + q * cells, d, n);
  }
  /* T_ijmn contraction (beta_elastic): kbe_ij = kmask sum_mn T_ijmn kb_mn */
#pragma omp parallel for schedule(static)
  for (long long idx = 0; idx < cells; ++idx) {
    int ix, iy, iz;
    double k[3];
    if (d == 2) {
      ix = (int)(idx / n);
      iy = (int)(idx % n);
      k[0] = n * ident(2.0 * M_PI * ix / (double)n);
      k[1] = n * ident(2.0 * M_PI * iy / (double)n);
      k[2] = 0.0;
      iz = 0;
    } else {
      ix = (int)(idx / ((long long)n * n));
      iy = (int)((idx / n) % n);
      iz = (int)(idx % n);
      k[0] = n * ident(2.0 * M_PI * ix / (double)n);
      k[1] = n * ident(2.0 * M_PI * iy / (double)n);
      k[2] = n * ident(2.0 * M_PI * iz / (double)n);
    }
    /* the half spectrum only: irfftn reads last-axis indices 0..n/2 */
    if ((d == 2 ? iy : iz) > n / 2) continue;
    const double kSq0 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    const double kSq = kSq0 + ME, kSqSq = kSq0 * kSq0 + ME;
    const double kmask = (ix == 0 && iy == 0 && iz == 0) ? 0.0 : 1.0;
    /* T_ijmn = A k_i k_j / K2 d_mn + k_i / K2 (k_m d_jn + k_n d_jm) - C k_i k_j k_m k_n / K4,
     * each factor evaluated in that order (the sub-products that do not
     * depend on m, n are formed once) */
    const double A = lamb / (lamb + 2.0 * mu), C = 2.0 * (lamb + mu) / (lamb + 2.0 * mu);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double t1 = A * k[i] * k[j] / kSq, t2 = k[i] / kSq, kij = k[i] * k[j];
        double complex acc = 0.0;
        for (int m = 0; m < 3; ++m)
          for (int nn = 0; nn < 3; ++nn) {
            const double T = t1 * (double)(m == nn) + t2 * ((k[m] * (double)(j == nn)) + (k[nn] * (double)(j == m))) -
                             C * (kij * k[m] * k[nn]) / kSqSq;
            acc = acc + kb[CI(m, nn) * cells + idx] * T;
          }
        ke[CI(i, j) * cells + idx] = acc * kmask;
      }
  }
  for (int q = 0; q < 9; ++q) irfftn_half(ke + q * cells, d, n);
  /* load shift (GetSigma overrides) */
  double mean_bzz = 0.0;
  if (c->load_kind == 3) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long long i = 0; i < cells; ++i) s += beta[8 * cells + i];
    mean_bzz = s / (double)cells;
  }
  double shift[9] = {0};
  if (c->load_kind == 1) {
    double e[3];
    for (int i = 0; i < 3; ++i) e[i] = c->init[i] + c->rate[i] * (t - c->t0);
    const double tr = e[0] + e[1] + e[2];
    const double smu = c->load_mu > 0 ? c->load_mu : mu;
    const double snu = c->load_mu > 0 ? c->load_nu : c->nu;
    const double slam = 2.0 * smu * snu / (1.0 - 2.0 * snu);
    for (int i = 0; i < 3; ++i) shift[CI(i, i)] = slam * tr + 2.0 * smu * e[i];
  } else if (c->load_kind == 2) {
    for (int i = 0; i < 3; ++i) shift[CI(i, i)] = c->init[i] + c->rate[i] * (t - c->t0);
  } else if (c->load_kind == 3) {
    const double bzz = c->uniaxial_rate * t - (mean_bzz - c->init_strain);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) shift[CI(i, j)] = (lamb * (double)(i == j) + 2.0 * mu * (double)(i == ZZ && j == ZZ)) * bzz;
  } else if (c->load_kind == 4) {
    for (int q = 0; q < 9; ++q) shift[q] = c->init[q] + c->rate[q] * (t - c->t0);
  }
  /* betaE = IFFT - beta; sigma = lam tr(betaE) d + mu (betaE + betaE^T) */
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < cells; ++i) {
    double be[9];
    for (int q = 0; q < 9; ++q) be[q] = creal(ke[q * cells + i]) - beta[q * cells + i];
    const double tr = be[0] + be[4] + be[8];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double s = lamb * tr * (double)(a == b) + mu * (be[CI(a, b)] + be[CI(b, a)]);
        if (c->load_kind != 0) s += shift[CI(a, b)];
        sig[CI(a, b) * cells + i] = s;
      }
  }
}

/* ------------------------------------------------------------ limiter */
static double minmod3(double a, double b, double c) { /* NumericalMethods.py:367-378 */
  const double ca = (double)(a > 0), cb = (double)(b > 0), cc = (double)(c > 0);
  const double cab = (double)(a < b), cbc = (double)(b < c), cac = (double)(a < c);
  const double cabc = ca * cb * cc;
  const double cnot = (1 - ca) * (1 - cb) * (1 - cc);
  return cabc * (cab * cac * a + (1 - cab) * cbc * b + (1 - cbc) * (1 - cac) * c) +
         cnot * ((1 - cab) * (1 - cac) * a + cab * (1 - cbc) * b + cbc * cac * c);
}
static double max3(double a, double b, double c) {
  const double cab = (double)(a > b), cbc = (double)(b > c), cac = (double)(a > c);
  return cab * cac * a + (1 - cab) * cbc * b + (1 - cbc) * (1 - cac) * c;
}
static double min3(double a, double b, double c) {
  const double cab = (double)(a < b), cbc = (double)(b < c), cac = (double)(a < c);
  return cab * cac * a + (1 - cab) * cbc * b + (1 - cbc) * (1 - cac) * c;
}
static double max5(const double* v) { return max3(v[0], v[1], max3(v[2], v[3], v[4])); }
static double min5(const double* v) { return min3(v[0], v[1], min3(v[2], v[3], v[4])); }
static double max9(const double* v) {
  double t[5] = {v[4], v[5], v[6], v[7], v[8]};
  double u[5] = {v[0], v[1], v[2], v[3], max5(t)};
  return max5(u);
}
static double min9(const double* v) {
  double t[5] = {v[4], v[5], v[6], v[7], v[8]};
  double u[5] = {v[0], v[1], v[2], v[3], min5(t)};
  return min5(u);
}

/* FindDerivatives at one cell from f[i-2..i+2] (CentralUpwindHJ3D.py:11-34) */
static void derivs(const double f[5], double Dx, double* dp, double* dm) {
  const double diff_m2 = f[1] - f[0], diff_m1 = f[2] - f[1], diff_0 = f[3] - f[2], diff_p1 = f[4] - f[3];
  /* diff2 at i and at i-1 */
  const double d21_i = 1.0 * (diff_p1 - diff_0) / Dx / Dx;
  const double d22_i = (diff_p1 - diff_m1) / Dx / Dx * 0.5;
  const double d21_im1 = 1.0 * (diff_0 - diff_m1) / Dx / Dx;
  const double d23_i = d21_im1;
  const double d22_im1 = (diff_0 - diff_m2) / Dx / Dx * 0.5;
  const double d23_im1 = 1.0 * (diff_m1 - diff_m2) / Dx / Dx;
  const double m_i = minmod3(d21_i, d22_i, d23_i);
  const double m_im1 = minmod3(d21_im1, d22_im1, d23_im1);
  *dp = diff_0 / Dx - 0.5 * Dx * m_i;
  *dm = diff_m1 / Dx + 0.5 * Dx * m_im1;
}

/* derivs over n cells whose five stencil values are rows r0..r4 (the same
 * operations as derivs, in a form the compiler vectorises) */
static void deriv_rows(const double* restrict r0, const double* restrict r1, const double* restrict r2,
                       const double* restrict r3, const double* restrict r4, double* restrict dpo,
                       double* restrict dmo, int n, double Dx) {
  for (int z = 0; z < n; ++z) {
    const double diff_m2 = r1[z] - r0[z], diff_m1 = r2[z] - r1[z], diff_0 = r3[z] - r2[z], diff_p1 = r4[z] - r3[z];
    const double d21_i = 1.0 * (diff_p1 - diff_0) / Dx / Dx;
    const double d22_i = (diff_p1 - diff_m1) / Dx / Dx * 0.5;
    const double d21_im1 = 1.0 * (diff_0 - diff_m1) / Dx / Dx;
    const double d23_i = d21_im1;
    const double d22_im1 = (diff_0 - diff_m2) / Dx / Dx * 0.5;
    const double d23_im1 = 1.0 * (diff_m1 - diff_m2) / Dx / Dx;
    const double m_i = minmod3(d21_i, d22_i, d23_i);
    const double m_im1 = minmod3(d21_im1, d22_im1, d23_im1);
    dpo[z] = diff_0 / Dx - 0.5 * Dx * m_i;
    dmo[z] = diff_m1 / Dx + 0.5 * Dx * m_im1;
  }
}

/* ------------------------------------------------------------ velocities */
static double rhomod(int d, double us[3][9]) { /* BetaP{2,3}D_RhoMod */
  double rho2 = ME;
  for (int c = 0; c < 9; ++c)
    for (int a = 0; a < d; ++a) rho2 = rho2 + us[a][c] * us[a][c];
  for (int j = 0; j < 3; ++j) {
    rho2 = rho2 - us[0][CI(XX, j)] * us[0][CI(XX, j)];
    rho2 = rho2 - us[1][CI(YY, j)] * us[1][CI(YY, j)];
    if (d == 3) rho2 = rho2 - us[2][CI(ZZ, j)] * us[2][CI(ZZ, j)];
    rho2 = rho2 - 2 * us[0][CI(YY, j)] * us[1][CI(XX, j)];
    if (d == 3) {
      rho2 = rho2 - 2 * us[1][CI(ZZ, j)] * us[2][CI(YY, j)];
      rho2 = rho2 - 2 * us[2][CI(XX, j)] * us[0][CI(ZZ, j)];
    }
  }
  return sqrt(rho2);
}

static void climb_v(int d, double us[3][9], const double* sg, double v[3]) {
  v[0] = v[1] = v[2] = 0.0;
  for (int n = 0; n < 3; ++n)
    for (int l = 0; l < 3; ++l) {
      for (int a = 0; a < d; ++a) v[l] = v[l] + us[a][CI(l, n)] * sg[CI(a, n)];
      for (int a = 0; a < d; ++a) v[a] = v[a] - us[a][CI(l, n)] * sg[CI(l, n)];
    }
}

static void velocity_gc(int d, double us[3][9], const double* sg, double Lambda, double v[3]) {
  climb_v(d, us, sg, v);
  if (Lambda != 0) {
    const double sigma_tr = (sg[0] + sg[4] + sg[8]) / 3.0 * Lambda;
    for (int a = 0; a < d; ++a) v[a] = v[a] + sigma_tr * (us[a][0] + us[a][4] + us[a][8]);
    for (int l = 0; l < 3; ++l) {
      double acc = us[0][CI(l, XX)];
      for (int a = 1; a < d; ++a) acc = acc + us[a][CI(l, a)];
      v[l] = v[l] - sigma_tr * acc;
    }
  }
  const double rm = rhomod(d, us);
  for (int l = 0; l < 3; ++l) v[l] = v[l] / rm;
}

static void velocity_lvp(int d, double us[3][9], const double* sg, double v[3]) {
  climb_v(d, us, sg, v);
  const double* ux = us[0];
  const double* uy = us[1];
  if (d == 2) {
    double rhorho = ME;
    const double uxTr = ux[0] + ux[4] + ux[8], uyTr = uy[0] + uy[4] + uy[8];
    rhorho = rhorho + uxTr * uxTr + uyTr * uyTr;
    rhorho = rhorho - 2 * uxTr * (ux[CI(XX, XX)] + uy[CI(XX, YY)]);
    rhorho = rhorho - 2 * uyTr * (ux[CI(YY, XX)] + uy[CI(YY, YY)]);
    for (int n = 0; n < 3; ++n) {
      const double t = ux[CI(n, XX)] + uy[CI(n, YY)];
      rhorho = rhorho + t * t;
    }
    double rhov = ME;
    rhov = rhov + v[0] * (ux[CI(ZZ, ZZ)] + ux[CI(YY, YY)] - uy[CI(XX, YY)]);
    rhov = rhov + v[1] * (uy[CI(XX, XX)] + uy[CI(ZZ, ZZ)] - ux[CI(YY, XX)]);
    rhov = rhov + -v[2] * (ux[CI(ZZ, XX)] + uy[CI(ZZ, YY)]);
    v[0] = v[0] + (ux[CI(YY, YY)] + ux[CI(ZZ, ZZ)] - uy[CI(XX, YY)]) * rhov / rhorho;
    v[1] = v[1] + (uy[CI(XX, XX)] + uy[CI(ZZ, ZZ)] - ux[CI(YY, XX)]) * rhov / rhorho;
    v[2] = v[2] - (ux[CI(ZZ, XX)] + uy[CI(ZZ, YY)]) * rhov / rhorho;
  } else {
    const double* uz = us[2];
    double rhorho = ME;
    const double uxTr = ux[0] + ux[4] + ux[8], uyTr = uy[0] + uy[4] + uy[8], uzTr = uz[0] + uz[4] + uz[8];
    rhorho = rhorho + uxTr * uxTr + uyTr * uyTr + uzTr * uzTr;
    rhorho = rhorho - 2 * uxTr * (ux[CI(XX, XX)] + uy[CI(XX, YY)] + uz[CI(XX, ZZ)]);
    rhorho = rhorho - 2 * uyTr * (ux[CI(YY, XX)] + uy[CI(YY, YY)] + uz[CI(YY, ZZ)]);
    rhorho = rhorho - 2 * uzTr * (ux[CI(ZZ, XX)] + uy[CI(ZZ, YY)] + uz[CI(ZZ, ZZ)]);
    for (int n = 0; n < 3; ++n) {
      const double t = ux[CI(n, XX)] + uy[CI(n, YY)] + uz[CI(n, ZZ)];
      rhorho = rhorho + t * t;
    }
    double rhov = ME;
    rhov = rhov + v[0] * (ux[CI(ZZ, ZZ)] + ux[CI(YY, YY)] - uy[CI(XX, YY)] - uz[CI(XX, ZZ)]);
    rhov = rhov + v[1] * (uy[CI(XX, XX)] + uy[CI(ZZ, ZZ)] - ux[CI(YY, XX)] - uz[CI(YY, ZZ)]);
    rhov = rhov + v[2] * (uz[CI(XX, XX)] + uz[CI(YY, YY)] - ux[CI(ZZ, XX)] - uy[CI(ZZ, YY)]);
    const double factor = rhov / rhorho;
    v[0] = v[0] + (ux[CI(YY, YY)] + ux[CI(ZZ, ZZ)] - uy[CI(XX, YY)] - uz[CI(XX, ZZ)]) * factor;
    v[1] = v[1] + (uy[CI(XX, XX)] + uy[CI(ZZ, ZZ)] - ux[CI(YY, XX)] - uz[CI(YY, ZZ)]) * factor;
    v[2] = v[2] + (uz[CI(XX, XX)] + uz[CI(YY, YY)] - ux[CI(ZZ, XX)] - uy[CI(ZZ, YY)]) * factor;
  }
  const double rm = rhomod(d, us);
  for (int l = 0; l < 3; ++l) v[l] = v[l] / rm;
}

static void hamiltonian(int d, double us[3][9], const double V[3], double Lambda, double H[9]) {
  for (int c = 0; c < 9; ++c) {
    double h = us[0][c] * V[0];
    for (int a = 1; a < d; ++a) h = h + us[a][c] * V[a];
    H[c] = h;
  }
  for (int j = 0; j < 3; ++j)
    for (int l = 0; l < 3; ++l)
      for (int a = 0; a < d; ++a) H[CI(a, j)] = H[CI(a, j)] - V[l] * us[a][CI(l, j)];
  if (Lambda != 0) {
    double J;
    const double* ux = us[0];
    const double* uy = us[1];
    if (d == 2) {
      J = -Lambda / 3.0 *
          (V[0] * (ux[CI(YY, YY)] + ux[CI(ZZ, ZZ)] - uy[CI(XX, YY)]) + V[1] * (uy[CI(XX, XX)] + uy[CI(ZZ, ZZ)] - ux[CI(YY, XX)]) -
           V[2] * (ux[CI(ZZ, XX)] + uy[CI(ZZ, YY)]));
    } else {
      const double* uz = us[2];
      J = -Lambda / 3.0 *
          (V[0] * (ux[CI(YY, YY)] + ux[CI(ZZ, ZZ)] - uy[CI(XX, YY)] - uz[CI(XX, ZZ)]) +
           V[1] * (uy[CI(XX, XX)] + uy[CI(ZZ, ZZ)] - ux[CI(YY, XX)] - uz[CI(YY, ZZ)]) +
           V[2] * (uz[CI(XX, XX)] + uz[CI(YY, YY)] - ux[CI(ZZ, XX)] - uy[CI(ZZ, YY)]));
    }
    for (int i = 0; i < 3; ++i) H[CI(i, i)] = H[CI(i, i)] + J;
  }
}

static void velocity(const oc_cfg* c, double us[3][9], const double* sg, double v[3]) {
  if (c->dyn == 1) velocity_lvp(c->d, us, sg, v);
  else velocity_gc(c->d, us, sg, c->Lambda, v);
}

/* ------------------------------------------------------------ flux */
/* The limited one-sided derivatives of every component along every axis,
 * DP/DM[(a 9 + q) cells + idx]: each a function of the cell's five stencil
 * values only, computed line by line along the contiguous axis so that the
 * five source rows stream (the values do not depend on the loop order). */
static void deriv_fields(int d, int n, long long cells, const double* beta, double* DP, double* DM) {
  const double Dx = 1.0 / n;
  const long long lines = cells / n;
  for (int a = 0; a < d; ++a) {
    const long long str = a == d - 1 ? 1 : (a == 0 ? cells / n : (long long)n);
#pragma omp parallel for schedule(static) collapse(2)
    for (int q = 0; q < 9; ++q)
      for (long long l = 0; l < lines; ++l) {
        const double* src = beta + q * cells;
        const long long idx0 = l * n;
        double* dpo = DP + (long long)(a * 9 + q) * cells + idx0;
        double* dmo = DM + (long long)(a * 9 + q) * cells + idx0;
        if (a == d - 1) {
          const double* row = src + idx0;
          for (int z = 0; z < n; ++z) {
            double f[5];
            for (int k = 0; k < 5; ++k) f[k] = row[(z + k - 2 + n) % n];
            derivs(f, Dx, &dpo[z], &dmo[z]);
          }
        } else {
          const int co = (int)((idx0 / str) % n);
          const double* rows[5];
          for (int k = 0; k < 5; ++k) {
            const int p = (co + k - 2 + n) % n;
            rows[k] = src + idx0 + (long long)(p - co) * str;
          }
          deriv_rows(rows[0], rows[1], rows[2], rows[3], rows[4], dpo, dmo, n, Dx);
        }
      }
  }
}

/* central_upwind_flux: rhs (9 planes) and speed (1 plane; may be NULL) */
void oc_flux(const oc_cfg* c, const double* beta, const double* sig, double* rhs, double* speed) {
  const int d = c->d, n = c->n;
  const long long cells = d == 3 ? (long long)n * n * n : (long long)n * n;
  const double lam = c->dyn == 1 ? 0.0 : c->Lambda;
  /* derivative fields kept between calls (one host thread drives the oracle) */
  static double* dws = NULL;
  static long long dcap = 0;
  if (dcap < 2LL * d * 9 * cells) {
    free(dws);
    dws = (double*)malloc(sizeof(double) * 2 * d * 9 * cells);
    dcap = 2LL * d * 9 * cells;
  }
  double* DP = dws;
  double* DM = dws + (long long)d * 9 * cells;
  deriv_fields(d, n, cells, beta, DP, DM);
#pragma omp parallel for schedule(static)
  for (long long idx = 0; idx < cells; ++idx) {
    double dp[3][9], dm[3][9], sg[9];
    for (int a = 0; a < d; ++a)
      for (int q = 0; q < 9; ++q) {
        dp[a][q] = DP[(long long)(a * 9 + q) * cells + idx];
        dm[a][q] = DM[(long long)(a * 9 + q) * cells + idx];
      }
    for (int q = 0; q < 9; ++q) sg[q] = sig[q * cells + idx];
    double out[9], spd;
    if (d == 2) {
      double us[3][9];
      double v_mm[3], v_mp[3], v_pm[3], v_pp[3];
      memcpy(us[0], dm[0], sizeof(us[0])); memcpy(us[1], dm[1], sizeof(us[1])); velocity(c, us, sg, v_mm);
      memcpy(us[0], dm[0], sizeof(us[0])); memcpy(us[1], dp[1], sizeof(us[1])); velocity(c, us, sg, v_mp);
      memcpy(us[0], dp[0], sizeof(us[0])); memcpy(us[1], dm[1], sizeof(us[1])); velocity(c, us, sg, v_pm);
      memcpy(us[0], dp[0], sizeof(us[0])); memcpy(us[1], dp[1], sizeof(us[1])); velocity(c, us, sg, v_pp);
      double w0[5] = {v_mm[0], v_mp[0], v_pm[0], v_pp[0], 0.0}, w1[5] = {v_mm[1], v_mp[1], v_pm[1], v_pp[1], 0.0};
      const double a_p = max5(w0), a_m = fabs(min5(w0)), b_p = max5(w1), b_m = fabs(min5(w1));
      const double a_tot = a_p + a_m + ME, b_tot = b_p + b_m + ME;
      double H_pp[9], H_pm[9], H_mp[9], H_mm[9];
      memcpy(us[0], dp[0], sizeof(us[0])); memcpy(us[1], dp[1], sizeof(us[1])); hamiltonian(2, us, v_pp, lam, H_pp);
      memcpy(us[0], dp[0], sizeof(us[0])); memcpy(us[1], dm[1], sizeof(us[1])); hamiltonian(2, us, v_pm, lam, H_pm);
      memcpy(us[0], dm[0], sizeof(us[0])); memcpy(us[1], dp[1], sizeof(us[1])); hamiltonian(2, us, v_mp, lam, H_mp);
      memcpy(us[0], dm[0], sizeof(us[0])); memcpy(us[1], dm[1], sizeof(us[1])); hamiltonian(2, us, v_mm, lam, H_mm);
      for (int q = 0; q < 9; ++q)
        out[q] = (-(H_pp[q] * a_m * b_m + H_pm[q] * a_m * b_p + H_mp[q] * a_p * b_m + H_mm[q] * a_p * b_p) / a_tot / b_tot +
                  ((dp[0][q] - dm[0][q]) * a_m * a_p / a_tot)) +
                 ((dp[1][q] - dm[1][q]) * b_m * b_p / b_tot);
      spd = a_tot + b_tot;
    } else {
      /* opts '---','--+',...,'+++': bit 2 = x, 1 = y, 0 = z; '+' = 1 */
      double V[8][3];
      for (int o = 0; o < 8; ++o) {
        double us[3][9];
        memcpy(us[0], (o & 4) ? dp[0] : dm[0], sizeof(us[0]));
        memcpy(us[1], (o & 2) ? dp[1] : dm[1], sizeof(us[1]));
        memcpy(us[2], (o & 1) ? dp[2] : dm[2], sizeof(us[2]));
        velocity(c, us, sg, V[o]);
      }
      double w[3][9];
      for (int a = 0; a < 3; ++a) {
        for (int o = 0; o < 8; ++o) w[a][o] = V[o][a];
        w[a][8] = 0.0;
      }
      const double a_p = max9(w[0]), a_m = fabs(min9(w[0]));
      const double b_p = max9(w[1]), b_m = fabs(min9(w[1]));
      const double c_p = max9(w[2]), c_m = fabs(min9(w[2]));
      const double a_tot = a_p + a_m + ME, b_tot = b_p + b_m + ME, c_tot = c_p + c_m + ME;
      double deriv[9];
      /* summation order '+++', '++-', '+-+', '+--', '-++', '-+-', '--+', '---' */
      for (int s = 0; s < 8; ++s) {
        const int o = 7 - s;
        double us[3][9], H[9];
        memcpy(us[0], (o & 4) ? dp[0] : dm[0], sizeof(us[0]));
        memcpy(us[1], (o & 2) ? dp[1] : dm[1], sizeof(us[1]));
        memcpy(us[2], (o & 1) ? dp[2] : dm[2], sizeof(us[2]));
        hamiltonian(3, us, V[o], lam, H);
        const double wa = (o & 4) ? a_m : a_p, wb = (o & 2) ? b_m : b_p, wc = (o & 1) ? c_m : c_p;
        for (int q = 0; q < 9; ++q) {
          const double term = H[q] * wa * wb * wc;
          deriv[q] = s == 0 ? -term : deriv[q] - term;
        }
      }
      for (int q = 0; q < 9; ++q) {
        const double tempa = (dp[0][q] - dm[0][q]) * a_m * a_p / a_tot;
        const double tempb = (dp[1][q] - dm[1][q]) * b_m * b_p / b_tot;
        const double tempc = (dp[2][q] - dm[2][q]) * c_m * c_p / c_tot;
        out[q] = deriv[q] / a_tot / b_tot / c_tot + tempa + tempb + tempc;
      }
      spd = a_tot + b_tot + c_tot;
    }
    for (int q = 0; q < 9; ++q) rhs[q * cells + idx] = out[q];
    if (speed) speed[idx] = spd;
  }
}

/* ------------------------------------------------------------ stepping */
typedef struct {
  const oc_cfg* c;
  double *sig, *rhs, *speed;
  long long cells;
} ws_t;

static void L_of(ws_t* w, double t, const double* u, int want_speed) {
  oc_sigma(w->c, u, t, w->sig);
  oc_flux(w->c, u, w->sig, w->rhs, want_speed ? w->speed : NULL);
}

/* Advance beta (9 planes, in place) by up to nsteps steps from t0 (stopping
 * at t_end); dts receives each step's dt.  Returns the steps taken; *t_out
 * the final time.  Every stage is evaluated at the step's start time. */
int oc_run(const oc_cfg* c, double* beta, double t0, int nsteps, double t_end, double* t_out, double* dts) {
  const int d = c->d, n = c->n;
  const long long cells = d == 3 ? (long long)n * n * n : (long long)n * n;
  const long long F = 9 * cells;
  ws_t w = {c, (double*)malloc(sizeof(double) * F), (double*)malloc(sizeof(double) * F),
            (double*)malloc(sizeof(double) * cells), cells};
  double* L0 = (double*)malloc(sizeof(double) * F);
  double* F0 = (double*)malloc(sizeof(double) * F);
  double* F1 = (double*)malloc(sizeof(double) * F);
  double* F2 = (double*)malloc(sizeof(double) * F);
  double t = t0;
  int s = 0;
  for (; s < nsteps; ++s) {
    L_of(&w, t, beta, 1);
    double mx = w.speed[0];
    for (long long i = 1; i < cells; ++i) mx = w.speed[i] > mx ? w.speed[i] : mx;
    double dt = c->fixed_dt > 0 ? c->fixed_dt : c->cfl / ((double)n * mx); /* FieldMover.py:94-96 */
    if (t + dt > t_end) dt = t_end - t;
    if (c->order == 3) {
      /* tvd_rk3_step, FieldMover.py:115-135 */
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < F; ++i) {
        L0[i] = dt * w.rhs[i];
        F0[i] = 1.0 * beta[i] + 1.0 * L0[i];
        F2[i] = 1.0 / 3.0 * beta[i] + 0. * L0[i] + 0. * F0[i];
      }
      L_of(&w, t, F0, 0);
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < F; ++i) {
        const double L1 = dt * w.rhs[i];
        F1[i] = 3.0 / 4.0 * beta[i] + 0. * L0[i] + 1.0 / 4.0 * F0[i] + 1.0 / 4.0 * L1;
        F2[i] = F2[i] + 0. * L1 + 2.0 / 3.0 * F1[i];
      }
      L_of(&w, t, F1, 0);
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < F; ++i) beta[i] = F2[i] + 2.0 / 3.0 * (dt * w.rhs[i]);
    } else {
      /* SSP-RK2: u1 = u + dt L(u); u' = u/2 + (u1 + dt L(u1))/2 */
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < F; ++i) F0[i] = beta[i] + dt * w.rhs[i];
      L_of(&w, t, F0, 0);
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < F; ++i) beta[i] = 0.5 * beta[i] + 0.5 * (F0[i] + dt * w.rhs[i]);
    }
    if (dts) dts[s] = dt;
    t += dt;
    if (t >= t_end) {
      ++s;
      break;
    }
  }
  free(w.sig);
  free(w.rhs);
  free(w.speed);
  free(L0);
  free(F0);
  free(F1);
  free(F2);
  if (t_out) *t_out = t;
  return s;
}
