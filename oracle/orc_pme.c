/* Oracle: Ewald coefficient, PME grid sizes, smooth PME and the plain Ewald sum in FP64 — TEST INFRASTRUCTURE ONLY.
 *
 * Restates core/src/kspace.cpp:35-52 and core/fortran/caps_kspace.f90 (good_size :23-36, green :518-564,
 * bspline :568-598, bspline_moduli :601-620, caps_f_pme :625-715, caps_f_ewald :718-764). The FFT is a plain
 * recursive mixed-radix Cooley–Tukey transform (any length), X(m) = Σ x(k) exp(sign 2πi mk/n), unnormalised.
 * Pinned by: the direct DFT (tests), the Madelung constant of rock salt, PME against the plain Ewald sum, the
 * partition of unity of the B-splines, finite differences of the energy, and the reference build in oracle/_ref.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "caps_oracle.h"

static const double kPi = 3.14159265358979323846;

double orc_ewald_beta(double cutoff, double rtol) {
  double lo = 0, hi = 20.0 / cutoff;
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (erfc(mid * cutoff) > rtol) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

int32_t orc_good_size(int32_t n) {
  int32_t m = n > 1 ? n : 1;
  for (;;) {
    int32_t k = m;
    while (k % 2 == 0) k /= 2;
    while (k % 3 == 0) k /= 3;
    while (k % 5 == 0) k /= 5;
    while (k % 7 == 0) k /= 7;
    if (k == 1) return m;
    ++m;
  }
}

void orc_pme_grid(const double cell[12], double spacing, int32_t order, int32_t k[3]) {
  for (int d = 0; d < 3; ++d) {
    const double* e = cell + 3 + 3 * d;
    const double len = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    const int32_t need = (int32_t)ceil(len / spacing);
    k[d] = orc_good_size(need > order + 1 ? need : order + 1);
  }
}

void orc_bspline(double w, int32_t p, double* theta, double* dtheta) {
  for (int j = 0; j < p; ++j) theta[j] = 0;
  theta[1] = w;
  theta[0] = 1 - w;
  for (int k = 3; k <= p - 1; ++k) {
    const double div = 1.0 / (k - 1);
    theta[k - 1] = div * w * theta[k - 2];
    for (int j = 1; j <= k - 2; ++j) theta[k - j - 1] = div * ((w + j) * theta[k - j - 2] + (k - j - w) * theta[k - j - 1]);
    theta[0] = div * (1 - w) * theta[0];
  }
  dtheta[0] = -theta[0];
  for (int j = 2; j <= p; ++j) dtheta[j - 1] = theta[j - 2] - theta[j - 1];
  const double div = 1.0 / (p - 1);
  theta[p - 1] = div * w * theta[p - 2];
  for (int j = 1; j <= p - 2; ++j) theta[p - j - 1] = div * ((w + j) * theta[p - j - 2] + (p - j - w) * theta[p - j - 1]);
  theta[0] = div * (1 - w) * theta[0];
}

void orc_bspline_moduli(int32_t k, int32_t p, double* bsp) {
  double arr[32], darr[32];
  orc_bspline(0.0, p, arr, darr);
  for (int m = 0; m < k; ++m) {
    double sr = 0, si = 0;
    for (int j = 0; j <= p - 2; ++j) {
      const double arg = 2 * kPi * m * j / (double)k;
      sr += arr[j] * cos(arg);
      si += arr[j] * sin(arg);
    }
    bsp[m] = sr * sr + si * si;
  }
  for (int m = 0; m < k; ++m)
    if (bsp[m] < 1e-7) bsp[m] = 0.5 * (bsp[(m - 1 + k) % k] + bsp[(m + 1) % k]);
}

/* recursive mixed radix: n = p·m with p the smallest prime factor */
static void fft_rec(int32_t n, int32_t stride, int32_t sign, const double* ire, const double* iim, double* ore, double* oim) {
  if (n == 1) { ore[0] = ire[0]; oim[0] = iim[0]; return; }
  int32_t p = 2;
  while (n % p) ++p;
  const int32_t m = n / p;
  double* sre = (double*)malloc(sizeof(double) * (size_t)n);
  double* sim = (double*)malloc(sizeof(double) * (size_t)n);
  for (int32_t r = 0; r < p; ++r) fft_rec(m, stride * p, sign, ire + (size_t)r * stride, iim + (size_t)r * stride, sre + (size_t)r * m, sim + (size_t)r * m);
  for (int32_t kk = 0; kk < n; ++kk) {
    double acc_re = 0, acc_im = 0;
    for (int32_t r = 0; r < p; ++r) {
      const double ang = sign * 2 * kPi * (double)((int64_t)r * kk % n) / n;
      const double c = cos(ang), s = sin(ang);
      const double a = sre[(size_t)r * m + kk % m], b = sim[(size_t)r * m + kk % m];
      acc_re += a * c - b * s;
      acc_im += a * s + b * c;
    }
    ore[kk] = acc_re;
    oim[kk] = acc_im;
  }
  free(sre);
  free(sim);
}

void orc_fft1d(int32_t n, int32_t sign, const double* in_re, const double* in_im, double* out_re, double* out_im) {
  fft_rec(n, 1, sign, in_re, in_im, out_re, out_im);
}

void orc_fft3d(int32_t k1, int32_t k2, int32_t k3, int32_t sign, double* re, double* im) {
  const int32_t kmax = k1 > k2 ? (k1 > k3 ? k1 : k3) : (k2 > k3 ? k2 : k3);
  double* lr = (double*)malloc(sizeof(double) * (size_t)kmax);
  double* li = (double*)malloc(sizeof(double) * (size_t)kmax);
  double* orr = (double*)malloc(sizeof(double) * (size_t)kmax);
  double* oi = (double*)malloc(sizeof(double) * (size_t)kmax);
  const int32_t ks[3] = {k1, k2, k3};
  const size_t st[3] = {1, (size_t)k1, (size_t)k1 * k2};
  for (int ax = 0; ax < 3; ++ax) {
    const int32_t n = ks[ax];
    const int32_t o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
    for (int32_t p = 0; p < ks[o1]; ++p)
      for (int32_t q = 0; q < ks[o2]; ++q) {
        const size_t base = (size_t)p * st[o1] + (size_t)q * st[o2];
        for (int32_t t = 0; t < n; ++t) lr[t] = re[base + t * st[ax]], li[t] = im[base + t * st[ax]];
        fft_rec(n, 1, sign, lr, li, orr, oi);
        for (int32_t t = 0; t < n; ++t) re[base + t * st[ax]] = orr[t], im[base + t * st[ax]] = oi[t];
      }
  }
  free(lr); free(li); free(orr); free(oi);
}

/* h⁻¹ rows = reciprocal vectors (caps_kspace.f90:484-500): cell rows a, b, c are the columns of h */
static void cell_inverse(const double cell[12], double hi[3][3], double* vol) {
  double h[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) h[r][c] = cell[3 + 3 * c + r];
  const double v = h[0][0] * (h[1][1] * h[2][2] - h[2][1] * h[1][2]) - h[0][1] * (h[1][0] * h[2][2] - h[2][0] * h[1][2]) +
                   h[0][2] * (h[1][0] * h[2][1] - h[2][0] * h[1][1]);
  hi[0][0] = (h[1][1] * h[2][2] - h[1][2] * h[2][1]) / v;
  hi[0][1] = (h[0][2] * h[2][1] - h[0][1] * h[2][2]) / v;
  hi[0][2] = (h[0][1] * h[1][2] - h[0][2] * h[1][1]) / v;
  hi[1][0] = (h[1][2] * h[2][0] - h[1][0] * h[2][2]) / v;
  hi[1][1] = (h[0][0] * h[2][2] - h[0][2] * h[2][0]) / v;
  hi[1][2] = (h[0][2] * h[1][0] - h[0][0] * h[1][2]) / v;
  hi[2][0] = (h[1][0] * h[2][1] - h[1][1] * h[2][0]) / v;
  hi[2][1] = (h[0][1] * h[2][0] - h[0][0] * h[2][1]) / v;
  hi[2][2] = (h[0][0] * h[1][1] - h[0][1] * h[1][0]) / v;
  *vol = fabs(v);
}

static void green_at(int32_t a, int32_t b, int32_t c, const int32_t k[3], double hi[3][3], double vol, double fac, double kcoul,
                     const double* b1, const double* b2, const double* b3, double* g, double w[6]) {
  const int32_t idx[3] = {a, b, c};
  int32_t m[3], nyq[3], nn = 0;
  for (int d = 0; d < 3; ++d) {
    m[d] = idx[d];
    if (m[d] >= (k[d] + 1) / 2) m[d] -= k[d];
    nyq[d] = (k[d] % 2 == 0) && 2 * idx[d] == k[d];
    nn += nyq[d];
  }
  const double bm = b1[a] * b2[b] * b3[c];
  const int ns = 1 << nn;
  *g = 0;
  for (int q = 0; q < 6; ++q) w[q] = 0;
  for (int s = 0; s < ns; ++s) {
    int sel = 0;
    double mv[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
      double sg = 1;
      if (nyq[d]) {
        if ((s >> sel) & 1) sg = -1;
        ++sel;
      }
      for (int q = 0; q < 3; ++q) mv[q] += sg * m[d] * hi[d][q];
    }
    const double m2 = mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2];
    const double e = kcoul * exp(-fac * m2) / (kPi * vol * m2 * bm);
    const double ss = 2 * (1 + fac * m2) / m2;
    *g += e;
    w[0] += e * (1 - ss * mv[0] * mv[0]);
    w[1] += e * (1 - ss * mv[1] * mv[1]);
    w[2] += e * (1 - ss * mv[2] * mv[2]);
    w[3] -= e * ss * mv[0] * mv[1];
    w[4] -= e * ss * mv[0] * mv[2];
    w[5] -= e * ss * mv[1] * mv[2];
  }
  *g /= ns;
  for (int q = 0; q < 6; ++q) w[q] /= ns;
}

void orc_pme_green(int32_t a, int32_t b, int32_t c, const int32_t k[3], const double cell[12], int32_t order, double beta,
                   double kcoul, double* g, double w[6]) {
  double hi[3][3], vol;
  cell_inverse(cell, hi, &vol);
  double* b1 = (double*)malloc(sizeof(double) * (size_t)k[0]);
  double* b2 = (double*)malloc(sizeof(double) * (size_t)k[1]);
  double* b3 = (double*)malloc(sizeof(double) * (size_t)k[2]);
  orc_bspline_moduli(k[0], order, b1);
  orc_bspline_moduli(k[1], order, b2);
  orc_bspline_moduli(k[2], order, b3);
  green_at(a, b, c, k, hi, vol, kPi * kPi / (beta * beta), kcoul, b1, b2, b3, g, w);
  free(b1); free(b2); free(b3);
}

double orc_pme(const double* x, const double* q, int32_t n, const double cell[12], const int32_t k[3], int32_t p,
               double beta, double kcoul, double* f, double* vir) {
  double hi[3][3], vol;
  cell_inverse(cell, hi, &vol);
  const size_t K = (size_t)k[0] * k[1] * k[2];
  double* th = (double*)malloc(sizeof(double) * (size_t)p * 3 * (size_t)n);
  double* dth = (double*)malloc(sizeof(double) * (size_t)p * 3 * (size_t)n);
  int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)p * 3 * (size_t)n);
  for (int32_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      const double s = hi[d][0] * x[3 * i] + hi[d][1] * x[3 * i + 1] + hi[d][2] * x[3 * i + 2];
      const double u = (s - floor(s)) * k[d], w = u - floor(u);
      const size_t o = ((size_t)i * 3 + d) * (size_t)p;
      orc_bspline(w, p, th + o, dth + o);
      for (int a = 1; a <= p; ++a) {
        int32_t t = ((int32_t)floor(u) - p + a) % k[d];
        if (t < 0) t += k[d];
        idx[o + a - 1] = t;
      }
    }
  double* gr = (double*)calloc(K, sizeof(double));
  double* gi = (double*)calloc(K, sizeof(double));
#define AT(a, b, c) ((size_t)(a) + (size_t)k[0] * ((size_t)(b) + (size_t)k[1] * (size_t)(c)))
  for (int32_t i = 0; i < n; ++i) {
    const size_t o0 = ((size_t)i * 3) * p, o1 = o0 + p, o2 = o1 + p;
    for (int c = 0; c < p; ++c)
      for (int b = 0; b < p; ++b) {
        const double st = q[i] * th[o1 + b] * th[o2 + c];
        for (int a = 0; a < p; ++a) gr[AT(idx[o0 + a], idx[o1 + b], idx[o2 + c])] += st * th[o0 + a];
      }
  }
  orc_fft3d(k[0], k[1], k[2], +1, gr, gi);
  double* b1 = (double*)malloc(sizeof(double) * (size_t)k[0]);
  double* b2 = (double*)malloc(sizeof(double) * (size_t)k[1]);
  double* b3 = (double*)malloc(sizeof(double) * (size_t)k[2]);
  orc_bspline_moduli(k[0], p, b1);
  orc_bspline_moduli(k[1], p, b2);
  orc_bspline_moduli(k[2], p, b3);
  const double fac = kPi * kPi / (beta * beta);
  double energy = 0;
  for (int q6 = 0; q6 < 6; ++q6) vir[q6] = 0;
  for (int c = 0; c < k[2]; ++c)
    for (int b = 0; b < k[1]; ++b)
      for (int a = 0; a < k[0]; ++a) {
        const size_t at = AT(a, b, c);
        if (a == 0 && b == 0 && c == 0) { gr[at] = gi[at] = 0; continue; }
        double g, wv[6];
        green_at(a, b, c, k, hi, vol, fac, kcoul, b1, b2, b3, &g, wv);
        const double em = 0.5 * (gr[at] * gr[at] + gi[at] * gi[at]);
        energy += em * g;
        for (int q6 = 0; q6 < 6; ++q6) vir[q6] += em * wv[q6];
        gr[at] *= g;
        gi[at] *= g;
      }
  orc_fft3d(k[0], k[1], k[2], -1, gr, gi);
  for (int32_t i = 0; i < n; ++i) {
    const size_t o0 = ((size_t)i * 3) * p, o1 = o0 + p, o2 = o1 + p;
    double gs[3] = {0, 0, 0};
    for (int c = 0; c < p; ++c)
      for (int b = 0; b < p; ++b)
        for (int a = 0; a < p; ++a) {
          const double st = gr[AT(idx[o0 + a], idx[o1 + b], idx[o2 + c])];
          gs[0] += dth[o0 + a] * th[o1 + b] * th[o2 + c] * st;
          gs[1] += th[o0 + a] * dth[o1 + b] * th[o2 + c] * st;
          gs[2] += th[o0 + a] * th[o1 + b] * dth[o2 + c] * st;
        }
    for (int m = 0; m < 3; ++m) {
      const double grad = k[0] * gs[0] * hi[0][m] + k[1] * gs[1] * hi[1][m] + k[2] * gs[2] * hi[2][m];
      f[3 * i + m] -= q[i] * grad;
    }
  }
#undef AT
  free(th); free(dth); free(idx); free(gr); free(gi); free(b1); free(b2); free(b3);
  return energy;
}

double orc_ewald_recip(const double* x, const double* q, int32_t n, const double cell[12], const int32_t kmax[3],
                       double beta, double kcoul, double* f, double* vir) {
  double hi[3][3], vol;
  cell_inverse(cell, hi, &vol);
  const double fac = kPi * kPi / (beta * beta);
  double energy = 0;
  for (int q6 = 0; q6 < 6; ++q6) vir[q6] = 0;
  double* cs = (double*)malloc(sizeof(double) * (size_t)n);
  double* sn = (double*)malloc(sizeof(double) * (size_t)n);
  for (int m1 = -kmax[0]; m1 <= kmax[0]; ++m1)
    for (int m2 = -kmax[1]; m2 <= kmax[1]; ++m2)
      for (int m3 = -kmax[2]; m3 <= kmax[2]; ++m3) {
        if (!m1 && !m2 && !m3) continue;
        double mv[3];
        for (int d = 0; d < 3; ++d) mv[d] = m1 * hi[0][d] + m2 * hi[1][d] + m3 * hi[2][d];
        const double m2l = mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2];
        const double a = kcoul * exp(-fac * m2l) / (2 * kPi * vol * m2l);
        double sr = 0, si = 0;
        for (int32_t i = 0; i < n; ++i) {
          const double arg = 2 * kPi * (mv[0] * x[3 * i] + mv[1] * x[3 * i + 1] + mv[2] * x[3 * i + 2]);
          cs[i] = cos(arg), sn[i] = sin(arg);
          sr += q[i] * cs[i];
          si += q[i] * sn[i];
        }
        const double em = a * (sr * sr + si * si);
        energy += em;
        const double ss = 2 * (1 + fac * m2l) / m2l;
        vir[0] += em * (1 - ss * mv[0] * mv[0]);
        vir[1] += em * (1 - ss * mv[1] * mv[1]);
        vir[2] += em * (1 - ss * mv[2] * mv[2]);
        vir[3] -= em * ss * mv[0] * mv[1];
        vir[4] -= em * ss * mv[0] * mv[2];
        vir[5] -= em * ss * mv[1] * mv[2];
        for (int32_t i = 0; i < n; ++i)
          for (int d = 0; d < 3; ++d) f[3 * i + d] += 4 * kPi * a * q[i] * mv[d] * (sr * sn[i] - si * cs[i]);
      }
  free(cs);
  free(sn);
  return energy;
}
