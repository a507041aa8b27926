/* Oracle: pair, bonded and tail terms in FP64 — TEST INFRASTRUCTURE ONLY (see caps_oracle.h).
 *
 * orc_pairs_direct: every atom pair and periodic image within the cut-off, summed directly, as the reference's
 *   tests/nbnxm_check.cpp:78-119 does to check the cluster kernel; excluded pairs are left out at their nearest image
 *   only (core/src/nbnxm.cpp:241-249) and, when asked, contribute their Ewald / DSF correction (the Coulomb term less
 *   the part of the bare 1/r they drop, core/src/nbnxm.cpp:497-526).
 * orc_bonded: harmonic bonds and angles (core/src/field.cpp:1980-2003) and FENE + WCA (core/src/field.cpp:2150-2180),
 *   with the virial of each term (core/src/field.cpp:1866-1870).
 * orc_tail: the LJ tail correction (core/src/field.cpp:2501-2537).
 * Pinned by finite differences of the energies (forces, virial) and by the reference's cluster kernel in oracle/_ref.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "caps_oracle.h"

static const double kPi = 3.14159265358979323846;

static void frac_of(const double c[12], const double d[3], double f[3]) {
  const double *a = c + 3, *b = c + 6, *cc = c + 9;
  const double bc[3] = {b[1] * cc[2] - b[2] * cc[1], b[2] * cc[0] - b[0] * cc[2], b[0] * cc[1] - b[1] * cc[0]};
  const double ca[3] = {cc[1] * a[2] - cc[2] * a[1], cc[2] * a[0] - cc[0] * a[2], cc[0] * a[1] - cc[1] * a[0]};
  const double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double v = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
  f[0] = (d[0] * bc[0] + d[1] * bc[1] + d[2] * bc[2]) / v;
  f[1] = (d[0] * ca[0] + d[1] * ca[1] + d[2] * ca[2]) / v;
  f[2] = (d[0] * ab[0] + d[1] * ab[1] + d[2] * ab[2]) / v;
}

/* Cell::minimum_image (core/src/system.cpp:27-45): rounded fractional image, then the 27 neighbours of it */
static void min_image(const double c[12], double d[3]) {
  const double *a = c + 3, *b = c + 6, *cc = c + 9;
  double f[3];
  frac_of(c, d, f);
  for (int k = 0; k < 3; ++k) f[k] -= round(f[k]);
  double r[3];
  for (int k = 0; k < 3; ++k) r[k] = a[k] * f[0] + b[k] * f[1] + cc[k] * f[2];
  double best = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], bt[3] = {r[0], r[1], r[2]};
  const double r2 = best;
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k) {
        double t[3];
        for (int m = 0; m < 3; ++m) t[m] = r[m] + a[m] * i + b[m] * j + cc[m] * k;
        const double q = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        if (q < best - 1e-12) best = q, bt[0] = t[0], bt[1] = t[1], bt[2] = t[2];
      }
  if (best == r2) bt[0] = r[0], bt[1] = r[1], bt[2] = r[2];
  d[0] = bt[0], d[1] = bt[1], d[2] = bt[2];
}

static int in_csr(const int32_t* off, const int32_t* idx, int32_t i, int32_t j) {
  if (!off) return 0;
  for (int32_t e = off[i]; e < off[i + 1]; ++e)
    if (idx[e] == j) return 1;
  return 0;
}

static void add_w(double* w, double fr, const double d[3]) {
  w[0] += fr * d[0] * d[0]; w[1] += fr * d[1] * d[1]; w[2] += fr * d[2] * d[2];
  w[3] += fr * d[0] * d[1]; w[4] += fr * d[0] * d[2]; w[5] += fr * d[1] * d[2];
}

int orc_pairs_direct(const double* x, int32_t n, const double cell[12], const int32_t* type, int32_t nt, const double* c12,
                     const double* c6, int32_t lj, const double* q, double kcoul, int32_t coul, double kappa, double dsf_e0,
                     double dsf_f0, int32_t shift, double rc, const int32_t* excl_off, const int32_t* excl_idx,
                     const int32_t* p14_off, const int32_t* p14_idx, double f14, int32_t with_excl, double* e, double* f,
                     double* w) {
  const double *A = cell + 3, *B = cell + 6, *C = cell + 9;
  const double pw = lj == 2 ? 9.0 : 12.0, sp = 2 / sqrt(kPi), rc2 = rc * rc;
  /* image range: enough lattice vectors along each axis to reach rc from anywhere in the cell */
  double wd[3];
  {
    const double bc[3] = {B[1] * C[2] - B[2] * C[1], B[2] * C[0] - B[0] * C[2], B[0] * C[1] - B[1] * C[0]};
    const double ca[3] = {C[1] * A[2] - C[2] * A[1], C[2] * A[0] - C[0] * A[2], C[0] * A[1] - C[1] * A[0]};
    const double ab[3] = {A[1] * B[2] - A[2] * B[1], A[2] * B[0] - A[0] * B[2], A[0] * B[1] - A[1] * B[0]};
    const double v = fabs(A[0] * bc[0] + A[1] * bc[1] + A[2] * bc[2]);
    wd[0] = v / sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
    wd[1] = v / sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]);
    wd[2] = v / sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
  }
  int nim[3];
  for (int k = 0; k < 3; ++k) nim[k] = (int)ceil(rc / wd[k]) + 1;
  double ev = 0, ec = 0;
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = i; j < n; ++j) {
      const int excluded = j != i && excl_off && in_csr(excl_off, excl_idx, i, j);
      double dmin2 = 0;
      if (excluded) {
        double dm[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
        min_image(cell, dm);
        dmin2 = dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2];
      }
      for (int ia = -nim[0]; ia <= nim[0]; ++ia)
        for (int ib = -nim[1]; ib <= nim[1]; ++ib)
          for (int ic = -nim[2]; ic <= nim[2]; ++ic) {
            if (j == i && (ia > 0 || (ia == 0 && (ib > 0 || (ib == 0 && ic >= 0))))) continue;
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = x[3 * j + k] - x[3 * i + k] + ia * A[k] + ib * B[k] + ic * C[k];
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 >= rc2) continue;
            const double r = sqrt(r2);
            if (excluded && r2 <= dmin2 + 1e-6) {
              if (coul && with_excl) {
                const double qq = kcoul * q[i] * q[j], er = erfc(kappa * r);
                const double keep = in_csr(p14_off, p14_idx, i, j) ? f14 : 0.0;
                double E = er / r - (1 - keep) / r;
                double F = (er / r + sp * kappa * exp(-kappa * kappa * r2)) / r2 - (1 - keep) / (r2 * r);
                if (coul == 2) E += -dsf_e0 + dsf_f0 * (r - rc), F -= dsf_f0 / r;
                ec += qq * E;
                const double fr = qq * F;
                for (int k = 0; k < 3; ++k) f[3 * j + k] += fr * d[k], f[3 * i + k] -= fr * d[k];
                add_w(w, fr, d);
              }
              continue;
            }
            const size_t tp = (size_t)type[i] * (size_t)nt + (size_t)type[j];
            const double a12 = c12[tp], a6 = c6[tp];
            ev += a12 * pow(r, -pw) - a6 * pow(r, -6.0);
            if (shift) ev -= a12 * pow(rc, -pw) - a6 * pow(rc, -6.0);
            double fr = (pw * a12 * pow(r, -pw) - 6 * a6 * pow(r, -6.0)) / r2;
            if (coul) {
              const double qq = kcoul * q[i] * q[j], er = erfc(kappa * r), E = er / r;
              const double F = (er / r + sp * kappa * exp(-kappa * kappa * r2)) / r2;
              if (coul == 1) ec += qq * E, fr += qq * F;
              else ec += qq * (E - dsf_e0 + dsf_f0 * (r - rc)), fr += qq * (F - dsf_f0 / r);
            }
            for (int k = 0; k < 3; ++k) f[3 * j + k] += fr * d[k], f[3 * i + k] -= fr * d[k];
            add_w(w, fr, d);
          }
    }
  e[0] += ev;
  e[1] += ec;
  return 0;
}

/* the minimum image of the bonded terms (core/src/field.cpp:1328-1339): per axis for an orthogonal cell, else the
 * general one */
static void mi_bonded(const double c[12], double d[3]) {
  const double *a = c + 3, *b = c + 6, *cc = c + 9;
  if (a[1] == 0 && a[2] == 0 && b[0] == 0 && b[2] == 0 && cc[0] == 0 && cc[1] == 0) {
    const double L[3] = {a[0], b[1], cc[2]};
    for (int k = 0; k < 3; ++k) d[k] -= L[k] * round(d[k] / L[k]);
  } else {
    min_image(c, d);
  }
}

static void vir_term(double* e_vir, double* w, const double d[3], const double fv[3]) {
  *e_vir += d[0] * fv[0] + d[1] * fv[1] + d[2] * fv[2];
  w[0] += d[0] * fv[0]; w[1] += d[1] * fv[1]; w[2] += d[2] * fv[2];
  w[3] += 0.5 * (d[0] * fv[1] + d[1] * fv[0]); w[4] += 0.5 * (d[0] * fv[2] + d[2] * fv[0]);
  w[5] += 0.5 * (d[1] * fv[2] + d[2] * fv[1]);
}

int orc_bonded(const double* x, int32_t n, const double cell[12], int32_t nb, const int32_t* bij, const double* bkr,
               int32_t na, const int32_t* aijk, const double* akt, int32_t nf, const int32_t* fij, const double* fp,
               double* e, double* f, double* w) {
  (void)n;
  for (int32_t t = 0; t < nb; ++t) {
    const int32_t i = bij[2 * t], j = bij[2 * t + 1];
    double d[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
    mi_bonded(cell, d);
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), dr = r - bkr[2 * t + 1], k = bkr[2 * t];
    e[0] += k * dr * dr;
    double fj[3] = {0, 0, 0};
    if (r > 1e-12)
      for (int m = 0; m < 3; ++m) fj[m] = d[m] * (-2 * k * dr / r);
    for (int m = 0; m < 3; ++m) f[3 * j + m] += fj[m], f[3 * i + m] -= fj[m];
    vir_term(&e[2], w, d, fj);
  }
  for (int32_t t = 0; t < na; ++t) {
    const int32_t i = aijk[3 * t], j = aijk[3 * t + 1], k = aijk[3 * t + 2];
    double u[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
    double v[3] = {x[3 * k] - x[3 * j], x[3 * k + 1] - x[3 * j + 1], x[3 * k + 2] - x[3 * j + 2]};
    mi_bonded(cell, u);
    mi_bonded(cell, v);
    const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), lv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double c = (u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) / (lu * lv);
    c = c < -1 ? -1 : c > 1 ? 1 : c;
    const double th = acos(c), sn0 = sqrt(1 - c * c), sn = sn0 > 1e-8 ? sn0 : 1e-8;
    const double dth = th - akt[2 * t + 1], kt = akt[2 * t];
    e[1] += kt * dth * dth;
    const double g = 2 * kt * dth / sn;
    double fi[3], fk[3];
    for (int m = 0; m < 3; ++m) {
      fi[m] = (v[m] * (1 / (lu * lv)) - u[m] * (c / (lu * lu))) * g;
      fk[m] = (u[m] * (1 / (lu * lv)) - v[m] * (c / (lv * lv))) * g;
      f[3 * i + m] += fi[m];
      f[3 * k + m] += fk[m];
      f[3 * j + m] -= fi[m] + fk[m];
    }
    vir_term(&e[2], w, u, fi);
    vir_term(&e[2], w, v, fk);
  }
  for (int32_t t = 0; t < nf; ++t) {
    const int32_t i = fij[2 * t], j = fij[2 * t + 1];
    const double K = fp[4 * t], R0 = fp[4 * t + 1], eps = fp[4 * t + 2], sg = fp[4 * t + 3];
    double d[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
    mi_bonded(cell, d);
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), qq = r / R0;
    if (qq >= 1) return -1;
    double en = -0.5 * K * R0 * R0 * log(1 - qq * qq), dEdr = K * r / (1 - qq * qq);
    if (eps > 0 && r < pow(2.0, 1.0 / 6) * sg) {
      const double s6 = pow(sg / r, 6);
      en += 4 * eps * (s6 * s6 - s6) + eps;
      dEdr += -24 * eps * (2 * s6 * s6 - s6) / r;
    }
    e[0] += en;
    double fj[3] = {0, 0, 0};
    if (r > 1e-12)
      for (int m = 0; m < 3; ++m) fj[m] = d[m] * (-dEdr / r);
    for (int m = 0; m < 3; ++m) f[3 * j + m] += fj[m], f[3 * i + m] -= fj[m];
    vir_term(&e[2], w, d, fj);
  }
  return 0;
}

void orc_tail(int32_t nt, const double* eps, const double* sig6, const double* count, int32_t lj, double rc, double volume,
              double out[2]) {
  const double rc2 = rc * rc, rc3 = rc2 * rc, rc9 = rc3 * rc3 * rc3;
  double et = 0, wt = 0;
  for (int32_t a = 0; a < nt; ++a)
    for (int32_t b = 0; b < nt; ++b) {
      const size_t tp = (size_t)a * (size_t)nt + (size_t)b;
      const double ep = eps[tp], s6 = sig6[tp], s12 = s6 * s6, nn = count[a] * count[b];
      if (lj == 2) {
        const double s3 = sqrt(s6), rc6 = rc3 * rc3;
        et += nn * 2 * kPi * ep * s6 * (s3 - 3 * rc3) / (3 * rc6);
        wt += nn * 3 * 2 * kPi * ep * s6 * (s3 - 2 * rc3) / rc6;
        continue;
      }
      et += nn * 2 * kPi * 4 * ep * (s12 / (9 * rc9) - s6 / (3 * rc3));
      wt += nn * 2 * kPi * 4 * ep * (4 * s12 / (3 * rc9) - 2 * s6 / rc3);
    }
  out[0] = et / volume;
  out[1] = wt / volume;
}
