/* Oracle: the cluster pair list — TEST INFRASTRUCTURE ONLY (see caps_oracle.h).
 *
 * A restatement of core/src/nbnxm.cpp:22-390 in plain C with the same floating-point operations in the same order, so
 * the result is bit-identical: FP64 binning and bounding boxes without FMA contraction (this file is compiled with
 * -ffp-contract=off), and the FP32 4x4 distance masks with SSE2's unfused multiply and add (core/src/simd4.hpp:79-80):
 * d² = dx·dx + (dy·dy + dz·dz), each operation rounded.
 * Pinned against the reference's own List::build (oracle/_ref) and by completeness invariants (tests/test_oracle_list.py).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "caps_oracle.h"

static char g_err[256];
const char* orc_last_error(void) { return g_err; }

/* Cell helpers with the reference's operation order (core/src/system.cpp:14-23, core/include/caps/system.hpp:20-25) */
typedef struct { double v[3]; } v3;
static v3 mk(double a, double b, double c) { v3 r = {{a, b, c}}; return r; }
static double dot3(v3 a, v3 b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
static v3 cross3(v3 a, v3 b) {
  return mk(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
static v3 add3(v3 a, v3 b) { return mk(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
static v3 sub3(v3 a, v3 b) { return mk(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
static v3 scl3(v3 a, double s) { return mk(a.v[0] * s, a.v[1] * s, a.v[2] * s); }
static double norm3(v3 a) { return sqrt(dot3(a, a)); }

typedef struct { v3 o, a, b, c; } cellt;
static cellt cell_of(const double c[12]) {
  cellt r;
  for (int k = 0; k < 3; ++k) r.o.v[k] = c[k], r.a.v[k] = c[3 + k], r.b.v[k] = c[6 + k], r.c.v[k] = c[9 + k];
  return r;
}
static double cell_volume(const cellt* c) { return fabs(dot3(c->a, cross3(c->b, c->c))); }
static v3 to_fractional(const cellt* c, v3 r) {
  const v3 d = sub3(r, c->o);
  const double v = dot3(c->a, cross3(c->b, c->c));
  return mk(dot3(d, cross3(c->b, c->c)) / v, dot3(d, cross3(c->c, c->a)) / v, dot3(d, cross3(c->a, c->b)) / v);
}
static double clampd(double x, double lo, double hi) { return x < lo ? lo : x > hi ? hi : x; }
static int clampi(int x, int lo, int hi) { return x < lo ? lo : x > hi ? hi : x; }

struct orc_list {
  int32_t n, nu, nv;
  int64_t nclusters, ns, nentries, ncodes;
  uint32_t *perm, *iperm;
  int64_t* first;
  uint32_t* cj;
  uint16_t *code, *mask, *ex, *ex14;
  double *org, *off, *qvec;
  int32_t *ioff, *qint;
  float *bx, *by, *bz;
};

void orc_list_free(orc_list* L) {
  if (!L) return;
  free(L->perm); free(L->iperm); free(L->first); free(L->cj); free(L->code); free(L->mask); free(L->ex); free(L->ex14);
  free(L->org); free(L->off); free(L->qvec); free(L->ioff); free(L->qint); free(L->bx); free(L->by); free(L->bz);
  free(L);
}

/* per-column sort key (nbnxm.cpp:70-71): f_c, then the atom index */
static const double* g_fc;
static int cmp_col(const void* p, const void* r) {
  const uint32_t a = *(const uint32_t*)p, b = *(const uint32_t*)r;
  if (g_fc[a] != g_fc[b]) return g_fc[a] < g_fc[b] ? -1 : 1;
  return a < b ? -1 : a > b ? 1 : 0;
}
typedef struct { uint32_t cj; uint16_t bit, is14; } xitem;
static int cmp_x(const void* p, const void* r) {
  const xitem* a = (const xitem*)p; const xitem* b = (const xitem*)r;
  if (a->cj != b->cj) return a->cj < b->cj ? -1 : 1;
  return a->bit < b->bit ? -1 : a->bit > b->bit ? 1 : 0;
}
static int has(const int32_t* off, const int32_t* idx, uint32_t i, uint32_t j) {
  int32_t lo = off[i], hi = off[i + 1];
  while (lo < hi) {
    const int32_t mid = (lo + hi) / 2;
    if ((uint32_t)idx[mid] < j) lo = mid + 1; else hi = mid;
  }
  return lo < off[i + 1] && (uint32_t)idx[lo] == j;
}

typedef struct { uint32_t cj; uint16_t code, mask, ex, ex14; } entry;

orc_list* orc_list_build(const double* x, int32_t n, const double cellv[12], double rl, const int32_t* excl_off,
                         const int32_t* excl_idx, const int32_t* p14_off, const int32_t* p14_idx, int32_t keep_excluded) {
  const cellt cell = cell_of(cellv);
  const double v = cell_volume(&cell);
  if (!(v > 1e-9)) { snprintf(g_err, sizeof g_err, "the list needs a periodic cell"); return NULL; }
  orc_list* L = (orc_list*)calloc(1, sizeof(orc_list));
  L->n = n;
  const double w[3] = {v / norm3(cross3(cell.b, cell.c)), v / norm3(cross3(cell.c, cell.a)), v / norm3(cross3(cell.a, cell.b))};
  const double side = cbrt(4.0 * v / (n > 1 ? (double)n : 1.0));
  const int nu = clampi((int)(w[0] / side), 1, 1024), nv = clampi((int)(w[1] / side), 1, 1024);
  L->nu = nu, L->nv = nv;
  v3* fr = (v3*)malloc(sizeof(v3) * (size_t)(n > 0 ? n : 1));
  uint32_t* col = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  const size_t ncols = (size_t)nu * (size_t)nv;
  int64_t* cstart = (int64_t*)calloc(ncols + 1, sizeof(int64_t));
  for (int32_t i = 0; i < n; ++i) {
    v3 f = to_fractional(&cell, mk(x[3 * i], x[3 * i + 1], x[3 * i + 2]));
    for (int k = 0; k < 3; ++k) f.v[k] = clampd(f.v[k] - floor(f.v[k]), 0.0, 1.0 - 1e-12);
    fr[i] = f;
    const double fu = clampd(f.v[0], 0.0, 1.0 - 1e-12), fv = clampd(f.v[1], 0.0, 1.0 - 1e-12);
    int cu = (int)(fu * nu), cv = (int)(fv * nv);
    if (cu > nu - 1) cu = nu - 1;
    if (cv > nv - 1) cv = nv - 1;
    col[i] = (uint32_t)(cu * nv + cv);
    ++cstart[col[i] + 1];
  }
  for (size_t c = 0; c < ncols; ++c) cstart[c + 1] += cstart[c];
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  {
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * ncols);
    memcpy(fill, cstart, sizeof(int64_t) * ncols);
    for (int32_t i = 0; i < n; ++i) order[fill[col[i]]++] = (uint32_t)i;
    free(fill);
  }
  double* fc = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int32_t i = 0; i < n; ++i) fc[i] = fr[i].v[2];
  g_fc = fc;
  int64_t* ccl = (int64_t*)calloc(ncols + 1, sizeof(int64_t));
  for (size_t c = 0; c < ncols; ++c) {
    qsort(order + cstart[c], (size_t)(cstart[c + 1] - cstart[c]), sizeof(uint32_t), cmp_col);
    ccl[c + 1] = ccl[c] + (cstart[c + 1] - cstart[c] + 3) / 4;
  }
  free(fc);
  const int64_t ncl = ccl[ncols], ns = 4 * ncl;
  L->nclusters = ncl, L->ns = ns;
  L->perm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(ns > 0 ? ns : 1));
  L->iperm = (uint32_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(uint32_t));
  for (int64_t s = 0; s < ns; ++s) L->perm[s] = ORC_PAD;
  uint32_t* ccol = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(ncl > 0 ? ncl : 1));
  for (size_t c = 0; c < ncols; ++c)
    for (int64_t k = cstart[c]; k < cstart[c + 1]; ++k) {
      const int64_t s = 4 * ccl[c] + (k - cstart[c]);
      L->perm[s] = order[k];
      L->iperm[order[k]] = (uint32_t)s;
    }
  for (size_t c = 0; c < ncols; ++c)
    for (int64_t k = ccl[c]; k < ccl[c + 1]; ++k) ccol[k] = (uint32_t)c;
  /* wrapped Cartesian positions by slot (padding: the cluster's first atom), boxes and c ranges (nbnxm.cpp:87-105) */
  v3* cw = (v3*)malloc(sizeof(v3) * (size_t)(ns > 0 ? ns : 1));
  v3* frs = (v3*)malloc(sizeof(v3) * (size_t)(ns > 0 ? ns : 1));
  for (int64_t s = 0; s < ns; ++s) {
    const uint32_t a = L->perm[s] != ORC_PAD ? L->perm[s] : L->perm[s & ~(int64_t)3];
    frs[s] = fr[a];
    cw[s] = add3(add3(add3(cell.o, scl3(cell.a, fr[a].v[0])), scl3(cell.b, fr[a].v[1])), scl3(cell.c, fr[a].v[2]));
  }
  v3* blo = (v3*)malloc(sizeof(v3) * (size_t)(ncl > 0 ? ncl : 1));
  v3* bhi = (v3*)malloc(sizeof(v3) * (size_t)(ncl > 0 ? ncl : 1));
  double* wlo = (double*)malloc(sizeof(double) * (size_t)(ncl > 0 ? ncl : 1));
  double* whi = (double*)malloc(sizeof(double) * (size_t)(ncl > 0 ? ncl : 1));
  for (int64_t c = 0; c < ncl; ++c) {
    blo[c] = bhi[c] = cw[4 * c];
    wlo[c] = whi[c] = frs[4 * c].v[2];
    for (int l = 1; l < 4; ++l)
      for (int k = 0; k < 3; ++k) {
        if (cw[4 * c + l].v[k] < blo[c].v[k]) blo[c].v[k] = cw[4 * c + l].v[k];
        if (cw[4 * c + l].v[k] > bhi[c].v[k]) bhi[c].v[k] = cw[4 * c + l].v[k];
      }
    for (int l = 1; l < 4; ++l) {
      if (frs[4 * c + l].v[2] < wlo[c]) wlo[c] = frs[4 * c + l].v[2];
      if (frs[4 * c + l].v[2] > whi[c]) whi[c] = frs[4 * c + l].v[2];
    }
  }
  L->org = (double*)malloc(sizeof(double) * (size_t)(3 * ncl > 0 ? 3 * ncl : 1));
  for (int64_t c = 0; c < ncl; ++c)
    for (int k = 0; k < 3; ++k) L->org[3 * c + k] = 0.5 * (blo[c].v[k] + bhi[c].v[k]);
  L->bx = (float*)malloc(sizeof(float) * (size_t)(ns > 0 ? ns : 1));
  L->by = (float*)malloc(sizeof(float) * (size_t)(ns > 0 ? ns : 1));
  L->bz = (float*)malloc(sizeof(float) * (size_t)(ns > 0 ? ns : 1));
  uint8_t* real = (uint8_t*)calloc((size_t)(ncl > 0 ? ncl : 1), 1);
  for (int64_t s = 0; s < ns; ++s) {
    const double* o = &L->org[3 * (s / 4)];
    L->bx[s] = (float)(cw[s].v[0] - o[0]), L->by[s] = (float)(cw[s].v[1] - o[1]), L->bz[s] = (float)(cw[s].v[2] - o[2]);
    if (L->perm[s] != ORC_PAD) real[s / 4] |= (uint8_t)(1u << (s % 4));
  }
  /* excluded atom pairs by cluster pair, merged per j cluster (nbnxm.cpp:118-145) */
  int64_t* xfirst = (int64_t*)calloc((size_t)ncl + 1, sizeof(int64_t));
  int64_t xcap = 1024, nxm = 0;
  uint32_t* xcj = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)xcap);
  uint16_t* xbits = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)xcap);
  uint16_t* x14 = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)xcap);
  {
    int64_t tcap = 64;
    xitem* tmp = (xitem*)malloc(sizeof(xitem) * (size_t)tcap);
    for (int64_t c = 0; c < ncl; ++c) {
      int64_t nt = 0;
      for (int a = 0; a < 4; ++a) {
        const uint32_t ai = L->perm[4 * c + a];
        if (ai == ORC_PAD) continue;
        for (int32_t e = excl_off[ai]; e < excl_off[ai + 1]; ++e) {
          const uint32_t aj = (uint32_t)excl_idx[e];
          const uint32_t sj = L->iperm[aj];
          const uint16_t bit = (uint16_t)(1u << (4 * a + (int)(sj % 4)));
          const int k14 = keep_excluded && p14_off && has(p14_off, p14_idx, ai, aj);
          if (nt == tcap) tcap *= 2, tmp = (xitem*)realloc(tmp, sizeof(xitem) * (size_t)tcap);
          tmp[nt].cj = sj / 4, tmp[nt].bit = bit, tmp[nt].is14 = (uint16_t)(k14 ? bit : 0);
          ++nt;
        }
      }
      qsort(tmp, (size_t)nt, sizeof(xitem), cmp_x);
      for (int64_t t = 0; t < nt; ++t) {
        if (nxm > xfirst[c] && xcj[nxm - 1] == tmp[t].cj) {
          xbits[nxm - 1] |= tmp[t].bit, x14[nxm - 1] |= tmp[t].is14;
        } else {
          if (nxm == xcap) {
            xcap *= 2;
            xcj = (uint32_t*)realloc(xcj, sizeof(uint32_t) * (size_t)xcap);
            xbits = (uint16_t*)realloc(xbits, sizeof(uint16_t) * (size_t)xcap);
            x14 = (uint16_t*)realloc(x14, sizeof(uint16_t) * (size_t)xcap);
          }
          xcj[nxm] = tmp[t].cj, xbits[nxm] = tmp[t].bit, x14[nxm] = tmp[t].is14;
          ++nxm;
        }
      }
      xfirst[c + 1] = nxm;
    }
    free(tmp);
  }
  /* reach in columns and images along c; image codes (nbnxm.cpp:146-162) */
  const int ru = (int)floor(rl * nu / w[0]) + 1, rv = (int)floor(rl * nv / w[1]) + 1, rq = (int)ceil(rl / w[2]) + 1;
  const double dw = rl / w[2];
  int qlo[3], span[3];
  qlo[0] = (int)floor((double)(-ru) / nu);
  qlo[1] = (int)floor((double)(-rv) / nv);
  qlo[2] = -rq;
  span[0] = (int)floor((double)(nu - 1 + ru) / nu) - qlo[0] + 1;
  span[1] = (int)floor((double)(nv - 1 + rv) / nv) - qlo[1] + 1;
  span[2] = 2 * rq + 1;
  if ((int64_t)span[0] * span[1] * span[2] > 65536) {
    snprintf(g_err, sizeof g_err, "the cell is too small for the cut-off: more than 65536 periodic images in range");
    orc_list_free(L);
    L = NULL;
    goto done;
  }
  L->ncodes = (int64_t)span[0] * span[1] * span[2];
  L->qint = (int32_t*)malloc(sizeof(int32_t) * (size_t)(3 * L->ncodes));
  {
    int64_t q = 0;
    for (int a = 0; a < span[0]; ++a)
      for (int b = 0; b < span[1]; ++b)
        for (int c = 0; c < span[2]; ++c) L->qint[q++] = qlo[0] + a, L->qint[q++] = qlo[1] + b, L->qint[q++] = qlo[2] + c;
  }
  {
    const double rl2 = rl * rl;
    const float rl2f = (float)(rl2 * (1 + 1e-5));
    int64_t cap = 1 << 16, ne = 0;
    entry* E = (entry*)malloc(sizeof(entry) * (size_t)cap);
    L->first = (int64_t*)calloc((size_t)ncl + 1, sizeof(int64_t));
    for (int64_t ci = 0; ci < ncl; ++ci) {
      const int cu = (int)ccol[ci] / nv, cv = (int)ccol[ci] % nv;
      for (int du = -ru; du <= ru; ++du) {
        int U = cu + du;
        const int qa = (int)floor((double)U / nu);
        U -= qa * nu;
        for (int dvv = -rv; dvv <= rv; ++dvv) {
          int V = cv + dvv;
          const int qb = (int)floor((double)V / nv);
          V -= qb * nv;
          const size_t cc = (size_t)U * nv + V;
          int qc0 = (int)floor(wlo[ci] - dw), qc1 = (int)floor(whi[ci] + dw);
          if (qc0 < -rq) qc0 = -rq;
          if (qc1 > rq) qc1 = rq;
          for (int qc = qc0; qc <= qc1; ++qc) {
            const v3 S = add3(add3(scl3(cell.a, (double)qa), scl3(cell.b, (double)qb)), scl3(cell.c, (double)qc));
            const int zero = qa == 0 && qb == 0 && qc == 0;
            const int positive = qa > 0 || (qa == 0 && (qb > 0 || (qb == 0 && qc > 0)));
            const uint16_t code = (uint16_t)(((qa - qlo[0]) * span[1] + (qb - qlo[1])) * span[2] + (qc - qlo[2]));
            const double lo = wlo[ci] - dw - qc, hi = whi[ci] + dw - qc;
            /* first cluster of the column whose high end is not below lo */
            int64_t b0 = ccl[cc], b1 = ccl[cc + 1];
            while (b0 < b1) {
              const int64_t mid = b0 + (b1 - b0) / 2;
              if (whi[mid] < lo) b0 = mid + 1; else b1 = mid;
            }
            for (int64_t cjx = b0; cjx < ccl[cc + 1] && wlo[cjx] <= hi; ++cjx) {
              if (cjx != ci) {
                if ((cjx > ci) != (((ci + cjx) & 1) == 0)) continue;
              } else if (!zero && !positive) {
                continue;
              }
              double d2 = 0;
              for (int k = 0; k < 3; ++k) {
                const double g1 = blo[cjx].v[k] + S.v[k] - bhi[ci].v[k], g2 = blo[ci].v[k] - (bhi[cjx].v[k] + S.v[k]);
                const double g = g1 > 0 ? g1 : g2 > 0 ? g2 : 0.0;
                d2 += g * g;
              }
              if (d2 >= rl2) continue;
              const double* oi = &L->org[3 * ci];
              const double* oj = &L->org[3 * cjx];
              const float sxf = (float)(S.v[0] + oj[0] - oi[0]), syf = (float)(S.v[1] + oj[1] - oi[1]),
                          szf = (float)(S.v[2] + oj[2] - oi[2]);
              float xj[4], yj[4], zj[4];
              for (int l = 0; l < 4; ++l) xj[l] = L->bx[4 * cjx + l] + sxf, yj[l] = L->by[4 * cjx + l] + syf, zj[l] = L->bz[4 * cjx + l] + szf;
              unsigned mask = 0;
              for (int a = 0; a < 4; ++a) {
                if (!((real[ci] >> a) & 1u)) continue;
                const int64_t sa = 4 * ci + a;
                unsigned row = 0;
                for (int l = 0; l < 4; ++l) {
                  const float dx = xj[l] - L->bx[sa], dy = yj[l] - L->by[sa], dz = zj[l] - L->bz[sa];
                  const float zz = dz * dz;
                  const float yy = dy * dy + zz;
                  const float d2f = dx * dx + yy;
                  if (d2f < rl2f) row |= 1u << l;
                }
                row &= real[cjx];
                if (cjx == ci && zero) row &= (0xFu << (a + 1)) & 0xFu;
                if (!row) continue;
                mask |= row << (4 * a);
              }
              unsigned exm = 0, e14 = 0;
              if (mask) {
                int64_t lo2 = xfirst[ci], hi2 = xfirst[ci + 1];
                while (lo2 < hi2) {
                  const int64_t mid = (lo2 + hi2) / 2;
                  if (xcj[mid] < (uint32_t)cjx) lo2 = mid + 1; else hi2 = mid;
                }
                if (lo2 < xfirst[ci + 1] && xcj[lo2] == (uint32_t)cjx && (xbits[lo2] & mask)) {
                  const unsigned exb = xbits[lo2] & mask, k14 = x14[lo2];
                  for (int b = 0; b < 16; ++b) {
                    if (!((exb >> b) & 1u)) continue;
                    const int64_t sa = 4 * ci + b / 4, sb = 4 * cjx + b % 4;
                    if (L->perm[sa] == L->perm[sb]) continue;
                    const v3 d = sub3(add3(cw[sb], S), cw[sa]);
                    v3 m = sub3(frs[sb], frs[sa]);
                    for (int k = 0; k < 3; ++k) m.v[k] -= round(m.v[k]);
                    const v3 dm = add3(add3(scl3(cell.a, m.v[0]), scl3(cell.b, m.v[1])), scl3(cell.c, m.v[2]));
                    if (dot3(d, d) <= dot3(dm, dm) + 1e-6) {
                      mask &= ~(1u << b);
                      if (keep_excluded) exm |= 1u << b, e14 |= k14 & (1u << b);
                    }
                  }
                }
              }
              if (mask | exm) {
                if (ne == cap) cap *= 2, E = (entry*)realloc(E, sizeof(entry) * (size_t)cap);
                E[ne].cj = (uint32_t)cjx, E[ne].code = code, E[ne].mask = (uint16_t)mask, E[ne].ex = (uint16_t)exm,
                E[ne].ex14 = (uint16_t)e14;
                ++ne;
              }
            }
          }
        }
      }
      L->first[ci + 1] = ne;
    }
    L->nentries = ne;
    const size_t m = (size_t)(ne > 0 ? ne : 1);
    L->cj = (uint32_t*)malloc(4 * m);
    L->code = (uint16_t*)malloc(2 * m);
    L->mask = (uint16_t*)malloc(2 * m);
    L->ex = (uint16_t*)malloc(2 * m);
    L->ex14 = (uint16_t*)malloc(2 * m);
    for (int64_t e = 0; e < ne; ++e)
      L->cj[e] = E[e].cj, L->code[e] = E[e].code, L->mask[e] = E[e].mask, L->ex[e] = E[e].ex, L->ex14[e] = E[e].ex14;
    free(E);
  }
  /* offsets to the binned positions (nbnxm.cpp:267-277) and image vectors (:278-282) */
  L->off = (double*)calloc((size_t)(3 * ns > 0 ? 3 * ns : 1), sizeof(double));
  L->ioff = (int32_t*)calloc((size_t)(3 * ns > 0 ? 3 * ns : 1), sizeof(int32_t));
  for (int64_t s = 0; s < ns; ++s) {
    if (L->perm[s] == ORC_PAD) continue;
    for (int k = 0; k < 3; ++k) L->off[3 * s + k] = cw[s].v[k] - x[3 * (size_t)L->perm[s] + k];
    const v3 o = to_fractional(&cell, add3(cell.o, mk(L->off[3 * s], L->off[3 * s + 1], L->off[3 * s + 2])));
    for (int k = 0; k < 3; ++k) L->ioff[3 * s + k] = (int32_t)lround(o.v[k]);
  }
  L->qvec = (double*)malloc(sizeof(double) * (size_t)(3 * L->ncodes));
  for (int64_t c = 0; c < 3 * L->ncodes; c += 3) {
    const v3 sv = add3(add3(scl3(cell.a, L->qint[c]), scl3(cell.b, L->qint[c + 1])), scl3(cell.c, L->qint[c + 2]));
    for (int k = 0; k < 3; ++k) L->qvec[c + k] = sv.v[k];
  }
done:
  free(fr); free(col); free(cstart); free(order); free(ccl); free(ccol); free(cw); free(frs); free(blo); free(bhi);
  free(wlo); free(whi); free(real); free(xfirst); free(xcj); free(xbits); free(x14);
  return L;
}

void orc_list_sizes(const orc_list* L, int64_t out[7]) {
  out[0] = L->n, out[1] = L->nclusters, out[2] = L->ns, out[3] = L->nentries, out[4] = L->ncodes, out[5] = L->nu,
  out[6] = L->nv;
}

#define CP(dst, src, cnt) if (dst) memcpy(dst, src, sizeof(*(src)) * (size_t)(cnt))
void orc_list_get(const orc_list* L, uint32_t* perm, uint32_t* iperm, int64_t* first, uint32_t* cj, uint16_t* code,
                  uint16_t* mask, uint16_t* ex, uint16_t* ex14, double* org, double* off, int32_t* ioff, int32_t* qint,
                  double* qvec, float* bx, float* by, float* bz) {
  CP(perm, L->perm, L->ns); CP(iperm, L->iperm, L->n); CP(first, L->first, L->nclusters + 1);
  CP(cj, L->cj, L->nentries); CP(code, L->code, L->nentries); CP(mask, L->mask, L->nentries);
  CP(ex, L->ex, L->nentries); CP(ex14, L->ex14, L->nentries); CP(org, L->org, 3 * L->nclusters);
  CP(off, L->off, 3 * L->ns); CP(ioff, L->ioff, 3 * L->ns); CP(qint, L->qint, 3 * L->ncodes);
  CP(qvec, L->qvec, 3 * L->ncodes); CP(bx, L->bx, L->ns); CP(by, L->by, L->ns); CP(bz, L->bz, L->ns);
}

void orc_list_positions(const orc_list* L, const double* x, float* X, float* Y, float* Z) {
  for (int64_t s = 0; s < L->ns; ++s) {
    const int64_t r = L->perm[s] != ORC_PAD ? s : (s & ~(int64_t)3);
    const size_t a = 3 * (size_t)L->perm[r];
    const double* o = &L->org[3 * (s / 4)];
    X[s] = (float)(x[a] + L->off[3 * r] - o[0]);
    Y[s] = (float)(x[a + 1] + L->off[3 * r + 1] - o[1]);
    Z[s] = (float)(x[a + 2] + L->off[3 * r + 2] - o[2]);
  }
}

void orc_list_prune(const orc_list* L, const float* X, const float* Y, const float* Z, double radius, uint16_t* pmask,
                    uint16_t* pex, uint16_t* pex14) {
  const float r2 = (float)(radius * radius);
  for (int64_t ci = 0; ci < L->nclusters; ++ci) {
    const int64_t si = 4 * ci;
    for (int64_t e = L->first[ci]; e < L->first[ci + 1]; ++e) {
      const int64_t sj = 4 * (int64_t)L->cj[e];
      const double* sq = &L->qvec[3 * (size_t)L->code[e]];
      const double* oi = &L->org[3 * ci];
      const double* oj = &L->org[3 * (size_t)L->cj[e]];
      const float sx = (float)(sq[0] + oj[0] - oi[0]), sy = (float)(sq[1] + oj[1] - oi[1]), sz = (float)(sq[2] + oj[2] - oi[2]);
      unsigned in = 0;
      const unsigned both = (unsigned)L->mask[e] | (unsigned)L->ex[e];
      for (int a = 0; a < 4; ++a) {
        const unsigned rowbits = (both >> (4 * a)) & 15u;
        if (!rowbits) continue;
        unsigned row = 0;
        for (int l = 0; l < 4; ++l) {
          const float xj = X[sj + l] + sx, yj = Y[sj + l] + sy, zj = Z[sj + l] + sz;
          const float dx = xj - X[si + a], dy = yj - Y[si + a], dz = zj - Z[si + a];
          const float zz = dz * dz;
          const float yy = dy * dy + zz;
          const float d2 = dx * dx + yy;
          if (d2 < r2) row |= 1u << l;
        }
        in |= (row & rowbits) << (4 * a);
      }
      pmask[e] = (uint16_t)(L->mask[e] & in);
      pex[e] = (uint16_t)(L->ex[e] & in);
      pex14[e] = (uint16_t)(L->ex14[e] & in);
    }
  }
}
