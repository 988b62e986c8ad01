/* ORACLE — test infrastructure only (never linked into rascil_amd).
 *
 * CPU reference of the w-stacking gridder's visibility ordering: per-visibility
 * (w plane, uv tile) bucket keys, in-tile origin digits ((first cell mod tile) along u,
 * then v) and a stable sort by (bucket key, origin digit). This is the checker the GPU
 * bucket sort must match bit for bit. The coordinate contract (one IEEE fma, one
 * subtraction and one floor per axis, FP64 or FP32) is the one documented in
 * include/hipng.h for hipng_bucket_sort; the hermitian flip w < 0 -> (-u, -v, -w)
 * restates the ducc0 wgridder's published treatment of negative w (ducc0 0.24, not
 * vendored), applied to the uvw RASCIL passes at rascil/processing_components/imaging/
 * ng.py:201-204 (u and w negated by the caller).
 *
 * Build: oracle/build.py (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t gx, gy, W, tile, ntx, nty, nplanes, wstack, f32coords;
  double w0, dw;
} oracle_plan;

static void item_pos(double u, double v, double w, double sxc, double syc, double swc,
                     const oracle_plan* P, long long* ax, long long* ay, int* s) {
  const double hw = 0.5 * P->W, hgx = 0.5 * P->gx, hgy = 0.5 * P->gy;
  const double mw0 = P->wstack ? -(P->w0 / P->dw) : 0.0;
  if (w < 0.0) {
    u = -u;
    v = -v;
    w = -w;
  }
  long long a, b;
  int sp = 0;
  if (!P->f32coords) {
    a = (long long)(floor(fma(u, sxc, hgx) - hw) + 1.0);
    b = (long long)(floor(fma(v, syc, hgy) - hw) + 1.0);
    if (P->wstack) sp = (int)(floor(fma(w, swc, mw0) - hw) + 1.0);
  } else {
    const float fhw = (float)hw;
    a = (long long)(floorf(fmaf((float)u, (float)sxc, (float)hgx) - fhw) + 1.0f);
    b = (long long)(floorf(fmaf((float)v, (float)syc, (float)hgy) - fhw) + 1.0f);
    if (P->wstack) sp = (int)(floorf(fmaf((float)w, (float)swc, (float)mw0) - fhw) + 1.0f);
  }
  a %= P->gx;
  if (a < 0) a += P->gx;
  b %= P->gy;
  if (b < 0) b += P->gy;
  const int ww = P->wstack ? P->W : 1;
  if (sp < 0) sp = 0;
  if (sp > P->nplanes - ww) sp = P->nplanes - ww;
  *ax = a;
  *ay = b;
  *s = sp;
}

/* keys[i] for i = row * nchan + chan; digits[i] (if not NULL) = (ax mod tile) * tile + ay mod tile */
void oracle_keys(const double* uvw, int64_t nrow, int32_t nchan, const double* sx,
                 const double* sy, const double* sw, const oracle_plan* P, uint32_t* keys,
                 uint32_t* digits) {
  const uint32_t ntiles = (uint32_t)P->ntx * (uint32_t)P->nty;
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < nrow; ++r) {
    for (int32_t c = 0; c < nchan; ++c) {
      long long ax, ay;
      int s;
      item_pos(uvw[3 * r], uvw[3 * r + 1], uvw[3 * r + 2], sx[c], sy[c], sw[c], P, &ax, &ay, &s);
      keys[r * nchan + c] = (uint32_t)s * ntiles + (uint32_t)(ax / P->tile) * (uint32_t)P->nty +
                            (uint32_t)(ay / P->tile);
      if (digits) digits[r * nchan + c] = (uint32_t)((ax % P->tile) * P->tile + ay % P->tile);
    }
  }
}

/* Stable counting sort of keys[0..n): sorted keys and the permutation (original indices). */
int oracle_counting_sort(const uint32_t* keys, int64_t n, uint64_t nkeys, uint32_t* keys_sorted,
                         uint32_t* perm) {
  uint64_t* start = (uint64_t*)calloc(nkeys + 1, sizeof(uint64_t));
  if (!start) return -1;
  for (int64_t i = 0; i < n; ++i) start[keys[i] + 1]++;
  for (uint64_t k = 0; k < nkeys; ++k) start[k + 1] += start[k];
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t pos = start[keys[i]]++;
    keys_sorted[pos] = keys[i];
    perm[pos] = (uint32_t)i;
  }
  free(start);
  return 0;
}

/* Stable sort by (keys, digits): a counting sort by digit (ndigits bins), then a stable
 * counting sort of that order by key. perm holds original indices, keys_sorted the keys. */
int oracle_bucket_sort(const uint32_t* keys, const uint32_t* digits, int64_t n, uint64_t nkeys,
                       uint64_t ndigits, uint32_t* keys_sorted, uint32_t* perm) {
  uint32_t* dk = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint32_t));
  uint32_t* p1 = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint32_t));
  uint32_t* p2 = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint32_t));
  if (!dk || !p1 || !p2) {
    free(dk);
    free(p1);
    free(p2);
    return -1;
  }
  int rc = oracle_counting_sort(digits, n, ndigits, dk, p1);
  if (rc == 0) {
    for (int64_t i = 0; i < n; ++i) dk[i] = keys[p1[i]];
    rc = oracle_counting_sort(dk, n, nkeys, keys_sorted, p2);
  }
  if (rc == 0)
    for (int64_t i = 0; i < n; ++i) perm[i] = p1[p2[i]];
  free(dk);
  free(p1);
  free(p2);
  return rc;
}

/* ---------------------------------------------------------------------------------
 * CPU port of the whole w-stacking gridder (FP64, OpenMP): sorted records, per-plane
 * scatter gridding and gather degridding. The per-plane FFTs and image-domain steps
 * are done by the Python driver (oracle/wstack_cpu.py) with scipy.fft.
 * --------------------------------------------------------------------------------- */

static inline double es_kernel(double z, double beta) {
  const double t = 1.0 - z * z;
  return t > 0.0 ? exp(beta * (sqrt(t) - 1.0)) : 0.0;
}

static void item_full(double u, double v, double w, double sxc, double syc, double swc,
                      const oracle_plan* P, int32_t* ax, int32_t* ay, int* s, double* f3, int* flip) {
  const double hw = 0.5 * P->W, hgx = 0.5 * P->gx, hgy = 0.5 * P->gy;
  const double mw0 = P->wstack ? -(P->w0 / P->dw) : 0.0;
  *flip = w < 0.0;
  if (*flip) {
    u = -u;
    v = -v;
    w = -w;
  }
  const double ux = fma(u, sxc, hgx), vy = fma(v, syc, hgy);
  const double a0 = floor(ux - hw) + 1.0, b0 = floor(vy - hw) + 1.0;
  f3[0] = a0 - ux;
  f3[1] = b0 - vy;
  long long a = (long long)a0 % P->gx, b = (long long)b0 % P->gy;
  if (a < 0) a += P->gx;
  if (b < 0) b += P->gy;
  int sp = 0;
  f3[2] = 0.0;
  if (P->wstack) {
    const double wh = fma(w, swc, mw0);
    const double s0 = floor(wh - hw) + 1.0;
    f3[2] = s0 - wh;
    sp = (int)s0;
  }
  const int ww = P->wstack ? P->W : 1;
  if (sp < 0) sp = 0;
  if (sp > P->nplanes - ww) sp = P->nplanes - ww;
  *ax = (int32_t)a;
  *ay = (int32_t)b;
  *s = sp;
}

/* Sorted records. vals (nullable): weighted values with flip-conjugation and the
 * exp(-2 pi i |w| f/c nshift) phase, [n][nb] complex as (re, im) pairs. */
void oracle_records(const double* uvw, int32_t nchan, const uint32_t* perm, int64_t n,
                    const double* sx, const double* sy, const double* sw, const double* pn,
                    const oracle_plan* P, const double* ms, const double* wgt, int32_t nb,
                    int32_t* ax, int32_t* ay, double* f3, double* vals) {
  const double twopi = 6.283185307179586476925286766559;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; ++j) {
    const int64_t i = perm[j];
    const int64_t r = i / nchan;
    const int c = (int)(i - r * nchan);
    int s, flip;
    item_full(uvw[3 * r], uvw[3 * r + 1], uvw[3 * r + 2], sx[c], sy[c], sw[c], P, &ax[j], &ay[j],
              &s, &f3[3 * j], &flip);
    if (vals) {
      double cs = 1.0, sn = 0.0;
      if (P->wstack) {
        double t = -fabs(uvw[3 * r + 2]) * pn[c];
        t -= nearbyint(t);
        cs = cos(twopi * t);
        sn = sin(twopi * t);
      }
      for (int b = 0; b < nb; ++b) {
        double re = ms[2 * (i * nb + b)], im = ms[2 * (i * nb + b) + 1];
        if (flip) im = -im;
        if (wgt) {
          re *= wgt[i * nb + b];
          im *= wgt[i * nb + b];
        }
        vals[2 * (j * nb + b)] = re * cs - im * sn;
        vals[2 * (j * nb + b) + 1] = re * sn + im * cs;
      }
    }
  }
}

/* plane [nb][gx][gy] complex (re, im) += contributions of items with start plane in
 * [p - Ww + 1, p] of the tiles listed in `tiles` (those with items in that plane range).
 * Tiles are processed in 3x3 colour classes so that concurrently processed tiles never
 * write the same cell (a footprint reaches one tile further). */
void oracle_grid_plane(int32_t p, const oracle_plan* P, double beta, const uint32_t* bstart,
                       const uint32_t* bend, const int32_t* ax, const int32_t* ay, const double* f3,
                       const double* vals, int32_t nb, const uint32_t* tiles, int64_t ntl,
                       double* plane) {
  const int W = P->W, ww = P->wstack ? P->W : 1;
  const int ntx = P->ntx, nty = P->nty;
  const uint32_t ntiles = (uint32_t)ntx * (uint32_t)nty;
  const int slo = p - (ww - 1) < 0 ? 0 : p - (ww - 1);
  const int shi = p < P->nplanes - ww ? p : P->nplanes - ww;
  const double sc = 2.0 / W;
  for (int cx = 0; cx < 3; ++cx)
    for (int cy = 0; cy < 3; ++cy) {
#pragma omp parallel for schedule(dynamic, 4)
      for (int64_t it = 0; it < ntl; ++it) {
        const uint32_t t = tiles[it];
        const int tx = (int)(t / nty), ty = (int)(t % nty);
        const int ccx = (ntx % 2 == 1 && tx == ntx - 1) ? 2 : tx % 2;
        const int ccy = (nty % 2 == 1 && ty == nty - 1) ? 2 : ty % 2;
        if (ccx != cx || ccy != cy) continue;
        double kx[16], ky[16];
        for (int s = slo; s <= shi; ++s) {
          const uint32_t key = (uint32_t)s * ntiles + t;
          const double iw = (double)(p - s);
          for (uint32_t j = bstart[key]; j < bend[key]; ++j) {
            const double* f = f3 + 3 * (int64_t)j;
            const double kw = P->wstack ? es_kernel((f[2] + iw) * sc, beta) : 1.0;
            for (int i = 0; i < W; ++i) {
              kx[i] = es_kernel((f[0] + i) * sc, beta) * kw;
              ky[i] = es_kernel((f[1] + i) * sc, beta);
            }
            for (int b = 0; b < nb; ++b) {
              const double vr = vals[2 * ((int64_t)j * nb + b)], vi = vals[2 * ((int64_t)j * nb + b) + 1];
              for (int ix = 0; ix < W; ++ix) {
                int gxi = ax[j] + ix;
                if (gxi >= P->gx) gxi -= P->gx;
                double* row = plane + 2 * (((int64_t)b * P->gx + gxi) * P->gy);
                const double wr = kx[ix] * vr, wi = kx[ix] * vi;
                for (int iy = 0; iy < W; ++iy) {
                  int gyi = ay[j] + iy;
                  if (gyi >= P->gy) gyi -= P->gy;
                  row[2 * gyi] += wr * ky[iy];
                  row[2 * gyi + 1] += wi * ky[iy];
                }
              }
            }
          }
        }
      }
    }
}

/* acc[j][b] += kw(p) * sum_cells k(x) k(y) plane[b][cell] for items j in [j0, j1) whose
 * start plane s satisfies p - Ww < s <= p (s given per item in splane). */
void oracle_degrid_plane(int32_t p, const oracle_plan* P, double beta, int64_t j0, int64_t j1,
                         const int32_t* splane, const int32_t* ax, const int32_t* ay,
                         const double* f3, const double* plane, int32_t nb, double* acc) {
  const int W = P->W;
  const double sc = 2.0 / W;
#pragma omp parallel for schedule(static)
  for (int64_t j = j0; j < j1; ++j) {
    const int s = splane[j];
    const double* f = f3 + 3 * j;
    const double kw = P->wstack ? es_kernel((f[2] + (double)(p - s)) * sc, beta) : 1.0;
    double kx[16], ky[16];
    for (int i = 0; i < W; ++i) {
      kx[i] = es_kernel((f[0] + i) * sc, beta);
      ky[i] = es_kernel((f[1] + i) * sc, beta);
    }
    for (int b = 0; b < nb; ++b) {
      double sr = 0.0, si = 0.0;
      for (int ix = 0; ix < W; ++ix) {
        int gxi = ax[j] + ix;
        if (gxi >= P->gx) gxi -= P->gx;
        const double* row = plane + 2 * (((int64_t)b * P->gx + gxi) * P->gy);
        double rr = 0.0, ri = 0.0;
        for (int iy = 0; iy < W; ++iy) {
          int gyi = ay[j] + iy;
          if (gyi >= P->gy) gyi -= P->gy;
          rr += ky[iy] * row[2 * gyi];
          ri += ky[iy] * row[2 * gyi + 1];
        }
        sr += kx[ix] * rr;
        si += kx[ix] * ri;
      }
      acc[2 * (j * nb + b)] += kw * sr;
      acc[2 * (j * nb + b) + 1] += kw * si;
    }
  }
}

/* acc[b][ix][iy] += Re(exp(-2 pi i wp ((n-1) - nshift)) (-1)^(tx+ty) plane[b][tx mod gx][ty mod gy]) */
void oracle_screen_acc(const oracle_plan* P, int32_t nx, int32_t ny, double px, double py,
                       double nshift, double wp, const double* plane, int32_t nb, double* acc) {
  const double twopi = 6.283185307179586476925286766559;
#pragma omp parallel for schedule(static)
  for (int ix = 0; ix < nx; ++ix) {
    const int tx = ix - nx / 2;
    const int gxi = tx < 0 ? tx + P->gx : tx;
    const double x = tx * px;
    for (int iy = 0; iy < ny; ++iy) {
      const int ty = iy - ny / 2;
      const int gyi = ty < 0 ? ty + P->gy : ty;
      double cs = 1.0, sn = 0.0;
      if (P->wstack) {
        const double y = ty * py, r2 = x * x + y * y;
        double t = -wp * (-r2 / (sqrt(1.0 - r2) + 1.0) - nshift);
        t -= nearbyint(t);
        cs = cos(twopi * t);
        sn = sin(twopi * t);
      }
      const double sg = ((tx + ty) & 1) ? -1.0 : 1.0;
      for (int b = 0; b < nb; ++b) {
        const double* F = plane + 2 * (((int64_t)b * P->gx + gxi) * P->gy + gyi);
        acc[((int64_t)b * nx + ix) * ny + iy] += sg * (F[0] * cs - F[1] * sn);
      }
    }
  }
}

/* plane[b][g] = (-1)^(tx+ty) dc[b][t] exp(+2 pi i wp ((n-1) - nshift)) on image pixels, 0 elsewhere */
void oracle_screen_fill(const oracle_plan* P, int32_t nx, int32_t ny, double px, double py,
                        double nshift, double wp, const double* dc, int32_t nb, double* plane) {
  const double twopi = 6.283185307179586476925286766559;
  const int hx = nx / 2, hy = ny / 2;
#pragma omp parallel for schedule(static)
  for (int gxi = 0; gxi < P->gx; ++gxi) {
    int tx = 0, inx = 1;
    if (gxi < hx) tx = gxi;
    else if (gxi >= P->gx - hx) tx = gxi - P->gx;
    else inx = 0;
    for (int gyi = 0; gyi < P->gy; ++gyi) {
      int ty = 0, in = inx;
      if (gyi < hy) ty = gyi;
      else if (gyi >= P->gy - hy) ty = gyi - P->gy;
      else in = 0;
      for (int b = 0; b < nb; ++b) {
        double* o = plane + 2 * (((int64_t)b * P->gx + gxi) * P->gy + gyi);
        if (!in) {
          o[0] = 0.0;
          o[1] = 0.0;
          continue;
        }
        double cs = 1.0, sn = 0.0;
        if (P->wstack) {
          const double x = tx * px, y = ty * py, r2 = x * x + y * y;
          double t = wp * (-r2 / (sqrt(1.0 - r2) + 1.0) - nshift);
          t -= nearbyint(t);
          cs = cos(twopi * t);
          sn = sin(twopi * t);
        }
        const double d = (((tx + ty) & 1) ? -1.0 : 1.0) * dc[((int64_t)b * nx + tx + hx) * ny + ty + hy];
        o[0] = d * cs;
        o[1] = d * sn;
      }
    }
  }
}
