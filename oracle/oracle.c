/* FP64 CPU oracle of the geometry hot path: test infrastructure only (see oracle.h for what it
 * restates and how each part is pinned). Built with -ffp-contract=off so that every sum runs in
 * the reference's order with separately rounded products. */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { PT = 0, IV = 1, TR = 2, QU = 3, TE = 4, HE = 5, PR = 6, PY = 7 };

/* ---- Reference cells (cell.rs:53-193) -------------------------------------------------------- */

static const double VX[8][8][3] = {
    {{0, 0, 0}},
    {{0, 0, 0}, {1, 0, 0}},
    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}},
    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}},
    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}},
    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}},
};
static const int NV[8] = {1, 2, 3, 4, 4, 8, 6, 5};
static const int TDIM[8] = {0, 1, 2, 2, 3, 3, 3, 3};

/* Sub-entity tables: per cell and dimension 1..2, a list of (type, nverts, v0..v3). The cell
 * itself (dimension tdim) is one entity of its own type over all vertices. */
typedef struct {
  int type, n, v[4];
} Ent;

static const Ent E_TR1[] = {{IV, 2, {0, 1}}, {IV, 2, {0, 2}}, {IV, 2, {1, 2}}};
static const Ent E_QU1[] = {{IV, 2, {0, 1}}, {IV, 2, {0, 2}}, {IV, 2, {1, 3}}, {IV, 2, {2, 3}}};
static const Ent E_TE1[] = {{IV, 2, {0, 1}}, {IV, 2, {0, 2}}, {IV, 2, {0, 3}},
                            {IV, 2, {1, 2}}, {IV, 2, {1, 3}}, {IV, 2, {2, 3}}};
static const Ent E_TE2[] = {
    {TR, 3, {0, 1, 2}}, {TR, 3, {0, 1, 3}}, {TR, 3, {0, 2, 3}}, {TR, 3, {1, 2, 3}}};
static const Ent E_HE1[] = {{IV, 2, {0, 1}}, {IV, 2, {0, 2}}, {IV, 2, {0, 4}}, {IV, 2, {1, 3}},
                            {IV, 2, {1, 5}}, {IV, 2, {2, 3}}, {IV, 2, {2, 6}}, {IV, 2, {3, 7}},
                            {IV, 2, {4, 5}}, {IV, 2, {4, 6}}, {IV, 2, {5, 7}}, {IV, 2, {6, 7}}};
static const Ent E_HE2[] = {{QU, 4, {0, 1, 2, 3}}, {QU, 4, {0, 1, 4, 5}}, {QU, 4, {0, 2, 4, 6}},
                            {QU, 4, {1, 3, 5, 7}}, {QU, 4, {2, 3, 6, 7}}, {QU, 4, {4, 5, 6, 7}}};
static const Ent E_PR1[] = {{IV, 2, {0, 1}}, {IV, 2, {0, 2}}, {IV, 2, {0, 3}},
                            {IV, 2, {1, 2}}, {IV, 2, {1, 4}}, {IV, 2, {2, 5}},
                            {IV, 2, {3, 4}}, {IV, 2, {3, 5}}, {IV, 2, {4, 5}}};
static const Ent E_PR2[] = {{TR, 3, {0, 1, 2}},
                            {QU, 4, {0, 1, 3, 4}},
                            {QU, 4, {0, 2, 3, 5}},
                            {QU, 4, {1, 2, 4, 5}},
                            {TR, 3, {3, 4, 5}}};
static const Ent E_PY1[] = {{IV, 2, {0, 1}}, {IV, 2, {0, 2}}, {IV, 2, {0, 4}}, {IV, 2, {1, 3}},
                            {IV, 2, {1, 4}}, {IV, 2, {2, 3}}, {IV, 2, {2, 4}}, {IV, 2, {3, 4}}};
static const Ent E_PY2[] = {
    {QU, 4, {0, 1, 2, 3}}, {TR, 3, {0, 1, 4}}, {TR, 3, {0, 2, 4}}, {TR, 3, {1, 3, 4}}, {TR, 3, {2, 3, 4}}};

static const Ent* ents(int cell, int dim, int* n) {
  *n = 0;
  switch (cell * 4 + dim) {
    case TR * 4 + 1: *n = 3; return E_TR1;
    case QU * 4 + 1: *n = 4; return E_QU1;
    case TE * 4 + 1: *n = 6; return E_TE1;
    case TE * 4 + 2: *n = 4; return E_TE2;
    case HE * 4 + 1: *n = 12; return E_HE1;
    case HE * 4 + 2: *n = 6; return E_HE2;
    case PR * 4 + 1: *n = 9; return E_PR1;
    case PR * 4 + 2: *n = 5; return E_PR2;
    case PY * 4 + 1: *n = 8; return E_PY1;
    case PY * 4 + 2: *n = 5; return E_PY2;
  }
  return NULL;
}

int or_tdim(int cell) { return (cell >= 0 && cell < 8) ? TDIM[cell] : -1; }

int or_sub_entity(int cell, int dim, int index, int* type, int* verts) {
  if (cell < 0 || cell > 7 || dim < 0 || dim > TDIM[cell] || index < 0) return -1;
  if (dim == 0) {
    if (index >= NV[cell]) return -1;
    *type = PT;
    verts[0] = index;
    return 1;
  }
  if (dim == TDIM[cell]) {
    if (index > 0) return -1;
    *type = cell;
    for (int v = 0; v < NV[cell]; ++v) verts[v] = v;
    return NV[cell];
  }
  int ne;
  const Ent* es = ents(cell, dim, &ne);
  if (index >= ne) return -1;
  *type = es[index].type;
  for (int v = 0; v < es[index].n; ++v) verts[v] = es[index].v[v];
  return es[index].n;
}

size_t or_num_nodes(int cell, int degree) {
  size_t p = (size_t)degree;
  switch (cell) {
    case PT: return 1;
    case IV: return p + 1;
    case TR: return (p + 1) * (p + 2) / 2;
    case QU: return (p + 1) * (p + 1);
    case TE: return (p + 1) * (p + 2) * (p + 3) / 6;
    case HE: return (p + 1) * (p + 1) * (p + 1);
    case PR: return (p + 1) * (p + 1) * (p + 2) / 2;
    case PY: return (p + 1) * (p + 2) * (2 * p + 3) / 6;
  }
  return 0;
}

/* ---- GLL points: roots of P'_{m-1} by Newton with the Legendre ODE ------------------------- */

static void legendre(int n, double x, double* p, double* dp) {
  double p0 = 1.0, p1 = x;
  if (n == 0) {
    *p = 1.0;
    *dp = 0.0;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *p = p1;
  /* P'_n = n (x P_n - P_{n-1}) / (x^2 - 1) */
  *dp = n * (x * p1 - p0) / (x * x - 1.0);
}

/* m GLL points on [0,1], ascending, exactly symmetric. */
static void gll01(int m, double* t) {
  int n = m - 1;
  t[0] = 0.0;
  t[n] = 1.0;
  for (int i = 1; i < n; ++i) {
    double x = -cos(M_PI * i / n);
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(n, x, &p, &dp);
      /* f = P'_n, f' = P''_n = (2 x P'_n - n(n+1) P_n) / (1 - x^2) */
      double d2 = (2.0 * x * dp - n * (n + 1.0) * p) / (1.0 - x * x);
      double dx = dp / d2;
      x -= dx;
      if (fabs(dx) < 1e-16) break;
    }
    t[i] = 0.5 * (1.0 + x);
  }
  for (int i = 1; i < m / 2; ++i) {
    double u = 1.0 - t[i];
    t[n - i] = u;
    t[i] = 1.0 - u;
  }
  if (m % 2 == 1) t[m / 2] = 0.5;
}

/* ---- A.5 node layout (lagrange.rs:163-266) ------------------------------------------------- */

static long push_pt(double* out, long n, int td, const double* v0, const double* const* ax,
                    const double* t, int na) {
  for (int j = 0; j < td; ++j) {
    double x = v0[j];
    for (int a = 0; a < na; ++a) x += t[a] * (ax[a][j] - v0[j]);
    out[n * td + j] = x;
  }
  return n + 1;
}

long or_lagrange_points(int cell, int degree, int family, double* out) {
  if (cell < 1 || cell > 7 || degree < 1) return -1;
  const int p = degree, td = TDIM[cell];
  double line[64];
  int nl = p - 1;
  if (p > 60) return -1;
  if (family == 0) {
    for (int i = 1; i < p; ++i) line[i - 1] = (double)i / p;
  } else {
    if (cell != IV && cell != QU && cell != HE) return -1;
    double g[64];
    gll01(p + 1, g);
    for (int i = 1; i < p; ++i) line[i - 1] = g[i];
  }
  long n = 0;
  for (int v = 0; v < NV[cell]; ++v) n = push_pt(out, n, td, VX[cell][v], NULL, NULL, 0);
  for (int dim = 1; dim <= td; ++dim) {
    int ne;
    const Ent* es = ents(cell, dim, &ne);
    Ent whole = {cell, NV[cell], {0, 1, 2, 3}};
    if (dim == td) {
      es = &whole;
      ne = 1;
    }
    for (int e = 0; e < ne; ++e) {
      const Ent* en = &es[e];
      const double* v0 = VX[cell][en->v[0]];
      const double* ax[3];
      double t[3];
      switch (en->type) {
        case IV:
          ax[0] = VX[cell][en->v[1]];
          for (int i = 0; i < nl; ++i) {
            t[0] = line[i];
            n = push_pt(out, n, td, v0, ax, t, 1);
          }
          break;
        case TR:
          ax[0] = VX[cell][en->v[1]];
          ax[1] = VX[cell][en->v[2]];
          for (int i1 = 1; i1 < p; ++i1)
            for (int i0 = 1; i0 + i1 < p; ++i0) {
              t[0] = (double)i0 / p;
              t[1] = (double)i1 / p;
              n = push_pt(out, n, td, v0, ax, t, 2);
            }
          break;
        case QU:
          ax[0] = VX[cell][en->v[1]];
          ax[1] = VX[cell][en->v[2]];
          for (int i1 = 0; i1 < nl; ++i1)
            for (int i0 = 0; i0 < nl; ++i0) {
              t[0] = line[i0];
              t[1] = line[i1];
              n = push_pt(out, n, td, v0, ax, t, 2);
            }
          break;
        case TE:
          ax[0] = VX[cell][1];
          ax[1] = VX[cell][2];
          ax[2] = VX[cell][3];
          for (int i2 = 1; i2 < p; ++i2)
            for (int i1 = 1; i1 + i2 < p; ++i1)
              for (int i0 = 1; i0 + i1 + i2 < p; ++i0) {
                t[0] = (double)i0 / p;
                t[1] = (double)i1 / p;
                t[2] = (double)i2 / p;
                n = push_pt(out, n, td, v0, ax, t, 3);
              }
          break;
        case HE:
          ax[0] = VX[cell][1];
          ax[1] = VX[cell][2];
          ax[2] = VX[cell][4];
          for (int i2 = 0; i2 < nl; ++i2)
            for (int i1 = 0; i1 < nl; ++i1)
              for (int i0 = 0; i0 < nl; ++i0) {
                t[0] = line[i0];
                t[1] = line[i1];
                t[2] = line[i2];
                n = push_pt(out, n, td, v0, ax, t, 3);
              }
          break;
        case PR:
          ax[0] = VX[cell][1];
          ax[1] = VX[cell][2];
          ax[2] = VX[cell][3];
          for (int i2 = 1; i2 < p; ++i2)
            for (int i1 = 1; i1 < p; ++i1)
              for (int i0 = 1; i0 + i1 < p; ++i0) {
                t[0] = (double)i0 / p;
                t[1] = (double)i1 / p;
                t[2] = (double)i2 / p;
                n = push_pt(out, n, td, v0, ax, t, 3);
              }
          break;
        case PY:
          for (int i2 = 1; i2 < p - 1; ++i2)
            for (int i1 = 1; i1 < p - i2; ++i1)
              for (int i0 = 1; i0 < p - i2; ++i0) {
                out[n * 3] = (double)i0 / p;
                out[n * 3 + 1] = (double)i1 / p;
                out[n * 3 + 2] = (double)i2 / p;
                ++n;
              }
          break;
      }
    }
  }
  return (size_t)n == or_num_nodes(cell, degree) ? n : -1;
}

/* ---- Closed-form bases ------------------------------------------------------------------- */

/* Silvester factor l_a(s) = prod_{m<a} (p s - m)/(m+1) and its derivative in s. */
static void silvester(int a, int p, double s, double* v, double* dv) {
  double val = 1.0, der = 0.0;
  for (int m = 0; m < a; ++m) {
    double f = (p * s - m) / (m + 1.0), df = p / (m + 1.0);
    der = der * f + val * df;
    val *= f;
  }
  *v = val;
  *dv = der;
}

/* 1D Lagrange polynomial i on nodes t[0..n-1] and its derivative. */
static void lag1d(const double* t, int n, int i, double x, double* v, double* dv) {
  double val = 1.0, der = 0.0;
  for (int m = 0; m < n; ++m) {
    if (m == i) continue;
    double f = (x - t[m]) / (t[i] - t[m]), df = 1.0 / (t[i] - t[m]);
    der = der * f + val * df;
    val *= f;
  }
  *v = val;
  *dv = der;
}

static int nearest(const double* t, int n, double x) {
  int best = 0;
  for (int i = 1; i < n; ++i)
    if (fabs(t[i] - x) < fabs(t[best] - x)) best = i;
  return best;
}

/* Long-double Gaussian elimination with partial pivoting: solves A X = B (n x n, n x m). */
static int solve_ld(long double* a, long double* b, int n, int m) {
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (fabsl(a[i * n + k]) > fabsl(a[piv * n + k])) piv = i;
    if (a[piv * n + k] == 0.0L) return -1;
    if (piv != k) {
      for (int j = 0; j < n; ++j) {
        long double t = a[k * n + j];
        a[k * n + j] = a[piv * n + j];
        a[piv * n + j] = t;
      }
      for (int j = 0; j < m; ++j) {
        long double t = b[k * m + j];
        b[k * m + j] = b[piv * m + j];
        b[piv * m + j] = t;
      }
    }
    for (int i = k + 1; i < n; ++i) {
      long double f = a[i * n + k] / a[k * n + k];
      for (int j = k; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
      for (int j = 0; j < m; ++j) b[i * m + j] -= f * b[k * m + j];
    }
  }
  for (int i = n - 1; i >= 0; --i)
    for (int j = 0; j < m; ++j) {
      long double s = b[i * m + j];
      for (int c = i + 1; c < n; ++c) s -= a[i * n + c] * b[c * m + j];
      b[i * m + j] = s / a[i * n + i];
    }
  return 0;
}

/* Pyramid rational monomials x^a y^b z^c / (1-z)^min(a,b), a+c <= p, b+c <= p: value and
 * gradient at x (long double). Returns the count. */
static int pyr_span(int p, const double* xp, long double* f, long double* fx, long double* fy,
                    long double* fz) {
  int n = 0;
  long double x = xp[0], y = xp[1], z = xp[2], s = 1.0L - z;
  for (int c = 0; c <= p; ++c)
    for (int b = 0; b + c <= p; ++b)
      for (int a = 0; a + c <= p; ++a) {
        int m = a < b ? a : b;
        long double xa = powl(x, a), yb = powl(y, b), zc = powl(z, c);
        long double dxa = a ? a * powl(x, a - 1) : 0.0L, dyb = b ? b * powl(y, b - 1) : 0.0L,
                    dzc = c ? c * powl(z, c - 1) : 0.0L;
        long double r = 1.0L, rz = 0.0L;
        if (m > 0) {
          if (s == 0.0L) {
            f[n] = fx[n] = fy[n] = fz[n] = 0.0L;
            ++n;
            continue;
          }
          r = powl(s, -m);
          rz = m * r / s;
        }
        f[n] = xa * yb * zc * r;
        fx[n] = dxa * yb * zc * r;
        fy[n] = xa * dyb * zc * r;
        fz[n] = xa * yb * (dzc * r + zc * rz);
        ++n;
      }
  return n;
}

int or_tabulate(int cell, int degree, int family, const double* pts, size_t nq, double* values,
                double* derivs) {
  const int p = degree, td = TDIM[cell];
  if (cell < 1 || cell > 7 || p < 1 || p > 40) return -1;
  const long k = (long)or_num_nodes(cell, p);
  double* nodes = (double*)malloc(sizeof(double) * k * 3);
  if (or_lagrange_points(cell, p, family, nodes) != k) {
    free(nodes);
    return -1;
  }
  double t1[64];
  if (family == 0)
    for (int i = 0; i <= p; ++i) t1[i] = (double)i / p;
  else
    gll01(p + 1, t1);
  int rc = 0;
  if (cell == PY && p > 1) {
    /* Vandermonde in long double: V[i][j] = f_j(node_i); phi(x) = V^{-T} f(x). */
    long double* V = (long double*)malloc(sizeof(long double) * k * k);
    long double* VT = (long double*)malloc(sizeof(long double) * k * k);
    long double* f = (long double*)malloc(sizeof(long double) * 4 * k);
    long double* rhs = (long double*)malloc(sizeof(long double) * k * 4);
    for (long i = 0; i < k; ++i) {
      pyr_span(p, &nodes[i * 3], f, f + k, f + 2 * k, f + 3 * k);
      for (long j = 0; j < k; ++j) V[i * k + j] = f[j];
    }
    for (size_t q = 0; q < nq && rc == 0; ++q) {
      for (long i = 0; i < k; ++i)
        for (long j = 0; j < k; ++j) VT[i * k + j] = V[j * k + i];
      pyr_span(p, &pts[q * 3], f, f + k, f + 2 * k, f + 3 * k);
      for (long j = 0; j < k; ++j)
        for (int d = 0; d < 4; ++d) rhs[j * 4 + d] = f[d * k + j];
      rc = solve_ld(VT, rhs, (int)k, 4);
      for (long i = 0; i < k; ++i) {
        if (values) values[q * k + i] = (double)rhs[i * 4];
        if (derivs)
          for (int d = 0; d < 3; ++d) derivs[(q * k + i) * 3 + d] = (double)rhs[i * 4 + 1 + d];
      }
    }
    free(V);
    free(VT);
    free(f);
    free(rhs);
    free(nodes);
    return rc;
  }
  for (size_t q = 0; q < nq; ++q) {
    const double* x = &pts[q * td];
    for (long i = 0; i < k; ++i) {
      const double* r = &nodes[i * td];
      double v = 1.0, g[3] = {0, 0, 0};
      if (cell == PY) {
        /* Degree-1 pyramid: phi0 = (1-x-z)(1-y-z)/s, phi1 = x(1-y-z)/s, phi2 = (1-x-z) y/s,
         * phi3 = x y / s, phi4 = z, with s = 1 - z. */
        double s = 1.0 - x[2], a0 = 1.0 - x[0] - x[2], b0 = 1.0 - x[1] - x[2];
        double A = (i == 0 || i == 2) ? a0 : x[0], B = (i == 0 || i == 1) ? b0 : x[1];
        double dAx = (i == 0 || i == 2) ? -1.0 : 1.0, dAz = (i == 0 || i == 2) ? -1.0 : 0.0;
        double dBy = (i == 0 || i == 1) ? -1.0 : 1.0, dBz = (i == 0 || i == 1) ? -1.0 : 0.0;
        if (s == 0.0) {
          /* The apex: values are the Kronecker limits; gradients the limits along the axis
           * x = y = s/2 (the rational functions have no limit over all directions). */
          static const double apex[5][3] = {{-0.5, -0.5, -0.75}, {0.5, -0.5, -0.25},
                                            {-0.5, 0.5, -0.25}, {0.5, 0.5, 0.25}, {0, 0, 1}};
          v = i == 4 ? 1.0 : 0.0;
          g[0] = apex[i][0];
          g[1] = apex[i][1];
          g[2] = apex[i][2];
        } else if (i == 4) {
          v = x[2];
          g[0] = 0.0;
          g[1] = 0.0;
          g[2] = 1.0;
        } else {
          v = A * B / s;
          g[0] = dAx * B / s;
          g[1] = A * dBy / s;
          g[2] = (dAz * B + A * dBz) / s + A * B / (s * s);
        }
      } else if (cell == QU || cell == HE || (cell == IV && family == 1)) {
        for (int d = 0; d < td; ++d) {
          int id = nearest(t1, p + 1, r[d]);
          double lv, ld;
          lag1d(t1, p + 1, id, x[d], &lv, &ld);
          for (int e = 0; e < td; ++e) g[e] = (e == d) ? g[e] * lv + 0.0 : g[e] * lv;
          g[d] = v * ld + g[d];
          /* product rule accumulated: g_e holds d/dx_e of the product so far */
          v *= lv;
        }
      } else if (cell == PR) {
        /* Triangle Silvester in (x, y) times 1D Lagrange in z. */
        double lam[3] = {1.0 - x[0] - x[1], x[0], x[1]};
        double lr[3] = {1.0 - r[0] - r[1], r[0], r[1]};
        double tv = 1.0, tg[2] = {0, 0};
        for (int jj = 0; jj < 3; ++jj) {
          int a = (int)lround(p * lr[jj]);
          double lv, ld;
          silvester(a, p, lam[jj], &lv, &ld);
          double dl[2] = {jj == 0 ? -1.0 : (jj == 1 ? 1.0 : 0.0), jj == 0 ? -1.0 : (jj == 2 ? 1.0 : 0.0)};
          for (int e = 0; e < 2; ++e) tg[e] = tg[e] * lv + tv * ld * dl[e];
          tv *= lv;
        }
        int iz = nearest(t1, p + 1, r[2]);
        double zv, zd;
        lag1d(t1, p + 1, iz, x[2], &zv, &zd);
        v = tv * zv;
        g[0] = tg[0] * zv;
        g[1] = tg[1] * zv;
        g[2] = tv * zd;
      } else {
        /* Simplices (equispaced interval, triangle, tetrahedron): Silvester's formula over the
         * barycentric coordinates lambda_0 = 1 - sum x, lambda_{d+1} = x_d. */
        double lam[4], lr[4];
        lam[0] = 1.0;
        lr[0] = 1.0;
        for (int d = 0; d < td; ++d) {
          lam[0] -= x[d];
          lam[d + 1] = x[d];
          lr[0] -= r[d];
          lr[d + 1] = r[d];
        }
        for (int jj = 0; jj <= td; ++jj) {
          int a = (int)lround(p * lr[jj]);
          double lv, ld;
          silvester(a, p, lam[jj], &lv, &ld);
          for (int e = 0; e < td; ++e) {
            double dl = jj == 0 ? -1.0 : (jj - 1 == e ? 1.0 : 0.0);
            g[e] = g[e] * lv + v * ld * dl;
          }
          v *= lv;
        }
      }
      if (values) values[q * k + i] = v;
      if (derivs)
        for (int d = 0; d < td; ++d) derivs[(q * k + i) * td + d] = g[d];
    }
  }
  free(nodes);
  return rc;
}

/* ---- Factors (reference.rs:117-354; layout.rs:138-369) ------------------------------------ */

static double det_of(const double* j, int td, int gd) {
  if (td == gd) {
    if (td == 1) return j[0];
    if (td == 2) return j[0] * j[3] - j[2] * j[1];
    return j[0] * (j[4] * j[8] - j[7] * j[5]) - j[3] * (j[1] * j[8] - j[7] * j[2]) +
           j[6] * (j[1] * j[5] - j[4] * j[2]);
  }
  double a = 0, b = 0, c = 0;
  for (int g = 0; g < gd; ++g) {
    a += j[g] * j[g];
    if (td == 2) {
      b += j[g] * j[gd + g];
      c += j[gd + g] * j[gd + g];
    }
  }
  return td == 1 ? sqrt(a) : sqrt(a * c - b * b);
}

/* K = J^-1 by the adjugate when tdim = gdim (layout.rs:219-246), else (J^T J)^-1 J^T via
 * G = J^T J and its inverse (reference.rs:197-256). */
static void pinv_of(const double* j, int td, int gd, double* k) {
  if (td == gd && td == 3) {
    double inv = 1.0 / det_of(j, 3, 3);
    k[0] = (j[4] * j[8] - j[7] * j[5]) * inv;
    k[3] = (j[6] * j[5] - j[3] * j[8]) * inv;
    k[6] = (j[3] * j[7] - j[6] * j[4]) * inv;
    k[1] = (j[7] * j[2] - j[1] * j[8]) * inv;
    k[4] = (j[0] * j[8] - j[6] * j[2]) * inv;
    k[7] = (j[6] * j[1] - j[0] * j[7]) * inv;
    k[2] = (j[1] * j[5] - j[4] * j[2]) * inv;
    k[5] = (j[3] * j[2] - j[0] * j[5]) * inv;
    k[8] = (j[0] * j[4] - j[3] * j[1]) * inv;
    return;
  }
  if (td == gd && td == 2) {
    double inv = 1.0 / (j[0] * j[3] - j[2] * j[1]);
    k[0] = j[3] * inv;
    k[2] = -j[2] * inv;
    k[1] = -j[1] * inv;
    k[3] = j[0] * inv;
    return;
  }
  double G[9] = {0}, Gi[9] = {0};
  for (int a = 0; a < td; ++a)
    for (int b = 0; b < td; ++b) {
      double s = 0;
      for (int g = 0; g < gd; ++g) s += j[a * gd + g] * j[b * gd + g];
      G[a * td + b] = s;
    }
  if (td == 1) {
    Gi[0] = 1.0 / G[0];
  } else if (td == 2) {
    double d = G[0] * G[3] - G[1] * G[2];
    Gi[0] = G[3] / d;
    Gi[1] = -G[1] / d;
    Gi[2] = -G[2] / d;
    Gi[3] = G[0] / d;
  } else {
    double d = det_of(G, 3, 3);
    Gi[0] = (G[4] * G[8] - G[5] * G[7]) / d;
    Gi[1] = (G[2] * G[7] - G[1] * G[8]) / d;
    Gi[2] = (G[1] * G[5] - G[2] * G[4]) / d;
    Gi[3] = (G[5] * G[6] - G[3] * G[8]) / d;
    Gi[4] = (G[0] * G[8] - G[2] * G[6]) / d;
    Gi[5] = (G[2] * G[3] - G[0] * G[5]) / d;
    Gi[6] = (G[3] * G[7] - G[4] * G[6]) / d;
    Gi[7] = (G[1] * G[6] - G[0] * G[7]) / d;
    Gi[8] = (G[0] * G[4] - G[1] * G[3]) / d;
  }
  /* K(t, g) = sum_s Gi(t, s) J(g, s), at g * td + t. */
  for (int g = 0; g < gd; ++g)
    for (int t = 0; t < td; ++t) {
      double s = 0;
      for (int u = 0; u < td; ++u) s += Gi[t * td + u] * j[u * gd + g];
      k[g * td + t] = s;
    }
}

int or_eval_block(int cell, int degree, int family, int gdim, const double* nodes,
                  size_t num_nodes, const uint32_t* dofmap, size_t num_cells,
                  const uint32_t* flags, const double* pts, size_t nq, const double* ref_normals,
                  double* x, double* j, double* detj, double* pinv, double* fnormals,
                  double* normals, int threads) {
  const int td = TDIM[cell], gd = gdim;
  const long k = (long)or_num_nodes(cell, degree);
  if (num_cells == 0 || nq == 0) return 0;
  double* phi = (double*)malloc(sizeof(double) * nq * k);
  double* dphi = (double*)malloc(sizeof(double) * nq * k * 3);
  if (or_tabulate(cell, degree, family, pts, nq, phi, dphi)) {
    free(phi);
    free(dphi);
    return -1;
  }
  const int JL = td * gd;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
#endif
  for (long c = 0; c < (long)num_cells; ++c) {
    double xc[729 * 3];
    for (long i = 0; i < k; ++i) {
      uint32_t nidx = dofmap[c * k + i];
      if (nidx >= num_nodes) nidx = (uint32_t)(num_nodes - 1);
      for (int g = 0; g < gd; ++g) xc[i * gd + g] = nodes[(size_t)nidx * gd + g];
    }
    for (size_t q = 0; q < nq; ++q) {
      const double* ph = &phi[q * k];
      const double* dp = &dphi[q * k * td];
      double J[9];
      for (int g = 0; g < gd; ++g) {
        double acc = xc[g] * ph[0];
        for (long i = 1; i < k; ++i) acc += xc[i * gd + g] * ph[i];
        if (x) x[((size_t)c * nq + q) * gd + g] = acc;
      }
      for (int d = 0; d < td; ++d)
        for (int g = 0; g < gd; ++g) {
          double acc = xc[g] * dp[d];
          for (long i = 1; i < k; ++i) acc += xc[i * gd + g] * dp[i * td + d];
          J[d * gd + g] = acc;
        }
      if (j)
        for (int i = 0; i < JL; ++i) j[((size_t)c * nq + q) * JL + i] = J[i];
      if (detj) detj[(size_t)c * nq + q] = det_of(J, td, gd);
      if (pinv || fnormals) {
        double K[9];
        pinv_of(J, td, gd, K);
        if (pinv)
          for (int i = 0; i < JL; ++i) pinv[((size_t)c * nq + q) * JL + i] = K[i];
        if (fnormals && ref_normals && td == gd) {
          double m[3], s2 = 0;
          for (int g = 0; g < gd; ++g) {
            double acc = 0;
            for (int t = 0; t < td; ++t) acc += K[g * td + t] * ref_normals[q * td + t];
            m[g] = acc;
            s2 += acc * acc;
          }
          for (int g = 0; g < gd; ++g) fnormals[((size_t)c * nq + q) * gd + g] = m[g] / sqrt(s2);
        }
      }
      if (normals && td + 1 == gd) {
        double n[3];
        if (gd == 2) {
          n[0] = J[1];
          n[1] = -J[0];
        } else {
          n[0] = J[1] * J[5] - J[2] * J[4];
          n[1] = J[2] * J[3] - J[0] * J[5];
          n[2] = J[0] * J[4] - J[1] * J[3];
        }
        double s2 = 0;
        for (int g = 0; g < gd; ++g) s2 += n[g] * n[g];
        double s = (flags && (flags[c] & 1u)) ? -1.0 : 1.0;
        for (int g = 0; g < gd; ++g) normals[((size_t)c * nq + q) * gd + g] = s * n[g] / sqrt(s2);
      }
    }
  }
  free(phi);
  free(dphi);
  return 0;
}

int or_cpu_factors(int tdim, int gdim, int k, const double* nodes, size_t num_nodes,
                   const double* origin, const uint32_t* dofmap, size_t num_cells,
                   const double* values, const double* derivs, size_t nq, int tile,
                   double* evector, double* x, double* j, double* detj, int threads) {
  const size_t padded = (num_cells + (size_t)tile - 1) / tile * tile;
  const size_t ntiles = padded / tile;
  const int td = tdim, gd = gdim, JL = td * gd;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  /* E-vector [k][g][cell], padding copies the last cell (evector.rs:197-236). */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long c = 0; c < (long)padded; ++c) {
    size_t src = (size_t)c < num_cells ? (size_t)c : num_cells - 1;
    for (int i = 0; i < k; ++i) {
      uint32_t n = dofmap[src * k + i];
      if (n >= num_nodes) n = (uint32_t)(num_nodes - 1);
      for (int g = 0; g < gd; ++g) evector[((size_t)i * gd + g) * padded + c] = nodes[(size_t)n * gd + g];
    }
  }
  /* Factors per tile, lanes = the tile's cells (layout.rs:382-446, direct strategy). */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long t = 0; t < (long)ntiles; ++t) {
    for (size_t q = 0; q < nq; ++q) {
      const double* ph = &values[q * k];
      const double* dp = &derivs[q * k * td];
      for (int lane = 0; lane < tile; ++lane) {
        size_t c = (size_t)t * tile + lane;
        double J[9];
        for (int g = 0; g < gd; ++g) {
          double acc = evector[(size_t)g * padded + c] * ph[0];
          for (int i = 1; i < k; ++i) acc += evector[((size_t)i * gd + g) * padded + c] * ph[i];
          x[(((size_t)t * nq + q) * gd + g) * tile + lane] = acc + origin[g];
        }
        for (int d = 0; d < td; ++d)
          for (int g = 0; g < gd; ++g) {
            double acc = evector[(size_t)g * padded + c] * dp[d];
            for (int i = 1; i < k; ++i)
              acc += evector[((size_t)i * gd + g) * padded + c] * dp[i * td + d];
            J[d * gd + g] = acc;
            j[(((size_t)t * nq + q) * JL + d * gd + g) * tile + lane] = acc;
          }
        detj[((size_t)t * nq + q) * tile + lane] = det_of(J, td, gd);
      }
    }
  }
  return 0;
}

double or_functional(int cell, int degree, int family, const double* nodes, size_t num_nodes,
                     const uint32_t* dofmap, size_t num_cells, const double* pts,
                     const double* wts, size_t nq, double beta, double* grad, int threads) {
  (void)threads;
  const long k = (long)or_num_nodes(cell, degree);
  double* phi = (double*)malloc(sizeof(double) * nq * k);
  double* dphi = (double*)malloc(sizeof(double) * nq * k * 3);
  if (TDIM[cell] != 3 || or_tabulate(cell, degree, family, pts, nq, phi, dphi)) {
    free(phi);
    free(dphi);
    return NAN;
  }
  if (grad) memset(grad, 0, sizeof(double) * num_nodes * 3);
  double total = 0.0;
  double xc[729 * 3];
  uint32_t idx[729];
  for (size_t c = 0; c < num_cells; ++c) {
    for (long i = 0; i < k; ++i) {
      uint32_t n = dofmap[c * k + i];
      if (n >= num_nodes) n = (uint32_t)(num_nodes - 1);
      idx[i] = n;
      for (int g = 0; g < 3; ++g) xc[i * 3 + g] = nodes[(size_t)n * 3 + g];
    }
    for (size_t q = 0; q < nq; ++q) {
      const double* dp = &dphi[q * k * 3];
      double J[9];
      for (int d = 0; d < 3; ++d)
        for (int g = 0; g < 3; ++g) {
          double acc = xc[g] * dp[d];
          for (long i = 1; i < k; ++i) acc += xc[i * 3 + g] * dp[i * 3 + d];
          J[d * 3 + g] = acc;
        }
      double det = det_of(J, 3, 3);
      double n2 = 0;
      for (int i = 0; i < 9; ++i) n2 += J[i] * J[i];
      double cr = cbrt(det);
      double inv23 = 1.0 / (cr * cr);
      total += wts[q] * det + beta * wts[q] * n2 * inv23;
      if (grad) {
        /* cof(J)(g,d) = d det / d J(g,d); J stored column-major at d*3+g. */
        double C[9];
#define M(g, d) J[(d) * 3 + (g)]
        C[0 * 3 + 0] = M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1);
        C[1 * 3 + 0] = -(M(1, 0) * M(2, 2) - M(1, 2) * M(2, 0));
        C[2 * 3 + 0] = M(1, 0) * M(2, 1) - M(1, 1) * M(2, 0);
        C[0 * 3 + 1] = -(M(0, 1) * M(2, 2) - M(0, 2) * M(2, 1));
        C[1 * 3 + 1] = M(0, 0) * M(2, 2) - M(0, 2) * M(2, 0);
        C[2 * 3 + 1] = -(M(0, 0) * M(2, 1) - M(0, 1) * M(2, 0));
        C[0 * 3 + 2] = M(0, 1) * M(1, 2) - M(0, 2) * M(1, 1);
        C[1 * 3 + 2] = -(M(0, 0) * M(1, 2) - M(0, 2) * M(1, 0));
        C[2 * 3 + 2] = M(0, 0) * M(1, 1) - M(0, 1) * M(1, 0);
#undef M
        /* C is stored as C[d*3+g] = cofactor of entry (g, d), same layout as J. */
        double G[9];
        for (int i = 0; i < 9; ++i)
          G[i] = wts[q] * (C[i] + beta * (2.0 * J[i] * inv23 -
                                          (2.0 / 3.0) * n2 * inv23 / det * C[i]));
        for (long i = 0; i < k; ++i)
          for (int g = 0; g < 3; ++g) {
            double s = 0;
            for (int d = 0; d < 3; ++d) s += G[d * 3 + g] * dp[i * 3 + d];
            grad[(size_t)idx[i] * 3 + g] += s;
          }
      }
    }
  }
  free(phi);
  free(dphi);
  return total;
}

/* ---- Pull-back (pull_back.rs:26-180) -------------------------------------------------------- */

int or_pull_back(int cell, int degree, int family, int gdim, const double* nodes,
                 size_t num_nodes, const uint32_t* dofmap, size_t num_cells, const double* x,
                 size_t nq, int max_it, double tol, double* xi, double* step) {
  if (cell < 1 || cell > 7) return -1;
  const int td = TDIM[cell], gd = gdim;
  const long k = (long)or_num_nodes(cell, degree);
  if (k <= 0 || k > 729 || gd < td || gd > 3) return -1;
  /* The start: the vertex centroid, rounded to f32 as the reference's centroid() returns it. */
  double start[3] = {0, 0, 0};
  for (int d = 0; d < td; ++d) {
    double acc = 0;
    for (int v = 0; v < NV[cell]; ++v) acc += VX[cell][v][d];
    start[d] = (double)(float)(acc / NV[cell]);
  }
  const double lo = (double)-0.1f, hi = (double)1.1f;
  double* phi = (double*)malloc(sizeof(double) * k);
  double* dphi = (double*)malloc(sizeof(double) * k * 3);
  double* xc = (double*)malloc(sizeof(double) * k * 3);
  int rc = 0;
  for (size_t c = 0; c < num_cells && rc == 0; ++c) {
    for (long i = 0; i < k; ++i) {
      uint32_t nidx = dofmap[c * k + i];
      if (nidx >= num_nodes) nidx = (uint32_t)(num_nodes - 1);
      for (int g = 0; g < gd; ++g) xc[i * gd + g] = nodes[(size_t)nidx * gd + g];
    }
    for (size_t q = 0; q < nq && rc == 0; ++q) {
      const double* y = &x[(c * nq + q) * gd];
      double p[3] = {start[0], start[1], start[2]};
      double last = DBL_MAX;
      int frozen = 0;
      for (int it = 0; it < max_it; ++it) {
        if (or_tabulate(cell, degree, family, p, 1, phi, dphi)) {
          rc = -1;
          break;
        }
        double X[3], J[9], K[9], delta[3];
        for (int g = 0; g < gd; ++g) {
          double acc = xc[g] * phi[0];
          for (long i = 1; i < k; ++i) acc += xc[i * gd + g] * phi[i];
          X[g] = acc;
        }
        for (int d = 0; d < td; ++d)
          for (int g = 0; g < gd; ++g) {
            double acc = xc[g] * dphi[d];
            for (long i = 1; i < k; ++i) acc += xc[i * gd + g] * dphi[i * td + d];
            J[d * gd + g] = acc;
          }
        pinv_of(J, td, gd, K);
        double s = 0;
        int bad = 0;
        for (int t = 0; t < td; ++t) {
          double acc = K[t] * (y[0] - X[0]);
          for (int g = 1; g < gd; ++g) acc += K[g * td + t] * (y[g] - X[g]);
          delta[t] = acc;
          const double m = fabs(acc);
          if (t == 0 || m > s) s = m;
          if (!(m <= DBL_MAX)) bad = 1;
        }
        if (bad) s = DBL_MAX;
        if (!frozen) {
          for (int d = 0; d < td; ++d) {
            double v = p[d] + delta[d];
            v = v < lo ? lo : v;
            v = v > hi ? hi : v;
            p[d] = v;
          }
          last = s;
          frozen = s <= tol;
        }
      }
      for (int d = 0; d < td; ++d) xi[(c * nq + q) * td + d] = p[d];
      step[c * nq + q] = last;
    }
  }
  free(phi);
  free(dphi);
  free(xc);
  return rc;
}

double or_rel_err(const double* got, const double* want, size_t n) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double d = fabs(got[i] - want[i]);
    if (d > num || d != d) num = d;
    if (fabs(want[i]) > den) den = fabs(want[i]);
  }
  return den > 0.0 ? num / den : num;
}
