/*
 * ed_twin.c -- CPU restatement of the reference's sector H.v and Lanczos recurrence.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the large-size checker and by the
 * cpu_baseline leg of bench.py; never linked into the product.
 *
 * Follows the reference's construction:
 *   sector states: ascending num_sites-bit integers of fixed popcount
 *     (cmpy/basis.py:655-666), here by next-combination enumeration;
 *   index lookup: left bisection of the sorted list (cmpy/operators.py:276-299);
 *   hop sign: (-1)^popcount(state & bits strictly between the sites & low `width`
 *     bits) (cmpy/operators.py:253-273, 425-433);
 *   hop terms: per state and bond, in bond order (cmpy/operators.py:436-527);
 *   diagonal: eps summed site by site + u_i n_up n_dn (cmpy/operators.py:226-422);
 *   H.v: y[target] += val * x[origin] over every term (cmpy/operators.py:626-630),
 *     evaluated matrix-free over the up (x) dn structure, one output row per task;
 *   Lanczos: normalised three-term recurrence (cmpy/exactdiag.py:324-347 in
 *     normalised form).
 * Build: make -C oracle  (gcc -O3 -fopenmp -shared).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

typedef struct {
  int ns, n_up, n_dn, nb, width;
  int64_t nu, nd;
  uint32_t *up, *dn;
  double *d_up, *d_dn;
  double u[32];
  int up_w, dn_w;
  int32_t *up_cnt, *up_col, *dn_cnt, *dn_col;
  double *up_val, *dn_val;
} twin_op;

static int64_t binom64(int n, int k) {
  if (k < 0 || k > n) return 0;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

static int64_t bisect_left(const uint32_t* a, int64_t n, uint32_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

static uint32_t* enumerate(int ns, int n, int64_t* count) {
  *count = binom64(ns, n);
  uint32_t* out = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(*count > 0 ? *count : 1));
  if (n == 0) {
    out[0] = 0;
    return out;
  }
  uint32_t s = (n == 32) ? 0xffffffffu : ((1u << n) - 1u);
  for (int64_t i = 0; i < *count; ++i) {
    out[i] = s;
    if (i + 1 == *count) break;
    uint32_t c = s & (0u - s), r = s + c;
    s = (((r ^ s) >> 2) / c) | r;
  }
  return out;
}

static int sign_of(uint32_t s, int width, int i, int j) {
  int cnt = 0;
  for (int b = i + 1; b < j; ++b)
    if (b < width && ((s >> b) & 1u)) ++cnt;
  return (cnt & 1) ? -1 : 1;
}

static void build_spin(const uint32_t* st, int64_t n, int nb, const int* bonds, const double* hop, int width,
                       int32_t** cnt, int32_t** col, double** val) {
  *cnt = (int32_t*)calloc((size_t)n, sizeof(int32_t));
  *col = (int32_t*)calloc((size_t)n * (nb > 0 ? nb : 1), sizeof(int32_t));
  *val = (double*)calloc((size_t)n * (nb > 0 ? nb : 1), sizeof(double));
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) {
    uint32_t s = st[k];
    int c = 0;
    for (int b = 0; b < nb; ++b) {
      int i = bonds[2 * b], j = bonds[2 * b + 1];
      int oi = (s >> i) & 1, oj = (s >> j) & 1;
      if (oi == oj) continue;
      uint32_t t = s ^ (1u << i) ^ (1u << j);
      (*col)[k * nb + c] = (int32_t)bisect_left(st, n, t);
      (*val)[k * nb + c] = sign_of(s, width, i, j) * hop[b];
      ++c;
    }
    (*cnt)[k] = c;
  }
}

static double weighted(uint32_t s, const double* v, int ns) {
  double acc = 0.0;
  for (int i = 0; i < ns; ++i)
    if ((s >> i) & 1u) acc += v[i];
  return acc;
}

twin_op* twin_create(int ns, int n_up, int n_dn, int nb, const int* bonds, const double* hop, const double* eps,
                     const double* u, int width) {
  if (ns < 1 || ns > 30 || n_up < 0 || n_up > ns || n_dn < 0 || n_dn > ns) return NULL;
  twin_op* op = (twin_op*)calloc(1, sizeof(twin_op));
  op->ns = ns;
  op->n_up = n_up;
  op->n_dn = n_dn;
  op->nb = nb;
  op->width = width;
  for (int i = 0; i < ns; ++i) op->u[i] = u[i];
  op->up = enumerate(ns, n_up, &op->nu);
  op->dn = enumerate(ns, n_dn, &op->nd);
  op->d_up = (double*)malloc(sizeof(double) * op->nu);
  op->d_dn = (double*)malloc(sizeof(double) * op->nd);
  for (int64_t k = 0; k < op->nu; ++k) op->d_up[k] = weighted(op->up[k], eps, ns);
  for (int64_t k = 0; k < op->nd; ++k) op->d_dn[k] = weighted(op->dn[k], eps, ns);
  op->up_w = op->dn_w = nb > 0 ? nb : 1;
  build_spin(op->up, op->nu, nb, bonds, hop, width, &op->up_cnt, &op->up_col, &op->up_val);
  build_spin(op->dn, op->nd, nb, bonds, hop, width, &op->dn_cnt, &op->dn_col, &op->dn_val);
  return op;
}

void twin_destroy(twin_op* op) {
  if (!op) return;
  free(op->up);
  free(op->dn);
  free(op->d_up);
  free(op->d_dn);
  free(op->up_cnt);
  free(op->up_col);
  free(op->up_val);
  free(op->dn_cnt);
  free(op->dn_col);
  free(op->dn_val);
  free(op);
}

void twin_dims(const twin_op* op, int64_t* nu, int64_t* nd) {
  *nu = op->nu;
  *nd = op->nd;
}

/* y = H x, one up row per task. */
void twin_matvec(const twin_op* op, const double* x, double* y) {
  const int64_t nd = op->nd;
  const int w = op->nb > 0 ? op->nb : 1;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t iu = 0; iu < op->nu; ++iu) {
    const uint32_t su = op->up[iu];
    const double du = op->d_up[iu];
    const double* xr = x + iu * nd;
    double* yr = y + iu * nd;
    for (int64_t id = 0; id < nd; ++id) {
      const uint32_t both = su & op->dn[id];
      double uu = 0.0;
      for (int i = 0; i < op->ns; ++i)
        if ((both >> i) & 1u) uu += op->u[i];
      double acc = (du + op->d_dn[id] + uu) * xr[id];
      const int c = op->dn_cnt[id];
      for (int k = 0; k < c; ++k) acc += op->dn_val[id * w + k] * xr[op->dn_col[id * w + k]];
      yr[id] = acc;
    }
    const int cu = op->up_cnt[iu];
    for (int k = 0; k < cu; ++k) {
      const double v = op->up_val[iu * w + k];
      const double* xs = x + (int64_t)op->up_col[iu * w + k] * nd;
      for (int64_t id = 0; id < nd; ++id) yr[id] += v * xs[id];
    }
  }
}

static double dot(const double* a, const double* b, int64_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Normalised three-vector Lanczos from v0 (or a seeded uniform vector) for
 * `iters` steps: alpha[iters], beta[iters] (beta[k] couples k and k+1).
 * Returns the number of steps done (fewer on breakdown). */
int twin_lanczos(const twin_op* op, int iters, const double* v0, uint64_t seed, double* alpha, double* beta) {
  const int64_t D = op->nu * op->nd;
  double* v = (double*)malloc(sizeof(double) * D);
  double* vp = (double*)calloc((size_t)D, sizeof(double));
  double* w = (double*)malloc(sizeof(double) * D);
  if (!v || !vp || !w) {
    free(v);
    free(vp);
    free(w);
    return -1;
  }
  uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
  for (int64_t i = 0; i < D; ++i) {
    if (v0) {
      v[i] = v0[i];
    } else {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      v[i] = (double)(st >> 11) * 0x1.0p-53;
    }
  }
  double nrm = sqrt(dot(v, v, D));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < D; ++i) v[i] /= nrm;
  double b = 0.0;
  int done = 0;
  for (int j = 0; j < iters; ++j) {
    twin_matvec(op, v, w);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < D; ++i) w[i] -= b * vp[i];
    const double a = dot(w, v, D);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < D; ++i) w[i] -= a * v[i];
    b = sqrt(dot(w, w, D));
    alpha[j] = a;
    beta[j] = b;
    done = j + 1;
    if (!(b > 1e-12 * fabs(a))) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < D; ++i) {
      vp[i] = v[i];
      v[i] = w[i] / b;
    }
  }
  free(v);
  free(vp);
  free(w);
  return done;
}

int twin_threads(void) { return omp_get_max_threads(); }
