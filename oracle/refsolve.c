/*
 * CPU baseline of goSPL's per-step path in "reference-solver mode": the
 * linear systems solved the way goSPL solves them on one MPI rank, where
 * PETSc's block Jacobi preconditioner is a single ILU(0) block.
 *
 * TEST/MEASUREMENT INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and the
 * tests): the product library never links or calls it.
 *
 *  - rs_flow_solve:  (I - W^T) q = b by FGMRES(30) + ILU(0), rtol 1e-8,
 *                    max_it 5000 (gospl/flow/flowplex.py:341-408, 38-46);
 *  - rs_spl_solve:   the stream-power matrix (SPL.py:45-114) likewise;
 *  - rs_hill_solve:  the diffusion operator by Richardson + ILU(0),
 *                    rtol 1e-8 (gospl/sed/hillslope.py:347-371).
 * Convergence is tested on the unpreconditioned residual, ||b - A x||_2 <=
 * rtol ||b||_2.  Matrices are assembled in CSR with sorted columns (the
 * natural, input vertex order: no reordering, as PETSc's default ILU).
 * `threads` OpenMP threads run the sparse products and vector updates; the
 * ILU(0) factorisation and its triangular solves are sequential, as PETSc's.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

typedef struct {
  int64_t n;
  int64_t* rp;   /* row pointers (n + 1) */
  int32_t* ci;   /* column indices, sorted in each row */
  double* v;     /* values */
  int64_t* dg;   /* position of the diagonal in each row */
} csr_t;

static void csr_free(csr_t* a) {
  free(a->rp);
  free(a->ci);
  free(a->v);
  free(a->dg);
}

/* sort the entries of each row by column (insertion sort: rows are short) */
static void csr_sort_rows(csr_t* a) {
  for (int64_t i = 0; i < a->n; ++i) {
    for (int64_t p = a->rp[i] + 1; p < a->rp[i + 1]; ++p) {
      const int32_t c = a->ci[p];
      const double v = a->v[p];
      int64_t q = p - 1;
      while (q >= a->rp[i] && a->ci[q] > c) {
        a->ci[q + 1] = a->ci[q];
        a->v[q + 1] = a->v[q];
        --q;
      }
      a->ci[q + 1] = c;
      a->v[q + 1] = v;
    }
    a->dg[i] = -1;
    for (int64_t p = a->rp[i]; p < a->rp[i + 1]; ++p)
      if (a->ci[p] == i) a->dg[i] = p;
  }
}

static int csr_alloc(csr_t* a, int64_t n, int64_t nnz) {
  a->n = n;
  a->rp = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  a->ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  a->v = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  a->dg = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  return (a->rp && a->ci && a->v && a->dg) ? 0 : 1;
}

/* y = A x */
static void spmv(const csr_t* a, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < a->n; ++i) {
    double s = 0.0;
    for (int64_t p = a->rp[i]; p < a->rp[i + 1]; ++p) s += a->v[p] * x[a->ci[p]];
    y[i] = s;
  }
}

static double dot(int64_t n, const double* x, const double* y) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

/* ILU(0) of A in place (values only; the pattern is A's).  Zero pivots are
   replaced by 1 (a row without a diagonal entry keeps the identity). */
static int ilu0(csr_t* f) {
  const int64_t n = f->n;
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  if (!pos) return 1;
  for (int64_t i = 0; i < n; ++i) pos[i] = -1;
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t p = f->rp[i]; p < f->rp[i + 1]; ++p) pos[f->ci[p]] = p;
    for (int64_t p = f->rp[i]; p < f->rp[i + 1]; ++p) {
      const int32_t k = f->ci[p];
      if (k >= i) break;
      const double dk = f->dg[k] >= 0 ? f->v[f->dg[k]] : 1.0;
      const double lik = f->v[p] / (dk != 0.0 ? dk : 1.0);
      f->v[p] = lik;
      for (int64_t q = (f->dg[k] >= 0 ? f->dg[k] : f->rp[k]) + 1; q < f->rp[k + 1]; ++q) {
        const int32_t j = f->ci[q];
        if (j <= k) continue;
        if (pos[j] >= 0) f->v[pos[j]] -= lik * f->v[q];
      }
    }
    for (int64_t p = f->rp[i]; p < f->rp[i + 1]; ++p) pos[f->ci[p]] = -1;
    if (f->dg[i] >= 0 && f->v[f->dg[i]] == 0.0) f->v[f->dg[i]] = 1.0;
  }
  free(pos);
  return 0;
}

/* z = (LU)^-1 r */
static void ilu_apply(const csr_t* f, const double* r, double* z) {
  const int64_t n = f->n;
  for (int64_t i = 0; i < n; ++i) {
    double s = r[i];
    for (int64_t p = f->rp[i]; p < f->rp[i + 1] && f->ci[p] < i; ++p) s -= f->v[p] * z[f->ci[p]];
    z[i] = s;
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    double s = z[i];
    const int64_t d = f->dg[i];
    for (int64_t p = (d >= 0 ? d + 1 : f->rp[i]); p < f->rp[i + 1]; ++p)
      if (f->ci[p] > i) s -= f->v[p] * z[f->ci[p]];
    z[i] = d >= 0 ? s / f->v[d] : s;
  }
}

/* FGMRES(m) with right ILU(0) preconditioning; x holds the initial guess.
   Returns iterations, or -1 if not converged within maxit. */
static int fgmres(const csr_t* a, const csr_t* f, const double* b, double* x, double rtol, int m, int maxit,
                  double* relres) {
  const int64_t n = a->n;
  double* V = (double*)malloc(sizeof(double) * (size_t)n * (m + 1));
  double* Z = (double*)malloc(sizeof(double) * (size_t)n * m);
  double* w = (double*)malloc(sizeof(double) * (size_t)n);
  double* H = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
  double *g = (double*)calloc(m + 1, sizeof(double)), *cs = (double*)calloc(m, sizeof(double)),
         *sn = (double*)calloc(m, sizeof(double)), *y = (double*)calloc(m, sizeof(double));
  int it = 0, conv = 0;
  const double bn = sqrt(dot(n, b, b));
  const double tol = rtol * (bn > 0.0 ? bn : 1.0);
  double res = 0.0;
  for (;;) {
    spmv(a, x, w);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
    res = sqrt(dot(n, w, w));
    if (res <= tol) {
      conv = 1;
      break;
    }
    if (it >= maxit) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) V[i] = w[i] / res;
    memset(g, 0, sizeof(double) * (m + 1));
    g[0] = res;
    int j = 0;
    for (; j < m && it < maxit; ++j) {
      ilu_apply(f, V + (size_t)j * n, Z + (size_t)j * n);
      spmv(a, Z + (size_t)j * n, w);
      for (int i = 0; i <= j; ++i) {
        const double hij = dot(n, w, V + (size_t)i * n);
        H[i * m + j] = hij;
        const double* vi = V + (size_t)i * n;
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < n; ++t) w[t] -= hij * vi[t];
      }
      const double hn = sqrt(dot(n, w, w));
      H[(j + 1) * m + j] = hn;
      if (hn > 0.0) {
        double* vn = V + (size_t)(j + 1) * n;
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < n; ++t) vn[t] = w[t] / hn;
      }
      for (int i = 0; i < j; ++i) {
        const double t0 = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
        H[(i + 1) * m + j] = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
        H[i * m + j] = t0;
      }
      const double den = hypot(H[j * m + j], H[(j + 1) * m + j]);
      cs[j] = den > 0.0 ? H[j * m + j] / den : 1.0;
      sn[j] = den > 0.0 ? H[(j + 1) * m + j] / den : 0.0;
      H[j * m + j] = den;
      H[(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++it;
      if (fabs(g[j + 1]) <= tol || hn == 0.0) {
        ++j;
        break;
      }
    }
    for (int i = j - 1; i >= 0; --i) {
      double s = g[i];
      for (int k = i + 1; k < j; ++k) s -= H[i * m + k] * y[k];
      y[i] = H[i * m + i] != 0.0 ? s / H[i * m + i] : 0.0;
    }
    for (int i = 0; i < j; ++i) {
      const double* zi = Z + (size_t)i * n;
      const double yi = y[i];
#pragma omp parallel for schedule(static)
      for (int64_t t = 0; t < n; ++t) x[t] += yi * zi[t];
    }
  }
  if (relres) *relres = res / (bn > 0.0 ? bn : 1.0);
  free(V);
  free(Z);
  free(w);
  free(H);
  free(g);
  free(cs);
  free(sn);
  free(y);
  return conv ? it : -1;
}

/* Richardson with ILU(0): x += (LU)^-1 (b - A x) */
static int richardson(const csr_t* a, const csr_t* f, const double* b, double* x, double rtol, int maxit,
                      double* relres) {
  const int64_t n = a->n;
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  const double bn = sqrt(dot(n, b, b));
  const double tol = rtol * (bn > 0.0 ? bn : 1.0);
  int it = 0, conv = 0;
  double res = 0.0;
  for (;;) {
    spmv(a, x, r);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    res = sqrt(dot(n, r, r));
    if (res <= tol) {
      conv = 1;
      break;
    }
    if (it >= maxit) break;
    ilu_apply(f, r, z);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] += z[i];
    ++it;
  }
  if (relres) *relres = res / (bn > 0.0 ? bn : 1.0);
  free(r);
  free(z);
  return conv ? it : -1;
}

static int solve(csr_t* a, const double* b, double* x, int krylov, double rtol, int maxit, int threads,
                 double* relres) {
  omp_set_num_threads(threads > 0 ? threads : 1);
  csr_t f;
  if (csr_alloc(&f, a->n, a->rp[a->n])) return -2;
  memcpy(f.rp, a->rp, sizeof(int64_t) * (size_t)(a->n + 1));
  memcpy(f.ci, a->ci, sizeof(int32_t) * (size_t)a->rp[a->n]);
  memcpy(f.v, a->v, sizeof(double) * (size_t)a->rp[a->n]);
  memcpy(f.dg, a->dg, sizeof(int64_t) * (size_t)a->n);
  int it = -2;
  if (!ilu0(&f))
    it = krylov ? fgmres(a, &f, b, x, rtol, 30, maxit, relres) : richardson(a, &f, b, x, rtol, maxit, relres);
  csr_free(&f);
  return it;
}

/* (I - W^T): row r holds 1 at r and -w_jk at column j for every donor j with
   receiver k == r (j != r, w != 0) -- matrixFlow (flowplex.py:429-478). */
int rs_flow_solve(int64_t n, int K, const int32_t* rcv, const double* wgt, const double* b, double* x,
                  double rtol, int maxit, int threads, double* relres) {
  csr_t a;
  int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  for (int64_t j = 0; j < n; ++j) {
    cnt[j] += 1;
    for (int k = 0; k < K; ++k) {
      const int r = rcv[j * K + k];
      if (r != j && r >= 0 && wgt[j * K + k] != 0.0) cnt[r] += 1;
    }
  }
  int64_t nnz = 0;
  for (int64_t i = 0; i < n; ++i) nnz += cnt[i];
  if (csr_alloc(&a, n, nnz)) return -2;
  for (int64_t i = 0; i < n; ++i) a.rp[i + 1] = a.rp[i] + cnt[i];
  memset(cnt, 0, sizeof(int64_t) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    a.ci[a.rp[i]] = (int32_t)i;
    a.v[a.rp[i]] = 1.0;
    cnt[i] = 1;
  }
  for (int64_t j = 0; j < n; ++j)
    for (int k = 0; k < K; ++k) {
      const int r = rcv[j * K + k];
      const double w = wgt[j * K + k];
      if (r != j && r >= 0 && w != 0.0) {
        const int64_t p = a.rp[r] + cnt[r]++;
        a.ci[p] = (int32_t)j;
        a.v[p] = -w;
      }
    }
  free(cnt);
  csr_sort_rows(&a);
  const int it = solve(&a, b, x, 1, rtol, maxit, threads, relres);
  csr_free(&a);
  return it;
}

/* diag_i x_i + sum_k off_ik x_rcv_ik = rhs_i (SPL.py:80-108) */
int rs_spl_solve(int64_t n, int K, const int32_t* rcv, const double* off, const double* diag, const double* rhs,
                 double* x, double rtol, int maxit, int threads, double* relres) {
  csr_t a;
  int64_t nnz = 0;
  for (int64_t i = 0; i < n; ++i) {
    nnz += 1;
    for (int k = 0; k < K; ++k)
      if (rcv[i * K + k] != i && rcv[i * K + k] >= 0 && off[i * K + k] != 0.0) nnz += 1;
  }
  if (csr_alloc(&a, n, nnz)) return -2;
  int64_t p = 0;
  for (int64_t i = 0; i < n; ++i) {
    a.ci[p] = (int32_t)i;
    a.v[p++] = diag[i];
    for (int k = 0; k < K; ++k) {
      const int r = rcv[i * K + k];
      if (r != i && r >= 0 && off[i * K + k] != 0.0) {
        a.ci[p] = r;
        a.v[p++] = off[i * K + k];
      }
    }
    a.rp[i + 1] = p;
  }
  csr_sort_rows(&a);
  const int it = solve(&a, rhs, x, 1, rtol, maxit, threads, relres);
  csr_free(&a);
  return it;
}

/* the diffusion operator from sethillslopecoeff's (n, 13) rows: column 0 the
   diagonal, column p + 1 the entry of neighbour p; rows with a zero diagonal
   become identity rows */
int rs_hill_solve(int64_t n, const int32_t* ngb, const int32_t* nnb, const double* coeff, const double* rhs,
                  double* x, double rtol, int maxit, int threads, double* relres) {
  csr_t a;
  int64_t nnz = 0;
  for (int64_t i = 0; i < n; ++i) nnz += 1 + nnb[i];
  if (csr_alloc(&a, n, nnz)) return -2;
  int64_t p = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double d = coeff[i * 13];
    a.ci[p] = (int32_t)i;
    a.v[p++] = d != 0.0 ? d : 1.0;
    if (d != 0.0)
      for (int q = 0; q < nnb[i]; ++q) {
        const double c = coeff[i * 13 + q + 1];
        if (c != 0.0 && ngb[i * 12 + q] >= 0) {
          a.ci[p] = ngb[i * 12 + q];
          a.v[p++] = c;
        }
      }
    a.rp[i + 1] = p;
  }
  csr_sort_rows(&a);
  const int it = solve(&a, rhs, x, 0, rtol, maxit, threads, relres);
  csr_free(&a);
  return it;
}
