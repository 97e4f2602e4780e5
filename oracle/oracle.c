/*
 * CPU restatement (FP64) of goSPL's per-timestep surface-process hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that the HIP kernels are compared
 * against (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg).
 * It is never linked into, called by or shipped with the product library.
 *
 * Arrays are in input-mesh vertex order.  Neighbour tables follow goSPL's
 * Fortran module state: ngb/elgt/vdist are (nb, 12) row-major, nnb (nb).
 * Every function cites the reference code it restates.
 *
 * Pinning (see tests/test_oracle_pins.py): mfdreceivers, the fill, label_pits,
 * fill_dir/nghb_dir, getpitvol and sethillslopecoeff are compared bit-for-bit
 * with goSPL's own Fortran compiled from source (oracle/_ref) on goSPL's test
 * meshes and on random inputs; the DAG solves are pinned by residuals and
 * brute-force dense solves; the hillslope solve by its residual and the
 * closed-form 1-D parabola.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXNB 12

/* ------------------------------------------------------------------------ */
/* binary min-heap on (Z, id) (functions.F90:105-182 PQpush/PQpop)          */
/* ------------------------------------------------------------------------ */
typedef struct { double z; int64_t id; } node_t;
typedef struct { node_t* buf; int64_t n, cap; } heap_t;

static void hpush(heap_t* h, double z, int64_t id) {
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 1024;
    h->buf = (node_t*)realloc(h->buf, (size_t)h->cap * sizeof(node_t));
  }
  int64_t i = h->n++;
  h->buf[i].z = z;
  h->buf[i].id = id;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h->buf[i].z < h->buf[p].z) {
      node_t t = h->buf[i]; h->buf[i] = h->buf[p]; h->buf[p] = t; i = p;
    } else break;
  }
}

static node_t hpop(heap_t* h) {
  node_t r = h->buf[0];
  h->buf[0] = h->buf[--h->n];
  int64_t p = 0;
  for (;;) {
    int64_t c = 2 * p + 1;
    if (c >= h->n) break;
    if (c + 1 < h->n && h->buf[c + 1].z < h->buf[c].z) c++;
    if (h->buf[p].z > h->buf[c].z) {
      node_t t = h->buf[p]; h->buf[p] = h->buf[c]; h->buf[c] = t; p = c;
    } else break;
  }
  return r;
}

/* ------------------------------------------------------------------------ */
/* quicksort/split (functions.F90:683-762), 1-based, including the left scan */
/* bound of 12 in split.                                                      */
/* ------------------------------------------------------------------------ */
static void f_split(double* a, int low, int high, int* mid, int* idx) {
  int left = low, right = high;
  double pivot = a[low];
  int ipivot = idx[low];
  for (;;) {
    if (left >= right) break;
    for (;;) {
      if (left >= right || a[right] < pivot) break;
      right--;
    }
    for (;;) {
      if (left > 12) break;
      if (a[left] > pivot) break;
      left++;
    }
    if (left < right) {
      double s = a[left]; a[left] = a[right]; a[right] = s;
      int t = idx[left]; idx[left] = idx[right]; idx[right] = t;
    }
  }
  a[low] = a[right];
  a[right] = pivot;
  *mid = right;
  idx[low] = idx[right];
  idx[right] = ipivot;
}

static void f_quicksort(double* a, int first, int last, int* idx) {
  if (first < last) {
    int mid;
    f_split(a, first, last, &mid, idx);
    f_quicksort(a, first, mid - 1, idx);
    f_quicksort(a, mid + 1, last, idx);
  }
}

/* ------------------------------------------------------------------------ */
/* mfdreceivers (functions.F90:1544-1671)                                     */
/* rcv/dist/wgt are (nb, nrcv) row-major.                                     */
/* ------------------------------------------------------------------------ */
void or_mfdreceivers(int nrcv, double fexp0, const double* elev, double sl,
                     const int32_t* gid, int64_t nb, const int32_t* ngb,
                     const int32_t* nnb, const double* elgt, int32_t* rcv,
                     double* dist, double* wgt) {
  for (int64_t k = 0; k < nb * nrcv; ++k) { rcv[k] = -1; dist[k] = 0.0; wgt[k] = 0.0; }
  double dst[13];
  for (int p = 0; p < 13; ++p) dst[p] = 0.0;
  for (int64_t k = 0; k < nb; ++k) {
    int32_t* r = rcv + k * nrcv;
    double* d = dist + k * nrcv;
    double* w = wgt + k * nrcv;
    if (elev[k] <= sl) {
      for (int p = 0; p < nrcv; ++p) r[p] = (int32_t)k;
      continue;
    }
    const int32_t* nk = ngb + k * MAXNB;
    const double* lk = elgt + k * MAXNB;
    double e = 0.0;
    for (int p = 0; p < nnb[k]; ++p) {
      int32_t n = nk[p];
      if (n >= 0 && lk[p] > 0.0) {
        double val = (elev[k] - elev[n]) / lk[p];
        e = (val > e) ? val : e;
      }
    }
    double fexp = (e > 0.0) ? 5.0 * (e < 1.0 ? e : 1.0) + fexp0 : 1.0;
    double slp[13], slope[13];
    int id[13];
    for (int p = 0; p < 13; ++p) { slp[p] = 0.0; id[p] = 0; slope[p] = 0.0; }
    int kk = 0;
    for (int p = 0; p < nnb[k]; ++p) {
      int32_t n = nk[p];
      if (n >= 0 && lk[p] > 0.0) {
        double val = pow(elev[k] - elev[n], fexp) / lk[p];
        if (val > 0.0) {
          kk++;
          slp[kk] = val * (1.0 - 1.0e-15 * (double)gid[n]);
          id[kk] = n;
          dst[kk] = lk[p];
        }
      }
    }
    if (kk == 0) {
      for (int p = 0; p < nrcv; ++p) r[p] = (int32_t)k;
    } else if (kk <= nrcv) {
      double val = 0.0;
      for (int p = 0; p < nrcv; ++p) r[p] = (int32_t)k;
      for (int p = 1; p <= kk; ++p) {
        r[p - 1] = id[p];
        d[p - 1] = dst[p];
        val = val + slp[p];
      }
      for (int p = 1; p <= nrcv; ++p) w[p - 1] = slp[p] / val;
    } else {
      for (int p = 0; p < nrcv; ++p) r[p] = (int32_t)k;
      f_quicksort(slp, 1, kk, id);
      int n = 0;
      double val = 0.0;
      for (int p = kk; p >= kk - nrcv + 1; --p) {
        n++;
        slope[n] = slp[p];
        r[n - 1] = id[p];
        d[n - 1] = dst[p];
        val = val + slp[p];
      }
      for (int p = 1; p <= nrcv; ++p) w[p - 1] = slope[p] / val;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Depression fill, one-rank path of PITFill._performFilling                 */
/* (gospl/flow/pitfilling.py:557-702).                                        */
/*                                                                            */
/* edge_tile(0, borders, dem) (functions.F90:2036-2080) marks seeds (cells    */
/* below the level touching a cell at/above it, and draining borders) and     */
/* excluded cells (below the level with no neighbour at/above it).            */
/* fill_tile (functions.F90:2082-2178) floods from the seeds with a priority  */
/* queue plus a plain queue for cells at the current level; fillz is then the */
/* minimax ("spill") elevation over paths from a seed.  With one rank every   */
/* label's spill-graph level is at or below the elevation of the seed that    */
/* created it, so fill_edges/fill_depressions (functions.F90:2180-2329)        */
/* return fillz for every reached cell above the level and leave the rest at  */
/* dem; that is what is written here.  tests/test_oracle_pins.py runs the     */
/* reference's full chain (edge_tile, fill_tile, graph_nodes, combine_edges,  */
/* fill_edges, fill_depressions and the pitfilling.py glue) and checks it is  */
/* bit-identical to this function.                                            */
/* ------------------------------------------------------------------------ */
int or_fill(int64_t nb, const int32_t* ngb, const int32_t* nnb, const double* dem,
            const int8_t* outlet, double* out, int8_t* reached, int8_t* ledge_out) {
  int8_t* ledge = (int8_t*)malloc((size_t)nb);
  int8_t* inids = (int8_t*)malloc((size_t)nb);
  int32_t* labels = (int32_t*)malloc((size_t)nb * sizeof(int32_t));
  double* fillz = (double*)malloc((size_t)nb * sizeof(double));
  if (!ledge || !inids || !labels || !fillz) return -1;
  const double lvl = 0.0;
  for (int64_t c = 0; c < nb; ++c) {
    int8_t le = -1;
    int border = outlet[c] ? 1 : -1;
    if (dem[c] < lvl) {
      double maxe = lvl - 1;
      for (int p = 0; p < nnb[c]; ++p) {
        int32_t nc = ngb[c * MAXNB + p];
        if (nc >= 0) {
          maxe = (dem[nc] > maxe) ? dem[nc] : maxe;
          if (maxe >= lvl) break;
        }
      }
      if (maxe >= lvl && border > -1) le = (int8_t)(border + 1);
      else if (maxe >= lvl) le = 0;
      else le = -2;
    } else if (border > -1) {
      le = (int8_t)(border + 1);
    }
    ledge[c] = le;
    inids[c] = (le == -2) ? 0 : 1;
    labels[c] = -1;
    fillz[c] = dem[c];
  }
  heap_t pq = {0, 0, 0};
  int64_t* plain = (int64_t*)malloc((size_t)(nb + 1) * sizeof(int64_t));
  int64_t np = 0;
  for (int64_t c = 0; c < nb; ++c)
    if (ledge[c] >= 0) hpush(&pq, fillz[c], c);
  int32_t nblab = 1;
  while (pq.n > 0 || np > 0) {
    int64_t i;
    if (np > 0) {          /* plain queue pop: first element, last moved in */
      i = plain[0];
      plain[0] = plain[--np];
    } else {
      i = hpop(&pq).id;
    }
    if (labels[i] <= 0) {
      for (int k = 0; k < nnb[i]; ++k) {
        int32_t c = ngb[i * MAXNB + k];
        if (c >= 0 && inids[c] > 0 && labels[c] > 0 && fillz[c] <= fillz[i]) labels[i] = labels[c];
      }
      if (labels[i] <= 0) labels[i] = nblab++;
    }
    for (int k = 0; k < nnb[i]; ++k) {
      int32_t c = ngb[i * MAXNB + k];
      if (c >= 0 && inids[c] > 0 && labels[c] <= 0) {
        labels[c] = labels[i];
        if (fillz[c] <= fillz[i]) {
          fillz[c] = fillz[i];
          plain[np++] = c;
        } else {
          hpush(&pq, fillz[c], c);
        }
      }
    }
  }
  for (int64_t c = 0; c < nb; ++c) {
    reached[c] = labels[c] > 0;
    out[c] = (dem[c] > lvl && labels[c] >= 0) ? fillz[c] : dem[c];
    if (ledge_out) ledge_out[c] = ledge[c];
  }
  free(pq.buf); free(plain); free(ledge); free(inids); free(labels); free(fillz);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* label_pits (functions.F90:2420-2499)                                       */
/* ------------------------------------------------------------------------ */
void or_label_pits(double lvl, const double* fill, int64_t nb, const int32_t* ngb,
                   const int32_t* nnb, int32_t* label) {
  int8_t* flag = (int8_t*)calloc((size_t)nb, 1);
  int64_t* q = (int64_t*)malloc((size_t)(nb + 1) * sizeof(int64_t));
  int32_t lbl = 0;
  for (int64_t i = 0; i < nb; ++i) label[i] = -1;
  for (int64_t i = 0; i < nb; ++i) {
    if (fill[i] > lvl && !flag[i]) {
      int noflow = 1;
      for (int k = 0; k < nnb[i]; ++k) {
        int32_t c = ngb[i * MAXNB + k];
        if (c >= 0 && fill[i] > fill[c]) { noflow = 0; break; }
      }
      if (noflow) {
        flag[i] = 1;
        lbl++;
        label[i] = lbl;
        int64_t nq = 0;
        q[nq++] = i;
        while (nq > 0) {
          int64_t p = q[0];
          q[0] = q[--nq];
          for (int k = 0; k < nnb[p]; ++k) {
            int32_t c = ngb[p * MAXNB + k];
            if (c >= 0 && fill[c] == fill[p] && !flag[c]) {
              q[nq++] = c;
              label[c] = lbl;
              flag[c] = 1;
            }
          }
        }
      }
    } else {
      flag[i] = 1;
    }
  }
  free(flag); free(q);
}

/* ------------------------------------------------------------------------ */
/* fill_dir (functions.F90:1845-1899) followed by nghb_dir relaxation         */
/* (:1901-1969) to its fixed point (pitfilling.py:400-477 at one rank).       */
/* ------------------------------------------------------------------------ */
void or_fill_dir(const int32_t* spill, int64_t m, const int32_t* pitids, const double* h,
                 int64_t n, const int32_t* ngb, const int32_t* nnb, int32_t* ptdir) {
  int8_t* flag = (int8_t*)calloc((size_t)n, 1);
  heap_t pq = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i) ptdir[i] = -1;
  for (int64_t c = 0; c < m; ++c) {
    int64_t i = spill[c];
    flag[i] = 1;
    ptdir[i] = 1;
    hpush(&pq, 1.0, i);
  }
  while (pq.n > 0) {
    node_t t = hpop(&pq);
    int64_t i = t.id;
    for (int k = 0; k < nnb[i]; ++k) {
      int32_t c = ngb[i * MAXNB + k];
      if (c >= 0 && !flag[c] && pitids[c] > -1 && h[c] >= h[i]) {
        flag[c] = 1;
        if (ptdir[c] < ptdir[i] + 1) {
          ptdir[c] = ptdir[i] + 1;
          hpush(&pq, t.z + 1.0, c);
        }
      }
    }
  }
  free(flag);
  free(pq.buf);
}

int or_nghb_dir(const int32_t* pitids, const double* h, const int32_t* ptdir, int64_t n,
                const int32_t* ngb, const int32_t* nnb, int32_t* nptdir) {
  heap_t pq = {0, 0, 0};
  int changed = 0;
  for (int64_t c = 0; c < n; ++c) {
    nptdir[c] = ptdir[c];
    if (nptdir[c] > -1) hpush(&pq, (double)nptdir[c], c);
  }
  while (pq.n > 0) {
    node_t t = hpop(&pq);
    int64_t i = t.id;
    int32_t di = nptdir[i];
    if ((int32_t)t.z > di) continue;
    for (int k = 0; k < nnb[i]; ++k) {
      int32_t c = ngb[i * MAXNB + k];
      if (c >= 0 && pitids[c] > -1 && h[c] >= h[i]) {
        if (nptdir[c] < 0 || di + 1 < nptdir[c]) {
          nptdir[c] = di + 1;
          hpush(&pq, (double)(di + 1), c);
          changed++;
        }
      }
    }
  }
  free(pq.buf);
  return changed;
}

/* ------------------------------------------------------------------------ */
/* getpitvol (functions.F90:1971-2003); vol is (npit, 4) row-major.          */
/* ------------------------------------------------------------------------ */
void or_getpitvol(const double* hlvl, const double* elev, const int32_t* pit,
                  const int32_t* ids, const double* area, int64_t m, int64_t npit, double* vol) {
  for (int64_t i = 0; i < npit * 4; ++i) vol[i] = 0.0;
  for (int64_t i = 0; i < m; ++i) {
    int32_t p = pit[i];
    if (p > -1 && ids[i] > 0) {
      for (int k = 0; k < 4; ++k) {
        if (hlvl[p * 4 + k] > elev[i]) vol[p * 4 + k] = vol[p * 4 + k] + (hlvl[p * 4 + k] - elev[i]) * area[i];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* sethillslopecoeff (functions.F90:901-938); dcoeff is (nb, 13) row-major.  */
/* ------------------------------------------------------------------------ */
void or_sethillslopecoeff(int64_t nb, const int32_t* ngb, const int32_t* nnb,
                          const double* area, const double* elgt, const double* vdist,
                          const double* kd, double* dcoeff) {
  for (int64_t i = 0; i < nb * 13; ++i) dcoeff[i] = 0.0;
  for (int64_t k = 0; k < nb; ++k) {
    double s1 = 0.0;
    if (area[k] > 0.0) {
      double ck = kd[k];
      for (int p = 0; p < nnb[k]; ++p) {
        if (vdist[k * MAXNB + p] > 0.0) {
          int32_t n = ngb[k * MAXNB + p];
          double cn = kd[n];
          double c = 0.5 * (ck + cn) / area[k];
          double v = c * vdist[k * MAXNB + p] / elgt[k * MAXNB + p];
          s1 = s1 + v;
          dcoeff[k * 13 + p + 1] = -v;
        }
      }
      dcoeff[k * 13] = 1.0 + s1;
    }
  }
}

#define OR_CYCLE_MAXIT 100000

/* ------------------------------------------------------------------------ */
/* Exact solve of (I - W)^T q = b for a receiver DAG (flow accumulation,     */
/* flowplex.py:429-478 matrix, solved by _solve_KSP :323-427 to rtol 1e-8;   */
/* here by a topological sweep).  rcv/wgt are (nb, K) row-major; a receiver  */
/* equal to the row or a zero weight adds nothing.  Returns the number of    */
/* vertices left unsolved (nonzero only if the receivers hold a cycle that   */
/* does not drain; see the fixed-point pass at the end).                     */
/* ------------------------------------------------------------------------ */
int64_t or_flowacc(int64_t nb, int K, const int32_t* rcv, const double* wgt,
                   const double* b, double* q, const int32_t* ngb, const int32_t* nnb) {
  /* Receivers are mesh neighbours, so the donors of i are found among its
     FV neighbours; each solved vertex sums its donors in neighbour order. */
  int32_t* pend = (int32_t*)calloc((size_t)nb, sizeof(int32_t));
  int64_t* st = (int64_t*)malloc((size_t)(nb + 1) * sizeof(int64_t));
  for (int64_t j = 0; j < nb; ++j)
    for (int k = 0; k < K; ++k) {
      int32_t r = rcv[j * K + k];
      if (r >= 0 && r != j && wgt[j * K + k] != 0.0) pend[r]++;
    }
  int64_t ns = 0, done = 0;
  for (int64_t j = 0; j < nb; ++j) if (pend[j] == 0) st[ns++] = j;
  while (ns > 0) {
    int64_t i = st[--ns];
    double s = b[i];
    for (int p = 0; p < nnb[i]; ++p) {
      int32_t j = ngb[i * MAXNB + p];
      if (j < 0) continue;
      for (int k = 0; k < K; ++k) {
        if (rcv[j * K + k] == i) {
          double w = wgt[j * K + k];
          if (w != 0.0) s = s + w * q[j];
        }
      }
    }
    q[i] = s;
    done++;
    for (int k = 0; k < K; ++k) {
      int32_t r = rcv[i * K + k];
      if (r >= 0 && r != i && wgt[i * K + k] != 0.0) {
        if (--pend[r] == 0) st[ns++] = r;
      }
    }
  }
  int64_t left = nb - done;
  if (left > 0) {
    /* Vertices left over: a cycle in the receivers (cascade passes route on
       waterFilled with flat receivers, flowplex.py:655-661) and everything
       below it.  goSPL's iterative solve converges there when the cycle
       drains and gives up otherwise, zeroing a non-fatal solve
       (flowplex.py:178-322, 410-424).  Restated as the fixed-point
       iteration q <- b + W^T q on the left-over vertices (Jacobi form, from
       q = 0, solved vertices fixed) until the largest change is at most
       1e-14 of the largest value; 0 is returned when it converges, the
       number of left-over vertices when it does not within
       OR_CYCLE_MAXIT passes or leaves the finite range. */
    int64_t* U = (int64_t*)malloc((size_t)left * sizeof(int64_t));
    double* xn = (double*)malloc((size_t)left * sizeof(double));
    int64_t nu = 0;
    for (int64_t j = 0; j < nb; ++j)
      if (pend[j] > 0) { U[nu++] = j; q[j] = 0.0; }
    int conv = 0;
    for (int it = 0; it < OR_CYCLE_MAXIT; ++it) {
      double dmax = 0.0, xmax = 0.0;
      int bad = 0;
      for (int64_t t = 0; t < nu; ++t) {
        const int64_t i = U[t];
        double s = b[i];
        for (int p = 0; p < nnb[i]; ++p) {
          int32_t j = ngb[i * MAXNB + p];
          if (j < 0) continue;
          for (int k = 0; k < K; ++k) {
            if (rcv[j * K + k] == i) {
              double w = wgt[j * K + k];
              if (w != 0.0) s = s + w * q[j];
            }
          }
        }
        xn[t] = s;
        if (!isfinite(s)) bad = 1;
        const double d = fabs(s - q[i]);
        if (d > dmax) dmax = d;
        if (fabs(s) > xmax) xmax = fabs(s);
      }
      for (int64_t t = 0; t < nu; ++t) q[U[t]] = xn[t];
      if (bad) break;
      if (dmax <= 1.0e-14 * xmax) { conv = 1; break; }
    }
    free(U); free(xn);
    left = conv ? 0 : nu;
  }
  free(pend); free(st);
  return left;
}

/* ------------------------------------------------------------------------ */
/* Exact solve of the implicit stream-power system (SPL.py:45-114 matrix,    */
/* solved in SPL.py:297): row i reads diag[i]*x_i + sum_k off[i,k]*x_rcv = h */
/* with receivers below the row, so a sweep from the sinks upward is exact.  */
/* ------------------------------------------------------------------------ */
int64_t or_spl_solve(int64_t nb, int K, const int32_t* rcv, const double* off,
                     const double* diag, const double* rhs, double* x) {
  int32_t* pend = (int32_t*)calloc((size_t)nb, sizeof(int32_t));
  int32_t* dstart = (int32_t*)calloc((size_t)nb + 1, sizeof(int32_t));
  int64_t* don = NULL;
  int64_t* st = (int64_t*)malloc((size_t)(nb + 1) * sizeof(int64_t));
  /* donor lists (reverse edges) */
  for (int64_t j = 0; j < nb; ++j)
    for (int k = 0; k < K; ++k) {
      int32_t r = rcv[j * K + k];
      if (r >= 0 && r != j && off[j * K + k] != 0.0) { pend[j]++; dstart[r + 1]++; }
    }
  for (int64_t j = 0; j < nb; ++j) dstart[j + 1] += dstart[j];
  don = (int64_t*)malloc((size_t)(dstart[nb] + 1) * sizeof(int64_t));
  int32_t* fillp = (int32_t*)calloc((size_t)nb, sizeof(int32_t));
  for (int64_t j = 0; j < nb; ++j)
    for (int k = 0; k < K; ++k) {
      int32_t r = rcv[j * K + k];
      if (r >= 0 && r != j && off[j * K + k] != 0.0) don[dstart[r] + fillp[r]++] = j;
    }
  int64_t ns = 0, done = 0;
  for (int64_t j = 0; j < nb; ++j) if (pend[j] == 0) st[ns++] = j;
  while (ns > 0) {
    int64_t i = st[--ns];
    double s = rhs[i];
    for (int k = 0; k < K; ++k) {
      int32_t r = rcv[i * K + k];
      if (r >= 0 && r != i && off[i * K + k] != 0.0) s -= off[i * K + k] * x[r];
    }
    x[i] = s / diag[i];
    done++;
    for (int32_t e = dstart[i]; e < dstart[i + 1]; ++e) {
      int64_t j = don[e];
      if (--pend[j] == 0) st[ns++] = j;
    }
  }
  for (int64_t j = 0; j < nb; ++j) if (pend[j] > 0) x[j] = NAN;
  free(pend); free(dstart); free(don); free(st); free(fillp);
  return nb - done;
}

/* ------------------------------------------------------------------------ */
/* Linear hillslope solve M x = rhs with M from sethillslopecoeff            */
/* (hillslope.py:171-209, Richardson+ILU(0) to rtol 1e-8 in the reference;   */
/* here Jacobi iteration to a fixed point, which the strongly diagonally     */
/* dominant M reaches in a few sweeps).  coeff is (nb,13): col 0 diagonal,   */
/* col p+1 the entry for neighbour p.  Rows with a zero diagonal are kept at */
/* rhs.  Returns the number of sweeps.                                        */
/* ------------------------------------------------------------------------ */
/* Chebyshev semi-iteration over the same Jacobi sweep: x1 = sweep(x0 = rhs),
 * then x_{k+1} = omega[k] (sweep(x_k) - x_{k-1}) + x_{k-1} for k = 1 ..
 * nsweep - 1, each row's sweep in the operation order of or_hillslope_solve. */
void or_cheb_smooth(int64_t nb, const int32_t* ngb, const int32_t* nnb, const double* coeff,
                    const double* rhs, double* x, int nsweep, const double* omega) {
  double* xp = (double*)malloc((size_t)nb * sizeof(double));
  double* y = (double*)malloc((size_t)nb * sizeof(double));
  for (int64_t i = 0; i < nb; ++i) xp[i] = rhs[i];
  for (int k = 0; k < nsweep; ++k) {
    const double* xc = k == 0 ? xp : x;
    for (int64_t i = 0; i < nb; ++i) {
      double d = coeff[i * 13];
      double v;
      if (d == 0.0) {
        v = rhs[i];
      } else {
        double s = rhs[i];
        for (int p = 0; p < nnb[i]; ++p) {
          double c = coeff[i * 13 + p + 1];
          if (c != 0.0) s -= c * xc[ngb[i * MAXNB + p]];
        }
        v = s / d;
      }
      y[i] = k == 0 ? v : omega[k] * (v - xp[i]) + xp[i];
    }
    if (k > 0)
      for (int64_t i = 0; i < nb; ++i) xp[i] = x[i];
    for (int64_t i = 0; i < nb; ++i) x[i] = y[i];
  }
  free(xp);
  free(y);
}

int or_hillslope_solve(int64_t nb, const int32_t* ngb, const int32_t* nnb,
                       const double* coeff, const double* rhs, double* x, int maxit) {
  double* y = (double*)malloc((size_t)nb * sizeof(double));
  for (int64_t i = 0; i < nb; ++i) x[i] = rhs[i];
  int it;
  for (it = 1; it <= maxit; ++it) {
    int64_t changed = 0;
    for (int64_t i = 0; i < nb; ++i) {
      double d = coeff[i * 13];
      if (d == 0.0) { y[i] = rhs[i]; continue; }
      double s = rhs[i];
      for (int p = 0; p < nnb[i]; ++p) {
        double c = coeff[i * 13 + p + 1];
        if (c != 0.0) s -= c * x[ngb[i * MAXNB + p]];
      }
      y[i] = s / d;
    }
    for (int64_t i = 0; i < nb; ++i) {
      if (y[i] != x[i]) changed++;
      x[i] = y[i];
    }
    if (changed == 0) break;
  }
  free(y);
  return it;
}
