// Address/undefined-behaviour sanitizer driver for the host code (test
// infrastructure only: `make -C oracle sanitize`).  Builds a structured
// triangulation of a small grid (flat, and the same grid wrapped on a sphere
// for the curved branch), runs the FV geometry of the library's host code
// (gospl_fv_geometry), then every kernel of the C oracle and the
// reference-solver-mode solvers on random surfaces, and checks a few
// invariants so that the sanitizers see every code path run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <utility>
#include <vector>

extern "C" {
int gospl_fv_geometry(const double* coords, int64_t nb, const int32_t* cells_nodes, const int32_t* cells_edges,
                      int64_t n, const int32_t* edges_nodes, int64_t m, const double* circumcenter, int curved,
                      int32_t* ngbID, int32_t* nnb, double* eLgt, double* vDist, double* area);
void or_mfdreceivers(int nrcv, double fexp0, const double* elev, double sl, const int32_t* gid, int64_t nb,
                     const int32_t* ngb, const int32_t* nnb, const double* elgt, int32_t* rcv, double* dist,
                     double* wgt);
int or_fill(int64_t nb, const int32_t* ngb, const int32_t* nnb, const double* dem, const int8_t* outlet,
            double* out, int8_t* reached, int8_t* ledge_out);
void or_label_pits(double lvl, const double* fill, int64_t nb, const int32_t* ngb, const int32_t* nnb,
                   int32_t* label);
void or_fill_dir(const int32_t* spill, int64_t m, const int32_t* pitids, const double* h, int64_t n,
                 const int32_t* ngb, const int32_t* nnb, int32_t* ptdir);
int or_nghb_dir(const int32_t* pitids, const double* h, const int32_t* ptdir, int64_t n, const int32_t* ngb,
                const int32_t* nnb, int32_t* nptdir);
void or_getpitvol(const double* hlvl, const double* elev, const int32_t* pit, const int32_t* ids,
                  const double* area, int64_t m, int64_t npit, double* vol);
void or_sethillslopecoeff(int64_t nb, const int32_t* ngb, const int32_t* nnb, const double* area,
                          const double* elgt, const double* vdist, const double* kd, double* dcoeff);
int64_t or_flowacc(int64_t nb, int K, const int32_t* rcv, const double* wgt, const double* b, double* q,
                   const int32_t* ngb, const int32_t* nnb);
int64_t or_spl_solve(int64_t nb, int K, const int32_t* rcv, const double* off, const double* diag,
                     const double* rhs, double* x);
int or_hillslope_solve(int64_t nb, const int32_t* ngb, const int32_t* nnb, const double* coeff, const double* rhs,
                       double* x, int maxit);
int rs_flow_solve(int64_t n, int K, const int32_t* rcv, const double* wgt, const double* b, double* x, double rtol,
                  int maxit, int threads, double* relres);
int rs_spl_solve(int64_t n, int K, const int32_t* rcv, const double* off, const double* diag, const double* rhs,
                 double* x, double rtol, int maxit, int threads, double* relres);
int rs_hill_solve(int64_t n, const int32_t* ngb, const int32_t* nnb, const double* coeff, const double* rhs,
                  double* x, double rtol, int maxit, int threads, double* relres);
}

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      ++fails;                                                     \
    }                                                              \
  } while (0)

static void run(int nx, int ny, bool curved, unsigned seed) {
  const int64_t N = (int64_t)nx * ny;
  std::vector<double> xyz(3 * N);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int64_t v = (int64_t)j * nx + i;
      if (!curved) {
        xyz[3 * v] = 100.0 * i + 13.0 * ((i * 7 + j * 3) % 5) / 5.0;
        xyz[3 * v + 1] = 100.0 * j + 11.0 * ((i * 3 + j * 5) % 7) / 7.0;
        xyz[3 * v + 2] = 0.0;
      } else {  // a patch of the sphere
        const double lon = -0.3 + 0.6 * i / (nx - 1), lat = -0.3 + 0.6 * j / (ny - 1), R = 6.371e6;
        xyz[3 * v] = R * std::cos(lat) * std::cos(lon);
        xyz[3 * v + 1] = R * std::cos(lat) * std::sin(lon);
        xyz[3 * v + 2] = R * std::sin(lat);
      }
    }
  std::vector<int32_t> cells;
  for (int j = 0; j + 1 < ny; ++j)
    for (int i = 0; i + 1 < nx; ++i) {
      const int32_t a = j * nx + i, b = a + 1, c = a + nx, d = c + 1;
      cells.insert(cells.end(), {a, b, d, a, d, c});
    }
  const int64_t T = (int64_t)cells.size() / 3;
  std::map<std::pair<int, int>, int> eid;
  std::vector<int32_t> edges, cell_edges(3 * T);
  for (int64_t t = 0; t < T; ++t)
    for (int k = 0; k < 3; ++k) {
      int a = cells[3 * t + k], b = cells[3 * t + (k + 1) % 3];
      if (a > b) std::swap(a, b);
      auto it = eid.find({a, b});
      if (it == eid.end()) {
        it = eid.emplace(std::make_pair(a, b), (int)(edges.size() / 2)).first;
        edges.push_back(a);
        edges.push_back(b);
      }
      cell_edges[3 * t + k] = it->second;
    }
  std::vector<double> cc(3 * T);
  for (int64_t t = 0; t < T; ++t)
    for (int d = 0; d < 3; ++d)
      cc[3 * t + d] = (xyz[3 * cells[3 * t] + d] + xyz[3 * cells[3 * t + 1] + d] + xyz[3 * cells[3 * t + 2] + d]) / 3.0;
  std::vector<int32_t> ngb(12 * N), nnb(N);
  std::vector<double> elgt(12 * N), vdist(12 * N), area(N);
  CHECK(gospl_fv_geometry(xyz.data(), N, cells.data(), cell_edges.data(), T, edges.data(),
                          (int64_t)edges.size() / 2, cc.data(), curved ? 1 : 0, ngb.data(), nnb.data(), elgt.data(),
                          vdist.data(), area.data()) == 0);
  for (int64_t i = 0; i < N; ++i) {
    CHECK(nnb[i] >= 2 && nnb[i] <= 12);
    if (!(area[i] > 0.0)) area[i] = 1.0;
  }
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<double> h(N);
  std::vector<int8_t> outlet(N, 0);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int64_t v = (int64_t)j * nx + i;
      const bool border = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
      outlet[v] = border;
      h[v] = border ? 0.0 : 5.0 + 20.0 * U(rng) + 0.5 * (std::abs(i - nx / 2) + std::abs(j - ny / 2));
    }
  const int K = 3;
  std::vector<int32_t> gid(N), rcv(K * N);
  std::vector<double> dist(K * N), wgt(K * N);
  for (int64_t i = 0; i < N; ++i) gid[i] = (int32_t)i;
  or_mfdreceivers(K, 1.1, h.data(), -1e9, gid.data(), N, ngb.data(), nnb.data(), elgt.data(), rcv.data(), dist.data(),
                  wgt.data());
  std::vector<double> fill(N);
  std::vector<int8_t> reached(N), ledge(N);
  CHECK(or_fill(N, ngb.data(), nnb.data(), h.data(), outlet.data(), fill.data(), reached.data(), ledge.data()) >= 0);
  for (int64_t i = 0; i < N; ++i) CHECK(fill[i] >= h[i]);
  std::vector<int32_t> lab(N), pd(N), pd2(N);
  or_label_pits(-1.0, fill.data(), N, ngb.data(), nnb.data(), lab.data());
  std::vector<int32_t> spill;
  int32_t npit = 0;
  for (int64_t i = 0; i < N; ++i) {
    npit = std::max(npit, lab[i] + 1);
    if (outlet[i] && spill.size() < 4) spill.push_back((int32_t)i);
  }
  or_fill_dir(spill.data(), (int64_t)spill.size(), lab.data(), fill.data(), N, ngb.data(), nnb.data(), pd.data());
  (void)or_nghb_dir(lab.data(), fill.data(), pd.data(), N, ngb.data(), nnb.data(), pd2.data());
  std::vector<double> lvl(4 * (npit + 1), 0.0), vol(4 * (npit + 1));
  std::vector<int32_t> ones(N, 1);
  for (int p = 0; p <= npit; ++p)
    for (int k = 0; k < 4; ++k) lvl[4 * p + k] = 10.0 + k;
  or_getpitvol(lvl.data(), h.data(), lab.data(), ones.data(), area.data(), N, npit + 1, vol.data());
  std::vector<double> kd(N, 50.0), coeff(13 * N);
  or_sethillslopecoeff(N, ngb.data(), nnb.data(), area.data(), elgt.data(), vdist.data(), kd.data(), coeff.data());
  std::vector<double> b(N), q(N), q2(N, 0.0), x(N), x2(N);
  for (int64_t i = 0; i < N; ++i) b[i] = area[i];
  CHECK(or_flowacc(N, K, rcv.data(), wgt.data(), b.data(), q.data(), ngb.data(), nnb.data()) == 0);
  double rr = 0.0;
  CHECK(rs_flow_solve(N, K, rcv.data(), wgt.data(), b.data(), q2.data(), 1e-10, 5000, 2, &rr) >= 0);
  double e = 0.0, qm = 0.0;
  for (int64_t i = 0; i < N; ++i) {
    e = std::max(e, std::fabs(q[i] - q2[i]));
    qm = std::max(qm, std::fabs(q[i]));
  }
  CHECK(e <= 1e-6 * qm);
  std::vector<double> off(K * N), diag(N, 1.0);
  for (int64_t i = 0; i < N; ++i)
    for (int k = 0; k < K; ++k) {
      off[K * i + k] = rcv[K * i + k] != i ? -0.3 * wgt[K * i + k] : 0.0;
      diag[i] -= off[K * i + k];
    }
  CHECK(or_spl_solve(N, K, rcv.data(), off.data(), diag.data(), h.data(), x.data()) == 0);
  x2 = h;
  CHECK(rs_spl_solve(N, K, rcv.data(), off.data(), diag.data(), h.data(), x2.data(), 1e-10, 5000, 2, &rr) >= 0);
  CHECK(or_hillslope_solve(N, ngb.data(), nnb.data(), coeff.data(), h.data(), x.data(), 100000) <= 100000);
  x2 = h;
  CHECK(rs_hill_solve(N, ngb.data(), nnb.data(), coeff.data(), h.data(), x2.data(), 1e-10, 100000, 2, &rr) >= 0);
}

int main() {
  run(14, 11, false, 1);
  run(9, 17, true, 2);
  run(30, 30, false, 3);
  if (fails) {
    std::fprintf(stderr, "%d checks failed\n", fails);
    return 1;
  }
  std::printf("sanitizer driver: all checks passed\n");
  return 0;
}
