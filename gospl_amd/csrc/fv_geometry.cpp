// Init-time finite-volume geometry of a triangulated surface (host C++).
//
// Restates goSPL's `definetin` (fortran/functions.F90:2715-2940) and the
// spherical helpers it calls (`spherical2lonlat` :316-335,
// `great_circle_distance` :337-366, `distance` :368-393, `slerpmidpt` :425-452,
// `striarea` :538-577, `gtriarea` :579-616) so that the neighbour order, edge
// lengths, Voronoi face lengths and Voronoi areas handed to the kernels are the
// ones goSPL would build for the same triangulation.  Runs once per mesh; the
// result is uploaded to the GPU by the Python host.
//
// Floating point is evaluated in the same operation order as the Fortran and
// without FMA contraction (the library is built with -ffp-contract=off).
// Fortran's NORM2 is restated as the scaled accumulation the ROCm flang
// runtime uses (largest magnitude first-seen scaling; sqrt(1+sum)).  Where the
// Fortran takes both the sine and cosine of one angle the compiler emits a
// single libm sincos call, whose results can differ from separate sin/cos in
// the last bit; the same pairs call sincos here.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kMaxNb = 12;                          // functions.F90:2733 (ngbID(nb,12))
constexpr double kPi = 3.1415926535897932;          // functions.F90:323

// Fortran NORM2 of a 3-vector as evaluated by the flang runtime.
inline double fnorm2(const double v[3]) {
  double mx = 0.0, s = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double a = std::fabs(v[i]);
    if (mx == 0.0) {
      mx = a;
      s = 0.0;
    } else if (a > mx) {
      const double t = mx / a;
      const double tsq = t * t;
      s = s * tsq + tsq;
      mx = a;
    } else {
      const double t = a / mx;
      s = s + t * t;
    }
  }
  return mx * std::sqrt(1.0 + s);
}

inline void lonlat(const double p[3], double* lon, double* lat) {
  const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  const double r2d = 180.0 / kPi;
  double la = std::asin(p[2] / r) * r2d;
  double lo = std::atan2(p[1], p[0]) * r2d;
  *lat = la * kPi / 180.0;
  *lon = lo * kPi / 180.0;
}

inline double great_circle(double lon1, double lat1, double lon2, double lat2,
                           double radius) {
  double c1, s1, c2, s2, clon, slon;
  ::sincos(lat1, &s1, &c1);
  ::sincos(lat2, &s2, &c2);
  const double dlon = lon1 - lon2;
  ::sincos(dlon, &slon, &clon);
  const double a = c2 * slon;
  const double b = c1 * s2 - s1 * c2 * clon;
  return radius * std::atan2(std::sqrt(a * a + b * b), s1 * s2 + c1 * c2 * clon);
}

// goSPL `distance`: euclidean when both points have z == 0, else great circle
// on the sphere through p1.
inline double fdistance(const double p1[3], const double p2[3]) {
  if (p2[2] == 0.0 && p1[2] == 0.0) {
    const double v[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    return fnorm2(v);
  }
  const double radius = fnorm2(p1);
  double lo1, la1, lo2, la2;
  lonlat(p1, &lo1, &la1);
  lonlat(p2, &lo2, &la2);
  return great_circle(lo1, la1, lo2, la2, radius);
}

inline void slerp_mid(const double p1[3], const double p2[3], double out[3]) {
  const double n1 = fnorm2(p1), n2 = fnorm2(p2);
  double v1[3], v2[3], pt[3];
  for (int i = 0; i < 3; ++i) { v1[i] = p1[i] / n1; v2[i] = p2[i] / n2; }
  double dot = v1[0] * v2[0];
  dot = dot + v1[1] * v2[1];
  dot = dot + v1[2] * v2[2];
  const double a = std::acos(dot);
  const double sh = std::sin(0.5 * a), sa = std::sin(a);
  for (int i = 0; i < 3; ++i) pt[i] = sh * v1[i] / sa + sh * v2[i] / sa;
  const double npt = fnorm2(pt);
  for (int i = 0; i < 3; ++i) out[i] = n1 * pt[i] / npt;
}

inline double sph_tri_area(const double v1[3], const double v2[3], const double v3[3]) {
  const double r = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  double lo1, la1, lo2, la2, lo3, la3;
  lonlat(v1, &lo1, &la1);
  lonlat(v2, &lo2, &la2);
  lonlat(v3, &lo3, &la3);
  double sl1, cl1, sl2, cl2, sl3, cl3;
  ::sincos(la1, &sl1, &cl1);
  ::sincos(la2, &sl2, &cl2);
  ::sincos(la3, &sl3, &cl3);
  const double a = std::acos(sl2 * sl3 + cl2 * cl3 * std::cos(lo3 - lo2));
  const double b = std::acos(sl1 * sl3 + cl1 * cl3 * std::cos(lo3 - lo1));
  const double c = std::acos(sl1 * sl2 + cl1 * cl2 * std::cos(lo2 - lo1));
  double sa, ca, sb, cb, sc, ccs;
  ::sincos(a, &sa, &ca);
  ::sincos(b, &sb, &cb);
  ::sincos(c, &sc, &ccs);
  const double alpha = std::acos((ca - cb * ccs) / (sb * sc));
  const double beta = std::acos((cb - ca * ccs) / (sa * sc));
  const double gamma = std::acos((ccs - ca * cb) / (sa * sb));
  const double E = alpha + beta + gamma - kPi;
  return E * (r * r);
}

inline double heron_area(const double p1[3], const double p2[3], const double p3[3]) {
  const double x1 = p1[0], x2 = p2[0], x3 = p3[0];
  const double y1 = p1[1], y2 = p2[1], y3 = p3[1];
  if ((x1 == x2 && y1 == y2) || (x1 == x3 && y1 == y3) || (x2 == x3 && y2 == y3)) return 0.0;
  const double l1 = std::sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
  const double l2 = std::sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3));
  const double l3 = std::sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1));
  const double p = (l1 + l2 + l3) / 2;
  return std::sqrt(p * (p - l1) * (p - l2) * (p - l3));
}

}  // namespace

extern "C" {

// See include/gospl_hip.h for the contract.
int gospl_fv_geometry(const double* coords, int64_t nb, const int32_t* cells_nodes,
                      const int32_t* cells_edges, int64_t n, const int32_t* edges_nodes,
                      int64_t m, const double* circumcenter, int curved, int32_t* ngbID,
                      int32_t* nnb, double* eLgt, double* vDist, double* area) {
  if (nb <= 0 || n <= 0 || m <= 0) return 1;
  std::vector<int32_t> cell_ids(static_cast<size_t>(nb) * kMaxNb, -1);
  std::vector<int32_t> edge(static_cast<size_t>(nb) * kMaxNb, -1);
  for (int64_t i = 0; i < nb * kMaxNb; ++i) {
    ngbID[i] = -1;
    eLgt[i] = 0.0;
    vDist[i] = 0.0;
  }
  for (int64_t i = 0; i < nb; ++i) { nnb[i] = 0; area[i] = 0.0; }

  double radius = 0.0;
  if (curved == 1) radius = fnorm2(coords);

  // Cells around each vertex, in cell order (at most 12 kept).
  for (int64_t i = 0; i < n; ++i) {
    for (int p = 0; p < 3; ++p) {
      const int32_t v = cells_nodes[3 * i + p];
      int32_t* row = &cell_ids[static_cast<size_t>(v) * kMaxNb];
      for (int k = 0; k < kMaxNb; ++k) {
        if (row[k] == static_cast<int32_t>(i)) break;
        if (row[k] == -1) { row[k] = static_cast<int32_t>(i); break; }
      }
    }
  }
  // Edges around each vertex, in edge order (at most 12 kept).
  for (int64_t i = 0; i < m; ++i) {
    for (int s = 0; s < 2; ++s) {
      const int32_t v = edges_nodes[2 * i + s];
      int32_t* row = &edge[static_cast<size_t>(v) * kMaxNb];
      for (int k = 0; k < kMaxNb; ++k) {
        if (row[k] == static_cast<int32_t>(i)) break;
        if (row[k] == -1) { row[k] = static_cast<int32_t>(i); nnb[v] += 1; break; }
      }
    }
  }

  for (int64_t k = 0; k < nb; ++k) {
    const double* c0 = coords + 3 * k;
    const int nbngbs = nnb[k];
    int32_t* ngb = ngbID + k * kMaxNb;
    for (int e = 0; e < nbngbs; ++e) {
      const int32_t eid = edge[static_cast<size_t>(k) * kMaxNb + e];
      const int32_t a = edges_nodes[2 * eid], b = edges_nodes[2 * eid + 1];
      const int32_t other = (a == k) ? b : a;
      ngb[e] = other;
      eLgt[k * kMaxNb + e] = fdistance(c0, coords + 3 * other);
    }
    for (int cid = 0; cid < kMaxNb; ++cid) {
      const int32_t cell = cell_ids[static_cast<size_t>(k) * kMaxNb + cid];
      if (cell == -1) break;
      for (int e = 0; e < 3; ++e) {
        const int32_t eg = cells_edges[3 * cell + e];
        const int32_t e1 = edges_nodes[2 * eg], e2 = edges_nodes[2 * eg + 1];
        if (!(e1 == k || e2 == k)) continue;
        double mid[3];
        if (radius > 0.0) {
          slerp_mid(coords + 3 * e1, coords + 3 * e2, mid);
        } else {
          for (int d = 0; d < 3; ++d) mid[d] = 0.5 * (coords[3 * e1 + d] + coords[3 * e2 + d]);
        }
        const int32_t want = (e1 == k) ? e2 : e1;
        int id = -1;
        for (int q = 0; q < nbngbs; ++q) {
          if (ngb[q] == want) { id = q; break; }
        }
        const double* cc = circumcenter + 3 * cell;
        const double dist = fdistance(mid, cc);
        if (id >= 0) vDist[k * kMaxNb + id] += dist;
        const double tri = (radius == 0.0) ? heron_area(c0, mid, cc) : sph_tri_area(c0, mid, cc);
        area[k] += tri;
      }
    }
  }
  return 0;
}

}  // extern "C"
