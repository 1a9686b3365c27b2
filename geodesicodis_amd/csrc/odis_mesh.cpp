// odis_mesh.cpp -- host assembly of the TRiSK C-grid operands of the LTE step.
//
// The floating-point recipes are the reference's (include/mathRoutines.h,
// src/mesh.cpp) so that the operators are the ones GeodesicODIS would build
// from the same grid: the face geometry of AssignFaces (src/mesh.cpp:524-666),
// the map-plane control-volume area (src/mesh.cpp:1828-1867), the TRiSK
// weights built from acos-form spherical triangle areas (src/mesh.cpp:677-910)
// and the G / Div / C coefficients (src/mesh.cpp:2808-3287).  Numbering and
// orientation follow the reference's first-visit order.  This file must be
// built with -ffp-contract=off so each product and quotient is rounded where
// it is written.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "odis_host.hpp"

namespace odis {

namespace {

// ---- spherical / stereographic primitives (include/mathRoutines.h) --------

// mapFactorAtPoint, include/mathRoutines.h:30-33
inline double map_factor(double lat1, double lat2, double lon1, double lon2) {
  return 2.0 / (1.0 + std::sin(lat1) * std::sin(lat2) +
                std::cos(lat1) * std::cos(lat2) * std::cos(lon2 - lon1));
}

struct MapPt {
  double m, x, y;
};
// mapAtPoint, include/mathRoutines.h:61-68; (lat0, lon0) is the projection centre
inline MapPt project(double lat0, double lon0, double lat, double lon, double r) {
  MapPt p;
  p.m = map_factor(lat0, lat, lon0, lon);
  p.x = p.m * r * (std::cos(lat) * std::sin(lon - lon0));
  p.y = p.m * r * (std::sin(lat) * std::cos(lat0) - std::cos(lat) * std::sin(lat0) * std::cos(lon - lon0));
  return p;
}

struct LL {
  double lat, lon;
};

// distanceBetweenSph (acos form), include/mathRoutines.h:133-137
inline double arc_acos(LL a, LL b, double r) {
  return r * std::acos(std::sin(a.lat) * std::sin(b.lat) +
                       std::cos(a.lat) * std::cos(b.lat) * std::cos(b.lon - a.lon));
}

// distanceBetweenSph2 (atan2 form on colatitude Cartesian), include/mathRoutines.h:139-160
inline double arc_atan2(LL a, LL b) {
  const double ca = kPi * 0.5 - a.lat, cb = kPi * 0.5 - b.lat;
  const double ax = 1.0 * std::sin(ca) * std::cos(a.lon), ay = 1.0 * std::sin(ca) * std::sin(a.lon),
               az = 1.0 * std::cos(ca);
  const double bx = 1.0 * std::sin(cb) * std::cos(b.lon), by = 1.0 * std::sin(cb) * std::sin(b.lon),
               bz = 1.0 * std::cos(cb);
  const double c0 = ay * bz - az * by, c1 = -(ax * bz - az * bx), c2 = ax * by - ay * bx;
  double cr = c0 * c0;
  cr += c1 * c1;
  cr += c2 * c2;
  cr = std::sqrt(cr);
  return std::atan2(cr, ax * bx + ay * by + az * bz);
}

struct XYZ {
  double x, y, z;
};
inline XYZ cart(LL a) {
  return {std::cos(a.lat) * std::cos(a.lon), std::cos(a.lat) * std::sin(a.lon), std::sin(a.lat)};
}

// midpointBetweenSph, include/mathRoutines.h:167-186
inline LL great_circle_mid(LL a, LL b) {
  const XYZ p = cart(a), q = cart(b);
  const double cx = 0.5 * (p.x + q.x), cy = 0.5 * (p.y + q.y), cz = 0.5 * (p.z + q.z);
  return {std::atan2(cz, std::sqrt(cx * cx + cy * cy)), std::atan2(cy, cx)};
}

// normalVectorBetweenMap, include/mathRoutines.h:208-243: unit normal of the
// chord a->b in the stereographic plane centred on the chord midpoint, rotated
// +90 degrees from the tangent.
inline void edge_normal(LL a, LL b, double *nx, double *ny) {
  const LL c = great_circle_mid(a, b);
  const MapPt pa = project(c.lat, c.lon, a.lat, a.lon, 1.0);
  const MapPt pb = project(c.lat, c.lon, b.lat, b.lon, 1.0);
  double tx = pb.x - pa.x, ty = pb.y - pa.y;
  const double mag = std::sqrt(tx * tx + ty * ty);
  tx /= mag;
  ty /= mag;
  *nx = -ty;
  *ny = tx;
}

// normalVectorBetweenXYZ, include/mathRoutines.h:246-283: unit vector along
// cart(a) x cart(b), the Cartesian normal of the face with vertices a, b
// (sph2cart on the colatitude, as the reference converts).
inline void edge_normal_xyz(LL a, LL b, double *n) {
  const double ca = kPi * 0.5 - a.lat, cb = kPi * 0.5 - b.lat;
  const double ax = std::sin(ca) * std::cos(a.lon), ay = std::sin(ca) * std::sin(a.lon), az = std::cos(ca);
  const double bx = std::sin(cb) * std::cos(b.lon), by = std::sin(cb) * std::sin(b.lon), bz = std::cos(cb);
  double x = ay * bz - az * by, y = -(ax * bz - az * bx), z = ax * by - ay * bx;
  const double mag = std::sqrt(x * x + y * y + z * z);
  n[0] = x / mag;
  n[1] = y / mag;
  n[2] = z / mag;
}

// intersectPointSph, include/mathRoutines.h:303-370: crossing of the great
// circles (a, b) and (c, d), on the side given by (a x b) x (c x d).
inline LL crossing(LL a, LL b, LL c, LL d) {
  const XYZ A = cart(a), B = cart(b), C = cart(c), D = cart(d);
  XYZ n1 = {A.y * B.z - A.z * B.y, -(A.x * B.z - A.z * B.x), A.x * B.y - A.y * B.x};
  XYZ n2 = {C.y * D.z - C.z * D.y, -(C.x * D.z - C.z * D.x), C.x * D.y - C.y * D.x};
  const double m1 = std::sqrt(n1.x * n1.x + n1.y * n1.y + n1.z * n1.z);
  const double m2 = std::sqrt(n2.x * n2.x + n2.y * n2.y + n2.z * n2.z);
  n1 = {n1.x / m1, n1.y / m1, n1.z / m1};
  n2 = {n2.x / m2, n2.y / m2, n2.z / m2};
  XYZ p = {n1.y * n2.z - n2.y * n1.z, n2.x * n1.z - n1.x * n2.z, n1.x * n2.y - n2.x * n1.y};
  const double mp = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
  p = {p.x / mp, p.y / mp, p.z / mp};
  LL out{std::asin(p.z / 1.0), std::atan2(p.y, p.x)};
  if (out.lon < 0.0) out.lon += 2 * kPi;
  return out;
}

// angleBetweenMap, include/mathRoutines.h:372-400, in degrees and shifted into
// (-360, 1e-8] as AssignFaces does (src/mesh.cpp:747-749): larger = closer
// anticlockwise to `from`.
inline double ccw_key(LL from, LL to, LL centre) {
  const MapPt p = project(centre.lat, centre.lon, from.lat, from.lon, 1.0);
  const MapPt q = project(centre.lat, centre.lon, to.lat, to.lon, 1.0);
  double ang = std::atan2(p.y - 0.0, p.x - 0.0) - std::atan2(q.y - 0.0, q.x - 0.0);
  if (ang > kPi) ang -= 2 * kPi;
  else if (ang < -kPi) ang += 2 * kPi;
  double deg = ang * 180. / kPi;
  if (deg > 0.0 + 1e-8) deg -= 360.0;
  return deg;
}

// triangularArea (map plane), include/mathRoutines.h:403-408
inline double planar_tri(double xc, double yc, double x1, double x2, double y1, double y2) {
  return 0.5 * std::fabs(xc * (y1 - y2) + x1 * (y2 - yc) + x2 * (yc - y1));
}

// triangularAreaSph (spherical excess from acos side lengths and angles),
// include/mathRoutines.h:410-430
inline double sph_tri(LL p1, LL p2, LL p3, double r) {
  auto side = [](LL u, LL v) {
    return std::fabs(std::acos(std::sin(u.lat) * std::sin(v.lat) +
                               std::cos(u.lat) * std::cos(v.lat) * std::cos(std::fabs(v.lon - u.lon))));
  };
  const double c = side(p1, p2), a = side(p2, p3), b = side(p3, p1);
  const double A = std::fabs(std::acos((std::cos(a) - std::cos(b) * std::cos(c)) / (std::sin(b) * std::sin(c))));
  const double B = std::fabs(std::acos((std::cos(b) - std::cos(a) * std::cos(c)) / (std::sin(a) * std::sin(c))));
  const double C = std::fabs(std::acos((std::cos(c) - std::cos(b) * std::cos(a)) / (std::sin(b) * std::sin(a))));
  return (r * r) * std::fabs((A + B + C) - kPi);
}

struct Keyed {
  int id;
  double key;
};
// stable descending order on <= 6 items
inline void sort_ccw(Keyed *a, int n) {
  std::stable_sort(a, a + n, [](const Keyed &l, const Keyed &r) { return l.key > r.key; });
}

}  // namespace

MeshOps build_mesh(const Grid &g, double r) {
  MeshOps M;
  const int N = g.n;
  const int E = 3 * N - 6;
  M.N = N;
  M.E = E;
  M.radius = r;
  M.lat = g.lat;
  M.lon = g.lon;
  auto node = [&](int i) { return LL{g.lat[i], g.lon[i]}; };
  auto corner = [&](int i, int j) { return LL{g.clat[i * 6 + j], g.clon[i * 6 + j]}; };
  std::vector<int> deg(N);
  for (int i = 0; i < N; i++) deg[i] = g.friends[i * 6 + 5] < 0 ? 5 : 6;

  // -- control-volume area in the stereographic plane of each node -----------
  M.cv_area.assign(N, 0.0);
  for (int i = 0; i < N; i++) {
    const MapPt c = project(g.lat[i], g.lon[i], g.lat[i], g.lon[i], r);
    MapPt v[6];
    for (int j = 0; j < deg[i]; j++) v[j] = project(g.lat[i], g.lon[i], g.clat[i * 6 + j], g.clon[i * 6 + j], r);
    double a = 0.0;
    for (int j = 0; j < deg[i]; j++) {
      const MapPt &p = v[j], &q = v[(j + 1) % deg[i]];
      a += planar_tri(c.x, c.y, p.x, q.x, p.y, q.y);
    }
    M.cv_area[i] = a;
  }

  // -- Voronoi vertex ids: first listing wins, others matched within 1e-9 rad
  //    (CalcControlVolumeVertexR, src/mesh.cpp:436-506)
  std::vector<int> vert(size_t(N) * 6, -1);
  for (int i = 0; i < N; i++) {
    for (int j = 0; j < deg[i]; j++) {
      if (vert[i * 6 + j] >= 0) continue;
      const int vid = int(M.vlat.size());
      const LL v = corner(i, j);
      M.vlat.push_back(v.lat);
      M.vlon.push_back(v.lon);
      vert[i * 6 + j] = vid;
      const int owners[2] = {g.friends[i * 6 + j], g.friends[i * 6 + (j + 1) % deg[i]]};
      for (int f : owners)
        for (int k = 0; k < deg[f]; k++) {
          const LL w = corner(f, k);
          if (w.lat < v.lat + 1e-9 && w.lat > v.lat - 1e-9 && w.lon < v.lon + 1e-9 && w.lon > v.lon - 1e-9)
            vert[f * 6 + k] = vid;
        }
    }
  }
  M.V = int(M.vlat.size());
  if (M.V != 2 * N - 4)
    throw std::runtime_error("grid vertices do not match between neighbours (" + std::to_string(M.V) +
                             " vertices, expected " + std::to_string(2 * N - 4) + ")");

  // -- faces in first-visit order with their geometry (src/mesh.cpp:524-666) --
  M.faces.assign(size_t(N) * 6, -1);
  M.node_face_dir.assign(size_t(N) * 6, 0);
  M.face_nodes.assign(size_t(E) * 2, -1);
  M.face_len.assign(E, 0.0);
  M.face_node_dist.assign(E, 0.0);
  M.face_lat.assign(E, 0.0);
  M.face_lon.assign(E, 0.0);
  M.face_m.assign(size_t(E) * 2, 0.0);
  M.face_normal.assign(size_t(E) * 2, 0.0);
  M.face_normal_xyz.assign(size_t(E) * 3, 0.0);
  M.face_intercept.assign(size_t(E) * 2, 0.0);
  int e = 0;
  for (int i = 0; i < N; i++) {
    for (int j = 0; j < deg[i]; j++) {
      if (M.faces[i * 6 + j] >= 0) continue;
      if (e >= E) throw std::runtime_error("grid has more faces than 3N-6");
      const int f = g.friends[i * 6 + j];
      const LL a = corner(i, (j + deg[i] - 1) % deg[i]), b = corner(i, j);
      M.faces[i * 6 + j] = e;
      M.node_face_dir[i * 6 + j] = 1;
      M.face_nodes[e * 2] = i;
      M.face_len[e] = std::fabs(arc_acos(a, b, r));
      LL c = great_circle_mid(a, b);
      if (c.lon < 0.0) c.lon += 2 * kPi;
      M.face_lat[e] = c.lat;
      M.face_lon[e] = c.lon;
      M.face_m[e * 2] = map_factor(c.lat, g.lat[i], c.lon, g.lon[i]);
      M.face_m[e * 2 + 1] = map_factor(c.lat, g.lat[f], c.lon, g.lon[f]);
      M.face_node_dist[e] = arc_atan2(node(i), node(f)) * r;
      edge_normal(a, b, &M.face_normal[e * 2], &M.face_normal[e * 2 + 1]);
      edge_normal_xyz(a, b, &M.face_normal_xyz[e * 3]);
      const LL x = crossing(a, b, node(i), node(f));
      M.face_intercept[e * 2] = x.lat;
      M.face_intercept[e * 2 + 1] = x.lon;
      for (int k = 0; k < deg[f]; k++)
        if (g.friends[f * 6 + k] == i) {
          M.faces[f * 6 + k] = e;
          M.face_nodes[e * 2 + 1] = f;
          M.node_face_dir[f * 6 + k] = -1;
        }
      e++;
    }
  }
  if (e != E) throw std::runtime_error("grid does not have 3N-6 faces");

  // -- TRiSK weights (src/mesh.cpp:677-910) -----------------------------------
  M.trisk_count.assign(E, 0);
  M.trisk_faces.assign(size_t(E) * 10, -1);
  M.trisk_w.assign(size_t(E) * 10, 0.0);
  auto intercept = [&](int face) { return LL{M.face_intercept[face * 2], M.face_intercept[face * 2 + 1]}; };
  for (int fe = 0; fe < E; fe++) {
    int slot = 0;
    const LL xe = intercept(fe);
    for (int side = 0; side < 2; side++) {
      const int c = M.face_nodes[fe * 2 + side];
      const int k = deg[c];
      const LL pc = node(c);
      Keyed fl[6], vl[6];
      for (int j = 0; j < k; j++) {
        const int id = M.faces[c * 6 + j];
        fl[j] = {id, ccw_key(xe, intercept(id), pc)};
        const int vid = vert[c * 6 + j];
        vl[j] = {vid, ccw_key(xe, LL{M.vlat[vid], M.vlon[vid]}, pc)};
      }
      sort_ccw(fl, k);
      sort_ccw(vl, k);
      double area = 0.0;
      for (int j = 0; j < k; j++) {
        const int v1 = vl[j].id, v2 = vl[(j + 1) % k].id;
        area += sph_tri(pc, LL{M.vlat[v1], M.vlon[v1]}, LL{M.vlat[v2], M.vlon[v2]}, r);
      }
      double Rk[6];
      for (int j = 0; j < k; j++) {
        const LL pv{M.vlat[vl[j].id], M.vlon[vl[j].id]};
        const double a1 = sph_tri(pc, pv, intercept(fl[j].id), r);
        const double a2 = sph_tri(pc, pv, intercept(fl[(j + 1) % k].id), r);
        Rk[j] = (a1 + a2) / area;
      }
      const int t_ev = M.face_nodes[fe * 2] == c ? 1 : -1;
      for (int j = 0; j < k; j++) {
        if (fl[j].id == fe) continue;
        if (slot >= 10) throw std::runtime_error("TRiSK stencil exceeds 10 faces");
        double w = 0.0;
        for (int q = 0; q < j; q++) w += Rk[q];
        w -= 0.5;
        for (int q = 0; q < k; q++)
          if (fl[j].id == M.faces[c * 6 + q]) w *= M.node_face_dir[c * 6 + q];
        w *= t_ev;
        M.trisk_faces[fe * 10 + slot] = fl[j].id;
        M.trisk_w[fe * 10 + slot] = w;
        slot++;
      }
    }
    M.trisk_count[fe] = slot;
  }

  // -- derived operator coefficients ------------------------------------------
  M.face_area.resize(E);
  M.grad_coef.resize(size_t(E) * 2);
  M.trisk_coef.assign(size_t(E) * 10, 0.0);
  for (int fe = 0; fe < E; fe++) {
    const double d = M.face_node_dist[fe];
    M.face_area[fe] = d * M.face_len[fe];                 // src/mesh.cpp:1093
    M.grad_coef[fe * 2] = (-M.face_m[fe * 2]) / d;         // src/mesh.cpp:3110-3114
    M.grad_coef[fe * 2 + 1] = (M.face_m[fe * 2 + 1]) / d;
    for (int j = 0; j < M.trisk_count[fe]; j++) {
      const int f = M.trisk_faces[fe * 10 + j];
      M.trisk_coef[fe * 10 + j] = M.trisk_w[fe * 10 + j] * M.face_len[f] / d;
    }
  }
  M.div_coef.assign(size_t(N) * 6, 0.0);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < deg[i]; j++) {
      const int f = M.faces[i * 6 + j];
      M.div_coef[i * 6 + j] = -M.node_face_dir[i * 6 + j] * M.face_len[f] / M.cv_area[i];  // src/mesh.cpp:3242-3247
    }
  build_rbf_operator(M);
  return M;
}

// RBF reconstruction of the Cartesian velocity at cell centres from the
// face-normal velocities (operatorRBFinterp = RBF Vinv node2faceAdj,
// src/mesh.cpp:2263-2361, with the per-cell matrices of src/mesh.cpp:360-430
// and the adjacency of src/mesh.cpp:2725-2763).  For cell i with faces f_j
// (j in friend order) and Cartesian face normals n_j:
//     A_i[j1][j2] = phi(arc(f_j1, f_j2)) n_j1 . n_j2,
//     row 3i + c, column f_j:  sum_m phi(arc(i, f_m)) n_m[c] (A_i^-1)[m][j],
// with phi(d) = exp(-d^2) (shape parameter 1, phi = 1 below 1e-8 rad,
// include/mathRoutines.h:432-439) and arcs between face centres and cell
// nodes on the unit sphere.  A_i^-1 by LU with partial pivoting (the
// reference's dynamic-size inverse).  Columns ascend within a row.
void build_rbf_operator(MeshOps &M) {
  const int N = M.N;
  M.rbf_row_ptr.assign(size_t(3) * N + 1, 0);
  M.rbf_col.clear();
  M.rbf_val.clear();
  auto rbf = [](double d) { return d < 1e-8 ? 1.0 : std::exp(-(d * d)); };
  // colatitude Cartesian points, as distanceBetweenSph2 converts them
  auto colat_xyz = [](double lat, double lon, double *p) {
    const double c = kPi * 0.5 - lat;
    p[0] = 1.0 * std::sin(c) * std::cos(lon);
    p[1] = 1.0 * std::sin(c) * std::sin(lon);
    p[2] = 1.0 * std::cos(c);
  };
  auto arc = [](const double *a, const double *b) {
    const double c0 = a[1] * b[2] - a[2] * b[1], c1 = -(a[0] * b[2] - a[2] * b[0]), c2 = a[0] * b[1] - a[1] * b[0];
    double cr = c0 * c0;
    cr += c1 * c1;
    cr += c2 * c2;
    cr = std::sqrt(cr);
    return std::atan2(cr, a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
  };
  std::vector<double> fc(size_t(M.E) * 3);
  for (int e = 0; e < M.E; e++) colat_xyz(M.face_lat[e], M.face_lon[e], &fc[size_t(e) * 3]);
  for (int i = 0; i < N; i++) {
    int k = 0;
    while (k < 6 && M.faces[i * 6 + k] >= 0) k++;
    int f[6];
    double n[6][3], A[6][6], Ai[6][6], phin[6], pi[3];
    colat_xyz(M.lat[i], M.lon[i], pi);
    for (int j = 0; j < k; j++) {
      f[j] = M.faces[i * 6 + j];
      for (int c = 0; c < 3; c++) n[j][c] = M.face_normal_xyz[size_t(f[j]) * 3 + c];
    }
    for (int j1 = 0; j1 < k; j1++) {
      const double *a = &fc[size_t(f[j1]) * 3];
      for (int j2 = 0; j2 < k; j2++) {
        const double *b = &fc[size_t(f[j2]) * 3];
        A[j1][j2] = rbf(arc(a, b)) * (n[j1][0] * n[j2][0] + n[j1][1] * n[j2][1] + n[j1][2] * n[j2][2]);
      }
      phin[j1] = rbf(arc(pi, a));
    }
    // inverse: LU with row pivoting, then one solve per unit column
    int perm[6];
    for (int j = 0; j < k; j++) perm[j] = j;
    for (int c = 0; c < k; c++) {
      int piv = c;
      for (int r = c + 1; r < k; r++)
        if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
      if (A[piv][c] == 0.0) throw std::runtime_error("RBF matrix of cell " + std::to_string(i) + " is singular");
      if (piv != c) {
        for (int q = 0; q < k; q++) std::swap(A[c][q], A[piv][q]);
        std::swap(perm[c], perm[piv]);
      }
      for (int r = c + 1; r < k; r++) {
        A[r][c] /= A[c][c];
        for (int q = c + 1; q < k; q++) A[r][q] -= A[r][c] * A[c][q];
      }
    }
    for (int col = 0; col < k; col++) {
      double y[6];
      for (int r = 0; r < k; r++) {
        double t = perm[r] == col ? 1.0 : 0.0;
        for (int q = 0; q < r; q++) t -= A[r][q] * y[q];
        y[r] = t;
      }
      for (int r = k - 1; r >= 0; r--) {
        double t = y[r];
        for (int q = r + 1; q < k; q++) t -= A[r][q] * Ai[q][col];
        Ai[r][col] = t / A[r][r];
      }
    }
    int order[6];
    for (int j = 0; j < k; j++) order[j] = j;
    std::sort(order, order + k, [&](int x, int y) { return f[x] < f[y]; });
    for (int c = 0; c < 3; c++) {
      for (int o = 0; o < k; o++) {
        const int j = order[o];
        double w = 0.0;
        for (int m = 0; m < k; m++) w += phin[m] * n[m][c] * Ai[m][j];
        M.rbf_col.push_back(f[j]);
        M.rbf_val.push_back(w);
      }
      M.rbf_row_ptr[size_t(3) * i + c + 1] = int(M.rbf_col.size());
    }
  }
}

}  // namespace odis
