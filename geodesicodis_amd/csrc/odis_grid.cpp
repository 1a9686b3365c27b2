// odis_grid.cpp -- grid ingest, writer and generator.
//
// Ingest follows Mesh::ReadMeshFile (src/mesh.cpp:4016-4101): one header
// line, then per node "id lat lon {f0,..,f5}, {(lat,lon), ... x6}" in degrees,
// converted with atof(str) * pi/180 exactly as the reference does.
//
// The reference ships grids only up to 10242 cells (grid_l6.txt); the
// 40962- and 163842-cell grids are not in the repository.  generate_grid
// builds them: icosahedron, recursive edge bisection onto the sphere, then
// fixed-topology Lloyd sweeps towards a spherical centroidal Voronoi
// tessellation (the property of the shipped grids: each node sits at its
// cell's centroid, vertices are triangle circumcentres).  Nodes are numbered
// pentagons first (0..11, as in the shipped files) and then along a 3-D
// Hilbert curve, so that neighbouring cells get nearby indices.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "odis_host.hpp"

namespace odis {

namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 unit(V3 a) { return (1.0 / std::sqrt(dot(a, a))) * a; }

V3 from_latlon(double lat, double lon) {
  return {std::cos(lat) * std::cos(lon), std::cos(lat) * std::sin(lon), std::sin(lat)};
}
void to_latlon(V3 p, double *lat, double *lon) {
  *lat = std::atan2(p.z, std::sqrt(p.x * p.x + p.y * p.y));
  double l = std::atan2(p.y, p.x);
  if (l < 0.0) l += 2.0 * kPi;
  if (l >= 2.0 * kPi) l -= 2.0 * kPi;
  *lon = l;
}

// Circumcentre of a spherical triangle (unit sphere), on the triangle's side.
V3 circumcentre(V3 a, V3 b, V3 c) {
  V3 n = unit(cross(b - a, c - a));
  if (dot(n, a + b + c) < 0.0) n = -1.0 * n;
  return n;
}

// 3-D Hilbert key (Skilling, "Programming the Hilbert curve", AIP 2004).
uint64_t hilbert3(uint32_t X0, uint32_t X1, uint32_t X2, int bits) {
  uint32_t X[3] = {X0, X1, X2};
  uint32_t M = 1u << (bits - 1), P, Q, t;
  for (Q = M; Q > 1; Q >>= 1) {
    P = Q - 1;
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
  t = 0;
  for (Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; i++) X[i] ^= t;
  uint64_t key = 0;
  for (int b = bits - 1; b >= 0; b--)
    for (int i = 0; i < 3; i++) key = (key << 1) | ((X[i] >> b) & 1u);
  return key;
}

uint64_t hilbert_key(V3 p) {
  const int bits = 20;
  const double s = double((1u << bits) - 1);
  auto q = [&](double v) { return uint32_t(std::lround((v + 1.0) * 0.5 * s)); };
  return hilbert3(q(p.x), q(p.y), q(p.z), bits);
}

// Order the neighbours of node p clockwise as seen from outside the sphere.
void order_clockwise(V3 p, const std::vector<V3> &pts, std::vector<int> &nb) {
  V3 ref = std::fabs(p.z) < 0.9 ? V3{0, 0, 1} : V3{1, 0, 0};
  V3 t1 = unit(cross(ref, p));
  V3 t2 = cross(p, t1);  // (t1, t2, p) right-handed: angle increases anticlockwise
  std::vector<std::pair<double, int>> a;
  for (int q : nb) {
    V3 d = pts[q] - p;
    a.push_back({std::atan2(dot(d, t2), dot(d, t1)), q});
  }
  std::sort(a.begin(), a.end(), [](const auto &l, const auto &r) { return l.first > r.first; });
  for (size_t k = 0; k < nb.size(); k++) nb[k] = a[k].second;
}

}  // namespace

// ---------------------------------------------------------------------------
// reader / writer
// ---------------------------------------------------------------------------

Grid read_grid_file(const std::string &path) {
  std::ifstream in(path);
  if (!in.is_open()) throw std::runtime_error("ERROR: GRID FILE NOT FOUND AT " + path);
  std::string line;
  std::getline(in, line);  // header
  struct Row {
    double lat, lon;
    int fr[6];
    double cl[6][2];
  };
  std::map<int, Row> rows;
  while (std::getline(in, line)) {
    if (line.find('{') == std::string::npos) continue;
    const char *s = line.c_str();
    char *end = nullptr;
    Row r{};
    long id = std::strtol(s, &end, 10);
    s = end;
    // lat / lon tokens: atof of the whitespace-delimited token, as the reference
    r.lat = std::strtod(s, &end) * kRadConv;
    s = end;
    r.lon = std::strtod(s, &end) * kRadConv;
    s = std::strchr(end, '{');
    if (!s) throw std::runtime_error("malformed grid line: " + line);
    s++;
    for (int i = 0; i < 6; i++) {
      r.fr[i] = int(std::strtol(s, &end, 10));
      s = end;
      while (*s == ',' || *s == ' ' || *s == '}') s++;
    }
    s = std::strchr(s, '{');
    if (!s) throw std::runtime_error("malformed grid line: " + line);
    for (int i = 0; i < 6; i++) {
      s = std::strchr(s, '(');
      if (!s) throw std::runtime_error("malformed grid line: " + line);
      s++;
      r.cl[i][0] = std::strtod(s, &end) * kRadConv;
      s = std::strchr(end, ',') + 1;
      r.cl[i][1] = std::strtod(s, &end) * kRadConv;
      s = end;
    }
    rows[int(id)] = r;
  }
  Grid g;
  g.n = int(rows.size());
  g.lat.resize(g.n);
  g.lon.resize(g.n);
  g.friends.resize(size_t(g.n) * 6);
  g.clat.resize(size_t(g.n) * 6);
  g.clon.resize(size_t(g.n) * 6);
  for (auto &kv : rows) {
    int i = kv.first;
    if (i < 0 || i >= g.n) throw std::runtime_error("grid node id out of range");
    g.lat[i] = kv.second.lat;
    g.lon[i] = kv.second.lon;
    for (int j = 0; j < 6; j++) {
      g.friends[i * 6 + j] = kv.second.fr[j];
      g.clat[i * 6 + j] = kv.second.cl[j][0];
      g.clon[i * 6 + j] = kv.second.cl[j][1];
    }
  }
  return g;
}

void write_grid_file(const Grid &g, const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fprintf(f, "ID    NODE LAT     NODE LON     FRIENDS LIST                           "
                  "CENTROID COORD LIST \n");
  for (int i = 0; i < g.n; i++) {
    std::fprintf(f, "%-5d %.16f %.16f {", i, g.lat[i] / kRadConv, g.lon[i] / kRadConv);
    for (int j = 0; j < 6; j++) std::fprintf(f, "%5d%s", g.friends[i * 6 + j], j < 5 ? "," : "");
    std::fprintf(f, "}, {");
    for (int j = 0; j < 6; j++) {
      double a = g.clat[i * 6 + j], b = g.clon[i * 6 + j];
      if (g.friends[i * 6 + j] < 0) {
        a = -kRadConv;
        b = -kRadConv;
      }
      std::fprintf(f, "(%.16f, %.16f)%s", a / kRadConv, b / kRadConv, j < 5 ? ", " : "");
    }
    std::fprintf(f, "} \n");
  }
  std::fclose(f);
}

// ---------------------------------------------------------------------------
// generator
// ---------------------------------------------------------------------------

Grid generate_grid(int level, int lloyd_iters) {
  if (level < 1 || level > 11) throw std::runtime_error("grid level must be in 1..11");
  std::vector<V3> P;
  std::vector<std::array<int, 3>> T;
  const double lat_ring = std::atan(0.5);
  P.push_back({0, 0, 1});
  P.push_back({0, 0, -1});
  for (int k = 0; k < 5; k++) P.push_back(from_latlon(lat_ring, (216.0 + 72.0 * k) * kRadConv));
  for (int k = 0; k < 5; k++)
    P.push_back(from_latlon(-lat_ring, (216.0 + 36.0 + 72.0 * k) * kRadConv));
  for (int k = 0; k < 5; k++) {
    int u0 = 2 + k, u1 = 2 + (k + 1) % 5, l0 = 7 + k, l1 = 7 + (k + 1) % 5;
    T.push_back({0, u0, u1});
    T.push_back({u0, l0, u1});
    T.push_back({u1, l0, l1});
    T.push_back({1, l1, l0});
  }
  for (int s = 1; s < level; s++) {
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int a, int b) {
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      P.push_back(unit(P[a] + P[b]));
      int id = int(P.size()) - 1;
      mid[key] = id;
      return id;
    };
    std::vector<std::array<int, 3>> T2;
    T2.reserve(T.size() * 4);
    for (auto &t : T) {
      int ab = midpoint(t[0], t[1]), bc = midpoint(t[1], t[2]), ca = midpoint(t[2], t[0]);
      T2.push_back({t[0], ab, ca});
      T2.push_back({ab, t[1], bc});
      T2.push_back({ca, bc, t[2]});
      T2.push_back({ab, bc, ca});
    }
    T.swap(T2);
  }
  const int n = int(P.size());
  // adjacency
  std::vector<std::vector<int>> nb(n);
  for (auto &t : T)
    for (int k = 0; k < 3; k++) {
      int a = t[k], b = t[(k + 1) % 3];
      if (std::find(nb[a].begin(), nb[a].end(), b) == nb[a].end()) nb[a].push_back(b);
      if (std::find(nb[b].begin(), nb[b].end(), a) == nb[b].end()) nb[b].push_back(a);
    }
  for (int i = 0; i < n; i++) {
    if (nb[i].size() != 5 && nb[i].size() != 6) throw std::runtime_error("bad valence");
    order_clockwise(P[i], P, nb[i]);
  }
  // fixed-topology Lloyd sweeps (Jacobi update)
  std::vector<V3> Q(n);
  for (int it = 0; it < lloyd_iters; it++) {
    for (int i = 0; i < n; i++) {
      const int k = int(nb[i].size());
      V3 c[6];
      for (int j = 0; j < k; j++) c[j] = circumcentre(P[i], P[nb[i][j]], P[nb[i][(j + 1) % k]]);
      V3 acc{0, 0, 0};
      for (int j = 0; j < k; j++) {
        V3 a = c[j], b = c[(j + 1) % k];
        double w = std::sqrt(dot(cross(a - P[i], b - P[i]), cross(a - P[i], b - P[i])));
        acc = acc + w * (P[i] + a + b);
      }
      Q[i] = unit(acc);
    }
    P.swap(Q);
  }
  // numbering: pentagons first, then along a Hilbert curve
  std::vector<int> order;
  for (int i = 0; i < 12; i++) order.push_back(i);
  std::vector<std::pair<uint64_t, int>> keyed;
  for (int i = 12; i < n; i++) keyed.push_back({hilbert_key(P[i]), i});
  std::sort(keyed.begin(), keyed.end());
  for (auto &kv : keyed) order.push_back(kv.second);
  std::vector<int> rank(n);
  for (int r = 0; r < n; r++) rank[order[r]] = r;

  Grid g;
  g.n = n;
  g.lat.resize(n);
  g.lon.resize(n);
  g.friends.assign(size_t(n) * 6, -1);
  g.clat.assign(size_t(n) * 6, -kRadConv);
  g.clon.assign(size_t(n) * 6, -kRadConv);
  // one circumcentre per triangle, shared bit-for-bit by its three nodes
  std::map<std::array<int, 3>, V3> cc;
  auto tri_cc = [&](int a, int b, int c) {
    std::array<int, 3> key = {a, b, c};
    std::sort(key.begin(), key.end());
    auto it = cc.find(key);
    if (it != cc.end()) return it->second;
    V3 v = circumcentre(P[key[0]], P[key[1]], P[key[2]]);
    cc[key] = v;
    return v;
  };
  for (int r = 0; r < n; r++) {
    int i = order[r];
    to_latlon(P[i], &g.lat[r], &g.lon[r]);
    const int k = int(nb[i].size());
    for (int j = 0; j < k; j++) {
      g.friends[r * 6 + j] = rank[nb[i][j]];
      V3 v = tri_cc(i, nb[i][j], nb[i][(j + 1) % k]);
      to_latlon(v, &g.clat[r * 6 + j], &g.clon[r * 6 + j]);
    }
  }
  return g;
}

}  // namespace odis
