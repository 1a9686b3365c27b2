// odis_compact.cpp -- the TRiSK tangential-velocity operator in cell-local
// ("compact") form, derived from and checked against the tabulated operator.
//
// The reference's TRiSK weights (src/mesh.cpp:677-910, restated in
// odis_mesh.cpp) for face e and a neighbour face e_j of one of its cells c
// are
//     w(e, e_j) = t_e dir_j (R_0 + ... + R_{j-1} - 1/2),
// where e_0 = e, e_1, ..., e_{k-1} are the faces of c in counter-clockwise
// order starting at e, R_q is the kite-area fraction of the Voronoi vertex
// between e_q and e_{q+1}, dir_j = +1 when c is the inner node of e_j and
// t_e = dir of e itself.  The coefficient applied to v[e_j] is
// w l_{e_j} / d_e (src/interpolation.cpp:43-46).  So the whole operator is
// fixed by one cyclic face order and k fractions per cell, plus the face
// lengths l and node distances d:
//     v_t(e) = (T(c0, e) - T(c1, e)) / d_e,
//     T(c, e_0) = sum_{j=1}^{k-1} (R_0 + ... + R_{j-1} - 1/2) dir_j l_j v_j,
// with c0 / c1 the inner / outer node of e.  A cell kernel that already
// reads dir_j l_j v_j for the divergence computes T for all k faces of its
// cell, and the face kernel needs two values per face instead of ten
// indices and ten coefficients.
//
// The cell order and the fractions are recovered from the tabulated
// operator the caller passes (trisk_faces / trisk_coef: the reference's
// face_interp_friends and weights), and the recovered form is checked
// against every tabulated coefficient and every divergence coefficient.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "odis_host.hpp"

namespace odis {

std::string build_compact(const odis_mesh_desc &m, CompactTrisk &c) {
  const int N = m.n_cells, E = m.n_faces;
  if (!m.face_len || !m.face_dist || !m.cell_area)
    return "mesh descriptor has no face_len / face_dist / cell_area";
  c.N = N;
  c.E = E;
  c.cyc.assign(size_t(6) * N, -1);
  c.R.assign(size_t(6) * N, 0.0);
  c.info.assign(N, 0);
  c.coef_err = c.div_err = c.area_err = 0.0;
  std::vector<int> deg(N, 0);
  for (int i = 0; i < N; i++) {
    int k = 0;
    while (k < 6 && m.cell_faces[6 * i + k] >= 0) k++;
    for (int j = k; j < 6; j++)
      if (m.cell_faces[6 * i + j] >= 0) return "cell_faces: padding before a face";
    if (k < 5) return "cell with fewer than 5 faces";
    deg[i] = k;
  }
  // side-c slots of face e: the first deg(inner) - 1 TRiSK slots belong to
  // the inner node, the next deg(outer) - 1 to the outer node
  auto side_slots = [&](int e, int cell, int *out) -> int {
    const int a = m.face_nodes[2 * e], b = m.face_nodes[2 * e + 1];
    if (m.trisk_count[e] != deg[a] - 1 + deg[b] - 1) return -1;
    int off, cnt;
    if (cell == a) {
      off = 0;
      cnt = deg[a] - 1;
    } else if (cell == b) {
      off = deg[a] - 1;
      cnt = deg[b] - 1;
    } else {
      return -1;
    }
    for (int j = 0; j < cnt; j++) out[j] = m.trisk_faces[10 * e + off + j];
    return off;
  };
  for (int e = 0; e < E; e++) {
    const double dl = m.face_dist[e] * m.face_len[e];
    if (!(m.face_dist[e] > 0) || !(m.face_len[e] > 0)) return "face_len / face_dist must be positive";
    c.area_err = std::max(c.area_err, std::fabs(dl - m.face_area[e]) / std::fabs(m.face_area[e]));
  }
  for (int i = 0; i < N; i++) {
    const int k = deg[i];
    int cyc[6], slots[6];
    cyc[0] = m.cell_faces[6 * i];
    if (side_slots(cyc[0], i, slots) < 0) return "TRiSK stencil does not hold the faces of both cells";
    for (int j = 1; j < k; j++) cyc[j] = slots[j - 1];
    // the cycle holds exactly the cell's faces
    for (int j = 0; j < k; j++) {
      bool found = false;
      for (int q = 0; q < k; q++) found = found || cyc[j] == m.cell_faces[6 * i + q];
      if (!found) return "TRiSK stencil of a cell is not its face set";
    }
    int dir[6], dbits = 0;
    for (int q = 0; q < k; q++) {
      dir[q] = m.face_nodes[2 * cyc[q]] == i ? 1 : -1;
      if (dir[q] < 0) dbits |= 1 << q;
    }
    // every face's stencil on this side is the cycle rotated to start at it
    double R[6];
    for (int q = 0; q < k; q++) {
      const int e = cyc[q];
      const int off = side_slots(e, i, slots);
      if (off < 0) return "TRiSK stencil does not hold the faces of both cells";
      for (int j = 1; j < k; j++)
        if (slots[j - 1] != cyc[(q + j) % k]) return "TRiSK stencils of one cell are not rotations of one cycle";
      const int n1 = cyc[(q + 1) % k];
      const double w = m.trisk_coef[10 * e + off] * m.face_dist[e] / m.face_len[n1];
      R[q] = w * double(dir[q] * dir[(q + 1) % k]) + 0.5;
    }
    // the recovered form reproduces every tabulated coefficient of the side
    for (int q = 0; q < k; q++) {
      const int e = cyc[q];
      const int off = side_slots(e, i, slots);
      double rowmax = 0.0;
      for (int j = 0; j < m.trisk_count[e]; j++) rowmax = std::max(rowmax, std::fabs(m.trisk_coef[10 * e + j]));
      double S = 0.0;
      for (int j = 1; j < k; j++) {
        S += R[(q + j - 1) % k];
        const int f = cyc[(q + j) % k];
        const double w = double(dir[q] * dir[(q + j) % k]) * (S - 0.5);
        const double cf = w * m.face_len[f] / m.face_dist[e];
        c.coef_err = std::max(c.coef_err, std::fabs(cf - m.trisk_coef[10 * e + off + j - 1]) / rowmax);
      }
    }
    // divergence: div_coef = -dir l / A in cell_faces order
    double dmax = 0.0;
    for (int j = 0; j < k; j++) dmax = std::max(dmax, std::fabs(m.div_coef[6 * i + j]));
    for (int j = 0; j < k; j++) {
      const int f = m.cell_faces[6 * i + j];
      const double d = -double(m.face_nodes[2 * f] == i ? 1 : -1) * m.face_len[f] / m.cell_area[i];
      c.div_err = std::max(c.div_err, std::fabs(d - m.div_coef[6 * i + j]) / dmax);
    }
    for (int q = 0; q < 6; q++) {
      c.cyc[size_t(i) * 6 + q] = cyc[q < k ? q : 0];
      c.R[size_t(i) * 6 + q] = q < k ? R[q] : 0.0;
    }
    c.info[i] = k | (dbits << 3);
  }
  const double tol = 1e-12;
  if (!(c.coef_err <= tol)) return "compact TRiSK form does not reproduce the tabulated coefficients";
  if (!(c.div_err <= tol)) return "divergence coefficients are not -dir l / A";
  if (!(c.area_err <= tol)) return "face_area is not face_dist * face_len";
  return std::string();
}

}  // namespace odis
