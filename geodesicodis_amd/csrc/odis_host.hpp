// odis_host.hpp -- host-side C++ types of geodesicodis_amd: grid, TRiSK mesh
// assembly and the Globals-equivalent configuration.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "odis.h"

namespace odis {

constexpr double kPi = 3.1415926535897932384626433832795028841971693993751058;
constexpr double kRadConv = kPi / 180.0;  // include/mathRoutines.h:10

// Grid in the reference's file layout (src/mesh.cpp:4016-4101): node lat/lon
// in radians, 6 neighbour ids clockwise seen from outside (-1 pads the 12
// pentagons), and the 6 Voronoi vertices; vertex j is the circumcentre of
// (node, friend j, friend j+1).
struct Grid {
  int n = 0;
  std::vector<double> lat, lon;      // [n]
  std::vector<int> friends;          // [n*6]
  std::vector<double> clat, clon;    // [n*6]
};

Grid read_grid_file(const std::string &path);      // throws std::runtime_error
void write_grid_file(const Grid &g, const std::string &path);
Grid generate_grid(int level, int lloyd_iters);

// Hot-path operands of the reference's Mesh, in reference numbering.
struct MeshOps {
  int N = 0, E = 0, V = 0;
  double radius = 0.0;
  // per cell
  std::vector<double> lat, lon;          // [N]
  std::vector<double> cv_area;           // [N] map-plane control-volume area
  std::vector<int> faces;                // [N*6] -1 pad
  std::vector<int> node_face_dir;        // [N*6] +1 inner, -1 outer, 0 pad
  std::vector<double> div_coef;          // [N*6]
  // per vertex
  std::vector<double> vlat, vlon;        // [V]
  // per face
  std::vector<int> face_nodes;           // [E*2]
  std::vector<double> face_len;          // [E] l_e (Voronoi edge)
  std::vector<double> face_node_dist;    // [E] d_e (node to node)
  std::vector<double> face_lat, face_lon;// [E] face centre
  std::vector<double> face_m;            // [E*2] map factors at the face centre
  std::vector<double> face_normal;       // [E*2] map-plane unit normal
  std::vector<double> face_intercept;    // [E*2]
  std::vector<double> face_area;         // [E] d_e l_e
  std::vector<double> grad_coef;         // [E*2]
  std::vector<int> trisk_count;          // [E]
  std::vector<int> trisk_faces;          // [E*10]
  std::vector<double> trisk_w;           // [E*10] TRiSK weights w_ee'
  std::vector<double> trisk_coef;        // [E*10] w_ee' l_e' / d_e
  std::vector<double> face_normal_xyz;   // [E*3] Cartesian unit normal (first vertex x second)
  // RBF velocity reconstruction operator, CSR [3N x E] (rows 3i + x/y/z)
  std::vector<int> rbf_row_ptr, rbf_col;
  std::vector<double> rbf_val;
};

MeshOps build_mesh(const Grid &g, double radius);  // throws std::runtime_error
// The reference's operatorRBFinterp of a built mesh (fills the rbf_* arrays).
void build_rbf_operator(MeshOps &M);

// The Globals fields the LTE step and its drivers read (include/globals.h).
struct Config {
  double radius = 2574.73e3, ang_vel = 4.56e-6, k2 = 0.120, h2 = 0.2;
  double love_reduct = 0.920012, h = 400, shell_thickness = 0.0, g = 1.35;
  double semimajor = 1221.87e6, ecc = 0.0288, obliquity = 0.32 * kPi / 180.;
  double time_step = 40.0, converge = 1e-7, alpha = 2.28e-7;
  int l_max = 10;
  double d_lat = 45, d_lon = 90;
  int geodesic_l = 1;
  double period = 16. * 24 * 60 * 60;
  double end_time = 80.;
  std::string potential = "ECC", friction = "QUADRATIC", surface = "FREE", solver = "AB3",
              init = "NONE";
  bool advection = false;
  double rbf_eps = 0.25;
  int freq = 0;
  double a20 = 0, a22 = 0, b22 = 0;
  bool diss_avg = true, kinetic_avg = false, work = false;
  bool out_displacement = false, out_velocity = false, out_velocity_cart = false,
       out_pressure = false, out_diss = false, out_kinetic = false, out_dummy1 = false,
       out_dummy2 = false, out_sh = false;
  int core_num = 1, total_iter = 1000, output_time = 1;
  std::string grav_coeff_file, forcing_coeff_file;
  // derived
  int tide = ODIS_ECC, surface_type = ODIS_FREE, fric_type = ODIS_QUADRATIC;
  int solver_type = 1, init_type = 0;  // solver: 0 EULER, 1 AB3, 2 RK4; init: 0 NONE 1 LOAD 2 ANALYTICAL
  std::vector<double> shell_beta;      // LID_* beta_l
  std::vector<std::string> warnings;
};

// Parse input.in with the reference semantics (src/globals.cpp:35-323,
// 325-477) and apply applySurfaceBCs (src/boundaryConditions.cpp:7-475).
// Throws std::runtime_error with the reference's error text on fatal errors.
Config read_config(const std::string &case_dir);
void finalize_config(Config &c);   // period rounding, enum mapping, BCs

// CalcMaxTimeStep (src/mesh.cpp:1566-1623): dt = P / N_s with N_s the
// smallest multiple of 100 such that P / N_s <= target.
void quantise_dt(double period, double target, double *dt, int *n_steps);

// analyticalLTE OBLIQ_WEST (src/analyticalLTE.cpp:133-180): o = (U, V,
// dU/dt, dV/dt, eta, deta/dt) at (lat, lon, t) (odis_solver.cpp).
void analytical_obliq_west(double R, double Om, double g, double h, double obl, double al, double lat,
                           double lon, double t, double *o);

// membraneNuBeta (src/membraneConstants.cpp:8-144).
void membrane_nu_beta(double radius, double hs, double ho, int l_max, std::vector<double> &nu,
                      std::vector<double> &beta, double *radius_ocean, double *grav_ocean);

// The TRiSK operator in cell-local form (odis_compact.cpp): per cell, its
// faces in counter-clockwise order (cyc), the kite fractions R of the
// vertices between consecutive faces, and info = k | dir_bits << 3 (bit q
// set: the cell is the outer node of cyc[q]).  Reference numbering, [N][6].
struct CompactTrisk {
  int N = 0, E = 0;
  std::vector<int> cyc;
  std::vector<double> R;
  std::vector<int> info;
  double coef_err = 0, div_err = 0, area_err = 0;  // largest relative deviations found
};
// Empty string on success, else why the descriptor has no compact form.
std::string build_compact(const odis_mesh_desc &m, CompactTrisk &c);

// Real 4pi-normalised spherical harmonics, k = l*l + l + m (odis_sh.cpp).
// Recurrence coefficients of P_lm (index sh_lm_index(l, m), m <= l):
//   P_mm = P_{m-1,m-1} sect[m] cos(lat) (P_11 = sect[1] cos(lat)),
//   P_{m+1,m} = f1[m] sin(lat) P_mm,
//   P_lm = a_lm sin(lat) P_{l-1,m} - b_lm P_{l-2,m}.
inline int sh_lm_index(int l, int m) { return l * (l + 1) / 2 + m; }
void sh_recurrence_tables(int L, std::vector<double> &a, std::vector<double> &b, std::vector<double> &sect,
                          std::vector<double> &f1);
void sh_basis_row(int L, double lat, double lon, double *Y);
void sh_basis(int L, const std::vector<double> &lat, const std::vector<double> &lon, std::vector<double> &Yt);
std::vector<double> surface_beta(int surface, int L, double radius, double g, double shell, double ocean);

}  // namespace odis
