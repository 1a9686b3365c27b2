// odis_solver.cpp -- host drivers with the shape of the reference's program
// flow: main (src/main.cpp:46-68) -> Globals -> Mesh -> solveODIS
// (src/solver.cpp:15-61) -> ab3Explicit (src/timeIntegrator.cpp:57-322),
// with the time loop running on the GPU through odis.h.
//
// Kept from the reference: input.in semantics, dt quantisation, initial
// conditions (NONE / LOAD / ANALYTICAL), output cadence
// out_freq = totalIter / outputTime with the "DUMPING DATA AT ... AVG DISS:
// ... GW<k>" line in DATA/OUTPUT.txt, the DATA/data.h5 fields the input
// enables (displacement, east / north velocity, RBF Cartesian velocity,
// per-face dissipation, average dissipation; src/outFiles.cpp:138-681 and
// src/timeIntegrator.cpp:140-152, 170-200, 280-304), and the end-of-run text
// checkpoint InitialConditions/{vel,pres}_init.txt in %1.6E, written beside
// an unrounded binary FP64 copy {vel,pres}_init.bin that INIT_LOAD prefers.
#include <sys/stat.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "odis.h"
#include "odis_host.hpp"

namespace odis {

// analyticalLTE OBLIQ_WEST (src/analyticalLTE.cpp:133-180): returns
// (U, V, dU/dt, dV/dt, eta, deta/dt) at (lat, lon, t) for the ocean of radius
// R, rotation Om, gravity g, thickness h, obliquity obl (rad) and drag al.
void analytical_obliq_west(double R, double Om, double g, double h, double obl, double al, double lat,
                           double lon, double t, double *o) {
  using cd = std::complex<double>;
  const cd J(0.0, 1.0);
  const double w = -Om;
  const double lam = w / (2 * Om), lambs = 4 * Om * Om * R * R / (g * h);
  const double p21 = (2 + 1) * (2 + 1) / (2.0 * (2 * 2 + 1)), q11 = 1.0 * (1 + 1 - 1) / ((1 + 1.0) * (2 * 1 + 1));
  const cd L1 = lam + 1.0 / (1 * 2) + J * al / (2 * Om);
  const cd K2 = lam + 1.0 / (2 * 3) - 2 * 3 / (lambs * lam) + J * al / (2 * Om);
  const cd U21 = 0.5 * Om * Om * R * R * obl;
  const cd Psi = U21 / (2 * Om) / (q11 - K2 * L1 / p21);
  const cd Phi = -J * L1 / p21 * Psi;
  const double col = kPi * 0.5 - lat;
  const cd eil = std::exp(J * 1.0 * lon);
  const cd Y31 = 1.5 * (5 * std::pow(std::cos(col), 2.0) - 1) * std::sin(col) * eil;
  const cd Y21 = 3 * std::cos(col) * std::sin(col) * eil;
  const cd Y11 = std::sin(col) * eil;
  const cd ew = std::exp(-J * w * t);
  cd U = (-1 / 3. * (Psi * Y21) + J * Phi * Y21) * ew, dU = -J * w * U;
  U = (U + std::conj(U)) / (2 * R * std::sin(col));
  dU = (dU + std::conj(dU)) / (2 * R * std::sin(col));
  cd V = -(J * Psi * Y11 + Phi * (4.0 * Y31 - 9.0 * Y11) / 5.) * ew, dV = -J * w * V;
  V = (V + std::conj(V)) / (2 * R * std::sin(col));
  dV = (dV + std::conj(dV)) / (2 * R * std::sin(col));
  cd ET = J * 6.0 / std::pow(R, 2.0) * h / w * Phi * Y21 * ew, dE = -J * w * ET;
  ET = (ET + std::conj(ET)) * 0.5;
  dE = (dE + std::conj(dE)) * 0.5;
  o[0] = U.real();
  o[1] = V.real();
  o[2] = dU.real();
  o[3] = dV.real();
  o[4] = ET.real();
  o[5] = dE.real();
}

namespace {

struct Log {
  std::string out, err;
  void open(const std::string &dir) {
    ::mkdir((dir + "/DATA").c_str(), 0775);
    out = dir + "/DATA/OUTPUT.txt";
    err = dir + "/DATA/ERROR.txt";
    std::remove(out.c_str());
    std::remove(err.c_str());
    std::ofstream(err) << "ODIS error file. Warnings and model termination errors will be written here."
                       << std::endl
                       << std::endl;
  }
  void msg(const std::string &s) const {
    std::ofstream f(out, std::ios::app);
    f << s << std::endl;
  }
  void error(const std::string &s) const {
    std::ofstream f(err, std::ios::app);
    f << s << std::endl;
  }
};

// loadInitialConditions (src/initialConditions.cpp:19-144): up to `rows`
// lines of "x, d0, d1, d2"; missing file -> false.
bool load_text(const std::string &path, int rows, std::vector<double> &x, std::vector<double> &ds) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  std::string line;
  int i = 0;
  while (i < rows && std::getline(f, line)) {
    std::istringstream ls(line);
    std::string tok;
    double vals[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; k++) {
      ls >> std::ws;
      std::getline(ls, tok, ' ');
      vals[k] = std::stod(tok);
    }
    x[i] = vals[0];
    ds[3 * i] = vals[1];
    ds[3 * i + 1] = vals[2];
    ds[3 * i + 2] = vals[3];
    i++;
  }
  return true;
}

// writeInitialConditions (src/initialConditions.cpp:209-276)
void write_text(const std::string &path, const std::vector<double> &x, const std::vector<double> &ds) {
  std::remove(path.c_str());
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  for (size_t i = 0; i < x.size(); i++)
    std::fprintf(f, "%1.6E, %1.6E, %1.6E, %1.6E\n", x[i], ds[3 * i], ds[3 * i + 1], ds[3 * i + 2]);
  std::fclose(f);
}

// Binary FP64 companion of the text checkpoint: the same rows (x, d0, d1,
// d2) unrounded, after a 16-byte header "ODISCKP1" + row count (int64,
// little-endian as the host).  INIT_LOAD prefers it to the %1.6E text when
// its row count matches, so a resumed run continues from the exact state.
const char kCkpMagic[8] = {'O', 'D', 'I', 'S', 'C', 'K', 'P', '1'};

void write_bin(const std::string &path, const std::vector<double> &x, const std::vector<double> &ds) {
  std::remove(path.c_str());
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  const long long rows = (long long)x.size();
  std::vector<double> buf(size_t(4) * x.size());
  for (size_t i = 0; i < x.size(); i++) {
    buf[4 * i] = x[i];
    buf[4 * i + 1] = ds[3 * i];
    buf[4 * i + 2] = ds[3 * i + 1];
    buf[4 * i + 3] = ds[3 * i + 2];
  }
  const bool ok = std::fwrite(kCkpMagic, 1, 8, f) == 8 && std::fwrite(&rows, sizeof(rows), 1, f) == 1 &&
                  std::fwrite(buf.data(), sizeof(double), buf.size(), f) == buf.size();
  std::fclose(f);
  if (!ok) throw std::runtime_error("short write " + path);
}

// false when the file is absent, not a checkpoint, or of another row count
bool load_bin(const std::string &path, int rows, std::vector<double> &x, std::vector<double> &ds) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  char magic[8];
  long long n = -1;
  std::vector<double> buf(size_t(4) * rows);
  const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kCkpMagic, 8) == 0 &&
                  std::fread(&n, sizeof(n), 1, f) == 1 && n == rows &&
                  std::fread(buf.data(), sizeof(double), buf.size(), f) == buf.size();
  std::fclose(f);
  if (!ok) return false;
  for (int i = 0; i < rows; i++) {
    x[i] = buf[4 * size_t(i)];
    ds[3 * size_t(i)] = buf[4 * size_t(i) + 1];
    ds[3 * size_t(i) + 1] = buf[4 * size_t(i) + 2];
    ds[3 * size_t(i) + 2] = buf[4 * size_t(i) + 3];
  }
  return true;
}

std::string dump_line(double t_over_p, double diss_gw, int count) {
  std::ostringstream s;
  s << std::fixed << "DUMPING DATA AT " << t_over_p;
  s << " AVG DISS: " << std::scientific << diss_gw << " GW" << count;
  return s.str();
}

struct CtxGuard {
  odis_ctx *c = nullptr;
  ~CtxGuard() {
    if (c) odis_destroy(c);
  }
};

}  // namespace

// ab3Explicit on the GPU; returns 0, 1 if stopped (as the reference), < 0 on error.
static int ab3_explicit_gpu(const Config &c, const MeshOps &M, int precision, const std::string &dir, const Log &log) {
  double dt;
  int n_iter;
  quantise_dt(c.period, c.time_step, &dt, &n_iter);
  odis_params p{};
  p.g = c.g;
  p.h = c.h;
  p.alpha = c.alpha;
  p.omega = c.ang_vel;
  p.radius = c.radius;
  p.love_reduct = c.love_reduct;
  p.ecc = c.ecc;
  p.obliquity = c.obliquity;
  p.dt = dt;
  p.shell_thickness = c.shell_thickness;
  p.semimajor = c.semimajor;
  p.tide = c.tide;
  p.surface = c.surface_type;
  p.friction = c.fric_type;
  p.init_load = c.init_type == 1;
  odis_mesh_desc d{};
  d.n_cells = M.N;
  d.n_faces = M.E;
  d.face_nodes = M.face_nodes.data();
  d.grad_coef = M.grad_coef.data();
  d.trisk_count = M.trisk_count.data();
  d.trisk_faces = M.trisk_faces.data();
  d.trisk_coef = M.trisk_coef.data();
  d.face_lat = M.face_lat.data();
  d.face_area = M.face_area.data();
  d.cell_faces = M.faces.data();
  d.div_coef = M.div_coef.data();
  d.cell_lat = M.lat.data();
  d.cell_lon = M.lon.data();
  CtxGuard g;
  int rc = odis_create(&d, &p, 1, precision, &g.c);
  if (rc) {
    log.error(std::string("ERROR: ") + odis_last_error(nullptr));
    return rc;
  }
  std::vector<double> v(M.E, 0.0), dv(size_t(3) * M.E, 0.0), eta(M.N, 0.0), de(size_t(3) * M.N, 0.0);
  if (c.init_type == 1) {
    const std::string base = dir + "/InitialConditions/";
    if (load_bin(base + "vel_init.bin", M.E, v, dv))
      log.msg("Loaded FP64 checkpoint " + base + "vel_init.bin");
    else if (!load_text(base + "vel_init.txt", M.E, v, dv))
      log.error("WARNING: NO INITIAL CONDITION FILE FOUND " + base + "vel_init.txt");
    if (load_bin(base + "pres_init.bin", M.N, eta, de))
      log.msg("Loaded FP64 checkpoint " + base + "pres_init.bin");
    else if (!load_text(base + "pres_init.txt", M.N, eta, de))
      log.error("WARNING: NO INITIAL CONDITION FILE FOUND " + base + "pres_init.txt");
  } else if (c.init_type == 2) {
    // analyticalInitialConditions (src/initialConditions.cpp:146-206)
    const double R0 = c.radius, Om0 = c.ang_vel;
    double o[6];
    for (int e = 0; e < M.E; e++) {
      const double nx = M.face_normal[2 * e], ny = M.face_normal[2 * e + 1];
      analytical_obliq_west(R0, Om0, c.g, c.h, c.obliquity, c.alpha, M.face_lat[e], M.face_lon[e], 0.0, o);
      v[e] = o[0] * nx + o[1] * ny;
      dv[3 * e] = o[2] * nx + o[3] * ny;
      analytical_obliq_west(R0, Om0, c.g, c.h, c.obliquity, c.alpha, M.face_lat[e], M.face_lon[e], -dt, o);
      dv[3 * e + 1] = o[2] * nx + o[3] * ny;
      analytical_obliq_west(R0, Om0, c.g, c.h, c.obliquity, c.alpha, M.face_lat[e], M.face_lon[e], -2 * dt, o);
      dv[3 * e + 2] = o[2] * nx + o[3] * ny;
    }
    for (int i = 0; i < M.N; i++) {
      analytical_obliq_west(R0, Om0, c.g, c.h, c.obliquity, c.alpha, M.lat[i], M.lon[i], 0.0, o);
      eta[i] = o[4];
      de[3 * i] = o[5];
      analytical_obliq_west(R0, Om0, c.g, c.h, c.obliquity, c.alpha, M.lat[i], M.lon[i], -dt, o);
      de[3 * i + 1] = o[5];
      analytical_obliq_west(R0, Om0, c.g, c.h, c.obliquity, c.alpha, M.lat[i], M.lon[i], -2 * dt, o);
      de[3 * i + 2] = o[5];
    }
  }
  if ((rc = odis_set_state(g.c, 0, v.data(), dv.data(), eta.data(), de.data(), 0))) return rc;

  const double r = c.radius, P = c.period;
  const double to_gw = 4 * kPi * r * r / 1e9;
  double diss = 0.0;
  if ((rc = odis_get_dissipation(g.c, &diss))) {
    log.error(std::string("ERROR: ") + odis_last_error(g.c));
    return rc;
  }
  // DATA/data.h5 (CreateHDF5Framework at start, DumpData per output)
  const int fields = (c.out_displacement ? ODIS_H5_DISPLACEMENT : 0) | (c.out_velocity ? ODIS_H5_VELOCITY : 0) |
                     (c.out_velocity_cart ? ODIS_H5_VELOCITY_XYZ : 0) | (c.out_diss ? ODIS_H5_DISSIPATION : 0) |
                     (c.diss_avg ? ODIS_H5_DISS_AVG : 0);
  struct H5Guard {
    odis_h5 *h = nullptr;
    ~H5Guard() {
      if (h) odis_h5_close(h);
    }
  } h5;
  if (fields) {
    ::mkdir((dir + "/DATA").c_str(), 0775);
    const int slices = int(std::lround(c.end_time)) * c.output_time + 1;
    rc = odis_h5_open((dir + "/DATA/data.h5").c_str(), M.N, M.E, std::max(1, slices), fields, M.face_lat.data(),
                      M.face_lon.data(), &h5.h);
    if (rc) {
      log.error(std::string("WARNING: DATA/data.h5 NOT WRITTEN: ") + odis_h5_last_error());
      h5.h = nullptr;
    } else if ((fields & ODIS_H5_VELOCITY_XYZ) &&
               (rc = odis_rbf_set(g.c, 3 * M.N, M.E, M.rbf_row_ptr.data(), M.rbf_col.data(), M.rbf_val.data()))) {
      log.error(std::string("ERROR: ") + odis_last_error(g.c));
      return rc;
    }
  }
  auto dump = [&](int k, double total_diss) -> int {
    if (!h5.h) return ODIS_OK;
    const int slice = k - 1;
    std::vector<double> uv(size_t(2) * M.E), xyz(size_t(3) * M.N), s_v(M.E), s_dv(size_t(3) * M.E),
        s_eta(M.N), s_de(size_t(3) * M.N);
    int r = ODIS_OK;
    long long it_now = 0;
    if ((fields & ODIS_H5_DISPLACEMENT) &&
        !(r = odis_get_state(g.c, 0, s_v.data(), s_dv.data(), s_eta.data(), s_de.data(), &it_now)))
      r = odis_h5_write(h5.h, ODIS_H5_DISPLACEMENT, slice, s_eta.data());
    if (!r && (fields & (ODIS_H5_VELOCITY | ODIS_H5_DISSIPATION)))
      r = odis_get_velocity_en(g.c, 0, M.face_normal.data(), uv.data());
    if (!r && (fields & ODIS_H5_VELOCITY)) r = odis_h5_write(h5.h, ODIS_H5_VELOCITY, slice, uv.data());
    if (!r && (fields & ODIS_H5_DISSIPATION)) {
      // updateEnergy per-face flux (src/energy.cpp:13-62)
      std::vector<double> e(M.E);
      for (int f = 0; f < M.E; f++) {
        const double q = uv[2 * f] * uv[2 * f] + uv[2 * f + 1] * uv[2 * f + 1];
        e[f] = c.fric_type == ODIS_LINEAR ? c.alpha * 1000.0 * c.h * q : c.alpha / c.h * std::sqrt(q) * q;
      }
      r = odis_h5_write(h5.h, ODIS_H5_DISSIPATION, slice, e.data());
    }
    if (!r && (fields & ODIS_H5_VELOCITY_XYZ) && !(r = odis_velocity_xyz(g.c, 0, xyz.data())))
      r = odis_h5_write(h5.h, ODIS_H5_VELOCITY_XYZ, slice, xyz.data());
    if (!r && (fields & ODIS_H5_DISS_AVG)) r = odis_h5_write(h5.h, ODIS_H5_DISS_AVG, slice, &total_diss);
    if (r == ODIS_E_RANGE) return ODIS_OK;  // more outputs than the file's slices (a partial last orbit)
    if (r) log.error(std::string("ERROR: DATA/data.h5 output ") + std::to_string(k));
    return r;
  };
  int out_count = 1;
  log.msg(dump_line(0.0, diss * to_gw, out_count));
  if ((rc = dump(out_count, diss))) return rc;
  out_count++;
  const long long total = (long long)std::ceil(double(n_iter) * c.end_time);
  const int out_freq = std::max(1, n_iter / std::max(1, c.output_time));
  long long iter = 0;
  int status = 0;
  while (iter < total) {
    const long long next = std::min(total, (iter / out_freq + 1) * out_freq);
    rc = odis_step(g.c, int(next - iter), ODIS_SYNC);
    if (rc == ODIS_E_STOPPED) {
      log.msg("Terminate signal caught...");
      status = 1;
      break;
    }
    if (rc) {
      log.error(std::string("ERROR: ") + odis_last_error(g.c));
      return rc;
    }
    iter = next;
    if (iter % out_freq == 0) {
      if ((rc = odis_get_dissipation(g.c, &diss))) {
        log.error(std::string("ERROR: ") + odis_last_error(g.c));
        return rc;
      }
      log.msg(dump_line(dt * double(iter) / P, diss * to_gw, out_count));
      if ((rc = dump(out_count, diss))) return rc;
      out_count++;
    }
  }
  long long it = 0;
  if ((rc = odis_get_state(g.c, 0, v.data(), dv.data(), eta.data(), de.data(), &it))) return rc;
  ::mkdir((dir + "/InitialConditions").c_str(), 0775);
  write_text(dir + "/InitialConditions/vel_init.txt", v, dv);
  write_text(dir + "/InitialConditions/pres_init.txt", eta, de);
  write_bin(dir + "/InitialConditions/vel_init.bin", v, dv);
  write_bin(dir + "/InitialConditions/pres_init.bin", eta, de);
  return status;
}

}  // namespace odis

extern "C" int odis_run_case(const char *case_dir, const char *grid_dir, int precision) {
  using namespace odis;
  const std::string dir = case_dir ? case_dir : ".";
  Log log;
  log.open(dir);
  Config c;
  try {
    c = read_config(dir);
  } catch (const std::exception &e) {
    log.error(e.what());
    log.error("TERMINATING ODIS.");
    return ODIS_E_INVALID;
  }
  for (auto &w : c.warnings) log.error(w);
  if (c.advection) {
    log.error("ERROR: advection = true (nonlinear LTE) is not implemented by this solver");
    return ODIS_E_UNSUPPORTED;
  }
  if (c.solver_type != 1) {
    // solveODIS dispatches only AB3 (src/solver.cpp:27-52)
    log.error("ERROR: NO SOLVER FOUND!");
    return ODIS_E_UNSUPPORTED;
  }
  MeshOps M;
  try {
    const std::string name = "grid_l" + std::to_string(c.geodesic_l) + ".txt";
    const std::string g1 = (grid_dir ? std::string(grid_dir) : dir + "/input_files") + "/" + name;
    Grid g;
    std::ifstream probe(g1);
    if (probe.is_open()) {
      log.msg("Found mesh file: " + g1);
      g = read_grid_file(g1);
    } else {
      log.msg("Generating geodesic grid level " + std::to_string(c.geodesic_l));
      g = generate_grid(c.geodesic_l, 64);
    }
    M = build_mesh(g, c.radius);
  } catch (const std::exception &e) {
    log.error(std::string("ERROR: ") + e.what());
    return ODIS_E_INVALID;
  }
  log.msg("Entering time solver AB3...");
  int rc;
  try {
    rc = ab3_explicit_gpu(c, M, precision, dir, log);
  } catch (const std::exception &e) {  // host staging buffers; the contexts are freed by their guards
    log.error(std::string("ERROR: ") + e.what());
    return ODIS_E_NOMEM;
  }
  if (rc > 0) log.msg("SOLVER RETURNED WITH AN ERROR...");
  else if (rc == 0) log.msg("Calculations appear to have finished!");
  return rc;
}
