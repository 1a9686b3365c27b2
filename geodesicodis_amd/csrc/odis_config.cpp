// odis_config.cpp -- input.in semantics of the reference's Globals.
//
// Parsing follows Globals::ReadGlobals (src/globals.cpp:325-477): the file is
// a stream of "key; value; comment;" triplets, keys are matched
// case-insensitively against the reference's string IDs, values are read
// with operator>> for their type (bools are the literal tokens true/false),
// unknown keys are ignored, unassigned keys keep the Titan defaults of
// SetDefault (src/globals.cpp:479-571) and produce a warning.  Afterwards the
// obliquity is converted to radians, the period is recomputed from the
// angular velocity and rounded to an even number of seconds
// (src/globals.cpp:209-214), the string options are mapped to enums
// (src/globals.cpp:233-292) and applySurfaceBCs runs
// (src/boundaryConditions.cpp:7-475).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "odis_host.hpp"

namespace odis {

namespace {

enum class Ty { D, I, B, S };
struct Key {
  const char *id;
  Ty ty;
  void *ptr;
  bool seen;
};

std::vector<Key> key_table(Config &c) {
  // Order and IDs of src/globals.cpp:56-216.  "fourier coeff file" is pushed
  // as a second entry for forcing_coeff_file in the reference
  // (src/globals.cpp:204-205), so that key never matches anything.
  return {
      {"radius", Ty::D, &c.radius, false},
      {"angular velocity", Ty::D, &c.ang_vel, false},
      {"k2", Ty::D, &c.k2, false},
      {"h2", Ty::D, &c.h2, false},
      {"love reduction factor", Ty::D, &c.love_reduct, false},
      {"ocean thickness", Ty::D, &c.h, false},
      {"shell thickness", Ty::D, &c.shell_thickness, false},
      {"surface gravity", Ty::D, &c.g, false},
      {"semimajor axis", Ty::D, &c.semimajor, false},
      {"eccentricity", Ty::D, &c.ecc, false},
      {"obliquity", Ty::D, &c.obliquity, false},
      {"time step", Ty::D, &c.time_step, false},
      {"converge", Ty::D, &c.converge, false},
      {"friction coefficient", Ty::D, &c.alpha, false},
      {"sh degree", Ty::I, &c.l_max, false},
      {"latitude spacing", Ty::D, &c.d_lat, false},
      {"longitude spacing", Ty::D, &c.d_lon, false},
      {"geodesic grid level", Ty::I, &c.geodesic_l, false},
      {"orbital period", Ty::D, &c.period, false},
      {"advection", Ty::B, &c.advection, false},
      {"simulation end time", Ty::D, &c.end_time, false},
      {"fourier a20", Ty::D, &c.a20, false},
      {"fourier a22", Ty::D, &c.a22, false},
      {"fourier b22", Ty::D, &c.b22, false},
      {"fourier freq", Ty::I, &c.freq, false},
      {"potential", Ty::S, &c.potential, false},
      {"friction type", Ty::S, &c.friction, false},
      {"surface type", Ty::S, &c.surface, false},
      {"solver type", Ty::S, &c.solver, false},
      {"initial conditions", Ty::S, &c.init, false},
      {"dissipation avg output", Ty::B, &c.diss_avg, false},
      {"kinetic avg output", Ty::B, &c.kinetic_avg, false},
      {"work flux output", Ty::B, &c.work, false},
      {"displacement output", Ty::B, &c.out_displacement, false},
      {"velocity cartesian output", Ty::B, &c.out_velocity_cart, false},
      {"velocity output", Ty::B, &c.out_velocity, false},
      {"pressure output", Ty::B, &c.out_pressure, false},
      {"dissipation output", Ty::B, &c.out_diss, false},
      {"kinetic output", Ty::B, &c.out_kinetic, false},
      {"dummy1 output", Ty::B, &c.out_dummy1, false},
      {"dummy2 output", Ty::B, &c.out_dummy2, false},
      {"sh coefficient output", Ty::B, &c.out_sh, false},
      {"output time", Ty::I, &c.output_time, false},
      {"total iterations", Ty::I, &c.total_iter, false},
      {"core number", Ty::I, &c.core_num, false},
      {"grav coeff file", Ty::S, &c.grav_coeff_file, false},
      {"forcing coeff file", Ty::S, &c.forcing_coeff_file, false},
      {"forcing coeff file", Ty::S, &c.forcing_coeff_file, false},
      {"rbf epsilon", Ty::D, &c.rbf_eps, false},
  };
}

}  // namespace

Config read_config(const std::string &case_dir) {
  Config c;
  std::ifstream in(case_dir + "/input.in");
  if (!in.is_open()) throw std::runtime_error("Unable to open 'input.in' file.");
  std::vector<Key> keys = key_table(c);
  std::string id, val;
  bool val_bool = true;  // persists across keys like the reference's valBool
  while (std::getline(in >> std::ws, id, ';')) {
    std::transform(id.begin(), id.end(), id.begin(), [](unsigned char ch) { return std::tolower(ch); });
    std::getline(in >> std::ws, val, ';');
    std::stringstream vs(val);
    for (auto &k : keys) {
      if (id != k.id) continue;
      switch (k.ty) {
        case Ty::D: {
          double d = 0.0;
          vs >> d;
          *static_cast<double *>(k.ptr) = d;
        } break;
        case Ty::I: {
          int v = 0;
          vs >> v;
          *static_cast<int *>(k.ptr) = v;
        } break;
        case Ty::B: {
          // src/globals.cpp:390-396: anything but "false"/"true" keeps the
          // previous parsed bool (initially true)
          std::string s;
          vs >> s;
          if (s == "false") val_bool = false;
          else if (s == "true") val_bool = true;
          *static_cast<bool *>(k.ptr) = val_bool;
        } break;
        case Ty::S: {
          std::string s;
          vs >> s;
          *static_cast<std::string *>(k.ptr) = s;
        } break;
      }
      k.seen = true;
      break;
    }
    std::string comment;
    std::getline(in, comment, ';');
  }
  c.obliquity = c.obliquity * kPi / 180.;
  c.period = 2. * kPi / c.ang_vel;
  for (auto &k : keys)
    if (!k.seen) c.warnings.push_back(std::string("WARNING: Unassigned global constant: ") + k.id + ".");
  finalize_config(c);
  return c;
}

void finalize_config(Config &c) {
  // period rounding, src/globals.cpp:209-214
  c.period = 2. * kPi / c.ang_vel;
  const int int_time = int(std::round(c.period / 2)) * 2;
  c.period = double(int_time);
  c.ang_vel = 2 * kPi / double(int_time);

  if (c.friction == "LINEAR") c.fric_type = ODIS_LINEAR;
  else if (c.friction == "QUADRATIC") c.fric_type = ODIS_QUADRATIC;
  else throw std::runtime_error("ERROR: NO DRAG MODEL FOUND!");

  if (c.solver == "EULER") c.solver_type = 0;
  else if (c.solver == "AB3") c.solver_type = 1;
  else if (c.solver == "RK4") c.solver_type = 2;
  else throw std::runtime_error("ERROR: NO SOLVER FOUND!");

  static const char *pots[] = {"OBLIQ", "OBLIQ_WEST", "OBLIQ_EAST", "ECC_RAD", "ECC_LIB", "ECC",
                               "ECC_WEST", "ECC_EAST", "FULL", "FULL2", "TOTAL", "ECC_W3",
                               "OBLIQ_W3", "PLANET", "PLANET_OBL", "GENERAL", "NONE"};
  c.tide = -1;
  for (int k = 0; k < 17; k++)
    if (c.potential == pots[k]) c.tide = k;
  if (c.tide < 0) throw std::runtime_error("ERROR: NO POTENTIAL FORCING FOUND!");

  static const char *surfs[] = {"FREE", "FREE_LOADING", "LID_LOVE", "LID_MEMBR", "LID_NUM", "LID_INF"};
  c.surface_type = -1;
  for (int k = 0; k < 6; k++)
    if (c.surface == surfs[k]) c.surface_type = k;
  if (c.surface_type < 0) throw std::runtime_error("ERROR: NO OCEAN SURFACE BOUNDARY CONDITION FOUND!");

  if (c.init == "NONE") c.init_type = 0;
  else if (c.init == "LOAD") c.init_type = 1;
  else if (c.init == "ANALYTICAL") c.init_type = 2;
  else
    throw std::runtime_error(
        "ERROR: INITIAL CONDITION MUST BE 0 (none), 1 (load from file), or 2 (analytical)!");

  // applySurfaceBCs, src/boundaryConditions.cpp:7-475
  switch (c.surface_type) {
    case ODIS_FREE:
      c.love_reduct = 1.0 + c.k2 - c.h2;
      break;
    case ODIS_FREE_LOADING: {
      const double den_ocean = 1000.0, bulk_den = 1609.22, rig = 40e9;
      const int l = 2;
      double eff = double(2 * l * l + 4 * l + 3) / double(l) * rig / (bulk_den * c.g * c.radius);
      c.love_reduct = 1.0 + 1.5 / (1. + eff) - 2.5 / (1. + eff);
      (void)den_ocean;
    } break;
    case ODIS_LID_MEMBR: {
      std::vector<double> nu, beta;
      double ro, go;
      membrane_nu_beta(c.radius, c.shell_thickness, c.h, c.l_max, nu, beta, &ro, &go);
      c.shell_beta = beta;
      c.love_reduct = nu.size() > 2 ? nu[2] : c.love_reduct;
      c.g = go;
      c.radius = ro;
    } break;
    case ODIS_LID_LOVE:
      // Needs input_files/beta.txt, which the reference does not ship; the
      // radius is reduced to the ocean top (src/boundaryConditions.cpp:126).
      c.radius = c.radius - c.shell_thickness;
      break;
    default:
      break;
  }
}

void quantise_dt(double period, double target, double *dt, int *n_steps) {
  int num = 100;
  double d = period;
  while (d > target) {
    d = period / num;
    num += 100;
  }
  *dt = d;
  *n_steps = num - 100;
}

void membrane_nu_beta(double radius, double hs, double ho, int l_max, std::vector<double> &nu,
                      std::vector<double> &beta, double *radius_ocean, double *grav_ocean) {
  const double G = 6.67384e-11, pois = 0.5, rig_shell = 3.5e9, rig_core = 40e9;
  const double rc = radius - (hs + ho), ro = radius - hs;
  const double mass_total = 1.08e20, den_ocean = 1000.0, den_shell = 940.0;
  const double grav_surf = G * mass_total / (radius * radius);
  const double vol_total = 4. / 3. * kPi * std::pow(radius, 3.0);
  const double vol_core = 4. / 3. * kPi * std::pow(rc, 3.0);
  const double vol_ocean = 4. / 3. * kPi * std::pow(ro, 3.0) - vol_core;
  const double vol_shell = vol_total - 4. / 3. * kPi * std::pow(ro, 3.0);
  const double den_bulk = mass_total / vol_total;
  const double mass_core = mass_total - (vol_ocean * den_ocean + vol_shell * den_shell);
  const double den_core = mass_core / vol_core;
  const double grav_core = G * mass_core / (rc * rc);
  *grav_ocean = G * (mass_core + vol_ocean * den_ocean) / (ro * ro);
  *radius_ocean = ro;
  nu.assign(l_max + 1, 0.0);
  beta.assign(l_max + 1, 0.0);
  for (int l = 0; l <= l_max; l++) {
    const double rf = 1. / (1. + double(2 * l * l + 4 * l + 3) / double(l) * rig_core /
                                     (den_core * grav_core * rc));
    double kt = rf * 3. / double(2 * (l - 1)), ht = rf * double(2 * l + 1) / double(2 * (l - 1));
    if (l == 1) kt = ht = 0.0;
    const double kl = -rf, hl = -rf * double(2 * l + 1) / 3.0;
    const double x = double((l - 1) * (l + 2));
    const double bend = rig_shell * std::pow(hs, 3.0) / (6. * (1. - pois));
    const double smem = 2. * x * (1. + pois) / (x + 1. + pois) * rig_shell / (den_ocean * grav_surf * radius) * hs / radius;
    const double sbend = x * x * (x + 2.) / (x + 1. + pois) * bend / (den_ocean * grav_surf * std::pow(radius, 4.0));
    const double spr = smem + sbend;
    const double xi = 3.0 / (2. * double(l) + 1.0) * (den_ocean / den_bulk);
    const double dspr = -spr * (1. - (1. + xi * hl) * (1. + xi * hl) / (1. + xi * (ht - hl) * spr));
    const double dgam = -spr * ((1. + xi * hl) * ht / (1. + xi * (ht - hl) * spr));
    beta[l] = 1. - xi * (1.0 + kl - hl) + spr + dspr;
    nu[l] = (1.0 + kt - ht) + dgam;
  }
}

}  // namespace odis
