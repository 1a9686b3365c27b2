// odis_capi.cpp -- extern "C" host entry points for grids and meshes.
#include <cmath>
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "odis.h"
#include "odis_host.hpp"

struct odis_grid {
  odis::Grid g;
};
struct odis_mesh {
  odis::MeshOps m;
};

namespace {
thread_local std::string g_host_error;
int host_fail(int code, const std::string &msg) {
  g_host_error = msg;
  return code;
}
}  // namespace

extern "C" const char *odis_host_last_error(void) { return g_host_error.c_str(); }

extern "C" int odis_grid_read(const char *path, odis_grid **out) {
  if (!path || !out) return host_fail(ODIS_E_INVALID, "null argument");
  try {
    auto *g = new odis_grid{odis::read_grid_file(path)};
    *out = g;
    return ODIS_OK;
  } catch (const std::exception &e) {
    return host_fail(ODIS_E_IO, e.what());
  }
}

extern "C" int odis_grid_generate(int level, int lloyd_iters, odis_grid **out) {
  if (!out || lloyd_iters < 0) return host_fail(ODIS_E_INVALID, "bad argument");
  try {
    auto *g = new odis_grid{odis::generate_grid(level, lloyd_iters)};
    *out = g;
    return ODIS_OK;
  } catch (const std::exception &e) {
    return host_fail(ODIS_E_INVALID, e.what());
  }
}

extern "C" int odis_grid_write(const odis_grid *g, const char *path) {
  if (!g || !path) return host_fail(ODIS_E_INVALID, "null argument");
  try {
    odis::write_grid_file(g->g, path);
    return ODIS_OK;
  } catch (const std::exception &e) {
    return host_fail(ODIS_E_IO, e.what());
  }
}

extern "C" int odis_grid_size(const odis_grid *g) { return g ? g->g.n : ODIS_E_INVALID; }

extern "C" int odis_grid_arrays(const odis_grid *g, double *lat, double *lon, int *friends, double *clat,
                                double *clon) {
  if (!g) return ODIS_E_INVALID;
  const size_t n = size_t(g->g.n);
  if (lat) std::memcpy(lat, g->g.lat.data(), n * sizeof(double));
  if (lon) std::memcpy(lon, g->g.lon.data(), n * sizeof(double));
  if (friends) std::memcpy(friends, g->g.friends.data(), n * 6 * sizeof(int));
  if (clat) std::memcpy(clat, g->g.clat.data(), n * 6 * sizeof(double));
  if (clon) std::memcpy(clon, g->g.clon.data(), n * 6 * sizeof(double));
  return ODIS_OK;
}

extern "C" int odis_grid_from_arrays(int n, const double *lat, const double *lon, const int *friends,
                                     const double *clat, const double *clon, odis_grid **out) {
  if (n < 12 || !lat || !lon || !friends || !clat || !clon || !out) return host_fail(ODIS_E_INVALID, "bad argument");
  // friend ids lie in [0, n); -1 (no sixth neighbour, pentagons) only in the last slot
  for (size_t k = 0; k < size_t(n) * 6; k++) {
    const int f = friends[k];
    if (f >= n || f < -1 || (f < 0 && k % 6 != 5)) return host_fail(ODIS_E_INVALID, "friend index out of range");
  }
  try {
    auto *g = new odis_grid;
    g->g.n = n;
    g->g.lat.assign(lat, lat + n);
    g->g.lon.assign(lon, lon + n);
    g->g.friends.assign(friends, friends + size_t(n) * 6);
    g->g.clat.assign(clat, clat + size_t(n) * 6);
    g->g.clon.assign(clon, clon + size_t(n) * 6);
    *out = g;
    return ODIS_OK;
  } catch (const std::bad_alloc &) {
    return host_fail(ODIS_E_NOMEM, "host allocation failed");
  }
}

extern "C" int odis_grid_free(odis_grid *g) {
  delete g;
  return ODIS_OK;
}

extern "C" int odis_mesh_build(const odis_grid *g, double radius, odis_mesh **out) {
  if (!g || !out || !(radius > 0)) return host_fail(ODIS_E_INVALID, "bad argument");
  try {
    auto *m = new odis_mesh{odis::build_mesh(g->g, radius)};
    *out = m;
    return ODIS_OK;
  } catch (const std::exception &e) {
    return host_fail(ODIS_E_INVALID, e.what());
  }
}

extern "C" int odis_mesh_sizes(const odis_mesh *m, int *n, int *e, int *v) {
  if (!m) return ODIS_E_INVALID;
  if (n) *n = m->m.N;
  if (e) *e = m->m.E;
  if (v) *v = m->m.V;
  return ODIS_OK;
}

extern "C" int odis_mesh_describe(const odis_mesh *mm, odis_mesh_desc *d) {
  if (!mm || !d) return ODIS_E_INVALID;
  const auto &m = mm->m;
  d->n_cells = m.N;
  d->n_faces = m.E;
  d->face_nodes = m.face_nodes.data();
  d->grad_coef = m.grad_coef.data();
  d->trisk_count = m.trisk_count.data();
  d->trisk_faces = m.trisk_faces.data();
  d->trisk_coef = m.trisk_coef.data();
  d->face_lat = m.face_lat.data();
  d->face_area = m.face_area.data();
  d->cell_faces = m.faces.data();
  d->div_coef = m.div_coef.data();
  d->cell_lat = m.lat.data();
  d->cell_lon = m.lon.data();
  d->face_len = m.face_len.data();
  d->face_dist = m.face_node_dist.data();
  d->cell_area = m.cv_area.data();
  return ODIS_OK;
}

extern "C" int odis_mesh_compact_check(const odis_mesh_desc *d, double *coef_err, double *div_err,
                                       double *area_err) {
  if (!d) return ODIS_E_INVALID;
  odis::CompactTrisk c;
  std::string why;
  try {
    why = odis::build_compact(*d, c);
  } catch (const std::exception &ex) {
    why = ex.what();
  }
  if (coef_err) *coef_err = c.coef_err;
  if (div_err) *div_err = c.div_err;
  if (area_err) *area_err = c.area_err;
  if (!why.empty()) return host_fail(ODIS_E_UNSUPPORTED, why);
  return ODIS_OK;
}

extern "C" int odis_mesh_get(const odis_mesh *mm, const char *name, void *dst) {
  if (!mm || !name) return ODIS_E_INVALID;
  const auto &m = mm->m;
  struct Entry {
    const char *nm;
    const void *src;
    size_t cnt, sz;
  };
  const Entry tab[] = {
      {"cv_area", m.cv_area.data(), m.cv_area.size(), sizeof(double)},
      {"faces", m.faces.data(), m.faces.size(), sizeof(int)},
      {"node_face_dir", m.node_face_dir.data(), m.node_face_dir.size(), sizeof(int)},
      {"div_coef", m.div_coef.data(), m.div_coef.size(), sizeof(double)},
      {"vertex_lat", m.vlat.data(), m.vlat.size(), sizeof(double)},
      {"vertex_lon", m.vlon.data(), m.vlon.size(), sizeof(double)},
      {"face_nodes", m.face_nodes.data(), m.face_nodes.size(), sizeof(int)},
      {"face_len", m.face_len.data(), m.face_len.size(), sizeof(double)},
      {"face_node_dist", m.face_node_dist.data(), m.face_node_dist.size(), sizeof(double)},
      {"face_centre_lat", m.face_lat.data(), m.face_lat.size(), sizeof(double)},
      {"face_centre_lon", m.face_lon.data(), m.face_lon.size(), sizeof(double)},
      {"face_centre_m", m.face_m.data(), m.face_m.size(), sizeof(double)},
      {"face_normal", m.face_normal.data(), m.face_normal.size(), sizeof(double)},
      {"face_intercept", m.face_intercept.data(), m.face_intercept.size(), sizeof(double)},
      {"face_area", m.face_area.data(), m.face_area.size(), sizeof(double)},
      {"grad_coef", m.grad_coef.data(), m.grad_coef.size(), sizeof(double)},
      {"interp_count", m.trisk_count.data(), m.trisk_count.size(), sizeof(int)},
      {"interp_friends", m.trisk_faces.data(), m.trisk_faces.size(), sizeof(int)},
      {"interp_weights", m.trisk_w.data(), m.trisk_w.size(), sizeof(double)},
      {"trisk_coef", m.trisk_coef.data(), m.trisk_coef.size(), sizeof(double)},
      {"face_normal_xyz", m.face_normal_xyz.data(), m.face_normal_xyz.size(), sizeof(double)},
      {"rbf_row_ptr", m.rbf_row_ptr.data(), m.rbf_row_ptr.size(), sizeof(int)},
      {"rbf_col", m.rbf_col.data(), m.rbf_col.size(), sizeof(int)},
      {"rbf_val", m.rbf_val.data(), m.rbf_val.size(), sizeof(double)},
      {"cell_lat", m.lat.data(), m.lat.size(), sizeof(double)},
      {"cell_lon", m.lon.data(), m.lon.size(), sizeof(double)},
  };
  for (const auto &t : tab)
    if (std::strcmp(t.nm, name) == 0) {
      if (dst) std::memcpy(dst, t.src, t.cnt * t.sz);
      return int(t.cnt);
    }
  return host_fail(ODIS_E_INVALID, std::string("unknown mesh array ") + name);
}

extern "C" int odis_mesh_free(odis_mesh *m) {
  delete m;
  return ODIS_OK;
}

extern "C" int odis_round_period(double ang_vel, double *period, double *omega) {
  if (!period || !omega || !(ang_vel > 0)) return host_fail(ODIS_E_INVALID, "bad argument");
  const double per = 2. * odis::kPi / ang_vel;
  if (!(per >= 2.0 && per < 2e9)) return host_fail(ODIS_E_INVALID, "period outside [2 s, 2e9 s)");
  const int it = int(std::round(per / 2)) * 2;
  *period = double(it);
  *omega = 2 * odis::kPi / double(it);
  return ODIS_OK;
}

extern "C" int odis_quantise_dt(double period, double target, double *dt, int *n) {
  if (!dt || !n || !(period > 0) || !(target > 0)) return host_fail(ODIS_E_INVALID, "bad argument");
  if (!(period / target < 1e8)) return host_fail(ODIS_E_INVALID, "more than 1e8 steps per period");
  odis::quantise_dt(period, target, dt, n);
  return ODIS_OK;
}

extern "C" int odis_config_read(const char *case_dir, odis_params *p, int *steps, double *end_time, int *level,
                                int *output_time) {
  if (!case_dir) return host_fail(ODIS_E_INVALID, "null case_dir");
  try {
    odis::Config c = odis::read_config(case_dir);
    if (!(c.period > 0) || !(c.time_step > 0) || !(c.period / c.time_step < 1e8))
      return host_fail(ODIS_E_INVALID, "time step: need 0 < period / time step < 1e8");
    double dt;
    int n;
    odis::quantise_dt(c.period, c.time_step, &dt, &n);
    if (p) {
      *p = odis_params{};
      p->g = c.g;
      p->h = c.h;
      p->alpha = c.alpha;
      p->omega = c.ang_vel;
      p->radius = c.radius;
      p->love_reduct = c.love_reduct;
      p->ecc = c.ecc;
      p->obliquity = c.obliquity;
      p->dt = dt;
      p->shell_thickness = c.shell_thickness;
      p->semimajor = c.semimajor;
      p->tide = c.tide;
      p->surface = c.surface_type;
      p->friction = c.fric_type;
      p->init_load = c.init_type == 1;
      p->quad_drag = 0;
    }
    if (steps) *steps = n;
    if (end_time) *end_time = c.end_time;
    if (level) *level = c.geodesic_l;
    if (output_time) *output_time = c.output_time;
    return ODIS_OK;
  } catch (const std::exception &e) {
    return host_fail(ODIS_E_INVALID, e.what());
  }
}

extern "C" int odis_surface_beta(int surface, int l_max, double radius, double g, double shell, double ocean,
                                 double *beta) {
  if (!beta || l_max < 0 || l_max > 1000) return host_fail(ODIS_E_INVALID, "bad argument");
  try {
    const std::vector<double> b = odis::surface_beta(surface, l_max, radius, g, shell, ocean);
    for (int l = 0; l <= l_max; l++) beta[l] = b[l];
    return ODIS_OK;
  } catch (const std::exception &e) {
    return host_fail(ODIS_E_INVALID, e.what());
  }
}

extern "C" int odis_analytical_obliq_west(const odis_params *p, double lat, double lon, double t, double *out) {
  if (!p || !out) return host_fail(ODIS_E_INVALID, "null argument");
  if (!(p->radius > 0) || !(p->omega > 0) || !(p->g > 0) || !(p->h > 0))
    return host_fail(ODIS_E_INVALID, "radius, omega, g and h must be positive");
  odis::analytical_obliq_west(p->radius, p->omega, p->g, p->h, p->obliquity, p->alpha, lat, lon, t, out);
  return ODIS_OK;
}

extern "C" int odis_sh_basis_eval(int l_max, int n, const double *lat, const double *lon, double *Y) {
  if (n < 0 || (n > 0 && (!lat || !lon || !Y)) || l_max < 0 || l_max > 200)
    return host_fail(ODIS_E_INVALID, "bad argument");
  try {
    const size_t K = size_t(l_max + 1) * size_t(l_max + 1);
    std::vector<double> a, b, sect, f1;
    for (int i = 0; i < n; i++) odis::sh_basis_row(l_max, lat[i], lon[i], Y + size_t(i) * K);
    return ODIS_OK;
  } catch (const std::exception &e) {
    return host_fail(ODIS_E_INVALID, e.what());
  }
}
