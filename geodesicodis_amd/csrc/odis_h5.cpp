// odis_h5.cpp -- the reference's DATA/data.h5 writer (OutFiles::
// CreateHDF5Framework, DumpGridData and DumpData, src/outFiles.cpp:138-681):
// one float32 dataset per output field of shape [time_slices][cells or
// faces] (time_slices = orbits x outputs per orbit + 1), the slice of output
// k (1-based) written at row k - 1, a 1-D [time_slices] dataset for the
// average dissipation, and the face centres in degrees as 1-D "face
// longitude" / "face latitude".  Dataset names are the reference's, so its
// post-processing (python_scripts/FieldPlot2.py) reads the file unchanged.
//
// The HDF5 C library is loaded with dlopen on the first open (ODIS_HDF5_LIB,
// else libhdf5.so.103, libhdf5.so, /opt/conda/lib/libhdf5.so[.103]), so the
// library itself does not depend on HDF5; without it the entry points return
// ODIS_E_UNSUPPORTED.
#include <dlfcn.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <string>
#include <vector>

#include "odis.h"
#include "odis_host.hpp"

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr hid_t kDefault = 0;        // H5P_DEFAULT
constexpr unsigned kAccTrunc = 2u;   // H5F_ACC_TRUNC
constexpr int kSelectSet = 0;        // H5S_SELECT_SET

struct H5Api {
  void *lib = nullptr;
  herr_t (*open)(void) = nullptr;
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t) = nullptr;
  herr_t (*fclose)(hid_t) = nullptr;
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *) = nullptr;
  herr_t (*sclose)(hid_t) = nullptr;
  herr_t (*select_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *,
                             const hsize_t *) = nullptr;
  hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *) = nullptr;
  herr_t (*dclose)(hid_t) = nullptr;
  hid_t *native_float = nullptr;  // H5T_NATIVE_FLOAT_g, valid after H5open
};

H5Api g_api;
std::string g_h5_err;

bool load_api() {
  if (g_api.lib) return true;
  std::vector<std::string> names;
  if (const char *e = std::getenv("ODIS_HDF5_LIB")) names.push_back(e);
  for (const char *n : {"libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so"})
    names.push_back(n);
  void *h = nullptr;
  for (const std::string &n : names)
    if ((h = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
  if (!h) {
    g_h5_err = "HDF5 C library not found (set ODIS_HDF5_LIB)";
    return false;
  }
  H5Api a;
  a.lib = h;
  bool ok = true;
  auto sym = [&](const char *nm) {
    void *p = dlsym(h, nm);
    if (!p) ok = false;
    return p;
  };
  a.open = reinterpret_cast<herr_t (*)(void)>(sym("H5open"));
  a.fcreate = reinterpret_cast<hid_t (*)(const char *, unsigned, hid_t, hid_t)>(sym("H5Fcreate"));
  a.fclose = reinterpret_cast<herr_t (*)(hid_t)>(sym("H5Fclose"));
  a.screate_simple = reinterpret_cast<hid_t (*)(int, const hsize_t *, const hsize_t *)>(sym("H5Screate_simple"));
  a.sclose = reinterpret_cast<herr_t (*)(hid_t)>(sym("H5Sclose"));
  a.select_hyperslab = reinterpret_cast<herr_t (*)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *,
                                                   const hsize_t *)>(sym("H5Sselect_hyperslab"));
  a.dcreate2 = reinterpret_cast<hid_t (*)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t)>(sym("H5Dcreate2"));
  a.dwrite = reinterpret_cast<herr_t (*)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *)>(sym("H5Dwrite"));
  a.dclose = reinterpret_cast<herr_t (*)(hid_t)>(sym("H5Dclose"));
  a.native_float = reinterpret_cast<hid_t *>(sym("H5T_NATIVE_FLOAT_g"));
  if (!ok || a.open() < 0) {
    dlclose(h);
    g_h5_err = "HDF5 C library lacks the entry points used";
    return false;
  }
  g_api = a;
  return true;
}

// One dataset: rank 1 ([slices]) or 2 ([slices][n]).
struct Dataset {
  hid_t set = -1, space = -1, mem = -1;
  int n = 0;
};

}  // namespace

struct odis_h5 {
  hid_t file = -1;
  int N = 0, E = 0, slices = 0, fields = 0;
  Dataset eta, u, v, ux, uy, uz, diss, diss_avg;
};

namespace {

bool make_set(odis_h5 *h, Dataset &d, const char *name, int n) {
  const H5Api &a = g_api;
  d.n = n;
  hsize_t dims[2] = {hsize_t(h->slices), hsize_t(n)}, one[2] = {1, hsize_t(n)};
  const int rank = n > 0 ? 2 : 1;
  d.space = a.screate_simple(rank, dims, nullptr);
  d.mem = a.screate_simple(rank, one, nullptr);
  if (d.space < 0 || d.mem < 0) return false;
  d.set = a.dcreate2(h->file, name, *a.native_float, d.space, kDefault, kDefault, kDefault);
  return d.set >= 0;
}

bool write_row(odis_h5 *h, Dataset &d, int slice, const std::vector<float> &row) {
  const H5Api &a = g_api;
  hsize_t start[2] = {hsize_t(slice), 0}, count[2] = {1, hsize_t(d.n)};
  if (a.select_hyperslab(d.space, kSelectSet, start, nullptr, count, nullptr) < 0) return false;
  return a.dwrite(d.set, *a.native_float, d.mem, d.space, kDefault, row.data()) >= 0;
}

void close_set(Dataset &d) {
  const H5Api &a = g_api;
  if (d.set >= 0) a.dclose(d.set);
  if (d.space >= 0) a.sclose(d.space);
  if (d.mem >= 0) a.sclose(d.mem);
  d = Dataset();
}

void close_all(odis_h5 *h) {
  for (Dataset *d : {&h->eta, &h->u, &h->v, &h->ux, &h->uy, &h->uz, &h->diss, &h->diss_avg}) close_set(*d);
  if (h->file >= 0) g_api.fclose(h->file);
  h->file = -1;
}

// DumpGridData (src/outFiles.cpp:464-520): face centres in degrees, the
// value cast to float before the scaling as the reference does.
bool write_face_positions(odis_h5 *h, const double *face_lat, const double *face_lon) {
  const H5Api &a = g_api;
  const double deg = 180. / odis::kPi;
  for (int q = 0; q < 2; q++) {
    std::vector<float> x(h->E);
    for (int e = 0; e < h->E; e++) x[e] = float(double(float(q == 0 ? face_lon[e] : face_lat[e])) * deg);
    hsize_t dims[1] = {hsize_t(h->E)};
    const hid_t sp = a.screate_simple(1, dims, nullptr);
    const hid_t ds = a.dcreate2(h->file, q == 0 ? "face longitude" : "face latitude", *a.native_float, sp, kDefault,
                                kDefault, kDefault);
    const bool ok = sp >= 0 && ds >= 0 && a.dwrite(ds, *a.native_float, sp, sp, kDefault, x.data()) >= 0;
    if (ds >= 0) a.dclose(ds);
    if (sp >= 0) a.sclose(sp);
    if (!ok) return false;
  }
  return true;
}

}  // namespace

static int h5_open_impl(const char *path, int n_cells, int n_faces, int n_slices, int fields,
                        const double *face_lat, const double *face_lon, odis_h5 **out) {
  if (!path || !out || n_cells < 1 || n_faces < 1 || n_slices < 1) return ODIS_E_INVALID;
  *out = nullptr;
  if (!load_api()) return ODIS_E_UNSUPPORTED;
  odis_h5 *h = new odis_h5;
  h->N = n_cells;
  h->E = n_faces;
  h->slices = n_slices;
  h->fields = fields;
  h->file = g_api.fcreate(path, kAccTrunc, kDefault, kDefault);
  bool ok = h->file >= 0;
  if (ok && (fields & ODIS_H5_VELOCITY))
    ok = make_set(h, h->u, "east velocity", n_faces) && make_set(h, h->v, "north velocity", n_faces);
  if (ok && (fields & ODIS_H5_VELOCITY_XYZ))
    ok = make_set(h, h->ux, "x velocity", n_cells) && make_set(h, h->uy, "y velocity", n_cells) &&
         make_set(h, h->uz, "z velocity", n_cells);
  if (ok && (fields & ODIS_H5_DISPLACEMENT)) ok = make_set(h, h->eta, "displacement", n_cells);
  if (ok && (fields & ODIS_H5_DISSIPATION)) ok = make_set(h, h->diss, "dissipated energy", n_faces);
  if (ok && (fields & ODIS_H5_DISS_AVG)) ok = make_set(h, h->diss_avg, "dissipation avg output", 0);
  if (ok && face_lat && face_lon) ok = write_face_positions(h, face_lat, face_lon);
  if (!ok) {
    close_all(h);
    delete h;
    g_h5_err = std::string("cannot create ") + path;
    return ODIS_E_IO;
  }
  *out = h;
  return ODIS_OK;
}

extern "C" int odis_h5_open(const char *path, int n_cells, int n_faces, int n_slices, int fields,
                            const double *face_lat, const double *face_lon, odis_h5 **out) {
  try {
    return h5_open_impl(path, n_cells, n_faces, n_slices, fields, face_lat, face_lon, out);
  } catch (const std::exception &) {  // host allocation; nothing is thrown across the C ABI
    return ODIS_E_NOMEM;
  }
}

static int h5_write_impl(odis_h5 *h, int field, int slice, const double *data) {
  if (!h || !data) return ODIS_E_INVALID;
  if (slice < 0 || slice >= h->slices) return ODIS_E_RANGE;
  if (!(h->fields & field)) return ODIS_E_INVALID;
  bool ok = true;
  if (field == ODIS_H5_VELOCITY) {  // data [E][2] (east, north)
    std::vector<float> e(h->E), n(h->E);
    for (int f = 0; f < h->E; f++) {
      e[f] = float(data[2 * f]);
      n[f] = float(data[2 * f + 1]);
    }
    ok = write_row(h, h->u, slice, e) && write_row(h, h->v, slice, n);
  } else if (field == ODIS_H5_VELOCITY_XYZ) {  // data [N][3]
    std::vector<float> x(h->N), y(h->N), z(h->N);
    for (int i = 0; i < h->N; i++) {
      x[i] = float(data[3 * i]);
      y[i] = float(data[3 * i + 1]);
      z[i] = float(data[3 * i + 2]);
    }
    ok = write_row(h, h->ux, slice, x) && write_row(h, h->uy, slice, y) && write_row(h, h->uz, slice, z);
  } else if (field == ODIS_H5_DISPLACEMENT || field == ODIS_H5_DISSIPATION) {
    Dataset &d = field == ODIS_H5_DISPLACEMENT ? h->eta : h->diss;
    std::vector<float> x(d.n);
    for (int i = 0; i < d.n; i++) x[i] = float(data[i]);
    ok = write_row(h, d, slice, x);
  } else if (field == ODIS_H5_DISS_AVG) {
    hsize_t start[1] = {hsize_t(slice)}, count[1] = {1};
    const float x = float(data[0]);
    ok = g_api.select_hyperslab(h->diss_avg.space, kSelectSet, start, nullptr, count, nullptr) >= 0 &&
         g_api.dwrite(h->diss_avg.set, *g_api.native_float, h->diss_avg.mem, h->diss_avg.space, kDefault, &x) >= 0;
  } else {
    return ODIS_E_INVALID;
  }
  return ok ? ODIS_OK : ODIS_E_IO;
}

extern "C" int odis_h5_write(odis_h5 *h, int field, int slice, const double *data) {
  try {
    return h5_write_impl(h, field, slice, data);
  } catch (const std::exception &) {  // the float staging vectors
    return ODIS_E_NOMEM;
  }
}

extern "C" int odis_h5_close(odis_h5 *h) {
  if (!h) return ODIS_E_INVALID;
  close_all(h);
  delete h;
  return ODIS_OK;
}

extern "C" const char *odis_h5_last_error(void) { return g_h5_err.c_str(); }
