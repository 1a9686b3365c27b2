"""geodesicodis_amd -- GeodesicODIS's Laplace-Tidal-Equation time step on MI355X.

Python binding (ctypes) of ``libgeodesicodis_amd.so``, whose C ABI is
declared in ``include/odis.h``.  The library holds the HIP kernels (gfx950)
and the host C++ that mirrors the reference's program flow
(Globals -> Mesh -> solveODIS -> ab3Explicit).

There is no CPU fallback: every stepping entry point runs the HIP kernels and
raises if the library or a GPU is unavailable.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgeodesicodis_amd.so")

FP64, FP32 = 0, 1
OBLIQ, OBLIQ_WEST, OBLIQ_EAST, ECC_RAD, ECC_LIB, ECC, ECC_WEST, ECC_EAST, FULL, FULL2, \
    TOTAL, ECC_W3, OBLIQ_W3, PLANET, PLANET_OBL, GENERAL, NONE = range(17)
FREE, FREE_LOADING, LID_LOVE, LID_MEMBR, LID_NUM, LID_INF = range(6)
LINEAR, QUADRATIC = range(2)

ERRORS = {0: "OK", -1: "ODIS_E_INVALID", -2: "ODIS_E_HIP", -3: "ODIS_E_NONFINITE",
          -4: "ODIS_E_UNSUPPORTED", -5: "ODIS_E_NOMEM", -6: "ODIS_E_RANGE", -7: "ODIS_E_STOPPED",
          -8: "ODIS_E_IO"}

# Symbols declared in include/odis.h (kept in sync by tests/test_host_and_abi.py).
EXPORTS = [
    "odis_create", "odis_destroy", "odis_last_error", "odis_set_state", "odis_get_state",
    "odis_step", "odis_sync", "odis_get_dissipation", "odis_get_dissipation_series",
    "odis_series_capacity", "odis_get_velocity_en", "odis_request_stop", "odis_enable_timing",
    "odis_get_timing", "odis_grid_read", "odis_grid_generate", "odis_grid_write", "odis_grid_size",
    "odis_grid_arrays", "odis_grid_from_arrays", "odis_grid_free", "odis_mesh_build",
    "odis_mesh_sizes", "odis_mesh_describe", "odis_mesh_get", "odis_mesh_free",
    "odis_round_period", "odis_quantise_dt", "odis_config_read", "odis_host_last_error",
    "odis_run_case", "odis_interp_ll_set", "odis_interp_ll", "odis_sh_set_basis",
    "odis_sh_project", "odis_sh_synthesize", "odis_surface_beta", "odis_set_option",
    "odis_get_option", "odis_mesh_compact_check", "odis_get_kernel_times", "odis_rbf_set",
    "odis_velocity_xyz", "odis_h5_open", "odis_h5_write", "odis_h5_close", "odis_h5_last_error",
    "odis_analytical_obliq_west", "odis_sh_project_noalloc", "odis_orbit_stats", "odis_retire",
    "odis_sh_basis_eval", "odis_mlp_create", "odis_mlp_destroy", "odis_mlp_param_count",
    "odis_mlp_get_params", "odis_mlp_set_params", "odis_mlp_forward", "odis_mlp_set_data",
    "odis_mlp_loss_grad", "odis_mlp_train", "odis_mlp_last_error",
]
H5_DISPLACEMENT, H5_VELOCITY, H5_VELOCITY_XYZ, H5_DISSIPATION, H5_DISS_AVG = 1, 2, 4, 8, 16
SH_OFF, SH_FP64, SH_SPLIT_BF16, SH_RECUR, SH_MFMA = 0, 1, 2, 3, 4
OPT_GRAPHS, OPT_COMPACT, OPT_ONE_KERNEL, OPT_PATCH_CELLS, OPT_STEP_BLOCKS, OPT_STEP_LANES, OPT_SH_CELLS = \
    1, 2, 3, 4, 5, 6, 7
OPT_INTERLEAVE, OPT_INTERLEAVE_ACTIVE, OPT_IL_UNROLL, OPT_IL_ORDER, OPT_IL_RECORDS = 8, 9, 10, 11, 12


class OdisError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Params(ctypes.Structure):
    """odis_params (include/odis.h)."""
    _fields_ = [(n, ctypes.c_double) for n in
                ("g", "h", "alpha", "omega", "radius", "love_reduct", "ecc", "obliquity", "dt",
                 "shell_thickness", "semimajor")] + \
               [(n, ctypes.c_int) for n in
                ("tide", "surface", "friction", "init_load", "quad_drag", "reserved")]

    def to_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MeshDesc(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int), ("n_faces", ctypes.c_int),
                ("face_nodes", ctypes.c_void_p), ("grad_coef", ctypes.c_void_p),
                ("trisk_count", ctypes.c_void_p), ("trisk_faces", ctypes.c_void_p),
                ("trisk_coef", ctypes.c_void_p), ("face_lat", ctypes.c_void_p),
                ("face_area", ctypes.c_void_p), ("cell_faces", ctypes.c_void_p),
                ("div_coef", ctypes.c_void_p), ("cell_lat", ctypes.c_void_p),
                ("cell_lon", ctypes.c_void_p), ("face_len", ctypes.c_void_p),
                ("face_dist", ctypes.c_void_p), ("cell_area", ctypes.c_void_p)]


_lib = None


def lib():
    """Load the in-tree library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("geodesicodis_amd: %s is missing; run __graft_entry__.build() "
                              "(make -C geodesicodis_amd)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        llp = ctypes.POINTER(ctypes.c_longlong)
        vpp = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "odis_create": ([ctypes.POINTER(MeshDesc), ctypes.POINTER(Params), ctypes.c_int,
                             ctypes.c_int, vpp], ctypes.c_int),
            "odis_destroy": ([vp], ctypes.c_int),
            "odis_last_error": ([vp], ctypes.c_char_p),
            "odis_set_state": ([vp, ctypes.c_int, dp, dp, dp, dp, ctypes.c_longlong], ctypes.c_int),
            "odis_get_state": ([vp, ctypes.c_int, dp, dp, dp, dp, llp], ctypes.c_int),
            "odis_step": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
            "odis_sync": ([vp], ctypes.c_int),
            "odis_get_dissipation": ([vp, dp], ctypes.c_int),
            "odis_get_dissipation_series": ([vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, dp],
                                            ctypes.c_int),
            "odis_series_capacity": ([vp], ctypes.c_int),
            "odis_orbit_stats": ([vp, ctypes.c_int, ip, ip, dp], ctypes.c_int),
            "odis_retire": ([vp, ip, ctypes.c_int], ctypes.c_int),
            "odis_get_velocity_en": ([vp, ctypes.c_int, dp, dp], ctypes.c_int),
            "odis_request_stop": ([vp], ctypes.c_int),
            "odis_enable_timing": ([vp, ctypes.c_int], ctypes.c_int),
            "odis_set_option": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
            "odis_get_option": ([vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
            "odis_mesh_compact_check": ([vp, vp, vp, vp], ctypes.c_int),
            "odis_get_kernel_times": ([vp, vp, ctypes.c_int], ctypes.c_int),
            "odis_get_timing": ([vp, dp, dp, dp], ctypes.c_int),
            "odis_grid_read": ([ctypes.c_char_p, vpp], ctypes.c_int),
            "odis_grid_generate": ([ctypes.c_int, ctypes.c_int, vpp], ctypes.c_int),
            "odis_grid_write": ([vp, ctypes.c_char_p], ctypes.c_int),
            "odis_grid_size": ([vp], ctypes.c_int),
            "odis_grid_arrays": ([vp, dp, dp, ip, dp, dp], ctypes.c_int),
            "odis_grid_from_arrays": ([ctypes.c_int, dp, dp, ip, dp, dp, vpp], ctypes.c_int),
            "odis_grid_free": ([vp], ctypes.c_int),
            "odis_mesh_build": ([vp, ctypes.c_double, vpp], ctypes.c_int),
            "odis_mesh_sizes": ([vp, ip, ip, ip], ctypes.c_int),
            "odis_mesh_describe": ([vp, ctypes.POINTER(MeshDesc)], ctypes.c_int),
            "odis_mesh_get": ([vp, ctypes.c_char_p, vp], ctypes.c_int),
            "odis_mesh_free": ([vp], ctypes.c_int),
            "odis_round_period": ([ctypes.c_double, dp, dp], ctypes.c_int),
            "odis_quantise_dt": ([ctypes.c_double, ctypes.c_double, dp, ip], ctypes.c_int),
            "odis_config_read": ([ctypes.c_char_p, ctypes.POINTER(Params), ip, dp, ip, ip],
                                 ctypes.c_int),
            "odis_host_last_error": ([], ctypes.c_char_p),
            "odis_run_case": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
            "odis_interp_ll_set": ([vp, ctypes.c_int, ctypes.c_int, ip, ip, dp], ctypes.c_int),
            "odis_interp_ll": ([vp, ctypes.c_int, dp], ctypes.c_int),
            "odis_rbf_set": ([vp, ctypes.c_int, ctypes.c_int, ip, ip, dp], ctypes.c_int),
            "odis_velocity_xyz": ([vp, ctypes.c_int, dp], ctypes.c_int),
            "odis_h5_open": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp,
                              ctypes.POINTER(vp)], ctypes.c_int),
            "odis_h5_write": ([vp, ctypes.c_int, ctypes.c_int, dp], ctypes.c_int),
            "odis_h5_close": ([vp], ctypes.c_int),
            "odis_h5_last_error": ([], ctypes.c_char_p),
            "odis_sh_set_basis": ([vp, ctypes.c_int, dp, ctypes.c_int], ctypes.c_int),
            "odis_sh_project": ([vp, dp, dp], ctypes.c_int),
            "odis_sh_project_noalloc": ([vp, dp, dp], ctypes.c_int),
            "odis_sh_synthesize": ([vp, dp, dp], ctypes.c_int),
            "odis_surface_beta": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, dp], ctypes.c_int),
            "odis_sh_basis_eval": ([ctypes.c_int, ctypes.c_int, dp, dp, dp], ctypes.c_int),
            "odis_mlp_create": ([ctypes.c_int] * 4 + [ctypes.c_ulonglong, vpp], ctypes.c_int),
            "odis_mlp_destroy": ([vp], ctypes.c_int),
            "odis_mlp_param_count": ([vp], ctypes.c_longlong),
            "odis_mlp_get_params": ([vp, fp], ctypes.c_int),
            "odis_mlp_set_params": ([vp, fp], ctypes.c_int),
            "odis_mlp_forward": ([vp, ctypes.c_int, fp, fp], ctypes.c_int),
            "odis_mlp_set_data": ([vp, ctypes.c_int, fp, fp], ctypes.c_int),
            "odis_mlp_loss_grad": ([vp, ctypes.c_int, ip, dp, fp], ctypes.c_int),
            "odis_mlp_train": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                ctypes.c_ulonglong, dp], ctypes.c_int),
            "odis_mlp_last_error": ([], ctypes.c_char_p),
            "odis_analytical_obliq_west": ([ctypes.POINTER(Params), ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, dp], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _host_check(rc):
    if rc < 0:
        raise OdisError(rc, lib().odis_host_last_error().decode())
    return rc


# ---------------------------------------------------------------------------
# grid and mesh (host)
# ---------------------------------------------------------------------------

class Grid:
    """Grid in the reference's file layout (src/mesh.cpp:4016-4101)."""

    def __init__(self, handle):
        self._h = handle
        self.n = lib().odis_grid_size(handle)

    @classmethod
    def generate(cls, level, lloyd_iters=64):
        h = ctypes.c_void_p()
        _host_check(lib().odis_grid_generate(int(level), int(lloyd_iters), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def read(cls, path):
        h = ctypes.c_void_p()
        _host_check(lib().odis_grid_read(str(path).encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, lat, lon, friends, clat, clon):
        lat = np.ascontiguousarray(lat, np.float64)
        lon = np.ascontiguousarray(lon, np.float64)
        fr = np.ascontiguousarray(friends, np.int32)
        clat = np.ascontiguousarray(clat, np.float64)
        clon = np.ascontiguousarray(clon, np.float64)
        h = ctypes.c_void_p()
        _host_check(lib().odis_grid_from_arrays(len(lat), _dp(lat), _dp(lon), _ip(fr), _dp(clat),
                                                _dp(clon), ctypes.byref(h)))
        return cls(h)

    def arrays(self):
        n = self.n
        lat, lon = np.zeros(n), np.zeros(n)
        fr = np.zeros((n, 6), np.int32)
        clat, clon = np.zeros((n, 6)), np.zeros((n, 6))
        lib().odis_grid_arrays(self._h, _dp(lat), _dp(lon), _ip(fr), _dp(clat), _dp(clon))
        return dict(lat=lat, lon=lon, friends=fr, clat=clat, clon=clon)

    def write(self, path):
        _host_check(lib().odis_grid_write(self._h, str(path).encode()))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().odis_grid_free(self._h)
            self._h = None


_MESH_INT = {"faces", "node_face_dir", "face_nodes", "interp_count", "interp_friends", "rbf_row_ptr",
             "rbf_col"}


class Mesh:
    """TRiSK C-grid operands built on the host (Mesh::Mesh, src/mesh.cpp:32-197)."""

    def __init__(self, grid, radius):
        h = ctypes.c_void_p()
        _host_check(lib().odis_mesh_build(grid._h, float(radius), ctypes.byref(h)))
        self._h = h
        n, e, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().odis_mesh_sizes(h, ctypes.byref(n), ctypes.byref(e), ctypes.byref(v))
        self.N, self.E, self.V = n.value, e.value, v.value
        self.radius = float(radius)

    def get(self, name):
        cnt = _host_check(lib().odis_mesh_get(self._h, name.encode(), None))
        a = np.zeros(cnt, np.int32 if name in _MESH_INT else np.float64)
        lib().odis_mesh_get(self._h, name.encode(), a.ctypes.data_as(ctypes.c_void_p))
        return a

    def describe(self):
        d = MeshDesc()
        lib().odis_mesh_describe(self._h, ctypes.byref(d))
        return d

    def rbf_operator(self):
        """The reference's operatorRBFinterp (src/mesh.cpp:2263-2361) as CSR
        (row_ptr [3N + 1], col, val): rows 3i + x/y/z, face columns."""
        return self.get("rbf_row_ptr"), self.get("rbf_col"), self.get("rbf_val")

    def compact_check(self, desc=None):
        """(ok, reason, coef_err, div_err, area_err) of the cell-local TRiSK form
        (odis_mesh_compact_check)."""
        d = desc if desc is not None else self.describe()
        errs = [ctypes.c_double() for _ in range(3)]
        rc = lib().odis_mesh_compact_check(ctypes.byref(d), *[ctypes.byref(e) for e in errs])
        reason = "" if rc == 0 else lib().odis_host_last_error().decode()
        return (rc == 0, reason) + tuple(e.value for e in errs)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().odis_mesh_free(self._h)
            self._h = None


# ---------------------------------------------------------------------------
# configuration
# ---------------------------------------------------------------------------

def round_period(ang_vel):
    p, o = ctypes.c_double(), ctypes.c_double()
    _host_check(lib().odis_round_period(float(ang_vel), ctypes.byref(p), ctypes.byref(o)))
    return p.value, o.value


def quantise_dt(period, target):
    d, n = ctypes.c_double(), ctypes.c_int()
    _host_check(lib().odis_quantise_dt(float(period), float(target), ctypes.byref(d), ctypes.byref(n)))
    return d.value, n.value


def read_config(case_dir):
    """Parse input.in (Globals semantics); returns (Params, info dict)."""
    p = Params()
    steps, lvl, outt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    end = ctypes.c_double()
    _host_check(lib().odis_config_read(str(case_dir).encode(), ctypes.byref(p), ctypes.byref(steps),
                                       ctypes.byref(end), ctypes.byref(lvl), ctypes.byref(outt)))
    return p, dict(steps_per_orbit=steps.value, end_time=end.value, grid_level=lvl.value,
                   output_time=outt.value)


def make_params(radius=252.1e3, g=0.113, ang_vel=5.307e-5, h=38e3, alpha=1e-8, ecc=0.0047,
                obliquity=0.0, tide=ECC, surface=FREE, friction=LINEAR, love_reduct=1.0,
                time_step=30.0, shell_thickness=0.0, semimajor=0.0, init_load=0, quad_drag=0):
    """Params with the reference's derived values: period rounding and dt
    quantisation (src/globals.cpp:209-214, src/mesh.cpp:1566-1623).
    Defaults are the Enceladus free-surface ocean of the plain-method case."""
    period, omega = round_period(ang_vel)
    dt, n = quantise_dt(period, time_step)
    p = Params(g=g, h=h, alpha=alpha, omega=omega, radius=radius, love_reduct=love_reduct,
               ecc=ecc, obliquity=obliquity, dt=dt, shell_thickness=shell_thickness,
               semimajor=semimajor, tide=tide, surface=surface, friction=friction,
               init_load=init_load, quad_drag=quad_drag)
    return p, n


class H5File:
    """DATA/data.h5 writer in the reference's layout (odis_h5_open / write /
    close; src/outFiles.cpp:138-681)."""

    def __init__(self, path, n_cells, n_faces, n_slices, fields, face_lat=None, face_lon=None):
        h = ctypes.c_void_p()
        fl = None if face_lat is None else np.ascontiguousarray(face_lat, np.float64)
        fo = None if face_lon is None else np.ascontiguousarray(face_lon, np.float64)
        rc = lib().odis_h5_open(str(path).encode(), int(n_cells), int(n_faces), int(n_slices), int(fields),
                                None if fl is None else _dp(fl), None if fo is None else _dp(fo), ctypes.byref(h))
        if rc:
            raise OdisError(rc, (lib().odis_h5_last_error() or b"").decode())
        self._h = h

    def write(self, field, slice_, data):
        d = np.ascontiguousarray(data, np.float64).ravel()
        rc = lib().odis_h5_write(self._h, int(field), int(slice_), _dp(d))
        if rc:
            raise OdisError(rc, "odis_h5_write")

    def close(self):
        if getattr(self, "_h", None):
            lib().odis_h5_close(self._h)
            self._h = None

    def __del__(self):
        self.close()


# ---------------------------------------------------------------------------
# device context (the time loop)
# ---------------------------------------------------------------------------

class Context:
    """GPU LTE stepper over one mesh and one or more instances
    (odis_create ... odis_destroy)."""

    def __init__(self, mesh, params, precision=FP64):
        if isinstance(params, Params):
            params = [params]
        self.mesh = mesh
        self.B = len(params)
        arr = (Params * self.B)(*params)
        self.params = list(params)
        d = mesh.describe()
        h = ctypes.c_void_p()
        rc = lib().odis_create(ctypes.byref(d), arr, self.B, int(precision), ctypes.byref(h))
        if rc:
            raise OdisError(rc, lib().odis_last_error(None).decode())
        self._h = h
        self.precision = precision
        self._ll_rows = 0

    def _check(self, rc):
        if rc:
            raise OdisError(rc, lib().odis_last_error(self._h).decode())

    def set_state(self, inst=0, v=None, dvdt=None, eta=None, detadt=None, iter=0):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64).ravel()
                for a in (v, dvdt, eta, detadt)]
        self._check(lib().odis_set_state(self._h, int(inst), *[_dp(a) for a in arrs], int(iter)))

    def get_state(self, inst=0):
        m = self.mesh
        v, dv = np.zeros(m.E), np.zeros(m.E * 3)
        eta, de = np.zeros(m.N), np.zeros(m.N * 3)
        it = ctypes.c_longlong()
        self._check(lib().odis_get_state(self._h, int(inst), _dp(v), _dp(dv), _dp(eta), _dp(de),
                                         ctypes.byref(it)))
        return dict(v=v, dvdt=dv.reshape(m.E, 3), eta=eta, detadt=de.reshape(m.N, 3),
                    iter=it.value)

    def step(self, n, sync=True):
        self._check(lib().odis_step(self._h, int(n), 1 if sync else 0))

    def sync(self):
        self._check(lib().odis_sync(self._h))

    def dissipation(self):
        out = np.zeros(self.B)
        self._check(lib().odis_get_dissipation(self._h, _dp(out)))
        return out

    def dissipation_series(self, inst, first, count):
        out = np.zeros(count)
        self._check(lib().odis_get_dissipation_series(self._h, int(inst), int(first), int(count),
                                                      _dp(out)))
        return out

    def orbit_stats(self, inst, steps_per_orbit):
        """(mean dissipation, v residual, eta residual) per selected instance
        whose orbit of steps_per_orbit steps ended at the current step
        (odis_orbit_stats); an (n, 3) array."""
        sel = np.ascontiguousarray(inst, np.int32)
        ns = np.ascontiguousarray(np.broadcast_to(np.asarray(steps_per_orbit, np.int32), sel.shape))
        out = np.zeros(3 * len(sel))
        self._check(lib().odis_orbit_stats(self._h, len(sel), _ip(sel), _ip(ns), _dp(out)))
        return out.reshape(-1, 3)

    def retire(self, keep):
        """Keep only the instances keep (increasing), renumbered 0.. in that
        order (odis_retire)."""
        k = np.ascontiguousarray(keep, np.int32)
        self._check(lib().odis_retire(self._h, _ip(k), len(k)))
        self.params = [self.params[i] for i in k]
        self.B = len(k)

    def velocity_en(self, inst=0):
        normal = np.ascontiguousarray(self.mesh.get("face_normal"))
        uv = np.zeros(self.mesh.E * 2)
        self._check(lib().odis_get_velocity_en(self._h, int(inst), _dp(normal), _dp(uv)))
        return uv.reshape(-1, 2)

    def set_interp_ll(self, row_ptr, col, val, n_cols=None):
        """Attach the geodesic -> lat-lon CSR matrix (3N columns, reference
        cell numbering, interleaved [eta, r deta/dlat, r deta/dlon])."""
        rp = np.ascontiguousarray(row_ptr, np.int32)
        c = np.ascontiguousarray(col, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        rows = len(rp) - 1
        self._check(lib().odis_interp_ll_set(self._h, rows, int(n_cols if n_cols is not None else 3 * self.mesh.N),
                                             _ip(rp), _ip(c), _dp(v)))
        self._ll_rows = rows  # the matrix in use: only once the library took it

    def interp_ll(self, inst=0):
        out = np.zeros(max(1, self._ll_rows))
        self._check(lib().odis_interp_ll(self._h, int(inst), _dp(out)))
        return out[:self._ll_rows]

    def set_rbf(self, row_ptr=None, col=None, val=None):
        """Attach the RBF Cartesian-velocity operator (3N x E CSR, reference
        numbering); by default the one the mesh builds (operatorRBFinterp)."""
        if row_ptr is None:
            row_ptr, col, val = self.mesh.rbf_operator()
        rp = np.ascontiguousarray(row_ptr, np.int32)
        c = np.ascontiguousarray(col, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        self._check(lib().odis_rbf_set(self._h, len(rp) - 1, int(self.mesh.E), _ip(rp), _ip(c), _dp(v)))

    def velocity_xyz(self, inst=0):
        """Cartesian velocity [N, 3] at cell centres (reference numbering)."""
        out = np.zeros(self.mesh.N * 3)
        self._check(lib().odis_velocity_xyz(self._h, int(inst), _dp(out)))
        return out.reshape(-1, 3)

    def set_sh(self, l_max, beta=None, mode=SH_FP64):
        """Self-gravity / shell coupling: beta is (n_inst, l_max + 1) or
        (l_max + 1,) per-degree factors; mode SH_OFF / SH_FP64 / SH_SPLIT_BF16 /
        SH_RECUR (basis evaluated on the fly) / SH_MFMA (batched split-BF16
        matrix products)."""
        b = None
        if beta is not None:
            b = np.ascontiguousarray(np.broadcast_to(np.asarray(beta, np.float64),
                                                     (self.B, l_max + 1)))
        self.sh_l_max = l_max
        self._check(lib().odis_sh_set_basis(self._h, int(l_max), _dp(b), int(mode)))

    def sh_project(self, field):
        f = np.ascontiguousarray(field, np.float64)
        c = np.zeros((self.sh_l_max + 1) ** 2)
        self._check(lib().odis_sh_project(self._h, _dp(f), _dp(c)))
        return c

    def sh_project_noalloc(self, field):
        """sh_project through odis_sh_project_noalloc (no device allocation;
        scratch in a static device array shared by every call on the GPU)."""
        f = np.ascontiguousarray(field, np.float64)
        c = np.zeros((self.sh_l_max + 1) ** 2)
        self._check(lib().odis_sh_project_noalloc(self._h, _dp(f), _dp(c)))
        return c

    def sh_synthesize(self, coeffs):
        c = np.ascontiguousarray(coeffs, np.float64)
        f = np.zeros(self.mesh.N)
        self._check(lib().odis_sh_synthesize(self._h, _dp(c), _dp(f)))
        return f

    def request_stop(self):
        self._check(lib().odis_request_stop(self._h))

    def enable_timing(self, stride=1):
        """Stamp the kernels of every stride-th step with HIP events (0 =
        off); timing() reads them."""
        self._check(lib().odis_enable_timing(self._h, int(stride)))

    def set_option(self, option, value):
        self._check(lib().odis_set_option(self._h, int(option), int(value)))

    def option(self, option):
        v = ctypes.c_int()
        self._check(lib().odis_get_option(self._h, int(option), ctypes.byref(v)))
        return v.value

    def use_graphs(self, on=True):
        """Replay 64-step chunks from captured HIP graphs (ODIS_OPT_GRAPHS)."""
        self.set_option(OPT_GRAPHS, 1 if on else 0)

    def use_compact(self, on=True):
        """Cell-local TRiSK form (default when available) or the tabulated
        operator (ODIS_OPT_COMPACT)."""
        self.set_option(OPT_COMPACT, 1 if on else 0)

    def use_one_kernel(self, on=True):
        """One kernel per step over cell patches (default with the compact
        form) or a face and a cell kernel per step (ODIS_OPT_ONE_KERNEL)."""
        self.set_option(OPT_ONE_KERNEL, 1 if on else 0)

    def timing(self):
        """Average kernel times (ms) of the sampled steps of the last step()
        call (odis_get_kernel_times; waits for the context stream)."""
        t = np.zeros(6)
        self._check(lib().odis_get_kernel_times(self._h, _dp(t), 6))
        return dict(face_ms=t[0], cell_ms=t[1], step_ms=t[2], sh_analysis_ms=t[3], sh_coef_ms=t[4],
                    sh_synth_ms=t[5])

    def close(self):
        if getattr(self, "_h", None):
            lib().odis_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """The warm start's network on the GPU (odis_mlp_*): n_in -> width x
    depth (tanh) -> n_out, FP32, trained by the library's kernels."""

    def __init__(self, n_in, width, depth, n_out, seed=0):
        h = ctypes.c_void_p()
        self._mcheck(lib().odis_mlp_create(int(n_in), int(width), int(depth), int(n_out), int(seed),
                                           ctypes.byref(h)))
        self._h = h
        self.shape = (int(n_in), int(width), int(depth), int(n_out))
        self.n_params = int(lib().odis_mlp_param_count(h))

    @staticmethod
    def _mcheck(rc):
        if rc:
            raise OdisError(rc, lib().odis_mlp_last_error().decode())

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        self._mcheck(lib().odis_mlp_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError("expected %d parameters" % self.n_params)
        self._mcheck(lib().odis_mlp_set_params(self._h, _fp(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), self.shape[3]), np.float32)
        self._mcheck(lib().odis_mlp_forward(self._h, len(x), _fp(x), _fp(y)))
        return y

    def set_data(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        self._mcheck(lib().odis_mlp_set_data(self._h, len(x), _fp(x), _fp(y)))
        self.n_data = len(x)

    def loss_grad(self, idx=None):
        idx = np.arange(self.n_data, dtype=np.int32) if idx is None else np.ascontiguousarray(idx, np.int32)
        loss = ctypes.c_double()
        g = np.zeros(self.n_params, np.float32)
        self._mcheck(lib().odis_mlp_loss_grad(self._h, len(idx), _ip(idx), ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, epochs, batch, lr_start=1e-3, lr_end=1e-5, seed=0):
        out = np.zeros(int(epochs))
        self._mcheck(lib().odis_mlp_train(self._h, int(epochs), int(batch), float(lr_start), float(lr_end),
                                          int(seed), _dp(out)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().odis_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def surface_beta(surface, l_max, top_radius, g, shell_thickness, ocean_thickness):
    """Per-degree pressure factors beta_l of a surface type (membraneNuBeta for
    LID_MEMBR, 1 - loading factor for FREE_LOADING, else 1)."""
    b = np.zeros(l_max + 1)
    _host_check(lib().odis_surface_beta(int(surface), int(l_max), float(top_radius), float(g),
                                        float(shell_thickness), float(ocean_thickness), _dp(b)))
    return b


def analytical_obliq_west(params, lat, lon, t):
    """(U, V, dU/dt, dV/dt, eta, deta/dt) of the OBLIQ_WEST analytic tide
    (analyticalLTE, src/analyticalLTE.cpp:133-180) at (lat, lon, t)."""
    o = np.zeros(6)
    _host_check(lib().odis_analytical_obliq_west(ctypes.byref(params), float(lat), float(lon), float(t), _dp(o)))
    return o


def sh_basis(l_max, lat, lon):
    """Real 4pi-normalised harmonics Y[i, l*l + l + m] at the points (lat,
    lon) radians (odis_sh_basis_eval): the basis of the SH coupling."""
    la = np.ascontiguousarray(lat, np.float64)
    lo = np.ascontiguousarray(lon, np.float64)
    Y = np.zeros((len(la), (l_max + 1) ** 2))
    _host_check(lib().odis_sh_basis_eval(int(l_max), len(la), _dp(la), _dp(lo), _dp(Y)))
    return Y


def run_case(case_dir, grid_dir=None, precision=FP64):
    """solveODIS-shaped driver over a case directory holding input.in."""
    rc = lib().odis_run_case(str(case_dir).encode(),
                             None if grid_dir is None else str(grid_dir).encode(), int(precision))
    if rc < 0:
        raise OdisError(rc, "see %s/DATA/ERROR.txt" % case_dir)
    return rc
