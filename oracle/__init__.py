"""ctypes binding of the CPU FP64 oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this module.  The product package never does.

The C source restates the reference's mesh assembly and AB3 step
(src/mesh.cpp, src/timeIntegrator.cpp, src/updateMomentum.cpp,
src/updateEta.cpp, src/temporalOperators.cpp, src/tidalPotentials.cpp,
src/interpolation.cpp:26-62, src/energy.cpp:13-62).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OBLIQ, OBLIQ_WEST, OBLIQ_EAST, ECC_RAD, ECC_LIB, ECC, ECC_WEST, ECC_EAST, FULL, FULL2, \
    TOTAL, ECC_W3, OBLIQ_W3, PLANET, PLANET_OBL, GENERAL, NONE = range(17)
FREE, FREE_LOADING, LID_LOVE, LID_MEMBR, LID_NUM, LID_INF = range(6)
LINEAR, QUADRATIC = range(2)


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("g", "h", "alpha", "omega", "radius", "love_reduct", "ecc", "obliquity",
                 "dt", "shell_thickness", "semimajor")] + \
               [(n, ctypes.c_int) for n in
                ("tide", "surface", "friction", "init_load", "quad_drag_momentum", "pad_")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.orc_mesh_build.restype = vp
        L.orc_mesh_build.argtypes = [ctypes.c_int, dp, dp, ip, dp, dp, ctypes.c_double, ip]
        L.orc_mesh_free.argtypes = [vp]
        L.orc_mesh_sizes.argtypes = [vp, ip, ip, ip]
        L.orc_mesh_get.argtypes = [vp, ctypes.c_char_p, vp]
        L.orc_mesh_get.restype = ctypes.c_int
        for nm in ("orc_mesh_grad_csr", "orc_mesh_div_csr"):
            getattr(L, nm).argtypes = [vp, ip, ip, dp]
        L.orc_mesh_coriolis_nnz.argtypes = [vp]
        L.orc_mesh_coriolis_csr.argtypes = [vp, ctypes.c_double, ip, ip, dp]
        L.orc_sim_create.restype = vp
        L.orc_sim_create.argtypes = [vp, ctypes.POINTER(OrcParams)]
        L.orc_sim_free.argtypes = [vp]
        L.orc_sim_set_state.argtypes = [vp, dp, dp, dp, dp, ctypes.c_int]
        L.orc_sim_get_state.argtypes = [vp, dp, dp, dp, dp, ip]
        L.orc_sim_step.argtypes = [vp, ctypes.c_int, dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_sim_dissipation.argtypes = [vp]
        L.orc_sim_dissipation.restype = ctypes.c_double
        L.orc_sim_forcing.argtypes = [vp, ctypes.c_double, dp]
        L.orc_sim_interp_velocity.argtypes = [vp, dp, dp]
        L.orc_sim_momentum.argtypes = [vp, dp, dp, dp]
        L.orc_sim_set_projector.argtypes = [vp, ctypes.c_int, dp, dp]
        L.orc_round_period.argtypes = [ctypes.c_double, dp, dp]
        L.orc_quantise_dt.argtypes = [ctypes.c_double, ctypes.c_double, dp, ip]
        L.orc_analytical_obliq_west.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, dp]
        L.orc_membrane_nu_beta.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, dp, dp, dp, dp]
        L.orc_mlp_forward.argtypes = [ctypes.c_int] * 4 + [dp, ctypes.c_int, dp, dp]
        L.orc_mlp_loss_grad.argtypes = [ctypes.c_int] * 4 + [dp, ctypes.c_int, dp, dp, dp]
        L.orc_mlp_loss_grad.restype = ctypes.c_double
        L.orc_membrane_nu_beta_gen.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] + \
            [ctypes.c_double] * 3 + [dp, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


_INT_ARRAYS = {"faces", "node_face_dir", "vertexes", "vertex_nodes", "face_nodes",
               "interp_count", "interp_friends"}


class Mesh:
    """Oracle mesh built from a grid (node lat/lon in radians, 6 friends,
    6 Voronoi vertices per node), src/mesh.cpp:32-197."""

    def __init__(self, grid, radius):
        L = lib()
        self.lat = np.ascontiguousarray(grid["lat"], dtype=np.float64)
        self.lon = np.ascontiguousarray(grid["lon"], dtype=np.float64)
        fr = np.ascontiguousarray(grid["friends"], dtype=np.int32)
        clat = np.ascontiguousarray(grid["clat"], dtype=np.float64)
        clon = np.ascontiguousarray(grid["clon"], dtype=np.float64)
        err = ctypes.c_int(0)
        self.h = L.orc_mesh_build(len(self.lat), _dp(self.lat), _dp(self.lon), _ip(fr),
                                  _dp(clat), _dp(clon), float(radius), ctypes.byref(err))
        if not self.h:
            raise RuntimeError("oracle mesh build failed: %d" % err.value)
        n, e, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.orc_mesh_sizes(self.h, ctypes.byref(n), ctypes.byref(e), ctypes.byref(v))
        self.N, self.E, self.V = n.value, e.value, v.value
        self.radius = float(radius)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None

    def get(self, name):
        L = lib()
        cnt = L.orc_mesh_get(self.h, name.encode(), None)
        if cnt < 0:
            raise KeyError(name)
        a = np.zeros(cnt, dtype=np.int32 if name in _INT_ARRAYS else np.float64)
        L.orc_mesh_get(self.h, name.encode(), a.ctypes.data_as(ctypes.c_void_p))
        return a

    def grad_csr(self):
        rp = np.zeros(self.E + 1, np.int32)
        col = np.zeros(2 * self.E, np.int32)
        val = np.zeros(2 * self.E)
        lib().orc_mesh_grad_csr(self.h, _ip(rp), _ip(col), _dp(val))
        return rp, col, val

    def div_csr(self):
        rp = np.zeros(self.N + 1, np.int32)
        col = np.zeros(6 * self.N, np.int32)
        val = np.zeros(6 * self.N)
        nnz = lib().orc_mesh_div_csr(self.h, _ip(rp), _ip(col), _dp(val))
        return rp, col[:nnz], val[:nnz]

    def coriolis_csr(self, omega):
        nnz = lib().orc_mesh_coriolis_nnz(self.h)
        rp = np.zeros(self.E + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        lib().orc_mesh_coriolis_csr(self.h, float(omega), _ip(rp), _ip(col), _dp(val))
        return rp, col, val


def make_params(**kw):
    p = OrcParams()
    defaults = dict(g=0.113, h=38e3, alpha=1e-8, omega=5.307e-5, radius=252.1e3,
                    love_reduct=1.0, ecc=0.0047, obliquity=0.0, dt=30.0,
                    shell_thickness=0.0, semimajor=0.0, tide=ECC, surface=FREE,
                    friction=LINEAR, init_load=0, quad_drag_momentum=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def set_threads(n):
    """OpenMP threads of the oracle's row loops (results do not depend on it)."""
    lib().orc_set_threads(int(n))


def threads():
    return int(lib().orc_get_threads())


class Sim:
    """Oracle AB3 stepper, src/timeIntegrator.cpp:57-322."""

    def __init__(self, mesh, params):
        self.mesh = mesh
        self.params = params
        self.h = lib().orc_sim_create(mesh.h, ctypes.byref(params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_free(self.h)
            self.h = None

    def set_state(self, v=None, dvdt=None, eta=None, detadt=None, iter=0):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (v, dvdt, eta, detadt)]
        lib().orc_sim_set_state(self.h, *[_dp(a) for a in arrs], int(iter))

    def get_state(self):
        m = self.mesh
        v, dv = np.zeros(m.E), np.zeros(m.E * 3)
        eta, de = np.zeros(m.N), np.zeros(m.N * 3)
        it = ctypes.c_int()
        lib().orc_sim_get_state(self.h, _dp(v), _dp(dv), _dp(eta), _dp(de), ctypes.byref(it))
        return dict(v=v, dvdt=dv.reshape(m.E, 3), eta=eta, detadt=de.reshape(m.N, 3),
                    iter=it.value)

    def step(self, n, record=False):
        d = np.zeros(n) if record else None
        lib().orc_sim_step(self.h, int(n), _dp(d))
        return d

    def dissipation(self):
        return lib().orc_sim_dissipation(self.h)

    def forcing(self, t):
        U = np.zeros(self.mesh.N)
        lib().orc_sim_forcing(self.h, float(t), _dp(U))
        return U

    def interp_velocity(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        uv = np.zeros(self.mesh.E * 2)
        lib().orc_sim_interp_velocity(self.h, _dp(v), _dp(uv))
        return uv.reshape(-1, 2)

    def set_projector(self, Pb, Pc):
        """Pressure term on eta + Pb @ (Pc @ eta); Pb (N, K), Pc (K, N)."""
        Pb = np.ascontiguousarray(Pb, dtype=np.float64)
        Pc = np.ascontiguousarray(Pc, dtype=np.float64)
        self._proj = (Pb, Pc)
        lib().orc_sim_set_projector(self.h, Pb.shape[1], _dp(Pb), _dp(Pc))

    def momentum(self, v, eta):
        v = np.ascontiguousarray(v, dtype=np.float64)
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        out = np.zeros(self.mesh.E)
        lib().orc_sim_momentum(self.h, _dp(v), _dp(eta), _dp(out))
        return out


def round_period(ang_vel):
    p, o = ctypes.c_double(), ctypes.c_double()
    lib().orc_round_period(float(ang_vel), ctypes.byref(p), ctypes.byref(o))
    return p.value, o.value


def quantise_dt(period, target):
    d, n = ctypes.c_double(), ctypes.c_int()
    lib().orc_quantise_dt(float(period), float(target), ctypes.byref(d), ctypes.byref(n))
    return d.value, n.value


def analytical_obliq_west(params, lat, lon, t):
    o = np.zeros(6)
    lib().orc_analytical_obliq_west(ctypes.byref(params), float(lat), float(lon), float(t), _dp(o))
    return o


def membrane_nu_beta(radius, hs, ho, l_max):
    nu, beta = np.zeros(l_max + 1), np.zeros(l_max + 1)
    ro, go = ctypes.c_double(), ctypes.c_double()
    lib().orc_membrane_nu_beta(float(radius), float(hs), float(ho), int(l_max), _dp(nu),
                               _dp(beta), ctypes.byref(ro), ctypes.byref(go))
    return nu, beta, ro.value, go.value


def membrane_nu_beta_gen(radius, hs, ho, l_max, mass_total, rig_core, rig_shell):
    """membraneNuBeta with the body mass and core / shell rigidities free."""
    nu, beta = np.zeros(l_max + 1), np.zeros(l_max + 1)
    ro, go = ctypes.c_double(), ctypes.c_double()
    lib().orc_membrane_nu_beta_gen(float(radius), float(hs), float(ho), int(l_max), float(mass_total),
                                   float(rig_core), float(rig_shell), _dp(nu), _dp(beta),
                                   ctypes.byref(ro), ctypes.byref(go))
    return nu, beta, ro.value, go.value


def mlp_forward(shape, params, x):
    """FP64 forward pass of the warm-start network shape (n_in, width, depth,
    n_out) with flat parameters (per layer W [out][in], then b)."""
    n_in, width, depth, n_out = shape
    x = np.ascontiguousarray(x, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    y = np.zeros((len(x), n_out))
    lib().orc_mlp_forward(n_in, width, depth, n_out, _dp(p), len(x), _dp(x), _dp(y))
    return y


def mlp_loss_grad(shape, params, x, t):
    """FP64 loss sum (y - t)^2 / (2 n n_out) and its gradient."""
    n_in, width, depth, n_out = shape
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    g = np.zeros_like(p)
    loss = lib().orc_mlp_loss_grad(n_in, width, depth, n_out, _dp(p), len(x), _dp(x), _dp(t), _dp(g))
    return loss, g


def read_reference_grid(path):
    """Parse a reference-format grid file (src/mesh.cpp:4016-4101) with the same
    deg->rad conversion (atof(str) * pi/180)."""
    rad = 3.1415926535897932384626433832795028841971693993751058 / 180.0
    rows = {}
    with open(path) as fh:
        fh.readline()
        for line in fh:
            if not line.strip():
                continue
            head, rest = line.split("{", 1)
            parts = head.split()
            nid = int(parts[0])
            fr_s, rest = rest.split("}", 1)
            fr = [int(x) for x in fr_s.split(",")]
            cen = rest.split("{", 1)[1].split("}", 1)[0]
            pts = [p.strip().strip("(").strip() for p in cen.split(")") if p.strip(" ,")]
            cl = []
            for p in pts:
                p = p.lstrip(",").strip().lstrip("(")
                a, b = p.split(",")
                cl.append((float(a) * rad, float(b) * rad))
            rows[nid] = (float(parts[1]) * rad, float(parts[2]) * rad, fr, cl)
    n = len(rows)
    lat = np.zeros(n)
    lon = np.zeros(n)
    friends = np.zeros((n, 6), np.int32)
    clat = np.zeros((n, 6))
    clon = np.zeros((n, 6))
    for i, (a, b, fr, cl) in rows.items():
        lat[i], lon[i] = a, b
        friends[i] = fr
        for j, (x, y) in enumerate(cl):
            clat[i, j], clon[i, j] = x, y
    return dict(lat=lat, lon=lon, friends=friends, clat=clat, clon=clon)


def real_sh_basis(L, lat, lon):
    """Real 4pi-normalised spherical harmonics on points, index l*l + l + m,
    from scipy's associated Legendre functions (independent of the product's
    recurrence): m >= 0 -> P_lm cos(m lon), m < 0 -> P_l|m| sin(|m| lon),
    no Condon-Shortley phase."""
    from scipy.special import gammaln, lpmv
    x = np.sin(lat)
    Y = np.zeros((len(lat), (L + 1) ** 2))
    for l in range(L + 1):
        for m in range(l + 1):
            norm = np.sqrt((2.0 - (m == 0)) * (2 * l + 1) * np.exp(gammaln(l - m + 1) - gammaln(l + m + 1)))
            p = norm * (-1.0) ** m * lpmv(m, l, x)
            Y[:, l * l + l + m] = p * np.cos(m * lon)
            if m > 0:
                Y[:, l * l + l - m] = p * np.sin(m * lon)
    return Y


def rbf_operator(grid, mesh):
    """operatorRBFinterp restated in numpy, the reference's RBF reconstruction
    of Cartesian velocity at cell centres from face-normal velocities
    (src/mesh.cpp:2263-2361 with the per-cell matrices of src/mesh.cpp:360-430
    and node2faceAdj of src/mesh.cpp:2725-2763).  Face normals are
    normalVectorBetweenXYZ (include/mathRoutines.h:246-283) of the face's two
    Voronoi vertices in the order its inner node lists them; arcs are
    distanceBetweenSph2 (include/mathRoutines.h:139-160); phi(d) = exp(-d^2)
    (include/mathRoutines.h:432-439); the per-cell inverse is
    numpy.linalg.inv.  Returns CSR (row_ptr [3N+1], col, val), rows 3i + c,
    columns ascending."""
    faces = mesh.get("faces").reshape(-1, 6)
    fnodes = mesh.get("face_nodes").reshape(-1, 2)
    fclat, fclon = mesh.get("face_centre_lat"), mesh.get("face_centre_lon")
    lat, lon = np.asarray(grid["lat"]), np.asarray(grid["lon"])
    clat, clon = np.asarray(grid["clat"]).reshape(-1, 6), np.asarray(grid["clon"]).reshape(-1, 6)
    N, E = faces.shape[0], fnodes.shape[0]
    deg = (faces >= 0).sum(axis=1)

    def xyz(la, lo):
        c = np.pi * 0.5 - la
        return np.stack([np.sin(c) * np.cos(lo), np.sin(c) * np.sin(lo), np.cos(c)], axis=-1)

    def arc(a, b):
        return np.arctan2(np.linalg.norm(np.cross(a, b), axis=-1), np.sum(a * b, axis=-1))

    n = np.zeros((E, 3))
    for e in range(E):
        i = fnodes[e, 0]
        k = deg[i]
        j = int(np.nonzero(faces[i, :k] == e)[0][0])
        a = xyz(clat[i, (j + k - 1) % k], clon[i, (j + k - 1) % k])
        b = xyz(clat[i, j], clon[i, j])
        c = np.cross(a, b)
        n[e] = c / np.linalg.norm(c)
    fc = xyz(fclat, fclon)
    pc = xyz(lat, lon)
    phi = lambda d: np.where(d < 1e-8, 1.0, np.exp(-d * d))
    rp, col, val = [0], [], []
    for i in range(N):
        f = faces[i, :deg[i]]
        A = phi(arc(fc[f][:, None, :], fc[f][None, :, :])) * (n[f] @ n[f].T)
        W = (phi(arc(pc[i][None, :], fc[f]))[:, None] * n[f]).T @ np.linalg.inv(A)
        order = np.argsort(f)
        for c in range(3):
            col.extend(f[order].tolist())
            val.extend(W[c, order].tolist())
            rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32), np.array(val)


def csr_apply_rowwise(rp, col, val, x):
    """y = R x with each row summed in column order from 0.0 (the order the
    device kernel uses)."""
    rp, col, val = np.asarray(rp), np.asarray(col), np.asarray(val)
    n = len(rp) - 1
    y = np.zeros(n)
    w = np.diff(rp).max() if n else 0
    for j in range(w):
        rows = np.nonzero(np.diff(rp) > j)[0]
        k = rp[rows] + j
        y[rows] = y[rows] + val[k] * x[col[k]]
    return y
