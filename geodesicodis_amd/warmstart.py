"""Learned warm start of the LTE spin-up: periodic states, their spectral
representation, and the network that predicts them.

A tidally forced, linearly damped ocean started from rest needs of order
1/alpha of simulated time before its orbit-averaged dissipation settles
(5 / (alpha P) orbits: 4 200 at alpha = 1e-8 s^-1 on Enceladus).  The
reference can only start from rest, from a file, or from the one analytic
solution (src/initialConditions.cpp:19-206).  This module starts instead
from a predicted periodic state:

- PeriodicSolver: the periodic response of the solver's own spatial
  operators (the product mesh's G, Div and TRiSK Coriolis coefficients, the
  momentum and continuity of src/updateMomentum.cpp:16-46 and
  src/updateEta.cpp:7-44 with the linear drag and forcing gradient of
  src/timeIntegrator.cpp:219,242) to a tide of one frequency, by one sparse
  complex solve.  The training set runs each such state through plain
  stepping on the GPU, so that its targets are converged runs of the solver.
- StateBasis: eta and the face-normal velocity as spherical-harmonic
  coefficients to degree L of eta, of the velocity potential chi and of the
  stream function psi (v_n = d(chi)/dn - d(psi)/ds), fitted by least squares
  on the grid; a coefficient vector is grid-independent.
- Mlp (odis_mlp_* in the library): the network, trained on the GPU by the
  library's forward and backward kernels, mapping ocean parameters to the
  coefficients of the periodic state at orbit phase 0.

The rest start stays the fallback; continued plain stepping from the
predicted state judges it (geodesicodis_amd.converge).
"""
import numpy as np

from . import ECC, FULL, OBLIQ, FREE, sh_basis

# C5 inputs: log10 thickness (m), log10 drag (s^-1), obliquity (degrees) and
# the tide (ECC or FULL)
TIDES = (ECC, FULL)


def _xyz(lat, lon):
    lat, lon = np.asarray(lat), np.asarray(lon)
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], -1)


def forcing_amplitude(p, lat, lon):
    """Complex amplitude U_hat of the tidal potential at the points, U(t) =
    Re(U_hat exp(i omega t)), for the single-frequency tides ECC, OBLIQ and
    FULL (src/tidalPotentials.cpp:82-131, 158-174): ECC's cos M and sin M
    terms and OBLIQ's cos M term."""
    lat, lon = np.asarray(lat), np.asarray(lon)
    base = p.love_reduct * p.omega ** 2 * p.radius ** 2
    sl2, cl2 = np.sin(lat) ** 2, np.cos(lat) ** 2
    s_cos = np.zeros_like(lat)
    s_sin = np.zeros_like(lat)
    if p.tide in (ECC, FULL):
        f = 0.75 * base * p.ecc
        s_cos = s_cos + f * ((1.0 - 3.0 * sl2) + cl2 * 3.0 * np.cos(2 * lon))
        s_sin = s_sin + f * cl2 * 4.0 * np.sin(2 * lon)
    if p.tide in (OBLIQ, FULL):
        s_cos = s_cos - 1.5 * base * p.obliquity * np.sin(2 * lat) * np.cos(lon)
    if p.tide not in (ECC, OBLIQ, FULL):
        raise ValueError("periodic solve: ECC, OBLIQ or FULL tide only")
    return s_cos - 1j * s_sin


class PeriodicSolver:
    """Periodic state of the semi-discrete linear LTE on one mesh.

    With v = Re(V e^{i w t}), eta = Re(H e^{i w t}) and the tide of
    frequency w, the step's tendencies dv/dt = -G(g eta - U) + C v - alpha v
    and d(eta)/dt = h Div v give
        ((i w + alpha) I - C + (g h / (i w)) G Div) V = G U_hat,
        H = h Div V / (i w),
    solved by a sparse LU (scipy's SuperLU).  Free surface, linear drag."""

    def __init__(self, mesh, omega):
        import scipy.sparse as sp
        N, E = mesh.N, mesh.E
        fn = mesh.get("face_nodes").reshape(E, 2)
        gc = mesh.get("grad_coef").reshape(E, 2)
        rows = np.repeat(np.arange(E), 2)
        self.G = sp.csr_matrix((gc.ravel(), (rows, fn.ravel())), shape=(E, N))
        faces = mesh.get("faces").reshape(N, 6)
        dc = mesh.get("div_coef").reshape(N, 6)
        ok = faces >= 0
        self.D = sp.csr_matrix((dc[ok], (np.nonzero(ok)[0], faces[ok])), shape=(N, E))
        tf = mesh.get("interp_friends").reshape(E, 10)
        tc = mesh.get("trisk_coef").reshape(E, 10)
        cnt = mesh.get("interp_count")
        lat_e = mesh.get("face_centre_lat")
        m = np.arange(10)[None, :] < cnt[:, None]
        r = np.nonzero(m)[0]
        self.C = sp.csr_matrix((-2.0 * omega * np.sin(lat_e[r]) * tc[m], (r, tf[m])), shape=(E, E))
        self.GD = (self.G @ self.D).tocsr()
        self.omega = float(omega)
        self.lat, self.lon = mesh.get("cell_lat"), mesh.get("cell_lon")
        self.N, self.E = N, E

    def solve(self, p):
        """(V, H): complex face-normal velocity and displacement amplitudes of
        the periodic state of the ocean p (Params); the state at time t is
        (Re(V e^{i w t}), Re(H e^{i w t}))."""
        import scipy.sparse as sp
        import scipy.sparse.linalg as spl
        if abs(p.omega - self.omega) > 1e-12 * self.omega:
            raise ValueError("solver built for another rotation rate")
        if p.surface != FREE or p.friction != 0 or p.quad_drag:
            raise ValueError("periodic solve: free surface and linear drag only")
        w = self.omega
        A = ((1j * w + p.alpha) * sp.identity(self.E, format="csr") - self.C
             + (p.g * p.h / (1j * w)) * self.GD).tocsc()
        U = forcing_amplitude(p, self.lat, self.lon)
        V = spl.splu(A).solve(self.G @ U)
        H = p.h * (self.D @ V) / (1j * w)
        return V, H

    def residual(self, p, V, H):
        """Relative residuals of the two periodic equations."""
        w = self.omega
        U = forcing_amplitude(p, self.lat, self.lon)
        rm = 1j * w * V - (-(self.G @ (p.g * H - U)) + self.C @ V - p.alpha * V)
        rc = 1j * w * H - p.h * (self.D @ V)
        return np.linalg.norm(rm) / np.linalg.norm(1j * w * V), np.linalg.norm(rc) / np.linalg.norm(1j * w * H)


class StateBasis:
    """(eta, v) <-> spherical-harmonic coefficients to degree L of eta, of
    the velocity potential chi and of the stream function psi.

    eta = Y c_eta on the cells; v_n = (chi(outer) - chi(inner)) / d_e -
    (psi(b) - psi(a)) / l_e on the faces, with a, b the face's Voronoi
    vertices ordered along s = r x n (n from the inner to the outer cell).
    Both fits are least squares (chi and psi from l = 1: a constant moves
    nothing).  Coefficients: [c_eta (K), c_chi (K - 1), c_psi (K - 1)],
    K = (L + 1)^2."""

    def __init__(self, mesh, grid_arrays, L=16):
        N, E = mesh.N, mesh.E
        K = (L + 1) ** 2
        self.L, self.K, self.N, self.E = L, K, N, E
        clat, clon = mesh.get("cell_lat"), mesh.get("cell_lon")
        self.Yc = sh_basis(L, clat, clon)
        fn = mesh.get("face_nodes").reshape(E, 2)
        d = mesh.get("face_node_dist")
        ln = mesh.get("face_len")
        faces = mesh.get("faces").reshape(N, 6)
        vlat = np.asarray(grid_arrays["clat"]).reshape(N, 6)
        vlon = np.asarray(grid_arrays["clon"]).reshape(N, 6)
        deg = (faces >= 0).sum(1)
        # the face's two Voronoi vertices in its inner cell's list: the face
        # between vertices j - 1 and j of that cell
        inner = fn[:, 0]
        pos = np.argmax(faces[inner] == np.arange(E)[:, None], axis=1)
        k = deg[inner]
        ja, jb = (pos + k - 1) % k, pos
        la_a, lo_a = vlat[inner, ja], vlon[inner, ja]
        la_b, lo_b = vlat[inner, jb], vlon[inner, jb]
        xa, xb = _xyz(la_a, lo_a), _xyz(la_b, lo_b)
        x1, x2 = _xyz(clat[fn[:, 0]], clon[fn[:, 0]]), _xyz(clat[fn[:, 1]], clon[fn[:, 1]])
        re = _xyz(mesh.get("face_centre_lat"), mesh.get("face_centre_lon"))
        s = np.cross(re, x2 - x1)
        sign = np.sign(np.sum((xb - xa) * s, axis=1))
        Ya, Yb = sh_basis(L, la_a, lo_a), sh_basis(L, la_b, lo_b)
        A_chi = (self.Yc[fn[:, 1], 1:] - self.Yc[fn[:, 0], 1:]) / d[:, None]
        A_psi = -sign[:, None] * (Yb[:, 1:] - Ya[:, 1:]) / ln[:, None]
        self.Av = np.concatenate([A_chi, A_psi], axis=1)
        del Ya, Yb, A_chi, A_psi
        self._qv, self._rv = np.linalg.qr(self.Av)
        self._qc, self._rc = np.linalg.qr(self.Yc)

    @property
    def size(self):
        return self.K + 2 * (self.K - 1)

    def project(self, v, eta):
        """Coefficients of a state (least squares)."""
        import scipy.linalg as sl
        ce = sl.solve_triangular(self._rc, self._qc.T @ eta)
        cv = sl.solve_triangular(self._rv, self._qv.T @ v)
        return np.concatenate([ce, cv])

    def synthesize(self, c):
        """(v, eta) of a coefficient vector."""
        c = np.asarray(c, np.float64)
        return self.Av @ c[self.K:], self.Yc @ c[:self.K]


# ---------------------------------------------------------------------------
# instances and periodic targets
# ---------------------------------------------------------------------------

TRAIN_SEED = 20261205
HELDOUT_SEED = 20261206
# largest gravity-wave frequency on the 40962-cell grid of a 38 km ocean; it
# scales as sqrt(h) and doubles per grid level
OMEGA_MAX_40962_38KM = 0.0386


def draw_instances(G, n, seed, tides=TIDES):
    """Warm-start instances: log10 thickness U(3, 5) m and log10 drag
    U(-11, -7) s^-1 (the sweep's ranges), obliquity U(0, 1) degrees and the
    tide from `tides`, drawn in that order per instance from
    numpy.random.default_rng(seed).  Free surface, e = 0.0047, linear drag.
    Returns the feature rows [log10 h, log10 alpha, obliquity_deg, tide]."""
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n):
        lh = rng.uniform(3.0, 5.0)
        la = rng.uniform(-11.0, -7.0)
        ob = rng.uniform(0.0, 1.0)
        td = tides[int(rng.integers(len(tides)))]
        rows.append((lh, la, ob if td != ECC else 0.0, td))
    return np.array(rows, np.float64)


def instance_params(G, row, cells):
    """Params of one feature row on a grid of `cells` cells: dt the quantised
    step below 0.35 / omega_max(h) of that grid."""
    lh, la, ob, td = row
    h = 10.0 ** lh
    w = OMEGA_MAX_40962_38KM * (cells / 40962.0) ** 0.5 * (h / 38e3) ** 0.5
    p, n = G.make_params(h=h, alpha=10.0 ** la, tide=int(td), obliquity=np.radians(ob), time_step=0.35 / w)
    return p, n


def features(rows):
    """Network inputs of feature rows: log10 h - 4, (log10 alpha + 9) / 2,
    2 obliquity - 1, and one-hot of the tide (ECC, FULL)."""
    rows = np.asarray(rows, np.float64)
    x = np.zeros((len(rows), 3 + len(TIDES)), np.float32)
    x[:, 0] = rows[:, 0] - 4.0
    x[:, 1] = (rows[:, 1] + 9.0) / 2.0
    x[:, 2] = 2.0 * rows[:, 2] - 1.0
    for k, t in enumerate(TIDES):
        x[:, 3 + k] = rows[:, 3] == t
    return x


_POOL_STATE = {}


def _pool_init(level, L):
    import geodesicodis_amd as G
    g = G.Grid.generate(level, 64)
    p0, _ = G.make_params()
    mesh = G.Mesh(g, p0.radius)
    _POOL_STATE["G"] = G
    _POOL_STATE["solver"] = PeriodicSolver(mesh, p0.omega)
    _POOL_STATE["basis"] = StateBasis(mesh, g.arrays(), L)
    _POOL_STATE["cells"] = mesh.N


def _pool_solve(row):
    G, ps, sb = _POOL_STATE["G"], _POOL_STATE["solver"], _POOL_STATE["basis"]
    p, _ = instance_params(G, row, _POOL_STATE["cells"])
    V, H = ps.solve(p)
    return np.concatenate([sb.project(V.real, H.real), sb.project(V.imag, H.imag)])


def periodic_coefficients(rows, level=6, L=16, workers=8):
    """Coefficients [phase 0 | phase pi/2] of the periodic states of the
    feature rows, solved on the generated grid of reference level `level`
    (10242 cells at 6) in `workers` processes: shape (n, 2 * basis size).
    The state at phase 0 is Re(V), Re(H); at phase pi/2 (a quarter period
    later, up to sign) Im(V), Im(H)."""
    import multiprocessing as mp
    import os
    rows = [tuple(r) for r in np.asarray(rows)]
    if workers <= 1:
        _pool_init(level, L)
        return np.array([_pool_solve(r) for r in rows])
    # fresh interpreters (no inherited thread pools), one BLAS / OpenMP thread each
    keep = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in keep:
        os.environ[k] = "1"
    try:
        with mp.get_context("spawn").Pool(workers, initializer=_pool_init, initargs=(level, L)) as pool:
            return np.array(pool.map(_pool_solve, rows, chunksize=4))
    finally:
        for k, v in keep.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
