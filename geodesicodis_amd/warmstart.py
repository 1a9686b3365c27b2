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
    Returns the feature rows [log10 h, log10 alpha, obliquity_deg, tide]
    (obliquity 0 for ECC)."""
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n):
        lh = rng.uniform(3.0, 5.0)
        la = rng.uniform(-11.0, -7.0)
        ob = rng.uniform(0.0, 1.0)
        td = tides[int(rng.integers(len(tides)))]
        rows.append((lh, la, ob if td != ECC else 0.0, td))
    return np.array(rows, np.float64)


ECC_DEFAULT = 0.0047


def instance_params(G, row, cells, ecc=ECC_DEFAULT):
    """Params of one feature row on a grid of `cells` cells: dt the quantised
    step below 0.35 / omega_max(h) of that grid.  Every row runs as FULL
    (src/tidalPotentials.cpp:158-174): an ECC row with zero obliquity, the
    same potential as ECC, so that all instances share a batch."""
    lh, la, ob, td = row
    h = 10.0 ** lh
    w = OMEGA_MAX_40962_38KM * (cells / 40962.0) ** 0.5 * (h / 38e3) ** 0.5
    obl = np.radians(ob) if int(td) == FULL else 0.0
    p, n = G.make_params(h=h, alpha=10.0 ** la, tide=FULL, ecc=ecc, obliquity=obl, time_step=0.35 / w)
    return p, n


def features(rows):
    """Network inputs of feature rows: log10 h - 4 and (log10 alpha + 9) / 2.
    The obliquity and the tide enter through the superposition of the
    eccentricity and obliquity responses (the LTE is linear in the tide):
    state = response(ECC, e) + obliquity_deg x response(OBLIQ, 1 degree)."""
    rows = np.asarray(rows, np.float64)
    return np.stack([rows[:, 0] - 4.0, (rows[:, 1] + 9.0) / 2.0], 1).astype(np.float32)


def combine(rows, pair):
    """Coefficients of the rows' states from [ECC | OBLIQ at 1 degree]
    response pairs (one block each, of any number of phases): ECC plus the
    row's obliquity in degrees times the obliquity response (0 for ECC)."""
    rows = np.asarray(rows, np.float64)
    pair = np.asarray(pair)
    half = pair.shape[1] // 2
    ob = np.where(rows[:, 3] == FULL, rows[:, 2], 0.0)
    return pair[:, :half] + ob[:, None] * pair[:, half:]


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


def _pool_solve(hrow):
    """[ECC phase 0, phase pi/2 | OBLIQ (1 degree) phase 0, phase pi/2]
    coefficients of the periodic states of thickness / drag row hrow."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    G, ps, sb = _POOL_STATE["G"], _POOL_STATE["solver"], _POOL_STATE["basis"]
    p, _ = instance_params(G, (hrow[0], hrow[1], 1.0, FULL), _POOL_STATE["cells"])
    w = ps.omega
    A = ((1j * w + p.alpha) * sp.identity(ps.E, format="csr") - ps.C + (p.g * p.h / (1j * w)) * ps.GD).tocsc()
    lu = spl.splu(A)
    out = []
    for tide, ecc, obl in ((ECC, ECC_DEFAULT, 0.0), (OBLIQ, 0.0, np.radians(1.0))):
        p.tide, p.ecc, p.obliquity = tide, ecc, obl
        V = lu.solve(ps.G @ forcing_amplitude(p, ps.lat, ps.lon))
        H = p.h * (ps.D @ V) / (1j * w)
        out += [sb.project(V.real, H.real), sb.project(V.imag, H.imag)]
    return np.concatenate(out)


def periodic_pairs(hrows, level=6, L=16, workers=8, log=None):
    """[ECC | OBLIQ at 1 degree] x [phase 0 | phase pi/2] coefficients of the
    periodic states of (log10 h, log10 alpha) rows, solved on the generated
    grid of reference level `level` (10242 cells at 6) in `workers`
    processes (one LU per row, two tides): shape (n, 4 x basis size).  The
    state at phase 0 is Re(V), Re(H); a quarter period later Im(V), Im(H)
    up to sign."""
    import multiprocessing as mp
    import os
    hrows = [tuple(r[:2]) for r in np.asarray(hrows)]
    if workers <= 1:
        _pool_init(level, L)
        return np.array([_pool_solve(r) for r in hrows])
    # fresh interpreters (no inherited thread pools), one BLAS / OpenMP thread each
    keep = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in keep:
        os.environ[k] = "1"
    try:
        with mp.get_context("spawn").Pool(workers, initializer=_pool_init, initargs=(level, L)) as pool:
            out = []
            for r in pool.imap(_pool_solve, hrows, chunksize=4):
                out.append(r)
                if log and len(out) % 200 == 0:
                    log("warm start: %d / %d periodic solves" % (len(out), len(hrows)))
            return np.array(out)
    finally:
        for k, v in keep.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def periodic_states(rows, level=6, workers=8, log=None):
    """[phase 0 | phase pi/2] coefficients of the periodic states of feature
    rows (the superposition of their tide's responses)."""
    pr = periodic_pairs(rows, level=level, workers=workers, log=log)
    q = pr.shape[1] // 4
    return combine(rows, np.concatenate([pr[:, :2 * q], pr[:, 2 * q:]], 1))


# ---------------------------------------------------------------------------
# converged targets, the model, and seeding
# ---------------------------------------------------------------------------

def _grid_mesh(G, level):
    g = G.Grid.generate(level, 64)
    p0, _ = G.make_params()
    return g, G.Mesh(g, p0.radius)


def seed_states(ctx, basis, coeffs, first=0):
    """Set instances first, first + 1, ... of ctx to the states of the
    coefficient rows (phase 0), at step 0 with the Euler start (the rest
    start's, src/temporalOperators.cpp:48-65), as the reference seeds its
    analytic start (src/initialConditions.cpp:146-206)."""
    for k, c in enumerate(np.asarray(coeffs)):
        v, eta = basis.synthesize(c)
        ctx.set_state(first + k, v, None, eta, None, 0)


def periodic_dissipation(G, mesh, params, basis, coeffs2, precision=None):
    """Orbit-mean dissipation D* of periodic states given as [phase 0 |
    phase pi/2] coefficient rows: the state is a cos t + b sin t, so the
    orbit mean of the quadratic dissipation (src/energy.cpp:13-62) is
    (D(a) + D(b)) / 2, evaluated by the GPU context on the mesh."""
    n = basis.size
    ctx = G.Context(mesh, list(params), precision=G.FP64 if precision is None else precision)
    out = np.zeros(len(params))
    for half in (0, 1):
        seed_states(ctx, basis, np.asarray(coeffs2)[:, half * n:(half + 1) * n])
        out += 0.5 * ctx.dissipation()
    ctx.close()
    return out


def converged_targets(G, hrows, pairs, level=7, orbits=2, batch=256, precision=None, log=None):
    """Training targets from converged runs of the solver: for every (h,
    alpha) row, its ECC response and its OBLIQ response (1 degree) are
    started from their periodic states (pairs, phase 0) on the generated
    grid of reference level `level` and stepped `orbits` whole orbits by the
    batched plain step (FP32 by default, both tides as FULL); each target is
    the coefficient vector of the state after the last orbit (phase 0
    again).  Returns (targets [ECC | OBLIQ], D_last, resid_v, resid_eta),
    the last three per run (2 per row)."""
    from .converge import run_to_convergence
    g, mesh = _grid_mesh(G, level)
    basis = StateBasis(mesh, g.arrays(), L=16)
    n = basis.size
    hrows = np.asarray(hrows)
    m = len(hrows)
    # runs: row r's ECC response is run 2r, its OBLIQ response run 2r + 1
    runs = []
    for r in hrows:
        runs.append(((r[0], r[1], 0.0, ECC), ECC_DEFAULT))
        runs.append(((r[0], r[1], 1.0, FULL), 0.0))
    q = np.asarray(pairs).shape[1] // 4
    start = np.zeros((2 * m, n))
    start[0::2] = np.asarray(pairs)[:, :n]
    start[1::2] = np.asarray(pairs)[:, 2 * q:2 * q + n]
    tg = np.zeros((2 * m, n))
    dl, rv, re = np.zeros(2 * m), np.zeros(2 * m), np.zeros(2 * m)
    prec = G.FP32 if precision is None else precision
    for s in range(0, 2 * m, batch):
        sub = runs[s:s + batch]
        ps, ns = zip(*[instance_params(G, row, mesh.N, ecc=e) for row, e in sub])
        ctx = G.Context(mesh, list(ps), precision=prec)
        seed_states(ctx, basis, start[s:s + batch])

        def grab(slot, b, ctx=ctx, s=s):
            st = ctx.get_state(slot)
            tg[s + b] = basis.project(st["v"], st["eta"])

        recs = run_to_convergence(ctx, ns, max_orbits=orbits, stable=10 ** 9, on_stop=grab)
        for b, r in enumerate(recs):
            dl[s + b], rv[s + b], re[s + b] = r.D[-1], r.resid_v[-1], r.resid_eta[-1]
        ctx.close()
        if log:
            log("warm start: converged targets %d / %d runs" % (s + len(sub), 2 * m))
    return np.concatenate([tg[0::2], tg[1::2]], 1), dl, rv, re


class WarmStartModel:
    """Network + target normalisation.  The network maps (log10 h, log10
    alpha) to the phase-0 coefficients of the ECC response and of the
    OBLIQ response at 1 degree; a row's state is their superposition
    (combine).  Each response's coefficient vector c is learned as its
    direction c / |c| (standardised by one scale over all components, so
    that the loss is its relative L2 error) and log |c| (standardised)."""

    def __init__(self, G, width=256, depth=4, size=865, seed=TRAIN_SEED):
        self.G = G
        self.size = size
        self.mlp = G.Mlp(2, width, depth, 2 * (size + 1), seed=seed)
        self.norm = None

    def _split(self, c):
        n = self.size
        return [c[:, :n], c[:, n:2 * n]]

    def fit(self, hrows, coeffs, epochs=1500, batch=250, lr_start=2e-3, lr_end=2e-5, seed=TRAIN_SEED):
        """coeffs: (n, 2 size) [ECC | OBLIQ 1 degree] phase-0 targets."""
        c = np.asarray(coeffs, np.float64)
        self.norm = {}
        z = []
        for k, part in enumerate(self._split(c)):
            nrm = np.linalg.norm(part, axis=1, keepdims=True)
            d = part / nrm
            ln = np.log(nrm[:, 0])
            self.norm.update({"dmu%d" % k: d.mean(0), "dsd%d" % k: float(d.std()) + 1e-12,
                              "lmu%d" % k: float(ln.mean()), "lsd%d" % k: float(ln.std()) + 1e-12})
            z += [(d - self.norm["dmu%d" % k]) / self.norm["dsd%d" % k],
                  ((ln - self.norm["lmu%d" % k]) / self.norm["lsd%d" % k])[:, None]]
        self.mlp.set_data(features(hrows), np.concatenate(z, 1).astype(np.float32))
        return self.mlp.train(epochs, batch, lr_start, lr_end, seed)

    def predict_pairs(self, rows):
        """(n, 2 size) predicted [ECC | OBLIQ 1 degree] phase-0 responses."""
        n = self.size
        z = self.mlp.forward(features(rows)).astype(np.float64)
        out = []
        for k in range(2):
            zk = z[:, k * (n + 1):(k + 1) * (n + 1)]
            d = zk[:, :n] * self.norm["dsd%d" % k] + self.norm["dmu%d" % k]
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            out.append(d * np.exp(zk[:, n:] * self.norm["lsd%d" % k] + self.norm["lmu%d" % k]))
        return np.concatenate(out, 1)

    def predict(self, rows):
        """(n, size) predicted phase-0 coefficients of the rows' states."""
        return combine(rows, self.predict_pairs(rows))

    def save(self, path):
        np.savez(path, params=self.mlp.params(), shape=np.array(self.mlp.shape), size=self.size,
                 **{k: np.asarray(v) for k, v in self.norm.items()})

    @classmethod
    def load(cls, G, path):
        f = np.load(path)
        shape = [int(x) for x in f["shape"]]
        m = cls(G, width=shape[1], depth=shape[2], size=int(f["size"]))
        m.mlp.set_params(f["params"])
        m.norm = {k: (float(f[k]) if f[k].ndim == 0 else f[k]) for k in f.files
                  if k[:3] in ("dmu", "dsd", "lmu", "lsd")}
        return m


def build_model(G, n_train=2000, solve_level=6, run_level=7, orbits=2, epochs=1500, batch=250, workers=16,
                seed=TRAIN_SEED, log=None):
    """The whole training pipeline: draw n_train instances (seed), solve the
    periodic ECC and OBLIQ responses of their (h, alpha) on the solve_level
    grid, run each two orbits on the run_level grid from that state
    (converged runs of the solver), and train the network on the GPU on the
    resulting phase-0 coefficients.  Returns (model, info)."""
    import time
    t0 = time.perf_counter()
    rows = draw_instances(G, n_train, seed)
    pairs = periodic_pairs(rows, level=solve_level, workers=workers, log=log)
    t1 = time.perf_counter()
    if log:
        log("warm start: %d periodic solves (2 tides) on level %d in %.1f s" % (n_train, solve_level, t1 - t0))
    tg, dl, rv, re = converged_targets(G, rows, pairs, level=run_level, orbits=orbits, log=log)
    t2 = time.perf_counter()
    model = WarmStartModel(G, seed=seed)
    loss = model.fit(rows, tg, epochs=epochs, batch=batch, seed=seed)
    t3 = time.perf_counter()
    n = model.size
    q = pairs.shape[1] // 4
    solved = np.concatenate([pairs[:, :n], pairs[:, 2 * q:2 * q + n]], 1)
    shift = np.linalg.norm(tg - solved, axis=1) / np.linalg.norm(solved, axis=1)
    fit = np.linalg.norm(model.predict_pairs(rows) - tg, axis=1) / np.linalg.norm(tg, axis=1)
    info = {"n_train": n_train, "seed": seed, "solve_level": solve_level, "run_level": run_level,
            "orbits": orbits, "epochs": epochs, "batch": batch, "loss_first": float(loss[0]),
            "loss_last": float(loss[-1]), "solve_s": t1 - t0, "runs_s": t2 - t1, "train_s": t3 - t2,
            "target_vs_solve_median": float(np.median(shift)), "target_vs_solve_max": float(shift.max()),
            "periodicity_resid_v_median": float(np.median(rv)), "periodicity_resid_v_max": float(np.max(rv)),
            "periodicity_resid_eta_median": float(np.median(re)),
            "train_fit_rel_err_median": float(np.median(fit)), "train_fit_rel_err_max": float(fit.max())}
    return model, info


def evaluate(G, model, rows, level=7, solve_level=6, cap_warm=30, cap_rest=20, band=0.01, stable=3,
             precision=None, workers=16, log=None):
    """Held-out check by continued plain stepping: every instance of `rows`
    runs batched from its predicted state and from rest until its orbit
    means stay within `band` of D* (the orbit mean of its periodic state,
    solved on the solve_level grid) for `stable` orbits, or to the caps.
    Also times one plain FP64 instance on the grid, for the sequential
    plain method's wall time: sum over instances of orbits-to-band x steps
    per orbit x that step time (a lower bound where the rest start capped).
    Returns a dict of per-instance orbits and the wall-time ratio."""
    import time
    from .converge import run_to_convergence
    prec = G.FP32 if precision is None else precision
    g, mesh = _grid_mesh(G, level)
    basis = StateBasis(mesh, g.arrays(), L=16)
    rows = np.asarray(rows)
    ps, ns = zip(*[instance_params(G, r, mesh.N) for r in rows])
    C2 = periodic_states(rows, level=solve_level, workers=workers, log=log)
    d_star = periodic_dissipation(G, mesh, ps, basis, C2)
    pred = model.predict(rows)
    n = basis.size
    pred_err = np.linalg.norm(pred - C2[:, :n], axis=1) / np.linalg.norm(C2[:, :n], axis=1)
    out = {"instances": len(rows), "cells": mesh.N, "band": band, "stable_orbits": stable,
           "pred_rel_err_median": float(np.median(pred_err)), "pred_rel_err_p90": float(np.percentile(pred_err, 90)),
           "pred_rel_err_max": float(pred_err.max())}
    res = {}
    for name, seed, cap in (("warm", True, cap_warm), ("rest", False, cap_rest)):
        ctx = G.Context(mesh, list(ps), precision=prec)
        if seed:
            seed_states(ctx, basis, pred)
        t0 = time.perf_counter()
        recs = run_to_convergence(ctx, ns, d_ref=d_star, band=band, stable=stable, max_orbits=cap, log=log)
        wall = time.perf_counter() - t0
        ctx.close()
        orb = np.array([r.orbits - stable + 1 if r.stop == "band" else r.orbits for r in recs])
        res[name] = (recs, orb, wall)
        out[name] = {"wall_s": wall, "cap_orbits": cap,
                     "in_band": int(sum(r.stop == "band" for r in recs)),
                     "orbits_to_band_median": float(np.median(orb)), "orbits_to_band_max": int(orb.max()),
                     "first_orbit_in_band": int(sum(r.stop == "band" and r.orbits == stable for r in recs))}
        if log:
            log("warm start: %s start, %d of %d in band, %.1f s" % (name, out[name]["in_band"], len(rows), wall))
    # one plain FP64 instance on this grid: step time
    hs = [p.h for p in ps]
    mid = int(np.argsort(hs)[len(hs) // 2])
    c1 = G.Context(mesh, ps[mid], precision=G.FP64)
    c1.step(50)
    t0 = time.perf_counter()
    c1.step(1000)
    t_fp64 = (time.perf_counter() - t0) / 1000
    c1.close()
    rest_orb = res["rest"][1]
    plain = float(np.sum(rest_orb * np.array(ns) * t_fp64))
    out["plain_fp64_step_s"] = t_fp64
    out["plain_fp64_sequential_s"] = plain
    out["plain_is_lower_bound"] = bool(any(r.stop != "band" for r in res["rest"][0]))
    out["speedup_vs_plain_fp64"] = plain / res["warm"][2]
    out["rest_batched_over_warm_batched"] = res["rest"][2] / res["warm"][2]
    return out


def _main(argv=None):
    import argparse
    import json
    import sys
    import time

    import geodesicodis_amd as G

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    ap = argparse.ArgumentParser(prog="python -m geodesicodis_amd.warmstart")
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build", help="train the network from converged runs and save it")
    b.add_argument("--out", required=True)
    b.add_argument("--n-train", type=int, default=2000)
    b.add_argument("--solve-level", type=int, default=6)
    b.add_argument("--run-level", type=int, default=7)
    b.add_argument("--epochs", type=int, default=2000)
    b.add_argument("--batch", type=int, default=250)
    b.add_argument("--workers", type=int, default=16)
    e = sub.add_parser("evaluate", help="held-out instances from the network's states and from rest")
    e.add_argument("--model", required=True)
    e.add_argument("--n", type=int, default=256)
    e.add_argument("--level", type=int, default=7)
    e.add_argument("--solve-level", type=int, default=6)
    e.add_argument("--cap-warm", type=int, default=30)
    e.add_argument("--cap-rest", type=int, default=20)
    e.add_argument("--workers", type=int, default=16)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    if a.cmd == "build":
        model, info = build_model(G, n_train=a.n_train, solve_level=a.solve_level, run_level=a.run_level,
                                  epochs=a.epochs, batch=a.batch, workers=a.workers, log=log)
        model.save(a.out)
        info["wall_s"] = time.perf_counter() - t0
        print(json.dumps(info), flush=True)
    else:
        model = WarmStartModel.load(G, a.model)
        rows = draw_instances(G, a.n, HELDOUT_SEED)
        ev = evaluate(G, model, rows, level=a.level, solve_level=a.solve_level, cap_warm=a.cap_warm,
                      cap_rest=a.cap_rest, workers=a.workers, log=log)
        ev["wall_s"] = time.perf_counter() - t0
        print(json.dumps(ev), flush=True)


if __name__ == "__main__":
    _main()
