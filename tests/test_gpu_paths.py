"""GPU product paths against the CPU FP64 oracle, through the C ABI: the
reconstructed (east, north) velocity, GENERAL forcing, the dissipation
series after a state is set, and the standalone driver's INIT_ANALYTICAL and
INIT_LOAD starts (text and binary checkpoints)."""
import os

import numpy as np
import pytest

from conftest import generated_grid, rel_l2

pytestmark = pytest.mark.gpu


def _oracle_params(O, p):
    return O.make_params(**{k: getattr(p, k) for k in
                            ("g", "h", "alpha", "omega", "radius", "love_reduct", "ecc", "obliquity", "dt",
                             "shell_thickness", "semimajor", "tide", "surface", "friction", "init_load")},
                         quad_drag_momentum=p.quad_drag)


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5)])
@pytest.mark.parametrize("tide", ["ECC", "OBLIQ_WEST"])
def test_velocity_en_matches_oracle(product, oracle, precision, tol, tide):
    """odis_get_velocity_en (k_velocity_en) is interpolateVelocity
    (src/interpolation.cpp:26-62): (east, north) = n v_n + t v_t with
    t = (n_y, -n_x) and v_t the TRiSK reconstruction.  Compared with the
    oracle's reconstruction of the oracle's own state after 150 steps, per
    component, so a swapped or sign-flipped component fails."""
    G, O = product, oracle
    g, arrs = generated_grid(4)
    p, _ = G.make_params(tide=getattr(G, tide), obliquity=0.02, time_step=60)
    mesh = G.Mesh(g, p.radius)
    ctx = G.Context(mesh, p, precision=precision)
    sim = O.Sim(O.Mesh(arrs, p.radius), _oracle_params(O, p))
    ctx.step(150)
    sim.step(150)
    uv = ctx.velocity_en(0)
    want = sim.interp_velocity(sim.get_state()["v"])
    for c in range(2):
        e = rel_l2(uv[:, c], want[:, c])
        assert e <= tol, (c, e)
    # the normal component is v itself: n . (u, v) = v_n
    nrm = mesh.get("face_normal").reshape(-1, 2)
    vn = ctx.get_state()["v"]
    assert rel_l2((uv * nrm).sum(1), vn) <= (1e-14 if precision == 0 else 1e-6)
    ctx.close()


def test_general_forcing_is_zero(product, oracle):
    """GENERAL (src/tidalPotentials.cpp:227-281) scales Pbar_20 / Pbar_22,
    which the reference allocates zero-filled and never sets
    (src/mesh.cpp:63-64), so its potential is zero: a rest start stays at
    rest, as in the oracle."""
    G, O = product, oracle
    g, arrs = generated_grid(3)
    p, _ = G.make_params(tide=G.GENERAL)
    ctx = G.Context(G.Mesh(g, p.radius), p)
    ctx.step(40)
    s = ctx.get_state()
    assert np.all(s["v"] == 0) and np.all(s["eta"] == 0) and ctx.dissipation()[0] == 0.0
    sim = O.Sim(O.Mesh(arrs, p.radius), _oracle_params(O, p))
    sim.step(40)
    assert np.all(sim.get_state()["v"] == 0)
    ctx.close()


def test_dissipation_series_after_set_state(product, oracle):
    """After odis_set_state with iter = k the series ring holds nothing of the
    earlier trajectory: reading below k is ODIS_E_RANGE.  Stepping from k
    fills entries k, k+1, ...: entry k is the dissipation of the state that
    was set, and the rest match the oracle resumed from the same state."""
    G, O = product, oracle
    g, arrs = generated_grid(4)
    p, _ = G.make_params(tide=G.ECC, time_step=60)
    mesh = G.Mesh(g, p.radius)
    ctx = G.Context(mesh, p)
    ctx.step(60)
    ck = ctx.get_state()
    assert ck["iter"] == 60
    ctx2 = G.Context(mesh, p)
    ctx2.step(90)                 # an earlier trajectory fills entries 0..89
    ctx2.set_state(0, ck["v"], ck["dvdt"], ck["eta"], ck["detadt"], iter=60)
    with pytest.raises(G.OdisError) as ei:
        ctx2.dissipation_series(0, 59, 1)
    assert ei.value.code == -6
    with pytest.raises(G.OdisError):
        ctx2.dissipation_series(0, 60, 1)   # nothing stepped yet
    d_set = ctx2.dissipation()[0]
    ctx2.step(25)
    series = ctx2.dissipation_series(0, 60, 25)
    assert abs(series[0] - d_set) <= 1e-14 * abs(d_set)
    sim = O.Sim(O.Mesh(arrs, p.radius), _oracle_params(O, p))
    sim.set_state(ck["v"], ck["dvdt"], ck["eta"], ck["detadt"], 60)
    d_orc = sim.step(25, record=True)     # d_orc[j]: state 61 + j
    assert rel_l2(series[1:], d_orc[:-1]) <= 1e-12
    ctx.close()
    ctx2.close()


CASE = """radius; 252.1e3; top radius;
k2; 0.0; k2;
h2; 0.0; h2;
love reduction factor; 1.0; set by FREE;
ANGULAR VELOCITY; 5.307e-5; mean motion;
surface gravity; 0.113; g;
eccentricity; 0.0047; e;
obliquity; {obl}; degrees;
ocean thickness; 38e3; h_o;
friction coefficient; 1e-7; alpha;
friction type; LINEAR; friction;
potential; {tide}; potential;
surface type; FREE; surface;
advection; false; advection;
solver type; AB3; solver;
geodesic grid level; 4; level;
output time; 4; outputs per orbit;
initial conditions; {init}; ic;
simulation end time; {end}; orbits;
time step; 100; dt target;
"""


def _case(path, tide="ECC", obl=0.0, init="NONE", end=1.0):
    path.mkdir(parents=True, exist_ok=True)
    (path / "input.in").write_text(CASE.format(tide=tide, obl=obl, init=init, end=end))
    return path


def _read_bin(path):
    with open(path, "rb") as fh:
        assert fh.read(8) == b"ODISCKP1"
        rows = int(np.frombuffer(fh.read(8), np.int64)[0])
        a = np.frombuffer(fh.read(), np.float64)
    assert a.size == 4 * rows
    return a.reshape(rows, 4)


def _checkpoint(case):
    v = _read_bin(case / "InitialConditions" / "vel_init.bin")
    e = _read_bin(case / "InitialConditions" / "pres_init.bin")
    return dict(v=v[:, 0], dvdt=v[:, 1:], eta=e[:, 0], detadt=e[:, 1:])


def test_run_case_analytical_start_matches_oracle(product, oracle, tmp_path):
    """odis_run_case with `initial conditions; ANALYTICAL` seeds v, eta and
    the three-level AB3 history from the product's OBLIQ_WEST solution
    (src/initialConditions.cpp:146-206) and steps; its FP64 checkpoint after
    60 steps matches the oracle seeded from the oracle's own closed form
    (src/analyticalLTE.cpp:133-180) and stepped alike."""
    G, O = product, oracle
    case = _case(tmp_path / "a", tide="OBLIQ_WEST", obl=0.1, init="ANALYTICAL", end=0.05)
    p, info = G.read_config(case)
    assert info["steps_per_orbit"] == 1200
    assert G.run_case(case) == 0
    got = _checkpoint(case)
    g, arrs = generated_grid(4)
    mesh = G.Mesh(g, p.radius)
    op = _oracle_params(O, p)
    flat, flon = mesh.get("face_centre_lat"), mesh.get("face_centre_lon")
    nrm = mesh.get("face_normal").reshape(-1, 2)
    clat, clon = mesh.get("cell_lat"), mesh.get("cell_lon")
    ts = (0.0, -p.dt, -2 * p.dt)
    v, dv = np.zeros(mesh.E), np.zeros((mesh.E, 3))
    eta, de = np.zeros(mesh.N), np.zeros((mesh.N, 3))
    for e in range(mesh.E):
        for k, t in enumerate(ts):
            o = O.analytical_obliq_west(op, flat[e], flon[e], t)
            if k == 0:
                v[e] = o[0] * nrm[e, 0] + o[1] * nrm[e, 1]
            dv[e, k] = o[2] * nrm[e, 0] + o[3] * nrm[e, 1]
    for i in range(mesh.N):
        for k, t in enumerate(ts):
            o = O.analytical_obliq_west(op, clat[i], clon[i], t)
            if k == 0:
                eta[i] = o[4]
            de[i, k] = o[5]
    sim = O.Sim(O.Mesh(arrs, p.radius), op)
    sim.set_state(v, dv, eta, de, 0)
    sim.step(60)
    want = sim.get_state()
    for k in ("v", "eta", "dvdt", "detadt"):
        assert rel_l2(got[k], want[k]) <= 1e-12, k


def test_run_case_load_resume(product, oracle, tmp_path):
    """INIT_LOAD through the standalone driver (src/initialConditions.cpp:
    19-144, AB3 from step 0 with the loaded history): two orbits in one run
    equal one orbit, then one more orbit resumed from the FP64 checkpoint
    (time restarts at 0, a whole orbit later the tide is the same) to
    rounding; resumed from the %1.6E text checkpoint alone, the run matches
    the oracle resumed from the same parsed values."""
    import shutil
    G, O = product, oracle
    full = _case(tmp_path / "full", end=2.0)
    assert G.run_case(full) == 0
    first = _case(tmp_path / "first", end=1.0)
    assert G.run_case(first) == 0
    second = _case(tmp_path / "second", init="LOAD", end=1.0)
    shutil.copytree(first / "InitialConditions", second / "InitialConditions")
    assert G.run_case(second) == 0
    assert "vel_init.bin" in (second / "DATA" / "OUTPUT.txt").read_text()
    a, b = _checkpoint(full), _checkpoint(second)
    for k in ("v", "eta", "dvdt", "detadt"):
        assert rel_l2(b[k], a[k]) <= 1e-12, (k, rel_l2(b[k], a[k]))
    # the text path: only the %1.6E files
    text = _case(tmp_path / "text", init="LOAD", end=0.25)
    (text / "InitialConditions").mkdir()
    for nm in ("vel_init.txt", "pres_init.txt"):
        shutil.copy(first / "InitialConditions" / nm, text / "InitialConditions" / nm)
    assert G.run_case(text) == 0
    vt = np.loadtxt(first / "InitialConditions" / "vel_init.txt", delimiter=",")
    et = np.loadtxt(first / "InitialConditions" / "pres_init.txt", delimiter=",")
    p, info = G.read_config(text)
    assert p.init_load == 1
    g, arrs = generated_grid(4)
    sim = O.Sim(O.Mesh(arrs, p.radius), _oracle_params(O, p))
    sim.set_state(vt[:, 0], vt[:, 1:], et[:, 0], et[:, 1:], 0)
    sim.step(300)
    got, want = _checkpoint(text), sim.get_state()
    for k in ("v", "eta", "dvdt", "detadt"):
        assert rel_l2(got[k], want[k]) <= 1e-10, k
