"""Host side of the learned warm start (no GPU): the periodic-state solver
on the product's operators against the oracle's own tendencies, the
spectral state basis, and the forcing superposition the network relies
on."""
import numpy as np
import pytest

from conftest import generated_grid, rel_l2


@pytest.fixture(scope="module")
def level4(product):
    from geodesicodis_amd.warmstart import PeriodicSolver, StateBasis
    G = product
    g, arrs = generated_grid(4)
    p0, _ = G.make_params()
    mesh = G.Mesh(g, p0.radius)
    return mesh, arrs, PeriodicSolver(mesh, p0.omega), StateBasis(mesh, arrs, L=8)


def test_periodic_solver_operators_are_the_oracles(product, oracle, level4):
    """The solver's G, Div and Coriolis matrices, assembled from the
    product mesh's arrays, equal the oracle's CSR operators
    (src/mesh.cpp:2808-3287) to the last bits."""
    import scipy.sparse as sp
    mesh, arrs, ps, _ = level4
    om = oracle.Mesh(arrs, mesh.radius)
    rp, col, val = om.grad_csr()
    assert abs(ps.G - sp.csr_matrix((val, col, rp), shape=(om.E, om.N))).max() == 0.0
    rp, col, val = om.div_csr()
    D = sp.csr_matrix((val, col, rp), shape=(om.N, om.E))
    assert abs(ps.D - D).max() <= 1e-15 * abs(D).max()
    rp, col, val = om.coriolis_csr(ps.omega)
    C = sp.csr_matrix((val, col, rp), shape=(om.E, om.E))
    assert abs(ps.C - C).max() <= 1e-15 * abs(C).max()


@pytest.mark.parametrize("h,alpha,tide,obl", [(38e3, 1e-8, "ECC", 0.0), (2e3, 1e-11, "FULL", 0.6),
                                              (90e3, 1e-7, "OBLIQ", 0.3)])
def test_periodic_state_satisfies_the_oracle_tendencies(product, oracle, level4, h, alpha, tide, obl):
    """v(t) = Re(V e^{iwt}), eta(t) = Re(H e^{iwt}) from the sparse solve
    satisfy the oracle's own semi-discrete equations at any t: dv/dt =
    momentum(v, eta) - alpha v + G U(t) (src/updateMomentum.cpp:16-46,
    src/timeIntegrator.cpp:219,242) and d(eta)/dt = h Div v
    (src/updateEta.cpp:7-44), with U(t) the oracle's forcing."""
    G, O = product, oracle
    mesh, arrs, ps, _ = level4
    p, _ = G.make_params(h=h, alpha=alpha, tide=getattr(G, tide), obliquity=np.radians(obl))
    V, H = ps.solve(p)
    om = O.Mesh(arrs, mesh.radius)
    sim = O.Sim(om, O.make_params(g=p.g, h=p.h, alpha=p.alpha, omega=p.omega, radius=p.radius, ecc=p.ecc,
                                  obliquity=p.obliquity, dt=p.dt, tide=p.tide))
    for t in (0.0, 23456.0, 90000.0):
        z = np.exp(1j * p.omega * t)
        v, eta = (V * z).real, (H * z).real
        dv = sim.momentum(v, eta) - p.alpha * v + ps.G @ sim.forcing(t)
        assert rel_l2(dv, (1j * p.omega * V * z).real) < 1e-9
        assert rel_l2(p.h * (ps.D @ v), (1j * p.omega * H * z).real) < 1e-12


def test_state_basis_round_trip(product, level4):
    """Synthesis then projection returns any coefficient vector (the fit is
    exact on its own span); the velocity part is not degenerate (chi and psi
    from degree 1)."""
    mesh, arrs, _, sb = level4
    rng = np.random.default_rng(2)
    c = rng.normal(size=sb.size)
    v, eta = sb.synthesize(c)
    assert rel_l2(sb.project(v, eta), c) < 1e-9
    assert np.linalg.matrix_rank(sb.Av) == 2 * (sb.K - 1)


def test_tide_superposition(product, level4):
    """The LTE is linear in the tide: the FULL response at obliquity theta
    is the ECC response plus theta / 1 degree times the OBLIQ response at
    1 degree, as the network's outputs are combined."""
    from geodesicodis_amd.warmstart import combine
    G = product
    mesh, arrs, ps, sb = level4
    kw = dict(h=7e3, alpha=3e-9)
    V_full, H_full = ps.solve(G.make_params(tide=G.FULL, obliquity=np.radians(0.7), **kw)[0])
    V_e, H_e = ps.solve(G.make_params(tide=G.ECC, **kw)[0])
    V_o, H_o = ps.solve(G.make_params(tide=G.OBLIQ, obliquity=np.radians(1.0), **kw)[0])
    pair = np.concatenate([sb.project(V_e.real, H_e.real), sb.project(V_o.real, H_o.real)])[None, :]
    row = np.array([[np.log10(7e3), np.log10(3e-9), 0.7, G.FULL]])
    want = sb.project(V_full.real, H_full.real)
    assert rel_l2(combine(row, pair)[0], want) < 1e-9
