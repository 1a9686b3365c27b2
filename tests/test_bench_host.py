"""The bench line's host logic without a GPU: the algorithmic flops and
bytes per launch that `bench.py` divides by the measured kernel times,
rebuilt here from their per-unit counts; the HBM traffic it reads from the
committed PMC summary; and the seeded draws of the sweep (256 instances)
and of the uneven batch (512)."""
import numpy as np

import bench

N, E, L = 163842, 491520, 32


def test_sh_flops_from_the_harmonics():
    """2 K N algorithmic flops per SH pass (one multiply-add per cell and
    harmonic); the recurrence issues, per cell and (l, m >= 0), a step of 3
    flops and the cos and (m > 0) sin multiply-adds."""
    K = sum(2 * l + 1 for l in range(L + 1))
    assert K == (L + 1) ** 2 == 1089
    assert bench.sh_alg_flops(N, K) == 2 * K * N == 356847876
    per_cell = sum(3 + 2 + (2 if m > 0 else 0) for l in range(L + 1) for m in range(l + 1))
    assert per_cell == 3861
    assert bench.sh_flops(N, L) == N * per_cell


def test_step_kernel_bytes_per_face_and_cell():
    """The one-kernel step's algorithmic bytes, item by item (p = 8 bytes per
    FP64 value, 2 forcing fields):
    per face 2 int32 gradient indices, 2 gradient coefficients, Coriolis,
    area, 5 state values (v, 2 history reads, history and v written), d_e
    and 2 TRiSK partial sums read, l_e, 2 partial sums written;
    per cell eta, 2 forcing fields, a 6-entry uint16 face cycle, 6 kite
    fractions, a 4-byte cycle record, area, eta written, 3 history values."""
    p = 8
    face = 2 * 4 + 2 * p + p + p + 5 * p + 3 * p + p + 2 * p
    cell = p + 2 * p + 6 * 2 + 6 * p + 4 + p + p + 3 * p
    assert (face, cell) == (128, 128)
    assert bench.step_kernel_bytes(E, N, 2, p) == face * E + cell * N == 83886336
    # FP32 state: the same items at p = 4
    assert bench.step_kernel_bytes(E, N, 2, 4) == (2 * 4 + 15 * 4) * E + (12 + 4 + 14 * 4) * N


def test_traffic_from_the_committed_pmc_summary():
    """The PMC bytes per launch of the headline's dominant kernel come from
    profiles/traffic.json: FETCH_SIZE x 1.9998 + WRITE_SIZE, in KiB."""
    t = bench.measured_traffic("k_sh_analysis_pair", "double", N)
    assert t == round((4532.7 * 1.9998 + 265.7 * 1.0) * 1024)
    assert bench.measured_traffic("k_sh_analysis_pair", "double", N + 1) is None


def test_sweep_instances_follow_the_seeded_draws(product):
    """The 256-instance sweep: thickness 10^U(3, 5) m and linear drag
    10^U(-11, -7) s^-1, drawn in that order per instance from
    default_rng(20261203); eccentricity tide; each dt the quantised step
    below 0.35 / omega_max(h)."""
    G = product
    ps = bench.sweep_params(G, 256)
    rng = np.random.default_rng(20261203)
    for p in ps:
        h = 10.0 ** rng.uniform(3.0, 5.0)
        a = 10.0 ** rng.uniform(-11.0, -7.0)
        assert (p.h, p.alpha, p.tide) == (h, a, G.ECC)
        assert p.dt * bench.OMEGA_MAX_40962_38KM * (h / 38e3) ** 0.5 <= 0.35
    hs = np.array([p.h for p in ps])
    assert 1e3 <= hs.min() and hs.max() <= 1e5 and len(set(hs)) == 256


def test_uneven_groups_follow_the_seeded_draws(product):
    """The 512-instance uneven batch: (level, obliquity, quadratic-drag
    coefficient, surface) drawn in that order per instance from
    default_rng(20261204), grouped by (level, shell); FULL tide, quadratic
    drag, the 20 km LID_MEMBR shell on half of them."""
    G = product
    groups, beta = bench.uneven_groups(G, 512)
    rng = np.random.default_rng(20261204)
    want = {}
    for _ in range(512):
        level = int(rng.choice([4, 5, 6, 7]))
        theta, cd, lid = rng.uniform(0.0, 1.0), rng.uniform(1e-3, 5e-3), bool(rng.uniform() < 0.5)
        want.setdefault((level, lid), []).append((theta, cd))
    assert sorted(groups) == sorted(want) and len(groups) == 8
    for key, ps in groups.items():
        got = np.array([(np.degrees(p.obliquity), p.alpha) for p in ps])
        assert got.shape == (len(want[key]), 2) and np.allclose(got, want[key], rtol=1e-14, atol=0)
        for p in ps:
            assert p.tide == G.FULL and p.friction == G.QUADRATIC and p.quad_drag == 1
            assert (p.surface == G.LID_MEMBR) == key[1]
            if key[1]:
                assert p.shell_thickness == 20e3
    assert sum(len(v) for v in groups.values()) == 512
    assert beta.shape == (bench.UNEVEN_SH_DEGREE + 1,) and np.all(np.diff(beta[2:]) > 0)


def test_more_than_one_gpu_is_refused():
    """`bench.py --gpus N` with N > 1 exits with status 2 and says that this
    build runs on one GPU, before touching any device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "exactly one GPU" in r.stderr and r.stdout == ""
