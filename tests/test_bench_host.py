"""The bench line's roofline arithmetic without a GPU: the algorithmic
flops and bytes per launch that `bench.py` divides by the measured kernel
times, rebuilt here from their per-unit counts, and the HBM traffic it reads
from the committed PMC summary."""
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
