// odis_sh.cpp -- real spherical-harmonic basis on the grid nodes and the
// Love-number factors of the self-gravity / ice-shell coupling.
//
// The reference's design (commented out: CalcLegendreFuncs,
// src/mesh.cpp:2154-2259, and pressureGradientSH, src/spatialOperators.cpp:
// 387-460) builds a node x harmonic matrix of 4pi-normalised associated
// Legendre functions times cos(m lon) / sin(m lon), weighted per degree by
// beta_l (ice shell) or the loading factor (self-gravity), and adds
// sh_matrix * coefficients to the surface height seen by the pressure
// gradient.  Here the coefficients come from a least-squares fit on the
// nodes (as SHExpandLSQ does, src/extractSHCoeffGG.f95:29).
//
// Index convention: k = l*l + l + m, m = -l..l; m >= 0 is P_lm(sin lat)
// cos(m lon), m < 0 is P_l|m|(sin lat) sin(|m| lon); P_lm is 4pi-normalised
// without the Condon-Shortley phase.
#include <cmath>
#include <stdexcept>
#include <vector>

#include "odis_host.hpp"

namespace odis {

void sh_recurrence_tables(int L, std::vector<double> &a, std::vector<double> &b, std::vector<double> &sect,
                          std::vector<double> &f1) {
  const size_t n = size_t(L + 1) * (L + 2) / 2;
  a.assign(n, 0.0);
  b.assign(n, 0.0);
  sect.assign(L + 1, 0.0);
  f1.assign(L + 1, 0.0);
  for (int m = 0; m <= L; m++) {
    sect[m] = m == 0 ? 1.0 : (m == 1 ? std::sqrt(3.0) : std::sqrt((2.0 * m + 1.0) / (2.0 * m)));
    f1[m] = std::sqrt(2.0 * m + 3.0);
    for (int l = m + 2; l <= L; l++) {
      const double ll = l, mm = m;
      a[sh_lm_index(l, m)] = std::sqrt((2 * ll + 1) * (2 * ll - 1) / ((ll - mm) * (ll + mm)));
      b[sh_lm_index(l, m)] =
          std::sqrt((2 * ll + 1) * (ll - mm - 1) * (ll + mm - 1) / ((2 * ll - 3) * (ll - mm) * (ll + mm)));
    }
  }
}

static void basis_row(int L, double lat, double lon, const std::vector<double> &a, const std::vector<double> &b,
                      const std::vector<double> &sect, const std::vector<double> &f1, double *Y) {
  const int K = (L + 1) * (L + 1);
  for (int k = 0; k < K; k++) Y[k] = 0.0;
  const double x = std::sin(lat), s = std::cos(lat);
  // P_mm by the sectoral recurrence, then upward in l at fixed m
  double pmm = 1.0;  // P_00
  for (int m = 0; m <= L; m++) {
    if (m == 1) pmm = sect[1] * s;
    else if (m > 1) pmm *= sect[m] * s;
    const double cm = std::cos(m * lon), sm = std::sin(m * lon);
    double p2 = 0.0, p1 = pmm;  // P_{l-2,m}, P_{l-1,m}
    for (int l = m; l <= L; l++) {
      double p;
      if (l == m) {
        p = pmm;
      } else if (l == m + 1) {
        p = f1[m] * x * pmm;
      } else {
        p = a[sh_lm_index(l, m)] * x * p1 - b[sh_lm_index(l, m)] * p2;
      }
      if (l > m) {
        p2 = p1;
        p1 = p;
      }
      Y[l * l + l + m] = p * cm;
      if (m > 0) Y[l * l + l - m] = p * sm;
    }
  }
}

void sh_basis_row(int L, double lat, double lon, double *Y) {
  std::vector<double> a, b, sect, f1;
  sh_recurrence_tables(L, a, b, sect, f1);
  basis_row(L, lat, lon, a, b, sect, f1, Y);
}

void sh_basis(int L, const std::vector<double> &lat, const std::vector<double> &lon, std::vector<double> &Yt) {
  if (L < 0 || L > 200) throw std::runtime_error("sh degree out of range");
  const int K = (L + 1) * (L + 1);
  const size_t N = lat.size();
  Yt.assign(size_t(K) * N, 0.0);
  std::vector<double> row(K), a, b, sect, f1;
  sh_recurrence_tables(L, a, b, sect, f1);
  for (size_t i = 0; i < N; i++) {
    basis_row(L, lat[i], lon[i], a, b, sect, f1, row.data());
    for (int k = 0; k < K; k++) Yt[size_t(k) * N + i] = row[k];
  }
}

// Per-degree factor beta_l of the pressure term for a surface type; 1 means
// "no change" (free surface without self-gravity).
//   FREE_LOADING: 1 - loading_factor[l] (src/boundaryConditions.cpp:29-77)
//   LID_MEMBR:    beta_l of membraneNuBeta (src/membraneConstants.cpp:126-128)
std::vector<double> surface_beta(int surface, int L, double radius, double g, double shell, double ocean) {
  std::vector<double> beta(L + 1, 1.0);
  if (surface == ODIS_FREE_LOADING) {
    const double ocean_den = 1000.0, bulk_den = 1609.22, rig = 40e9;
    for (int l = 1; l <= L; l++) {
      const double eff = double(2 * l * l + 4 * l + 3) / double(l) * rig / (bulk_den * g * radius);
      const double kl = -(1.0 / (1.0 + eff)), hl = -(1.0 / (1.0 + eff)) * (2.0 * l + 1.0) / 3.0;
      const double lf = (1.0 + kl - hl) * 3. * ocean_den / ((2. * l + 1.0) * bulk_den);
      beta[l] = 1.0 - lf;
    }
  } else if (surface == ODIS_LID_MEMBR) {
    std::vector<double> nu, b;
    double ro, go;
    membrane_nu_beta(radius, shell, ocean, L, nu, b, &ro, &go);
    for (int l = 0; l <= L; l++) beta[l] = b[l];
  }
  return beta;
}

}  // namespace odis
