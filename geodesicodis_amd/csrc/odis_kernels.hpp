// odis_kernels.hpp -- device kernels of the LTE step (gfx950).
//
// One step of the reference loop (src/timeIntegrator.cpp:207-277) is two
// kernels:
//
//   face kernel (one lane per face e, reads state n):
//     v_t   = sum_j c_ej v[f_j]                 TRiSK, <=10 terms (src/interpolation.cpp:43-46)
//             compact form: (T(inner, e) - T(outer, e)) / d_e, the partial
//             sums written by the previous cell kernel (odis_compact.cpp)
//     diss += A_e (v_e^2 + v_t^2)               dissipation of state n (src/energy.cpp:31-44)
//     f0    = -g (G eta)_e + fcor_e v_t         updateMomentum (src/updateMomentum.cpp:42)
//     drag  = -alpha v_e + sum_k c_k(t+dt) (G S_k)_e      D v + G U (src/timeIntegrator.cpp:219)
//     v'    = v + dt (a f0 + b f1 + c f2) + dt drag        AB3 + Euler drag (src/temporalOperators.cpp:36-47)
//   cell kernel (one lane per cell i, reads v'):
//     f0    = h (Div v')_i ; eta' = eta + dt (a f0 + b f1 + c f2)      updateEta + AB3
//     compact form: also T(i, e) of v' for each face e of the cell
//     workgroup 0 also sums the face kernel's dissipation partials of this
//     step in a fixed order (deterministic FP64 result).
//
// The step index n and the forcing time coefficients c_k(t_n + dt) are
// computed on the host (glibc cos/sin, as the reference's forcing does) and
// passed per step: n as a kernel argument, c_k through a small table that is
// uploaded once per chunk of steps.  No device-scope atomics or fences are
// used inside a step: kernel boundaries order the two kernels.
//
// Layout: structure-of-arrays, stencil slot-major ([slot][face]) so each
// slot load is one coalesced 64-lane access; AB3 history is a 3-slot ring
// indexed by step % 3 (no history shift traffic); v, eta and the TRiSK
// partial sums are double-buffered (state n read, state n+1 written) because
// neighbours read the old values.
#pragma once

#include <hip/hip_runtime.h>

#include "odis.h"
#include "odis_host.hpp"

namespace odis {

constexpr int kMaxForcing = 7;
constexpr int kBlock = 256;

struct TideDev {
  int tide, nfields;
  double omega, dt;
  double f_ecc, f_obl, f_west, f_full2, ecc, obl;  // amplitudes
  double pl_a1, pl_a2, pl_gm;                     // PLANET
  double radius;
};

template <typename T>
struct FaceArgs {
  int E, nb;                 // faces, workgroups per instance
  int B;                     // instances
  const int *tri_idx;        // [10][E]
  const T *tri_coef;         // [10][E]
  const int *g_idx;          // [2][E]
  const T *g_coef;           // [2][E]
  const T *fcor;             // [E]  -2 Omega sin(lat_e)
  const T *farea;            // [E]
  const T *fgs;              // [K][E] G S_k
  int nfields;
  // per instance (instance stride in elements)
  const T *v_cur;            // [B][E]
  T *v_nxt;                  // [B][E]
  T *ring;                   // [B][3][E]
  const T *eta;              // [B][N]
  long long v_stride, ring_stride, eta_stride;
  double *partials;          // [2][B][nb] by step parity
  const double *inst;        // [B][8] g, alpha, dt, init_load, quad_coef, h
  int quad_drag;
  // cell-local TRiSK form (COMPACT)
  int N;
  const T *fdist;            // [E] d_e
  const T *ts;               // [B][2][E] TRiSK partial sums T(inner, e), T(outer, e) of v_cur
  long long ts_stride;
  const T *cellS;            // [K][N] forcing fields at the cells
};

template <typename T>
struct CellArgs {
  int N, nb, nb_face;
  const int *div_idx;        // [6][N]
  const T *div_coef;         // [6][N]
  const T *v_new;            // [B][E]
  const T *eta;              // [B][N] eta^n
  T *eta_nxt;                // [B][N] eta^{n+1}
  T *ring;                   // [B][3][N]
  long long v_stride, eta_stride, ring_stride;
  const double *partials;    // [2][B][nb_face] by step parity
  double *series;            // [B][cap]
  int *nonfinite;            // [B]
  int cap;
  const double *inst;        // [B][8]
  const double *dconst;      // [B] dissipation constant
  // cell-local TRiSK form (COMPACT)
  int E;
  const int *cyc;            // [6][N] faces counter-clockwise (pentagons: slot 5 = slot 0)
  const T *kite;             // [6][N] kite fraction of the vertex after cyc[q]
  const int *info;           // [N] k | dir_bits << 3 (bit q: outer node of cyc[q])
  const T *area;             // [N] control-volume area A_i
  const T *flen;             // [E] l_e
  T *ts;                     // [B][2][E] TRiSK partial sums of v_new (written)
  long long ts_stride;
};

// Workgroup i runs on XCD i % 8; give each XCD a contiguous range of
// logical tiles so the TRiSK gathers of a tile hit that XCD's L2.
__device__ __forceinline__ int xcd_tile(int i, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = i & 7, k = i >> 3;
  return x * q + min(x, r) + k;
}

// Forcing time coefficients c_k(t) of the spatial fields S_k (see
// tide_fields in odis_device.hip for the field list of each potential).
// Formulas of src/tidalPotentials.cpp:82-225 regrouped by spatial field.
__device__ __host__ inline void tide_coeffs(const TideDev &p, double t, double *c) {
  const double M = p.omega * t;
  switch (p.tide) {
    case ODIS_ECC: {
      c[0] = p.f_ecc * cos(M);
      c[1] = p.f_ecc * sin(M);
    } break;
    case ODIS_OBLIQ:
      c[0] = p.f_obl * cos(M);
      break;
    case ODIS_OBLIQ_WEST:
      c[0] = 3 * p.f_west * cos(M);
      c[1] = -3 * p.f_west * sin(M);
      break;
    case ODIS_FULL:
      c[0] = p.f_ecc * cos(M);
      c[1] = p.f_ecc * sin(M);
      c[2] = p.f_obl * cos(M);
      break;
    case ODIS_FULL2: {
      const double e = p.ecc, o = p.obl, f = p.f_full2;
      const double cM = cos(M), sM = sin(M), c2 = cos(2 * M), c3 = cos(3 * M), c4 = cos(4 * M);
      double T1 = 3. * e * (4. - 7. * o * o) * cM + 6 * (o * o + e * e * (3 - 7 * o * o)) * c2;
      T1 += 3 * e * o * o * (7 * c3 + 17 * e * c4);
      c[0] = f * T1;
      c[1] = f * 24 * o * sM * (4 + 15 * e * e + 20 * e * cM + 43 * e * e * c2);
      c[2] = f * 24 * o * sM * (2 * e * (4 + 25 * e * cM) * sM);
      const double a3 = o * o * (2 + 3 * e * e + 6 * e * cM + 9 * e * e * c2);
      c[3] = f * 6 * a3 * cM;
      c[4] = f * 6 * a3 * sM;
      c[5] = f * 6 * (-(o * o - 2)) * (6 * e * cM + 17 * e * e * c2);
      c[6] = f * 6 * (-(o * o - 2)) * (2 * e * (4 + 17 * e * cM) * sM);
    } break;
    case ODIS_PLANET: {
      const double nij = 2.0 * p.omega - p.omega;
      const double cn = cos(nij * t), sn = sin(nij * t);
      const double a1 = p.pl_a1, a2 = p.pl_a2;
      const double pp = a1 * a1 + a2 * a2 - 2. * a1 * a2 * cn;
      const double cph = a1 - a2 * cn, sph = a2 * sn;
      const double f = 0.5 * p.pl_gm * (p.radius / pp) * (p.radius / pp) / sqrt(pp);
      c[0] = 3 * f * cph * cph;
      c[1] = 6 * f * cph * sph;
      c[2] = 3 * f * sph * sph;
      c[3] = -f * pp;
    } break;
    default:
      break;
  }
}

}  // namespace odis
