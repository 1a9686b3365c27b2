// odis_device.hip -- HIP kernels and the device context behind odis.h.
//
// Replaces the body of ab3Explicit (src/timeIntegrator.cpp:57-322) for one
// or more ocean instances sharing one grid.  See odis_kernels.hpp for the
// per-step dataflow.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "odis.h"
#include "odis_kernels.hpp"

namespace odis {

// ===========================================================================
// kernels
// ===========================================================================

// Fixed-shape block reduction (NW waves of 64): wave-level shuffle, then
// the wave sums in a fixed tree.  The result is valid on thread 0.
template <int NW>
__device__ __forceinline__ double block_sum(double x, double *sh) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = x;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    if (NW == 1) r = sh[0];
    else if (NW == 2) r = sh[0] + sh[1];
    else if (NW == 4) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    else r = ((sh[0] + sh[1]) + (sh[2] + sh[3])) + ((sh[4] + sh[5]) + (sh[6] + sh[7]));
  }
  return r;
}

// Sum count partials in a fixed order (one workgroup of NW waves).
template <int NW>
__device__ __forceinline__ double sum_partials(const double *p, int count, double *sh) {
  double acc = 0.0;
  for (int j = threadIdx.x; j < count; j += NW * 64) acc += p[j];
  return block_sum<NW>(acc, sh);
}

// Index of element i of instance b in a per-instance array of `stride`
// elements per instance: instance-major (b stride + i) or, for a batch held
// interleaved (IL), in groups of kIlGroup instances, each group
// element-major with its instances innermost
// (((b / kIlGroup) stride + i) kIlGroup + b % kIlGroup), so that a group's
// values of consecutive elements are contiguous.
constexpr int kIlGroup = 16;
template <bool IL>
__device__ __host__ __forceinline__ long long sidx(int b, long long stride, long long i) {
  return IL ? ((long long)(b / kIlGroup) * stride + i) * kIlGroup + b % kIlGroup : b * stride + i;
}

// Result of the momentum update of one face.
template <typename T>
struct FaceOut {
  T vn;           // v^{n+1}
  T f0;           // momentum tendency (history slot n % 3)
  double energy;  // dissipation integrand of state n: A_e (v_e^2 + v_t^2) (or ^{3/2})
};

// Tangential velocity, dissipation integrand and (unless DISS_ONLY) the AB3
// momentum update of face e of instance b:
//   f0 = -g (G eta)_e + fcor_e v_t [- c |u| v_e]   (src/updateMomentum.cpp:42)
//   v' = v + dt AB3(f0) + dt (-alpha v_e + (G U)_e)   (src/timeIntegrator.cpp:207-232)
// COMPACT: v_t from the TRiSK partial sums and G U from the forcing fields at
// the cells; else the tabulated stencil and per-face forcing gradients.
// IL (compact form only): the state is held interleaved; instances past
// a.B (padding of the last group) take the parameters of the last one.
template <typename T, bool DISS_ONLY, bool COMPACT, bool IL = false>
__device__ __forceinline__ FaceOut<T> face_update(const FaceArgs<T> &a, int e, int b, int energy_quad,
                                                  const double *table, int s) {
  static_assert(COMPACT || !IL, "the interleaved layout has the compact form only");
  const int E = a.E;
  const int bp = IL ? min(b, a.B - 1) : b;
  const double *ip = a.inst + bp * 8;
  FaceOut<T> out;
  T vt;
  if (COMPACT) {
    vt = (a.ts[sidx<IL>(b, a.ts_stride, e)] - a.ts[sidx<IL>(b, a.ts_stride, E + e)]) / a.fdist[e];
  } else {
    const T *v = a.v_cur + b * a.v_stride;
    vt = T(0);
#pragma unroll
    for (int j = 0; j < 10; ++j) vt += a.tri_coef[j * E + e] * v[a.tri_idx[j * E + e]];
  }
  const T ve = a.v_cur[sidx<IL>(b, a.v_stride, e)];
  const T sp = ve * ve + vt * vt;
  const T en = energy_quad ? sqrt(sp) * sp : sp;
  out.energy = double(a.farea[e]) * double(en);
  out.vn = T(0);
  out.f0 = T(0);
  if (!DISS_ONLY) {
    const long long n = (long long)table[0] + s;
    const double *coef = table + 8 + size_t(s) * a.B * 8;
    const T g = T(ip[0]), alpha = T(ip[1]), dt = T(ip[2]), qd = T(ip[4]);
    const bool ab3 = (n >= 2) || (ip[3] != 0.0);
    const int c0 = a.g_idx[e], c1 = a.g_idx[E + e];
    const T gc0 = a.g_coef[e], gc1 = a.g_coef[E + e];
    const T ge = gc0 * a.eta[sidx<IL>(b, a.eta_stride, c0)] + gc1 * a.eta[sidx<IL>(b, a.eta_stride, c1)];
    T f0 = -g * ge + a.fcor[e] * vt;
    if (a.quad_drag) f0 -= qd * sqrt(sp) * ve;
    // G U, U = sum_k c_k(t + dt) S_k (src/timeIntegrator.cpp:219)
    T fu;
    if (COMPACT) {
      T u0 = T(0), u1 = T(0);
      for (int k = 0; k < a.nfields; ++k) {
        const T ck = T(coef[bp * 8 + k]);
        u0 += ck * a.cellS[k * a.N + c0];
        u1 += ck * a.cellS[k * a.N + c1];
      }
      fu = gc0 * u0 + gc1 * u1;
    } else {
      fu = T(0);
      for (int k = 0; k < a.nfields; ++k) fu += T(coef[b * 8 + k]) * a.fgs[k * E + e];
    }
    const T drag = -alpha * ve + fu;
    const int s1 = int((n + 2) % 3), s2 = int((n + 1) % 3);
    T upd;
    if (ab3)
      upd = (T(23. / 12.) * f0 + T(-16. / 12.) * a.ring[sidx<IL>(b, a.ring_stride, (long long)s1 * E + e)] +
             T(5. / 12.) * a.ring[sidx<IL>(b, a.ring_stride, (long long)s2 * E + e)]) *
            dt;
    else
      upd = f0 * dt;
    out.f0 = f0;
    out.vn = (ve + upd) + dt * drag;
  }
  return out;
}

// Face kernel.  DISS_ONLY: evaluate the dissipation partials of the current
// state without stepping (output cadence / final state).
// The step index and forcing coefficients come from a host-filled table:
// table[0] = n0, then per step s and instance b 8 coefficients at
// table[8 + (s * B + b) * 8]; step s of the table is n = n0 + s.  Kernels
// captured in a graph therefore stay valid for every chunk of steps.
template <typename T, bool DISS_ONLY, bool COMPACT>
__global__ __launch_bounds__(kBlock) void k_face(FaceArgs<T> a, int energy_quad, const double *table, int s) {
  __shared__ double sh[4];
  const int b = blockIdx.y;
  const int tile = xcd_tile(blockIdx.x, a.nb);
  const int E = a.E;
  const int e = tile * kBlock + threadIdx.x;
  double local = 0.0;
  if (e < E) {
    const FaceOut<T> r = face_update<T, DISS_ONLY, COMPACT>(a, e, b, energy_quad, table, s);
    local = r.energy;
    if (!DISS_ONLY) {
      const long long n = (long long)table[0] + s;
      a.ring[b * a.ring_stride + int(n % 3) * E + e] = r.f0;
      a.v_nxt[b * a.v_stride + e] = r.vn;
    }
  }
  const double part = block_sum<kBlock / 64>(local, sh);
  const long long par = DISS_ONLY ? 0 : (((long long)table[0] + s) & 1);
  if (threadIdx.x == 0) a.partials[(par * gridDim.y + b) * a.nb + tile] = part;
}

// TRiSK partial sums of a cell with K faces: for each face p of the cycle,
// T_p = sum_{j=1}^{K-1} (R_p + ... + R_{p+j-1} - 1/2) F_{p+j}, F_q = dir_q l_q v_q
// (indices mod K).  The running sum of kite fractions starts at the face
// itself, in the order of the reference's weight loop (src/mesh.cpp:677-910).
template <typename T, int K, bool IL>
__device__ __forceinline__ void trisk_partials(const T *F, const T *R, const int *f, int db, T *ts, int E, int b,
                                               long long stride) {
#pragma unroll
  for (int p = 0; p < K; ++p) {
    T S = R[p];
    T acc = (S - T(0.5)) * F[(p + 1) % K];
#pragma unroll
    for (int j = 2; j < K; ++j) {
      S += R[(p + j - 1) % K];
      acc += (S - T(0.5)) * F[(p + j) % K];
    }
    ts[sidx<IL>(b, stride, (long long)((db >> p) & 1) * E + f[p])] = acc;
  }
}

// TRiSK partial sums of a cell with k faces.
template <typename T, bool IL = false>
__device__ __forceinline__ void cell_outputs(const CellArgs<T> &a, int b, int k, int db, const int *f, const T *F,
                                             const T *R) {
  if (k == 6)
    trisk_partials<T, 6, IL>(F, R, f, db, a.ts, a.E, b, a.ts_stride);
  else
    trisk_partials<T, 5, IL>(F, R, f, db, a.ts, a.E, b, a.ts_stride);
}

// Continuity + AB3 of cell i: eta^{n+1} = eta^n + dt AB3(h Div v^{n+1})
// (src/updateEta.cpp, src/temporalOperators.cpp:36-47), history slot n % 3.
// eta_old = eta^n, r1 / r2 = history slots (n+2) % 3 and (n+1) % 3, already
// loaded; h, dt and the AB3 flag are the instance's (continuity_loaded reads
// them).
template <typename T, bool IL = false>
__device__ __forceinline__ T continuity_with(const CellArgs<T> &a, int i, int b, long long n, T div, T eta_old,
                                             T r1, T r2, T h, T dt, bool ab3) {
  const T f0 = h * div;
  T upd;
  if (ab3)
    upd = (T(23. / 12.) * f0 + T(-16. / 12.) * r1 + T(5. / 12.) * r2) * dt;
  else
    upd = f0 * dt;
  a.ring[sidx<IL>(b, a.ring_stride, (long long)int(n % 3) * a.N + i)] = f0;
  const T eta_new = eta_old + upd;
  a.eta_nxt[sidx<IL>(b, a.eta_stride, i)] = eta_new;
  return eta_new;
}

template <typename T, bool IL = false>
__device__ __forceinline__ T continuity_loaded(const CellArgs<T> &a, int i, int b, long long n, T div,
                                               T eta_old, T r1, T r2) {
  const double *ip = a.inst + b * 8;
  const bool ab3 = (n >= 2) || (ip[3] != 0.0);
  return continuity_with<T, IL>(a, i, b, n, div, eta_old, r1, r2, T(ip[5]), T(ip[2]), ab3);
}

template <typename T>
__device__ __forceinline__ T continuity(const CellArgs<T> &a, int i, int b, long long n, T div) {
  const T *ring = a.ring + b * a.ring_stride;
  const int N = a.N;
  return continuity_loaded(a, i, b, n, div, a.eta[b * a.eta_stride + i], ring[int((n + 2) % 3) * N + i],
                           ring[int((n + 1) % 3) * N + i]);
}

// Divergence sum_q -dir_q l_q v_q / A_i (src/mesh.cpp:3242-3247) from the
// signed fluxes F_q = dir_q l_q v_q.
template <typename T>
__device__ __forceinline__ T divergence(const T *F, int k, T area) {
  T sum = F[0];
  sum += F[1];
  sum += F[2];
  sum += F[3];
  sum += F[4];
  if (k == 6) sum += F[5];
  return -sum / area;
}

// Compact-form cell update from the cell's signed fluxes F_q = dir_q l_q v_q
// (global face ids f): divergence, continuity (unless PREP), then the TRiSK
// partial sums of the cell's faces.
template <typename T, bool PREP>
__device__ __forceinline__ void cell_compact(const CellArgs<T> &a, int i, int b, long long n, int k, int db,
                                             const int *f, const T *F, const T *R) {
  if (!PREP) continuity(a, i, b, n, divergence(F, k, a.area[i]));
  cell_outputs(a, b, k, db, f, F, R);
}

// Cell kernel: continuity + AB3 into eta^{n+1}; the first workgroup of each
// instance also reduces the face kernel's dissipation partials of this step.
// COMPACT: divergence from the cell's face cycle, and the TRiSK partial sums
// of v_new for the next face kernel.  PREP (COMPACT only): the partial sums
// alone, for a velocity just uploaded.
template <typename T, bool COMPACT, bool PREP>
__global__ __launch_bounds__(kBlock) void k_cell(CellArgs<T> a, const double *table, int s_off) {
  __shared__ double sh[4];
  const int b = blockIdx.y;
  const int tile = xcd_tile(blockIdx.x, a.nb);
  const int N = a.N;
  const int i = tile * kBlock + threadIdx.x;
  const long long n = PREP ? 0 : (long long)table[0] + s_off;
  if (i < N) {
    const T *v = a.v_new + b * a.v_stride;
    if (COMPACT) {
      const int info = a.info[i];
      int f[6];
      T F[6], R[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        f[q] = a.cyc[q * N + i];
        R[q] = a.kite[q * N + i];
        const T x = a.flen[f[q]] * v[f[q]];
        F[q] = ((info >> (3 + q)) & 1) ? -x : x;
      }
      cell_compact<T, PREP>(a, i, b, n, info & 7, info >> 3, f, F, R);
    } else {
      T div = T(0);
#pragma unroll
      for (int j = 0; j < 6; ++j) div += a.div_coef[j * N + i] * v[a.div_idx[j * N + i]];
      continuity(a, i, b, n, div);
    }
  }
  if (PREP || blockIdx.x != 0) return;
  // the first-dispatched workgroup of each instance reduces the face
  // kernel's dissipation partials of this step (complete: previous kernel)
  const double tot =
      sum_partials<kBlock / 64>(a.partials + ((n & 1) * gridDim.y + b) * size_t(a.nb_face), a.nb_face, sh);
  if (threadIdx.x == 0) {
    const double d = tot * a.dconst[b];
    a.series[size_t(b) * a.cap + size_t(n % a.cap)] = d;
    if (!isfinite(d)) a.nonfinite[b] = 1;
  }
}

// ---------------------------------------------------------------------------
// one-kernel step over patches of cells (compact form)
// ---------------------------------------------------------------------------
// Workgroup j (4 P lanes) owns the P consecutive cells [j P, (j+1) P) of
// the Hilbert order and the faces whose lower-ranked cell is among them (a
// contiguous face range: faces are sorted by cell ranks).  It updates every
// face of its cells -- owned ones, and the "foreign" faces on the patch
// boundary owned by an earlier patch, which that patch computes identically
// -- keeps l_e v_e^{n+1} of all of them in LDS, writes the owned ones, then
// does continuity and the TRiSK partial sums of its cells from LDS.  Nothing
// written by one workgroup is read by another in the same launch: v, eta
// and the partial sums are double-buffered, the history slot written (n % 3)
// is not one read.  The dissipation partials of step n are summed by the
// first workgroup of the next launch (PENDING) or by k_series at the end of
// a chunk.
template <typename T>
struct StepArgs {
  FaceArgs<T> f;
  CellArgs<T> c;
  const int *cell_lo;           // [n_patch + 1] first cell of each patch
  const int *own_lo;            // [n_patch + 1] first owned face of each patch
  const int *for_ptr;           // [n_patch + 1] foreign face list offsets
  const int *for_idx;           // foreign faces
  const unsigned short *cyc_loc;  // [6][N] cycle faces, patch-local index
  int n_patch;
  int n_blocks;                 // workgroups per instance (= gridDim.x); each runs a contiguous run of patches
  int max_loc;                  // most faces (owned + foreign) of a patch
  const T *face_rec;            // [E][16] operand records of k_step_il (kRecSlots), or null
  const T *cell_rec;            // [N][16]
  int il_band;                  // k_step_il in band order: groups per band
};

// One patch of the one-kernel step; returns this lane's dissipation
// integrand sum over the owned faces it updated.
template <typename T, int P, int LPC>
__device__ __forceinline__ double step_patch(const StepArgs<T> &a, int j, int b, long long n, int energy_quad,
                                             const double *table, int s, T *lv, int *gid) {
  constexpr int NT = LPC * P;
  const int E = a.f.E, N = a.f.N;
  const int lo = a.own_lo[j], n_own = a.own_lo[j + 1] - lo;
  const int flo = a.for_ptr[j], n_loc = n_own + (a.for_ptr[j + 1] - flo);
  // the last P lanes (fewest faces) also own the patch cells; their cycle,
  // kite fractions and state are loaded before the face phase so that those
  // loads overlap it
  const int ci = int(threadIdx.x) - (NT - P);
  const int i = a.cell_lo[j] + ci;
  const bool cell = ci >= 0 && i < a.cell_lo[j + 1];
  int info = 0;
  unsigned short l[6] = {0, 0, 0, 0, 0, 0};
  T R[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  T area = T(1), eta_old = T(0), r1 = T(0), r2 = T(0);
  if (cell) {
    info = a.c.info[i];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      l[q] = a.cyc_loc[q * N + i];
      R[q] = a.c.kite[q * N + i];
    }
    area = a.c.area[i];
    eta_old = a.c.eta[b * a.c.eta_stride + i];
    const T *rg = a.c.ring + b * a.c.ring_stride;
    r1 = rg[int((n + 2) % 3) * N + i];
    r2 = rg[int((n + 1) % 3) * N + i];
  }
  double local = 0.0;
  T *ring = a.f.ring + b * a.f.ring_stride + int(n % 3) * E;
  T *vnxt = a.f.v_nxt + b * a.f.v_stride;
  for (int q = threadIdx.x; q < n_loc; q += NT) {
    const bool own = q < n_own;
    const int e = own ? lo + q : a.for_idx[flo + q - n_own];
    const FaceOut<T> r = face_update<T, false, true>(a.f, e, b, energy_quad, table, s);
    if (own) {
      local += r.energy;
      ring[e] = r.f0;
      vnxt[e] = r.vn;
    }
    lv[q] = a.c.flen[e] * r.vn;
    gid[q] = e;
  }
  __syncthreads();
  if (cell) {
    int f[6];
    T F[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      f[q] = gid[l[q]];
      const T x = lv[l[q]];
      F[q] = ((info >> (3 + q)) & 1) ? -x : x;
    }
    const int k = info & 7, db = info >> 3;
    continuity_loaded(a.c, i, b, n, divergence(F, k, area), eta_old, r1, r2);
    cell_outputs(a.c, b, k, db, f, F, R);
  }
  return local;
}

// MULTI: each workgroup runs a contiguous run of patches (fewer workgroups
// than patches); else one patch per workgroup.
template <typename T, bool PENDING, int P, int LPC, bool MULTI>
__global__ __launch_bounds__(LPC * P) void k_step(StepArgs<T> a, int energy_quad, const double *table, int s) {
  constexpr int NT = LPC * P;  // LPC lanes per cell; the faces of P cells are about 3.4 P
  __shared__ T lv[6 * P];
  __shared__ int gid[6 * P];
  __shared__ double sh[NT / 64];
  const int b = blockIdx.y;
  const long long n = (long long)table[0] + s;
  const int t = xcd_tile(blockIdx.x, a.n_blocks);
  double local = 0.0;
  if (MULTI) {
    const int iters = (a.n_patch + a.n_blocks - 1) / a.n_blocks;
    for (int it = 0; it < iters; ++it) {
      const int j = t * iters + it;
      if (j >= a.n_patch) break;    // uniform in the workgroup
      if (it > 0) __syncthreads();  // the previous patch's cell phase has read lv / gid
      local += step_patch<T, P, LPC>(a, j, b, n, energy_quad, table, s, lv, gid);
    }
  } else {
    local = step_patch<T, P, LPC>(a, t, b, n, energy_quad, table, s, lv, gid);
  }
  const double part = block_sum<NT / 64>(local, sh);
  if (threadIdx.x == 0) a.f.partials[((n & 1) * gridDim.y + b) * a.n_blocks + t] = part;
  if (!PENDING || blockIdx.x != 0) return;
  // dissipation of step n - 1: partials of the previous launch
  __syncthreads();
  const double tot = sum_partials<NT / 64>(
      a.f.partials + (((n - 1) & 1) * gridDim.y + b) * size_t(a.n_blocks), a.n_blocks, sh);
  if (threadIdx.x == 0) {
    const double d = tot * a.c.dconst[b];
    a.c.series[size_t(b) * a.c.cap + size_t((n - 1) % a.c.cap)] = d;
    if (!isfinite(d)) a.c.nonfinite[b] = 1;
  }
}

// Operand records of the interleaved step.  Each face and each cell keeps
// its operands in one record of kRecSlots values of the state type, so
// that the 16 lanes of a group (one DPP row: lanes 16r .. 16r + 15 serve
// the 16 instances of one face or cell) load the record with one vector
// load, lane k slot k, and take each operand from its slot with a row
// broadcast (v_mov_b32_dpp row_newbcast), instead of all 16 lanes loading
// every operand.  Integer operands are kept as bit patterns.
//   face e: 0, 1 gradient cells c0, c1; 2, 3 their coefficients; 4 Coriolis;
//           5 area; 6 d_e; 7 l_e; 8 + k forcing field k at c0, 12 + k at c1
//           (k < kRecFields)
//   cell i: 0 cycle info; 1-3 the six patch-local cycle faces, two 16-bit
//           indices per slot; 4-9 kite fractions; 10 area
constexpr int kRecSlots = 16;
constexpr int kRecFields = 4;  // forcing fields a face record holds

template <typename T>
__device__ __forceinline__ T int_slot(int x) {
  if constexpr (sizeof(T) == 8)
    return __longlong_as_double((long long)x);
  else
    return __int_as_float(x);
}

template <typename T>
__device__ __forceinline__ int slot_int(T x) {
  if constexpr (sizeof(T) == 8)
    return int(__double_as_longlong(x));
  else
    return __float_as_int(x);
}

// slot S of the lane's DPP row (every lane of the row active)
template <int S>
__device__ __forceinline__ float row_slot(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x150 + S, 0xF, 0xF, false));
}
template <int S>
__device__ __forceinline__ double row_slot(double x) {
  const long long v = __double_as_longlong(x);
  const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(v), 0x150 + S, 0xF, 0xF, false));
  const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(v >> 32), 0x150 + S, 0xF, 0xF, false));
  return __longlong_as_double((long long)((static_cast<unsigned long long>(hi) << 32) | lo));
}

// u0 += c_K S_K(c0), u1 += c_K S_K(c1) from a face record, for field K of
// the tide's nf (nf is uniform, so the row is whole for the broadcasts)
template <int K, typename T>
__device__ __forceinline__ void rec_forcing(T mine, const T *ck, int nf, T &u0, T &u1) {
  if (K < nf) {
    u0 += ck[K] * row_slot<8 + K>(mine);
    u1 += ck[K] * row_slot<12 + K>(mine);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pack_face_rec(int E, int N, int nf, const int *__restrict__ g_idx,
                                                          const T *__restrict__ g_coef, const T *__restrict__ fcor,
                                                          const T *__restrict__ farea, const T *__restrict__ fdist,
                                                          const T *__restrict__ flen, const T *__restrict__ cellS,
                                                          T *__restrict__ rec) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= E) return;
  const int c0 = g_idx[e], c1 = g_idx[E + e];
  T *r = rec + size_t(e) * kRecSlots;
  r[0] = int_slot<T>(c0);
  r[1] = int_slot<T>(c1);
  r[2] = g_coef[e];
  r[3] = g_coef[E + e];
  r[4] = fcor[e];
  r[5] = farea[e];
  r[6] = fdist[e];
  r[7] = flen[e];
  for (int k = 0; k < kRecFields; ++k) {
    r[8 + k] = k < nf ? cellS[size_t(k) * N + c0] : T(0);
    r[12 + k] = k < nf ? cellS[size_t(k) * N + c1] : T(0);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pack_cell_rec(int N, const int *__restrict__ info,
                                                          const unsigned short *__restrict__ cyc_loc,
                                                          const T *__restrict__ kite, const T *__restrict__ area,
                                                          T *__restrict__ rec) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  T *r = rec + size_t(i) * kRecSlots;
  r[0] = int_slot<T>(info[i]);
  for (int q = 0; q < 3; ++q)
    r[1 + q] = int_slot<T>(int(cyc_loc[size_t(2 * q) * N + i]) | (int(cyc_loc[size_t(2 * q + 1) * N + i]) << 16));
  for (int q = 0; q < 6; ++q) r[4 + q] = kite[size_t(q) * N + i];
  r[10] = area[i];
  for (int q = 11; q < kRecSlots; ++q) r[q] = T(0);
}

// One-kernel step of a batch held interleaved: every per-instance array is
// held in groups of IB instances, each group element-major with its
// instances innermost (sidx<true>).  Workgroup (patch j, group of IB
// instances); lane l serves instance l % IB of the group.  The patch's faces
// go NT / IB at a time, so each face's operators are loaded once for IB
// instances and its state is IB consecutive values; then its (cell,
// instance) pairs.  The arithmetic of a face and a cell is step_patch's, so
// the state is the same bit for bit; an instance's dissipation partial adds
// its faces in another grouping.  Padding instances (b >= B) compute with
// the last instance's parameters and store nothing.
// LDS (dynamic, il_lds_bytes): the dissipation partials, then per local
// face l_e v_e^{n+1} of the group and the face's id, sized for the largest
// patch (a.max_loc faces) rather than 6 P, so that more workgroups fit a CU.
// REC: operands from the records (a.face_rec, a.cell_rec; at most
// kRecFields forcing fields), else from the operand arrays.  The per-face
// arithmetic is face_update's, written out in the same order.
// GF: the groups in bands of a.il_band; a one-dimensional grid of tiles
// (band, patch, group of the band), each XCD a contiguous range of them, so
// that the band's groups of a patch (which read its operands) and
// neighbouring patches (which read each other's boundary state) run on one
// XCD close in time, while the XCD streams through the state of only a
// band's groups; a band of all groups is groups-fastest order.  Else a grid
// (n_blocks, groups) with each group's patches spread as contiguous ranges
// over the XCDs.
template <typename T, bool PENDING, int P, int IB, int FU, int CU, bool GF = false, bool REC = false>
__global__ __launch_bounds__(256) void k_step_il(StepArgs<T> a, int energy_quad, const double *table, int s) {
  static_assert(!REC || IB == 16, "a record is shared through one 16-lane DPP row");
  constexpr int NT = 256, SL = NT / IB;
  extern __shared__ __align__(16) unsigned char il_lds[];
  double *red = reinterpret_cast<double *>(il_lds);  // [SL * IB]
  double *sh = red + SL * IB;                         // [NT / 64]
  T *lv = reinterpret_cast<T *>(sh + NT / 64);        // [max_loc * IB]
  int *gid = reinterpret_cast<int *>(lv + size_t(a.max_loc) * IB);  // [max_loc]
  const int B = a.f.B, E = a.f.E, N = a.f.N;
  const int ib = threadIdx.x % IB, sl = threadIdx.x / IB;
  const int n_grp = (B + IB - 1) / IB;
  int j, grp;
  if (GF) {
    const int GB = a.il_band, per_band = a.n_blocks * GB;
    const int t = xcd_tile(blockIdx.x, per_band * ((n_grp + GB - 1) / GB));
    const int band = t / per_band, r = t - band * per_band;
    j = r / GB;
    grp = band * GB + (r - j * GB);
    if (grp >= n_grp) return;  // whole workgroup: a group the last band does not have
  } else {
    j = xcd_tile(blockIdx.x, a.n_blocks);
    grp = blockIdx.y;
  }
  const int b = grp * IB + ib;
  const bool live = b < B;
  const long long n = (long long)table[0] + s;
  const int lo = a.own_lo[j], n_own = a.own_lo[j + 1] - lo;
  const int flo = a.for_ptr[j], n_loc = n_own + (a.for_ptr[j + 1] - flo);
  double local = 0.0;
  // FU faces per lane are updated before any of them is stored, so that
  // their loads are in flight together (the stores could alias them)
  if constexpr (REC) {
    // the instance's parameters once (face_update reads them per face)
    const int bp = min(b, B - 1);
    const double *ip = a.f.inst + bp * 8;
    const double *coef = table + 8 + size_t(s) * B * 8;
    const T g = T(ip[0]), alpha = T(ip[1]), dt = T(ip[2]), qd = T(ip[4]);
    const bool ab3 = (n >= 2) || (ip[3] != 0.0);
    const int nf = a.f.nfields;
    T ck[kRecFields];
#pragma unroll
    for (int k = 0; k < kRecFields; ++k) ck[k] = k < nf ? T(coef[bp * 8 + k]) : T(0);
    const int s0 = int(n % 3), s1 = int((n + 2) % 3), s2 = int((n + 1) % 3);
    // the face ids and records of a lane's next FU faces are loaded one
    // iteration ahead, so that their latency overlaps this iteration's work
    // (past the end: the last face again, not stored)
    int e_n[FU];
    T rec_n[FU];
    if (n_loc > 0) {
#pragma unroll
      for (int u = 0; u < FU; ++u) {
        const int q = min(sl + u * SL, n_loc - 1);
        e_n[u] = q < n_own ? lo + q : a.for_idx[flo + q - n_own];
        rec_n[u] = a.face_rec[size_t(e_n[u]) * kRecSlots + ib];
      }
    }
    for (int q0 = sl; q0 < n_loc; q0 += SL * FU) {
      T vn[FU], f0o[FU], fl[FU];
      double en_u[FU];
      int e[FU];
      T rec[FU];
#pragma unroll
      for (int u = 0; u < FU; ++u) {
        e[u] = e_n[u];
        rec[u] = rec_n[u];
      }
      if (q0 + SL * FU < n_loc) {  // the same for the whole row
#pragma unroll
        for (int u = 0; u < FU; ++u) {
          const int q = min(q0 + SL * FU + u * SL, n_loc - 1);
          e_n[u] = q < n_own ? lo + q : a.for_idx[flo + q - n_own];
          rec_n[u] = a.face_rec[size_t(e_n[u]) * kRecSlots + ib];
        }
      }
#pragma unroll
      for (int u = 0; u < FU; ++u) {
        const T mine = rec[u];
        const int c0 = min(unsigned(slot_int(row_slot<0>(mine))), unsigned(N - 1));
        const int c1 = min(unsigned(slot_int(row_slot<1>(mine))), unsigned(N - 1));
        const T gc0 = row_slot<2>(mine), gc1 = row_slot<3>(mine), fcor = row_slot<4>(mine);
        const T farea = row_slot<5>(mine), fdist = row_slot<6>(mine);
        fl[u] = row_slot<7>(mine);
        // forcing sums at the two cells, broadcasting only the fields the tide has
        T u0 = T(0), u1 = T(0);
        rec_forcing<0>(mine, ck, nf, u0, u1);
        rec_forcing<1>(mine, ck, nf, u0, u1);
        rec_forcing<2>(mine, ck, nf, u0, u1);
        rec_forcing<3>(mine, ck, nf, u0, u1);
        // face_update<T, false, true, true>, operation for operation
        const T vt = (a.f.ts[sidx<true>(b, a.f.ts_stride, e[u])] -
                      a.f.ts[sidx<true>(b, a.f.ts_stride, E + e[u])]) / fdist;
        const T ve = a.f.v_cur[sidx<true>(b, a.f.v_stride, e[u])];
        const T sp = ve * ve + vt * vt;
        const T en = energy_quad ? sqrt(sp) * sp : sp;
        en_u[u] = double(farea) * double(en);
        const T ge = gc0 * a.f.eta[sidx<true>(b, a.f.eta_stride, c0)] + gc1 * a.f.eta[sidx<true>(b, a.f.eta_stride, c1)];
        T f0 = -g * ge + fcor * vt;
        if (a.f.quad_drag) f0 -= qd * sqrt(sp) * ve;
        const T fu = gc0 * u0 + gc1 * u1;
        const T drag = -alpha * ve + fu;
        T upd;
        if (ab3)
          upd = (T(23. / 12.) * f0 + T(-16. / 12.) * a.f.ring[sidx<true>(b, a.f.ring_stride, (long long)s1 * E + e[u])] +
                 T(5. / 12.) * a.f.ring[sidx<true>(b, a.f.ring_stride, (long long)s2 * E + e[u])]) *
                dt;
        else
          upd = f0 * dt;
        f0o[u] = f0;
        vn[u] = (ve + upd) + dt * drag;
      }
#pragma unroll
      for (int u = 0; u < FU; ++u) {
        const int q = q0 + u * SL;
        if (q >= n_loc) break;
        if (q < n_own && live) {
          local += en_u[u];
          a.f.ring[sidx<true>(b, a.f.ring_stride, (long long)s0 * E + e[u])] = f0o[u];
          a.f.v_nxt[sidx<true>(b, a.f.v_stride, e[u])] = vn[u];
        }
        lv[q * IB + ib] = fl[u] * vn[u];
        if (ib == 0) gid[q] = e[u];
      }
    }
  } else {
    for (int q0 = sl; q0 < n_loc; q0 += SL * FU) {
      FaceOut<T> r[FU];
      int e[FU];
  #pragma unroll
      for (int u = 0; u < FU; ++u) {
        const int q = min(q0 + u * SL, n_loc - 1);  // past the end: the last face again, not stored
        e[u] = q < n_own ? lo + q : a.for_idx[flo + q - n_own];
        r[u] = face_update<T, false, true, true>(a.f, e[u], b, energy_quad, table, s);
      }
  #pragma unroll
      for (int u = 0; u < FU; ++u) {
        const int q = q0 + u * SL;
        if (q >= n_loc) break;
        if (q < n_own && live) {
          local += r[u].energy;
          a.f.ring[sidx<true>(b, a.f.ring_stride, (long long)int(n % 3) * E + e[u])] = r[u].f0;
          a.f.v_nxt[sidx<true>(b, a.f.v_stride, e[u])] = r[u].vn;
        }
        lv[q * IB + ib] = a.c.flen[e[u]] * r[u].vn;
        if (ib == 0) gid[q] = e[u];
      }
    }
  }
  __syncthreads();
  const int c0 = a.cell_lo[j], nc = a.cell_lo[j + 1] - c0;
  // CU (cell, instance) pairs per lane loaded together, then updated
  if constexpr (REC) {
    // every lane of a row takes part in the broadcasts; padding instances
    // of the last group then skip their state
    const double *ipc = a.c.inst + min(b, B - 1) * 8;
    const T h_c = T(ipc[5]), dt_c = T(ipc[2]);
    const bool ab3_c = (n >= 2) || (ipc[3] != 0.0);
    for (int p0 = threadIdx.x; p0 < nc * IB; p0 += NT * CU) {  // p % IB == ib
      int info[CU], lq[CU][6];
      T R[CU][6], eta_old[CU], r1[CU], r2[CU], area[CU];
#pragma unroll
      for (int u = 0; u < CU; ++u) {
        const int i = c0 + min(p0 + u * NT, nc * IB - 1) / IB;
        const T mine = a.cell_rec[size_t(i) * kRecSlots + ib];
        info[u] = slot_int(row_slot<0>(mine));
        const int w0 = slot_int(row_slot<1>(mine)), w1 = slot_int(row_slot<2>(mine)), w2 = slot_int(row_slot<3>(mine));
        const int wl[3] = {w0, w1, w2};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          lq[u][2 * q] = min(wl[q] & 0xffff, a.max_loc - 1);
          lq[u][2 * q + 1] = min((wl[q] >> 16) & 0xffff, a.max_loc - 1);
        }
        R[u][0] = row_slot<4>(mine);
        R[u][1] = row_slot<5>(mine);
        R[u][2] = row_slot<6>(mine);
        R[u][3] = row_slot<7>(mine);
        R[u][4] = row_slot<8>(mine);
        R[u][5] = row_slot<9>(mine);
        area[u] = row_slot<10>(mine);
        eta_old[u] = r1[u] = r2[u] = T(0);
        if (live) {
          eta_old[u] = a.c.eta[sidx<true>(b, a.c.eta_stride, i)];
          r1[u] = a.c.ring[sidx<true>(b, a.c.ring_stride, (long long)int((n + 2) % 3) * N + i)];
          r2[u] = a.c.ring[sidx<true>(b, a.c.ring_stride, (long long)int((n + 1) % 3) * N + i)];
        }
      }
      if (live) {
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int p = p0 + u * NT;
          if (p >= nc * IB) break;
          const int i = c0 + p / IB;
          int f[6];
          T F[6];
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            const int l = lq[u][q];
            f[q] = gid[l];
            const T x = lv[l * IB + ib];
            F[q] = ((info[u] >> (3 + q)) & 1) ? -x : x;
          }
          const int k = info[u] & 7, db = info[u] >> 3;
          continuity_with<T, true>(a.c, i, b, n, divergence(F, k, area[u]), eta_old[u], r1[u], r2[u], h_c, dt_c,
                                   ab3_c);
          cell_outputs<T, true>(a.c, b, k, db, f, F, R[u]);
        }
      }
    }
  } else {
    if (live)
      for (int p0 = threadIdx.x; p0 < nc * IB; p0 += NT * CU) {  // p % IB == ib
        int info[CU], lq[CU][6];
        T R[CU][6], eta_old[CU], r1[CU], r2[CU], area[CU];
  #pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int i = c0 + min(p0 + u * NT, nc * IB - 1) / IB;
          info[u] = a.c.info[i];
  #pragma unroll
          for (int q = 0; q < 6; ++q) {
            lq[u][q] = a.cyc_loc[q * N + i];
            R[u][q] = a.c.kite[q * N + i];
          }
          area[u] = a.c.area[i];
          eta_old[u] = a.c.eta[sidx<true>(b, a.c.eta_stride, i)];
          r1[u] = a.c.ring[sidx<true>(b, a.c.ring_stride, (long long)int((n + 2) % 3) * N + i)];
          r2[u] = a.c.ring[sidx<true>(b, a.c.ring_stride, (long long)int((n + 1) % 3) * N + i)];
        }
  #pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int p = p0 + u * NT;
          if (p >= nc * IB) break;
          const int i = c0 + p / IB;
          int f[6];
          T F[6];
  #pragma unroll
          for (int q = 0; q < 6; ++q) {
            const int l = lq[u][q];
            f[q] = gid[l];
            const T x = lv[l * IB + ib];
            F[q] = ((info[u] >> (3 + q)) & 1) ? -x : x;
          }
          const int k = info[u] & 7, db = info[u] >> 3;
          continuity_loaded<T, true>(a.c, i, b, n, divergence(F, k, area[u]), eta_old[u], r1[u], r2[u]);
          cell_outputs<T, true>(a.c, b, k, db, f, F, R[u]);
        }
      }
  }
  // the patch's dissipation partial of each instance: its face slots in order
  red[sl * IB + ib] = local;
  __syncthreads();
  if (threadIdx.x < IB) {
    double sum = 0.0;
    for (int u = 0; u < SL; ++u) sum += red[u * IB + threadIdx.x];
    const int bb = grp * IB + threadIdx.x;
    if (bb < B) a.f.partials[((n & 1) * B + bb) * a.n_blocks + j] = sum;
  }
  if (!PENDING || j != 0) return;
  // dissipation of step n - 1 of the group's instances: partials of the previous launch
  for (int u = 0; u < IB; ++u) {
    const int bb = grp * IB + u;
    if (bb >= B) break;  // uniform in the workgroup
    __syncthreads();     // sh is free
    const double tot = sum_partials<NT / 64>(a.f.partials + (((n - 1) & 1) * B + bb) * size_t(a.n_blocks),
                                             a.n_blocks, sh);
    if (threadIdx.x == 0) {
      const double d = tot * a.c.dconst[bb];
      a.c.series[size_t(bb) * a.c.cap + size_t((n - 1) % a.c.cap)] = d;
      if (!isfinite(d)) a.c.nonfinite[bb] = 1;
    }
  }
}

// p[0 .. n) = 0, grid-stride
__global__ __launch_bounds__(kBlock) void k_zero_words(unsigned int *__restrict__ p, size_t n) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) p[i] = 0u;
}

// Per-instance array of n elements between the instance-major layout (std,
// [B][n]) and the interleaved one (il, sidx<true>): TO_IL copies std into
// il, else il into std.  Grid (elements / kBlock, B): workgroup y is one
// instance, so the instance-major side is contiguous.  Padding instances
// are not touched.
template <typename T, bool TO_IL>
__global__ __launch_bounds__(kBlock) void k_interleave(T *__restrict__ stdp, T *__restrict__ ilp, long long n) {
  const int b = blockIdx.y;
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (TO_IL)
    ilp[sidx<true>(b, n, i)] = stdp[b * n + i];
  else
    stdp[b * n + i] = ilp[sidx<true>(b, n, i)];
}

// Dissipation of the last step of a chunk of k_step launches.
__global__ __launch_bounds__(kBlock) void k_series(const double *partials, int count, const double *dconst,
                                                   double *series, int *nonfinite, int cap, const double *table,
                                                   int s) {
  __shared__ double sh[4];
  const int b = blockIdx.x;
  const long long n = (long long)table[0] + s;
  const double tot = sum_partials<kBlock / 64>(partials + ((n & 1) * gridDim.x + b) * size_t(count), count, sh);
  if (threadIdx.x == 0) {
    const double d = tot * dconst[b];
    series[size_t(b) * cap + size_t(n % cap)] = d;
    if (!isfinite(d)) nonfinite[b] = 1;
  }
}

// Sum dissipation partials produced by k_face<DISS_ONLY>.
__global__ __launch_bounds__(kBlock) void k_diss_final(const double *partials, int nb, const double *dconst,
                                                       double *out) {
  __shared__ double sh[4];
  const int b = blockIdx.x;
  const double t = sum_partials<kBlock / 64>(partials + size_t(b) * nb, nb, sh);
  if (threadIdx.x == 0) out[b] = t * dconst[b];
}

// (east, north) velocity per face: n v_n + t v_t with t = (n_y, -n_x)
// (src/interpolation.cpp:49-58).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_velocity_en(int E, const int *tri_idx, const T *tri_coef,
                                                        const T *v, const double *normal, double *uv) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= E) return;
  double vt = 0.0;
  for (int j = 0; j < 10; ++j) vt += double(tri_coef[j * E + e]) * double(v[tri_idx[j * E + e]]);
  const double nx = normal[2 * e], ny = normal[2 * e + 1], vn = double(v[e]);
  uv[2 * e] = nx * vn + ny * vt;
  uv[2 * e + 1] = ny * vn + (-nx) * vt;
}

// RBF reconstruction of the Cartesian velocity at cell centres
// (interpolateVelocityCartRBF, src/interpolation.cpp:64-174): per device
// cell the reference operator's 3 x k block (k <= 6 faces, slot-major, the
// CSR's column order), out [3][N].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_velocity_xyz(int N, const int *__restrict__ idx,
                                                         const double *__restrict__ w, const T *__restrict__ v,
                                                         double *out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  double u[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int f = idx[j * N + i];
    u[j] = f >= 0 ? double(v[f]) : 0.0;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) acc += w[(c * 6 + j) * N + i] * u[j];
    out[c * N + i] = acc;
  }
}

// CSR rows of the lat-lon interpolation: one wavefront per row, the eta
// columns already mapped to device cell indices on the host.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_interp_ll(int n_rows, const int *rp, const int *col,
                                                      const double *val, const T *eta, double *out) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (row >= n_rows) return;
  double acc = 0.0;
  for (int k = rp[row] + lane; k < rp[row + 1]; k += 64) acc += val[k] * double(eta[col[k]]);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) out[row] = acc;
}

// ---------------------------------------------------------------------------
// spherical-harmonic projection (self-gravity / ice-shell coupling)
// ---------------------------------------------------------------------------
// Basis Yt[k][i] (k = harmonic, i = device cell) stored basis-major.  The
// analysis b_k = sum_i Yt[k][i] x_i runs one workgroup per (harmonic, chunk
// of kShChunks cell ranges) and writes partial sums, which k_sh_bsum adds in
// a fixed order; the synthesis y_i = sum_k c_k Yt[k][i] runs 64 cells per
// workgroup, each of its 4 waves over a quarter of the harmonics, the
// quarters added in a fixed order through LDS.  Every basis entry is read
// once per pass, coalesced along the cells.  The split-BF16 form packs
// hi = bf16(Y) in the upper and lo = bf16(Y - hi) in the lower half of a
// 32-bit word.

constexpr int kShChunks = 8;

__device__ __forceinline__ float bf16_hi(unsigned int u) { return __uint_as_float(u & 0xFFFF0000u); }
__device__ __forceinline__ float bf16_lo(unsigned int u) { return __uint_as_float(u << 16); }

// grid (K * kShChunks, B): partial b of harmonic k over cell chunk q
template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_sh_analysis(int K, int N, const double *Yt, const unsigned int *Yp,
                                                        const T *x, long long x_stride, double *b_part) {
  __shared__ double sh[4];
  const int k = blockIdx.x / kShChunks, q = blockIdx.x % kShChunks, b = blockIdx.y;
  const int chunk = ((N + kShChunks - 1) / kShChunks + kBlock - 1) / kBlock * kBlock;
  const int lo = q * chunk, hi = min(N, lo + chunk);
  const T *xb = x + b * x_stride;
  double acc = 0.0;
  if (SPLIT) {
    float facc = 0.f;
    const unsigned int *row = Yp + size_t(k) * N;
    for (int i = lo + threadIdx.x; i < hi; i += kBlock) {
      const unsigned int u = row[i];
      const float xi = float(xb[i]);
      facc = fmaf(bf16_hi(u), xi, facc);
      facc = fmaf(bf16_lo(u), xi, facc);
    }
    acc = double(facc);
  } else {
    const double *row = Yt + size_t(k) * N;
    for (int i = lo + threadIdx.x; i < hi; i += kBlock) acc += row[i] * double(xb[i]);
  }
  const double t = block_sum<kBlock / 64>(acc, sh);
  if (threadIdx.x == 0) b_part[(size_t(b) * kShChunks + q) * K + k] = t;
}

// Store c_k (k = l l + l + m) at its m-major place: for m = 0..L, l = m..L
// the pair (c_{l,|m|}, c_{l,-|m|}) at 2 (m (L + 1) - m (m - 1) / 2 + l - m);
// the unused sine slot of m = 0 is zero.  chain (nullable): the degree
// table of the chain synthesis instead, 4 doubles per (l, m) at
// 4 (m (L + 1) - m (m - 1) / 2 + l - m): (beta_l, S_l c_{l,m}, S_l c_{l,-m}, 0),
// S_l the chain scale; beta and the padding are set with the basis.
__device__ __forceinline__ void put_mmajor(int K, int b, int k, double ck, const double *chain, double *cmaj) {
  int l = 0;
  while ((l + 1) * (l + 1) <= k) ++l;
  const int mm = k - l * l - l, m = mm < 0 ? -mm : mm;
  int L = 0;
  while ((L + 1) * (L + 1) < K) ++L;
  const int at = m * (L + 1) - m * (m - 1) / 2 + l - m;
  if (chain) {
    double *cb = cmaj + size_t(b) * 2 * (L + 1) * (L + 2);
    if (m == 0) cb[4 * at + 2] = 0.0;
    cb[4 * at + 1 + (mm < 0 ? 1 : 0)] = chain[2 * at + 1] * ck;
  } else {
    double *cb = cmaj + size_t(b) * (L + 1) * (L + 2);
    if (m == 0) cb[2 * l + 1] = 0.0;
    cb[2 * at + (mm < 0 ? 1 : 0)] = ck;
  }
}

// coefficient k of instance b -> its place in the fragment-ordered copy of
// the batched products, Cs[b][k / 32][(k / 8) % 4] = 8 hi, then 8 lo bf16
// of float(c)
__device__ __forceinline__ void put_split(int KB, int b, int k, double ck, unsigned short *cs) {
  const float f = float(ck);
  const __bf16 h = (__bf16)f;
  const __bf16 l = (__bf16)(f - (float)h);
  unsigned short *p = cs + ((size_t(b) * KB + (k >> 5)) * 4 + ((k >> 3) & 3)) * 16 + (k & 7);
  p[0] = __builtin_bit_cast(unsigned short, h);
  p[8] = __builtin_bit_cast(unsigned short, l);
}

// b_j = sum_q b_part[q][j] over the n_part analysis partials of instance b
// ([B][n_part][K]): a workgroup per 64 harmonics; wave w adds the partials
// q = w, w + 4, ... (their loads in flight together) and the four wave sums
// are added in order through LDS.  Written after the partials of every
// instance.
__global__ __launch_bounds__(kBlock) void k_sh_bsum(int K, int n_part, double *b_part) {
  constexpr int W = kBlock / 64;
  __shared__ double part[W][64];
  const int b = blockIdx.y, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  const double *p = b_part + size_t(b) * n_part * K + min(j, K - 1);  // lanes past K load a valid entry
  double t = 0.0;
  for (int q0 = w; q0 < n_part; q0 += W * 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = q0 + W * u < n_part ? p[size_t(q0 + W * u) * K] : 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (q0 + W * u < n_part) t += v[u];
  }
  part[w][lane] = t;
  __syncthreads();
  if (w != 0 || j >= K) return;
  double sum = 0.0;
  for (int q = 0; q < W; ++q) sum += part[q][lane];
  b_part[(size_t(gridDim.y) * n_part + b) * K + j] = sum;
}

// c_k = D_k sum_j Minv[k][j] b_j: one workgroup per (k, instance); thread t
// loads the row and b entries j = t, t + 256, ... (all of a row in flight
// at once), each wave sums its lanes and thread 0 adds the wave sums in
// order.  cmaj (nullable) also receives c in m-major order (put_mmajor),
// csplit (nullable) its split copy (put_split).
__global__ __launch_bounds__(kBlock) void k_sh_coef(int K, const double *__restrict__ Minv, const double *D,
                                                    const double *__restrict__ bsum, double *c, const double *chain,
                                                    double *cmaj, unsigned short *csplit, int KB) {
  __shared__ double part[kBlock / 64];
  const int b = blockIdx.y, k = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double *row = Minv + size_t(k) * K, *bv = bsum + size_t(b) * K;
  double acc = 0.0;
  for (int j0 = threadIdx.x; j0 < K; j0 += 8 * kBlock) {
    double r[8], x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool in = j0 + kBlock * u < K;
      r[u] = in ? row[j0 + kBlock * u] : 0.0;
      x[u] = in ? bv[j0 + kBlock * u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j0 + kBlock * u < K) acc += r[u] * x[u];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) part[w] = acc;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double sum = 0.0;
  for (int q = 0; q < kBlock / 64; ++q) sum += part[q];
  const double ck = (D ? D[size_t(b) * K + k] : 1.0) * sum;
  c[size_t(b) * K + k] = ck;
  if (cmaj) put_mmajor(K, b, k, ck, chain, cmaj);
  if (csplit) put_split(KB, b, k, ck, csplit);
}

// c (index k = l l + l + m) -> its m-major copy, one lane per k
__global__ __launch_bounds__(kBlock) void k_sh_to_mmajor(int K, const double *c, const double *chain,
                                                         double *cmaj) {
  const int b = blockIdx.y;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < K) put_mmajor(K, b, k, c[size_t(b) * K + k], chain, cmaj);
}

// y_i = base_i + sum_k c_k Yt[k][i]: 64 cells per workgroup, wave w sums the
// harmonics [w K/4, (w+1) K/4)
template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_sh_synth(int K, int N, const double *Yt, const unsigned int *Yp,
                                                     const double *c, const T *base, long long stride, T *out) {
  extern __shared__ double cs[];  // [K] coefficients, then [4][64] wave partials
  double *part = cs + K;
  const int b = blockIdx.y;
  for (int k = threadIdx.x; k < K; k += kBlock) cs[k] = c[size_t(b) * K + k];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int per = (K + 3) / 4;
  const int k0 = wave * per, k1 = min(K, k0 + per);
  double acc = 0.0;
  if (i < N) {
    if (SPLIT) {
      float facc = 0.f;
      for (int k = k0; k < k1; ++k) {
        const unsigned int u = Yp[size_t(k) * N + i];
        const float ck = float(cs[k]);
        facc = fmaf(bf16_hi(u), ck, facc);
        facc = fmaf(bf16_lo(u), ck, facc);
      }
      acc = double(facc);
    } else {
      for (int k = k0; k < k1; ++k) acc += cs[k] * Yt[size_t(k) * N + i];
    }
  }
  part[wave * 64 + lane] = acc;
  __syncthreads();
  if (wave != 0 || i >= N) return;
  const double sum = (part[lane] + part[64 + lane]) + (part[128 + lane] + part[192 + lane]);
  const T add = base ? base[b * stride + i] : T(0);
  out[b * stride + i] = T(double(add) + sum);
}

// ---------------------------------------------------------------------------
// the same projection with the basis evaluated on the fly
// ---------------------------------------------------------------------------
// Y_lm(i) = P_lm(sin lat_i) cos(m lon_i) (m >= 0) or P_l|m| sin(|m| lon_i),
// P_lm by the recurrence whose coefficients the host tabulates
// (sh_recurrence_tables: a, b per (l, m), sect and f1 per m; the host basis
// used for the Gram matrix comes from the same table), cos / sin(m lon) by
// rotation from cos / sin(lon).  No basis is read from memory: per pass the
// kernels read 4 values per cell and the coefficient table.
//   analysis, runtime degree: workgroup (cell chunks, m pair) -- m and
//     L - m, so every wave has L + 2 degrees -- CPT cells per lane, per
//     (l, +-m) a wave sum, partials per workgroup added in a fixed order by
//     k_sh_bsum; degrees 8, 16 and 32 have the pair analysis below;
//   synthesis: one lane per cell over an m-range.

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Analysis: workgroup = kShWaves waves on kShWaves consecutive cell chunks
// (64 CPT cells each) and one m pair (m, L - m), so every wave has L + 2
// degrees.  A wave runs the degrees of each m in blocks of 8: per lane the
// 16 (l, cos / sin) sums over its CPT cells, then a butterfly reduce-scatter
// over the 64 lanes (xor 1, 2, 4, 8 halving the slots, then xor 16, 32)
// leaves each of the 16 sums in one lane -- 17 shuffles for 16 sums.  The
// workgroup adds its waves' sums through LDS in a fixed order and writes one
// partial per (l, +-m).  Products are fused multiply-adds; the recurrence
// table is read m-major (sect, f1, then per m the (a, b) pairs of
// l = m+2..L) with wave-uniform indices.
constexpr int kShWaves = 8;

// 16 per-lane values -> the sum over the wave of value q in lane
// (q & 1) << 3 | (q & 2) << 1 | (q & 4) >> 1 | (q & 8) >> 3 (and its +16, +32, +48)
template <typename T>
__device__ __forceinline__ T reduce_scatter16(T *v, int lane) {
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const int d = 1 << st, half = 8 >> st;
    const bool hi = (lane & d) != 0;
#pragma unroll
    for (int j = 0; j < half; ++j) {
      const T keep = hi ? v[half + j] : v[j];
      const T send = hi ? v[j] : v[half + j];
      v[j] = keep + __shfl_xor(send, d, 64);
    }
  }
  T r = v[0];
  r += __shfl_xor(r, 16, 64);
  r += __shfl_xor(r, 32, 64);
  return r;
}

template <typename T, int LT, int CPT>
__global__ __launch_bounds__(64 * kShWaves) void k_sh_analysis_rec(int Lrt, int N, const double *__restrict__ tabm,
                                                                   const double *__restrict__ trig,
                                                                   const T *__restrict__ x, long long x_stride,
                                                                   int n_chunks, double *b_part) {
  extern __shared__ double red[];  // [kShWaves][2 (L + 2)]
  const int L = LT > 0 ? LT : Lrt;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n_pairs = L / 2 + 1, n_slot = 2 * (L + 2);
  const int cblk = blockIdx.x / n_pairs, pair = blockIdx.x % n_pairs;
  const int n_cblk = (n_chunks + kShWaves - 1) / kShWaves;
  const int b = blockIdx.y;
  const int chunk = cblk * kShWaves + w;
  const int K = (L + 1) * (L + 1);
  const double *tsect = tabm, *tf1 = tabm + (L + 1);
  const T *xb = x + b * x_stride;
  double *rw = red + w * n_slot;
  const int m0 = pair, m1 = L - pair;
  // lane of sum q after reduce_scatter16 is rev4(q); this lane holds q = rev4(lane & 15)
  const int q_lane = ((lane & 1) << 3) | ((lane & 2) << 1) | ((lane & 4) >> 1) | ((lane & 8) >> 3);
  if (chunk < n_chunks) {
    for (int h = 0; h < 2; ++h) {
      const int m = h == 0 ? m0 : m1;
      if (h == 1 && m1 == m0) break;
      double *slot = rw + (h == 0 ? 0 : 2 * (L - m0 + 1));
      int oab = 2 * (L + 1);
      for (int q = 0; q < m; ++q) oab += 2 * max(0, L - q - 1);
      const double *ab = tabm + oab;  // (a, b) of l = m + 2 + j at 2 j
      T mu[CPT], xc[CPT], xs[CPT], p1[CPT], p2[CPT];
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const int i = (chunk * CPT + j) * 64 + lane;
        const bool ok = i < N;
        mu[j] = ok ? T(trig[i]) : T(0);
        const T s = ok ? T(trig[N + i]) : T(0), cl = ok ? T(trig[2 * N + i]) : T(1),
                sl = ok ? T(trig[3 * N + i]) : T(0);
        const T xv = ok ? xb[i] : T(0);
        T cm = T(1), sm = T(0), pmm = T(1);
        for (int q = 1; q <= m; ++q) {
          const T c2 = fma(cm, cl, -(sm * sl));
          sm = fma(sm, cl, cm * sl);
          cm = c2;
          pmm = pmm * (T(tsect[q]) * s);
        }
        xc[j] = xv * cm;
        xs[j] = xv * sm;
        p1[j] = pmm;  // P_{l-1} for l = m + 1 after the first degree
        p2[j] = T(0);
      }
      const int nl = L - m + 1;  // degrees m .. L, index t = l - m
      const T f = T(tf1[m]);
      for (int t0 = 0; t0 < nl; t0 += 8) {
        T v[16];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int t = t0 + u;  // wave-uniform
          T ac = T(0), as = T(0);
          if (t < nl) {
            if (t == 0) {
#pragma unroll
              for (int j = 0; j < CPT; ++j) {
                ac = fma(p1[j], xc[j], ac);
                as = fma(p1[j], xs[j], as);
              }
            } else if (t == 1) {
#pragma unroll
              for (int j = 0; j < CPT; ++j) {
                const T p = f * mu[j] * p1[j];
                p2[j] = p1[j];
                p1[j] = p;
                ac = fma(p, xc[j], ac);
                as = fma(p, xs[j], as);
              }
            } else {
              const T a = T(ab[2 * (t - 2)]), bb = T(ab[2 * (t - 2) + 1]);
#pragma unroll
              for (int j = 0; j < CPT; ++j) {
                const T p = fma(a * mu[j], p1[j], -(bb * p2[j]));
                p2[j] = p1[j];
                p1[j] = p;
                ac = fma(p, xc[j], ac);
                as = fma(p, xs[j], as);
              }
            }
          }
          v[2 * u] = ac;
          v[2 * u + 1] = as;
        }
        const T r = reduce_scatter16(v, lane);
        // lane with (lane & 15) = rev4(q) writes sum q = 2 u + s: degree t0 + u
        if (lane < 16 && t0 + (q_lane >> 1) < nl) slot[2 * t0 + q_lane] = double(r);
      }
    }
  }
  __syncthreads();
  // fixed-order sum over the workgroup's waves, one lane per slot
  const int t = threadIdx.x;
  if (t >= n_slot) return;
  const int base1 = 2 * (L - m0 + 1);
  int m, l;
  if (t < base1) {
    m = m0;
    l = m0 + t / 2;
  } else {
    if (m1 == m0) return;
    m = m1;
    l = m1 + (t - base1) / 2;
    if (l > L) return;
  }
  const bool sn = (t & 1) != 0;
  if (sn && m == 0) return;
  const int n_w = min(kShWaves, n_chunks - cblk * kShWaves);
  double sum = 0.0;
  for (int q = 0; q < n_w; ++q) sum += red[q * n_slot + t];
  b_part[(size_t(b) * n_cblk + cblk) * K + l * l + l + (sn ? -m : m)] = sum;
}

// Analysis by m pairs, for a degree L fixed at compile time.  A workgroup of
// kPairWaves waves owns one pair (m0, m1) = (p, L - p) and a block of
// 64 kPairWaves R cells; each lane runs the two Legendre chains of its pair
// for its R cells and keeps all 2 (L + 2) sums (cos, sin per degree) in
// registers, so the cross-lane reduction happens once per block instead of
// once per degree.  The chains use the rescaled recurrence
//     q_m = P_mm,  q_l = mu q_{l-1} + beta_l q_{l-2},  P_lm = S_l q_l,
// (beta_l = -b_l / (a_l a_{l-1}), S_l = a_{m+1} ... a_l, a_{m+1} = f1_m), two
// operations per degree; S is applied to the block sums.  cos(m lon),
// sin(m lon) and cos^m(lat) come from binary powers of (cos lon, sin lon)
// and cos lat, and P_mm = C_m cos^m(lat) with C_m = sect_1 ... sect_m.
// chain (m-major): for m = 0..L and k = 0..L-m the pair (beta, S) of degree
// l = m + k at 2 (m (L + 1) - m (m - 1) / 2 + k) (k = 0: beta = 1, S = 1;
// k = 1: beta = 0), then C_0..C_L.
// Block sums go to b_part ([B][n_cblk][K]), added by k_sh_bsum.
constexpr int kPairWaves = 4;

template <int LT>
struct PairShape {
  static constexpr int NT = LT + 2;            // degrees of a pair
  static constexpr int NV = 2 * NT;            // sums per lane
  static constexpr int CH = (NV + 3) / 4;      // sums per reduction pass
  static constexpr int NPASS = (NV + CH - 1) / CH;
};

template <typename T>
__device__ __forceinline__ void cmul(T &c, T &s, T zc, T zs) {
  const T c2 = fma(c, zc, -(s * zs));
  s = fma(s, zc, c * zs);
  c = c2;
}

// z <- z w (complex), r <- r u
template <typename T>
__device__ __forceinline__ void pow_step(T &c, T &s, T &r, T zc, T zs, T zr, bool first) {
  if (first) {
    c = zc;
    s = zs;
    r = zr;
  } else {
    cmul(c, s, zc, zs);
    r *= zr;
  }
}

// (cos m lon, sin m lon, cos^m lat) for M0 and M1 (M1 >= M0) by binary powers
template <typename T, int M0, int M1>
__device__ __forceinline__ void pair_powers(T zc, T zs, T zr, T &ac, T &as, T &ar, T &bc, T &bs, T &br) {
  ac = T(1), as = T(0), ar = T(1), bc = T(1), bs = T(0), br = T(1);
  bool fa = true, fb = true;
#pragma unroll
  for (int k = 0; (1 << k) <= M1; ++k) {
    if ((M0 >> k) & 1) {
      pow_step(ac, as, ar, zc, zs, zr, fa);
      fa = false;
    }
    if ((M1 >> k) & 1) {
      pow_step(bc, bs, br, zc, zs, zr, fb);
      fb = false;
    }
    if ((2 << k) <= M1) {
      const T c2 = fma(zc, zc, -(zs * zs));
      zs = T(2) * (zc * zs);
      zc = c2;
      zr *= zr;
    }
  }
}

// One chain of NK degrees starting at table entry OFF for two cells (a, b)
// at once (two independent recurrences), sums into acc[2 T0 ..]
template <typename T, int T0, int NK, int OFF, int NV>
__device__ __forceinline__ void run_chain2(const double *__restrict__ chain, T mua, T mub, T pa, T pb, T xca, T xsa,
                                          T xcb, T xsb, T (&acc)[NV]) {
  T qa1 = T(0), qa2 = pa, qb1 = T(0), qb2 = pb;
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    T qa, qb;
    if (k == 0) {
      qa = pa;
      qb = pb;
    } else if (k == 1) {
      qa = mua * qa1;
      qb = mub * qb1;
    } else {
      const T be = T(chain[2 * (OFF + k)]);
      qa = fma(mua, qa1, be * qa2);
      qb = fma(mub, qb1, be * qb2);
    }
    qa2 = qa1;
    qa1 = qa;
    qb2 = qb1;
    qb1 = qb;
    acc[2 * (T0 + k)] = fma(qb, xcb, fma(qa, xca, acc[2 * (T0 + k)]));
    acc[2 * (T0 + k) + 1] = fma(qb, xsb, fma(qa, xsa, acc[2 * (T0 + k) + 1]));
  }
}

// operands of one cell (cells past N read cell N - 1 with x = 0)
struct PairCell {
  double mu, s, zc, zs;
};

// R cells per lane (R even) for the pair (P, LT - P), all shapes fixed at
// compile time; two cells per step, the next two loaded meanwhile
template <typename T, int LT, int P>
__device__ __forceinline__ void pair_cells(int N, const double *__restrict__ chain, const double *__restrict__ trig,
                                           const T *__restrict__ xb, int i0, int span, int R,
                                           T (&acc)[PairShape<LT>::NV]) {
  constexpr int L = LT, M0 = P, M1 = L - P, NA = L - M0 + 1, NB = M1 != M0 ? M0 + 1 : 0;
  constexpr int OFFA = M0 * (L + 1) - M0 * (M0 - 1) / 2, OFFB = M1 * (L + 1) - M1 * (M1 - 1) / 2;
  const double *cm_tab = chain + (L + 1) * (L + 2);
  const T CA = T(cm_tab[M0]), CB = T(cm_tab[M1]);
  auto load = [&](int r, PairCell &c, T &xv) {
    const int i = i0 + r * span, ii = min(i, N - 1);
    c.mu = trig[ii];
    c.s = trig[N + ii];
    c.zc = trig[2 * N + ii];
    c.zs = trig[3 * N + ii];
    xv = xb[ii];
    if (i >= N) xv = T(0);
  };
  PairCell na, nb;
  T nxa, nxb;
  load(0, na, nxa);
  load(1, nb, nxb);
  for (int r = 0; r < R; r += 2) {
    const PairCell ca = na, cb = nb;
    const T xa = nxa, xbv = nxb;
    if (r + 2 < R) {
      load(r + 2, na, nxa);
      load(r + 3, nb, nxb);
    }
    T aac, aas, aar, abc, abs_, abr, bac, bas, bar, bbc, bbs, bbr;
    pair_powers<T, M0, M1>(T(ca.zc), T(ca.zs), T(ca.s), aac, aas, aar, abc, abs_, abr);
    pair_powers<T, M0, M1>(T(cb.zc), T(cb.zs), T(cb.s), bac, bas, bar, bbc, bbs, bbr);
    const T mua = T(ca.mu), mub = T(cb.mu);
    run_chain2<T, 0, NA, OFFA>(chain, mua, mub, CA * aar, CA * bar, xa * aac, xa * aas, xbv * bac, xbv * bas, acc);
    if constexpr (NB > 0)
      run_chain2<T, NA, NB, OFFB>(chain, mua, mub, CB * abr, CB * bbr, xa * abc, xa * abs_, xbv * bbc, xbv * bbs,
                                  acc);
  }
}

template <typename T, int LT, int P>
__device__ __forceinline__ void pair_dispatch(int pair, int N, const double *__restrict__ chain,
                                              const double *__restrict__ trig, const T *__restrict__ xb, int i0,
                                              int span, int R, T (&acc)[PairShape<LT>::NV]) {
  if (pair == P) {
    pair_cells<T, LT, P>(N, chain, trig, xb, i0, span, R, acc);
  } else if constexpr (P < LT / 2) {
    pair_dispatch<T, LT, P + 1>(pair, N, chain, trig, xb, i0, span, R, acc);
  }
}

template <typename T, int LT>
__global__ __launch_bounds__(64 * kPairWaves) void k_sh_analysis_pair(int N, const double *__restrict__ chain,
                                                                      const double *__restrict__ trig,
                                                                      const T *__restrict__ x, long long x_stride,
                                                                      int R, int n_cblk, int n_logical,
                                                                      double *b_part) {
  using S = PairShape<LT>;
  constexpr int L = LT, K = (L + 1) * (L + 1), n_pairs = L / 2 + 1;
  constexpr int NTASK = kPairWaves * S::CH * 4;  // (wave, sum, lane group of 16)
  __shared__ double red[kPairWaves * 64 * (S::CH + 1)];
  __shared__ double red2[NTASK];
  __shared__ double tot[S::NV];
  // blocks that share cells run on one XCD (dispatch round-robins over 8)
  const int g = blockIdx.x, gpad = gridDim.x;
  const int logical = (g & 7) * (gpad >> 3) + (g >> 3);
  if (logical >= n_logical) return;  // whole workgroup
  const int cblk = logical / n_pairs, pair = logical % n_pairs;
  const int b = blockIdx.y;
  const int m0 = pair, m1 = L - pair;
  const bool two = m1 != m0;
  const int nA = L - m0 + 1, n_tot = nA + (two ? m0 + 1 : 0);
  const int offA = m0 * (L + 1) - m0 * (m0 - 1) / 2, offB = m1 * (L + 1) - m1 * (m1 - 1) / 2;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T acc[S::NV];
#pragma unroll
  for (int v = 0; v < S::NV; ++v) acc[v] = T(0);
  const int span = 64 * kPairWaves;
  pair_dispatch<T, LT, 0>(__builtin_amdgcn_readfirstlane(pair), N, chain, trig, x + b * x_stride,
                          cblk * span * R + w * 64 + lane, span, R, acc);
  // block sums: per pass, CH sums of every lane through LDS; tasks of 16
  // lanes, then a fixed-order sum over waves and lane groups
#pragma unroll
  for (int ps = 0; ps < S::NPASS; ++ps) {
    double *rw = red + (w * 64 + lane) * (S::CH + 1);
#pragma unroll
    for (int v = 0; v < S::CH; ++v)
      if (ps * S::CH + v < S::NV) rw[v] = double(acc[ps * S::CH + v]);
    __syncthreads();
    for (int task = threadIdx.x; task < NTASK; task += 64 * kPairWaves) {
      const int tw = task / (S::CH * 4), rem = task % (S::CH * 4), v = rem >> 2, lg = rem & 3;
      const double *src = red + (tw * 64 + lg * 16) * (S::CH + 1) + v;
      double u = 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q) u += src[q * (S::CH + 1)];
      red2[task] = u;
    }
    __syncthreads();
    if (threadIdx.x < S::CH && ps * S::CH + threadIdx.x < S::NV) {
      double u = 0.0;
      for (int tw = 0; tw < kPairWaves; ++tw)
        for (int lg = 0; lg < 4; ++lg) u += red2[tw * S::CH * 4 + threadIdx.x * 4 + lg];
      tot[ps * S::CH + threadIdx.x] = u;
    }
    __syncthreads();
  }
  const int v = threadIdx.x;
  if (v >= 2 * n_tot) return;
  const int t = v >> 1;
  const bool sn = (v & 1) != 0;
  const int m = t < nA ? m0 : m1, kk = t < nA ? t : t - nA, l = m + kk;
  if (sn && m == 0) return;
  const double scale = chain[2 * ((t < nA ? offA : offB) + kk) + 1];
  b_part[(size_t(b) * n_cblk + cblk) * K + l * l + l + (sn ? -m : m)] = scale * tot[v];
}

// Synthesis: workgroup = 64 cells x 4 waves; wave g sums the contiguous
// degrees [mr[g], mr[g+1]) (split on the host into near-equal (l, m) counts)
// for the 64 cells, so every table and coefficient index is wave-uniform
// and contiguous in l (scalar loads in batches): the coefficients in the
// m-major copy of k_sh_coef, the recurrence table m-major
// (tabm: sect[L+1], f1[L+1], then per m the pairs (a, b) for l = m+2..L).
// cos / sin(m lon) by rotation, P_mm by the sectoral product.  The four wave
// sums are added through LDS in a fixed order.
template <typename T, bool SIN>
__device__ __forceinline__ T synth_degree(int L, int m, const double *__restrict__ ab, const double *__restrict__ cc,
                                          T f1, T mu, T pmm, T cm, T sm, T acc) {
  T t = T(cc[0]) * cm;
  if (SIN) t = fma(T(cc[1]), sm, t);
  acc = fma(t, pmm, acc);
  if (m + 1 > L) return acc;
  T p2 = pmm, p1 = f1 * mu * pmm;
  T u = T(cc[2]) * cm;
  if (SIN) u = fma(T(cc[3]), sm, u);
  acc = fma(u, p1, acc);
  const int n = L - m - 1;  // degrees m + 2 .. L
#pragma unroll 4
  for (int j = 0; j < n; ++j) {
    const T p = fma(T(ab[2 * j]) * mu, p1, -(T(ab[2 * j + 1]) * p2));
    p2 = p1;
    p1 = p;
    T v = T(cc[2 * j + 4]) * cm;
    if (SIN) v = fma(T(cc[2 * j + 5]), sm, v);
    acc = fma(v, p, acc);
  }
  return acc;
}

template <typename T, int LT>
__global__ __launch_bounds__(kBlock) void k_sh_synth_rec(int Lrt, int N, const double *__restrict__ tabm,
                                                         const double *__restrict__ trig,
                                                         const double *__restrict__ cmaj, const T *base,
                                                         long long stride, T *out, int4 mr4) {
  __shared__ double part[kBlock];
  const int L = LT > 0 ? LT : Lrt;
  const int b = blockIdx.y;
  const double *tsect = tabm, *tf1 = tabm + (L + 1);
  const double *cb = cmaj + size_t(b) * (L + 1) * (L + 2);
  const int lane = threadIdx.x & 63;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = blockIdx.x * 64 + lane;
  const bool ok = i < N;
  const int mlo = g == 0 ? 0 : (g == 1 ? mr4.x : (g == 2 ? mr4.y : mr4.z));
  const int mhi = g == 0 ? mr4.x : (g == 1 ? mr4.y : (g == 2 ? mr4.z : mr4.w));
  const int ii = ok ? i : 0;
  const T mu = T(trig[ii]), s = T(trig[N + ii]), cl = T(trig[2 * N + ii]), sl = T(trig[3 * N + ii]);
  T acc = T(0), cm = T(1), sm = T(0), pmm = T(1);
  int oab = 0;  // offset of degree m's (a, b) pairs after sect and f1
  for (int q = 1; q < mlo; ++q) {  // advance to m = mlo - 1 ...
    const T c2 = fma(cm, cl, -(sm * sl));
    sm = fma(sm, cl, cm * sl);
    cm = c2;
    pmm = pmm * (T(tsect[q]) * s);
  }
  for (int q = 0; q < mlo; ++q) oab += 2 * max(0, L - q - 1);
#pragma unroll 1
  for (int m = mlo; m < mhi; ++m) {
    const double *ab = tabm + 2 * (L + 1) + oab;
    const double *cc = cb + 2 * (m * (L + 1) - m * (m - 1) / 2);
    if (m == 0) {
      acc = synth_degree<T, false>(L, m, ab, cc, T(tf1[m]), mu, pmm, cm, sm, acc);
    } else {  // ... and one more per degree
      const T c2 = fma(cm, cl, -(sm * sl));
      sm = fma(sm, cl, cm * sl);
      cm = c2;
      pmm = pmm * (T(tsect[m]) * s);
      acc = synth_degree<T, true>(L, m, ab, cc, T(tf1[m]), mu, pmm, cm, sm, acc);
    }
    oab += 2 * max(0, L - m - 1);
  }
  part[threadIdx.x] = double(acc);
  __syncthreads();
  if (threadIdx.x >= 64 || !ok) return;
  const double sum = (part[lane] + part[64 + lane]) + (part[128 + lane] + part[192 + lane]);
  const T add = base ? base[b * stride + i] : T(0);
  out[b * stride + i] = add + T(sum);
}

// Synthesis with the rescaled recurrence, for a degree L fixed at compile
// time: 128 cells x 4 waves per workgroup, each lane two cells (i, i + 64),
// wave g summing the contiguous m-range of synth_split (cos / sin (m lon)
// by rotation, P_mm = P_{m-1,m-1} sect_m cos(lat)).  Each m sums its
// degrees against the degree table (beta_l, S_l c_{l,m}, S_l c_{l,-m}, 0)
// written by put_mmajor, four operations per degree and cell:
//     q_l = mu q_{l-1} + beta_l q_{l-2},  A += (S c)_{l,m} q_l,  B += (S c)_{l,-m} q_l,
// and y = sum_m A_m cos(m lon) + B_m sin(m lon); the four waves' sums are
// added through LDS in a fixed order.
template <typename T>
struct SynthCell {
  T mu, s, cl, sl, cm, sm, pmm, y;
};

template <typename T>
__device__ __forceinline__ void synth_next_m(SynthCell<T> &c, double sect) {
  const T c2 = fma(c.cm, c.cl, -(c.sm * c.sl));
  c.sm = fma(c.sm, c.cl, c.cm * c.sl);
  c.cm = c2;
  c.pmm = c.pmm * (T(sect) * c.s);
}

template <typename T, int LT>
__global__ __launch_bounds__(kBlock) void k_sh_synth_chain(int N, const double *__restrict__ chain,
                                                           const double *__restrict__ tabm,
                                                           const double *__restrict__ trig,
                                                           const double *__restrict__ cmaj, const T *base,
                                                           long long stride, T *out, int4 mr4) {
  __shared__ double part[2][kBlock];
  constexpr int L = LT;
  const int b = blockIdx.y;
  const double *cb = cmaj + size_t(b) * 2 * (L + 1) * (L + 2);
  const int lane = threadIdx.x & 63;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = blockIdx.x * 128 + lane;
  const int mlo = g == 0 ? 0 : (g == 1 ? mr4.x : (g == 2 ? mr4.y : mr4.z));
  const int mhi = g == 0 ? mr4.x : (g == 1 ? mr4.y : (g == 2 ? mr4.z : mr4.w));
  SynthCell<T> u[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int ii = min(i0 + 64 * h, N - 1);
    u[h].mu = T(trig[ii]);
    u[h].s = T(trig[N + ii]);
    u[h].cl = T(trig[2 * N + ii]);
    u[h].sl = T(trig[3 * N + ii]);
    u[h].cm = T(1);
    u[h].sm = T(0);
    u[h].pmm = T(1);
    u[h].y = T(0);
  }
  for (int q = 1; q < mlo; ++q) {  // advance to m = mlo - 1
    synth_next_m(u[0], tabm[q]);
    synth_next_m(u[1], tabm[q]);
  }
#pragma unroll 1
  for (int m = mlo; m < mhi; ++m) {
    if (m > 0) {
      synth_next_m(u[0], tabm[m]);
      synth_next_m(u[1], tabm[m]);
    }
    const int off = m * (L + 1) - m * (m - 1) / 2, nk = L - m + 1;
    const double *e = cb + 4 * off;  // (beta, S c, S s, 0) per degree
    // degrees m and m + 1, then the recurrence (the sine slots of m = 0 are zero)
    const T c0 = T(e[1]), s0 = T(e[2]);
    T A0 = c0 * u[0].pmm, B0 = s0 * u[0].pmm, A1 = c0 * u[1].pmm, B1 = s0 * u[1].pmm;
    T p2 = u[0].pmm, p1 = u[0].pmm, r2 = u[1].pmm, r1 = u[1].pmm;
    if (nk > 1) {
      const T c1 = T(e[5]), s1 = T(e[6]);
      p1 = u[0].mu * p2;
      r1 = u[1].mu * r2;
      A0 = fma(c1, p1, A0);
      B0 = fma(s1, p1, B0);
      A1 = fma(c1, r1, A1);
      B1 = fma(s1, r1, B1);
    }
    e += 8;
#pragma unroll 4
    for (int k = 2; k < nk; ++k, e += 4) {
      const T be = T(e[0]), cc = T(e[1]), ss = T(e[2]);
      const T p = fma(u[0].mu, p1, be * p2);
      const T r = fma(u[1].mu, r1, be * r2);
      p2 = p1;
      p1 = p;
      r2 = r1;
      r1 = r;
      A0 = fma(cc, p, A0);
      B0 = fma(ss, p, B0);
      A1 = fma(cc, r, A1);
      B1 = fma(ss, r, B1);
    }
    u[0].y = fma(B0, u[0].sm, fma(A0, u[0].cm, u[0].y));
    u[1].y = fma(B1, u[1].sm, fma(A1, u[1].cm, u[1].y));
  }
  part[0][threadIdx.x] = double(u[0].y);
  part[1][threadIdx.x] = double(u[1].y);
  __syncthreads();
  if (threadIdx.x >= 128) return;
  const int h = threadIdx.x >> 6, i = i0 + 64 * h;
  if (i >= N) return;
  const double *pp = part[h];
  const double sum = (pp[lane] + pp[64 + lane]) + (pp[128 + lane] + pp[192 + lane]);
  const T add = base ? base[b * stride + i] : T(0);
  out[b * stride + i] = add + T(sum);
}

// Degree ranges of the 4 synthesis waves of a workgroup: near-equal numbers of
// (l, m) pairs (sum of L - m + 1) per lane.
static int4 synth_split(int L) {
  const int total = (L + 1) * (L + 2) / 2;
  int cut[3], g = 0, acc = 0;
  for (int m = 0; m <= L && g < 3; ++m) {
    acc += L - m + 1;
    while (g < 3 && acc * 4 >= total * (g + 1)) cut[g++] = m + 1;
  }
  while (g < 3) cut[g++] = L + 1;
  return make_int4(cut[0], std::max(cut[0], cut[1]), std::max(cut[1], cut[2]), L + 1);
}

// ---------------------------------------------------------------------------
// the projection of a batch as split-BF16 matrix products (MFMA)
// ---------------------------------------------------------------------------
// For nb instances on one grid the analysis B = Y X (K x nb, X = eta of the
// batch, N x nb) and the synthesis X_eff = X + Y^T C are products with one
// inner dimension N or Kp.  Each operand is split as x = hi + lo, hi and lo
// bf16, and a product is hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16
// with FP32 accumulation (the lo*lo term, below 2^-16 relative, is dropped).
//
// Operand fragments of a 16x16x32 product: lane l holds row (or column)
// l % 16 and eight consecutive entries 8 (l / 16) .. + 7 of the inner
// dimension, for A and B alike; the result tile holds, in lane l, rows
// 4 (l / 16) .. + 3 of column l % 16.  The basis is stored twice, each copy
// in the fragment order of its pass, as 16 bf16 per lane (8 hi, then 8 lo):
//   Ya[h][q][lane]  analysis B operand: harmonic 16 h + l % 16, cells
//                   32 q + 8 (l / 16) + i;
//   Ys[t][kb][lane] synthesis B operand: cell 16 t + l % 16, harmonics
//                   32 kb + 8 (l / 16) + i;
// harmonics k >= K (Kp = K rounded up to 32) and cells >= N are zero.  The
// coefficients go to Cs[j][kb][l / 16] (instance j, same harmonic order),
// written by k_sh_coef.  Both passes read the basis once per step for the
// whole batch; instance tiles past nb issue no products.

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma3(const bf16x8 &ahi, const bf16x8 &alo, const bf16x8 &bhi,
                                         const bf16x8 &blo, floatx4 acc) {
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi, acc, 0, 0, 0);
}

// hi = bf16(x) and lo = bf16(x - hi), both rounded to nearest even
__device__ __forceinline__ void split8(const float (&f)[8], bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const __bf16 h = (__bf16)f[i];
    hi[i] = h;
    lo[i] = (__bf16)(f[i] - (float)h);
  }
}

// eta of the batch -> the analysis A operands, Xa[it][q][lane] = 8 hi, then
// 8 lo bf16 of instance 16 it + l % 16 at cells 32 q + 8 (l / 16) .. + 7
// (zero past nb and N).  One wave per fragment (it, q): lane l reads its 8
// cells of one instance (16 rows of 128 contiguous bytes per wave) and the
// wave writes the fragment as 2 KB in one piece.
// IL: x held interleaved (sidx<true>).
template <typename T, bool IL>
__global__ __launch_bounds__(kBlock) void k_sh_pack_mfma(int N, int NB32, int n_frag, int nb,
                                                         const T *__restrict__ x, long long x_stride,
                                                         uint4 *__restrict__ Xa) {
  const int f = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (f >= n_frag) return;
  const int it = f / NB32, q = f % NB32;
  const int j = 16 * it + (lane & 15), c0 = 32 * q + 8 * (lane >> 4);
  const int jc = min(j, nb - 1);
  float v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = (j < nb && c0 + i < N) ? float(x[sidx<IL>(jc, x_stride, c0 + i)]) : 0.f;
  bf16x8 hi, lo;
  split8(v, hi, lo);
  uint4 *dst = Xa + (size_t(f) * 64 + lane) * 2;
  dst[0] = __builtin_bit_cast(uint4, hi);
  dst[1] = __builtin_bit_cast(uint4, lo);
}

// analysis: grid (n_rng cell ranges of qpr blocks of 32 cells, harmonic
// groups of 4 HPW tiles, instance groups of NIT tiles); wave w of a
// workgroup owns HPW harmonic tiles and the group's instance tiles, and
// writes its sums over the range as partial `rng` of b_part
// ([nb][n_part][K]).  The operands of block q + 1 are loaded while block q
// is multiplied.  Tiles past the last (harmonic or instance) repeat the
// last one's operands and are not stored, so the loop has no branch.
template <int NIT, int HPW>
__global__ __launch_bounds__(256) void k_sh_ana_mfma(int K, int KT, int NB32, int qpr, int n_part,
                                                     const uint4 *__restrict__ Ya, const uint4 *__restrict__ Xa,
                                                     int nb, double *__restrict__ b_part) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, r16 = lane & 15;
  const int rng = blockIdx.x;
  const int h0 = (blockIdx.y * 4 + w) * HPW;
  if (h0 >= KT) return;  // a whole wave without tiles (no barrier below)
  const int it0 = blockIdx.z * NIT, n_it = (nb + 15) / 16;
  const int q0 = rng * qpr, q1 = min(NB32, q0 + qpr);
  const uint4 *xp[NIT], *yp[HPW];
#pragma unroll
  for (int it = 0; it < NIT; ++it) xp[it] = Xa + (size_t(min(it0 + it, n_it - 1)) * NB32 * 64 + lane) * 2;
#pragma unroll
  for (int t = 0; t < HPW; ++t) yp[t] = Ya + (size_t(min(h0 + t, KT - 1)) * NB32 * 64 + lane) * 2;
  floatx4 acc[HPW][NIT];
#pragma unroll
  for (int t = 0; t < HPW; ++t)
#pragma unroll
    for (int it = 0; it < NIT; ++it) acc[t][it] = floatx4{0.f, 0.f, 0.f, 0.f};
  uint4 xa[NIT][2], ya[HPW][2];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    xa[it][0] = xp[it][size_t(q0) * 128];
    xa[it][1] = xp[it][size_t(q0) * 128 + 1];
  }
#pragma unroll
  for (int t = 0; t < HPW; ++t) {
    ya[t][0] = yp[t][size_t(q0) * 128];
    ya[t][1] = yp[t][size_t(q0) * 128 + 1];
  }
  for (int q = q0; q < q1; ++q) {
    uint4 xn[NIT][2], yn[HPW][2];
    const size_t on = size_t(min(q + 1, q1 - 1)) * 128;  // the last block loads itself again
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      xn[it][0] = xp[it][on];
      xn[it][1] = xp[it][on + 1];
    }
#pragma unroll
    for (int t = 0; t < HPW; ++t) {
      yn[t][0] = yp[t][on];
      yn[t][1] = yp[t][on + 1];
    }
#pragma unroll
    for (int t = 0; t < HPW; ++t) {
      const bf16x8 bhi = __builtin_bit_cast(bf16x8, ya[t][0]), blo = __builtin_bit_cast(bf16x8, ya[t][1]);
#pragma unroll
      for (int it = 0; it < NIT; ++it)
        acc[t][it] = mfma3(__builtin_bit_cast(bf16x8, xa[it][0]), __builtin_bit_cast(bf16x8, xa[it][1]), bhi,
                           blo, acc[t][it]);
    }
    // issue every load of block q + 1, then the products of block q

#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      xa[it][0] = xn[it][0];
      xa[it][1] = xn[it][1];
    }
#pragma unroll
    for (int t = 0; t < HPW; ++t) {
      ya[t][0] = yn[t][0];
      ya[t][1] = yn[t][1];
    }
  }
  // result tile: rows = instances 16 (it0 + it) + 4 g + r, column = harmonic 16 h + r16
#pragma unroll
  for (int t = 0; t < HPW; ++t) {
    const int h = h0 + t, k = 16 * h + r16;
#pragma unroll
    for (int it = 0; it < NIT; ++it)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = 16 * (it0 + it) + 4 * g + r;
        if (h < KT && j < nb && k < K) b_part[(size_t(j) * n_part + rng) * K + k] = double(acc[t][it][r]);
      }
  }
}

// synthesis: grid (cell tiles in groups of 4 CT, instance groups of NIT
// tiles); wave w owns CT tiles of 16 cells and the group's instance tiles:
// out = base + (C^T Y)^T over all Kp harmonics (the fragment of Cs is the A
// operand, that of Ys the B operand: both hold 8 harmonics of one row or
// column), the operands of harmonic block kb + 1 loaded while block kb is
// multiplied.  Tiles past the last repeat the last one's operands and are
// not stored (no branch in the loop).  IL: base and out held interleaved.
template <typename T, int NIT, int CT, bool IL>
__global__ __launch_bounds__(256) void k_sh_syn_mfma(int N, int NT16, int KB, const uint4 *__restrict__ Ys,
                                                     const uint4 *__restrict__ Cs, int nb, const T *base,
                                                     long long stride, T *out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, r16 = lane & 15;
  const int t0 = (blockIdx.x * 4 + w) * CT;
  if (t0 >= NT16) return;
  const int j0 = blockIdx.y * NIT * 16;
  const uint4 *cp[NIT], *yp[CT];
#pragma unroll
  for (int it = 0; it < NIT; ++it)  // rows past nb read the last instance's coefficients
    cp[it] = Cs + (size_t(min(j0 + it * 16 + r16, nb - 1)) * KB * 4 + g) * 2;
#pragma unroll
  for (int u = 0; u < CT; ++u) yp[u] = Ys + (size_t(min(t0 + u, NT16 - 1)) * KB * 64 + lane) * 2;
  floatx4 acc[CT][NIT];
#pragma unroll
  for (int u = 0; u < CT; ++u)
#pragma unroll
    for (int it = 0; it < NIT; ++it) acc[u][it] = floatx4{0.f, 0.f, 0.f, 0.f};
  uint4 cb[NIT][2], ab[CT][2];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    cb[it][0] = cp[it][0];
    cb[it][1] = cp[it][1];
  }
#pragma unroll
  for (int u = 0; u < CT; ++u) {
    ab[u][0] = yp[u][0];
    ab[u][1] = yp[u][1];
  }
  for (int kb = 0; kb < KB; ++kb) {
    uint4 cn[NIT][2], an[CT][2];
    const int kn = min(kb + 1, KB - 1);  // the last block loads itself again
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      cn[it][0] = cp[it][size_t(kn) * 8];
      cn[it][1] = cp[it][size_t(kn) * 8 + 1];
    }
#pragma unroll
    for (int u = 0; u < CT; ++u) {
      an[u][0] = yp[u][size_t(kn) * 128];
      an[u][1] = yp[u][size_t(kn) * 128 + 1];
    }
#pragma unroll
    for (int u = 0; u < CT; ++u) {
      const bf16x8 yhi = __builtin_bit_cast(bf16x8, ab[u][0]), ylo = __builtin_bit_cast(bf16x8, ab[u][1]);
#pragma unroll
      for (int it = 0; it < NIT; ++it)
        acc[u][it] = mfma3(__builtin_bit_cast(bf16x8, cb[it][0]), __builtin_bit_cast(bf16x8, cb[it][1]), yhi, ylo,
                           acc[u][it]);
    }
    // issue every load of block kb + 1, then the products of block kb

#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      cb[it][0] = cn[it][0];
      cb[it][1] = cn[it][1];
    }
#pragma unroll
    for (int u = 0; u < CT; ++u) {
      ab[u][0] = an[u][0];
      ab[u][1] = an[u][1];
    }
  }
  // result tile (C^T Y): rows = instances j0 + 16 it + 4 g + r, column =
  // cell 16 t + r16.  Held interleaved (FP32), a lane's four rows are four
  // consecutive instances of one group at one cell: 16 contiguous bytes,
  // loaded and stored whole.
  if constexpr (IL && sizeof(T) == 4) {
    const int last_grp = (nb - 1) / 16;
    float4 add4[CT][NIT];
#pragma unroll
    for (int u = 0; u < CT; ++u)
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int c = min(16 * (t0 + u) + r16, N - 1), grp = min((j0 + it * 16) / 16, last_grp);
        add4[u][it] = base ? *reinterpret_cast<const float4 *>(base + (size_t(grp) * stride + c) * 16 + 4 * g)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
    for (int u = 0; u < CT; ++u) {
      const int t = t0 + u, c = 16 * t + r16;
      if (t >= NT16 || c >= N) continue;
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int j = j0 + it * 16 + 4 * g;
        if (j >= nb) continue;
        const float4 a4 = add4[u][it];
        const float4 v = make_float4(float(double(a4.x) + double(acc[u][it][0])),
                                     float(double(a4.y) + double(acc[u][it][1])),
                                     float(double(a4.z) + double(acc[u][it][2])),
                                     float(double(a4.w) + double(acc[u][it][3])));
        float *dst = reinterpret_cast<float *>(out) + (size_t(j / 16) * stride + c) * 16 + 4 * g;
        if (j + 3 < nb) {
          *reinterpret_cast<float4 *>(dst) = v;
        } else {
          dst[0] = v.x;
          if (j + 1 < nb) dst[1] = v.y;
          if (j + 2 < nb) dst[2] = v.z;
        }
      }
    }
    return;
  }
  // the base values are loaded together first (clamped into range)
  T add[CT][NIT][4];
#pragma unroll
  for (int u = 0; u < CT; ++u)
#pragma unroll
    for (int it = 0; it < NIT; ++it)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int c = min(16 * (t0 + u) + r16, N - 1), j = min(j0 + it * 16 + 4 * g + r, nb - 1);
        add[u][it][r] = base ? base[sidx<IL>(j, stride, c)] : T(0);
      }
#pragma unroll
  for (int u = 0; u < CT; ++u) {
    const int t = t0 + u, c = 16 * t + r16;
#pragma unroll
    for (int it = 0; it < NIT; ++it)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = j0 + it * 16 + 4 * g + r;
        if (t < NT16 && c < N && j < nb) out[sidx<IL>(j, stride, c)] = T(double(add[u][it][r]) + double(acc[u][it][r]));
      }
  }
}

// c (index k = l l + l + m) -> its split copy, one lane per k
__global__ __launch_bounds__(kBlock) void k_sh_to_split(int K, int KB, const double *c, unsigned short *cs) {
  const int b = blockIdx.y;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < K) put_split(KB, b, k, c[size_t(b) * K + k], cs);
}

// Gram matrix M = Yt Yt^T (K x K), 32 x 32 output tiles, FP64.
__global__ __launch_bounds__(kBlock) void k_gram(int K, int N, const double *Yt, double *M) {
  __shared__ double A[32][33], Bm[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  if (c0 > r0 + 31) return;  // upper triangle filled by symmetry on the host
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  double acc[4] = {0, 0, 0, 0};
  for (int i0 = 0; i0 < N; i0 += 32) {
    for (int q = 0; q < 4; q++) {
      const int rr = ty + 8 * q;
      const int i = i0 + tx;
      A[rr][tx] = (r0 + rr < K && i < N) ? Yt[size_t(r0 + rr) * N + i] : 0.0;
      Bm[rr][tx] = (c0 + rr < K && i < N) ? Yt[size_t(c0 + rr) * N + i] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < 32; j++) {
      const double bv = Bm[tx][j];
      for (int q = 0; q < 4; q++) acc[q] += A[ty + 8 * q][j] * bv;
    }
    __syncthreads();
  }
  for (int q = 0; q < 4; q++) {
    const int r = r0 + ty + 8 * q, cc = c0 + tx;
    if (r < K && cc < K) M[size_t(r) * K + cc] = acc[q];
  }
}

}  // namespace odis

// ===========================================================================
// context
// ===========================================================================

using namespace odis;

namespace {

thread_local std::string g_last_error;

template <typename T>
struct Ops {
  int *tri_idx = nullptr;
  T *tri_coef = nullptr;
  int *g_idx = nullptr;
  T *g_coef = nullptr;
  T *fcor = nullptr, *farea = nullptr, *fgs = nullptr;
  int *div_idx = nullptr;
  T *div_coef = nullptr;
  T *v[2] = {nullptr, nullptr};
  T *ring_v = nullptr, *ring_e = nullptr;
  T *eta[2] = {nullptr, nullptr};
  // cell-local TRiSK form (compact kernels)
  int *cyc = nullptr, *info = nullptr;
  T *kite = nullptr, *area = nullptr, *flen = nullptr, *fdist = nullptr, *cellS = nullptr;
  T *ts[2] = {nullptr, nullptr};  // [B][2][E] TRiSK partial sums of v[.] from the inner / outer cell
  // patches of the one-kernel step
  int *cell_lo = nullptr, *own_lo = nullptr, *for_ptr = nullptr, *for_idx = nullptr;
  unsigned short *cyc_loc = nullptr;  // [6][N] cycle faces as patch-local indices
  // operand records of the interleaved step ([E][kRecSlots], [N][kRecSlots]), made on first use
  T *face_rec = nullptr, *cell_rec = nullptr;
};

// Device-order operands of the cell-local TRiSK form.
struct CompactHost {
  std::vector<int> cyc, info;                                // [6][N], [N]
  std::vector<double> kite, area, flen, fdist, cellS;        // [6][N], [N], [E], [E], [K][N]
  // patches of the one-kernel step (empty: not available)
  std::vector<int> cell_lo, own_lo, for_ptr, for_idx;        // [P + 1] x 3, [*]
  std::vector<unsigned short> cyc_loc;                       // [6][N]
  int max_loc = 0;                                           // most faces (owned + foreign) of a patch
};

uint64_t hilbert3(uint32_t X0, uint32_t X1, uint32_t X2, int bits) {
  // Skilling's transpose form of the Hilbert index, bits per axis
  uint32_t X[3] = {X0, X1, X2};
  uint32_t M = 1u << (bits - 1), P, Q, t;
  for (Q = M; Q > 1; Q >>= 1) {
    P = Q - 1;
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
  t = 0;
  for (Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; i++) X[i] ^= t;
  uint64_t key = 0;
  for (int b = bits - 1; b >= 0; b--)
    for (int i = 0; i < 3; i++) key = (key << 1) | ((X[i] >> b) & 1u);
  return key;
}

}  // namespace

// Event pairs (start, end) recorded for the kernels of sampled steps.
struct ChunkEvents {
  std::vector<int> mom;   // momentum kernel (k_face, or the one-kernel k_step) of each sampled step
  std::vector<int> cont;  // continuity kernel (k_cell) of each sampled face + cell step
  std::vector<int> sh;    // 3 event pairs (analysis, coefficients, synthesis) of each sampled SH step
};

struct odis_ctx {
  int N = 0, E = 0, B = 0, precision = ODIS_FP64;
  int nb_face = 0, nb_cell = 0;
  int part_cap = 0;            // partials per (step parity, instance): max(face tiles, patches of 64 cells)
  int nfields = 0;
  int energy_quad = 0, quad_drag = 0;
  hipStream_t stream = nullptr;
  std::vector<int> cell_perm, cell_inv, face_perm, face_inv;  // ref -> dev, dev -> ref
  Ops<double> d;
  Ops<float> f;
  int cur = 0;                 // v is v[cur] between calls
  int compact_ok = 0;          // the descriptor has a verified cell-local TRiSK form
  int compact = 0;             // use it (ODIS_OPT_COMPACT)
  int ts_valid = 1;            // Ops::ts holds the partial sums of v[cur]
  std::string compact_why;     // why compact_ok is 0
  int one_kernel_ok = 0;       // patches for the one-kernel step are built
  int one_kernel = 0;          // use it (ODIS_OPT_ONE_KERNEL; compact form)
  int n_patch = 0;
  int patch_max_loc = 0;       // most faces (owned + foreign) of a patch
  int patch = 64;              // cells per patch (ODIS_OPT_PATCH_CELLS: 64, 128, 256)
  int step_blocks = 0;         // workgroups of the one-kernel step (ODIS_OPT_STEP_BLOCKS; 0: one per patch)
  int step_lanes = 4;          // lanes per cell of the one-kernel step (ODIS_OPT_STEP_LANES: 2 or 4)
  int il_opt = 2;              // ODIS_OPT_INTERLEAVE: 0 off, 1 on, 2 auto (batches of >= kInterleaveMin)
  int il_unroll = 0;           // ODIS_OPT_IL_UNROLL: 0 default, 1 (1, 1), 2 (2, 1), 3 (4, 2) faces / cells per lane
  int il_order = 0;            // ODIS_OPT_IL_ORDER: 0 patches per group over the XCDs, 1 groups fastest, 2/4/8 bands
  int il_rec = 1;              // ODIS_OPT_IL_RECORDS: operand records of the interleaved step (when they fit)
  int il_Bp = 0;               // instances of the interleaved copy, padded to kInterleaveGroup (0: none yet)
  void *il_buf = nullptr;      // interleaved v[2], eta[2], ring_v, ring_e, ts[2]
  int il_fresh = 0;            // the interleaved copy holds the current state
  int std_fresh = 1;           // the instance-major arrays hold the current state
  std::string one_kernel_why;
  std::vector<int> gi_host, cyc_host;  // device-order face cells and cell cycles, to rebuild patches
  long long step = 0;          // next step index n, shared by all instances
  long long series_lo = 0;     // first step of the series ring written since the last set_state
  // orbit statistics: FP64 state snapshot [B][E + N] per instance at its last
  // orbit end, whether it holds one, and the small argument / result buffers
  double *snap = nullptr;
  std::vector<int> snap_valid;
  int *orb_sel = nullptr, *orb_ns = nullptr;
  double *orb_out = nullptr;
  double *partials = nullptr;  // [2][B][part_cap], by step parity
  double *series = nullptr;    // [B][cap]
  int *nonfinite = nullptr;    // [B]
  int cap = 1 << 16;
  double *inst = nullptr;      // [B][8]
  double *dconst = nullptr;    // [B]
  double *scratch_out = nullptr;
  // forcing coefficient tables: 2 device + 2 pinned host buffers of
  // chunk x B x 8 doubles, alternated per chunk
  int chunk = 0;
  double *coef_dev[2] = {nullptr, nullptr};
  double *coef_host[2] = {nullptr, nullptr};
  hipEvent_t coef_ev[2] = {nullptr, nullptr};
  int coef_next = 0, coef_last = 0;
  // graphs of graph_steps steps, one per (coefficient buffer, v parity)
  int graph_steps = 0, use_graphs = 0;
  hipGraphExec_t graphs[2][2] = {};
  std::vector<TideDev> tides;
  std::vector<odis_params> params;
  std::vector<std::vector<double>> init_hist_v, init_hist_e;  // history as last set ([M][3])
  std::vector<void *> allocs;
  volatile int stop_flag = 0;
  int timing = 0;
  std::vector<hipEvent_t> ev;
  double t_face = 0, t_cell = 0, t_step = 0;
  double t_sh[3] = {0, 0, 0};  // SH analysis, coefficients, synthesis
  std::vector<ChunkEvents> pending;  // sampled steps of the last odis_step call, read on demand
  // spherical-harmonic coupling
  int sh_mode = ODIS_SH_OFF, sh_L = 0, sh_K = 0;
  double *sh_Yt = nullptr, *sh_Minv = nullptr, *sh_D = nullptr, *sh_b = nullptr, *sh_c = nullptr;
  unsigned int *sh_Yp = nullptr;
  double *sh_tab = nullptr, *sh_trig = nullptr;  // recurrence table, [4][N] sin/cos lat, cos/sin lon
  double *sh_tabm = nullptr, *sh_cmaj = nullptr; // m-major table, m-major coefficients [B][2 (L+1)(L+2)]
  double *sh_chain = nullptr;                    // rescaled recurrence (beta, S) per (l, m), then C_m
  uint4 *sh_Ya = nullptr, *sh_Ys = nullptr;      // batched products: split basis in fragment order
  uint4 *sh_Xa = nullptr;                        // batched products: split eta of the batch (analysis operand)
  unsigned short *sh_Cs = nullptr;               // batched products: split coefficients [B][Kp / 32][4][16]
  int sh_Kp = 0, sh_qpr = 0, sh_rng = 0;         // batched products: padded K, 32-cell blocks per range, ranges
  int sh_part = 0;                               // capacity of analysis partials per instance
  int sh_cpt = 0;                                // cells per lane of the on-the-fly analysis (0: auto)
  void *eta_eff = nullptr;  // [B][N] of the stencil precision
  int *sh_inv = nullptr;    // [N] cell_inv on the device (reference index of device cell r)
  std::vector<void *> sh_allocs;
  std::vector<double> cell_lat, cell_lon;  // reference order
  // lat-lon interpolation operator (eta columns only, device cell indices)
  int ll_rows = 0;
  int *ll_rp = nullptr, *ll_col = nullptr;
  double *ll_val = nullptr, *ll_out = nullptr;
  std::vector<void *> ll_allocs;
  // RBF Cartesian velocity operator: [6][N] device face ids (-1 pad),
  // [3][6][N] coefficients, [3][N] output
  int *rbf_idx = nullptr;
  double *rbf_w = nullptr, *rbf_out = nullptr;
  std::vector<void *> rbf_allocs;
  std::string err;
};

#define CK(x)                                                                              \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      return fail(ctx, ODIS_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));        \
    }                                                                                      \
  } while (0)

static int fail(odis_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  g_last_error = msg;
  return code;
}

// Runs an entry point's body; a host exception (a staging vector that
// cannot be allocated, mostly) ends as an error code, so nothing is thrown
// across the C ABI.
template <typename F>
static int guarded(odis_ctx *ctx, F &&body) {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    try {
      return fail(ctx, ODIS_E_NOMEM, "host allocation failed");
    } catch (...) {
      return ODIS_E_NOMEM;
    }
  } catch (const std::exception &ex) {
    try {
      return fail(ctx, ODIS_E_INVALID, ex.what());
    } catch (...) {
      return ODIS_E_INVALID;
    }
  }
}

template <typename T>
static int dev_alloc(odis_ctx *ctx, T **p, size_t count) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return fail(ctx, ODIS_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return ODIS_OK;
}

template <typename T>
static int upload(odis_ctx *ctx, T **p, const std::vector<T> &h) {
  int rc = dev_alloc(ctx, p, h.size());
  if (rc) return rc;
  CK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return ODIS_OK;
}

template <typename T>
static std::vector<T> cast_vec(const std::vector<double> &x) {
  return std::vector<T>(x.begin(), x.end());
}

// Spatial fields S_k of each potential, from the reference's trig tables
// (src/mesh.cpp:2122-2152); U(x, t) = sum_k c_k(t) S_k(x) with c_k from
// tide_coeffs.  Regrouping of src/tidalPotentials.cpp:82-225.
static int tide_fields(int tide, double lat, double lon, double *S) {
  const double cl = std::cos(lat), sl = std::sin(lat), co = std::cos(lon), so = std::sin(lon);
  const double c2l = std::cos(2.0 * lat), s2l = std::sin(2.0 * lat);
  const double c2o = std::cos(2.0 * lon), s2o = std::sin(2.0 * lon);
  const double cq = std::pow(std::cos(lat), 2.0), sq = std::pow(std::sin(lat), 2.0);
  const double coq = std::pow(std::cos(lon), 2.0), soq = std::pow(std::sin(lon), 2.0);
  switch (tide) {
    case ODIS_ECC:
      S[0] = (1. - 3. * sq) + cq * (3. * c2o);
      S[1] = cq * (4. * s2o);
      return 2;
    case ODIS_OBLIQ:
      S[0] = s2l * co;
      return 1;
    case ODIS_OBLIQ_WEST:
      S[0] = sl * cl * co;
      S[1] = sl * cl * so;
      return 2;
    case ODIS_FULL:
      S[0] = (1. - 3. * sq) + cq * (3. * c2o);
      S[1] = cq * (4. * s2o);
      S[2] = s2l * co;
      return 3;
    case ODIS_FULL2:
      S[0] = -(1 - 3 * c2l);
      S[1] = cl * sl * co;
      S[2] = cl * sl * so;
      S[3] = cq * co;
      S[4] = cq * so;
      S[5] = cq * c2o;
      S[6] = cq * s2o;
      return 7;
    case ODIS_PLANET:
      S[0] = cl * cl * coq;
      S[1] = cl * cl * so * co;
      S[2] = cl * cl * soq;
      S[3] = 1.0;
      return 4;
    default:
      return 0;
  }
}

static TideDev make_tide(const odis_params &p) {
  TideDev t{};
  t.tide = p.tide;
  t.omega = p.omega;
  t.dt = p.dt;
  double R = p.radius;
  if (p.surface == ODIS_LID_LOVE || p.surface == ODIS_LID_MEMBR) R += p.shell_thickness;
  const double base = p.love_reduct * std::pow(p.omega, 2.0) * std::pow(R, 2.0);
  t.f_ecc = 0.75 * base * p.ecc;
  t.f_obl = -3. / 2. * base * p.obliquity;
  t.f_west = 0.5 * base * p.obliquity;
  t.f_full2 = 1 / 32. * base;
  t.ecc = p.ecc;
  t.obl = p.obliquity;
  t.pl_a1 = 421800000.0;
  t.pl_a2 = p.semimajor;
  t.pl_gm = 6.67408e-11 * 8.931938e+22;
  t.radius = R;
  return t;
}

template <typename T>
static int build_ops(odis_ctx *ctx, Ops<T> &o, const std::vector<double> &tri_c, const std::vector<int> &tri_i,
                     const std::vector<double> &gc, const std::vector<int> &gi, const std::vector<double> &fcor,
                     const std::vector<double> &farea, const std::vector<double> &fgs,
                     const std::vector<double> &dc, const std::vector<int> &di) {
  int rc;
  if ((rc = upload(ctx, &o.tri_idx, tri_i))) return rc;
  if ((rc = upload(ctx, &o.tri_coef, cast_vec<T>(tri_c)))) return rc;
  if ((rc = upload(ctx, &o.g_idx, gi))) return rc;
  if ((rc = upload(ctx, &o.g_coef, cast_vec<T>(gc)))) return rc;
  if ((rc = upload(ctx, &o.fcor, cast_vec<T>(fcor)))) return rc;
  if ((rc = upload(ctx, &o.farea, cast_vec<T>(farea)))) return rc;
  if ((rc = upload(ctx, &o.fgs, cast_vec<T>(fgs)))) return rc;
  if ((rc = upload(ctx, &o.div_idx, di))) return rc;
  if ((rc = upload(ctx, &o.div_coef, cast_vec<T>(dc)))) return rc;
  const size_t E = ctx->E, N = ctx->N, B = ctx->B;
  if ((rc = dev_alloc(ctx, &o.v[0], B * E))) return rc;
  if ((rc = dev_alloc(ctx, &o.v[1], B * E))) return rc;
  if ((rc = dev_alloc(ctx, &o.ring_v, B * 3 * E))) return rc;
  if ((rc = dev_alloc(ctx, &o.eta[0], B * N))) return rc;
  if ((rc = dev_alloc(ctx, &o.eta[1], B * N))) return rc;
  if ((rc = dev_alloc(ctx, &o.ring_e, B * 3 * N))) return rc;
  CK(hipMemset(o.v[0], 0, B * E * sizeof(T)));
  CK(hipMemset(o.v[1], 0, B * E * sizeof(T)));
  CK(hipMemset(o.ring_v, 0, B * 3 * E * sizeof(T)));
  CK(hipMemset(o.eta[0], 0, B * N * sizeof(T)));
  CK(hipMemset(o.eta[1], 0, B * N * sizeof(T)));
  CK(hipMemset(o.ring_e, 0, B * 3 * N * sizeof(T)));
  return ODIS_OK;
}


// Patches of at most P consecutive device cells (ceil(N / P) ranges of
// near-equal size): owned face ranges, foreign face lists and patch-local
// cycle indices.  Empty string on success.
static std::string build_patches(int N, int E, const std::vector<int> &gi, int P, CompactHost &h) {
  const int n_patch = (N + P - 1) / P;
  h.cell_lo.resize(n_patch + 1);
  for (int j = 0; j <= n_patch; j++) h.cell_lo[j] = int((long long)j * N / n_patch);
  std::vector<int> minr(E);
  for (int r = 0; r < E; r++) {
    minr[r] = std::min(gi[r], gi[size_t(E) + r]);
    if (r > 0 && minr[r] < minr[r - 1]) return "faces are not sorted by their lower cell rank";
  }
  h.own_lo.assign(n_patch + 1, E);
  h.for_ptr.assign(n_patch + 1, 0);
  h.for_idx.clear();
  h.cyc_loc.assign(size_t(6) * N, 0);
  for (int j = 0, r = 0; j < n_patch; j++) {
    while (r < E && minr[r] < h.cell_lo[j]) r++;
    h.own_lo[j] = r;
  }
  std::vector<int> loc(E, -1), foreign;
  h.max_loc = 0;
  for (int j = 0; j < n_patch; j++) {
    const int c_lo = h.cell_lo[j], c_hi = h.cell_lo[j + 1];
    if (c_hi - c_lo > P) return "patch larger than its workgroup";
    const int lo = h.own_lo[j], hi = h.own_lo[j + 1];
    foreign.clear();
    for (int c = c_lo; c < c_hi; c++)
      for (int q = 0; q < 6; q++) {
        const int f = h.cyc[size_t(q) * N + c];
        if (minr[f] < c_lo) foreign.push_back(f);
        else if (f < lo || f >= hi) return "a face of a patch cell is neither owned nor foreign";
      }
    std::sort(foreign.begin(), foreign.end());
    foreign.erase(std::unique(foreign.begin(), foreign.end()), foreign.end());
    const int n_loc = (hi - lo) + int(foreign.size());
    if (n_loc > 6 * P) return "patch has more faces than its local store";
    h.max_loc = std::max(h.max_loc, n_loc);
    for (int f = lo; f < hi; f++) loc[f] = f - lo;
    for (size_t q = 0; q < foreign.size(); q++) loc[foreign[q]] = (hi - lo) + int(q);
    for (int c = c_lo; c < c_hi; c++)
      for (int q = 0; q < 6; q++) {
        const int l = loc[h.cyc[size_t(q) * N + c]];
        if (l < 0 || l >= n_loc) return "patch-local face index out of range";
        h.cyc_loc[size_t(q) * N + c] = (unsigned short)l;
      }
    for (int f = lo; f < hi; f++) loc[f] = -1;
    for (int f : foreign) loc[f] = -1;
    h.for_idx.insert(h.for_idx.end(), foreign.begin(), foreign.end());
    h.for_ptr[j + 1] = int(h.for_idx.size());
  }
  if (h.for_idx.empty()) h.for_idx.push_back(0);  // keep the device array non-empty
  return std::string();
}

// (Re)place the patch arrays of the one-kernel step.
template <typename T>
static int upload_patches(odis_ctx *ctx, Ops<T> &o, const CompactHost &h) {
  // the cell records hold patch-local indices: both records are made again on next use
  for (void *p : {(void *)o.cell_lo, (void *)o.own_lo, (void *)o.for_ptr, (void *)o.for_idx, (void *)o.cyc_loc,
                  (void *)o.face_rec, (void *)o.cell_rec}) {
    if (!p) continue;
    (void)hipFree(p);
    ctx->allocs.erase(std::remove(ctx->allocs.begin(), ctx->allocs.end(), p), ctx->allocs.end());
  }
  o.cell_lo = o.own_lo = o.for_ptr = o.for_idx = nullptr;
  o.cyc_loc = nullptr;
  o.face_rec = o.cell_rec = nullptr;
  int rc;
  if ((rc = upload(ctx, &o.cell_lo, h.cell_lo))) return rc;
  if ((rc = upload(ctx, &o.own_lo, h.own_lo))) return rc;
  if ((rc = upload(ctx, &o.for_ptr, h.for_ptr))) return rc;
  if ((rc = upload(ctx, &o.for_idx, h.for_idx))) return rc;
  if ((rc = upload(ctx, &o.cyc_loc, h.cyc_loc))) return rc;
  return ODIS_OK;
}

template <typename T>
static int build_compact_ops(odis_ctx *ctx, Ops<T> &o, const CompactHost &h) {
  int rc;
  if ((rc = upload(ctx, &o.cyc, h.cyc))) return rc;
  if ((rc = upload(ctx, &o.info, h.info))) return rc;
  if ((rc = upload(ctx, &o.kite, cast_vec<T>(h.kite)))) return rc;
  if ((rc = upload(ctx, &o.area, cast_vec<T>(h.area)))) return rc;
  if ((rc = upload(ctx, &o.flen, cast_vec<T>(h.flen)))) return rc;
  if ((rc = upload(ctx, &o.fdist, cast_vec<T>(h.fdist)))) return rc;
  if ((rc = upload(ctx, &o.cellS, cast_vec<T>(h.cellS)))) return rc;
  const size_t n = size_t(ctx->B) * 2 * ctx->E;
  for (int k = 0; k < 2; k++) {  // v = 0 at creation: partial sums 0
    if ((rc = dev_alloc(ctx, &o.ts[k], n))) return rc;
    CK(hipMemset(o.ts[k], 0, n * sizeof(T)));
  }
  if (ctx->one_kernel_ok && (rc = upload_patches(ctx, o, h))) return rc;
  return ODIS_OK;
}

extern "C" int odis_destroy(odis_ctx *ctx) {
  if (!ctx) return ODIS_E_INVALID;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &e : ctx->ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 2; k++)
    for (int c = 0; c < 2; c++)
      if (ctx->graphs[k][c]) (void)hipGraphExecDestroy(ctx->graphs[k][c]);
  for (int k = 0; k < 2; k++) {
    if (ctx->coef_ev[k]) (void)hipEventDestroy(ctx->coef_ev[k]);
    if (ctx->coef_host[k]) (void)hipHostFree(ctx->coef_host[k]);
  }
  for (void *p : ctx->allocs) (void)hipFree(p);
  for (void *p : ctx->rbf_allocs) (void)hipFree(p);
  for (void *p : ctx->ll_allocs) (void)hipFree(p);
  for (void *p : ctx->sh_allocs) (void)hipFree(p);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return ODIS_OK;
}

// *building holds the context while it is assembled, for odis_create to free
// if a host exception ends the assembly
static int create_impl(const odis_mesh_desc *m, const odis_params *params, int n_inst, int precision,
                       odis_ctx **out, odis_ctx **building) {
  if (!m || !params || !out || n_inst < 1) return fail(nullptr, ODIS_E_INVALID, "null argument or n_inst < 1");
  if (precision != ODIS_FP64 && precision != ODIS_FP32) return fail(nullptr, ODIS_E_INVALID, "bad precision");
  const int N = m->n_cells, E = m->n_faces;
  if (N < 12 || E != 3 * N - 6) return fail(nullptr, ODIS_E_INVALID, "mesh sizes: need E = 3N-6");
  if (!m->face_nodes || !m->grad_coef || !m->trisk_count || !m->trisk_faces || !m->trisk_coef ||
      !m->face_lat || !m->face_area || !m->cell_faces || !m->div_coef || !m->cell_lat || !m->cell_lon)
    return fail(nullptr, ODIS_E_INVALID, "mesh descriptor has a null array");
  for (int b = 0; b < n_inst; b++) {
    const odis_params &p = params[b];
    if (p.tide < 0 || p.tide > ODIS_NONE) return fail(nullptr, ODIS_E_INVALID, "unknown potential");
    if (p.tide != params[0].tide || p.omega != params[0].omega || p.friction != params[0].friction ||
        p.quad_drag != params[0].quad_drag)
      return fail(nullptr, ODIS_E_UNSUPPORTED, "instances of one context must share tide type, omega and drag law");
    if (!(p.dt > 0) || !(p.h > 0) || !(p.radius > 0))
      return fail(nullptr, ODIS_E_INVALID, "dt, h and radius must be positive");
  }
  // validate every index on the host before anything reaches a kernel
  for (int e = 0; e < E; e++) {
    for (int k = 0; k < 2; k++)
      if (m->face_nodes[2 * e + k] < 0 || m->face_nodes[2 * e + k] >= N)
        return fail(nullptr, ODIS_E_INVALID, "face_nodes out of range");
    const int c = m->trisk_count[e];
    if (c < 0 || c > 10) return fail(nullptr, ODIS_E_INVALID, "trisk_count out of range");
    for (int j = 0; j < c; j++)
      if (m->trisk_faces[10 * e + j] < 0 || m->trisk_faces[10 * e + j] >= E)
        return fail(nullptr, ODIS_E_INVALID, "trisk_faces out of range");
  }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < 6; j++) {
      const int f = m->cell_faces[6 * i + j];
      if (f >= E || (f < 0 && j < 5)) return fail(nullptr, ODIS_E_INVALID, "cell_faces out of range");
    }

  odis_ctx *ctx = new odis_ctx();
  *building = ctx;
  ctx->N = N;
  ctx->E = E;
  ctx->B = n_inst;
  ctx->precision = precision;
  ctx->params.assign(params, params + n_inst);
  ctx->energy_quad = params[0].friction == ODIS_QUADRATIC;
  ctx->quad_drag = params[0].quad_drag != 0;
  ctx->init_hist_v.assign(n_inst, std::vector<double>(size_t(3) * E, 0.0));
  ctx->init_hist_e.assign(n_inst, std::vector<double>(size_t(3) * N, 0.0));
  ctx->cell_lat.assign(m->cell_lat, m->cell_lat + N);
  ctx->cell_lon.assign(m->cell_lon, m->cell_lon + N);

  // -- locality renumbering: cells along a 3-D Hilbert curve, faces by the
  //    device ranks of their two cells
  std::vector<std::pair<uint64_t, int>> ck(N);
  for (int i = 0; i < N; i++) {
    const double la = m->cell_lat[i], lo = m->cell_lon[i];
    const double x = std::cos(la) * std::cos(lo), y = std::cos(la) * std::sin(lo), z = std::sin(la);
    const double s = double((1u << 20) - 1);
    auto q = [&](double v) { return uint32_t(std::lround((v + 1.0) * 0.5 * s)); };
    ck[i] = {hilbert3(q(x), q(y), q(z), 20), i};
  }
  std::stable_sort(ck.begin(), ck.end());
  ctx->cell_perm.resize(N);
  ctx->cell_inv.resize(N);
  for (int r = 0; r < N; r++) {
    ctx->cell_inv[r] = ck[r].second;
    ctx->cell_perm[ck[r].second] = r;
  }
  std::vector<std::pair<std::pair<int, int>, int>> fk(E);
  for (int e = 0; e < E; e++) {
    const int a = ctx->cell_perm[m->face_nodes[2 * e]], b = ctx->cell_perm[m->face_nodes[2 * e + 1]];
    fk[e] = {{std::min(a, b), std::max(a, b)}, e};
  }
  std::stable_sort(fk.begin(), fk.end());
  ctx->face_perm.resize(E);
  ctx->face_inv.resize(E);
  for (int r = 0; r < E; r++) {
    ctx->face_inv[r] = fk[r].second;
    ctx->face_perm[fk[r].second] = r;
  }

  // -- operands in device order ----------------------------------------------
  const odis_params &p0 = params[0];
  std::vector<double> tri_c(size_t(10) * E, 0.0), gc(size_t(2) * E), fcor(E), farea(E);
  std::vector<int> tri_i(size_t(10) * E), gi(size_t(2) * E);
  double S[kMaxForcing];
  const int K = tide_fields(p0.tide, 0.0, 0.0, S);
  ctx->nfields = K;
  std::vector<double> cellS(size_t(K) * N);
  for (int i = 0; i < N; i++) {
    tide_fields(p0.tide, m->cell_lat[i], m->cell_lon[i], S);
    for (int k = 0; k < K; k++) cellS[size_t(k) * N + i] = S[k];
  }
  std::vector<double> fgs(size_t(std::max(K, 1)) * E, 0.0);
  for (int r = 0; r < E; r++) {
    const int e = ctx->face_inv[r];
    const int c = m->trisk_count[e];
    for (int j = 0; j < 10; j++) {
      if (j < c) {
        tri_i[size_t(j) * E + r] = ctx->face_perm[m->trisk_faces[10 * e + j]];
        tri_c[size_t(j) * E + r] = m->trisk_coef[10 * e + j];
      } else {
        tri_i[size_t(j) * E + r] = r;  // padded slot: own face, zero weight
        tri_c[size_t(j) * E + r] = 0.0;
      }
    }
    const int n0 = m->face_nodes[2 * e], n1 = m->face_nodes[2 * e + 1];
    gi[r] = ctx->cell_perm[n0];
    gi[size_t(E) + r] = ctx->cell_perm[n1];
    gc[r] = m->grad_coef[2 * e];
    gc[size_t(E) + r] = m->grad_coef[2 * e + 1];
    fcor[r] = -2.0 * p0.omega * std::sin(m->face_lat[e]);
    farea[r] = m->face_area[e];
    for (int k = 0; k < K; k++)
      fgs[size_t(k) * E + r] = m->grad_coef[2 * e] * cellS[size_t(k) * N + n0] +
                               m->grad_coef[2 * e + 1] * cellS[size_t(k) * N + n1];
  }
  std::vector<double> dc(size_t(6) * N, 0.0);
  std::vector<int> di(size_t(6) * N, 0);
  for (int r = 0; r < N; r++) {
    const int i = ctx->cell_inv[r];
    const int f0 = ctx->face_perm[m->cell_faces[6 * i]];
    for (int j = 0; j < 6; j++) {
      const int f = m->cell_faces[6 * i + j];
      if (f >= 0) {
        di[size_t(j) * N + r] = ctx->face_perm[f];
        dc[size_t(j) * N + r] = m->div_coef[6 * i + j];
      } else {
        di[size_t(j) * N + r] = f0;
        dc[size_t(j) * N + r] = 0.0;
      }
    }
  }

  // -- cell-local TRiSK form, when the descriptor supports it ----------------
  CompactHost ch;
  {
    CompactTrisk ct;
    std::string why;
    try {
      why = build_compact(*m, ct);
    } catch (const std::exception &ex) {
      why = ex.what();
    }
    ctx->compact_ok = why.empty() ? 1 : 0;
    ctx->compact = ctx->compact_ok;
    ctx->compact_why = why;
    if (ctx->compact_ok) {
      ch.cyc.resize(size_t(6) * N);
      ch.kite.resize(size_t(6) * N);
      ch.info.resize(N);
      ch.area.resize(N);
      ch.cellS.assign(size_t(std::max(K, 1)) * N, 0.0);
      for (int r = 0; r < N; r++) {
        const int i = ctx->cell_inv[r];
        for (int q = 0; q < 6; q++) {
          ch.cyc[size_t(q) * N + r] = ctx->face_perm[ct.cyc[size_t(i) * 6 + q]];
          ch.kite[size_t(q) * N + r] = ct.R[size_t(i) * 6 + q];
        }
        ch.info[r] = ct.info[i];
        ch.area[r] = m->cell_area[i];
        for (int k = 0; k < K; k++) ch.cellS[size_t(k) * N + r] = cellS[size_t(k) * N + i];
      }
      ch.flen.resize(E);
      ch.fdist.resize(E);
      for (int r = 0; r < E; r++) {
        ch.flen[r] = m->face_len[ctx->face_inv[r]];
        ch.fdist[r] = m->face_dist[ctx->face_inv[r]];
      }
      const std::string pw = build_patches(N, E, gi, ctx->patch, ch);
      ctx->one_kernel_ok = pw.empty() ? 1 : 0;
      ctx->one_kernel = ctx->one_kernel_ok;
      ctx->one_kernel_why = pw;
      ctx->n_patch = (N + ctx->patch - 1) / ctx->patch;
      ctx->patch_max_loc = ch.max_loc;
      ctx->gi_host = gi;
      ctx->cyc_host = ch.cyc;
    } else {
      ctx->one_kernel_why = "needs the compact TRiSK form";
    }
  }

  int rc = ODIS_OK;
  // A blocking stream: the synchronous hipMemset / hipMemcpy calls of set-up
  // and state transfer run on the null stream, and only a blocking stream is
  // ordered after them (a hipMemset may still be running on the device when
  // it returns).  Contexts' streams still run concurrently with each other.
  hipError_t he = hipStreamCreateWithFlags(&ctx->stream, hipStreamDefault);
  if (he != hipSuccess) rc = fail(ctx, ODIS_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  if (!rc) {
    if (precision == ODIS_FP64)
      rc = build_ops(ctx, ctx->d, tri_c, tri_i, gc, gi, fcor, farea, fgs, dc, di);
    else
      rc = build_ops(ctx, ctx->f, tri_c, tri_i, gc, gi, fcor, farea, fgs, dc, di);
  }
  if (!rc && ctx->compact_ok)
    rc = precision == ODIS_FP64 ? build_compact_ops(ctx, ctx->d, ch) : build_compact_ops(ctx, ctx->f, ch);
  ctx->nb_face = (E + kBlock - 1) / kBlock;
  ctx->nb_cell = (N + kBlock - 1) / kBlock;
  // per-instance parameters
  std::vector<double> inst(size_t(8) * n_inst, 0.0), dconst(n_inst);
  ctx->tides.resize(n_inst);
  for (int b = 0; b < n_inst; b++) {
    const odis_params &p = params[b];
    double *q = &inst[size_t(8) * b];
    q[0] = p.g;
    q[1] = p.quad_drag ? 0.0 : p.alpha;
    q[2] = p.dt;
    q[3] = p.init_load ? 1.0 : 0.0;
    q[4] = p.quad_drag ? p.alpha / p.h : 0.0;
    q[5] = p.h;
    const double area = 4 * kPi * std::pow(p.radius, 2.0);
    dconst[b] = (p.friction == ODIS_QUADRATIC ? p.alpha / p.h : p.alpha * 1000.0 * p.h) / area;
    ctx->tides[b] = make_tide(p);
    ctx->tides[b].nfields = K;
  }
  if (!rc) rc = upload(ctx, &ctx->inst, inst);
  if (!rc) rc = upload(ctx, &ctx->dconst, dconst);
  ctx->part_cap = std::max(ctx->nb_face, (N + 63) / 64);
  if (!rc) rc = dev_alloc(ctx, &ctx->partials, size_t(2) * n_inst * ctx->part_cap);
  if (!rc) rc = dev_alloc(ctx, &ctx->series, size_t(n_inst) * ctx->cap);
  if (!rc) rc = dev_alloc(ctx, &ctx->nonfinite, size_t(n_inst));
  if (!rc) rc = dev_alloc(ctx, &ctx->scratch_out, size_t(n_inst));
  // coefficient tables: chunk steps x B x 8 doubles, <= 4 MiB per buffer
  ctx->chunk = int(std::max<long long>(2, std::min<long long>(4096, (4LL << 20) / (8LL * 8 * n_inst))));
  ctx->chunk &= ~1;
  ctx->graph_steps = std::min(64, ctx->chunk);
  const size_t tab = 8 + size_t(ctx->chunk) * n_inst * 8;
  for (int k = 0; k < 2 && !rc; k++) {
    rc = dev_alloc(ctx, &ctx->coef_dev[k], tab);
    if (!rc) {
      he = hipHostMalloc(reinterpret_cast<void **>(&ctx->coef_host[k]), tab * sizeof(double), hipHostMallocDefault);
      if (he == hipSuccess) he = hipEventCreateWithFlags(&ctx->coef_ev[k], hipEventDisableTiming);
      if (he == hipSuccess) he = hipEventRecord(ctx->coef_ev[k], ctx->stream);
      if (he != hipSuccess) rc = fail(ctx, ODIS_E_HIP, std::string("coef table: ") + hipGetErrorString(he));
    }
  }
  if (!rc) {
    he = hipMemset(ctx->series, 0, sizeof(double) * n_inst * ctx->cap);
    if (he == hipSuccess) he = hipMemset(ctx->nonfinite, 0, sizeof(int) * n_inst);
    if (he == hipSuccess) he = hipDeviceSynchronize();
    if (he != hipSuccess) rc = fail(ctx, ODIS_E_HIP, hipGetErrorString(he));
  }
  *building = nullptr;
  if (rc) {
    const std::string msg = ctx->err;
    odis_destroy(ctx);
    g_last_error = msg;
    return rc;
  }
  *out = ctx;
  return ODIS_OK;
}

extern "C" int odis_create(const odis_mesh_desc *m, const odis_params *params, int n_inst, int precision,
                           odis_ctx **out) {
  odis_ctx *building = nullptr;
  const int rc = guarded(nullptr, [&] { return create_impl(m, params, n_inst, precision, out, &building); });
  if (building) odis_destroy(building);  // assembly ended by an exception; keeps the error text
  return rc;
}

extern "C" const char *odis_last_error(const odis_ctx *ctx) {
  if (ctx && !ctx->err.empty()) return ctx->err.c_str();
  return g_last_error.c_str();
}

// ---------------------------------------------------------------------------
// state in / out (reference numbering and Array2D history layout)
// ---------------------------------------------------------------------------
//
// The AB3 history lives in a 3-slot ring: step n writes its tendency f_n to
// slot n % 3 and reads f_{n-1}, f_{n-2} from slots (n+2) % 3 and (n+1) % 3.
// The reference keeps ds(:,0..2) and shifts them (src/temporalOperators.cpp:
// 36-65); after step m >= 1 it holds (f_m, f_m, f_{m-1}), except after an
// Euler step 0 (f_0, ds1 as set, f_0).  set/get translate between the two.

static int il_out_any(odis_ctx *ctx);  // the instance-major arrays up to date (batches held interleaved)

template <typename T>
static int put_state(odis_ctx *ctx, Ops<T> &o, int b, const double *v, const double *dv, const double *eta,
                     const double *de, long long k) {
  const int E = ctx->E, N = ctx->N;
  const bool load = ctx->params[b].init_load != 0;
  const size_t bE = size_t(b) * E, bN = size_t(b) * N;
  if (v) {
    std::vector<T> h(E);
    for (int r = 0; r < E; r++) h[r] = T(v[ctx->face_inv[r]]);
    CK(hipMemcpy(o.v[ctx->cur] + bE, h.data(), E * sizeof(T), hipMemcpyHostToDevice));
    ctx->ts_valid = 0;
  }
  if (eta) {
    std::vector<T> h(N);
    for (int r = 0; r < N; r++) h[r] = T(eta[ctx->cell_inv[r]]);
    CK(hipMemcpy(o.eta[ctx->cur] + bN, h.data(), N * sizeof(T), hipMemcpyHostToDevice));
  }
  auto put_hist = [&](T *ring, int M, const std::vector<int> &inv, const double *ds,
                      std::vector<double> &keep) -> int {
    std::vector<T> r1(M), r2(M);
    for (int r = 0; r < M; r++) {
      r1[r] = T(ds[3 * inv[r] + 1]);
      r2[r] = T(ds[3 * inv[r] + 2]);
    }
    keep.assign(ds, ds + size_t(3) * M);
    const size_t base = size_t(b) * 3 * M;
    if (k == 1 && !load) {
      CK(hipMemcpy(ring + base, r2.data(), M * sizeof(T), hipMemcpyHostToDevice));
    } else {
      const int s1 = int((k + 2) % 3), s2 = int((k + 1) % 3);
      CK(hipMemcpy(ring + base + size_t(s1) * M, r1.data(), M * sizeof(T), hipMemcpyHostToDevice));
      CK(hipMemcpy(ring + base + size_t(s2) * M, r2.data(), M * sizeof(T), hipMemcpyHostToDevice));
    }
    return ODIS_OK;
  };
  int rc;
  if (dv && (rc = put_hist(o.ring_v, E, ctx->face_inv, dv, ctx->init_hist_v[b]))) return rc;
  if (de && (rc = put_hist(o.ring_e, N, ctx->cell_inv, de, ctx->init_hist_e[b]))) return rc;
  return ODIS_OK;
}

static int set_state_impl(odis_ctx *ctx, int inst, const double *v, const double *dvdt, const double *eta,
                          const double *detadt, long long iter) {
  if (!ctx) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  if (iter < 0) return fail(ctx, ODIS_E_INVALID, "iter < 0");
  int rc = il_out_any(ctx);  // the other instances' state, when the batch stepped interleaved
  if (rc) return rc;
  CK(hipStreamSynchronize(ctx->stream));
  ctx->step = iter;  // the step counter is shared by all instances of a context
  ctx->series_lo = iter;  // ring entries below iter belong to an earlier trajectory
  ctx->il_fresh = 0;
  return ctx->precision == ODIS_FP64 ? put_state(ctx, ctx->d, inst, v, dvdt, eta, detadt, iter)
                                     : put_state(ctx, ctx->f, inst, v, dvdt, eta, detadt, iter);
}

extern "C" int odis_set_state(odis_ctx *ctx, int inst, const double *v, const double *dvdt, const double *eta,
                              const double *detadt, long long iter) {
  return guarded(ctx, [&] { return set_state_impl(ctx, inst, v, dvdt, eta, detadt, iter); });
}

template <typename T>
static int get_state_t(odis_ctx *ctx, Ops<T> &o, int b, double *v, double *dv, double *eta, double *de,
                       long long k) {
  const int E = ctx->E, N = ctx->N;
  const bool load = ctx->params[b].init_load != 0;
  const size_t bE = size_t(b) * E, bN = size_t(b) * N;
  if (v) {
    std::vector<T> h(E);
    CK(hipMemcpy(h.data(), o.v[ctx->cur] + bE, E * sizeof(T), hipMemcpyDeviceToHost));
    for (int r = 0; r < E; r++) v[ctx->face_inv[r]] = double(h[r]);
  }
  if (eta) {
    std::vector<T> h(N);
    CK(hipMemcpy(h.data(), o.eta[ctx->cur] + bN, N * sizeof(T), hipMemcpyDeviceToHost));
    for (int r = 0; r < N; r++) eta[ctx->cell_inv[r]] = double(h[r]);
  }
  auto get_hist = [&](T *ring, int M, const std::vector<int> &inv, double *ds,
                      const std::vector<double> &keep) -> int {
    std::vector<T> all(size_t(3) * M);
    CK(hipMemcpy(all.data(), ring + size_t(b) * 3 * M, all.size() * sizeof(T), hipMemcpyDeviceToHost));
    for (int r = 0; r < M; r++) {
      const int i = inv[r];
      if (k == 0) {  // no step since set_state: the history as set
        for (int q = 0; q < 3; q++) ds[3 * i + q] = keep[3 * i + q];
        continue;
      }
      const double last = double(all[size_t((k - 1) % 3) * M + r]);
      ds[3 * i + 0] = last;
      if (k == 1 && !load) {
        ds[3 * i + 1] = keep[3 * i + 1];
        ds[3 * i + 2] = last;
      } else {
        ds[3 * i + 1] = last;
        ds[3 * i + 2] = double(all[size_t((k + 1) % 3) * M + r]);
      }
    }
    return ODIS_OK;
  };
  int rc;
  if (dv && (rc = get_hist(o.ring_v, E, ctx->face_inv, dv, ctx->init_hist_v[b]))) return rc;
  if (de && (rc = get_hist(o.ring_e, N, ctx->cell_inv, de, ctx->init_hist_e[b]))) return rc;
  return ODIS_OK;
}

static int get_state_impl(odis_ctx *ctx, int inst, double *v, double *dvdt, double *eta, double *detadt,
                          long long *iter) {
  if (!ctx) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  const int rc = il_out_any(ctx);
  if (rc) return rc;
  CK(hipStreamSynchronize(ctx->stream));
  if (iter) *iter = ctx->step;
  return ctx->precision == ODIS_FP64 ? get_state_t(ctx, ctx->d, inst, v, dvdt, eta, detadt, ctx->step)
                                     : get_state_t(ctx, ctx->f, inst, v, dvdt, eta, detadt, ctx->step);
}

extern "C" int odis_get_state(odis_ctx *ctx, int inst, double *v, double *dvdt, double *eta, double *detadt,
                              long long *iter) {
  return guarded(ctx, [&] { return get_state_impl(ctx, inst, v, dvdt, eta, detadt, iter); });
}

// ---------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------

template <typename T>
static FaceArgs<T> face_args(odis_ctx *ctx, Ops<T> &o) {
  FaceArgs<T> a;
  a.E = ctx->E;
  a.nb = ctx->nb_face;
  a.B = ctx->B;
  a.tri_idx = o.tri_idx;
  a.tri_coef = o.tri_coef;
  a.g_idx = o.g_idx;
  a.g_coef = o.g_coef;
  a.fcor = o.fcor;
  a.farea = o.farea;
  a.fgs = o.fgs;
  a.nfields = ctx->nfields;
  a.v_cur = o.v[ctx->cur];
  a.v_nxt = o.v[ctx->cur ^ 1];
  a.ring = o.ring_v;
  a.eta = ctx->sh_mode != ODIS_SH_OFF ? static_cast<const T *>(ctx->eta_eff) : o.eta[ctx->cur];
  a.v_stride = ctx->E;
  a.ring_stride = 3LL * ctx->E;
  a.eta_stride = ctx->N;
  a.partials = ctx->partials;
  a.inst = ctx->inst;
  a.quad_drag = ctx->quad_drag;
  a.N = ctx->N;
  a.fdist = o.fdist;
  a.ts = o.ts[ctx->cur];
  a.ts_stride = 2LL * ctx->E;
  a.cellS = o.cellS;
  return a;
}

template <typename T>
static CellArgs<T> cell_args(odis_ctx *ctx, Ops<T> &o) {
  CellArgs<T> a;
  a.N = ctx->N;
  a.nb = ctx->nb_cell;
  a.nb_face = ctx->nb_face;
  a.div_idx = o.div_idx;
  a.div_coef = o.div_coef;
  a.v_new = o.v[ctx->cur ^ 1];
  a.eta = o.eta[ctx->cur];
  a.eta_nxt = o.eta[ctx->cur ^ 1];
  a.ring = o.ring_e;
  a.v_stride = ctx->E;
  a.eta_stride = ctx->N;
  a.ring_stride = 3LL * ctx->N;
  a.partials = ctx->partials;
  a.series = ctx->series;
  a.nonfinite = ctx->nonfinite;
  a.cap = ctx->cap;
  a.inst = ctx->inst;
  a.dconst = ctx->dconst;
  a.E = ctx->E;
  a.cyc = o.cyc;
  a.kite = o.kite;
  a.info = o.info;
  a.area = o.area;
  a.flen = o.flen;
  a.ts = o.ts[ctx->cur ^ 1];
  a.ts_stride = 2LL * ctx->E;
  return a;
}

// Fill and upload the forcing coefficients of steps [n0, n0 + count) into the
// next table buffer; returns its device pointer.
static int stage_coefficients(odis_ctx *ctx, long long n0, int count, const double **dev) {
  const int k = ctx->coef_next;
  ctx->coef_next ^= 1;
  // the previous upload from this pinned buffer must have been consumed
  CK(hipEventSynchronize(ctx->coef_ev[k]));
  double *h = ctx->coef_host[k];
  const int B = ctx->B;
  for (int q = 0; q < 8; q++) h[q] = 0.0;
  h[0] = double(n0);  // exact for n0 < 2^53
  for (int s = 0; s < count; s++)
    for (int b = 0; b < B; b++) {
      double *c = h + 8 + (size_t(s) * B + b) * 8;
      for (int q = 0; q < 8; q++) c[q] = 0.0;
      const double dt = ctx->params[b].dt;
      // forcing at current_time + dt, current_time = dt * iter (src/timeIntegrator.cpp:218)
      tide_coeffs(ctx->tides[b], dt * double(n0 + s) + dt, c);
    }
  CK(hipMemcpyAsync(ctx->coef_dev[k], h, sizeof(double) * (8 + size_t(count) * B * 8), hipMemcpyHostToDevice,
                    ctx->stream));
  CK(hipEventRecord(ctx->coef_ev[k], ctx->stream));
  *dev = ctx->coef_dev[k];
  ctx->coef_last = k;
  return ODIS_OK;
}

// On-the-fly analysis with CPT cells per lane: the degree-32 kernel with its
// loops unrolled at compile time, or the runtime-degree one.
template <typename T, int CPT>
static void launch_analysis_rec(odis_ctx *ctx, int nb, int L, const T *x, long long stride, int n_chunks,
                                hipEvent_t e0, hipEvent_t e1) {
  const int n_cblk = (n_chunks + kShWaves - 1) / kShWaves;
  const dim3 ga(n_cblk * (L / 2 + 1), nb), blk(64 * kShWaves);
  const size_t lds = sizeof(double) * kShWaves * 2 * (L + 2);
  if (L == 32)
    hipExtLaunchKernelGGL((k_sh_analysis_rec<T, 32, CPT>), ga, blk, lds, ctx->stream, e0, e1, 0, L, ctx->N,
                          ctx->sh_tabm, ctx->sh_trig, x, stride, n_chunks, ctx->sh_b);
  else
    hipExtLaunchKernelGGL((k_sh_analysis_rec<T, 0, CPT>), ga, blk, lds, ctx->stream, e0, e1, 0, L, ctx->N,
                          ctx->sh_tabm, ctx->sh_trig, x, stride, n_chunks, ctx->sh_b);
}

// Degrees with a compiled pair analysis (the others use the per-degree one).
static bool pair_degree(int L) { return L == 8 || L == 16 || L == 32; }

// Cells per lane of the pair analysis when none is set: the grid
// (ceil(N / (64 kPairWaves R)) cell blocks x pairs x instances) fits the
// workgroups the device holds at once (one round, no tail), R even and >= 4.
template <typename T>
static int pair_auto_cells(odis_ctx *ctx, int L, int nb) {
  int per_cu = 0, n_cu = 0, dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  const void *k = L == 32 ? reinterpret_cast<const void *>(&k_sh_analysis_pair<T, 32>)
                          : (L == 16 ? reinterpret_cast<const void *>(&k_sh_analysis_pair<T, 16>)
                                     : reinterpret_cast<const void *>(&k_sh_analysis_pair<T, 8>));
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64 * kPairWaves, 0);
  if (e != hipSuccess || per_cu <= 0 || n_cu <= 0) return 16;
  const int n_pairs = L / 2 + 1;
  const long long slots = (long long)per_cu * n_cu;
  const long long n_cblk = std::max(1LL, slots / ((long long)n_pairs * nb));
  const long long span = 64LL * kPairWaves * n_cblk;
  int R = int((ctx->N + span - 1) / span);
  R += R & 1;
  return std::max(4, R);
}

template <typename T, int LT>
static void launch_analysis_pair(odis_ctx *ctx, int nb, const T *x, long long stride, int cpt, int n_cblk,
                                 hipEvent_t e0, hipEvent_t e1) {
  const int n_logical = n_cblk * (LT / 2 + 1);
  const dim3 ga((n_logical + 7) / 8 * 8, nb), blk(64 * kPairWaves);
  hipExtLaunchKernelGGL((k_sh_analysis_pair<T, LT>), ga, blk, 0, ctx->stream, e0, e1, 0, ctx->N, ctx->sh_chain,
                        ctx->sh_trig, x, stride, cpt, n_cblk, n_logical, ctx->sh_b);
}

// Batched products: NIT instance tiles per wave (all of the batch up to
// 128 instances) and HPW harmonic or CT cell tiles per wave, fewer as NIT
// grows, so that operands of two blocks and the accumulators stay in
// registers.  The analysis ranges depend on N only, and every kernel sums
// an instance's products in the same order whatever NIT, so an instance's
// results do not depend on the rest of its batch.
template <typename T, int NIT, int CT, bool IL>
static void launch_syn_mfma(odis_ctx *ctx, int nb, const T *base, long long stride, T *out, hipEvent_t e0,
                            hipEvent_t e1) {
  const int NT16 = (ctx->N + 15) / 16, n_it = (nb + 15) / 16;
  const dim3 grid((NT16 + 4 * CT - 1) / (4 * CT), (n_it + NIT - 1) / NIT);
  hipExtLaunchKernelGGL((k_sh_syn_mfma<T, NIT, CT, IL>), grid, dim3(256), 0, ctx->stream, e0, e1, 0, ctx->N, NT16,
                        ctx->sh_Kp / 32, ctx->sh_Ys, reinterpret_cast<const uint4 *>(ctx->sh_Cs), nb, base, stride,
                        out);
}

template <typename T, int NIT, bool IL>
static void launch_mfma_t(odis_ctx *ctx, bool ana, int nb, const T *x, const T *base, long long stride, T *out,
                          hipEvent_t e0, hipEvent_t e1) {
  // HPW of the analysis: 3 at NIT 5 re-reads the eta fragments in 2 harmonic
  // groups instead of 3 (163842 cells, 69 instances: 77 against 87 us)
  constexpr int W = NIT <= 3 ? 4 : (NIT <= 5 ? 3 : (NIT <= 6 ? 2 : 1));
  constexpr int CT = NIT == 1 ? 8 : (NIT == 2 ? 6 : (NIT == 3 ? 4 : (NIT == 4 ? 3 : (NIT <= 6 ? 2 : 1))));
  const int N = ctx->N, NB32 = (N + 31) / 32, NT16 = (N + 15) / 16, n_it = (nb + 15) / 16;
  if (ana) {
    const int KT = ctx->sh_Kp / 16;
    const int n_frag = n_it * NB32;
    hipLaunchKernelGGL((k_sh_pack_mfma<T, IL>), dim3((n_frag + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), 0,
                       ctx->stream, N, NB32, n_frag, nb, x, stride, ctx->sh_Xa);
    const dim3 grid(ctx->sh_rng, (KT + 4 * W - 1) / (4 * W), (n_it + NIT - 1) / NIT);
    hipExtLaunchKernelGGL((k_sh_ana_mfma<NIT, W>), grid, dim3(256), 0, ctx->stream, e0, e1, 0, ctx->sh_K, KT, NB32,
                          ctx->sh_qpr, ctx->sh_rng, ctx->sh_Ya, ctx->sh_Xa, nb, ctx->sh_b);
  } else if constexpr (NIT == 5) {
    // cell tiles per wave by grid size (measured on the uneven batch's
    // groups: 4 at 163842 cells, 3 at 40962, 2 at 2562 cells, where fewer
    // tiles per wave keep more workgroups)
    if (NT16 >= 8192)
      launch_syn_mfma<T, NIT, 4, IL>(ctx, nb, base, stride, out, e0, e1);
    else if (NT16 >= 1024)
      launch_syn_mfma<T, NIT, 3, IL>(ctx, nb, base, stride, out, e0, e1);
    else
      launch_syn_mfma<T, NIT, 2, IL>(ctx, nb, base, stride, out, e0, e1);
  } else {
    launch_syn_mfma<T, NIT, CT, IL>(ctx, nb, base, stride, out, e0, e1);
  }
}

template <typename T, bool IL>
static void launch_mfma_l(odis_ctx *ctx, bool ana, int nb, const T *x, const T *base, long long stride, T *out,
                          hipEvent_t e0, hipEvent_t e1) {
  switch (std::min(8, (nb + 15) / 16)) {
    case 1: launch_mfma_t<T, 1, IL>(ctx, ana, nb, x, base, stride, out, e0, e1); break;
    case 2: launch_mfma_t<T, 2, IL>(ctx, ana, nb, x, base, stride, out, e0, e1); break;
    case 3: launch_mfma_t<T, 3, IL>(ctx, ana, nb, x, base, stride, out, e0, e1); break;
    case 4: launch_mfma_t<T, 4, IL>(ctx, ana, nb, x, base, stride, out, e0, e1); break;
    case 5: launch_mfma_t<T, 5, IL>(ctx, ana, nb, x, base, stride, out, e0, e1); break;
    case 6: launch_mfma_t<T, 6, IL>(ctx, ana, nb, x, base, stride, out, e0, e1); break;
    case 7: launch_mfma_t<T, 7, IL>(ctx, ana, nb, x, base, stride, out, e0, e1); break;
    default: launch_mfma_t<T, 8, IL>(ctx, ana, nb, x, base, stride, out, e0, e1); break;
  }
}

// il: x, base and out held interleaved (a batch stepping interleaved)
template <typename T>
static void launch_mfma(odis_ctx *ctx, bool ana, int nb, const T *x, const T *base, long long stride, T *out,
                        hipEvent_t e0, hipEvent_t e1, bool il) {
  if (il)
    launch_mfma_l<T, true>(ctx, ana, nb, x, base, stride, out, e0, e1);
  else
    launch_mfma_l<T, false>(ctx, ana, nb, x, base, stride, out, e0, e1);
}

// out = base + Yt^T D Minv Yt x for nb instances (x, base, out with instance
// stride `stride`; D null: plain LSQ coefficients; base null: synthesis
// alone).  ana / syn: run the analysis + coefficients / the synthesis.
// ev (nullable): 6 dispatch-stamped events (analysis, coefficients,
// synthesis).
// il (batched products only): x, base and out held interleaved.
template <typename T>
static int sh_apply(odis_ctx *ctx, const T *x, long long stride, int nb, const double *D, const T *base, T *out,
                    double *c, bool ana, bool syn, hipEvent_t *ev, bool il = false) {
  const int K = ctx->sh_K, N = ctx->N, L = ctx->sh_L;
  const dim3 blk(kBlock);
  const bool split = ctx->sh_mode == ODIS_SH_SPLIT_BF16;
  // on the fly: chunks of 64 CPT cells, partials per workgroup of kShWaves chunks
  const bool by_pair = ctx->sh_mode == ODIS_SH_RECUR && pair_degree(L);
  const bool mfma = ctx->sh_mode == ODIS_SH_MFMA;
  // cells per lane: the option, else for the pair analysis the fewest that
  // fit its grid in one round of resident workgroups, 8 for the per-degree one
  const int cpt = ctx->sh_cpt ? ctx->sh_cpt : (by_pair ? pair_auto_cells<T>(ctx, L, nb) : 8);
  const int n_chunks = (N + 64 * cpt - 1) / (64 * cpt);
  // cmaj holds S_l c for the compiled-degree synthesis
  const double *chain = by_pair ? ctx->sh_chain : nullptr;
  const int n_cblk = (N + 64 * kPairWaves * cpt - 1) / (64 * kPairWaves * cpt);
  const int n_part =
      mfma ? ctx->sh_rng
           : (by_pair ? n_cblk
                      : (ctx->sh_mode == ODIS_SH_RECUR ? (n_chunks + kShWaves - 1) / kShWaves : kShChunks));
  if (n_part > ctx->sh_part) return fail(ctx, ODIS_E_INVALID, "analysis partials exceed their buffer");
  if (mfma) {
    if (ana)
      launch_mfma<T>(ctx, true, nb, x, nullptr, stride, nullptr, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr, il);
  } else if (by_pair) {
    if (ana) {
      hipEvent_t e0 = ev ? ev[0] : nullptr, e1 = ev ? ev[1] : nullptr;
      if (L == 32)
        launch_analysis_pair<T, 32>(ctx, nb, x, stride, cpt, n_cblk, e0, e1);
      else if (L == 16)
        launch_analysis_pair<T, 16>(ctx, nb, x, stride, cpt, n_cblk, e0, e1);
      else
        launch_analysis_pair<T, 8>(ctx, nb, x, stride, cpt, n_cblk, e0, e1);
    }
  } else if (ctx->sh_mode == ODIS_SH_RECUR) {
    if (ana) {
      hipEvent_t e0 = ev ? ev[0] : nullptr, e1 = ev ? ev[1] : nullptr;
      if (cpt == 4)
        launch_analysis_rec<T, 4>(ctx, nb, L, x, stride, n_chunks, e0, e1);
      else if (cpt == 16)
        launch_analysis_rec<T, 16>(ctx, nb, L, x, stride, n_chunks, e0, e1);
      else
        launch_analysis_rec<T, 8>(ctx, nb, L, x, stride, n_chunks, e0, e1);
    }
  } else if (ana) {
    const dim3 ga(K * kShChunks, nb);
    if (split)
      hipExtLaunchKernelGGL((k_sh_analysis<T, true>), ga, blk, 0, ctx->stream, ev ? ev[0] : nullptr,
                            ev ? ev[1] : nullptr, 0, K, N, ctx->sh_Yt, ctx->sh_Yp, x, stride, ctx->sh_b);
    else
      hipExtLaunchKernelGGL((k_sh_analysis<T, false>), ga, blk, 0, ctx->stream, ev ? ev[0] : nullptr,
                            ev ? ev[1] : nullptr, 0, K, N, ctx->sh_Yt, ctx->sh_Yp, x, stride, ctx->sh_b);
  }
  if (ana) {
    // partial sums, then c = D Minv b; events around both
    const double *bsum = ctx->sh_b + size_t(nb) * n_part * K;
    hipExtLaunchKernelGGL(k_sh_bsum, dim3((K + 63) / 64, nb), blk, 0, ctx->stream,
                          ev ? ev[2] : nullptr, nullptr, 0, K, n_part, ctx->sh_b);
    hipExtLaunchKernelGGL(k_sh_coef, dim3(K, nb), blk, 0, ctx->stream, nullptr, ev ? ev[3] : nullptr,
                          0, K, ctx->sh_Minv, D, bsum, c, chain,
                          ctx->sh_mode == ODIS_SH_RECUR ? ctx->sh_cmaj : static_cast<double *>(nullptr),
                          mfma ? ctx->sh_Cs : static_cast<unsigned short *>(nullptr), ctx->sh_Kp / 32);
  }
  if (syn && mfma) {
    if (!ana)  // coefficients from the caller: their split copy first
      hipLaunchKernelGGL(k_sh_to_split, dim3((K + kBlock - 1) / kBlock, nb), blk, 0, ctx->stream, K,
                         ctx->sh_Kp / 32, c, ctx->sh_Cs);
    launch_mfma<T>(ctx, false, nb, nullptr, base, stride, out, ev ? ev[4] : nullptr, ev ? ev[5] : nullptr, il);
  } else if (syn && ctx->sh_mode == ODIS_SH_RECUR) {
    if (!ana)  // coefficients from the caller: their m-major copy first
      hipLaunchKernelGGL(k_sh_to_mmajor, dim3((K + kBlock - 1) / kBlock, nb), blk, 0, ctx->stream, K, c, chain,
                         ctx->sh_cmaj);
    const dim3 gs((N + 63) / 64, nb), gs2((N + 127) / 128, nb);
    const int4 mr = synth_split(L);
    hipEvent_t e4 = ev ? ev[4] : nullptr, e5 = ev ? ev[5] : nullptr;
    if (L == 32)
      hipExtLaunchKernelGGL((k_sh_synth_chain<T, 32>), gs2, blk, 0, ctx->stream, e4, e5, 0, N, ctx->sh_chain,
                            ctx->sh_tabm, ctx->sh_trig, ctx->sh_cmaj, base, stride, out, mr);
    else if (L == 16)
      hipExtLaunchKernelGGL((k_sh_synth_chain<T, 16>), gs2, blk, 0, ctx->stream, e4, e5, 0, N, ctx->sh_chain,
                            ctx->sh_tabm, ctx->sh_trig, ctx->sh_cmaj, base, stride, out, mr);
    else if (L == 8)
      hipExtLaunchKernelGGL((k_sh_synth_chain<T, 8>), gs2, blk, 0, ctx->stream, e4, e5, 0, N, ctx->sh_chain,
                            ctx->sh_tabm, ctx->sh_trig, ctx->sh_cmaj, base, stride, out, mr);
    else
      hipExtLaunchKernelGGL((k_sh_synth_rec<T, 0>), gs, blk, 0, ctx->stream, e4, e5, 0, L, N, ctx->sh_tabm,
                            ctx->sh_trig, ctx->sh_cmaj, base, stride, out, mr);
  } else if (syn) {
    const dim3 gs((N + 63) / 64, nb);
    const size_t lds_s = sizeof(double) * (K + kBlock);
    if (split)
      hipExtLaunchKernelGGL((k_sh_synth<T, true>), gs, blk, lds_s, ctx->stream, ev ? ev[4] : nullptr,
                            ev ? ev[5] : nullptr, 0, K, N, ctx->sh_Yt, ctx->sh_Yp, c, base, stride, out);
    else
      hipExtLaunchKernelGGL((k_sh_synth<T, false>), gs, blk, lds_s, ctx->stream, ev ? ev[4] : nullptr,
                            ev ? ev[5] : nullptr, 0, K, N, ctx->sh_Yt, ctx->sh_Yp, c, base, stride, out);
  }
  CK(hipGetLastError());
  return ODIS_OK;
}

// eta_eff = eta + Yt^T D Minv Yt eta for every instance (three kernels)
template <typename T>
static int sh_effective_height(odis_ctx *ctx, Ops<T> &o, hipEvent_t *ev) {
  return sh_apply<T>(ctx, o.eta[ctx->cur], ctx->N, ctx->B, ctx->sh_D, o.eta[ctx->cur],
                     static_cast<T *>(ctx->eta_eff), ctx->sh_c, true, true, ev);
}


static int reserve_events(odis_ctx *ctx, size_t n) {
  while (ctx->ev.size() < n) {
    hipEvent_t e;
    CK(hipEventCreate(&e));
    ctx->ev.push_back(e);
  }
  return ODIS_OK;
}

// TRiSK partial sums of the current velocity, when the last state upload
// left them stale (compact form only).
template <typename T>
static int ensure_ts(odis_ctx *ctx, Ops<T> &o) {
  if (!ctx->compact || ctx->ts_valid) return ODIS_OK;
  CellArgs<T> ca = cell_args(ctx, o);
  ca.v_new = o.v[ctx->cur];
  ca.ts = o.ts[ctx->cur];
  hipLaunchKernelGGL((k_cell<T, true, true>), dim3(ctx->nb_cell, ctx->B), dim3(kBlock), 0, ctx->stream, ca,
                     static_cast<const double *>(nullptr), 0);
  CK(hipGetLastError());
  ctx->ts_valid = 1;
  return ODIS_OK;
}

// Workgroups per instance of the one-kernel step.
static int step_blocks(const odis_ctx *ctx) {
  return ctx->step_blocks > 0 ? std::min(ctx->step_blocks, ctx->n_patch) : ctx->n_patch;
}

template <typename T>
static StepArgs<T> step_args(odis_ctx *ctx, Ops<T> &o) {
  StepArgs<T> a;
  a.f = face_args(ctx, o);
  a.c = cell_args(ctx, o);
  a.cell_lo = o.cell_lo;
  a.own_lo = o.own_lo;
  a.for_ptr = o.for_ptr;
  a.for_idx = o.for_idx;
  a.cyc_loc = o.cyc_loc;
  a.n_patch = ctx->n_patch;
  a.n_blocks = step_blocks(ctx);
  a.max_loc = ctx->patch_max_loc;
  a.face_rec = a.cell_rec = nullptr;
  a.il_band = 1;
  return a;
}

// ---------------------------------------------------------------------------
// batches held interleaved (k_step_il)
// ---------------------------------------------------------------------------
// A batch may step with its state interleaved (groups of kIlGroup
// instances, sidx<true>); the instance-major arrays are then brought up to
// date only when something reads or writes them (state in / out,
// dissipation, outputs, the other step kernels).  std_fresh / il_fresh say
// which copy holds the current state (at least one always does).
constexpr int kInterleaveGroup = kIlGroup;  // instances per workgroup of k_step_il (one group of the layout)
// faces and (cell, instance) pairs per lane in flight in k_step_il, the
// fastest measured: FP64 4 and 2, FP32 2 and 1 (256 instances at 40962
// cells: 0.548 ms per step against 0.556 with 1 and 1, 0.62 with 4 and 2)
template <typename T>
struct IlUnroll {
  static constexpr int faces = sizeof(T) == 8 ? 4 : 2, cells = sizeof(T) == 8 ? 2 : 1;
};
constexpr int kInterleaveMin = 16;    // batch size from which the automatic choice interleaves
constexpr int kInterleaveCells = 20000;  // ... on grids of at least this many cells (smaller: slower)

static bool il_wanted(const odis_ctx *ctx) {
  const bool on =
      ctx->il_opt == 1 || (ctx->il_opt == 2 && ctx->B >= kInterleaveMin && ctx->N >= kInterleaveCells);
  return on && ctx->compact && ctx->one_kernel && ctx->patch == 64 &&
         (ctx->sh_mode == ODIS_SH_OFF || ctx->sh_mode == ODIS_SH_MFMA) && step_blocks(ctx) == ctx->n_patch;
}

template <typename T>
struct IlOps {
  T *v[2], *eta[2], *ring_v, *ring_e, *ts[2];
};

template <typename T>
static IlOps<T> il_ops(odis_ctx *ctx) {
  const size_t E = ctx->E, N = ctx->N, Bp = ctx->il_Bp;
  T *p = static_cast<T *>(ctx->il_buf);
  IlOps<T> q;
  q.v[0] = p;
  q.v[1] = q.v[0] + E * Bp;
  q.eta[0] = q.v[1] + E * Bp;
  q.eta[1] = q.eta[0] + N * Bp;
  q.ring_v = q.eta[1] + N * Bp;
  q.ring_e = q.ring_v + 3 * E * Bp;
  q.ts[0] = q.ring_e + 3 * N * Bp;
  q.ts[1] = q.ts[0] + 2 * E * Bp;
  return q;
}

template <typename T>
static int il_alloc(odis_ctx *ctx) {
  if (ctx->il_buf) return ODIS_OK;
  const int Bp = (ctx->B + kInterleaveGroup - 1) / kInterleaveGroup * kInterleaveGroup;
  const size_t bytes = (size_t(9) * ctx->E + size_t(5) * ctx->N) * Bp * sizeof(T);
  void *p = nullptr;
  CK(hipMalloc(&p, bytes));
  ctx->allocs.push_back(p);              // freed with the context either way
  // padding instances read zeros; cleared on the context stream, so ordered
  // before the layout conversion that fills the live instances
  const size_t words = bytes / 4;
  hipLaunchKernelGGL(k_zero_words, dim3(unsigned(std::min<size_t>((words + kBlock - 1) / kBlock, 65535))),
                     dim3(kBlock), 0, ctx->stream, static_cast<unsigned int *>(p), words);
  CK(hipGetLastError());
  ctx->il_buf = p;                       // in use only once set up
  ctx->il_Bp = Bp;
  return ODIS_OK;
}

// Operand records for the interleaved step: used when on and the tide has
// at most kRecFields forcing fields.
static bool il_rec_wanted(const odis_ctx *ctx) { return ctx->il_rec == 1 && ctx->nfields <= kRecFields; }

template <typename T>
static int il_records(odis_ctx *ctx, Ops<T> &o) {
  if (!il_rec_wanted(ctx) || (o.face_rec && o.cell_rec)) return ODIS_OK;
  const int E = ctx->E, N = ctx->N;
  int rc;
  if ((rc = dev_alloc(ctx, &o.face_rec, size_t(E) * kRecSlots))) return rc;
  if ((rc = dev_alloc(ctx, &o.cell_rec, size_t(N) * kRecSlots))) return rc;
  hipLaunchKernelGGL(k_pack_face_rec<T>, dim3((E + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, E, N,
                     ctx->nfields, o.g_idx, o.g_coef, o.fcor, o.farea, o.fdist, o.flen, o.cellS, o.face_rec);
  hipLaunchKernelGGL(k_pack_cell_rec<T>, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, N, o.info,
                     o.cyc_loc, o.kite, o.area, o.cell_rec);
  CK(hipGetLastError());
  return ODIS_OK;
}

// the current state between the layouts (v, eta and ts of parity cur, both rings)
template <typename T>
static int il_convert(odis_ctx *ctx, Ops<T> &o, bool to_il) {
  const IlOps<T> q = il_ops<T>(ctx);
  const int c = ctx->cur, B = ctx->B;
  const long long E = ctx->E, N = ctx->N;
  // (instance-major array [B][n], its interleaved copy, n)
  struct {
    T *stdp, *ilp;
    long long n;
  } arr[5] = {{o.v[c], q.v[c], E}, {o.eta[c], q.eta[c], N}, {o.ring_v, q.ring_v, 3 * E},
              {o.ring_e, q.ring_e, 3 * N}, {o.ts[c], q.ts[c], 2 * E}};
  for (const auto &x : arr) {
    if (!x.stdp) continue;  // no partial sums without the compact form
    const dim3 g(unsigned((x.n + kBlock - 1) / kBlock), B);
    if (to_il)
      hipLaunchKernelGGL((k_interleave<T, true>), g, dim3(kBlock), 0, ctx->stream, x.stdp, x.ilp, x.n);
    else
      hipLaunchKernelGGL((k_interleave<T, false>), g, dim3(kBlock), 0, ctx->stream, x.stdp, x.ilp, x.n);
  }
  CK(hipGetLastError());
  return ODIS_OK;
}

// bring the instance-major arrays up to date (before anything uses them)
template <typename T>
static int il_out(odis_ctx *ctx, Ops<T> &o) {
  if (ctx->std_fresh) return ODIS_OK;
  const int rc = il_convert(ctx, o, false);
  if (rc) return rc;
  ctx->std_fresh = 1;
  return ODIS_OK;
}

static int il_out_any(odis_ctx *ctx) {
  return ctx->precision == ODIS_FP64 ? il_out(ctx, ctx->d) : il_out(ctx, ctx->f);
}

// step arguments on the interleaved copy
template <typename T>
static StepArgs<T> step_args_il(odis_ctx *ctx, Ops<T> &o) {
  StepArgs<T> a = step_args(ctx, o);
  const IlOps<T> q = il_ops<T>(ctx);
  const int c = ctx->cur;
  // the strides stay the elements per instance (sidx<true> groups the instances)
  a.f.v_cur = q.v[c];
  a.f.v_nxt = q.v[c ^ 1];
  a.f.ring = q.ring_v;
  a.f.eta = ctx->sh_mode != ODIS_SH_OFF ? static_cast<const T *>(ctx->eta_eff) : q.eta[c];  // eta_eff interleaved
  a.f.ts = q.ts[c];
  a.c.v_new = q.v[c ^ 1];
  a.c.eta = q.eta[c];
  a.c.eta_nxt = q.eta[c ^ 1];
  a.c.ring = q.ring_e;
  a.c.ts = q.ts[c ^ 1];
  if (il_rec_wanted(ctx)) {
    a.face_rec = o.face_rec;
    a.cell_rec = o.cell_rec;
  }
  const int n_grp = (ctx->B + kInterleaveGroup - 1) / kInterleaveGroup;
  a.il_band = ctx->il_order == 1 ? n_grp : std::max(1, std::min(ctx->il_order, n_grp));
  return a;
}

template <typename T, int P, int LPC, bool MULTI>
static void launch_step_kernel(odis_ctx *ctx, const StepArgs<T> &a, const double *table, int s, hipEvent_t *ev) {
  const dim3 gp(a.n_blocks, ctx->B), bp(LPC * P);
  if (s == 0) {
    if (ev)
      hipExtLaunchKernelGGL((k_step<T, false, P, LPC, MULTI>), gp, bp, 0, ctx->stream, ev[0], ev[1], 0, a,
                            ctx->energy_quad, table, s);
    else
      hipLaunchKernelGGL((k_step<T, false, P, LPC, MULTI>), gp, bp, 0, ctx->stream, a, ctx->energy_quad, table, s);
  } else {
    if (ev)
      hipExtLaunchKernelGGL((k_step<T, true, P, LPC, MULTI>), gp, bp, 0, ctx->stream, ev[0], ev[1], 0, a,
                            ctx->energy_quad, table, s);
    else
      hipLaunchKernelGGL((k_step<T, true, P, LPC, MULTI>), gp, bp, 0, ctx->stream, a, ctx->energy_quad, table, s);
  }
}

template <typename T, int P>
static void launch_step_lanes(odis_ctx *ctx, const StepArgs<T> &a, const double *table, int s, hipEvent_t *ev) {
  const bool multi = a.n_blocks < a.n_patch;
  if (ctx->step_lanes == 2) {
    if (multi)
      launch_step_kernel<T, P, 2, true>(ctx, a, table, s, ev);
    else
      launch_step_kernel<T, P, 2, false>(ctx, a, table, s, ev);
  } else {
    if (multi)
      launch_step_kernel<T, P, 4, true>(ctx, a, table, s, ev);
    else
      launch_step_kernel<T, P, 4, false>(ctx, a, table, s, ev);
  }
}

template <typename T, bool C>
static void launch_pair(odis_ctx *ctx, const FaceArgs<T> &fa, const CellArgs<T> &ca, const double *table, int s,
                        hipEvent_t *evf, hipEvent_t *evc) {
  const dim3 gf(ctx->nb_face, ctx->B), gc(ctx->nb_cell, ctx->B), blk(kBlock);
  if (evf)
    hipExtLaunchKernelGGL((k_face<T, false, C>), gf, blk, 0, ctx->stream, evf[0], evf[1], 0, fa, ctx->energy_quad,
                          table, s);
  else
    hipLaunchKernelGGL((k_face<T, false, C>), gf, blk, 0, ctx->stream, fa, ctx->energy_quad, table, s);
  if (evc)
    hipExtLaunchKernelGGL((k_cell<T, C, false>), gc, blk, 0, ctx->stream, evc[0], evc[1], 0, ca, table, s);
  else
    hipLaunchKernelGGL((k_cell<T, C, false>), gc, blk, 0, ctx->stream, ca, table, s);
}

// Enqueue the steps [n0, n0 + count) of a staged coefficient table
// (n0 = ctx->step): per step the SH projection kernels (when a basis is
// set), the face kernel and the cell kernel.  evs: dispatch-stamped events on
// the kernels of every ctx->timing-th step (null: none).
// One interleaved step launch: FU faces and CU (cell, instance) pairs per
// lane in flight; the first launch of a chunk has no pending dissipation.
// Dynamic LDS of k_step_il (its layout is in the kernel's comment).
template <typename T>
static size_t il_lds_bytes(int max_loc) {
  constexpr int IB = kInterleaveGroup, NT = 256;
  return sizeof(double) * (NT + NT / 64) + size_t(max_loc) * (IB * sizeof(T) + sizeof(int));
}

template <typename T, int FU, int CU, bool GF, bool REC>
static void launch_step_il_order(odis_ctx *ctx, const StepArgs<T> &a, const double *table, int s, hipEvent_t *ev) {
  constexpr int IB = kInterleaveGroup;
  const int n_grp = (ctx->B + IB - 1) / IB;
  const unsigned n_band = unsigned((n_grp + a.il_band - 1) / a.il_band);
  const dim3 gp = GF ? dim3(unsigned(a.n_blocks) * unsigned(a.il_band) * n_band, 1) : dim3(a.n_blocks, n_grp),
             bp(256);
  const size_t lds = il_lds_bytes<T>(a.max_loc);
  if (s == 0) {
    if (ev)
      hipExtLaunchKernelGGL((k_step_il<T, false, 64, IB, FU, CU, GF, REC>), gp, bp, lds, ctx->stream, ev[0], ev[1],
                            0, a, ctx->energy_quad, table, s);
    else
      hipLaunchKernelGGL((k_step_il<T, false, 64, IB, FU, CU, GF, REC>), gp, bp, lds, ctx->stream, a,
                         ctx->energy_quad, table, s);
  } else {
    if (ev)
      hipExtLaunchKernelGGL((k_step_il<T, true, 64, IB, FU, CU, GF, REC>), gp, bp, lds, ctx->stream, ev[0], ev[1],
                            0, a, ctx->energy_quad, table, s);
    else
      hipLaunchKernelGGL((k_step_il<T, true, 64, IB, FU, CU, GF, REC>), gp, bp, lds, ctx->stream, a,
                         ctx->energy_quad, table, s);
  }
}

template <typename T, int FU, int CU>
static void launch_step_il(odis_ctx *ctx, const StepArgs<T> &a, const double *table, int s, hipEvent_t *ev) {
  const bool rec = a.face_rec != nullptr && a.cell_rec != nullptr;
  if (ctx->il_order) {
    if (rec)
      launch_step_il_order<T, FU, CU, true, true>(ctx, a, table, s, ev);
    else
      launch_step_il_order<T, FU, CU, true, false>(ctx, a, table, s, ev);
  } else {
    if (rec)
      launch_step_il_order<T, FU, CU, false, true>(ctx, a, table, s, ev);
    else
      launch_step_il_order<T, FU, CU, false, false>(ctx, a, table, s, ev);
  }
}

template <typename T>
static int launch_chunk(odis_ctx *ctx, Ops<T> &o, const double *table, int count, ChunkEvents *evs,
                        size_t *ev_used) {
  const long long n0 = ctx->step;
  if (evs) {
    const size_t sampled = size_t(count / ctx->timing + 2);
    const int rc = reserve_events(ctx, *ev_used + 10 * sampled);
    if (rc) return rc;
  }
  const bool one = ctx->compact && ctx->one_kernel;
  const bool il = il_wanted(ctx);
  if (one && step_blocks(ctx) > ctx->part_cap)
    return fail(ctx, ODIS_E_INVALID, "more workgroups than dissipation partials");
  for (int s = 0; s < count; s++) {
    const bool stamp = evs && (n0 + s) % ctx->timing == 0;
    if (ctx->sh_mode != ODIS_SH_OFF) {
      hipEvent_t *evsh = nullptr;
      if (stamp) {
        evs->sh.push_back(int(*ev_used));
        evsh = &ctx->ev[*ev_used];
        *ev_used += 6;
      }
      int rc;
      if (il) {  // the batched products on the interleaved eta
        const T *eta = il_ops<T>(ctx).eta[ctx->cur];
        rc = sh_apply<T>(ctx, eta, ctx->N, ctx->B, ctx->sh_D, eta, static_cast<T *>(ctx->eta_eff), ctx->sh_c, true,
                         true, evsh, true);
      } else {
        rc = sh_effective_height(ctx, o, evsh);
      }
      if (rc) return rc;
    }
    if (one && il) {
      StepArgs<T> a = step_args_il(ctx, o);
      hipEvent_t *ev = nullptr;
      if (stamp) {
        evs->mom.push_back(int(*ev_used));
        ev = &ctx->ev[*ev_used];
        *ev_used += 2;
      }
      // faces and (cell, instance) pairs per lane in flight (ODIS_OPT_IL_UNROLL)
      switch (ctx->il_unroll) {
        case 1: launch_step_il<T, 1, 1>(ctx, a, table, s, ev); break;
        case 2: launch_step_il<T, 2, 1>(ctx, a, table, s, ev); break;
        case 3: launch_step_il<T, 4, 2>(ctx, a, table, s, ev); break;
        default: launch_step_il<T, IlUnroll<T>::faces, IlUnroll<T>::cells>(ctx, a, table, s, ev); break;
      }
    } else if (one) {
      // the face phase reads eta_eff (face_args) when a basis is set
      StepArgs<T> a = step_args(ctx, o);
      hipEvent_t *ev = nullptr;
      if (stamp) {
        evs->mom.push_back(int(*ev_used));
        ev = &ctx->ev[*ev_used];
        *ev_used += 2;
      }
      if (ctx->patch == 64)
        launch_step_lanes<T, 64>(ctx, a, table, s, ev);
      else if (ctx->patch == 256)
        launch_step_lanes<T, 256>(ctx, a, table, s, ev);
      else
        launch_step_lanes<T, 128>(ctx, a, table, s, ev);
    } else {
      FaceArgs<T> fa = face_args(ctx, o);
      CellArgs<T> ca = cell_args(ctx, o);
      hipEvent_t *evf = nullptr, *evc = nullptr;
      if (stamp) {
        evs->mom.push_back(int(*ev_used));
        evs->cont.push_back(int(*ev_used + 2));
        evf = &ctx->ev[*ev_used];
        evc = &ctx->ev[*ev_used + 2];
        *ev_used += 4;
      }
      if (ctx->compact)
        launch_pair<T, true>(ctx, fa, ca, table, s, evf, evc);
      else
        launch_pair<T, false>(ctx, fa, ca, table, s, evf, evc);
    }
    CK(hipGetLastError());
    ctx->cur ^= 1;
  }
  if (one) {
    hipLaunchKernelGGL(k_series, dim3(ctx->B), dim3(kBlock), 0, ctx->stream, ctx->partials, step_blocks(ctx),
                       ctx->dconst, ctx->series, ctx->nonfinite, ctx->cap, table, count - 1);
    CK(hipGetLastError());
  }
  return ODIS_OK;
}

// Graph of graph_steps steps reading coefficient buffer k, one per entry
// parity of v (graph_steps is even: a replay restores the parity).
template <typename T>
static int get_graph(odis_ctx *ctx, Ops<T> &o, int k, hipGraphExec_t *out) {
  hipGraphExec_t &g = ctx->graphs[k][ctx->cur];
  if (g) {
    *out = g;
    return ODIS_OK;
  }
  const int cur0 = ctx->cur;
  hipGraph_t graph = nullptr;
  CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_chunk(ctx, o, ctx->coef_dev[k], ctx->graph_steps, nullptr, nullptr);
  hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  ctx->cur = cur0;
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, ODIS_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    g = nullptr;
    return fail(ctx, ODIS_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  *out = g;
  return ODIS_OK;
}

static void drop_graphs(odis_ctx *ctx) {
  for (int k = 0; k < 2; k++)
    for (int c = 0; c < 2; c++)
      if (ctx->graphs[k][c]) {
        (void)hipGraphExecDestroy(ctx->graphs[k][c]);
        ctx->graphs[k][c] = nullptr;
      }
}

// Average kernel times of the sampled steps of the last odis_step call (its
// events, read once: after the step's work, never inside it).
static int read_kernel_times(odis_ctx *ctx) {
  if (ctx->pending.empty()) return ODIS_OK;
  const std::vector<ChunkEvents> sampled = std::move(ctx->pending);
  ctx->pending.clear();
  {
    CK(hipStreamSynchronize(ctx->stream));
    double tm = 0, tc = 0, ts = 0;
    long long ns = 0;
    for (const ChunkEvents &c : sampled)
      for (size_t j = 0; j < c.mom.size(); j++) {
        float a = 0, b = 0, w = 0;
        CK(hipEventElapsedTime(&a, ctx->ev[c.mom[j]], ctx->ev[c.mom[j] + 1]));
        w = a;
        if (j < c.cont.size()) {  // face + cell kernel steps
          CK(hipEventElapsedTime(&b, ctx->ev[c.cont[j]], ctx->ev[c.cont[j] + 1]));
          CK(hipEventElapsedTime(&w, ctx->ev[c.mom[j]], ctx->ev[c.cont[j] + 1]));
        }
        tm += a;
        tc += b;
        ts += w;
        ns++;
      }
    ctx->t_face = ns ? tm / ns : 0.0;
    ctx->t_cell = ns ? tc / ns : 0.0;
    ctx->t_step = ns ? ts / ns : 0.0;
    double tsh[3] = {0, 0, 0};
    long long nsh = 0;
    for (const ChunkEvents &c : sampled)
      for (int base : c.sh) {
        for (int q = 0; q < 3; q++) {
          float a = 0;
          CK(hipEventElapsedTime(&a, ctx->ev[base + 2 * q], ctx->ev[base + 2 * q + 1]));
          tsh[q] += a;
        }
        nsh++;
      }
    for (int q = 0; q < 3; q++) ctx->t_sh[q] = nsh ? tsh[q] / nsh : 0.0;
    if (nsh) {  // a sampled SH step runs from the analysis to the end of the step kernel or cell kernel
      ctx->t_step = 0.0;
      long long k = 0;
      for (const ChunkEvents &c : sampled)
        for (size_t j = 0; j < c.sh.size() && j < c.mom.size(); j++) {
          float w = 0;
          const int end = j < c.cont.size() ? c.cont[j] + 1 : c.mom[j] + 1;
          CK(hipEventElapsedTime(&w, ctx->ev[c.sh[j]], ctx->ev[end]));
          ctx->t_step += w;
          k++;
        }
      ctx->t_step = k ? ctx->t_step / k : 0.0;
    }
  }
  return ODIS_OK;
}

template <typename T>
static int step_t(odis_ctx *ctx, Ops<T> &o, int n_steps) {
  // Timing mode (ctx->timing = T > 0): the kernels of every T-th step get
  // dispatch-stamped start/end events (chunks with a sampled step are not
  // replayed from graphs).  The events are read by read_kernel_times when
  // the times are asked for, or before this call reuses them.
  int rc = read_kernel_times(ctx);
  if (rc) return rc;
  std::vector<ChunkEvents> sampled;
  size_t ev_used = 0;
  int done = 0;
  const bool il = il_wanted(ctx);
  if (il) {
    if (!ctx->il_fresh) {  // the instance-major arrays are current: copy them in
      if ((rc = ensure_ts(ctx, o)) || (rc = il_alloc<T>(ctx)) || (rc = il_convert(ctx, o, true))) return rc;
      ctx->il_fresh = 1;
    }
    if ((rc = il_records(ctx, o))) return rc;
    ctx->std_fresh = 0;
  } else {
    if ((rc = il_out(ctx, o)) || (rc = ensure_ts(ctx, o))) return rc;
    ctx->il_fresh = 0;
  }
  while (done < n_steps) {
    if (ctx->stop_flag) {
      ctx->stop_flag = 0;
      return fail(ctx, ODIS_E_STOPPED, "stop requested");
    }
    const int left = n_steps - done;
    const bool graph_fits = ctx->use_graphs && left >= ctx->graph_steps;
    const int count = graph_fits ? ctx->graph_steps : std::min(ctx->chunk, left);
    // does [step, step + count) hold a multiple of the timing stride?
    const bool sample = ctx->timing > 0 && ((ctx->step + ctx->timing - 1) / ctx->timing) * ctx->timing <
                                               ctx->step + count;
    const bool graph = graph_fits && !sample;
    const double *table = nullptr;
    if ((rc = stage_coefficients(ctx, ctx->step, count, &table))) return rc;
    if (graph) {
      hipGraphExec_t g = nullptr;
      if ((rc = get_graph(ctx, o, ctx->coef_last, &g))) return rc;
      CK(hipGraphLaunch(g, ctx->stream));
      if (count & 1) ctx->cur ^= 1;
    } else {
      ChunkEvents *evs = nullptr;
      if (sample) {
        sampled.emplace_back();
        evs = &sampled.back();
      }
      if ((rc = launch_chunk(ctx, o, table, count, evs, &ev_used))) return rc;
    }
    ctx->step += count;
    done += count;
  }
  // the events are read when the kernel times are asked for, outside the
  // caller's timed steps
  ctx->pending = std::move(sampled);
  return ODIS_OK;
}

static int step_impl(odis_ctx *ctx, int n_steps, int flags) {
  if (!ctx) return ODIS_E_INVALID;
  if (n_steps < 0) return fail(ctx, ODIS_E_INVALID, "n_steps < 0");
  int rc = ctx->precision == ODIS_FP64 ? step_t(ctx, ctx->d, n_steps) : step_t(ctx, ctx->f, n_steps);
  if (rc) return rc;
  if (flags & ODIS_SYNC) CK(hipStreamSynchronize(ctx->stream));
  return ODIS_OK;
}

extern "C" int odis_step(odis_ctx *ctx, int n_steps, int flags) {
  return guarded(ctx, [&] { return step_impl(ctx, n_steps, flags); });
}

extern "C" int odis_sync(odis_ctx *ctx) {
  if (!ctx) return ODIS_E_INVALID;
  CK(hipStreamSynchronize(ctx->stream));
  return ODIS_OK;
}

template <typename T>
static int diss_t(odis_ctx *ctx, Ops<T> &o) {
  int rc = il_out(ctx, o);
  if (rc || (rc = ensure_ts(ctx, o))) return rc;
  FaceArgs<T> fa = face_args(ctx, o);
  const dim3 gf(ctx->nb_face, ctx->B), blk(kBlock);
  if (ctx->compact)
    hipLaunchKernelGGL((k_face<T, true, true>), gf, blk, 0, ctx->stream, fa, ctx->energy_quad,
                       static_cast<const double *>(nullptr), 0);
  else
    hipLaunchKernelGGL((k_face<T, true, false>), gf, blk, 0, ctx->stream, fa, ctx->energy_quad,
                       static_cast<const double *>(nullptr), 0);
  hipLaunchKernelGGL(k_diss_final, dim3(ctx->B), blk, 0, ctx->stream, ctx->partials, ctx->nb_face, ctx->dconst,
                     ctx->scratch_out);
  CK(hipGetLastError());
  return ODIS_OK;
}

static int get_dissipation_impl(odis_ctx *ctx, double *per_inst) {
  if (!ctx || !per_inst) return ODIS_E_INVALID;
  int rc = ctx->precision == ODIS_FP64 ? diss_t(ctx, ctx->d) : diss_t(ctx, ctx->f);
  if (rc) return rc;
  CK(hipMemcpyAsync(per_inst, ctx->scratch_out, sizeof(double) * ctx->B, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < ctx->B; b++)
    if (!std::isfinite(per_inst[b])) return fail(ctx, ODIS_E_NONFINITE, "non-finite dissipation");
  return ODIS_OK;
}

extern "C" int odis_get_dissipation(odis_ctx *ctx, double *per_inst) {
  return guarded(ctx, [&] { return get_dissipation_impl(ctx, per_inst); });
}

extern "C" int odis_series_capacity(const odis_ctx *ctx) { return ctx ? ctx->cap : ODIS_E_INVALID; }

static int get_dissipation_series_impl(odis_ctx *ctx, int inst, long long first, int count, double *out) {
  if (!ctx || !out || count < 0) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  CK(hipStreamSynchronize(ctx->stream));
  // series[k] exists for max(series_lo, step - cap) <= k < step
  if (first < ctx->series_lo || first + count > ctx->step || first < ctx->step - ctx->cap)
    return fail(ctx, ODIS_E_RANGE, "series range not available");
  std::vector<double> all(ctx->cap);
  CK(hipMemcpy(all.data(), ctx->series + size_t(inst) * ctx->cap, sizeof(double) * ctx->cap,
               hipMemcpyDeviceToHost));
  for (int k = 0; k < count; k++) out[k] = all[size_t((first + k) % ctx->cap)];
  return ODIS_OK;
}

extern "C" int odis_get_dissipation_series(odis_ctx *ctx, int inst, long long first, int count, double *out) {
  return guarded(ctx, [&] { return get_dissipation_series_impl(ctx, inst, first, count, out); });
}

// ---------------------------------------------------------------------------
// orbit statistics and retirement (batched runs to converged dissipation)
// ---------------------------------------------------------------------------

// Workgroup j: instance sel[j] whose orbit of ns[j] steps ended at `step`:
// out[3j] = mean of its dissipation series over the orbit, out[3j + 1],
// out[3j + 2] = ||v - v_snap|| / ||v||, ||eta - eta_snap|| / ||eta|| in FP64
// against its snapshot, which then takes the current state.  Fixed-order
// sums (strided per lane, then the wave tree): deterministic.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_orbit_stats(const int *__restrict__ sel, const int *__restrict__ ns,
                                                        long long step, int cap, const double *__restrict__ series,
                                                        const T *__restrict__ v, const T *__restrict__ eta,
                                                        long long E, long long N, double *__restrict__ snap,
                                                        double *__restrict__ out) {
  __shared__ double sh[kBlock / 64];
  const int j = blockIdx.x, b = sel[j], m = ns[j];
  double acc = 0.0;
  for (int k = threadIdx.x; k < m; k += kBlock) acc += series[size_t(b) * cap + size_t((step - m + k) % cap)];
  const double tot = block_sum<kBlock / 64>(acc, sh);
  double dv = 0.0, nv = 0.0, de = 0.0, ne = 0.0;
  double *sp = snap + size_t(b) * size_t(E + N);
  const T *vb = v + size_t(b) * E, *eb = eta + size_t(b) * N;
  for (long long i = threadIdx.x; i < E; i += kBlock) {
    const double x = double(vb[i]), d = x - sp[i];
    dv += d * d;
    nv += x * x;
    sp[i] = x;
  }
  for (long long i = threadIdx.x; i < N; i += kBlock) {
    const double x = double(eb[i]), d = x - sp[E + i];
    de += d * d;
    ne += x * x;
    sp[E + i] = x;
  }
  __syncthreads();
  const double sdv = block_sum<kBlock / 64>(dv, sh);
  __syncthreads();
  const double snv = block_sum<kBlock / 64>(nv, sh);
  __syncthreads();
  const double sde = block_sum<kBlock / 64>(de, sh);
  __syncthreads();
  const double sne = block_sum<kBlock / 64>(ne, sh);
  if (threadIdx.x == 0) {
    out[3 * j] = tot / double(m);
    out[3 * j + 1] = snv > 0.0 ? sqrt(sdv / snv) : (sdv > 0.0 ? HUGE_VAL : 0.0);
    out[3 * j + 2] = sne > 0.0 ? sqrt(sde / sne) : (sde > 0.0 ? HUGE_VAL : 0.0);
  }
}

template <typename T>
static int orbit_stats_t(odis_ctx *ctx, Ops<T> &o, int n, const int *inst, const int *ns, double *out) {
  const int B = ctx->B;
  const long long E = ctx->E, N = ctx->N;
  int rc = il_out(ctx, o);
  if (rc) return rc;
  if (!ctx->snap) {
    if ((rc = dev_alloc(ctx, &ctx->snap, size_t(B) * size_t(E + N)))) return rc;
    ctx->snap_valid.assign(B, 0);
  }
  if (!ctx->orb_sel) {  // sized for the batch at the first call; retirement only shrinks it
    if ((rc = dev_alloc(ctx, &ctx->orb_sel, size_t(B))) || (rc = dev_alloc(ctx, &ctx->orb_ns, size_t(B))) ||
        (rc = dev_alloc(ctx, &ctx->orb_out, size_t(3) * B)))
      return rc;
  }
  CK(hipMemcpyAsync(ctx->orb_sel, inst, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  CK(hipMemcpyAsync(ctx->orb_ns, ns, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL((k_orbit_stats<T>), dim3(n), dim3(kBlock), 0, ctx->stream, ctx->orb_sel, ctx->orb_ns,
                     ctx->step, ctx->cap, ctx->series, o.v[ctx->cur], o.eta[ctx->cur], E, N, ctx->snap,
                     ctx->orb_out);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(out, ctx->orb_out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  for (int j = 0; j < n; j++) {
    if (!ctx->snap_valid[inst[j]]) out[3 * j + 1] = out[3 * j + 2] = std::nan("");
    ctx->snap_valid[inst[j]] = 1;
  }
  return ODIS_OK;
}

static int orbit_stats_impl(odis_ctx *ctx, int n, const int *inst, const int *ns, double *out) {
  if (!ctx || n < 0 || (n > 0 && (!inst || !ns || !out))) return ODIS_E_INVALID;
  if (n > ctx->B) return fail(ctx, ODIS_E_RANGE, "more instances than the batch holds");
  std::vector<char> seen(ctx->B, 0);
  for (int j = 0; j < n; j++) {
    if (inst[j] < 0 || inst[j] >= ctx->B || seen[inst[j]]) return fail(ctx, ODIS_E_RANGE, "bad or repeated instance");
    seen[inst[j]] = 1;
    if (ns[j] < 1 || ns[j] > ctx->cap || ctx->step - ns[j] < ctx->series_lo)
      return fail(ctx, ODIS_E_RANGE, "orbit longer than the series held since the last set_state");
  }
  if (n == 0) return ODIS_OK;
  return ctx->precision == ODIS_FP64 ? orbit_stats_t(ctx, ctx->d, n, inst, ns, out)
                                     : orbit_stats_t(ctx, ctx->f, n, inst, ns, out);
}

extern "C" int odis_orbit_stats(odis_ctx *ctx, int n, const int *inst, const int *steps_per_orbit, double *out) {
  return guarded(ctx, [&] { return orbit_stats_impl(ctx, n, inst, steps_per_orbit, out); });
}

// Move instance keep[j] to slot j of a per-instance device array of `per`
// elements per instance (keep strictly increasing, so a source slot is never
// overwritten before it is read).
template <typename U>
static int compact_slots(odis_ctx *ctx, U *base, size_t per, const int *keep, int n_keep) {
  if (!base) return ODIS_OK;
  for (int j = 0; j < n_keep; j++)
    if (keep[j] != j)
      CK(hipMemcpyAsync(base + size_t(j) * per, base + size_t(keep[j]) * per, per * sizeof(U),
                        hipMemcpyDeviceToDevice, ctx->stream));
  return ODIS_OK;
}

template <typename T>
static int retire_t(odis_ctx *ctx, Ops<T> &o, const int *keep, int n_keep) {
  const size_t E = ctx->E, N = ctx->N;
  int rc = il_out(ctx, o);  // the instance-major arrays hold the state
  if (rc) return rc;
  for (int k = 0; k < 2 && !rc; k++) {
    if (!rc) rc = compact_slots(ctx, o.v[k], E, keep, n_keep);
    if (!rc) rc = compact_slots(ctx, o.eta[k], N, keep, n_keep);
    if (!rc) rc = compact_slots(ctx, o.ts[k], 2 * E, keep, n_keep);
  }
  if (!rc) rc = compact_slots(ctx, o.ring_v, 3 * E, keep, n_keep);
  if (!rc) rc = compact_slots(ctx, o.ring_e, 3 * N, keep, n_keep);
  if (!rc) rc = compact_slots(ctx, ctx->inst, size_t(8), keep, n_keep);
  if (!rc) rc = compact_slots(ctx, ctx->dconst, size_t(1), keep, n_keep);
  if (!rc) rc = compact_slots(ctx, ctx->series, size_t(ctx->cap), keep, n_keep);
  if (!rc) rc = compact_slots(ctx, ctx->nonfinite, size_t(1), keep, n_keep);
  if (!rc) rc = compact_slots(ctx, ctx->snap, E + N, keep, n_keep);
  if (!rc && ctx->sh_mode != ODIS_SH_OFF) rc = compact_slots(ctx, ctx->sh_D, size_t(ctx->sh_K), keep, n_keep);
  if (rc) return rc;
  auto pick = [&](auto &vec) {
    for (int j = 0; j < n_keep; j++) vec[j] = vec[keep[j]];
    vec.resize(n_keep);
  };
  pick(ctx->params);
  pick(ctx->tides);
  pick(ctx->init_hist_v);
  pick(ctx->init_hist_e);
  if (!ctx->snap_valid.empty()) pick(ctx->snap_valid);
  ctx->B = n_keep;
  drop_graphs(ctx);
  if (ctx->il_buf) {  // re-laid out from the compacted arrays at the next interleaved step
    ctx->il_Bp = (n_keep + kInterleaveGroup - 1) / kInterleaveGroup * kInterleaveGroup;
    const size_t words = (size_t(9) * E + size_t(5) * N) * size_t(ctx->il_Bp) * sizeof(T) / 4;
    hipLaunchKernelGGL(k_zero_words, dim3(unsigned(std::min<size_t>((words + kBlock - 1) / kBlock, 65535))),
                       dim3(kBlock), 0, ctx->stream, static_cast<unsigned int *>(ctx->il_buf), words);
    CK(hipGetLastError());
  }
  ctx->il_fresh = 0;
  ctx->std_fresh = 1;
  CK(hipStreamSynchronize(ctx->stream));
  return ODIS_OK;
}

static int retire_impl(odis_ctx *ctx, const int *keep, int n_keep) {
  if (!ctx || (n_keep > 0 && !keep)) return ODIS_E_INVALID;
  if (n_keep < 1 || n_keep > ctx->B) return fail(ctx, ODIS_E_RANGE, "n_keep outside [1, B]");
  for (int j = 0; j < n_keep; j++)
    if (keep[j] < 0 || keep[j] >= ctx->B || (j > 0 && keep[j] <= keep[j - 1]))
      return fail(ctx, ODIS_E_INVALID, "keep must be strictly increasing instance indices");
  if (n_keep == ctx->B) return ODIS_OK;
  CK(hipStreamSynchronize(ctx->stream));
  return ctx->precision == ODIS_FP64 ? retire_t(ctx, ctx->d, keep, n_keep) : retire_t(ctx, ctx->f, keep, n_keep);
}

extern "C" int odis_retire(odis_ctx *ctx, const int *keep, int n_keep) {
  return guarded(ctx, [&] { return retire_impl(ctx, keep, n_keep); });
}

template <typename T>
static int vel_t(odis_ctx *ctx, Ops<T> &o, int b, const double *normal, double *uv) {
  const int E = ctx->E;
  const int rc = il_out(ctx, o);
  if (rc) return rc;
  std::vector<double> nd(size_t(2) * E);
  for (int r = 0; r < E; r++) {
    nd[2 * r] = normal[2 * ctx->face_inv[r]];
    nd[2 * r + 1] = normal[2 * ctx->face_inv[r] + 1];
  }
  double *dn = nullptr, *duv = nullptr;
  CK(hipMalloc(&dn, sizeof(double) * 2 * E));
  hipError_t e = hipMalloc(&duv, sizeof(double) * 2 * E);
  if (e != hipSuccess) {
    (void)hipFree(dn);
    return fail(ctx, ODIS_E_NOMEM, "hipMalloc");
  }
  std::vector<double> h(size_t(2) * E);
  e = hipMemcpy(dn, nd.data(), sizeof(double) * 2 * E, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL((k_velocity_en<T>), dim3((E + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, E,
                       o.tri_idx, o.tri_coef, o.v[ctx->cur] + size_t(b) * E, dn, duv);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(h.data(), duv, sizeof(double) * 2 * E, hipMemcpyDeviceToHost);
  (void)hipFree(dn);
  (void)hipFree(duv);
  if (e != hipSuccess) return fail(ctx, ODIS_E_HIP, hipGetErrorString(e));
  for (int r = 0; r < E; r++) {
    uv[2 * ctx->face_inv[r]] = h[2 * r];
    uv[2 * ctx->face_inv[r] + 1] = h[2 * r + 1];
  }
  return ODIS_OK;
}

static int get_velocity_en_impl(odis_ctx *ctx, int inst, const double *face_normal, double *uv) {
  if (!ctx || !face_normal || !uv) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  CK(hipStreamSynchronize(ctx->stream));
  return ctx->precision == ODIS_FP64 ? vel_t(ctx, ctx->d, inst, face_normal, uv)
                                     : vel_t(ctx, ctx->f, inst, face_normal, uv);
}

extern "C" int odis_get_velocity_en(odis_ctx *ctx, int inst, const double *face_normal, double *uv) {
  return guarded(ctx, [&] { return get_velocity_en_impl(ctx, inst, face_normal, uv); });
}

extern "C" int odis_request_stop(odis_ctx *ctx) {
  if (!ctx) return ODIS_E_INVALID;
  ctx->stop_flag = 1;
  return ODIS_OK;
}

extern "C" int odis_enable_timing(odis_ctx *ctx, int stride) {
  if (!ctx || stride < 0) return ODIS_E_INVALID;
  ctx->timing = stride;
  // the events of a whole chunk's sampled steps, created here rather than
  // inside the steps they time
  if (stride > 0) return guarded(ctx, [&] { return reserve_events(ctx, 10 * size_t(ctx->chunk / stride + 2)); });
  return ODIS_OK;
}

static int set_option_impl(odis_ctx *ctx, int option, int value) {
  if (!ctx) return ODIS_E_INVALID;
  if (option == ODIS_OPT_GRAPHS) {
    ctx->use_graphs = value ? 1 : 0;
    return ODIS_OK;
  }
  if (option == ODIS_OPT_COMPACT) {
    const int c = value ? 1 : 0;
    if (c && !ctx->compact_ok) return fail(ctx, ODIS_E_UNSUPPORTED, "no compact TRiSK form: " + ctx->compact_why);
    if (c != ctx->compact) {
      CK(hipStreamSynchronize(ctx->stream));
      drop_graphs(ctx);
      ctx->compact = c;
      ctx->ts_valid = 0;
    }
    return ODIS_OK;
  }
  if (option == ODIS_OPT_PATCH_CELLS) {
    if (value != 64 && value != 128 && value != 256)
      return fail(ctx, ODIS_E_INVALID, "patch cells must be 64, 128 or 256");
    if (!ctx->one_kernel_ok) return fail(ctx, ODIS_E_UNSUPPORTED, "no one-kernel step: " + ctx->one_kernel_why);
    if (value == ctx->patch) return ODIS_OK;
    CompactHost h;
    h.cyc = ctx->cyc_host;
    const std::string why = build_patches(ctx->N, ctx->E, ctx->gi_host, value, h);
    if (!why.empty()) return fail(ctx, ODIS_E_UNSUPPORTED, why);
    CK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    const int rc = ctx->precision == ODIS_FP64 ? upload_patches(ctx, ctx->d, h) : upload_patches(ctx, ctx->f, h);
    if (rc) return rc;
    ctx->patch = value;
    ctx->n_patch = (ctx->N + value - 1) / value;
    ctx->patch_max_loc = h.max_loc;
    return ODIS_OK;
  }
  if (option == ODIS_OPT_SH_CELLS) {
    if (value != 0 && value != 4 && value != 8 && value != 16)
      return fail(ctx, ODIS_E_INVALID, "SH cells per lane: 0 (automatic), 4, 8 or 16");
    CK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->sh_cpt = value;
    return ODIS_OK;
  }
  if (option == ODIS_OPT_INTERLEAVE) {
    if (value < 0 || value > 2) return fail(ctx, ODIS_E_INVALID, "interleave: 0 (off), 1 (on) or 2 (automatic)");
    CK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->il_opt = value;
    return ODIS_OK;
  }
  if (option == ODIS_OPT_IL_ORDER) {
    if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
      return fail(ctx, ODIS_E_INVALID, "interleaved order: 0, 1 (all groups), or bands of 2, 4 or 8 groups");
    CK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->il_order = value;
    return ODIS_OK;
  }
  if (option == ODIS_OPT_IL_RECORDS) {
    if (value < 0 || value > 1) return fail(ctx, ODIS_E_INVALID, "interleaved records: 0 or 1");
    CK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->il_rec = value;
    return ODIS_OK;
  }
  if (option == ODIS_OPT_IL_UNROLL) {
    if (value < 0 || value > 3) return fail(ctx, ODIS_E_INVALID, "interleaved unroll: 0 (default), 1, 2 or 3");
    CK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->il_unroll = value;
    return ODIS_OK;
  }
  if (option == ODIS_OPT_STEP_LANES) {
    if (value != 2 && value != 4) return fail(ctx, ODIS_E_INVALID, "step lanes must be 2 or 4");
    CK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->step_lanes = value;
    return ODIS_OK;
  }
  if (option == ODIS_OPT_STEP_BLOCKS) {
    if (value < 0) return fail(ctx, ODIS_E_INVALID, "step blocks < 0");
    CK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->step_blocks = value;
    return ODIS_OK;
  }
  if (option == ODIS_OPT_ONE_KERNEL) {
    const int c = value ? 1 : 0;
    if (c && !ctx->one_kernel_ok)
      return fail(ctx, ODIS_E_UNSUPPORTED, "no one-kernel step: " + ctx->one_kernel_why);
    if (c != ctx->one_kernel) {
      CK(hipStreamSynchronize(ctx->stream));
      drop_graphs(ctx);
      ctx->one_kernel = c;
    }
    return ODIS_OK;
  }
  return fail(ctx, ODIS_E_INVALID, "unknown option");
}

extern "C" int odis_set_option(odis_ctx *ctx, int option, int value) {
  return guarded(ctx, [&] { return set_option_impl(ctx, option, value); });
}

extern "C" int odis_get_option(odis_ctx *ctx, int option, int *value) {
  if (!ctx || !value) return ODIS_E_INVALID;
  if (option == ODIS_OPT_GRAPHS) *value = ctx->use_graphs;
  else if (option == ODIS_OPT_COMPACT) *value = ctx->compact;
  else if (option == ODIS_OPT_ONE_KERNEL) *value = ctx->one_kernel;
  else if (option == ODIS_OPT_PATCH_CELLS) *value = ctx->patch;
  else if (option == ODIS_OPT_STEP_BLOCKS) *value = step_blocks(ctx);
  else if (option == ODIS_OPT_STEP_LANES) *value = ctx->step_lanes;
  else if (option == ODIS_OPT_SH_CELLS) *value = ctx->sh_cpt;
  else if (option == ODIS_OPT_INTERLEAVE) *value = ctx->il_opt;
  else if (option == ODIS_OPT_INTERLEAVE_ACTIVE) *value = il_wanted(ctx) ? 1 : 0;
  else if (option == ODIS_OPT_IL_UNROLL) *value = ctx->il_unroll;
  else if (option == ODIS_OPT_IL_ORDER) *value = ctx->il_order;
  else if (option == ODIS_OPT_IL_RECORDS) *value = ctx->il_rec;
  else return fail(ctx, ODIS_E_INVALID, "unknown option");
  return ODIS_OK;
}

extern "C" int odis_get_kernel_times(odis_ctx *ctx, double *ms, int n) {
  if (!ctx || !ms || n < 0) return ODIS_E_INVALID;
  const int rc = guarded(ctx, [&] { return read_kernel_times(ctx); });
  if (rc) return rc;
  const double v[ODIS_T_COUNT] = {ctx->t_face, ctx->t_cell, ctx->t_step, ctx->t_sh[0], ctx->t_sh[1], ctx->t_sh[2]};
  for (int q = 0; q < n && q < ODIS_T_COUNT; q++) ms[q] = v[q];
  return ODIS_OK;
}

extern "C" int odis_get_timing(odis_ctx *ctx, double *face_ms, double *cell_ms, double *step_ms) {
  if (!ctx) return ODIS_E_INVALID;
  const int rc = guarded(ctx, [&] { return read_kernel_times(ctx); });
  if (rc) return rc;
  if (face_ms) *face_ms = ctx->t_face;
  if (cell_ms) *cell_ms = ctx->t_cell;
  if (step_ms) *step_ms = ctx->t_step;
  return ODIS_OK;
}

// ---------------------------------------------------------------------------
// geodesic -> lat-lon interpolation (src/interpolation.cpp:208-287)
// ---------------------------------------------------------------------------

static int interp_ll_set_impl(odis_ctx *ctx, int n_rows, int n_cols, const int *row_ptr, const int *col,
                              const double *val) {
  if (!ctx || !row_ptr || !col || !val || n_rows < 1) return ODIS_E_INVALID;
  if (n_cols != 3 * ctx->N) return fail(ctx, ODIS_E_INVALID, "interpolation matrix must have 3N columns");
  if (row_ptr[0] != 0) return fail(ctx, ODIS_E_INVALID, "row_ptr[0] must be 0");
  for (int r = 0; r < n_rows; r++)
    if (row_ptr[r + 1] < row_ptr[r]) return fail(ctx, ODIS_E_INVALID, "row_ptr not monotone");
  const int nnz = row_ptr[n_rows];
  std::vector<int> rp(n_rows + 1, 0), cols;
  std::vector<double> vals;
  cols.reserve(nnz / 3 + 1);
  vals.reserve(nnz / 3 + 1);
  for (int r = 0; r < n_rows; r++) {
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; k++) {
      const int c = col[k];
      if (c < 0 || c >= n_cols) return fail(ctx, ODIS_E_INVALID, "interpolation column out of range");
      if (c % 3 != 0) continue;  // gradient entries of x are zero in the reference
      cols.push_back(ctx->cell_perm[c / 3]);
      vals.push_back(val[k]);
    }
    rp[r + 1] = int(cols.size());
  }
  CK(hipStreamSynchronize(ctx->stream));
  for (void *p : ctx->ll_allocs) (void)hipFree(p);
  ctx->ll_allocs.clear();
  ctx->ll_rows = 0;  // no matrix until the new one is in place
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    CK(hipMalloc(dst, std::max<size_t>(bytes, 8)));
    ctx->ll_allocs.push_back(*dst);
    if (src) CK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return ODIS_OK;
  };
  int rc;
  if ((rc = up(reinterpret_cast<void **>(&ctx->ll_rp), rp.data(), rp.size() * sizeof(int)))) return rc;
  if ((rc = up(reinterpret_cast<void **>(&ctx->ll_col), cols.data(), cols.size() * sizeof(int)))) return rc;
  if ((rc = up(reinterpret_cast<void **>(&ctx->ll_val), vals.data(), vals.size() * sizeof(double)))) return rc;
  if ((rc = up(reinterpret_cast<void **>(&ctx->ll_out), nullptr, size_t(n_rows) * sizeof(double)))) return rc;
  ctx->ll_rows = n_rows;
  return ODIS_OK;
}

extern "C" int odis_interp_ll_set(odis_ctx *ctx, int n_rows, int n_cols, const int *row_ptr, const int *col,
                                  const double *val) {
  return guarded(ctx, [&] { return interp_ll_set_impl(ctx, n_rows, n_cols, row_ptr, col, val); });
}

static int rbf_set_impl(odis_ctx *ctx, int n_rows, int n_cols, const int *row_ptr, const int *col,
                        const double *val) {
  if (!ctx || !row_ptr || !col || !val) return ODIS_E_INVALID;
  const int N = ctx->N, E = ctx->E;
  if (n_rows != 3 * N || n_cols != E) return fail(ctx, ODIS_E_INVALID, "RBF operator must be 3N x E");
  if (row_ptr[0] != 0) return fail(ctx, ODIS_E_INVALID, "row_ptr[0] must be 0");
  std::vector<int> idx(size_t(6) * N, -1);
  std::vector<double> w(size_t(18) * N, 0.0);
  for (int i = 0; i < N; i++) {
    const int r = ctx->cell_perm[i];
    const int k0 = row_ptr[3 * i], len = row_ptr[3 * i + 1] - k0;
    if (len < 0 || len > 6) return fail(ctx, ODIS_E_INVALID, "RBF row must hold 0..6 faces");
    for (int c = 0; c < 3; c++) {
      const int kc = row_ptr[3 * i + c];
      if (row_ptr[3 * i + c + 1] - kc != len)
        return fail(ctx, ODIS_E_INVALID, "RBF rows of a cell must share their faces");
      for (int j = 0; j < len; j++) {
        const int f = col[kc + j];
        if (f < 0 || f >= E) return fail(ctx, ODIS_E_INVALID, "RBF column out of range");
        if (f != col[k0 + j]) return fail(ctx, ODIS_E_INVALID, "RBF rows of a cell must share their faces");
        idx[size_t(j) * N + r] = ctx->face_perm[f];
        w[(size_t(c) * 6 + j) * N + r] = val[kc + j];
      }
    }
  }
  CK(hipStreamSynchronize(ctx->stream));
  for (void *p : ctx->rbf_allocs) (void)hipFree(p);
  ctx->rbf_allocs.clear();
  ctx->rbf_idx = nullptr;
  ctx->rbf_w = ctx->rbf_out = nullptr;
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    void *p = nullptr;
    CK(hipMalloc(&p, bytes));
    ctx->rbf_allocs.push_back(p);
    if (src) CK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    *dst = p;
    return ODIS_OK;
  };
  // the operator is in use (rbf_idx set) only once all of it is in place
  int *d_idx = nullptr;
  double *d_w = nullptr, *d_out = nullptr;
  int rc;
  if ((rc = up(reinterpret_cast<void **>(&d_idx), idx.data(), idx.size() * sizeof(int)))) return rc;
  if ((rc = up(reinterpret_cast<void **>(&d_w), w.data(), w.size() * sizeof(double)))) return rc;
  if ((rc = up(reinterpret_cast<void **>(&d_out), nullptr, size_t(3) * N * sizeof(double)))) return rc;
  ctx->rbf_w = d_w;
  ctx->rbf_out = d_out;
  ctx->rbf_idx = d_idx;
  return ODIS_OK;
}

extern "C" int odis_rbf_set(odis_ctx *ctx, int n_rows, int n_cols, const int *row_ptr, const int *col,
                            const double *val) {
  return guarded(ctx, [&] { return rbf_set_impl(ctx, n_rows, n_cols, row_ptr, col, val); });
}

static int velocity_xyz_impl(odis_ctx *ctx, int inst, double *xyz) {
  if (!ctx || !xyz) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  if (!ctx->rbf_idx) return fail(ctx, ODIS_E_INVALID, "no RBF operator set");
  const int rc = il_out_any(ctx);
  if (rc) return rc;
  const int N = ctx->N;
  const dim3 grid((N + kBlock - 1) / kBlock), blk(kBlock);
  if (ctx->precision == ODIS_FP64)
    hipLaunchKernelGGL((k_velocity_xyz<double>), grid, blk, 0, ctx->stream, N, ctx->rbf_idx, ctx->rbf_w,
                       ctx->d.v[ctx->cur] + size_t(inst) * ctx->E, ctx->rbf_out);
  else
    hipLaunchKernelGGL((k_velocity_xyz<float>), grid, blk, 0, ctx->stream, N, ctx->rbf_idx, ctx->rbf_w,
                       ctx->f.v[ctx->cur] + size_t(inst) * ctx->E, ctx->rbf_out);
  CK(hipGetLastError());
  std::vector<double> h(size_t(3) * N);
  CK(hipMemcpyAsync(h.data(), ctx->rbf_out, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  for (int r = 0; r < N; r++) {
    const int i = ctx->cell_inv[r];
    for (int c = 0; c < 3; c++) xyz[3 * i + c] = h[size_t(c) * N + r];
  }
  return ODIS_OK;
}

extern "C" int odis_velocity_xyz(odis_ctx *ctx, int inst, double *xyz) {
  return guarded(ctx, [&] { return velocity_xyz_impl(ctx, inst, xyz); });
}

static int interp_ll_impl(odis_ctx *ctx, int inst, double *ll) {
  if (!ctx || !ll) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  if (!ctx->ll_rows) return fail(ctx, ODIS_E_INVALID, "no interpolation matrix set");
  const int rc = il_out_any(ctx);
  if (rc) return rc;
  const dim3 grid((ctx->ll_rows + kBlock / 64 - 1) / (kBlock / 64)), blk(kBlock);
  if (ctx->precision == ODIS_FP64)
    hipLaunchKernelGGL((k_interp_ll<double>), grid, blk, 0, ctx->stream, ctx->ll_rows, ctx->ll_rp, ctx->ll_col,
                       ctx->ll_val, ctx->d.eta[ctx->cur] + size_t(inst) * ctx->N, ctx->ll_out);
  else
    hipLaunchKernelGGL((k_interp_ll<float>), grid, blk, 0, ctx->stream, ctx->ll_rows, ctx->ll_rp, ctx->ll_col,
                       ctx->ll_val, ctx->f.eta[ctx->cur] + size_t(inst) * ctx->N, ctx->ll_out);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(ll, ctx->ll_out, sizeof(double) * ctx->ll_rows, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return ODIS_OK;
}

extern "C" int odis_interp_ll(odis_ctx *ctx, int inst, double *ll) {
  return guarded(ctx, [&] { return interp_ll_impl(ctx, inst, ll); });
}

// ---------------------------------------------------------------------------
// spherical-harmonic basis set-up and stand-alone projection
// ---------------------------------------------------------------------------

static uint16_t to_bf16(float f) {
  // round to nearest even
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return uint16_t(u >> 16);
}
static float from_bf16(uint16_t h) {
  const uint32_t u = uint32_t(h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// In-place inverse of a symmetric positive definite matrix (Cholesky).
static bool spd_inverse(std::vector<double> &A, int n) {
  std::vector<double> L(size_t(n) * n, 0.0);
  for (int j = 0; j < n; j++) {
    double d = A[size_t(j) * n + j];
    for (int k = 0; k < j; k++) d -= L[size_t(j) * n + k] * L[size_t(j) * n + k];
    if (!(d > 0)) return false;
    const double ljj = std::sqrt(d);
    L[size_t(j) * n + j] = ljj;
    for (int i = j + 1; i < n; i++) {
      double s = A[size_t(i) * n + j];
      for (int k = 0; k < j; k++) s -= L[size_t(i) * n + k] * L[size_t(j) * n + k];
      L[size_t(i) * n + j] = s / ljj;
    }
  }
  // inverse of L (lower), then A^-1 = L^-T L^-1
  std::vector<double> Li(size_t(n) * n, 0.0);
  for (int j = 0; j < n; j++) {
    Li[size_t(j) * n + j] = 1.0 / L[size_t(j) * n + j];
    for (int i = j + 1; i < n; i++) {
      double s = 0.0;
      for (int k = j; k < i; k++) s -= L[size_t(i) * n + k] * Li[size_t(k) * n + j];
      Li[size_t(i) * n + j] = s / L[size_t(i) * n + i];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= i; j++) {
      double s = 0.0;
      for (int k = i; k < n; k++) s += Li[size_t(k) * n + i] * Li[size_t(k) * n + j];
      A[size_t(i) * n + j] = s;
      A[size_t(j) * n + i] = s;
    }
  return true;
}

static int sh_set_basis_impl(odis_ctx *ctx, int l_max, const double *beta, int mode) {
  if (!ctx) return ODIS_E_INVALID;
  if (mode != ODIS_SH_OFF && mode != ODIS_SH_FP64 && mode != ODIS_SH_SPLIT_BF16 && mode != ODIS_SH_RECUR &&
      mode != ODIS_SH_MFMA)
    return fail(ctx, ODIS_E_INVALID, "bad sh mode");
  CK(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  for (void *p : ctx->sh_allocs) (void)hipFree(p);
  ctx->sh_allocs.clear();
  ctx->sh_mode = ODIS_SH_OFF;
  ctx->sh_L = ctx->sh_K = 0;
  ctx->sh_Yt = ctx->sh_Minv = ctx->sh_D = ctx->sh_b = ctx->sh_c = ctx->sh_tab = ctx->sh_trig = nullptr;
  ctx->sh_tabm = ctx->sh_cmaj = ctx->sh_chain = nullptr;
  ctx->sh_Yp = nullptr;
  ctx->sh_Ya = ctx->sh_Ys = ctx->sh_Xa = nullptr;
  ctx->sh_Cs = nullptr;
  ctx->sh_Kp = ctx->sh_qpr = ctx->sh_rng = 0;
  ctx->eta_eff = nullptr;
  ctx->sh_inv = nullptr;
  if (mode == ODIS_SH_OFF) return ODIS_OK;
  if (l_max < 0 || l_max > 100) return fail(ctx, ODIS_E_INVALID, "l_max out of range");
  const int K = (l_max + 1) * (l_max + 1), N = ctx->N;
  if (K > N) return fail(ctx, ODIS_E_INVALID, "(l_max+1)^2 exceeds the number of cells");
  // the synthesis keeps the K coefficients and 4 x 64 wave partials (stored
  // basis), the recurrence table (on the fly) or nothing (batched products)
  const size_t lds_words = mode == ODIS_SH_RECUR  ? size_t(K) + size_t(l_max + 1) * (l_max + 2) + 2 * (l_max + 1)
                           : mode == ODIS_SH_MFMA ? 0
                                                  : size_t(K) + kBlock;
  if (lds_words * sizeof(double) > 64 * 1024) return fail(ctx, ODIS_E_INVALID, "l_max too large for LDS");
  // basis on the host in reference order, then columns in device order
  std::vector<double> Yref;
  odis::sh_basis(l_max, ctx->cell_lat, ctx->cell_lon, Yref);
  std::vector<double> Y(size_t(K) * N);
  for (int k = 0; k < K; k++)
    for (int r = 0; r < N; r++) Y[size_t(k) * N + r] = Yref[size_t(k) * N + ctx->cell_inv[r]];
  std::vector<double>().swap(Yref);
  auto alloc = [&](void **p, size_t bytes) -> int {
    CK(hipMalloc(p, std::max<size_t>(bytes, 8)));
    ctx->sh_allocs.push_back(*p);
    return ODIS_OK;
  };
  int rc;
  if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_Yt), Y.size() * sizeof(double)))) return rc;
  CK(hipMemcpy(ctx->sh_Yt, Y.data(), Y.size() * sizeof(double), hipMemcpyHostToDevice));
  // Gram matrix on the GPU, inverse on the host
  double *dM = nullptr;
  if ((rc = alloc(reinterpret_cast<void **>(&dM), size_t(K) * K * sizeof(double)))) return rc;
  const int tiles = (K + 31) / 32;
  hipLaunchKernelGGL(k_gram, dim3(tiles, tiles), dim3(kBlock), 0, ctx->stream, K, N, ctx->sh_Yt, dM);
  CK(hipGetLastError());
  std::vector<double> M(size_t(K) * K);
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipMemcpy(M.data(), dM, M.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int r = 0; r < K; r++)
    for (int c = r + 1; c < K; c++) M[size_t(r) * K + c] = M[size_t(c) * K + r];
  if (!spd_inverse(M, K)) return fail(ctx, ODIS_E_INVALID, "basis Gram matrix is singular");
  CK(hipMemcpy(dM, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
  ctx->sh_Minv = dM;
  ctx->sh_part = kShChunks;
  if (mode == ODIS_SH_RECUR) {
    // the stored basis was only needed for the Gram matrix
    (void)hipFree(ctx->sh_Yt);
    ctx->sh_allocs.erase(std::remove(ctx->sh_allocs.begin(), ctx->sh_allocs.end(), (void *)ctx->sh_Yt),
                         ctx->sh_allocs.end());
    ctx->sh_Yt = nullptr;
    std::vector<double> a, bt, sect, f1;
    odis::sh_recurrence_tables(l_max, a, bt, sect, f1);
    std::vector<double> tab;
    tab.insert(tab.end(), a.begin(), a.end());
    tab.insert(tab.end(), bt.begin(), bt.end());
    tab.insert(tab.end(), sect.begin(), sect.end());
    tab.insert(tab.end(), f1.begin(), f1.end());
    std::vector<double> trig(size_t(4) * N);
    for (int r = 0; r < N; r++) {
      const int i = ctx->cell_inv[r];
      trig[r] = std::sin(ctx->cell_lat[i]);
      trig[size_t(N) + r] = std::cos(ctx->cell_lat[i]);
      trig[size_t(2) * N + r] = std::cos(ctx->cell_lon[i]);
      trig[size_t(3) * N + r] = std::sin(ctx->cell_lon[i]);
    }
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_tab), tab.size() * sizeof(double)))) return rc;
    CK(hipMemcpy(ctx->sh_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    // m-major: sect, f1, then per m the pairs (a, b) for l = m + 2 .. L
    std::vector<double> tabm(sect.begin(), sect.end());
    tabm.insert(tabm.end(), f1.begin(), f1.end());
    for (int m = 0; m <= l_max; m++)
      for (int l = m + 2; l <= l_max; l++) {
        tabm.push_back(a[odis::sh_lm_index(l, m)]);
        tabm.push_back(bt[odis::sh_lm_index(l, m)]);
      }
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_tabm), tabm.size() * sizeof(double)))) return rc;
    CK(hipMemcpy(ctx->sh_tabm, tabm.data(), tabm.size() * sizeof(double), hipMemcpyHostToDevice));
    // rescaled chains: q_l = mu q_{l-1} + beta_l q_{l-2}, P_lm = S_l q_l
    std::vector<double> chain;
    for (int m = 0; m <= l_max; m++) {
      double S = 1.0, a_prev = f1[m];
      chain.push_back(1.0);  // k = 0: q = P_mm through q_{-2}
      chain.push_back(1.0);
      if (m < l_max) {
        S = f1[m];
        chain.push_back(0.0);  // k = 1: q = mu P_mm
        chain.push_back(S);
      }
      for (int l = m + 2; l <= l_max; l++) {
        const double al = a[odis::sh_lm_index(l, m)], bl = bt[odis::sh_lm_index(l, m)];
        S *= al;
        chain.push_back(-bl / (al * a_prev));
        chain.push_back(S);
        a_prev = al;
      }
    }
    double C = 1.0;
    for (int m = 0; m <= l_max; m++) {
      if (m > 0) C *= sect[m];
      chain.push_back(C);
    }
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_chain), chain.size() * sizeof(double)))) return rc;
    CK(hipMemcpy(ctx->sh_chain, chain.data(), chain.size() * sizeof(double), hipMemcpyHostToDevice));
    // coefficients: pairs (runtime degrees) or the 4-double degree table of
    // the chain synthesis, with beta and the padding set here
    const size_t nlm = size_t(l_max + 1) * (l_max + 2) / 2;
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_cmaj), size_t(ctx->B) * 4 * nlm * sizeof(double))))
      return rc;
    {
      std::vector<double> ct(size_t(ctx->B) * 4 * nlm, 0.0);
      for (int b = 0; b < ctx->B; b++)
        for (size_t q = 0; q < nlm; q++) ct[(size_t(b) * nlm + q) * 4] = chain[2 * q];
      CK(hipMemcpy(ctx->sh_cmaj, ct.data(), ct.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_trig), trig.size() * sizeof(double)))) return rc;
    CK(hipMemcpy(ctx->sh_trig, trig.data(), trig.size() * sizeof(double), hipMemcpyHostToDevice));
    // the most partials: 4 cells per lane
    ctx->sh_part = std::max(((N + 64 * 4 - 1) / (64 * 4) + kShWaves - 1) / kShWaves,
                            (N + 64 * kPairWaves * 4 - 1) / (64 * kPairWaves * 4));
  }
  if (mode == ODIS_SH_MFMA) {
    // the FP64 basis was only needed for the Gram matrix
    (void)hipFree(ctx->sh_Yt);
    ctx->sh_allocs.erase(std::remove(ctx->sh_allocs.begin(), ctx->sh_allocs.end(), (void *)ctx->sh_Yt),
                         ctx->sh_allocs.end());
    ctx->sh_Yt = nullptr;
    const int Kp = (K + 31) / 32 * 32, KT = Kp / 16, KB = Kp / 32;
    const int NB32 = (N + 31) / 32, NT16 = (N + 15) / 16;
    auto yv = [&](int k, int c) { return k < K && c < N ? Y[size_t(k) * N + c] : 0.0; };
    auto put = [](uint16_t *f, int i, double y) {
      const uint16_t hi = to_bf16(float(y));
      f[i] = hi;
      f[8 + i] = to_bf16(float(y - double(from_bf16(hi))));
    };
    std::vector<uint16_t> F(size_t(KT) * NB32 * 64 * 16);
    for (int h = 0; h < KT; h++)
      for (int q = 0; q < NB32; q++)
        for (int l = 0; l < 64; l++) {
          uint16_t *f = F.data() + ((size_t(h) * NB32 + q) * 64 + l) * 16;
          for (int i = 0; i < 8; i++) put(f, i, yv(16 * h + (l & 15), 32 * q + 8 * (l >> 4) + i));
        }
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_Ya), F.size() * sizeof(uint16_t)))) return rc;
    CK(hipMemcpy(ctx->sh_Ya, F.data(), F.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    F.assign(size_t(NT16) * KB * 64 * 16, 0);
    for (int t = 0; t < NT16; t++)
      for (int kb = 0; kb < KB; kb++)
        for (int l = 0; l < 64; l++) {
          uint16_t *f = F.data() + ((size_t(t) * KB + kb) * 64 + l) * 16;
          for (int i = 0; i < 8; i++) put(f, i, yv(32 * kb + 8 * (l >> 4) + i, 16 * t + (l & 15)));
        }
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_Ys), F.size() * sizeof(uint16_t)))) return rc;
    CK(hipMemcpy(ctx->sh_Ys, F.data(), F.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    std::vector<uint16_t>().swap(F);
    // split coefficients, their padding zero for good
    const size_t cs_bytes = size_t(ctx->B) * KB * 4 * 16 * sizeof(uint16_t);
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_Cs), cs_bytes))) return rc;
    CK(hipMemset(ctx->sh_Cs, 0, cs_bytes));
    // split eta of every instance tile
    const size_t xa_bytes = size_t((ctx->B + 15) / 16) * NB32 * 64 * 32;
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_Xa), xa_bytes))) return rc;
    // analysis ranges: at most 128 of at least 4 blocks of 32 cells
    ctx->sh_Kp = Kp;
    ctx->sh_qpr = std::max(4, (NB32 + 127) / 128);
    ctx->sh_rng = (NB32 + ctx->sh_qpr - 1) / ctx->sh_qpr;
    ctx->sh_part = ctx->sh_rng;
  }
  if (mode == ODIS_SH_SPLIT_BF16) {
    std::vector<unsigned int> P(Y.size());
    for (size_t q = 0; q < Y.size(); q++) {
      const float f = float(Y[q]);
      const uint16_t hi = to_bf16(f);
      const uint16_t lo = to_bf16(float(Y[q] - double(from_bf16(hi))));
      P[q] = (unsigned(hi) << 16) | unsigned(lo);
    }
    if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_Yp), P.size() * sizeof(unsigned int)))) return rc;
    CK(hipMemcpy(ctx->sh_Yp, P.data(), P.size() * sizeof(unsigned int), hipMemcpyHostToDevice));
  }
  // per-instance degree factors (beta_l - 1 for 2 <= l <= L, 0 otherwise)
  std::vector<double> D(size_t(ctx->B) * K, 0.0);
  if (beta)
    for (int b = 0; b < ctx->B; b++)
      for (int l = 2; l <= l_max; l++)
        for (int m = -l; m <= l; m++) D[size_t(b) * K + l * l + l + m] = beta[size_t(b) * (l_max + 1) + l] - 1.0;
  if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_D), D.size() * sizeof(double)))) return rc;
  CK(hipMemcpy(ctx->sh_D, D.data(), D.size() * sizeof(double), hipMemcpyHostToDevice));
  // analysis partials [B][sh_part][K], then their sums [B][K]
  if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_b),
                  size_t(ctx->B) * (ctx->sh_part + 1) * K * sizeof(double))))
    return rc;
  if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_c), size_t(ctx->B) * K * sizeof(double)))) return rc;
  const size_t tsz = ctx->precision == ODIS_FP64 ? sizeof(double) : sizeof(float);
  // eta_eff of every instance, padded to whole groups of kIlGroup for a batch held interleaved
  const size_t eff_bytes = size_t((ctx->B + kIlGroup - 1) / kIlGroup * kIlGroup) * N * tsz;
  if ((rc = alloc(&ctx->eta_eff, eff_bytes))) return rc;
  CK(hipMemset(ctx->eta_eff, 0, eff_bytes));  // padding instances read zeros
  // device-order permutation, for the projection entry point that allocates nothing
  if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_inv), size_t(N) * sizeof(int)))) return rc;
  CK(hipMemcpy(ctx->sh_inv, ctx->cell_inv.data(), size_t(N) * sizeof(int), hipMemcpyHostToDevice));
  ctx->sh_L = l_max;
  ctx->sh_K = K;
  ctx->sh_mode = mode;
  return ODIS_OK;
}

extern "C" int odis_sh_set_basis(odis_ctx *ctx, int l_max, const double *beta, int mode) {
  return guarded(ctx, [&] { return sh_set_basis_impl(ctx, l_max, beta, mode); });
}

// field (reference order) -> LSQ coefficients with the context's basis
static int sh_project_impl(odis_ctx *ctx, const double *field, double *coeffs) {
  if (!ctx || !field || !coeffs) return ODIS_E_INVALID;
  if (ctx->sh_mode == ODIS_SH_OFF) return fail(ctx, ODIS_E_INVALID, "no spherical-harmonic basis set");
  const int K = ctx->sh_K, N = ctx->N;
  std::vector<double> x(N);
  for (int r = 0; r < N; r++) x[r] = field[ctx->cell_inv[r]];
  double *dx = nullptr, *dc = nullptr;
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipMalloc(&dx, sizeof(double) * N));
  hipError_t e = hipMalloc(&dc, sizeof(double) * K);
  if (e == hipSuccess) e = hipMemcpy(dx, x.data(), sizeof(double) * N, hipMemcpyHostToDevice);
  int rc = ODIS_OK;
  if (e == hipSuccess)
    rc = sh_apply<double>(ctx, dx, N, 1, nullptr, nullptr, nullptr, dc, true, false, nullptr);
  if (e == hipSuccess && !rc) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && !rc) e = hipMemcpy(coeffs, dc, sizeof(double) * K, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  (void)hipFree(dc);
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, ODIS_E_HIP, hipGetErrorString(e));
  return ODIS_OK;
}

extern "C" int odis_sh_project(odis_ctx *ctx, const double *field, double *coeffs) {
  return guarded(ctx, [&] { return sh_project_impl(ctx, field, coeffs); });
}

// ---------------------------------------------------------------------------
// Projection that allocates nothing.  odis_sh_project allocates the field in
// device order and the coefficients per call; odis_sh_project_noalloc keeps
// both in `scratch`, a file-scope device array of kNoAllocScratch doubles
// (8 MiB per GPU, held until the process ends): [0, N) the field as given,
// [N, 2N) the field in device order, [2N, 2N + K) the coefficients.  It
// rejects 2N + K > kNoAllocScratch.  Host code never uses the array's name
// as an address: each device's address comes from hipGetSymbolAddress, once
// per device (that device current) at the entry point's first call on it.
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU (two host threads, two contexts) race; each call writes every word it
// reads before reading it, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
static constexpr size_t kNoAllocScratch = size_t(1) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratch];

static constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// the current device's address of `scratch` (nullptr and an error text if
// it cannot be had)
static double *scratch_for_current_device(odis_ctx *ctx) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    fail(ctx, ODIS_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    fail(ctx, ODIS_E_RANGE, "device index beyond the scratch table");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      fail(ctx, ODIS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  return g_scratch_ptr[dev];
}

// y[r] = x[inv[r]]: a cell field in reference order to device order
__global__ __launch_bounds__(kBlock) void k_to_device_order(int n, const int *__restrict__ inv,
                                                            const double *__restrict__ x, double *__restrict__ y) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) y[r] = x[inv[r]];
}

static int sh_project_noalloc_impl(odis_ctx *ctx, const double *field, double *coeffs) {
  if (!ctx || !field || !coeffs) return ODIS_E_INVALID;
  if (ctx->sh_mode == ODIS_SH_OFF) return fail(ctx, ODIS_E_INVALID, "no spherical-harmonic basis set");
  const int K = ctx->sh_K, N = ctx->N;
  if (2 * size_t(N) + size_t(K) > kNoAllocScratch)
    return fail(ctx, ODIS_E_RANGE,
                "odis_sh_project_noalloc: 2 N + K = " + std::to_string(2 * size_t(N) + size_t(K)) +
                    " doubles exceed its scratch of " + std::to_string(kNoAllocScratch));
  double *ptr = scratch_for_current_device(ctx);
  if (!ptr) return ODIS_E_HIP;
  double *x_ref = ptr, *x_dev = ptr + N, *c = ptr + 2 * size_t(N);
  CK(hipMemcpyAsync(x_ref, field, sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_to_device_order, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, N,
                     ctx->sh_inv, x_ref, x_dev);
  CK(hipGetLastError());
  const int rc = sh_apply<double>(ctx, x_dev, N, 1, nullptr, nullptr, nullptr, c, true, false, nullptr);
  if (rc) return rc;
  CK(hipMemcpyAsync(coeffs, c, sizeof(double) * K, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return ODIS_OK;
}

extern "C" int odis_sh_project_noalloc(odis_ctx *ctx, const double *field, double *coeffs) {
  return guarded(ctx, [&] { return sh_project_noalloc_impl(ctx, field, coeffs); });
}

static int sh_synthesize_impl(odis_ctx *ctx, const double *coeffs, double *field) {
  if (!ctx || !field || !coeffs) return ODIS_E_INVALID;
  if (ctx->sh_mode == ODIS_SH_OFF) return fail(ctx, ODIS_E_INVALID, "no spherical-harmonic basis set");
  const int K = ctx->sh_K, N = ctx->N;
  double *dy = nullptr, *dc = nullptr;
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipMalloc(&dy, sizeof(double) * N));
  hipError_t e = hipMalloc(&dc, sizeof(double) * K);
  if (e == hipSuccess) e = hipMemcpy(dc, coeffs, sizeof(double) * K, hipMemcpyHostToDevice);
  int rc = ODIS_OK;
  if (e == hipSuccess)
    rc = sh_apply<double>(ctx, static_cast<const double *>(nullptr), N, 1, nullptr, nullptr, dy, dc, false, true,
                          nullptr);
  std::vector<double> y(N);
  if (e == hipSuccess && !rc) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && !rc) e = hipMemcpy(y.data(), dy, sizeof(double) * N, hipMemcpyDeviceToHost);
  (void)hipFree(dy);
  (void)hipFree(dc);
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, ODIS_E_HIP, hipGetErrorString(e));
  for (int r = 0; r < N; r++) field[ctx->cell_inv[r]] = y[r];
  return ODIS_OK;
}

extern "C" int odis_sh_synthesize(odis_ctx *ctx, const double *coeffs, double *field) {
  return guarded(ctx, [&] { return sh_synthesize_impl(ctx, coeffs, field); });
}
