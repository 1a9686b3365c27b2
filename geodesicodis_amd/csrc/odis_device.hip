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
#include <string>
#include <vector>

#include "odis.h"
#include "odis_kernels.hpp"

namespace odis {

// ===========================================================================
// kernels
// ===========================================================================

template <typename T>
__device__ __forceinline__ double block_sum(double x, double *sh) {
  // fixed-shape tree: wave-level shuffle then LDS across the 4 waves
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = x;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  return r;
}

// Face kernel.  DISS_ONLY: evaluate the dissipation partials of the current
// state without stepping (output cadence / final state).
// The step index and forcing coefficients come from a host-filled table:
// table[0] = n0, then per step s and instance b 8 coefficients at
// table[8 + (s * B + b) * 8]; step s of the table is n = n0 + s.  Kernels
// captured in a graph therefore stay valid for every chunk of steps.
template <typename T, bool DISS_ONLY>
__global__ __launch_bounds__(kBlock) void k_face(FaceArgs<T> a, int energy_quad, const double *table, int s) {
  __shared__ double sh[4];
  const int b = blockIdx.y;
  const int tile = xcd_tile(blockIdx.x, a.nb);
  const int E = a.E;
  const int e = tile * kBlock + threadIdx.x;
  const double *ip = a.inst + b * 8;
  double local = 0.0;
  if (e < E) {
    const T *v = a.v_cur + b * a.v_stride;
    T vt = T(0);
#pragma unroll
    for (int j = 0; j < 10; ++j) vt += a.tri_coef[j * E + e] * v[a.tri_idx[j * E + e]];
    const T ve = v[e];
    const T sp = ve * ve + vt * vt;
    const T en = energy_quad ? sqrt(sp) * sp : sp;
    local = double(a.farea[e]) * double(en);
    if (!DISS_ONLY) {
      const long long n = (long long)table[0] + s;
      const double *coef = table + 8 + size_t(s) * gridDim.y * 8;
      const T g = T(ip[0]), alpha = T(ip[1]), dt = T(ip[2]), qd = T(ip[4]);
      const bool ab3 = (n >= 2) || (ip[3] != 0.0);
      const T *eta = a.eta + b * a.eta_stride;
      const T ge = a.g_coef[e] * eta[a.g_idx[e]] + a.g_coef[E + e] * eta[a.g_idx[E + e]];
      T f0 = -g * ge + a.fcor[e] * vt;
      if (a.quad_drag) f0 -= qd * sqrt(sp) * ve;
      T fu = T(0);
      for (int k = 0; k < a.nfields; ++k) fu += T(coef[b * 8 + k]) * a.fgs[k * E + e];
      const T drag = -alpha * ve + fu;
      T *ring = a.ring + b * a.ring_stride;
      const int s0 = int(n % 3), s1 = int((n + 2) % 3), s2 = int((n + 1) % 3);
      T upd;
      if (ab3)
        upd = (T(23. / 12.) * f0 + T(-16. / 12.) * ring[s1 * E + e] + T(5. / 12.) * ring[s2 * E + e]) * dt;
      else
        upd = f0 * dt;
      ring[s0 * E + e] = f0;
      a.v_nxt[b * a.v_stride + e] = (ve + upd) + dt * drag;
    }
  }
  const double part = block_sum<T>(local, sh);
  const long long par = DISS_ONLY ? 0 : (((long long)table[0] + s) & 1);
  if (threadIdx.x == 0) a.partials[(par * gridDim.y + b) * a.nb + tile] = part;
}

// Sum the face kernel's partials of instance b in a fixed order.
__device__ __forceinline__ double sum_partials(const double *p, int count, double *sh) {
  double acc = 0.0;
  for (int j = threadIdx.x; j < count; j += kBlock) acc += p[j];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = acc;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cell(CellArgs<T> a, const double *table, int s_off) {
  __shared__ double sh[4];
  const long long n = (long long)table[0] + s_off;
  const int b = blockIdx.y;
  const int tile = xcd_tile(blockIdx.x, a.nb);
  const int N = a.N;
  const int i = tile * kBlock + threadIdx.x;
  const double *ip = a.inst + b * 8;
  if (i < N) {
    const T *v = a.v_new + b * a.v_stride;
    T s = T(0);
#pragma unroll
    for (int j = 0; j < 6; ++j) s += a.div_coef[j * N + i] * v[a.div_idx[j * N + i]];
    const T h = T(ip[5]), dt = T(ip[2]);
    const bool ab3 = (n >= 2) || (ip[3] != 0.0);
    const T f0 = h * s;
    T *ring = a.ring + b * a.ring_stride;
    const int s0 = int(n % 3), s1 = int((n + 2) % 3), s2 = int((n + 1) % 3);
    T upd;
    if (ab3)
      upd = (T(23. / 12.) * f0 + T(-16. / 12.) * ring[s1 * N + i] + T(5. / 12.) * ring[s2 * N + i]) * dt;
    else
      upd = f0 * dt;
    ring[s0 * N + i] = f0;
    T *eta = a.eta + b * a.eta_stride;
    eta[i] = eta[i] + upd;
  }
  // the first-dispatched workgroup of each instance reduces the face
  // kernel's dissipation partials of this step (complete: previous kernel)
  if (blockIdx.x != 0) return;
  const double tot = sum_partials(a.partials + ((n & 1) * gridDim.y + b) * size_t(a.nb_face), a.nb_face, sh);
  if (threadIdx.x == 0) {
    const double d = tot * a.dconst[b];
    a.series[size_t(b) * a.cap + size_t(n % a.cap)] = d;
    if (!isfinite(d)) a.nonfinite[b] = 1;
  }
}

// Sum dissipation partials produced by k_face<DISS_ONLY>.
__global__ __launch_bounds__(kBlock) void k_diss_final(const double *partials, int nb, const double *dconst,
                                                       double *out) {
  __shared__ double sh[4];
  const int b = blockIdx.x;
  const double t = sum_partials(partials + size_t(b) * nb, nb, sh);
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

// ---------------------------------------------------------------------------
// fused step: one kernel per step
// ---------------------------------------------------------------------------
// Step n's kernel first completes the continuity update of step n-1 for the
// two cells of its face (eta^n = eta^{n-1} + dt AB3(h Div v^n)), with the
// same arithmetic as k_cell; the face that owns a cell (one per cell) writes
// eta^n and the history slot, the others only use the value.  Then it does
// the momentum update of step n with eta^n.  The last step of a chunk is
// completed by k_cell.  eta is double-buffered (read n-1, write n).
template <typename T>
struct FusedArgs {
  FaceArgs<T> f;                // eta = eta^{n-1} buffer when PENDING, else eta^n
  int N;
  const int *div_idx;           // [6][N]
  const T *div_coef;            // [6][N]
  T *eta_nxt;                   // [B][N] eta^n (owner writes)
  T *ring_e;                    // [B][3][N]
  long long ring_e_stride;
  const unsigned char *owner;   // [E] bit 0: owns cell g_idx[e], bit 1: owns g_idx[E + e]
  double *series;
  int *nonfinite;
  int cap;
  const double *dconst;
};

template <typename T>
__device__ __forceinline__ T complete_continuity(const FusedArgs<T> &a, int c, const T *v, const T *eta, T *ringe,
                                                 long long m, bool ab3, T h, T dt, bool own, int b) {
  const int N = a.N;
  T acc = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) acc += a.div_coef[j * N + c] * v[a.div_idx[j * N + c]];
  const T f0 = h * acc;
  const int s0 = int(m % 3), s1 = int((m + 2) % 3), s2 = int((m + 1) % 3);
  T upd;
  if (ab3)
    upd = (T(23. / 12.) * f0 + T(-16. / 12.) * ringe[s1 * N + c] + T(5. / 12.) * ringe[s2 * N + c]) * dt;
  else
    upd = f0 * dt;
  const T val = eta[c] + upd;
  if (own) {
    a.eta_nxt[b * a.f.eta_stride + c] = val;
    ringe[s0 * N + c] = f0;
  }
  return val;
}

template <typename T, bool PENDING>
__global__ __launch_bounds__(kBlock) void k_fused(FusedArgs<T> a, int energy_quad, const double *table, int s) {
  __shared__ double sh[4];
  __shared__ double sh2[4];
  const FaceArgs<T> &F = a.f;
  const int b = blockIdx.y;
  const int tile = xcd_tile(blockIdx.x, F.nb);
  const int E = F.E;
  const int e = tile * kBlock + threadIdx.x;
  const double *ip = F.inst + b * 8;
  const long long n = (long long)table[0] + s;
  const double *coef = table + 8 + size_t(s) * gridDim.y * 8;
  const bool load = ip[3] != 0.0;
  double local = 0.0;
  if (e < E) {
    const T *v = F.v_cur + b * F.v_stride;
    T vt = T(0);
#pragma unroll
    for (int j = 0; j < 10; ++j) vt += F.tri_coef[j * E + e] * v[F.tri_idx[j * E + e]];
    const T ve = v[e];
    const T sp = ve * ve + vt * vt;
    const T en = energy_quad ? sqrt(sp) * sp : sp;
    local = double(F.farea[e]) * double(en);
    const T g = T(ip[0]), alpha = T(ip[1]), dt = T(ip[2]), qd = T(ip[4]), h = T(ip[5]);
    const T *eta = F.eta + b * F.eta_stride;
    const int c0 = F.g_idx[e], c1 = F.g_idx[E + e];
    T eta0, eta1;
    if (PENDING) {
      const long long m = n - 1;
      const bool ab3m = (m >= 2) || load;
      T *ringe = a.ring_e + b * a.ring_e_stride;
      const unsigned int own = a.owner[e];
      eta0 = complete_continuity(a, c0, v, eta, ringe, m, ab3m, h, dt, (own & 1u) != 0, b);
      eta1 = complete_continuity(a, c1, v, eta, ringe, m, ab3m, h, dt, (own & 2u) != 0, b);
    } else {
      eta0 = eta[c0];
      eta1 = eta[c1];
    }
    const bool ab3 = (n >= 2) || load;
    const T ge = F.g_coef[e] * eta0 + F.g_coef[E + e] * eta1;
    T f0 = -g * ge + F.fcor[e] * vt;
    if (F.quad_drag) f0 -= qd * sqrt(sp) * ve;
    T fu = T(0);
    for (int k = 0; k < F.nfields; ++k) fu += T(coef[b * 8 + k]) * F.fgs[k * E + e];
    const T drag = -alpha * ve + fu;
    T *ring = F.ring + b * F.ring_stride;
    const int s0 = int(n % 3), s1 = int((n + 2) % 3), s2 = int((n + 1) % 3);
    T upd;
    if (ab3)
      upd = (T(23. / 12.) * f0 + T(-16. / 12.) * ring[s1 * E + e] + T(5. / 12.) * ring[s2 * E + e]) * dt;
    else
      upd = f0 * dt;
    ring[s0 * E + e] = f0;
    F.v_nxt[b * F.v_stride + e] = (ve + upd) + dt * drag;
  }
  const double part = block_sum<T>(local, sh);
  if (threadIdx.x == 0) F.partials[((n & 1) * gridDim.y + b) * F.nb + tile] = part;
  // dissipation of state n-1: partials of the previous kernel
  if (!PENDING || blockIdx.x != 0) return;
  const double tot = sum_partials(F.partials + (((n - 1) & 1) * gridDim.y + b) * size_t(F.nb), F.nb, sh2);
  if (threadIdx.x == 0) {
    const double d = tot * a.dconst[b];
    a.series[size_t(b) * a.cap + size_t((n - 1) % a.cap)] = d;
    if (!isfinite(d)) a.nonfinite[b] = 1;
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
// Basis Yt[k][i] (k = harmonic, i = device cell) stored basis-major so that
// the analysis b_k = sum_i Yt[k][i] x_i streams row k (one workgroup per k)
// and the synthesis y_i = sum_k c_k Yt[k][i] streams the rows with one lane
// per cell.  The split-BF16 form packs hi = bf16(Y) in the upper and
// lo = bf16(Y - hi) in the lower half of a 32-bit word.

__device__ __forceinline__ float bf16_hi(unsigned int u) { return __uint_as_float(u & 0xFFFF0000u); }
__device__ __forceinline__ float bf16_lo(unsigned int u) { return __uint_as_float(u << 16); }

template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_sh_analysis(int K, int N, const double *Yt, const unsigned int *Yp,
                                                        const T *x, long long x_stride, double *b_out) {
  __shared__ double sh[4];
  const int k = blockIdx.x, b = blockIdx.y;
  const T *xb = x + b * x_stride;
  double acc = 0.0;
  if (SPLIT) {
    float facc = 0.f;
    const unsigned int *row = Yp + size_t(k) * N;
    for (int i = threadIdx.x; i < N; i += kBlock) {
      const unsigned int u = row[i];
      const float xi = float(xb[i]);
      facc = fmaf(bf16_hi(u), xi, facc);
      facc = fmaf(bf16_lo(u), xi, facc);
    }
    acc = double(facc);
  } else {
    const double *row = Yt + size_t(k) * N;
    for (int i = threadIdx.x; i < N; i += kBlock) acc += row[i] * double(xb[i]);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) b_out[size_t(b) * K + k] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// c_k = D_k sum_j Minv[k][j] b_j : one wavefront per k
__global__ __launch_bounds__(kBlock) void k_sh_coef(int K, const double *Minv, const double *D, const double *bvec,
                                                    double *c) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int b = blockIdx.y;
  if (k >= K) return;
  const double *bb = bvec + size_t(b) * K;
  const double *row = Minv + size_t(k) * K;
  double acc = 0.0;
  for (int j = lane; j < K; j += 64) acc += row[j] * bb[j];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) c[size_t(b) * K + k] = (D ? D[size_t(b) * K + k] : 1.0) * acc;
}

// y_i = base_i + sum_k c_k Yt[k][i]
template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_sh_synth(int K, int N, const double *Yt, const unsigned int *Yp,
                                                     const double *c, const T *base, long long stride, T *out) {
  extern __shared__ double cs[];
  const int b = blockIdx.y;
  for (int k = threadIdx.x; k < K; k += kBlock) cs[k] = c[size_t(b) * K + k];
  __syncthreads();
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  double acc = 0.0;
  if (SPLIT) {
    float facc = 0.f;
    for (int k = 0; k < K; ++k) {
      const unsigned int u = Yp[size_t(k) * N + i];
      const float ck = float(cs[k]);
      facc = fmaf(bf16_hi(u), ck, facc);
      facc = fmaf(bf16_lo(u), ck, facc);
    }
    acc = double(facc);
  } else {
    for (int k = 0; k < K; ++k) acc += cs[k] * Yt[size_t(k) * N + i];
  }
  const T add = base ? base[b * stride + i] : T(0);
  out[b * stride + i] = T(double(add) + acc);
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
  T *eta_buf[2] = {nullptr, nullptr};  // canonical eta is eta_buf[ctx->ecur]
  unsigned char *owner = nullptr;
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

struct odis_ctx {
  int N = 0, E = 0, B = 0, precision = ODIS_FP64;
  int nb_face = 0, nb_cell = 0;
  int nfields = 0;
  int energy_quad = 0, quad_drag = 0;
  hipStream_t stream = nullptr;
  std::vector<int> cell_perm, cell_inv, face_perm, face_inv;  // ref -> dev, dev -> ref
  Ops<double> d;
  Ops<float> f;
  int cur = 0, ecur = 0;        // v is v[cur], eta is eta_buf[ecur] between calls
  int fused = 0;               // one kernel per step (SH coupling off)
  long long step = 0;          // next step index n, shared by all instances
  double *partials = nullptr;  // [2][B][nb_face], by step parity
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
  // graphs of graph_steps steps, one per (coefficient buffer, v parity x 2 + eta parity)
  int graph_steps = 0, use_graphs = 0;
  hipGraphExec_t graphs[2][4] = {};
  std::vector<TideDev> tides;
  std::vector<odis_params> params;
  std::vector<std::vector<double>> init_hist_v, init_hist_e;  // history as last set ([M][3])
  std::vector<void *> allocs;
  volatile int stop_flag = 0;
  int timing = 0;
  std::vector<hipEvent_t> ev;
  double t_face = 0, t_cell = 0, t_step = 0;
  // spherical-harmonic coupling
  int sh_mode = ODIS_SH_OFF, sh_L = 0, sh_K = 0;
  double *sh_Yt = nullptr, *sh_Minv = nullptr, *sh_D = nullptr, *sh_b = nullptr, *sh_c = nullptr;
  unsigned int *sh_Yp = nullptr;
  void *eta_eff = nullptr;  // [B][N] of the stencil precision
  std::vector<void *> sh_allocs;
  std::vector<double> cell_lat, cell_lon;  // reference order
  // lat-lon interpolation operator (eta columns only, device cell indices)
  int ll_rows = 0;
  int *ll_rp = nullptr, *ll_col = nullptr;
  double *ll_val = nullptr, *ll_out = nullptr;
  std::vector<void *> ll_allocs;
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
                     const std::vector<double> &dc, const std::vector<int> &di,
                     const std::vector<unsigned char> &owner) {
  int rc;
  if ((rc = upload(ctx, &o.owner, owner))) return rc;
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
  if ((rc = dev_alloc(ctx, &o.eta_buf[0], B * N))) return rc;
  if ((rc = dev_alloc(ctx, &o.eta_buf[1], B * N))) return rc;
  if ((rc = dev_alloc(ctx, &o.ring_e, B * 3 * N))) return rc;
  CK(hipMemset(o.v[0], 0, B * E * sizeof(T)));
  CK(hipMemset(o.v[1], 0, B * E * sizeof(T)));
  CK(hipMemset(o.ring_v, 0, B * 3 * E * sizeof(T)));
  CK(hipMemset(o.eta_buf[0], 0, B * N * sizeof(T)));
  CK(hipMemset(o.eta_buf[1], 0, B * N * sizeof(T)));
  CK(hipMemset(o.ring_e, 0, B * 3 * N * sizeof(T)));
  return ODIS_OK;
}

extern "C" int odis_destroy(odis_ctx *ctx) {
  if (!ctx) return ODIS_E_INVALID;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &e : ctx->ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 2; k++)
    for (int c = 0; c < 4; c++)
      if (ctx->graphs[k][c]) (void)hipGraphExecDestroy(ctx->graphs[k][c]);
  for (int k = 0; k < 2; k++) {
    if (ctx->coef_ev[k]) (void)hipEventDestroy(ctx->coef_ev[k]);
    if (ctx->coef_host[k]) (void)hipHostFree(ctx->coef_host[k]);
  }
  for (void *p : ctx->allocs) (void)hipFree(p);
  for (void *p : ctx->ll_allocs) (void)hipFree(p);
  for (void *p : ctx->sh_allocs) (void)hipFree(p);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return ODIS_OK;
}

extern "C" int odis_create(const odis_mesh_desc *m, const odis_params *params, int n_inst, int precision,
                           odis_ctx **out) {
  if (!m || !params || !out || n_inst < 1) return fail(nullptr, ODIS_E_INVALID, "null argument or n_inst < 1");
  if (precision != ODIS_FP64 && precision != ODIS_FP32) return fail(nullptr, ODIS_E_INVALID, "bad precision");
  const int N = m->n_cells, E = m->n_faces;
  if (N < 12 || E != 3 * N - 6) return fail(nullptr, ODIS_E_INVALID, "mesh sizes: need E = 3N-6");
  if (!m->face_nodes || !m->grad_coef || !m->trisk_count || !m->trisk_faces || !m->trisk_coef ||
      !m->face_lat || !m->face_area || !m->cell_faces || !m->div_coef || !m->cell_lat || !m->cell_lon)
    return fail(nullptr, ODIS_E_INVALID, "mesh descriptor has a null array");
  for (int b = 0; b < n_inst; b++) {
    const odis_params &p = params[b];
    if (p.tide == ODIS_GENERAL)
      return fail(nullptr, ODIS_E_UNSUPPORTED,
                  "GENERAL forcing needs the Legendre tables of CalcLegendreFuncs, which the reference leaves "
                  "commented out");
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

  // one owner face per cell (its lowest device face index) writes the cell's
  // continuity update in the fused step kernel
  std::vector<unsigned char> owner(E, 0);
  for (int r = 0; r < N; r++) {
    int f = di[r];
    for (int j = 1; j < 6; j++) f = std::min(f, di[size_t(j) * N + r]);
    if (gi[f] == r) {
      owner[f] |= 1u;
    } else if (gi[size_t(E) + f] == r) {
      owner[f] |= 2u;
    } else {
      delete ctx;
      return fail(nullptr, ODIS_E_INVALID, "cell_faces disagrees with face_nodes");
    }
  }
  int rc = ODIS_OK;
  hipError_t he = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (he != hipSuccess) rc = fail(ctx, ODIS_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  if (!rc) {
    if (precision == ODIS_FP64)
      rc = build_ops(ctx, ctx->d, tri_c, tri_i, gc, gi, fcor, farea, fgs, dc, di, owner);
    else
      rc = build_ops(ctx, ctx->f, tri_c, tri_i, gc, gi, fcor, farea, fgs, dc, di, owner);
  }
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
  if (!rc) rc = dev_alloc(ctx, &ctx->partials, size_t(2) * n_inst * ctx->nb_face);
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
  if (rc) {
    const std::string msg = ctx->err;
    odis_destroy(ctx);
    g_last_error = msg;
    return rc;
  }
  *out = ctx;
  return ODIS_OK;
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
  }
  if (eta) {
    std::vector<T> h(N);
    for (int r = 0; r < N; r++) h[r] = T(eta[ctx->cell_inv[r]]);
    CK(hipMemcpy(o.eta_buf[ctx->ecur] + bN, h.data(), N * sizeof(T), hipMemcpyHostToDevice));
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

extern "C" int odis_set_state(odis_ctx *ctx, int inst, const double *v, const double *dvdt, const double *eta,
                              const double *detadt, long long iter) {
  if (!ctx) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  if (iter < 0) return fail(ctx, ODIS_E_INVALID, "iter < 0");
  CK(hipStreamSynchronize(ctx->stream));
  ctx->step = iter;  // the step counter is shared by all instances of a context
  return ctx->precision == ODIS_FP64 ? put_state(ctx, ctx->d, inst, v, dvdt, eta, detadt, iter)
                                     : put_state(ctx, ctx->f, inst, v, dvdt, eta, detadt, iter);
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
    CK(hipMemcpy(h.data(), o.eta_buf[ctx->ecur] + bN, N * sizeof(T), hipMemcpyDeviceToHost));
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

extern "C" int odis_get_state(odis_ctx *ctx, int inst, double *v, double *dvdt, double *eta, double *detadt,
                              long long *iter) {
  if (!ctx) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  CK(hipStreamSynchronize(ctx->stream));
  if (iter) *iter = ctx->step;
  return ctx->precision == ODIS_FP64 ? get_state_t(ctx, ctx->d, inst, v, dvdt, eta, detadt, ctx->step)
                                     : get_state_t(ctx, ctx->f, inst, v, dvdt, eta, detadt, ctx->step);
}

// ---------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------

template <typename T>
static FaceArgs<T> face_args(odis_ctx *ctx, Ops<T> &o) {
  FaceArgs<T> a;
  a.E = ctx->E;
  a.nb = ctx->nb_face;
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
  a.eta = ctx->sh_mode != ODIS_SH_OFF ? static_cast<const T *>(ctx->eta_eff) : o.eta_buf[ctx->ecur];
  a.v_stride = ctx->E;
  a.ring_stride = 3LL * ctx->E;
  a.eta_stride = ctx->N;
  a.partials = ctx->partials;
  a.inst = ctx->inst;
  a.quad_drag = ctx->quad_drag;
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
  a.eta = o.eta_buf[ctx->ecur];
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

// eta_eff = eta + Yt^T D Minv Yt eta for every instance (three kernels)
template <typename T>
static int sh_effective_height(odis_ctx *ctx, Ops<T> &o) {
  const int K = ctx->sh_K, N = ctx->N;
  const dim3 blk(kBlock);
  if (ctx->sh_mode == ODIS_SH_SPLIT_BF16)
    hipLaunchKernelGGL((k_sh_analysis<T, true>), dim3(K, ctx->B), blk, 0, ctx->stream, K, N, ctx->sh_Yt,
                       ctx->sh_Yp, o.eta_buf[ctx->ecur], (long long)N, ctx->sh_b);
  else
    hipLaunchKernelGGL((k_sh_analysis<T, false>), dim3(K, ctx->B), blk, 0, ctx->stream, K, N, ctx->sh_Yt,
                       ctx->sh_Yp, o.eta_buf[ctx->ecur], (long long)N, ctx->sh_b);
  hipLaunchKernelGGL(k_sh_coef, dim3((K + 3) / 4, ctx->B), blk, 0, ctx->stream, K, ctx->sh_Minv, ctx->sh_D,
                     ctx->sh_b, ctx->sh_c);
  const size_t lds = sizeof(double) * K;
  if (ctx->sh_mode == ODIS_SH_SPLIT_BF16)
    hipLaunchKernelGGL((k_sh_synth<T, true>), dim3(ctx->nb_cell, ctx->B), blk, lds, ctx->stream, K, N, ctx->sh_Yt,
                       ctx->sh_Yp, ctx->sh_c, o.eta_buf[ctx->ecur], (long long)N, static_cast<T *>(ctx->eta_eff));
  else
    hipLaunchKernelGGL((k_sh_synth<T, false>), dim3(ctx->nb_cell, ctx->B), blk, lds, ctx->stream, K, N,
                       ctx->sh_Yt, ctx->sh_Yp, ctx->sh_c, o.eta_buf[ctx->ecur], (long long)N,
                       static_cast<T *>(ctx->eta_eff));
  CK(hipGetLastError());
  return ODIS_OK;
}

template <typename T>
static FusedArgs<T> fused_args(odis_ctx *ctx, Ops<T> &o) {
  FusedArgs<T> a;
  a.f = face_args(ctx, o);
  a.N = ctx->N;
  a.div_idx = o.div_idx;
  a.div_coef = o.div_coef;
  a.eta_nxt = o.eta_buf[ctx->ecur ^ 1];
  a.ring_e = o.ring_e;
  a.ring_e_stride = 3LL * ctx->N;
  a.owner = o.owner;
  a.series = ctx->series;
  a.nonfinite = ctx->nonfinite;
  a.cap = ctx->cap;
  a.dconst = ctx->dconst;
  return a;
}

// Event pairs (start, end) recorded for the kernels of sampled chunks.
struct ChunkEvents {
  std::vector<int> mom;   // momentum kernel of each step (k_face or k_fused)
  std::vector<int> cont;  // continuity kernels (k_cell)
};

static int reserve_events(odis_ctx *ctx, size_t n) {
  while (ctx->ev.size() < n) {
    hipEvent_t e;
    CK(hipEventCreate(&e));
    ctx->ev.push_back(e);
  }
  return ODIS_OK;
}

// Enqueue the steps [n0, n0 + count) of a staged coefficient table
// (n0 = ctx->step).  With the fused kernel a step is one launch plus one
// k_cell at the end of the chunk; with the SH coupling it is the projection
// kernels, k_face and k_cell.  evs: dispatch-stamped events on the kernels of
// every ctx->timing-th step (null: none).
template <typename T>
static int launch_chunk(odis_ctx *ctx, Ops<T> &o, const double *table, int count, ChunkEvents *evs,
                        size_t *ev_used) {
  const dim3 gf(ctx->nb_face, ctx->B), gc(ctx->nb_cell, ctx->B), blk(kBlock);
  const long long n0 = ctx->step;
  if (evs) {
    const size_t sampled = size_t(count / ctx->timing + 2);
    const int rc = reserve_events(ctx, *ev_used + 4 * sampled + 2);
    if (rc) return rc;
  }
  bool any = false;
  std::vector<int> dummy_rec;
  auto take = [&](std::vector<int> &rec, int s) -> hipEvent_t * {
    if (!evs || (s >= 0 && (n0 + s) % ctx->timing != 0) || (s < 0 && !any)) return nullptr;
    any = true;
    rec.push_back(int(*ev_used));
    hipEvent_t *p = &ctx->ev[*ev_used];
    *ev_used += 2;
    return p;
  };
  if (ctx->fused && ctx->sh_mode == ODIS_SH_OFF) {
    for (int s = 0; s < count; s++) {
      FusedArgs<T> a = fused_args(ctx, o);
      hipEvent_t *ev = take(evs ? evs->mom : dummy_rec, s);
      if (s == 0) {
        if (ev)
          hipExtLaunchKernelGGL((k_fused<T, false>), gf, blk, 0, ctx->stream, ev[0], ev[1], 0, a,
                                ctx->energy_quad, table, s);
        else
          hipLaunchKernelGGL((k_fused<T, false>), gf, blk, 0, ctx->stream, a, ctx->energy_quad, table, s);
      } else {
        if (ev)
          hipExtLaunchKernelGGL((k_fused<T, true>), gf, blk, 0, ctx->stream, ev[0], ev[1], 0, a,
                                ctx->energy_quad, table, s);
        else
          hipLaunchKernelGGL((k_fused<T, true>), gf, blk, 0, ctx->stream, a, ctx->energy_quad, table, s);
        ctx->ecur ^= 1;
      }
      ctx->cur ^= 1;
    }
    // continuity of the chunk's last step, in place on eta_buf[ecur]
    CellArgs<T> ca = cell_args(ctx, o);
    ca.v_new = o.v[ctx->cur];
    hipEvent_t *ev = take(evs ? evs->cont : dummy_rec, -1);
    if (ev)
      hipExtLaunchKernelGGL((k_cell<T>), gc, blk, 0, ctx->stream, ev[0], ev[1], 0, ca, table, count - 1);
    else
      hipLaunchKernelGGL((k_cell<T>), gc, blk, 0, ctx->stream, ca, table, count - 1);
    CK(hipGetLastError());
    return ODIS_OK;
  }
  for (int s = 0; s < count; s++) {
    if (ctx->sh_mode != ODIS_SH_OFF) {
      const int rc = sh_effective_height(ctx, o);
      if (rc) return rc;
    }
    FaceArgs<T> fa = face_args(ctx, o);
    CellArgs<T> ca = cell_args(ctx, o);
    hipEvent_t *ev = take(evs ? evs->mom : dummy_rec, s);
    const bool sampled = ev != nullptr;
    if (ev)
      hipExtLaunchKernelGGL((k_face<T, false>), gf, blk, 0, ctx->stream, ev[0], ev[1], 0, fa, ctx->energy_quad,
                            table, s);
    else
      hipLaunchKernelGGL((k_face<T, false>), gf, blk, 0, ctx->stream, fa, ctx->energy_quad, table, s);
    ev = sampled ? take(evs->cont, s) : nullptr;
    if (ev)
      hipExtLaunchKernelGGL((k_cell<T>), gc, blk, 0, ctx->stream, ev[0], ev[1], 0, ca, table, s);
    else
      hipLaunchKernelGGL((k_cell<T>), gc, blk, 0, ctx->stream, ca, table, s);
    CK(hipGetLastError());
    ctx->cur ^= 1;
  }
  return ODIS_OK;
}

// Graph of graph_steps steps reading coefficient buffer k, one per entry
// parity of (v, eta); replaying it leaves the parities as the captured
// launches did.
template <typename T>
static int get_graph(odis_ctx *ctx, Ops<T> &o, int k, hipGraphExec_t *out) {
  const int key = ctx->cur * 2 + ctx->ecur;
  hipGraphExec_t &g = ctx->graphs[k][key];
  if (g) {
    *out = g;
    return ODIS_OK;
  }
  const int cur0 = ctx->cur, ecur0 = ctx->ecur;
  hipGraph_t graph = nullptr;
  CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_chunk(ctx, o, ctx->coef_dev[k], ctx->graph_steps, nullptr, nullptr);
  hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  // replays advance the parities (advance_parities); restore them here
  ctx->cur = cur0;
  ctx->ecur = ecur0;
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
    for (int c = 0; c < 4; c++)
      if (ctx->graphs[k][c]) {
        (void)hipGraphExecDestroy(ctx->graphs[k][c]);
        ctx->graphs[k][c] = nullptr;
      }
}

// parities after launching count steps from (cur, ecur)
static void advance_parities(odis_ctx *ctx, int count) {
  if (count & 1) ctx->cur ^= 1;
  if (ctx->fused && ctx->sh_mode == ODIS_SH_OFF && ((count - 1) & 1)) ctx->ecur ^= 1;
}

template <typename T>
static int step_t(odis_ctx *ctx, Ops<T> &o, int n_steps) {
  // Timing mode (ctx->timing = T > 0): the kernels of every T-th step get
  // dispatch-stamped start/end events (chunks with a sampled step are not
  // replayed from graphs).  Events are read once, after the last chunk.
  std::vector<ChunkEvents> sampled;
  size_t ev_used = 0;
  int done = 0;
  while (done < n_steps) {
    if (ctx->stop_flag) {
      ctx->stop_flag = 0;
      return fail(ctx, ODIS_E_STOPPED, "stop requested");
    }
    const int left = n_steps - done;
    const bool graph_fits = ctx->use_graphs && left >= ctx->graph_steps;
    int count = graph_fits ? ctx->graph_steps : std::min(ctx->chunk, left);
    // does [step, step + count) hold a multiple of the timing stride?
    const bool sample = ctx->timing > 0 && ((ctx->step + ctx->timing - 1) / ctx->timing) * ctx->timing <
                                               ctx->step + count;
    const bool graph = graph_fits && !sample;
    const double *table = nullptr;
    int rc = stage_coefficients(ctx, ctx->step, count, &table);
    if (rc) return rc;
    if (graph) {
      hipGraphExec_t g = nullptr;
      if ((rc = get_graph(ctx, o, ctx->coef_last, &g))) return rc;
      CK(hipGraphLaunch(g, ctx->stream));
      advance_parities(ctx, count);
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
  if (!sampled.empty()) {
    CK(hipStreamSynchronize(ctx->stream));
    double tm = 0, tc = 0, ts = 0;
    long long nm = 0, nc = 0, ns = 0;
    for (const ChunkEvents &c : sampled) {
      for (size_t j = 0; j < c.mom.size(); j++) {
        float a = 0;
        CK(hipEventElapsedTime(&a, ctx->ev[c.mom[j]], ctx->ev[c.mom[j] + 1]));
        tm += a;
        nm++;
      }
      for (int base : c.cont) {
        float a = 0;
        CK(hipEventElapsedTime(&a, ctx->ev[base], ctx->ev[base + 1]));
        tc += a;
        nc++;
      }
      // step time: momentum start to continuity end of a sampled split
      // step (momentum kernel alone when fused)
      for (size_t j = 0; j < c.mom.size(); j++) {
        const bool pair = !ctx->fused || ctx->sh_mode != ODIS_SH_OFF;
        const int end = pair && j < c.cont.size() ? c.cont[j] : c.mom[j];
        float a = 0;
        CK(hipEventElapsedTime(&a, ctx->ev[c.mom[j]], ctx->ev[end + 1]));
        ts += a;
        ns++;
      }
    }
    ctx->t_face = nm ? tm / nm : 0.0;
    ctx->t_cell = nc ? tc / nc : 0.0;
    ctx->t_step = ns ? ts / ns : 0.0;
  }
  return ODIS_OK;
}

extern "C" int odis_step(odis_ctx *ctx, int n_steps, int flags) {
  if (!ctx) return ODIS_E_INVALID;
  if (n_steps < 0) return fail(ctx, ODIS_E_INVALID, "n_steps < 0");
  int rc = ctx->precision == ODIS_FP64 ? step_t(ctx, ctx->d, n_steps) : step_t(ctx, ctx->f, n_steps);
  if (rc) return rc;
  if (flags & ODIS_SYNC) CK(hipStreamSynchronize(ctx->stream));
  return ODIS_OK;
}

extern "C" int odis_sync(odis_ctx *ctx) {
  if (!ctx) return ODIS_E_INVALID;
  CK(hipStreamSynchronize(ctx->stream));
  return ODIS_OK;
}

template <typename T>
static int diss_t(odis_ctx *ctx, Ops<T> &o) {
  FaceArgs<T> fa = face_args(ctx, o);
  const dim3 gf(ctx->nb_face, ctx->B), blk(kBlock);
  hipLaunchKernelGGL((k_face<T, true>), gf, blk, 0, ctx->stream, fa, ctx->energy_quad,
                     static_cast<const double *>(nullptr), 0);
  hipLaunchKernelGGL(k_diss_final, dim3(ctx->B), blk, 0, ctx->stream, ctx->partials, ctx->nb_face, ctx->dconst,
                     ctx->scratch_out);
  CK(hipGetLastError());
  return ODIS_OK;
}

extern "C" int odis_get_dissipation(odis_ctx *ctx, double *per_inst) {
  if (!ctx || !per_inst) return ODIS_E_INVALID;
  int rc = ctx->precision == ODIS_FP64 ? diss_t(ctx, ctx->d) : diss_t(ctx, ctx->f);
  if (rc) return rc;
  CK(hipMemcpyAsync(per_inst, ctx->scratch_out, sizeof(double) * ctx->B, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < ctx->B; b++)
    if (!std::isfinite(per_inst[b])) return fail(ctx, ODIS_E_NONFINITE, "non-finite dissipation");
  return ODIS_OK;
}

extern "C" int odis_series_capacity(const odis_ctx *ctx) { return ctx ? ctx->cap : ODIS_E_INVALID; }

extern "C" int odis_get_dissipation_series(odis_ctx *ctx, int inst, long long first, int count, double *out) {
  if (!ctx || !out || count < 0) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  CK(hipStreamSynchronize(ctx->stream));
  // series[k] exists for max(0, step - cap) <= k < step
  if (first < 0 || first + count > ctx->step || first < ctx->step - ctx->cap)
    return fail(ctx, ODIS_E_RANGE, "series range not available");
  std::vector<double> all(ctx->cap);
  CK(hipMemcpy(all.data(), ctx->series + size_t(inst) * ctx->cap, sizeof(double) * ctx->cap,
               hipMemcpyDeviceToHost));
  for (int k = 0; k < count; k++) out[k] = all[size_t((first + k) % ctx->cap)];
  return ODIS_OK;
}

template <typename T>
static int vel_t(odis_ctx *ctx, Ops<T> &o, int b, const double *normal, double *uv) {
  const int E = ctx->E;
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

extern "C" int odis_get_velocity_en(odis_ctx *ctx, int inst, const double *face_normal, double *uv) {
  if (!ctx || !face_normal || !uv) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  CK(hipStreamSynchronize(ctx->stream));
  return ctx->precision == ODIS_FP64 ? vel_t(ctx, ctx->d, inst, face_normal, uv)
                                     : vel_t(ctx, ctx->f, inst, face_normal, uv);
}

extern "C" int odis_request_stop(odis_ctx *ctx) {
  if (!ctx) return ODIS_E_INVALID;
  ctx->stop_flag = 1;
  return ODIS_OK;
}

extern "C" int odis_enable_timing(odis_ctx *ctx, int stride) {
  if (!ctx || stride < 0) return ODIS_E_INVALID;
  ctx->timing = stride;
  return ODIS_OK;
}

extern "C" int odis_use_graphs(odis_ctx *ctx, int on) {
  if (!ctx) return ODIS_E_INVALID;
  ctx->use_graphs = on ? 1 : 0;
  return ODIS_OK;
}

extern "C" int odis_use_fused(odis_ctx *ctx, int on) {
  if (!ctx) return ODIS_E_INVALID;
  const int f = on ? 1 : 0;
  if (f != ctx->fused) {
    CK(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->fused = f;
  }
  return ODIS_OK;
}

extern "C" int odis_get_timing(odis_ctx *ctx, double *face_ms, double *cell_ms, double *step_ms) {
  if (!ctx) return ODIS_E_INVALID;
  if (face_ms) *face_ms = ctx->t_face;
  if (cell_ms) *cell_ms = ctx->t_cell;
  if (step_ms) *step_ms = ctx->t_step;
  return ODIS_OK;
}

// ---------------------------------------------------------------------------
// geodesic -> lat-lon interpolation (src/interpolation.cpp:208-287)
// ---------------------------------------------------------------------------

extern "C" int odis_interp_ll_set(odis_ctx *ctx, int n_rows, int n_cols, const int *row_ptr, const int *col,
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

extern "C" int odis_interp_ll(odis_ctx *ctx, int inst, double *ll) {
  if (!ctx || !ll) return ODIS_E_INVALID;
  if (inst < 0 || inst >= ctx->B) return fail(ctx, ODIS_E_RANGE, "instance out of range");
  if (!ctx->ll_rows) return fail(ctx, ODIS_E_INVALID, "no interpolation matrix set");
  const dim3 grid((ctx->ll_rows + kBlock / 64 - 1) / (kBlock / 64)), blk(kBlock);
  if (ctx->precision == ODIS_FP64)
    hipLaunchKernelGGL((k_interp_ll<double>), grid, blk, 0, ctx->stream, ctx->ll_rows, ctx->ll_rp, ctx->ll_col,
                       ctx->ll_val, ctx->d.eta_buf[ctx->ecur] + size_t(inst) * ctx->N, ctx->ll_out);
  else
    hipLaunchKernelGGL((k_interp_ll<float>), grid, blk, 0, ctx->stream, ctx->ll_rows, ctx->ll_rp, ctx->ll_col,
                       ctx->ll_val, ctx->f.eta_buf[ctx->ecur] + size_t(inst) * ctx->N, ctx->ll_out);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(ll, ctx->ll_out, sizeof(double) * ctx->ll_rows, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return ODIS_OK;
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

extern "C" int odis_sh_set_basis(odis_ctx *ctx, int l_max, const double *beta, int mode) {
  if (!ctx) return ODIS_E_INVALID;
  if (mode != ODIS_SH_OFF && mode != ODIS_SH_FP64 && mode != ODIS_SH_SPLIT_BF16)
    return fail(ctx, ODIS_E_INVALID, "bad sh mode");
  CK(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  for (void *p : ctx->sh_allocs) (void)hipFree(p);
  ctx->sh_allocs.clear();
  ctx->sh_mode = ODIS_SH_OFF;
  ctx->sh_L = ctx->sh_K = 0;
  if (mode == ODIS_SH_OFF) return ODIS_OK;
  if (l_max < 0 || l_max > 100) return fail(ctx, ODIS_E_INVALID, "l_max out of range");
  const int K = (l_max + 1) * (l_max + 1), N = ctx->N;
  if (K > N) return fail(ctx, ODIS_E_INVALID, "(l_max+1)^2 exceeds the number of cells");
  if (size_t(K) * sizeof(double) > 64 * 1024) return fail(ctx, ODIS_E_INVALID, "l_max too large for LDS");
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
  if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_b), size_t(ctx->B) * K * sizeof(double)))) return rc;
  if ((rc = alloc(reinterpret_cast<void **>(&ctx->sh_c), size_t(ctx->B) * K * sizeof(double)))) return rc;
  const size_t tsz = ctx->precision == ODIS_FP64 ? sizeof(double) : sizeof(float);
  if ((rc = alloc(&ctx->eta_eff, size_t(ctx->B) * N * tsz))) return rc;
  ctx->sh_L = l_max;
  ctx->sh_K = K;
  ctx->sh_mode = mode;
  return ODIS_OK;
}

// field (reference order) -> LSQ coefficients with the context's basis
extern "C" int odis_sh_project(odis_ctx *ctx, const double *field, double *coeffs) {
  if (!ctx || !field || !coeffs) return ODIS_E_INVALID;
  if (ctx->sh_mode == ODIS_SH_OFF) return fail(ctx, ODIS_E_INVALID, "no spherical-harmonic basis set");
  const int K = ctx->sh_K, N = ctx->N;
  std::vector<double> x(N);
  for (int r = 0; r < N; r++) x[r] = field[ctx->cell_inv[r]];
  double *dx = nullptr, *db = nullptr, *dc = nullptr;
  CK(hipMalloc(&dx, sizeof(double) * N));
  hipError_t e = hipMalloc(&db, sizeof(double) * K);
  if (e == hipSuccess) e = hipMalloc(&dc, sizeof(double) * K);
  if (e == hipSuccess) e = hipMemcpy(dx, x.data(), sizeof(double) * N, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    if (ctx->sh_mode == ODIS_SH_SPLIT_BF16)
      hipLaunchKernelGGL((k_sh_analysis<double, true>), dim3(K, 1), dim3(kBlock), 0, ctx->stream, K, N, ctx->sh_Yt,
                         ctx->sh_Yp, dx, (long long)N, db);
    else
      hipLaunchKernelGGL((k_sh_analysis<double, false>), dim3(K, 1), dim3(kBlock), 0, ctx->stream, K, N,
                         ctx->sh_Yt, ctx->sh_Yp, dx, (long long)N, db);
    hipLaunchKernelGGL(k_sh_coef, dim3((K + 3) / 4, 1), dim3(kBlock), 0, ctx->stream, K, ctx->sh_Minv,
                       static_cast<const double *>(nullptr), db, dc);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(coeffs, dc, sizeof(double) * K, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  (void)hipFree(db);
  (void)hipFree(dc);
  if (e != hipSuccess) return fail(ctx, ODIS_E_HIP, hipGetErrorString(e));
  return ODIS_OK;
}

extern "C" int odis_sh_synthesize(odis_ctx *ctx, const double *coeffs, double *field) {
  if (!ctx || !field || !coeffs) return ODIS_E_INVALID;
  if (ctx->sh_mode == ODIS_SH_OFF) return fail(ctx, ODIS_E_INVALID, "no spherical-harmonic basis set");
  const int K = ctx->sh_K, N = ctx->N;
  double *dy = nullptr, *dc = nullptr;
  CK(hipMalloc(&dy, sizeof(double) * N));
  hipError_t e = hipMalloc(&dc, sizeof(double) * K);
  if (e == hipSuccess) e = hipMemcpy(dc, coeffs, sizeof(double) * K, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const size_t lds = sizeof(double) * K;
    if (ctx->sh_mode == ODIS_SH_SPLIT_BF16)
      hipLaunchKernelGGL((k_sh_synth<double, true>), dim3(ctx->nb_cell, 1), dim3(kBlock), lds, ctx->stream, K, N,
                         ctx->sh_Yt, ctx->sh_Yp, dc, static_cast<const double *>(nullptr), (long long)N, dy);
    else
      hipLaunchKernelGGL((k_sh_synth<double, false>), dim3(ctx->nb_cell, 1), dim3(kBlock), lds, ctx->stream, K,
                         N, ctx->sh_Yt, ctx->sh_Yp, dc, static_cast<const double *>(nullptr), (long long)N, dy);
    e = hipGetLastError();
  }
  std::vector<double> y(N);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(y.data(), dy, sizeof(double) * N, hipMemcpyDeviceToHost);
  (void)hipFree(dy);
  (void)hipFree(dc);
  if (e != hipSuccess) return fail(ctx, ODIS_E_HIP, hipGetErrorString(e));
  for (int r = 0; r < N; r++) field[ctx->cell_inv[r]] = y[r];
  return ODIS_OK;
}
