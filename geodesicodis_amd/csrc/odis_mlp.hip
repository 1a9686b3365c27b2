// odis_mlp.hip -- the learned warm start's network: a fully connected
// tanh network (n_in -> width x depth -> n_out, linear output) trained and
// evaluated on the GPU.  No counterpart in the reference, which starts from
// rest, a file or its analytic solution (src/initialConditions.cpp:19-206).
//
// All products (forward X W^T, weight gradients dZ^T X, input gradients
// dZ W) run on one FP32 matrix-core kernel, k_mlp_gemm: a 64 x 64 tile per
// workgroup, 4 waves of 32 x 32, v_mfma_f32_16x16x4f32 over K in LDS stages
// of 16, with the bias and tanh fused into the forward epilogue.  Loss
// (mean squared error / 2), bias gradients, Adam and the minibatch gather
// are small elementwise kernels.  Sums that give the loss and the bias
// gradients run in a fixed order, so a training run is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <exception>
#include <new>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "odis.h"

namespace {

thread_local std::string g_mlp_error;

int mlp_fail(int code, const std::string &msg) {
  g_mlp_error = msg;
  return code;
}

#define MCK(x)                                                                      \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) return mlp_fail(ODIS_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kTile = 64, kKs = 16, kThreads = 256;

// C[m][n] = sum_k A(m, k) B(k, n) (+ bias[n], then tanh when ACT).
// A(m, k) = TA ? A[k * lda + m] : A[m * lda + k]; B(k, n) = TB ? B[n * ldb + k]
// : B[k * ldb + n].  Out-of-range operands read as 0, out-of-range results
// are not stored.  Wave w of the workgroup computes rows 32 (w / 2) and
// columns 32 (w % 2) of the tile as 2 x 2 blocks of 16 x 16; lane l of a
// block holds A(l % 16, l / 16), B(l / 16, l % 16) per product and
// C(4 (l / 16) + r, l % 16), r = 0..3, of the result.
template <bool TA, bool TB, int EPI>  // EPI 0: store, 1: + bias, 2: + bias, tanh
__global__ __launch_bounds__(kThreads) void k_mlp_gemm(int M, int N, int K, const float *__restrict__ A, int lda,
                                                       const float *__restrict__ B, int ldb,
                                                       const float *__restrict__ bias, float *__restrict__ C,
                                                       int ldc) {
  __shared__ float As[kTile][kKs + 1];
  __shared__ float Bs[kKs][kTile + 1];
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  typedef float f4 __attribute__((ext_vector_type(4)));
  f4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += kKs) {
    // 1024 elements of each operand tile, 4 per thread
    for (int q = threadIdx.x; q < kTile * kKs; q += kThreads) {
      const int r = TA ? q % kTile : q / kKs, c = TA ? q / kTile : q % kKs;  // coalesced along memory
      const int gm = m0 + r, gk = k0 + c;
      As[r][c] = (gm < M && gk < K) ? (TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk]) : 0.f;
      const int rb = TB ? q % kKs : q / kTile, cb = TB ? q / kKs : q % kTile;
      const int bk = k0 + rb, bn = n0 + cb;
      Bs[rb][cb] = (bk < K && bn < N) ? (TB ? B[(size_t)bn * ldb + bk] : B[(size_t)bk * ldb + bn]) : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kKs; kk += 4) {
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[wm + 16 * i + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kk + (lane >> 4)][wn + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn + 16 * j + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        if (row < M && col < N) {
          float x = acc[i][j][r];
          if (EPI >= 1) x += bias[col];
          if (EPI == 2) x = tanhf(x);
          C[(size_t)row * ldc + col] = x;
        }
      }
    }
}

// rows idx[m] of the data set into a contiguous minibatch
__global__ __launch_bounds__(kThreads) void k_mlp_gather(int n, int w, const int *__restrict__ idx,
                                                          const float *__restrict__ src, float *__restrict__ dst) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < (size_t)n * w) dst[i] = src[(size_t)idx[i / w] * w + i % w];
}

// dY = (Y - T) / count and a fixed-order partial of the squared error per
// workgroup: loss = sum (Y - T)^2 / (2 count)
__global__ __launch_bounds__(kThreads) void k_mlp_mse(size_t total, double count, const float *__restrict__ Y,
                                                       const float *__restrict__ T, float *__restrict__ dY,
                                                       double *__restrict__ part) {
  __shared__ double sh[kThreads];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < total; i += (size_t)gridDim.x * kThreads) {
    const double d = double(Y[i]) - double(T[i]);
    s += d * d;
    dY[i] = float(d / count);
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] / (2.0 * count);
}

// acc[0] += sum of the partials (one thread, fixed order)
__global__ void k_mlp_sum(int n, const double *__restrict__ part, double *__restrict__ acc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    acc[0] += s;
  }
}

// dZ = dA (1 - A^2), the tanh derivative through its output
__global__ __launch_bounds__(kThreads) void k_mlp_dtanh(size_t total, const float *__restrict__ A,
                                                         float *__restrict__ dZ) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < total) dZ[i] *= 1.f - A[i] * A[i];
}

// db[j] = sum over rows of dZ[., j] (fixed order per column)
__global__ __launch_bounds__(kThreads) void k_mlp_colsum(int M, int N, const float *__restrict__ dZ,
                                                          float *__restrict__ db) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dZ[(size_t)m * N + j];
  db[j] = s;
}

// Adam (Kingma & Ba) with bias-corrected moments
__global__ __launch_bounds__(kThreads) void k_mlp_adam(int n, float lr, float b1, float b2, float eps, float c1,
                                                        float c2, const float *__restrict__ g, float *__restrict__ p,
                                                        float *__restrict__ m, float *__restrict__ v) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

struct odis_mlp {
  int n_in = 0, width = 0, depth = 0, n_out = 0;
  std::vector<int> din, dout;         // per layer
  std::vector<size_t> woff, boff;     // offsets of W_l and b_l in the flat parameter vector
  size_t n_params = 0;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  int cap = 0;                        // rows the activation buffers hold
  std::vector<float *> act;           // act[0] = input batch, act[l + 1] = output of layer l
  float *dA = nullptr, *dB = nullptr; // gradient buffers [cap][max width]
  float *tgt = nullptr;               // target batch [cap][n_out]
  double *part = nullptr, *acc = nullptr;
  int n_data = 0;
  float *X = nullptr, *Y = nullptr;   // data set on the device
  int *idx = nullptr;
  long long adam_t = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
};

namespace {

constexpr int kPartBlocks = 64;

template <typename T>
int mlp_alloc(odis_mlp *m, T **p, size_t count) {
  void *q = nullptr;
  MCK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
  m->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return ODIS_OK;
}

void free_all(odis_mlp *m) {
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void *q : m->allocs) (void)hipFree(q);
  m->allocs.clear();
  if (m->stream) (void)hipStreamDestroy(m->stream);
  m->stream = nullptr;
}

int ensure_capacity(odis_mlp *m, int rows) {
  if (rows <= m->cap) return ODIS_OK;
  // activation buffers are reallocated for the larger batch (only at a
  // first use or a larger batch; the stream is idle between entry points)
  for (float *a : m->act)
    if (a) {
      auto it = std::find(m->allocs.begin(), m->allocs.end(), static_cast<void *>(a));
      if (it != m->allocs.end()) m->allocs.erase(it);
      (void)hipFree(a);
    }
  for (float **b : {&m->dA, &m->dB, &m->tgt})
    if (*b) {
      auto it = std::find(m->allocs.begin(), m->allocs.end(), static_cast<void *>(*b));
      if (it != m->allocs.end()) m->allocs.erase(it);
      (void)hipFree(*b);
      *b = nullptr;
    }
  m->act.assign(m->depth + 2, nullptr);
  const int wmax = std::max({m->n_in, m->width, m->n_out});
  int rc = mlp_alloc(m, &m->act[0], size_t(rows) * m->n_in);
  for (int l = 0; l <= m->depth && !rc; ++l) rc = mlp_alloc(m, &m->act[l + 1], size_t(rows) * m->dout[l]);
  if (!rc) rc = mlp_alloc(m, &m->dA, size_t(rows) * wmax);
  if (!rc) rc = mlp_alloc(m, &m->dB, size_t(rows) * wmax);
  if (!rc) rc = mlp_alloc(m, &m->tgt, size_t(rows) * m->n_out);
  if (rc) return rc;
  m->cap = rows;
  return ODIS_OK;
}

template <bool TA, bool TB, int EPI>
void gemm(odis_mlp *m, int M, int N, int K, const float *A, int lda, const float *B, int ldb, const float *bias,
          float *C, int ldc) {
  const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
  hipLaunchKernelGGL((k_mlp_gemm<TA, TB, EPI>), grid, dim3(kThreads), 0, m->stream, M, N, K, A, lda, B, ldb, bias,
                     C, ldc);
}

// forward pass of the n rows in act[0]; act[l + 1] = layer l's output
void forward(odis_mlp *m, int n) {
  for (int l = 0; l <= m->depth; ++l) {
    const float *W = m->p + m->woff[l], *b = m->p + m->boff[l];
    if (l < m->depth)
      gemm<false, true, 2>(m, n, m->dout[l], m->din[l], m->act[l], m->din[l], W, m->din[l], b, m->act[l + 1],
                           m->dout[l]);
    else
      gemm<false, true, 1>(m, n, m->dout[l], m->din[l], m->act[l], m->din[l], W, m->din[l], b, m->act[l + 1],
                           m->dout[l]);
  }
}

// loss of act[depth + 1] against tgt added to acc[0], gradients into g
void backward(odis_mlp *m, int n) {
  const size_t tot = size_t(n) * m->n_out;
  hipLaunchKernelGGL(k_mlp_mse, dim3(kPartBlocks), dim3(kThreads), 0, m->stream, tot, double(tot),
                     m->act[m->depth + 1], m->tgt, m->dA, m->part);
  hipLaunchKernelGGL(k_mlp_sum, dim3(1), dim3(64), 0, m->stream, kPartBlocks, m->part, m->acc);
  float *dcur = m->dA, *dnext = m->dB;  // dcur: gradient of layer l's pre-activation output
  for (int l = m->depth; l >= 0; --l) {
    const int no = m->dout[l], ni = m->din[l];
    if (l < m->depth) {
      const size_t t2 = size_t(n) * no;
      hipLaunchKernelGGL(k_mlp_dtanh, dim3(unsigned((t2 + kThreads - 1) / kThreads)), dim3(kThreads), 0, m->stream,
                         t2, m->act[l + 1], dcur);
    }
    // dW_l = dZ^T act[l]  ([no][ni]), db_l = column sums of dZ
    gemm<true, false, 0>(m, no, ni, n, dcur, no, m->act[l], ni, nullptr, m->g + m->woff[l], ni);
    hipLaunchKernelGGL(k_mlp_colsum, dim3((no + kThreads - 1) / kThreads), dim3(kThreads), 0, m->stream, n, no,
                       dcur, m->g + m->boff[l]);
    if (l > 0) {  // d act[l] = dZ W_l  ([n][ni])
      gemm<false, false, 0>(m, n, ni, no, dcur, no, m->p + m->woff[l], ni, nullptr, dnext, ni);
      std::swap(dcur, dnext);
    }
  }
}

int create_impl(int n_in, int width, int depth, int n_out, unsigned long long seed, odis_mlp **out) {
  if (!out || n_in < 1 || width < 1 || depth < 0 || n_out < 1 || n_in > 4096 || width > 4096 || n_out > 65536 ||
      depth > 16)
    return mlp_fail(ODIS_E_INVALID, "bad network shape");
  odis_mlp *m = new odis_mlp();
  m->n_in = n_in;
  m->width = width;
  m->depth = depth;
  m->n_out = n_out;
  size_t off = 0;
  for (int l = 0; l <= depth; ++l) {
    const int ni = l == 0 ? n_in : width, no = l == depth ? n_out : width;
    m->din.push_back(ni);
    m->dout.push_back(no);
    m->woff.push_back(off);
    off += size_t(ni) * no;
    m->boff.push_back(off);
    off += size_t(no);
  }
  m->n_params = off;
  int rc = ODIS_OK;
  hipError_t he = hipStreamCreateWithFlags(&m->stream, hipStreamDefault);
  if (he != hipSuccess) rc = mlp_fail(ODIS_E_HIP, hipGetErrorString(he));
  if (!rc) rc = mlp_alloc(m, &m->p, off);
  if (!rc) rc = mlp_alloc(m, &m->g, off);
  if (!rc) rc = mlp_alloc(m, &m->m1, off);
  if (!rc) rc = mlp_alloc(m, &m->m2, off);
  if (!rc) rc = mlp_alloc(m, &m->part, kPartBlocks);
  if (!rc) rc = mlp_alloc(m, &m->acc, 1);
  if (!rc) {
    // Glorot-uniform weights, zero biases, from a seeded 64-bit Mersenne twister
    std::vector<float> h(off, 0.f);
    std::mt19937_64 gen(seed);
    for (int l = 0; l <= depth; ++l) {
      const double a = std::sqrt(6.0 / double(m->din[l] + m->dout[l]));
      std::uniform_real_distribution<double> u(-a, a);
      for (size_t i = 0; i < size_t(m->din[l]) * m->dout[l]; ++i) h[m->woff[l] + i] = float(u(gen));
    }
    he = hipMemcpy(m->p, h.data(), off * sizeof(float), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemset(m->m1, 0, off * sizeof(float));
    if (he == hipSuccess) he = hipMemset(m->m2, 0, off * sizeof(float));
    if (he == hipSuccess) he = hipDeviceSynchronize();
    if (he != hipSuccess) rc = mlp_fail(ODIS_E_HIP, hipGetErrorString(he));
  }
  if (rc) {
    free_all(m);
    delete m;
    return rc;
  }
  *out = m;
  return ODIS_OK;
}

template <typename F>
int mlp_guarded(F &&body) {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return mlp_fail(ODIS_E_NOMEM, "host allocation failed");
  } catch (const std::exception &e) {
    return mlp_fail(ODIS_E_INVALID, e.what());
  }
}

}  // namespace

extern "C" const char *odis_mlp_last_error(void) { return g_mlp_error.c_str(); }

extern "C" int odis_mlp_create(int n_in, int width, int depth, int n_out, unsigned long long seed, odis_mlp **out) {
  return mlp_guarded([&] { return create_impl(n_in, width, depth, n_out, seed, out); });
}

extern "C" int odis_mlp_destroy(odis_mlp *m) {
  if (!m) return ODIS_E_INVALID;
  free_all(m);
  delete m;
  return ODIS_OK;
}

extern "C" long long odis_mlp_param_count(const odis_mlp *m) { return m ? (long long)m->n_params : ODIS_E_INVALID; }

extern "C" int odis_mlp_get_params(odis_mlp *m, float *params) {
  if (!m || !params) return mlp_fail(ODIS_E_INVALID, "null argument");
  MCK(hipStreamSynchronize(m->stream));
  MCK(hipMemcpy(params, m->p, m->n_params * sizeof(float), hipMemcpyDeviceToHost));
  return ODIS_OK;
}

extern "C" int odis_mlp_set_params(odis_mlp *m, const float *params) {
  if (!m || !params) return mlp_fail(ODIS_E_INVALID, "null argument");
  MCK(hipStreamSynchronize(m->stream));
  MCK(hipMemcpy(m->p, params, m->n_params * sizeof(float), hipMemcpyHostToDevice));
  MCK(hipMemset(m->m1, 0, m->n_params * sizeof(float)));
  MCK(hipMemset(m->m2, 0, m->n_params * sizeof(float)));
  MCK(hipDeviceSynchronize());
  m->adam_t = 0;
  return ODIS_OK;
}

extern "C" int odis_mlp_forward(odis_mlp *m, int n, const float *x, float *y) {
  return mlp_guarded([&]() -> int {
    if (!m || n < 0 || (n > 0 && (!x || !y))) return mlp_fail(ODIS_E_INVALID, "bad argument");
    if (n == 0) return ODIS_OK;
    int rc = ensure_capacity(m, n);
    if (rc) return rc;
    MCK(hipMemcpyAsync(m->act[0], x, size_t(n) * m->n_in * sizeof(float), hipMemcpyHostToDevice, m->stream));
    forward(m, n);
    MCK(hipGetLastError());
    MCK(hipMemcpyAsync(y, m->act[m->depth + 1], size_t(n) * m->n_out * sizeof(float), hipMemcpyDeviceToHost,
                       m->stream));
    MCK(hipStreamSynchronize(m->stream));
    return ODIS_OK;
  });
}

extern "C" int odis_mlp_set_data(odis_mlp *m, int n, const float *x, const float *y) {
  return mlp_guarded([&]() -> int {
    if (!m || n < 1 || !x || !y) return mlp_fail(ODIS_E_INVALID, "bad argument");
    MCK(hipStreamSynchronize(m->stream));
    for (void **b : {reinterpret_cast<void **>(&m->X), reinterpret_cast<void **>(&m->Y),
                     reinterpret_cast<void **>(&m->idx)})
      if (*b) {
        auto it = std::find(m->allocs.begin(), m->allocs.end(), *b);
        if (it != m->allocs.end()) m->allocs.erase(it);
        (void)hipFree(*b);
        *b = nullptr;
      }
    m->n_data = 0;
    int rc = mlp_alloc(m, &m->X, size_t(n) * m->n_in);
    if (!rc) rc = mlp_alloc(m, &m->Y, size_t(n) * m->n_out);
    if (!rc) rc = mlp_alloc(m, &m->idx, size_t(n));
    if (rc) return rc;
    MCK(hipMemcpy(m->X, x, size_t(n) * m->n_in * sizeof(float), hipMemcpyHostToDevice));
    MCK(hipMemcpy(m->Y, y, size_t(n) * m->n_out * sizeof(float), hipMemcpyHostToDevice));
    m->n_data = n;
    return ODIS_OK;
  });
}

namespace {

// loss and gradients of the rows idx_dev[0..n) of the data set; gradients in m->g
int batch_grad(odis_mlp *m, int n, const int *idx_dev) {
  const unsigned gx = unsigned((size_t(n) * m->n_in + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_mlp_gather, dim3(gx), dim3(kThreads), 0, m->stream, n, m->n_in, idx_dev, m->X, m->act[0]);
  const unsigned gy = unsigned((size_t(n) * m->n_out + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_mlp_gather, dim3(gy), dim3(kThreads), 0, m->stream, n, m->n_out, idx_dev, m->Y, m->tgt);
  forward(m, n);
  backward(m, n);
  MCK(hipGetLastError());
  return ODIS_OK;
}

}  // namespace

extern "C" int odis_mlp_loss_grad(odis_mlp *m, int n, const int *idx, double *loss, float *grad) {
  return mlp_guarded([&]() -> int {
    if (!m || n < 1 || !idx || !loss) return mlp_fail(ODIS_E_INVALID, "bad argument");
    if (!m->n_data) return mlp_fail(ODIS_E_INVALID, "no data set");
    for (int i = 0; i < n; ++i)
      if (idx[i] < 0 || idx[i] >= m->n_data) return mlp_fail(ODIS_E_RANGE, "sample index out of range");
    if (n > m->n_data) return mlp_fail(ODIS_E_RANGE, "more rows than the data set");
    int rc = ensure_capacity(m, n);
    if (rc) return rc;
    MCK(hipMemcpyAsync(m->idx, idx, size_t(n) * sizeof(int), hipMemcpyHostToDevice, m->stream));
    MCK(hipMemsetAsync(m->acc, 0, sizeof(double), m->stream));
    if ((rc = batch_grad(m, n, m->idx))) return rc;
    MCK(hipMemcpyAsync(loss, m->acc, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    if (grad) MCK(hipMemcpyAsync(grad, m->g, m->n_params * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MCK(hipStreamSynchronize(m->stream));
    return ODIS_OK;
  });
}

extern "C" int odis_mlp_train(odis_mlp *m, int epochs, int batch, float lr_start, float lr_end,
                              unsigned long long seed, double *epoch_loss) {
  return mlp_guarded([&]() -> int {
    if (!m || epochs < 1 || batch < 1 || !(lr_start > 0) || !(lr_end > 0))
      return mlp_fail(ODIS_E_INVALID, "bad argument");
    if (!m->n_data) return mlp_fail(ODIS_E_INVALID, "no data set");
    const int n = m->n_data, bs = std::min(batch, n);
    int rc = ensure_capacity(m, bs);
    if (rc) return rc;
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 gen(seed);
    const int nb = (n + bs - 1) / bs;
    const long long total = (long long)epochs * nb;
    long long it = 0;
    const unsigned ga = unsigned((m->n_params + kThreads - 1) / kThreads);
    for (int e = 0; e < epochs; ++e) {
      std::shuffle(perm.begin(), perm.end(), gen);
      MCK(hipMemcpyAsync(m->idx, perm.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice, m->stream));
      MCK(hipMemsetAsync(m->acc, 0, sizeof(double), m->stream));
      for (int b = 0; b < nb; ++b, ++it) {
        const int rows = std::min(bs, n - b * bs);
        if ((rc = batch_grad(m, rows, m->idx + size_t(b) * bs))) return rc;
        // cosine decay from lr_start to lr_end over all minibatches
        const double fr = total > 1 ? double(it) / double(total - 1) : 1.0;
        const float lr = float(lr_end + 0.5 * (lr_start - lr_end) * (1.0 + std::cos(M_PI * fr)));
        m->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = float(1.0 - std::pow(0.9, double(m->adam_t)));
        const float c2 = float(1.0 - std::pow(0.999, double(m->adam_t)));
        hipLaunchKernelGGL(k_mlp_adam, dim3(ga), dim3(kThreads), 0, m->stream, int(m->n_params), lr, b1, b2, 1e-8f,
                           c1, c2, m->g, m->p, m->m1, m->m2);
      }
      MCK(hipGetLastError());
      if (epoch_loss) {
        double l = 0.0;
        MCK(hipMemcpyAsync(&l, m->acc, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        MCK(hipStreamSynchronize(m->stream));
        epoch_loss[e] = l / nb;  // mean minibatch loss of the epoch
      }
      // the pinned-free host permutation is reused next epoch: the copy must be done
      MCK(hipStreamSynchronize(m->stream));
    }
    return ODIS_OK;
  });
}
