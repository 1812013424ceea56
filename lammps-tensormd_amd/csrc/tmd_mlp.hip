// Per-element MLP forward and backward (dE/dG) as MFMA GEMMs over
// element-sorted row tiles, plus the small reduction/integration kernels.
//
// Replaces forward_resnet_GPU / backward_resnet_GPU (gpu/cuda/nnp.cu:149-281,
// cuBLAS GEMMs + axpy_* bias/activation kernels) and the CPU dense path
// (src/ML-TENSORMD/nnp.cpp:246-366) with one launch per layer for all
// elements: a 64-row tile never straddles two elements, and the tile's
// element selects the weights.
//
// FP32 mode uses v_mfma_f32_16x16x4_f32, BF16 mode v_mfma_f32_16x16x32_bf16
// (gfx950), both with FP32 accumulation; the epilogue adds the bias, applies
// the activation, adds the resnet input and stores act'(z) for the backward.
#include "tmd_device.h"

#include <algorithm>

namespace tmd {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef long fp8x8;

__device__ inline float bf16_to_f32(uint16_t v) { return __uint_as_float(((uint32_t) v) << 16); }
__device__ inline uint16_t f32_to_bf16(float f)
{
  // round to nearest even (no NaN inputs on this path)
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t) (u >> 16);
}

// activation and derivative (nnp.cu:9-49, activation.hpp:158-261)
template <int ACT>
__device__ inline float act_f(float z, float &dz)
{
  if (ACT == 1) {  // softplus, linear above -(log(eps)+2) (nnp.cu:44-45)
    float e = __expf(z);
    const float thr = 13.942385f;  // -(log(1.19209e-07) + 2)
    float h = (z < thr) ? log1pf(e) : z;
    dz = 1.0f / (1.0f + 1.0f / e);
    return h;
  } else if (ACT == 2) {
    float h = tanhf(z);
    dz = 1.0f - h * h;
    return h;
  } else if (ACT == 3) {
    float s = sqrtf(z * z + 4.0f);
    float h = 0.5f * (z + s);
    dz = h / s;
    return h;
  } else {
    dz = z >= 0.0f ? 1.0f : 0.0f;
    return z >= 0.0f ? z : 0.0f;
  }
}

__device__ inline int tile_elem(const RowLayout &rl, int row0)
{
  int e = 0;
  while (e < kMaxElts - 1 && row0 >= rl.row_off[e + 1]) e++;
  return e;
}

template <typename T>
__device__ inline float ld_f(const T *p, size_t i);
template <>
__device__ inline float ld_f<float>(const float *p, size_t i) { return p[i]; }
template <>
__device__ inline float ld_f<uint16_t>(const uint16_t *p, size_t i) { return bf16_to_f32(p[i]); }

template <typename T>
__device__ inline void st_f(T *p, size_t i, float v);
template <>
__device__ inline void st_f<float>(float *p, size_t i, float v) { p[i] = v; }
template <>
__device__ inline void st_f<uint16_t>(uint16_t *p, size_t i, float v) { p[i] = f32_to_bf16(v); }

// ---------------------------------------------------------------------------
// C[64 x 64 tile] = Aeff[rows x K] * B[K x N]
//   PRO = 0: Aeff = A                (forward)
//   PRO = 1: Aeff = Gout .* D         (backward; Gout row-major [M][Kd] where Kd = K)
//   PRO = 2: Aeff = vec[k] * D        (backward of the last hidden layer)
// B is [K][N] row-major fp32 (FP32 mode) or bf16 (BF16 mode).
// 256 threads: 4 waves in 2x2, each wave 32x32 = 2x2 MFMA tiles of 16x16.
constexpr int BM = 64, BN = 64;

template <int PRO, typename TA>
__device__ inline float load_a(const TA *A, const float *D, const float *vec, size_t row, int k,
                               int ldk)
{
  if (PRO == 0) return ld_f<TA>(A, row * ldk + k);
  if (PRO == 1) return ld_f<TA>(A, row * ldk + k) * D[row * ldk + k];
  return vec[k] * D[row * ldk + k];
}

// ---- FP32 -------------------------------------------------------------------
template <int PRO, typename TA>
__device__ inline void gemm_f32(const TA *A, const float *D, const float *vec, int K, int N,
                                const float *B, int row0, int col0, f32x4 acc[2][2])
{
  constexpr int BK = 16;
  __shared__ float sA[BM][BK + 1];
  __shared__ float sB[BK][BN + 4];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += BK) {
    // A tile: 64 x 16, 4 per thread
    {
      int r = t >> 2, c = (t & 3) * 4;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int k = k0 + c + q;
        sA[r][c + q] = (k < K) ? load_a<PRO, TA>(A, D, vec, (size_t) (row0 + r), k, K) : 0.0f;
      }
    }
    // B tile: 16 x 64, 4 per thread
    {
      int r = t >> 4, c = (t & 15) * 4;
      int k = k0 + r;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int n = col0 + c + q;
        sB[r][c + q] = (k < K && n < N) ? B[(size_t) k * N + n] : 0.0f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; a++) av[a] = sA[wm * 32 + a * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int b = 0; b < 2; b++) bv[b] = sB[kk + (lane >> 4)][wn * 32 + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
}

// ---- BF16 -------------------------------------------------------------------
template <int PRO, typename TA>
__device__ inline void gemm_bf16(const TA *A, const float *D, const float *vec, int K, int N,
                                 const uint16_t *B, int row0, int col0, f32x4 acc[2][2])
{
  constexpr int BK = 32;
  // k-contiguous rows so one lane's 8 consecutive k are one 16-byte read
  __shared__ __align__(16) uint16_t sA[BM][BK + 8];
  __shared__ __align__(16) uint16_t sBt[BN][BK + 8];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += BK) {
    {
      // A: 64 x 32, 8 per thread
      int r = t >> 2, c = (t & 3) * 8;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        int k = k0 + c + q;
        float v = (k < K) ? load_a<PRO, TA>(A, D, vec, (size_t) (row0 + r), k, K) : 0.0f;
        sA[r][c + q] = f32_to_bf16(v);
      }
    }
    {
      // B: 32 x 64 -> transposed into sBt[n][k], 8 per thread
      int r = t >> 3, c = (t & 7) * 8;
      int k = k0 + r;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        int n = col0 + c + q;
        sBt[c + q][r] = (k < K && n < N) ? B[(size_t) k * N + n] : (uint16_t) 0;
      }
    }
    __syncthreads();
    bf16x8 av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; a++)
      av[a] = *reinterpret_cast<const bf16x8 *>(&sA[wm * 32 + a * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
    for (int b = 0; b < 2; b++)
      bv[b] = *reinterpret_cast<const bf16x8 *>(&sBt[wn * 32 + b * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
}

// Forward hidden layer: H = act(A W + b) (+ A if res), D = act'(.)
template <int MODE, int ACT, typename TA, typename TH>
__global__ __launch_bounds__(256) void k_dense_fwd(RowLayout rl, LayerParams lp, const TA *A, TH *H,
                                                   float *D)
{
  const int row0 = blockIdx.x * BM, col0 = blockIdx.y * BN;
  if (row0 >= rl.nrows) return;
  const int e = tile_elem(rl, row0);
  f32x4 acc[2][2];
  if (MODE == 0)
    gemm_f32<0, TA>(A, nullptr, nullptr, lp.K, lp.N, lp.W + e * lp.w_stride, row0, col0, acc);
  else
    gemm_bf16<0, TA>(A, nullptr, nullptr, lp.K, lp.N, lp.Wb + e * lp.w_stride, row0, col0, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const float *bias = lp.b ? lp.b + e * lp.b_stride : nullptr;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int r = row0 + wm * 32 + a * 16 + (lane >> 4) * 4 + q;
        int c = col0 + wn * 32 + b * 16 + (lane & 15);
        if (c >= lp.N) continue;
        float z = acc[a][b][q] + (bias ? bias[c] : 0.0f);
        float dz;
        float h = act_f<ACT>(z, dz);
        if (lp.res) h += ld_f<TA>(A, (size_t) r * lp.K + c);
        st_f<TH>(H, (size_t) r * lp.N + c, h);
        D[(size_t) r * lp.N + c] = dz;
      }
}

// Backward: Gin = Aeff W^T (+ Gout if res).  B = W^T stored [N][K] -> here
// the GEMM inner dimension is the layer's N and the output width is K.
template <int MODE, int PRO, typename TG, typename TO>
__global__ __launch_bounds__(256) void k_dense_bwd(RowLayout rl, LayerParams lp, const TG *Gout,
                                                   const float *gvec, int vec_stride,
                                                   const float *D, TO *Gin)
{
  const int row0 = blockIdx.x * BM, col0 = blockIdx.y * BN;
  if (row0 >= rl.nrows) return;
  const int e = tile_elem(rl, row0);
  const float *vec = gvec ? gvec + (size_t) e * vec_stride : nullptr;
  f32x4 acc[2][2];
  // GEMM: [M x N] * [N x K]
  if (MODE == 0)
    gemm_f32<PRO, TG>(Gout, D, vec, lp.N, lp.K, lp.Wt + e * lp.w_stride, row0, col0, acc);
  else
    gemm_bf16<PRO, TG>(Gout, D, vec, lp.N, lp.K, lp.Wtb + e * lp.w_stride, row0, col0, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int r = row0 + wm * 32 + a * 16 + (lane >> 4) * 4 + q;
        int c = col0 + wn * 32 + b * 16 + (lane & 15);
        if (c >= lp.K) continue;
        float v = acc[a][b][q];
        if (lp.res) v += (PRO == 2) ? vec[c] : ld_f<TG>(Gout, (size_t) r * lp.N + c);
        st_f<TO>(Gin, (size_t) r * lp.K + c, v);
      }
}

// Output layer (width 1): one wave per row group, E = H . w + b
template <typename TH>
__global__ void k_output(RowLayout rl, LayerParams lp, const TH *H, float *E)
{
  int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (row >= rl.nrows) return;
  int e = tile_elem(rl, row & ~(kRowTile - 1));
  const float *w = lp.W + e * lp.w_stride;
  float s = 0.0f;
  for (int k = lane; k < lp.K; k += 64) s += ld_f<TH>(H, (size_t) row * lp.K + k) * w[k];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) E[row] = s + (lp.b ? lp.b[e * lp.b_stride] : 0.0f);
}

// ---------------------------------------------------------------------------
template <int MODE, typename TA, typename TH>
hipError_t fwd_act(int act, const RowLayout &rl, const LayerParams &lp, const TA *A, TH *H, float *D,
                   hipStream_t s)
{
  dim3 grid(rl.nrows / BM, (lp.N + BN - 1) / BN);
  switch (act) {
    case 0: k_dense_fwd<MODE, 0, TA, TH><<<grid, 256, 0, s>>>(rl, lp, A, H, D); break;
    case 1: k_dense_fwd<MODE, 1, TA, TH><<<grid, 256, 0, s>>>(rl, lp, A, H, D); break;
    case 2: k_dense_fwd<MODE, 2, TA, TH><<<grid, 256, 0, s>>>(rl, lp, A, H, D); break;
    default: k_dense_fwd<MODE, 3, TA, TH><<<grid, 256, 0, s>>>(rl, lp, A, H, D); break;
  }
  return hipGetLastError();
}

// reductions --------------------------------------------------------------------
__global__ void k_zero(uint4 *p, size_t n16, unsigned char *tail, size_t ntail)
{
  size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for (size_t q = i; q < n16; q += stride) p[q] = make_uint4(0, 0, 0, 0);
  if (i < ntail) tail[i] = 0;
}

// fixed-order two-level reduction: each block writes a partial, the last
// kernel sums partials in order (deterministic)
__global__ void k_reduce_ev_part(const float *E, const double *wrow, const int *row2atom, int nrows,
                                 double *part)
{
  __shared__ double sh[7][256];
  int t = threadIdx.x;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int r = blockIdx.x * 256 + t; r < nrows; r += gridDim.x * 256) {
    if (row2atom[r] < 0) continue;
    v[0] += (double) E[r];
    for (int q = 0; q < 6; q++) v[1 + q] += wrow[(size_t) r * 6 + q];
  }
  for (int q = 0; q < 7; q++) sh[q][t] = v[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int q = 0; q < 7; q++) sh[q][t] += sh[q][t + o];
    __syncthreads();
  }
  if (t < 7) part[(size_t) blockIdx.x * 7 + t] = sh[t][0];
}

__global__ void k_reduce_ev_final(const double *part, int nparts, double scale, double *out)
{
  int t = threadIdx.x;
  if (t < 7) {
    double s = 0.0;
    for (int p = 0; p < nparts; p++) s += part[(size_t) p * 7 + t];
    out[t] = (t == 0) ? s * scale : s;
  }
}

__global__ void k_vv_first(int n, double4 *pos, double *v, const double *f, const double *dtfm,
                           double dt)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 p = pos[i];
  double c = dtfm[(int) p.w];
  double vx = v[3 * i] + c * f[3 * i];
  double vy = v[3 * i + 1] + c * f[3 * i + 1];
  double vz = v[3 * i + 2] + c * f[3 * i + 2];
  v[3 * i] = vx;
  v[3 * i + 1] = vy;
  v[3 * i + 2] = vz;
  p.x += dt * vx;
  p.y += dt * vy;
  p.z += dt * vz;
  pos[i] = p;
}

__global__ void k_vv_second(int n, const double4 *pos, double *v, const double *f, const double *dtfm)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double c = dtfm[(int) pos[i].w];
  v[3 * i] += c * f[3 * i];
  v[3 * i + 1] += c * f[3 * i + 1];
  v[3 * i + 2] += c * f[3 * i + 2];
}

__global__ void k_ekin_part(int n, const double4 *pos, const double *v, const double *mass,
                            double *part)
{
  __shared__ double sh[256];
  int t = threadIdx.x;
  double s = 0.0;
  for (int i = blockIdx.x * 256 + t; i < n; i += gridDim.x * 256) {
    double m = mass[(int) pos[i].w];
    s += m * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  }
  sh[t] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_ekin_final(const double *part, int nparts, double *out)
{
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int p = 0; p < nparts; p++) s += part[p];
    out[0] = 0.5 * 1.0364269e-4 * s;  // mvv2e (metal units)
  }
}

__global__ void k_copy_scalar(const double *src, double *dst)
{
  if (threadIdx.x == 0) dst[0] = src[0];
}

constexpr int kRedParts = 240;

}  // namespace

hipError_t launch_dense_fwd(int mode, const RowLayout &rl, const LayerParams &lp, const void *A,
                            int a_is_fp32, void *H, float *D, hipStream_t s)
{
  if (mode == 0)
    return fwd_act<0, float, float>(lp.act, rl, lp, (const float *) A, (float *) H, D, s);
  if (a_is_fp32)
    return fwd_act<1, float, uint16_t>(lp.act, rl, lp, (const float *) A, (uint16_t *) H, D, s);
  return fwd_act<1, uint16_t, uint16_t>(lp.act, rl, lp, (const uint16_t *) A, (uint16_t *) H, D, s);
}

hipError_t launch_output(const RowLayout &rl, const LayerParams &lp, const void *H, int h_is_fp32,
                         float *E, hipStream_t s)
{
  int blocks = (rl.nrows + 3) / 4;
  if (h_is_fp32)
    k_output<float><<<blocks, 256, 0, s>>>(rl, lp, (const float *) H, E);
  else
    k_output<uint16_t><<<blocks, 256, 0, s>>>(rl, lp, (const uint16_t *) H, E);
  return hipGetLastError();
}

hipError_t launch_dense_bwd(int mode, const RowLayout &rl, const LayerParams &lp, const void *Gout,
                            const float *gout_vec, int vec_stride, const float *D, void *Gin,
                            int gin_fp32, hipStream_t s)
{
  dim3 grid(rl.nrows / BM, (lp.K + BN - 1) / BN);
  if (mode == 0) {
    if (gout_vec)
      k_dense_bwd<0, 2, float, float><<<grid, 256, 0, s>>>(rl, lp, nullptr, gout_vec, vec_stride, D, (float *) Gin);
    else
      k_dense_bwd<0, 1, float, float><<<grid, 256, 0, s>>>(rl, lp, (const float *) Gout, nullptr, 0, D, (float *) Gin);
  } else {
    // BF16 mode: gradients between layers are kept in FP32 (they are
    // rounded to BF16 only as MFMA operands)
    if (gout_vec)
      k_dense_bwd<1, 2, float, float><<<grid, 256, 0, s>>>(rl, lp, nullptr, gout_vec, vec_stride, D, (float *) Gin);
    else
      k_dense_bwd<1, 1, float, float><<<grid, 256, 0, s>>>(rl, lp, (const float *) Gout, nullptr, 0, D, (float *) Gin);
  }
  (void) gin_fp32;
  return hipGetLastError();
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t s)
{
  if (bytes == 0) return hipSuccess;
  size_t n16 = bytes / 16, tail = bytes - n16 * 16;
  int blocks = (int) std::min<size_t>(4096, (n16 + 255) / 256 + 1);
  k_zero<<<blocks, 256, 0, s>>>((uint4 *) p, n16, (unsigned char *) p + n16 * 16, tail);
  return hipGetLastError();
}

hipError_t launch_reduce_ev(const float *E, const double *wrow, const int *row2atom, int nrows,
                            double scale, double *out, hipStream_t s)
{
  // out must have room for 7 + kRedParts*7 doubles: partials live after out[7]
  double *part = out + 8;
  k_reduce_ev_part<<<kRedParts, 256, 0, s>>>(E, wrow, row2atom, nrows, part);
  k_reduce_ev_final<<<1, 64, 0, s>>>(part, kRedParts, scale, out);
  return hipGetLastError();
}

hipError_t launch_vv_first(int n, double4 *pos, double *v, const double *f, const double *dtfm,
                           double dt, hipStream_t s)
{
  k_vv_first<<<(n + 255) / 256, 256, 0, s>>>(n, pos, v, f, dtfm, dt);
  return hipGetLastError();
}

hipError_t launch_vv_second(int n, const double4 *pos, double *v, const double *f,
                            const double *dtfm, hipStream_t s)
{
  k_vv_second<<<(n + 255) / 256, 256, 0, s>>>(n, pos, v, f, dtfm);
  return hipGetLastError();
}

hipError_t launch_ekin(int n, const double4 *pos, const double *v, const double *mass, double *out,
                       hipStream_t s)
{
  double *part = out + 8;
  k_ekin_part<<<kRedParts, 256, 0, s>>>(n, pos, v, mass, part);
  k_ekin_final<<<1, 64, 0, s>>>(part, kRedParts, out);
  return hipGetLastError();
}

hipError_t launch_copy_scalar(const double *src, double *dst, hipStream_t s)
{
  k_copy_scalar<<<1, 64, 0, s>>>(src, dst);
  return hipGetLastError();
}

}  // namespace tmd
