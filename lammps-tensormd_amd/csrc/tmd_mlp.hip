// Per-element MLP forward + backward (dE/dG) in ONE kernel (gfx950 MFMA).
//
// Replaces forward_resnet_GPU / backward_resnet_GPU (gpu/cuda/nnp.cu:149-281:
// one cuBLAS GEMM + one bias/activation kernel per layer, activations in
// HBM) and the CPU dense path (src/ML-TENSORMD/nnp.cpp:246-366).
//
// A workgroup (4 waves) owns R rows of one element (rows are element-sorted
// and each element group starts at a multiple of 64, so R | 64 keeps a tile
// inside one element).  Every hidden activation H_l and derivative
// D_l = act'(z_l) of the tile stays in LDS; the waves split each layer's
// output columns 16 at a time, reading A from LDS (layer 0: the FP32
// descriptor rows from HBM) and streaming the element's weights from L2
// straight into MFMA B operands.  The backward pass reuses the D_l slots for
// G_l = dE/dH_l * D_l and writes dE/dG (FP32) to HBM.  HBM traffic per atom
// is the descriptor read plus the gradient write.
//
// FP32 mode: v_mfma_f32_16x16x4_f32; BF16 mode: v_mfma_f32_16x16x32_bf16
// (operands rounded to BF16, FP32 accumulation; H/D kept as BF16 in LDS).
// Split-BF16 mode (fast kernel only): every operand x is carried as
// hi = bf16(x) and lo = bf16(x - hi), and each product as the three BF16
// MFMAs hi*hi + hi*lo + lo*hi with FP32 accumulation (the dropped lo*lo term
// and the split leave ~2^-16 relative error per product, against 2^-8 for
// plain BF16); H/D stay FP32 in LDS.
// Operand layouts (16x16 tiles): A lane l holds row l%16, k-group l/16;
// B lane l holds column l%16, k-group l/16; C lane l holds rows
// 4*(l/16)+i, column l%16.
#include "tmd_device.h"

namespace tmd {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16) f); }
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
__device__ inline float bf2f(uint16_t v) { return __uint_as_float(((uint32_t) v) << 16); }

// activation value and derivative (nnp.cu:9-49, activation.hpp:158-261).
// FAST (BF16 and FP8 modes, whose operands are rounded to 8 significant bits
// or fewer): tanh from one exp and one 1-ulp hardware reciprocal,
// tanh z = 1 - 2 / (exp(2z) + 1) (a correctly rounded reciprocal is a
// ten-instruction division sequence).
template <int ACT, bool FAST = false>
__device__ inline float act_f(float z, float &dz)
{
  if (ACT == 2 && FAST) {
    float h = 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * z) + 1.0f);
    dz = 1.0f - h * h;
    return h;
  }
  if (ACT == 1) {  // softplus; linear above -(log(eps_f32)+2) (nnp.cu:44-45)
    float ez = __expf(z);
    float h = (z < 13.942385f) ? log1pf(ez) : z;
    dz = 1.0f / (1.0f + 1.0f / ez);
    return h;
  } else if (ACT == 2) {
    float h = tanhf(z);
    dz = 1.0f - h * h;
    return h;
  } else if (ACT == 3) {
    // squareplus (z + sqrt(z^2 + 4)) / 2; for z < 0 the same value as
    // 2 / (sqrt(z^2 + 4) - z), which does not cancel in FP32
    float s = sqrtf(z * z + 4.0f);
    float h = z >= 0.0f ? 0.5f * (z + s) : 2.0f / (s - z);
    dz = h / s;
    return h;
  }
  dz = z >= 0.0f ? 1.0f : 0.0f;
  return z >= 0.0f ? z : 0.0f;
}

template <int MODE>
struct Traits;
template <>
struct Traits<0> {
  using T = float;
  static constexpr int KS = 4;  // k per MFMA
  static constexpr int CT = 4;  // column tiles per pass per wave
};
template <>
struct Traits<1> {
  using T = uint16_t;
  static constexpr int KS = 32;
  static constexpr int CT = 4;
};
template <>  // FP8 operands, BF16 activations in LDS
struct Traits<2> {
  using T = uint16_t;
  static constexpr int KS = 32;
  static constexpr int CT = 4;
};
template <>  // split-BF16 operands, FP32 activations in LDS
struct Traits<3> {
  using T = float;
  static constexpr int KS = 32;
  static constexpr int CT = 4;
};

template <typename T>
__device__ inline float to_f(T v);
template <>
__device__ inline float to_f<float>(float v) { return v; }
template <>
__device__ inline float to_f<uint16_t>(uint16_t v) { return bf2f(v); }
template <typename T>
__device__ inline T from_f(float v);
template <>
__device__ inline float from_f<float>(float v) { return v; }
template <>
__device__ inline uint16_t from_f<uint16_t>(float v) { return f2bf(v); }

__device__ inline int tile_elem(const RowLayout &rl, int row0)
{
  int e = 0;
  while (e < kMaxElts - 1 && row0 >= rl.row_off[e + 1]) e++;
  return e;
}

// ---- one MFMA step ------------------------------------------------------------
// A operand: from LDS tile (row stride ld) or from the FP32 descriptor rows.
template <int MODE, bool AGLOBAL>
__device__ inline void load_a(const typename Traits<MODE>::T *lds, int ld, const float *G, int gw,
                              int row, int k, int K, float &af, bf16x8 &ab)
{
  const int lane = threadIdx.x & 63;
  if (MODE == 0) {
    int kk = k + (lane >> 4);
    if (AGLOBAL)
      af = (kk < K) ? G[(size_t) row * gw + kk] : 0.0f;
    else
      af = (kk < K) ? ((const float *) lds)[row * ld + kk] : 0.0f;
  } else {
    int kk = k + (lane >> 4) * 8;
    if (AGLOBAL) {
      u16x8 v;
      if (kk + 8 <= K) {
        const float4 *p = reinterpret_cast<const float4 *>(G + (size_t) row * gw + kk);
        float4 x0 = p[0], x1 = p[1];
        f32x8 xv = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
        ab = __builtin_convertvector(xv, bf16x8);
        return;
      } else {
        for (int q = 0; q < 8; q++) v[q] = (kk + q < K) ? f2bf(G[(size_t) row * gw + kk + q]) : 0;
      }
      ab = __builtin_bit_cast(bf16x8, v);
    } else {
      const uint16_t *p = (const uint16_t *) lds + row * ld + kk;
      if (kk + 8 <= K) {
        ab = *reinterpret_cast<const bf16x8 *>(p);
      } else {
        u16x8 v;
        for (int q = 0; q < 8; q++) v[q] = (kk + q < K) ? p[q] : 0;
        ab = __builtin_bit_cast(bf16x8, v);
      }
    }
  }
}

// B operand for output column `col`, inner index k:
//   FP32: B row-major [K][N] (bf[k*N + col]); BF16: B stored [N][K] (k contiguous)
template <int MODE>
__device__ inline void load_b(const void *B, int K, int N, int col, int k, float &bf, bf16x8 &bb)
{
  const int lane = threadIdx.x & 63;
  if (MODE == 0) {
    int kk = k + (lane >> 4);
    bf = (kk < K && col < N) ? ((const float *) B)[(size_t) kk * N + col] : 0.0f;
  } else {
    int kk = k + (lane >> 4) * 8;
    const uint16_t *p = (const uint16_t *) B + (size_t) col * K + kk;
    if (col < N && kk + 8 <= K) {
      bb = *reinterpret_cast<const bf16x8 *>(p);
    } else {
      u16x8 v;
      for (int q = 0; q < 8; q++) v[q] = (col < N && kk + q < K) ? p[q] : 0;
      bb = __builtin_bit_cast(bf16x8, v);
    }
  }
}

// Wave-level GEMM over all R rows for up to CT column tiles (tiles[j] < 0 =
// unused): acc[rt][j] = sum_k A[rt*16.., k] B[k, tiles[j]*16..]
template <int MODE, int R, bool AGLOBAL>
__device__ inline void wave_gemm(const typename Traits<MODE>::T *lds, int ld, const float *G, int gw,
                                 int grow0, const void *B, int K, int N, const int *tiles,
                                 f32x4 acc[R / 16][Traits<MODE>::CT])
{
  constexpr int CT = Traits<MODE>::CT, KS = Traits<MODE>::KS, RT = R / 16;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < RT; r++)
#pragma unroll
    for (int j = 0; j < CT; j++) acc[r][j] = f32x4{0, 0, 0, 0};
  // A and B operands of step k+KS are loaded while step k's MFMAs run
  float bfc[CT], bfn[CT], afc[RT], afn[RT];
  bf16x8 bbc[CT], bbn[CT], abc[RT], abn[RT];
#pragma unroll
  for (int j = 0; j < CT; j++)
    if (tiles[j] >= 0) load_b<MODE>(B, K, N, tiles[j] * 16 + (lane & 15), 0, bfc[j], bbc[j]);
#pragma unroll
  for (int r = 0; r < RT; r++) {
    int row = r * 16 + (lane & 15);
    load_a<MODE, AGLOBAL>(lds, ld, G, gw, AGLOBAL ? grow0 + row : row, 0, K, afc[r], abc[r]);
  }
  for (int k = 0; k < K; k += KS) {
    const bool more = k + KS < K;
    if (more) {
#pragma unroll
      for (int j = 0; j < CT; j++)
        if (tiles[j] >= 0) load_b<MODE>(B, K, N, tiles[j] * 16 + (lane & 15), k + KS, bfn[j], bbn[j]);
#pragma unroll
      for (int r = 0; r < RT; r++) {
        int row = r * 16 + (lane & 15);
        load_a<MODE, AGLOBAL>(lds, ld, G, gw, AGLOBAL ? grow0 + row : row, k + KS, K, afn[r], abn[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < RT; r++) {
#pragma unroll
      for (int j = 0; j < CT; j++) {
        if (tiles[j] < 0) continue;
        if (MODE == 0)
          acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(afc[r], bfc[j], acc[r][j], 0, 0, 0);
        else
          acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abc[r], bbc[j], acc[r][j], 0, 0, 0);
      }
    }
    if (more) {
#pragma unroll
      for (int j = 0; j < CT; j++) {
        bfc[j] = bfn[j];
        bbc[j] = bbn[j];
      }
#pragma unroll
      for (int r = 0; r < RT; r++) {
        afc[r] = afn[r];
        abc[r] = abn[r];
      }
    }
  }
}

}  // namespace

template <int MODE, int ACT, int R>
__global__ __launch_bounds__(256) void k_mlp(RowLayout rl_cap, MlpParams mp, const float *Gd, double *E,
                                             float *dEdG, const BatchDims *dyn, int rbase)
{
  const RowLayout rl = dyn ? dyn->rl : rl_cap;
  using T = typename Traits<MODE>::T;
  constexpr int CT = Traits<MODE>::CT, RT = R / 16;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *sm = reinterpret_cast<T *>(smem_raw);
  const int row0 = rbase + (int) blockIdx.x * R;
  if (row0 >= rl.nrows) return;
  const int e = tile_elem(rl, row0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int gw = mp.sizes[0];
  const int L = mp.nl;

  // ---------------- forward ----------------
  for (int l = 0; l < L - 1; l++) {
    const int K = mp.sizes[l], N = mp.sizes[l + 1];
    const void *B = MODE == 0 ? (const void *) (mp.Wf32[l] + (size_t) e * mp.wstride[l])
                              : (const void *) (mp.Wtb[l] + (size_t) e * mp.wstride[l]);
    const float *bias = mp.b[l] ? mp.b[l] + (size_t) e * mp.bstride[l] : nullptr;
    T *H = sm + mp.offH[l];
    T *D = sm + mp.offD[l];
    const int ldo = mp.ld[l + 1];
    const T *Hin = l > 0 ? sm + mp.offH[l - 1] : nullptr;
    const int ldi = mp.ld[l];
    const int NT = (N + 15) / 16;
    for (int t0 = wave; t0 < NT; t0 += 4 * CT) {
      int tiles[CT];
#pragma unroll
      for (int j = 0; j < CT; j++) tiles[j] = (t0 + 4 * j < NT) ? t0 + 4 * j : -1;
      f32x4 acc[RT][CT];
      if (l == 0)
        wave_gemm<MODE, R, true>(nullptr, 0, Gd, gw, row0, B, K, N, tiles, acc);
      else
        wave_gemm<MODE, R, false>(Hin, ldi, nullptr, gw, row0, B, K, N, tiles, acc);
#pragma unroll
      for (int j = 0; j < CT; j++) {
        if (tiles[j] < 0) continue;
        int col = tiles[j] * 16 + (lane & 15);
        if (col >= N) continue;
        float bv = bias ? bias[col] : 0.0f;
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            int row = r * 16 + (lane >> 4) * 4 + i;
            float dz;
            float h = act_f<ACT, MODE == 1>(acc[r][j][i] + bv, dz);
            if (mp.res[l]) h += to_f<T>(Hin[row * ldi + col]);
            H[row * ldo + col] = from_f<T>(h);
            if (mp.storeD) D[row * ldo + col] = from_f<T>(dz);
          }
      }
    }
    __syncthreads();
  }

  // ---------------- output layer: E = H w + b (FP64 sum) ----------------
  {
    const int K = mp.sizes[L - 1];
    const T *H = sm + mp.offH[L - 2];
    const int ld = mp.ld[L - 1];
    const float *w = mp.wout + (size_t) e * mp.wstride[L - 1];
    const float bo = mp.bout ? mp.bout[(size_t) e * mp.bstride[L - 1]] : 0.0f;
    for (int r = wave; r < R; r += 4) {
      double s = 0.0;
      for (int k = lane; k < K; k += 64) s += (double) (to_f<T>(H[r * ld + k]) * w[k]);
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (lane == 0) E[row0 + r] = s + (double) bo;
    }
  }
  // the seed below may overwrite H_{L-2} (when D aliases H): all output dots first
  __syncthreads();
  // seed: G_{L-2} = w_out * D_{L-2} (in the D slot); gradH = w_out for a resnet layer
  {
    const int N = mp.sizes[L - 1];
    const float *w = mp.wout + (size_t) e * mp.wstride[L - 1];
    T *D = sm + mp.offD[L - 2];
    T *gH = sm + mp.offGH;
    const int ld = mp.ld[L - 1];
    for (int q = threadIdx.x; q < R * N; q += 256) {
      int r = q / N, c = q - r * N;
      float v = to_f<T>(D[r * ld + c]);  // D, or H when D is not stored (tanh: D = 1 - h^2)
      D[r * ld + c] = from_f<T>(w[c] * (mp.storeD ? v : 1.0f - v * v));
      if (mp.res[L - 2]) gH[r * mp.ldGH + c] = from_f<T>(w[c]);
    }
  }
  __syncthreads();

  // ---------------- backward ----------------
  for (int l = L - 2; l >= 0; l--) {
    const int K = mp.sizes[l], N = mp.sizes[l + 1];
    // Gin[R x K] = G_l[R x N] W_l^T : inner N, output columns K
    const void *B = MODE == 0 ? (const void *) (mp.Wt32[l] + (size_t) e * mp.wstride[l])
                              : (const void *) (mp.Wb16[l] + (size_t) e * mp.wstride[l]);
    const T *Gl = sm + mp.offD[l];
    const int ldg = mp.ld[l + 1];
    T *gH = sm + mp.offGH;
    const int NT = (K + 15) / 16;
    for (int t0 = wave; t0 < NT; t0 += 4 * CT) {
      int tiles[CT];
#pragma unroll
      for (int j = 0; j < CT; j++) tiles[j] = (t0 + 4 * j < NT) ? t0 + 4 * j : -1;
      f32x4 acc[RT][CT];
      wave_gemm<MODE, R, false>(Gl, ldg, nullptr, gw, row0, B, N, K, tiles, acc);
#pragma unroll
      for (int j = 0; j < CT; j++) {
        if (tiles[j] < 0) continue;
        int col = tiles[j] * 16 + (lane & 15);
        if (col >= K) continue;
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            int row = r * 16 + (lane >> 4) * 4 + i;
            float v = acc[r][j][i];
            if (mp.res[l]) v += to_f<T>(gH[row * mp.ldGH + col]);  // dE/dH_l skip path
            if (l == 0) {
              dEdG[(size_t) (row0 + row) * gw + col] = v;
            } else {
              T *D = sm + mp.offD[l - 1];
              const int ld = mp.ld[l];
              float dv = to_f<T>(D[row * ld + col]);
              D[row * ld + col] = from_f<T>(v * (mp.storeD ? dv : 1.0f - dv * dv));
              if (mp.res[l - 1]) gH[row * mp.ldGH + col] = from_f<T>(v);
            }
          }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fast path, used when every hidden width is a multiple of 16 and every GEMM
// inner dimension a multiple of 2 KS (all production models).  The weights
// are the MFMA A operand and the activations the B operand, so the
// accumulator of a lane holds 4 consecutive output columns of one row: the
// epilogue reads and writes them with one vector access.  Operands are
// double-buffered in registers (no copies) and loaded without bounds checks.
namespace {

template <int MODE>
struct Frag;
// FP32: four consecutive inner values per lane, consumed by four 16x16x4
// MFMAs (lane group g holds inner values 4g .. 4g+3 of a block of 16: the
// MFMA q sums the values 4g + q, so the four cover the block)
template <>
struct Frag<0> {
  using A = f32x4;
};
// inner values per fast-path step: 16 (FP32: four MFMAs), 32 (BF16), or
// 128 (FP8: one K=128 MFMA)
template <int MODE>
constexpr int kFastKS = MODE == 0 ? 16 : (MODE == 2 ? 128 : 32);
template <>
struct Frag<1> {
  using A = bf16x8;
};
template <>
struct Frag<2> {
  using A = i32x8;  // 32 FP8 e4m3 values
};
// split BF16: the high and low parts of 8 values
struct SplitBf16 {
  bf16x8 hi, lo;
};
template <>
struct Frag<3> {
  using A = SplitBf16;
};

__device__ inline SplitBf16 split8(const f32x8 &x)
{
  SplitBf16 r;
  r.hi = __builtin_convertvector(x, bf16x8);
  r.lo = __builtin_convertvector(x - __builtin_convertvector(r.hi, f32x8), bf16x8);
  return r;
}

// 8 floats -> 8 FP8 e4m3 (round to nearest even), first value in the low byte
__device__ inline long pack_fp8(const float *v)
{
  int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
  lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
  int hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4], v[5], 0, false);
  hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6], v[7], hi, true);
  return (long) (((unsigned long long) (unsigned) hi << 32) | (unsigned) lo);
}

// A operand (weights) of output tile t, inner step s (of nst), stored in
// fragment order (MlpParams::Af/Ab): one contiguous block per wave.
template <int MODE>
__device__ inline typename Frag<MODE>::A ld_a(const void *A, int t, int s, int nst)
{
  const int lane = threadIdx.x & 63;
  const size_t idx = ((size_t) t * nst + s) * 64 + lane;
  if constexpr (MODE == 0)
    return reinterpret_cast<const f32x4 *>(A)[idx];
  else if constexpr (MODE == 1)
    return reinterpret_cast<const bf16x8 *>(A)[idx];
  else if constexpr (MODE == 3)
    return reinterpret_cast<const SplitBf16 *>(A)[idx];
  else
    return reinterpret_cast<const i32x8 *>(A)[idx];
}

// B operand (activations of rows r0..r0+15) from a row-major LDS tile;
// FP8: scaled by the row's inverse scale ri
template <int MODE, typename T>
__device__ inline typename Frag<MODE>::A ld_x(const T *X, int ld, int r0, int k, float ri)
{
  const int lane = threadIdx.x & 63;
  if constexpr (MODE == 0) {
    return *reinterpret_cast<const f32x4 *>(X + (r0 + (lane & 15)) * ld + k + (lane >> 4) * 4);
  } else if constexpr (MODE == 1) {
    return *reinterpret_cast<const bf16x8 *>(X + (r0 + (lane & 15)) * ld + k + (lane >> 4) * 8);
  } else if constexpr (MODE == 3) {
    return split8(*reinterpret_cast<const f32x8 *>(X + (r0 + (lane & 15)) * ld + k + (lane >> 4) * 8));
  } else {  // FP8: 32 bytes of the quantized rows (X is the byte buffer Q)
    const uint4 *p = reinterpret_cast<const uint4 *>(reinterpret_cast<const uint8_t *>(X) +
                                                     (r0 + (lane & 15)) * ld + k + (lane >> 4) * 32);
    const uint4 u0 = p[0], u1 = p[1];
    return i32x8{(int) u0.x, (int) u0.y, (int) u0.z, (int) u0.w, (int) u1.x, (int) u1.y, (int) u1.z, (int) u1.w};
  }
}

// 8 consecutive values of row r at column c: the FP32 descriptor rows in
// HBM (XG) or a BF16 LDS tile
template <bool XG>
__device__ inline void ld8_row(const uint16_t *X, int ld, const float *G, int gw, int grow0, int r, int c,
                               float v[8])
{
  if constexpr (XG) {
    const float4 *p = reinterpret_cast<const float4 *>(G + (size_t) (grow0 + r) * gw + c);
    const float4 a = p[0], b = p[1];
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
  } else {
    const uint4 u = *reinterpret_cast<const uint4 *>(X + r * ld + c);
    v[0] = __uint_as_float(u.x << 16), v[1] = __uint_as_float(u.x & 0xffff0000u);
    v[2] = __uint_as_float(u.y << 16), v[3] = __uint_as_float(u.y & 0xffff0000u);
    v[4] = __uint_as_float(u.z << 16), v[5] = __uint_as_float(u.z & 0xffff0000u);
    v[6] = __uint_as_float(u.w << 16), v[7] = __uint_as_float(u.w & 0xffff0000u);
  }
}

// FP8, layer 0: the row scales of the descriptor rows and their quantized
// copy in one read of HBM: each thread holds its up to 8 x 8 values of the
// row in registers (all loads in flight) while the threads of the row find
// the maximum.  Returns false (nothing done) when a row does not fit.
__device__ inline bool quantize_g_rows(const float *G, int gw, int grow0, int width, int R, float *rs, uint8_t *Q,
                                       int ldq)
{
  constexpr int kMaxChunks = 8;
  const int tpr = 256 / R;
  if (width > kMaxChunks * tpr * 8) return false;
  const int r = threadIdx.x / tpr, part = threadIdx.x - r * tpr;
  float v[kMaxChunks][8];
  float m = 0.0f;
#pragma unroll
  for (int i = 0; i < kMaxChunks; i++) {
    const int c = (part + i * tpr) * 8;
    if (c < width) {
      ld8_row<true>(nullptr, 0, G, gw, grow0, r, c, v[i]);
#pragma unroll
      for (int j = 0; j < 8; j++) m = fmaxf(m, fabsf(v[i][j]));
    }
  }
  for (int o = 1; o < tpr; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
  const float sc = m > 0.0f ? m * (1.0f / 440.0f) : 1.0f, ri = 1.0f / sc;
  if (part == 0) {
    rs[r] = sc;
    rs[R + r] = ri;
  }
#pragma unroll
  for (int i = 0; i < kMaxChunks; i++) {
    const int c = (part + i * tpr) * 8;
    if (c < width) {
#pragma unroll
      for (int j = 0; j < 8; j++) v[i][j] *= ri;
      *reinterpret_cast<long *>(Q + r * ldq + c) = pack_fp8(v[i]);
    }
  }
  return true;
}

// FP8: the BF16 LDS rows [0, R) of width w quantized into Q with the row
// scales taken from the maxima that the producing epilogue recorded (rin,
// float bits; |x| >= 0 orders as unsigned); rs receives the scales and their
// inverses for the GEMM's epilogue, and rnext (the other buffer, written by
// this GEMM's epilogue) is cleared.
__device__ inline void quantize_rows_max(const uint16_t *X, int ld, int w, int R, const unsigned *rin,
                                         unsigned *rnext, float *rs, uint8_t *Q, int ldq)
{
  const int tpr = 256 / R;
  const int r = threadIdx.x / tpr, part = threadIdx.x - r * tpr;
  const float m = __uint_as_float(rin[r]);
  const float sc = m > 0.0f ? m * (1.0f / 440.0f) : 1.0f, ri = 1.0f / sc;
  if (part == 0) {
    rs[r] = sc;
    rs[R + r] = ri;
    rnext[r] = 0u;
  }
#pragma unroll 4
  for (int c = part * 8; c < w; c += tpr * 8) {
    float v[8];
    ld8_row<false>(X, ld, nullptr, 0, 0, r, c, v);
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] *= ri;
    *reinterpret_cast<long *>(Q + r * ldq + c) = pack_fp8(v);
  }
}

__device__ inline float max_abs4(const float v[4])
{
  return fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3])));
}

// A wave's per-lane row maxima (row r * 16 + lane % 16 of each row tile r)
// into the recorded maxima: the four lanes of a row combine first, then one
// LDS atomic per row and wave.
template <int RT>
__device__ inline void flush_max(unsigned *rmax, const float (&mx)[RT])
{
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < RT; r++) {
    float m = fmaxf(mx[r], __shfl_xor(mx[r], 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    if (lane < 16) atomicMax(rmax + r * 16 + lane, __float_as_uint(m));
  }
}

// FP8: rows [0, R) of width w (a multiple of 8) divided by their row scale
// (rinv = 1 / scale) and rounded to e4m3 into Q [R][ldq], once per GEMM, so
// that the GEMM's B operands are plain 32-byte LDS reads.  The source is
// the BF16 LDS tile X or, with XG, the FP32 descriptor rows in HBM.
template <bool XG>
__device__ inline void quantize_rows(const uint16_t *X, int ld, const float *G, int gw, int grow0, int w, int R,
                                     const float *rinv, uint8_t *Q, int ldq)
{
  const int tpr = 256 / R;
  const int r = threadIdx.x / tpr, part = threadIdx.x - r * tpr;
  const float ri = rinv[r];
#pragma unroll 4
  for (int c = part * 8; c < w; c += tpr * 8) {
    float v[8];
    ld8_row<XG>(X, ld, G, gw, grow0, r, c, v);
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] *= ri;
    *reinterpret_cast<long *>(Q + r * ldq + c) = pack_fp8(v);
  }
}

// B operand from the FP32 descriptor rows in HBM (layer 0), in two parts so
// that a prefetch ring holds the raw values and converts them only at use:
// ld_graw loads 8 (4 for FP32) values of row grow + lane%16, cvt_graw makes
// the MFMA operand (BF16; FP8 after scaling by the row's inverse scale ri).
template <int MODE>
struct GRaw {
  float4 lo, hi;
};
template <>
struct GRaw<0> {
  f32x4 v;
};

// GB: the descriptor rows are stored as BF16 (the BF16 mode's operand
// values, rounded once by the descriptor kernel), read 16 bytes per lane
template <int MODE, bool GB = false>
__device__ inline GRaw<MODE> ld_graw(const float *G, int gw, int grow, int k)
{
  const int lane = threadIdx.x & 63;
  GRaw<MODE> g;
  if constexpr (GB) {
    static_assert(MODE == 1, "BF16 descriptor rows feed the BF16 mode only");
    const uint16_t *pb = reinterpret_cast<const uint16_t *>(G) + (size_t) (grow + (lane & 15)) * gw + k;
    g.lo = *reinterpret_cast<const float4 *>(pb + (lane >> 4) * 8);
    return g;
  }
  const float *p = G + (size_t) (grow + (lane & 15)) * gw + k;
  if constexpr (MODE == 0) {
    g.v = *reinterpret_cast<const f32x4 *>(p + (lane >> 4) * 4);
  } else {
    const float4 *q = reinterpret_cast<const float4 *>(p + (lane >> 4) * 8);
    g.lo = q[0];
    g.hi = q[1];
  }
  return g;
}

template <int MODE, bool GB = false>
__device__ inline typename Frag<MODE>::A cvt_graw(const GRaw<MODE> &g, float ri)
{
  if constexpr (GB) {
    return __builtin_bit_cast(bf16x8, g.lo);
  } else if constexpr (MODE == 0) {
    return g.v;
  } else if constexpr (MODE == 1) {
    f32x8 xv = {g.lo.x, g.lo.y, g.lo.z, g.lo.w, g.hi.x, g.hi.y, g.hi.z, g.hi.w};
    return __builtin_convertvector(xv, bf16x8);
  } else if constexpr (MODE == 3) {
    f32x8 xv = {g.lo.x, g.lo.y, g.lo.z, g.lo.w, g.hi.x, g.hi.y, g.hi.z, g.hi.w};
    return split8(xv);
  } else {
    static_assert(MODE != 2, "FP8 layer-0 operands are quantized into LDS first");
    return {};
  }
}

template <int MODE>
__device__ inline f32x4 mma(typename Frag<MODE>::A a, typename Frag<MODE>::A b, f32x4 c)
{
  if constexpr (MODE == 0) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], b[2], c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], b[3], c, 0, 0, 0);
  } else if constexpr (MODE == 1) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  } else if constexpr (MODE == 3) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
  } else {
    // e4m3 A and B (format 0), unit E8M0 block scales (127 = 2^0): the
    // per-channel and per-row FP32 scales are applied in the epilogue
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, 0x7f7f7f7f, 0, 0x7f7f7f7f);
  }
}

// FP8: per-row scale of the next GEMM's B operand (the rows of X, or of the
// descriptor rows G): rs[r] = max_c |x[r][c]| / 440 (1 for a zero row),
// rs[R + r] = 1 / rs[r]; every value then fits e4m3 (|x| <= 448).
// Threads per row (256 / R consecutive threads) each take 8 consecutive
// values at a time; the loops are unrolled so that several loads are in
// flight (the descriptor rows come from HBM).
template <bool XG>
__device__ inline void row_scales(const uint16_t *X, int ld, const float *G, int gw, int grow0,
                                  int width, int R, float *rs)
{
  const int tpr = 256 / R;  // threads per row: 16, 8 or 4 consecutive lanes
  const int r = threadIdx.x / tpr, part = threadIdx.x - r * tpr;
  float m = 0.0f;
#pragma unroll 4
  for (int c = part * 8; c < width; c += tpr * 8) {
    float v[8];
    ld8_row<XG>(X, ld, G, gw, grow0, r, c, v);
#pragma unroll
    for (int i = 0; i < 8; i++) m = fmaxf(m, fabsf(v[i]));
  }
  for (int o = 1; o < tpr; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (part == 0) {
    const float sc = m > 0.0f ? m * (1.0f / 440.0f) : 1.0f;
    rs[r] = sc;
    rs[R + r] = 1.0f / sc;
  }
}

// 4 consecutive values of a row-major LDS tile (8- or 16-byte aligned)
__device__ inline void ld4(const float *p, float v[4])
{
  float4 x = *reinterpret_cast<const float4 *>(p);
  v[0] = x.x;
  v[1] = x.y;
  v[2] = x.z;
  v[3] = x.w;
}
__device__ inline void ld4(const uint16_t *p, float v[4])
{
  uint2 x = *reinterpret_cast<const uint2 *>(p);
  v[0] = __uint_as_float(x.x << 16);
  v[1] = __uint_as_float(x.x & 0xffff0000u);
  v[2] = __uint_as_float(x.y << 16);
  v[3] = __uint_as_float(x.y & 0xffff0000u);
}
__device__ inline void st4(float *p, const float v[4])
{
  *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ inline void st4(uint16_t *p, const float v[4])
{
  uint2 x;
  x.x = (uint32_t) f2bf(v[0]) | ((uint32_t) f2bf(v[1]) << 16);
  x.y = (uint32_t) f2bf(v[2]) | ((uint32_t) f2bf(v[3]) << 16);
  *reinterpret_cast<uint2 *>(p) = x;
}

// acc[j][r] = 16 output columns (tile t0 + 4 j) x 16 rows (r * 16 ..) of
// W X^T, summed over the inner dimension K, for the wave's NTW tiles.
//
// Global operands (the weights, and the descriptor rows of layer 0) go
// through a register ring D steps deep, D = (fragments budget) / NTW (2 for
// layer 0, whose ring also holds its raw rows), so a load has D - 1 steps of
// MFMA work to arrive in.  Every step issues the
// same number of loads (past the last step they repeat it), which lets the
// compiler wait for exactly the step being consumed.  Activations in LDS
// are read at the step that uses them.
template <int MODE, int R, bool XG, int NTW, typename T, bool GB = false>
__device__ inline void gemm_t(const void *W, int t0, const T *X, int ldx, const float *G, int gw,
                              int grow0, int K, const float *rinv, f32x4 (&acc)[4][R / 16])
{
  constexpr int KS = kFastKS<MODE>, RT = R / 16;
  constexpr int DW = 8 / NTW;
  constexpr int D = XG ? 2 : (DW >= 2 ? DW : 2);  // layer 0 also rings its raw rows
  using F = typename Frag<MODE>::A;
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int r = 0; r < RT; r++) acc[j][r] = f32x4{0, 0, 0, 0};
  float ri[RT];
#pragma unroll
  for (int r = 0; r < RT; r++) ri[r] = MODE == 2 ? rinv[r * 16 + (threadIdx.x & 15)] : 1.0f;
  const int nst = K / KS;
  F a[D][NTW];
  GRaw<MODE> g[XG ? D : 1][RT];
  auto issue = [&](F(&as)[NTW], GRaw<MODE>(&gs)[RT], int s) {
    const int sc = min(s, nst - 1), k = sc * KS;
#pragma unroll
    for (int j = 0; j < NTW; j++) as[j] = ld_a<MODE>(W, t0 + 4 * j, sc, nst);
    if constexpr (XG) {
#pragma unroll
      for (int r = 0; r < RT; r++) gs[r] = ld_graw<MODE, GB>(G, gw, grow0 + r * 16, k);
    }
  };
#pragma unroll
  for (int i = 0; i < D - 1; i++) issue(a[i], g[XG ? i : 0], i);
  for (int s0 = 0; s0 < nst; s0 += D) {
#pragma unroll
    for (int i = 0; i < D; i++) {
      const int nxt = (i + D - 1) % D;
      issue(a[nxt], g[XG ? nxt : 0], s0 + i + D - 1);
      const int s = s0 + i;
      if (s < nst) {
        F b[RT];
#pragma unroll
        for (int r = 0; r < RT; r++) {
          if constexpr (XG)
            b[r] = cvt_graw<MODE, GB>(g[i][r], ri[r]);
          else
            b[r] = ld_x<MODE>(X, ldx, r * 16, s * KS, ri[r]);
        }
#pragma unroll
        for (int j = 0; j < NTW; j++)
#pragma unroll
          for (int r = 0; r < RT; r++) acc[j][r] = mma<MODE>(a[i][j], b[r], acc[j][r]);
      }
    }
  }
}

// gemm_t for a runtime tile count ntw in 1..4 (wave-uniform)
template <int MODE, int R, bool XG, typename T, bool GB = false>
__device__ inline void gemm_n(const void *W, int t0, int ntw, const T *X, int ldx, const float *G,
                              int gw, int grow0, int K, const float *rinv, f32x4 (&acc)[4][R / 16])
{
  if (ntw >= 4)
    gemm_t<MODE, R, XG, 4, T, GB>(W, t0, X, ldx, G, gw, grow0, K, rinv, acc);
  else if (ntw == 3)
    gemm_t<MODE, R, XG, 3, T, GB>(W, t0, X, ldx, G, gw, grow0, K, rinv, acc);
  else if (ntw == 2)
    gemm_t<MODE, R, XG, 2, T, GB>(W, t0, X, ldx, G, gw, grow0, K, rinv, acc);
  else
    gemm_t<MODE, R, XG, 1, T, GB>(W, t0, X, ldx, G, gw, grow0, K, rinv, acc);
}

}  // namespace

// Waves per SIMD asked of the fast kernel (more tiles in flight hide the L2
// latency of the weight operands): 4 for BF16 and 16-row tiles, 3 for FP32
// and FP8 32-row tiles, 2 for 64-row tiles (their accumulators and rings
// need up to 256 VGPRs).
template <int MODE, int R>
constexpr int kMlpWaves = R == 64 ? 2 : ((MODE == 1 || R == 16) ? 4 : 3);

// bytes per weight of the fast kernel's A operands
template <int MODE>
constexpr size_t kAsz = MODE == 0 ? 4 : (MODE == 1 ? 2 : (MODE == 2 ? 1 : 4));

// Layer widths known at compile time (SH > 0): the layer loops unroll and
// every tile count, inner step count and ring shape is a constant.
template <int SH>
struct MlpShape {
  static constexpr int nl = 0;
  static constexpr int w(int) { return 0; }
};
template <>
struct MlpShape<1> {  // 384-256-128-64-1
  static constexpr int nl = 4;
  static constexpr int w(int l) { return l == 0 ? 384 : (l == 1 ? 256 : (l == 2 ? 128 : (l == 3 ? 64 : 1))); }
};
template <>
struct MlpShape<2> {  // 384-256-256-128-1
  static constexpr int nl = 4;
  static constexpr int w(int l) { return l == 0 ? 384 : (l == 1 ? 256 : (l == 2 ? 256 : (l == 3 ? 128 : 1))); }
};

// GB: Gd holds BF16 descriptor rows (MODE 1 only)
template <int MODE, int ACT, int R, int SH = 0, bool GB = false>
__global__ __launch_bounds__(256, (kMlpWaves<MODE, R>)) void k_mlp_fast(RowLayout rl_cap, MlpParams mp, const float *Gd,
                                                  double *E, float *dEdG, const BatchDims *dyn, int rbase)
{
  const RowLayout rl = dyn ? dyn->rl : rl_cap;
  using T = typename Traits<MODE>::T;
  constexpr int RT = R / 16;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *sm = reinterpret_cast<T *>(smem_raw);
  const int row0 = rbase + (int) blockIdx.x * R;
  if (row0 >= rl.nrows) return;
  const int e = tile_elem(rl, row0);
  // wave index made wave-uniform for the compiler: tile bounds and guards
  // derived from it are scalar branches
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  auto width = [&](int l) { return SH ? MlpShape<SH>::w(l) : mp.sizes[l]; };
  constexpr int kLayerUnroll = SH ? 8 : 1;
  const int gw = width(0);
  const int L = SH ? MlpShape<SH>::nl : mp.nl;
  const int cq = 4 * (lane >> 4);  // first of the lane's 4 output columns in a tile
  const int rq = lane & 15;        // the lane's row in a row tile
  float *rs = reinterpret_cast<float *>(smem_raw + mp.offRS);  // FP8 row scales [2][R]
  uint8_t *Q = smem_raw + mp.offQ;                              // FP8 B operand rows [R][ldq]
  // FP8 row maxima [2][R]: GEMM g (forward layer l: g = l; backward layer
  // l: g = 2 L - 3 - l) reads buffer g % 2, its epilogue fills the other
  unsigned *rmax = reinterpret_cast<unsigned *>(smem_raw + mp.offRS) + 2 * R;
  if constexpr (MODE == 2) {
    if (threadIdx.x < 2 * R) rmax[threadIdx.x] = 0u;
  }

  // ---------------- forward ----------------
#pragma unroll kLayerUnroll
  for (int l = 0; l < L - 1; l++) {
    const int K = width(l), N = width(l + 1);
    constexpr size_t esz = kAsz<MODE>;
    const void *W = (const unsigned char *) mp.Af[l] + (size_t) e * mp.wstride[l] * esz;
    const float *bias = mp.b[l] ? mp.b[l] + (size_t) e * mp.bstride[l] : nullptr;
    T *H = sm + mp.offH[l];
    T *D = sm + mp.offD[l];
    const int ldo = mp.ld[l + 1];
    const T *Hin = l > 0 ? sm + mp.offH[l - 1] : nullptr;
    const int ldi = mp.ld[l];
    const int NT = N / 16;
    unsigned *rnext = rmax + ((l + 1) & 1) * R;  // FP8: maxima of this layer's output
    if constexpr (MODE == 2) {
      // layer 0: a pass over the descriptor rows; later layers: the maxima
      // the previous epilogue recorded
      if (l > 0) {
        quantize_rows_max((const uint16_t *) Hin, ldi, K, R, rmax + (l & 1) * R, rnext, rs, Q, mp.ldq);
      } else if (!quantize_g_rows(Gd, gw, row0, K, R, rs, Q, mp.ldq)) {
        row_scales<true>(nullptr, 0, Gd, gw, row0, K, R, rs);
        __syncthreads();
        quantize_rows<true>(nullptr, 0, Gd, gw, row0, K, R, rs + R, Q, mp.ldq);
      }
      __syncthreads();
    }
    const float *wsc = MODE == 2 ? mp.swf[l] + (size_t) e * N : nullptr;
    for (int tb = 0; tb < NT; tb += 16) {
      const int t0 = tb + wave;
      const int ntw = t0 < NT ? min(4, (NT - t0 + 3) / 4) : 0;
      if (ntw == 0) continue;
      f32x4 acc[4][RT];
      float mx[RT];  // FP8: the lane's largest |h| per row tile
#pragma unroll
      for (int r = 0; r < RT; r++) mx[r] = 0.0f;
      if constexpr (MODE == 2)
        gemm_n<MODE, R, false, uint8_t>(W, t0, ntw, Q, mp.ldq, nullptr, gw, row0, K, rs + R, acc);
      else if (l == 0)
        gemm_n<MODE, R, true, T, GB>(W, t0, ntw, nullptr, 0, Gd, gw, row0, K, rs + R, acc);
      else
        gemm_n<MODE, R, false, T>(W, t0, ntw, Hin, ldi, nullptr, gw, row0, K, rs + R, acc);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (j >= ntw) continue;
        const int c0 = (t0 + 4 * j) * 16 + cq;
        float bv[4] = {0.f, 0.f, 0.f, 0.f};
        if (bias) ld4(bias + c0, bv);
        float ws[4] = {1.f, 1.f, 1.f, 1.f};
        if (MODE == 2) ld4(wsc + c0, ws);
#pragma unroll
        for (int r = 0; r < RT; r++) {
          const int row = r * 16 + rq;
          const float sr = MODE == 2 ? rs[row] : 1.0f;
          float h[4], dz[4];
#pragma unroll
          for (int i = 0; i < 4; i++)
            h[i] = act_f<ACT, MODE == 1 || MODE == 2>(acc[j][r][i] * (sr * ws[i]) + bv[i], dz[i]);
          if (mp.res[l]) {
            float x[4];
            ld4(Hin + row * ldi + c0, x);
#pragma unroll
            for (int i = 0; i < 4; i++) h[i] += x[i];
          }
          st4(H + row * ldo + c0, h);
          if (mp.storeD) st4(D + row * ldo + c0, dz);
          if constexpr (MODE == 2) mx[r] = fmaxf(mx[r], max_abs4(h));
        }
      }
      if constexpr (MODE == 2) {
        if (l < L - 2) flush_max<RT>(rnext, mx);
      }
    }
    __syncthreads();
  }

  // ---------------- output layer: E = H w + b (FP64 sum) ----------------
  {
    const int K = width(L - 1);
    const T *H = sm + mp.offH[L - 2];
    const int ld = mp.ld[L - 1];
    const float *w = mp.wout + (size_t) e * mp.wstride[L - 1];
    const float bo = mp.bout ? mp.bout[(size_t) e * mp.bstride[L - 1]] : 0.0f;
    for (int r = wave; r < R; r += 4) {
      double s = 0.0;
      for (int k = lane; k < K; k += 64) s += (double) (to_f<T>(H[r * ld + k]) * w[k]);
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (lane == 0) E[row0 + r] = s + (double) bo;
    }
  }
  __syncthreads();
  // seed: G_{L-2} = w_out * D_{L-2} (in the D slot); dE/dH = w_out for a resnet layer
  {
    const int N = width(L - 1);
    const float *w = mp.wout + (size_t) e * mp.wstride[L - 1];
    T *D = sm + mp.offD[L - 2];
    T *gH = sm + mp.offGH;
    const int ld = mp.ld[L - 1];
    for (int q = threadIdx.x; q < R * N; q += 256) {
      int r = q / N, c = q - r * N;
      float v = to_f<T>(D[r * ld + c]);
      D[r * ld + c] = from_f<T>(w[c] * (mp.storeD ? v : 1.0f - v * v));
      if (mp.res[L - 2]) gH[r * mp.ldGH + c] = from_f<T>(w[c]);
    }
  }
  __syncthreads();

  // ---------------- backward ----------------
#pragma unroll kLayerUnroll
  for (int l = L - 2; l >= 0; l--) {
    const int K = width(l), N = width(l + 1);
    // Gin^T [K x R] = W_l [K x N] G_l^T [N x R]
    constexpr size_t esz = kAsz<MODE>;
    const void *W = (const unsigned char *) mp.Ab[l] + (size_t) e * mp.wstride[l] * esz;
    const T *Gl = sm + mp.offD[l];
    const int ldg = mp.ld[l + 1];
    T *gH = sm + mp.offGH;
    const int NT = K / 16;
    const int g = 2 * L - 3 - l;                  // FP8: this GEMM's index
    unsigned *rnext = rmax + ((g + 1) & 1) * R;  // FP8: maxima of G_{l-1}
    if constexpr (MODE == 2) {
      // the seeded gradient: a pass over its rows; later ones: the maxima
      // the previous backward epilogue recorded
      if (l == L - 2) {
        row_scales<false>((const uint16_t *) Gl, ldg, nullptr, 0, 0, N, R, rs);
        if (threadIdx.x < R) rnext[threadIdx.x] = 0u;
        __syncthreads();
        quantize_rows<false>((const uint16_t *) Gl, ldg, nullptr, 0, 0, N, R, rs + R, Q, mp.ldq);
      } else {
        quantize_rows_max((const uint16_t *) Gl, ldg, N, R, rmax + (g & 1) * R, rnext, rs, Q, mp.ldq);
      }
      __syncthreads();
    }
    const float *wsc = MODE == 2 ? mp.swb[l] + (size_t) e * K : nullptr;
    for (int tb = 0; tb < NT; tb += 16) {
      const int t0 = tb + wave;
      const int ntw = t0 < NT ? min(4, (NT - t0 + 3) / 4) : 0;
      if (ntw == 0) continue;
      f32x4 acc[4][RT];
      float mx[RT];  // FP8: the lane's largest |G| per row tile
#pragma unroll
      for (int r = 0; r < RT; r++) mx[r] = 0.0f;
      if constexpr (MODE == 2)
        gemm_n<MODE, R, false, uint8_t>(W, t0, ntw, Q, mp.ldq, nullptr, gw, row0, N, rs + R, acc);
      else
        gemm_n<MODE, R, false, T>(W, t0, ntw, Gl, ldg, nullptr, gw, row0, N, rs + R, acc);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (j >= ntw) continue;
        const int c0 = (t0 + 4 * j) * 16 + cq;
        float ws[4] = {1.f, 1.f, 1.f, 1.f};
        if (MODE == 2) ld4(wsc + c0, ws);
#pragma unroll
        for (int r = 0; r < RT; r++) {
          const int row = r * 16 + rq;
          const float sr = MODE == 2 ? rs[row] : 1.0f;
          float v[4];
#pragma unroll
          for (int i = 0; i < 4; i++) v[i] = acc[j][r][i] * (sr * ws[i]);
          if (mp.res[l]) {  // dE/dH_l skip path
            float x[4];
            ld4(gH + row * mp.ldGH + c0, x);
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] += x[i];
          }
          if (l == 0) {
            *reinterpret_cast<float4 *>(dEdG + (size_t) (row0 + row) * gw + c0) =
                make_float4(v[0], v[1], v[2], v[3]);
          } else {
            T *D = sm + mp.offD[l - 1];
            const int ld = mp.ld[l];
            float d[4], o[4];
            ld4(D + row * ld + c0, d);
#pragma unroll
            for (int i = 0; i < 4; i++) o[i] = v[i] * (mp.storeD ? d[i] : 1.0f - d[i] * d[i]);
            st4(D + row * ld + c0, o);
            if constexpr (MODE == 2) mx[r] = fmaxf(mx[r], max_abs4(o));
            if (mp.res[l - 1]) st4(gH + row * mp.ldGH + c0, v);
          }
        }
      }
      if constexpr (MODE == 2) {
        if (l > 0) flush_max<RT>(rnext, mx);
      }
    }
    __syncthreads();
  }
}

namespace {
// the compile-time shape of the model's widths, 0 if none
int mlp_shape(const MlpParams &mp)
{
  for (int sh = 1; sh <= 2; sh++) {
    bool eq = mp.nl == 4;
    for (int l = 0; eq && l <= 4; l++)
      eq = mp.sizes[l] == (sh == 1 ? MlpShape<1>::w(l) : MlpShape<2>::w(l));
    if (eq) return sh;
  }
  return 0;
}

// k_mlp_fast with the model's compile-time shape where one exists (tanh
// models at each mode's default tile), else the generic widths.  Measured
// (128k- and 1M-atom systems): BF16 MLP -8% and -7%, split BF16 -5%, FP8
// -36% (78 instead of 128 VGPRs), FP32 -2.5% on the first shape; FP32 on
// the second shape spills (168 VGPRs) and runs 8% slower, so it keeps the
// generic kernel.
template <int MODE, int A, int R>
hipError_t launch_fast(dim3 grid, size_t lds, hipStream_t s, const RowLayout &rl, const MlpParams &mp,
                       const float *G, double *E, float *dEdG, const BatchDims *dyn, int rb, bool gb = false)
{
  auto go = [&](auto kern) -> hipError_t {
    if (hipError_t ae = hipFuncSetAttribute((const void *) kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
        ae != hipSuccess)
      return ae;
    kern<<<grid, 256, lds, s>>>(rl, mp, G, E, dEdG, dyn, rb);
    return hipGetLastError();
  };
  if constexpr (MODE == 1 && A == 2 && R == 32) {
    if (gb) {
      const int sh = mlp_shape(mp);
      if (sh == 1) return go(k_mlp_fast<MODE, A, R, 1, true>);
      if (sh == 2) return go(k_mlp_fast<MODE, A, R, 2, true>);
      return hipErrorInvalidValue;
    }
  }
  if (gb) return hipErrorInvalidValue;
  if constexpr (A == 2 && R == (MODE == 2 ? 16 : 32)) {
    const int sh = mlp_shape(mp);
    if (sh == 1) return go(k_mlp_fast<MODE, A, R, 1>);
    if (sh == 2 && MODE != 0) return go(k_mlp_fast<MODE, A, R, 2>);
  }
  return go(k_mlp_fast<MODE, A, R, 0>);
}

template <int MODE, int R>
hipError_t launch_r(const RowLayout &rl, const MlpParams &mp, const float *G, double *E, float *dEdG,
                    size_t lds, hipStream_t s, const BatchDims *dyn, int rb, int re, bool gb = false)
{
  if (rb % R || re % R) return hipErrorInvalidValue;
  dim3 grid((re - rb) / R);
  const bool fast = mlp_fast_ok(mp, MODE);
  if constexpr (MODE == 2 || MODE == 3) {  // FP8 and split BF16 exist on the fast path only
    if (!fast) return hipErrorInvalidConfiguration;
#define TMD_MLP8_CASE(A)                                                                         \
  case A:                                                                                        \
    return launch_fast<MODE, A, R>(grid, lds, s, rl, mp, G, E, dEdG, dyn, rb, gb);
    switch (mp.act) {
      TMD_MLP8_CASE(0)
      TMD_MLP8_CASE(1)
      TMD_MLP8_CASE(2)
      default:
        TMD_MLP8_CASE(3)
    }
#undef TMD_MLP8_CASE
    return hipGetLastError();
  } else {
#define TMD_MLP_CASE(A)                                                                          \
  case A:                                                                                        \
    if (fast) {                                                                                  \
      return launch_fast<MODE, A, R>(grid, lds, s, rl, mp, G, E, dEdG, dyn, rb, gb);             \
    } else {                                                                                     \
      if (gb) return hipErrorInvalidValue;                                                       \
      if (hipError_t ae = hipFuncSetAttribute((const void *) k_mlp<MODE, A, R>,                  \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);            \
          ae != hipSuccess)                                                                      \
        return ae;                                                                               \
      k_mlp<MODE, A, R><<<grid, 256, lds, s>>>(rl, mp, G, E, dEdG, dyn, rb);                         \
    }                                                                                            \
    break;
    switch (mp.act) {
      TMD_MLP_CASE(0)
      TMD_MLP_CASE(1)
      TMD_MLP_CASE(2)
      default:
        TMD_MLP_CASE(3)
    }
#undef TMD_MLP_CASE
    return hipGetLastError();
  }
}

__global__ void k_quant_fp8(const float *W, const float *swf, const float *swb, uint8_t *Af,
                            uint8_t *Ab, int K, int N, int nelt)
{
  const size_t per = (size_t) K * N;
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per * nelt) return;
  const int el = (int) (i / per);
  const size_t rem = i - (size_t) el * per;
  const int k = (int) (rem / N), n = (int) (rem - (size_t) k * N);
  const float w = W[i];
  // forward operand: output n (tile n/16, lane row n%16), inner k (step
  // k/128, lane group (k%128)/32, byte k%32); backward: output k, inner n.
  // The B operands (quantize_rows, ld_x) place inner index k of a step at
  // the same lane group and byte, so the MFMA pairs equal k.
  const size_t f = (((size_t) (n / 16) * (K / 128) + k / 128) * 64 + ((k % 128) / 32) * 16 + n % 16) * 32 + k % 32;
  const size_t b = (((size_t) (k / 16) * (N / 128) + n / 128) * 64 + ((n % 128) / 32) * 16 + k % 16) * 32 + n % 32;
  Af[(size_t) el * per + f] =
      (uint8_t) (__builtin_amdgcn_cvt_pk_fp8_f32(w / swf[(size_t) el * N + n], 0.0f, 0, false) & 0xff);
  Ab[(size_t) el * per + b] =
      (uint8_t) (__builtin_amdgcn_cvt_pk_fp8_f32(w / swb[(size_t) el * K + k], 0.0f, 0, false) & 0xff);
}
}  // namespace

// every hidden width a multiple of 16 and every GEMM inner size a multiple of
// two fast-path steps (FP8: of one 128-wide step)
bool mlp_fast_ok(const MlpParams &mp, int mode)
{
  const int ks2 = mode == 0 ? 32 : (mode == 2 ? 128 : 64);
  for (int l = 0; l < mp.nl - 1; l++) {
    if (mp.sizes[l + 1] % 16 || mp.sizes[l] % 16) return false;
    if (mp.sizes[l] % ks2 || mp.sizes[l + 1] % ks2) return false;
  }
  return true;
}

hipError_t launch_quant_fp8(const float *W, const float *swf, const float *swb, uint8_t *Af,
                            uint8_t *Ab, int K, int N, int nelt, hipStream_t stream)
{
  size_t total = (size_t) K * N * nelt;
  if (total == 0) return hipSuccess;
  if (K % 32 || N % 32) return hipErrorInvalidValue;
  k_quant_fp8<<<(unsigned) ((total + 255) / 256), 256, 0, stream>>>(W, swf, swb, Af, Ab, K, N, nelt);
  return hipGetLastError();
}

// LDS layout for R rows: H_l and D_l per hidden layer (+ dE/dH for resnet).
size_t mlp_lds_bytes(const MlpParams &mp, int mode, int R, MlpParams *out)
{
  MlpParams m = mp;
  // FP32 activations in LDS for the FP32 and split-BF16 modes, BF16 for
  // the BF16 and FP8 modes
  const bool f32lds = mode == 0 || mode == 3;
  size_t esz = f32lds ? 4 : 2;
  size_t off = 0;
  int maxw = 0;
  bool anyres = false;
  for (int l = 0; l <= mp.nl - 1; l++) {
    int w = mp.sizes[l];
    m.ld[l] = f32lds ? w + 4 : w + 8;
  }
  for (int l = 0; l < mp.nl - 1; l++)
    if (mp.res[l]) anyres = true;
  // tanh without skips: D = 1 - h^2 is recomputed from H, so G_l overwrites H_l
  m.storeD = !(mp.act == 2 && !anyres);
  for (int l = 0; l < mp.nl - 1; l++) {
    int N = mp.sizes[l + 1];
    maxw = std::max(maxw, N);
    m.offH[l] = (int) off;
    off += (size_t) R * m.ld[l + 1];
    m.offD[l] = m.offH[l];
    if (m.storeD) {
      m.offD[l] = (int) off;
      off += (size_t) R * m.ld[l + 1];
    }
  }
  m.ldGH = f32lds ? maxw + 4 : maxw + 8;
  m.offGH = (int) off;
  if (anyres) off += (size_t) R * m.ldGH;
  size_t bytes = (off * esz + 15) & ~(size_t) 15;
  m.offRS = (int) bytes;
  m.offQ = 0;
  m.ldq = 0;
  if (mode == 2) {
    bytes += 4 * (size_t) R * sizeof(float);  // scales, inverse scales, two buffers of recorded maxima
    // quantized B rows: the widest GEMM inner size, rows 16-byte aligned
    int maxk = 0;
    for (int l = 0; l < mp.nl; l++) maxk = std::max(maxk, mp.sizes[l]);
    m.ldq = maxk + 16;
    m.offQ = (int) ((bytes + 15) & ~(size_t) 15);
    bytes = (size_t) m.offQ + (size_t) R * m.ldq;
  }
  if (out) *out = m;
  return bytes;
}

bool mlp_bf16_rows(const MlpParams &mp, int mode, int rtile)
{
  if (mode != 1 || !mlp_fast_ok(mp, 1) || mp.act != 2 || mlp_shape(mp) == 0) return false;
  if (rtile != 0 && rtile != 32) return false;
  return mlp_lds_bytes(mp, 1, 32, nullptr) <= 160 * 1024;  // the 32-row tile is the one taken
}

hipError_t launch_mlp(int mode, const RowLayout &rl, const MlpParams &mp, const float *G, double *E,
                      float *dEdG, hipStream_t s, int rtile, const BatchDims *dyn, int row_begin,
                      int row_end, bool g_bf16)
{
  if (g_bf16 && !mlp_bf16_rows(mp, mode, rtile)) return hipErrorInvalidValue;
  if (row_end < 0 || row_end > rl.nrows) row_end = rl.nrows;
  if (row_begin < 0 || row_begin > row_end) return hipErrorInvalidValue;
  if (row_end == row_begin) return hipSuccess;
  const size_t limit = 160 * 1024;
  // 32-row tiles by default: several workgroups per CU hide the LDS/L2
  // latency better than one 64-row tile, at twice the (L2-resident) weight
  // traffic.  FP8: 16-row tiles (its 32-byte K=128 operands need 152 VGPRs
  // at 32 rows, 3 waves per SIMD; 128 and 4 waves at 16 rows, 2.5% faster
  // on the 1M-atom C4 system).  rtile (16/32/64) overrides the first choice.
  int Rs[3] = {mode == 2 ? 16 : 32, 16, 16};
  if (rtile == 16 || rtile == 32 || rtile == 64) Rs[0] = rtile;
  for (int q = 0; q < 3; q++) {
    int R = Rs[q];
    MlpParams m;
    size_t lds = mlp_lds_bytes(mp, mode, R, &m);
    if (lds > limit) continue;
    if (mode == 0) {
      if (R == 64) return launch_r<0, 64>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end);
      if (R == 32) return launch_r<0, 32>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end);
      return launch_r<0, 16>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end);
    }
    if (mode == 3) {
      if (R == 64) return launch_r<3, 64>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end);
      if (R == 32) return launch_r<3, 32>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end);
      return launch_r<3, 16>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end);
    }
    if (mode == 2) {
      if (R == 64) return launch_r<2, 64>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end);
      if (R == 32) return launch_r<2, 32>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end);
      return launch_r<2, 16>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end);
    }
    if (R == 64) return launch_r<1, 64>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end, g_bf16);
    if (R == 32) return launch_r<1, 32>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end, g_bf16);
    return launch_r<1, 16>(rl, m, G, E, dEdG, lds, s, dyn, row_begin, row_end, g_bf16);
  }
  return hipErrorInvalidConfiguration;
}

}  // namespace tmd
