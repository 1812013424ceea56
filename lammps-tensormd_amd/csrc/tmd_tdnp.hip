// TDNP: the temperature-dependent networks that replace the per-element MLP
// of a tdnp = 1 model (TDNP::forward/backward, src/ML-TENSORMD/tdnp.cpp:150-218;
// set_element_tdnp_device and backward_tdnp_GPU, gpu/cuda/nnp.cu:294-316).
//
// Per atom row, one wavefront:
//   y = H(G), y[ydim - 1] = T (the electron temperature, eV)
//   u = U(y), s = S(y) (times T for a Sommerfeld model), a = u - T s
//   dA/dG = H^T-backward of (dU/dy - T dS/dy) (T^2 for Sommerfeld) with
//   the temperature column's gradient zeroed.
// E[row] = u is the atom's energy and dA/dG drives the forces, as in the
// reference.  Layer activations and their derivatives stay in LDS; weights
// are read from HBM/L2 (W for the forward pass, its transpose for the
// backward pass, so that the lanes of a wave read consecutive words).
// FP32 arithmetic with the scalar outputs summed in FP64.  TDNP models are
// not shipped with the reference; this kernel favours generality over the
// MFMA tiling of the main MLP.
#include "tmd_device.h"

#include <cmath>

namespace tmd {

namespace {

// activation value and derivative, as the FP32 MLP kernel evaluates them
// (nnp.cu:9-49, activation.hpp:158-261)
__device__ inline float tdnp_act(int act, float z, float &dz)
{
  if (act == 1) {  // softplus; linear above -(log(eps_f32) + 2)
    const float ez = __expf(z);
    dz = 1.0f / (1.0f + 1.0f / ez);
    return z < 13.942385f ? log1pf(ez) : z;
  }
  if (act == 2) {
    const float h = tanhf(z);
    dz = 1.0f - h * h;
    return h;
  }
  if (act == 3) {  // squareplus, rationalised for z < 0
    const float s = sqrtf(z * z + 4.0f);
    const float h = z >= 0.0f ? 0.5f * (z + s) : 2.0f / (s - z);
    dz = h / s;
    return h;
  }
  dz = z >= 0.0f ? 1.0f : 0.0f;
  return z >= 0.0f ? z : 0.0f;
}

__device__ inline double wave_sum(double v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// LDS layout of one row (floats): input, hidden activations h and their
// derivatives d of each network, y, and three gradient vectors.
struct TdnpLds {
  int x;                        // G row [gw]
  int h[3][kMaxLayers];         // hidden layer l output
  int d[3][kMaxLayers];         // its activation derivative
  int y;                        // H output plus the temperature [ydim]
  int ga, gb;                   // backward ping-pong [maxw]
  int gu;                       // dU/dy [ydim]
  int one;                      // the seed 1 of the U and S backward passes
  int total;
};

// Forward pass of network q for element e from LDS vector `in`.  Hidden
// outputs and derivatives go to their LDS slots; the last layer's outputs
// go to `out` (LDS) when it is given, else its single output is returned
// as an FP64 sum.
__device__ double net_forward(const TdnpNet &n, const TdnpLds &L, int q, int e, float *sm, const float *in,
                              float *out)
{
  const int lane = threadIdx.x & 63;
  const float *x = in;
  double res = 0.0;
  for (int l = 0; l < n.nl; l++) {
    const int K = n.sizes[l], N = n.sizes[l + 1];
    const float *W = n.W[l] + (size_t) e * K * N;
    const float *b = n.b[l] ? n.b[l] + (size_t) e * N : nullptr;
    if (l == n.nl - 1 && !out) {
      double s = 0.0;
      for (int i = lane; i < K; i += 64) s += (double) (x[i] * W[i]);
      res = wave_sum(s) + (b ? (double) b[0] : 0.0);
      break;
    }
    float *ho = l == n.nl - 1 ? out : sm + L.h[q][l];
    float *dz = sm + L.d[q][l];
    for (int j = lane; j < N; j += 64) {
      float z = b ? b[j] : 0.0f;
      for (int i = 0; i < K; i++) z = fmaf(x[i], W[(size_t) i * N + j], z);
      if (l == n.nl - 1) {
        ho[j] = z;  // linear output layer
      } else {
        float dv;
        float v = tdnp_act(n.act, z, dv);
        if (n.res[l]) v += x[j];
        ho[j] = v;
        dz[j] = dv;
      }
    }
    __syncthreads();
    x = ho;
  }
  return res;
}

// Backward pass of network q: gx = d(g_out . output)/d(input), g_out at the
// output layer (width sizes[nl]); ga and gb are scratch vectors.
__device__ void net_backward(const TdnpNet &n, const TdnpLds &L, int q, int e, float *sm, const float *g_out,
                             float *gx)
{
  const int lane = threadIdx.x & 63;
  const float *g = g_out;
  for (int l = n.nl - 1; l >= 0; l--) {
    const int K = n.sizes[l], N = n.sizes[l + 1];
    const float *Wt = n.Wt[l] + (size_t) e * K * N;
    const float *dz = l < n.nl - 1 ? sm + L.d[q][l] : nullptr;
    float *gin = l == 0 ? gx : sm + ((n.nl - 1 - l) % 2 ? L.gb : L.ga);
    for (int i = lane; i < K; i += 64) {
      float s = 0.0f;
      for (int j = 0; j < N; j++) s = fmaf(Wt[(size_t) j * K + i], dz ? g[j] * dz[j] : g[j], s);
      if (l < n.nl - 1 && n.res[l]) s += g[i];
      gin[i] = s;
    }
    __syncthreads();
    g = gin;
  }
}

__global__ __launch_bounds__(64) void k_tdnp(RowLayout rl, TdnpParams tp, TdnpLds L, const int *row2atom,
                                             const float *G, double *E, double *sa, float *dEdG,
                                             const BatchDims *dyn)
{
  extern __shared__ float sm[];
  const RowLayout &lay = dyn ? dyn->rl : rl;
  const int row = blockIdx.x;
  if (row >= rl.nrows || row >= lay.nrows) return;
  const int lane = threadIdx.x & 63;
  const int gw = tp.net[0].sizes[0];
  float *grow = dEdG + (size_t) row * gw;
  const int atom = row2atom[row];
  if (atom < 0) {  // padding row
    for (int i = lane; i < gw; i += 64) grow[i] = 0.0f;
    if (lane == 0) {
      E[row] = 0.0;
      sa[2 * (size_t) row] = 0.0;
      sa[2 * (size_t) row + 1] = 0.0;
    }
    return;
  }
  int e = 0;
  while (e + 1 < kMaxElts && row >= lay.row_off[e + 1]) e++;
  const double T = *tp.etemp;
  const int yd = tp.ydim;
  float *x = sm + L.x;
  for (int i = lane; i < gw; i += 64) x[i] = G[(size_t) row * gw + i];
  __syncthreads();
  float *y = sm + L.y;
  net_forward(tp.net[0], L, 0, e, sm, x, y);
  if (lane == 0) y[yd - 1] = (float) T;
  __syncthreads();
  const double u = net_forward(tp.net[1], L, 1, e, sm, y, nullptr);
  double s = net_forward(tp.net[2], L, 2, e, sm, y, nullptr);
  if (tp.sommerfeld) s *= T;
  // dU/dy into gu, dS/dy into gs -> gy = dU/dy - T' dS/dy in gu
  float *gu = sm + L.gu;
  float *gs = sm + L.x;  // the G row is no longer needed
  if (lane == 0) sm[L.one] = 1.0f;
  __syncthreads();
  net_backward(tp.net[1], L, 1, e, sm, sm + L.one, gu);
  net_backward(tp.net[2], L, 2, e, sm, sm + L.one, gs);
  const float sc = (float) (tp.sommerfeld ? T * T : T);
  for (int j = lane; j < yd; j += 64) gu[j] = j == yd - 1 ? 0.0f : fmaf(-sc, gs[j], gu[j]);
  __syncthreads();
  net_backward(tp.net[0], L, 0, e, sm, gu, gs);
  for (int i = lane; i < gw; i += 64) grow[i] = gs[i];
  if (lane == 0) {
    E[row] = u;
    sa[2 * (size_t) row] = s;
    sa[2 * (size_t) row + 1] = u - T * s;
  }
}

}  // namespace

hipError_t launch_tdnp(const RowLayout &rl, const TdnpParams &tp, const int *row2atom, const float *G,
                       double *E, double *sa, float *dEdG, hipStream_t stream, const BatchDims *dyn)
{
  if (rl.nrows <= 0) return hipSuccess;
  TdnpLds L{};
  int o = 0;
  auto take = [&](int n) {
    const int r = o;
    o += (n + 3) & ~3;
    return r;
  };
  int maxw = 1;
  for (int q = 0; q < 3; q++)
    for (int l = 0; l <= tp.net[q].nl; l++) maxw = maxw > tp.net[q].sizes[l] ? maxw : tp.net[q].sizes[l];
  L.x = take(maxw);  // the G row, later dS/dy and dA/dG
  for (int q = 0; q < 3; q++)
    for (int l = 0; l + 1 < tp.net[q].nl; l++) {
      L.h[q][l] = take(tp.net[q].sizes[l + 1]);
      L.d[q][l] = take(tp.net[q].sizes[l + 1]);
    }
  L.y = take(tp.ydim);
  L.ga = take(maxw);
  L.gb = take(maxw);
  L.gu = take(maxw);
  L.one = take(1);
  L.total = o;
  const size_t lds = (size_t) o * 4;
  if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
  if (hipError_t e = hipFuncSetAttribute((const void *) k_tdnp, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int) lds);
      e != hipSuccess)
    return e;
  k_tdnp<<<rl.nrows, 64, lds, stream>>>(rl, tp, L, row2atom, G, E, sa, dEdG, dyn);
  return hipGetLastError();
}

}  // namespace tmd
