// The radial interpolation table built on the device at model load.
//
// Replaces forward_fnn_GPU (src/ML-TENSORMD/gpu/cuda/nnp.cu:59-116: the
// radial network over the grid, one GEMM per layer) and the host steps of
// FeaturePotential::setup_ration (fnn.cpp:77-105: cutoff, Kerley set-up of
// BatchRation1d, ration.hpp:166-225).  Two kernels:
//   k_table_values   one workgroup per grid point r_i: the radial network
//                    (or the analytic functions) times the cutoff -> y
//   k_table_kerley   one thread per (point, filter): slope s and the
//                    Kerley coefficients c1, c2 from the neighboring y
// Every product and sum is a separate FP64 operation in the host build's
// order (no contraction), so the two builds differ only where the device and
// host elementary functions (exp, log1p, tanh, cos, pow) round differently.
#include <algorithm>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "tmd_model.h"

namespace tmd {

namespace {

constexpr int kTabLayers = 16;
constexpr int kTabWidth = 1024;

struct TableArgs {
  int n, k;
  double xdx, rc;
  int fctype;
  int use_fnn;
  // radial network
  int nl, act, resnet, maxw;
  int sizes[kTabLayers + 1];
  const double *W[kTabLayers];
  const double *B[kTabLayers];  // nullptr: no bias
  double thr;                   // softplus threshold
  // analytic functions: [parameter][filter]
  int radial_fn;
  const double *p0, *p1, *p2;
};

__device__ double act_value(int act, double z, double thr)
{
#pragma clang fp contract(off)
  switch (act) {
    case ACT_SOFTPLUS: return (z < -thr) ? log1p(exp(z)) : z;
    case ACT_SQUAREPLUS: return 0.5 * (z + sqrt(z * z + 4.0));
    case ACT_TANH: return tanh(z);
    default: return z >= 0 ? z : 0.0;
  }
}

__device__ double analytic(const TableArgs &a, double r, int k)
{
#pragma clang fp contract(off)
  switch (a.radial_fn) {
    case RADIAL_PEXP: return exp(-pow(r / a.p0[k], a.p1[k]));
    case RADIAL_MORSE: {
      const double g = a.p1[k] * (r - a.p2[k]);
      return a.p0[k] * (exp(-2.0 * g) - 2.0 * exp(-g));
    }
    case RADIAL_DENSITY: return a.p0[k] * exp(-a.p1[k] * ((r - a.p2[k]) / a.p2[k]));
    default: {
      const double d = r - a.p1[k];
      return exp(-a.p0[k] * d * d / (a.rc * a.rc));
    }
  }
}

__device__ double cutoff(const TableArgs &a, double r)
{
#pragma clang fp contract(off)
  if (a.fctype == 0) {
    const double Z = fmin(r / a.rc, 1.0) * M_PI;
    return 0.5 + 0.5 * cos(Z);
  }
  const double gamma = 5.0;
  const double Z = fmin(r / a.rc, M_PI);
  const double Z4 = Z * Z * Z * Z, Z5 = Z4 * Z;
  return 1.0 + gamma * Z5 * Z - (gamma + 1.0) * Z5;
}

// y(r_i) of every filter -> tab[i][j][0]; LDS holds two activation rows
__global__ __launch_bounds__(64) void k_table_values(TableArgs a, const double *xs, double *tab)
{
#pragma clang fp contract(off)
  extern __shared__ double sh[];
  const int i = blockIdx.x;
  const double r = xs[i];
  double *h = sh, *hn = sh + a.maxw;
  if (a.use_fnn) {
    if (threadIdx.x == 0) h[0] = r;
    __syncthreads();
    for (int l = 0; l < a.nl; l++) {
      const int K = a.sizes[l], N = a.sizes[l + 1];
      const bool last = l == a.nl - 1;
      const bool res = a.resnet && l > 0 && K == N;
      const double *W = a.W[l], *B = a.B[l];
      for (int j = threadIdx.x; j < N; j += 64) {
        double z = 0.0;
        for (int q = 0; q < K; q++) z += h[q] * W[(size_t) q * N + j];
        if (B) z += B[j];
        hn[j] = last ? z : act_value(a.act, z, a.thr) + (res ? h[j] : 0.0);
      }
      __syncthreads();
      double *t = h;
      h = hn;
      hn = t;
    }
  } else {
    for (int j = threadIdx.x; j < a.k; j += 64) h[j] = analytic(a, r, j);
    __syncthreads();
  }
  const double fc = cutoff(a, r);
  for (int j = threadIdx.x; j < a.k; j += 64) tab[((size_t) i * a.k + j) * 4] = h[j] * fc;
}

// slope s_i = (y_{i+1} - y_i) / dx and the Kerley coefficients of cell i
// (the last grid point keeps zeros), as the host build orders them
__global__ void k_table_kerley(TableArgs a, double *tab)
{
#pragma clang fp contract(off)
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t) a.n * a.k) return;
  const int i = (int) (t / a.k), j = (int) (t % a.k), n = a.n;
  const double xdx = a.xdx;
  auto y = [&](int q) { return tab[((size_t) q * a.k + j) * 4]; };
  auto slope = [&](int q) { return q < n - 1 ? (y(q + 1) - y(q)) * xdx : 0.0; };
  double *o = tab + t * 4;
  if (i == n - 1) {
    o[1] = o[2] = o[3] = 0.0;
    return;
  }
  const double s = slope(i);
  double c1, c2;
  if (i == 0) {
    const double sp = slope(i + 1);
    c2 = (s * (s - (sp - s) / 2) <= 0) ? s * xdx : (sp - s) * xdx / 2;
    c1 = 0;
  } else if (i == n - 2) {
    c1 = (s - slope(i - 1)) * xdx / 2;
    c2 = 0;
  } else {
    const double sm = slope(i - 1), sp = slope(i + 1);
    c1 = (i == 1 && sm * (sm - (s - sm) / 2) <= 0) ? (s - 2 * sm) * xdx : (s - sm) * xdx / 2;
    c2 = (sp - s) * xdx / 2;
  }
  o[1] = s;
  o[2] = c1;
  o[3] = c2;
}

// device buffers of one build, freed on every return path
struct Scratch {
  std::vector<void *> p;
  ~Scratch()
  {
    for (void *q : p) hipFree(q);
  }
  double *up(const std::vector<double> &v, hipError_t &e)
  {
    void *q = nullptr;
    if (e != hipSuccess) return nullptr;
    if ((e = hipMalloc(&q, sizeof(double) * std::max<size_t>(v.size(), 1))) != hipSuccess) return nullptr;
    p.push_back(q);
    if (!v.empty()) e = hipMemcpy(q, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice);
    return (double *) q;
  }
};

}  // namespace

bool build_table_device(Model &m, double delta, std::string &err)
{
  std::vector<double> xs;
  table_grid(m, delta, xs);
  TableArgs a{};
  a.n = m.tab_n;
  a.k = m.k;
  a.xdx = m.xdx;
  a.rc = m.rc;
  a.fctype = m.fctype;
  a.use_fnn = m.use_fnn ? 1 : 0;
  a.thr = softplus_threshold();
  a.radial_fn = m.radial_fn;
  Scratch sc;
  hipError_t e = hipSuccess;
  a.maxw = m.k;
  if (m.use_fnn) {
    const DenseNet &net = m.fnn;
    if (net.nlayers < 1 || net.nlayers > kTabLayers) {
      err = "device table build: the radial network has more than 16 layers";
      return false;
    }
    a.nl = net.nlayers;
    a.act = net.act;
    a.resnet = net.resnet ? 1 : 0;
    for (int l = 0; l <= net.nlayers; l++) {
      a.sizes[l] = net.sizes[l];
      a.maxw = std::max(a.maxw, net.sizes[l]);
    }
    for (int l = 0; l < net.nlayers; l++) {
      a.W[l] = sc.up(net.W[l], e);
      a.B[l] = net.B[l].empty() ? nullptr : sc.up(net.B[l], e);
    }
  } else {
    const size_t need = (m.radial_fn == RADIAL_PEXP || m.radial_fn == RADIAL_SF) ? 2 : 3;
    if (m.rparams.size() < need) {
      err = "device table build: missing analytic parameters";
      return false;
    }
    a.p0 = sc.up(m.rparams[0], e);
    a.p1 = sc.up(m.rparams[1], e);
    a.p2 = m.rparams.size() > 2 ? sc.up(m.rparams[2], e) : nullptr;
  }
  if (a.maxw > kTabWidth) {
    err = "device table build: a radial layer is wider than 1024";
    return false;
  }
  double *dx = sc.up(xs, e);
  double *tab = nullptr;
  if (e == hipSuccess && (e = hipMalloc((void **) &tab, sizeof(double) * (size_t) a.n * a.k * 4)) == hipSuccess)
    sc.p.push_back(tab);
  if (e != hipSuccess) {
    err = std::string("device table build: ") + hipGetErrorString(e);
    return false;
  }
  k_table_values<<<a.n, 64, 2 * sizeof(double) * a.maxw, 0>>>(a, dx, tab);
  if ((e = hipGetLastError()) == hipSuccess) {
    const size_t cells = (size_t) a.n * a.k;
    k_table_kerley<<<(unsigned) ((cells + 255) / 256), 256, 0, 0>>>(a, tab);
    e = hipGetLastError();
  }
  m.tab.assign((size_t) a.n * a.k * 4, 0.0);
  if (e == hipSuccess) e = hipMemcpy(m.tab.data(), tab, sizeof(double) * m.tab.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    err = std::string("device table build: ") + hipGetErrorString(e);
    return false;
  }
  return true;
}

}  // namespace tmd
