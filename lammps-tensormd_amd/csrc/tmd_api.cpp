// C ABI implementation: contexts (model + device parameters), the
// Pair::compute drop-in over LAMMPS neighbor lists, and device-resident
// periodic systems with velocity-Verlet.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tensormd_amd.h"
#include "tmd_device.h"
#include "tmd_internal.h"
#include "tmd_model.h"

using namespace tmd;

namespace {
thread_local std::string g_thread_err;
// set on library-owned threads (the serving worker): their errors stay in
// the thread's own message and never touch a context's message, which the
// caller's thread reads and writes without a lock
thread_local bool g_library_thread = false;
}

namespace tmd {

void set_err(tmd_ctx *ctx, const std::string &m)
{
  g_thread_err = m;
  if (ctx && !g_library_thread) ctx->err = m;
}

void mark_library_thread() { g_library_thread = true; }

const std::string &thread_error() { return g_thread_err; }

int hip_fail(tmd_ctx *ctx, hipError_t e, const char *what)
{
  set_err(ctx, std::string(what) + ": " + hipGetErrorString(e));
  return TMD_ERR_DEVICE;
}

DeviceBuffer::~DeviceBuffer() { release(); }

void DeviceBuffer::release()
{
  if (p && owned) hipFree(p);
  p = nullptr;
  bytes = 0;
  owned = true;
}

hipError_t DeviceBuffer::ensure(size_t nbytes)
{
  if (nbytes <= bytes && p) return hipSuccess;
  if (!owned) return hipErrorOutOfMemory;  // a fixed window never grows
  release();
  size_t want = std::max<size_t>(nbytes, 256);
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    return e;
  }
  bytes = want;
  return hipSuccess;
}

PinnedBuffer::~PinnedBuffer() { release(); }

void PinnedBuffer::release()
{
  if (p) hipHostFree(p);
  p = nullptr;
  bytes = 0;
}

hipError_t PinnedBuffer::ensure(size_t nbytes)
{
  if (nbytes <= bytes && p) return hipSuccess;
  release();
  size_t want = std::max<size_t>(nbytes + nbytes / 4, 4096);
  hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    p = nullptr;
    return e;
  }
  bytes = want;
  return hipSuccess;
}

static uint16_t bf16_bits(float f)
{
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t) (u >> 16);
}

static float bf16_float(uint16_t h)
{
  const uint32_t u = (uint32_t) h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// TDNP networks H, U, S: per layer, all elements packed, W [K][N], its
// transpose and the biases in FP32; the electron temperature as a device
// scalar.
int upload_tdnp(tmd_ctx *ctx)
{
  const Model &m = ctx->model;
  const std::vector<DenseNet> *nets[3] = {&m.tdH, &m.tdU, &m.tdS};
  size_t total = 0;
  auto al = [](size_t x) { return (x + 255) & ~(size_t) 255; };
  size_t off[3][kMaxLayers][3];
  for (int q = 0; q < 3; q++) {
    const DenseNet &n0 = (*nets[q])[0];
    if (n0.nlayers > kMaxLayers) return (set_err(ctx, "TDNP network with more than 8 layers"), TMD_ERR_MODEL);
    for (int l = 0; l < n0.nlayers; l++) {
      const size_t per = (size_t) n0.sizes[l] * n0.sizes[l + 1];
      off[q][l][0] = total;
      total = al(total + per * m.nelt * 4);
      off[q][l][1] = total;
      total = al(total + per * m.nelt * 4);
      off[q][l][2] = total;
      total = al(total + (size_t) n0.sizes[l + 1] * m.nelt * 4);
    }
  }
  hipError_t e = ctx->tdnp_params.ensure(total);
  if (e != hipSuccess) return hip_fail(ctx, e, "TDNP parameter alloc");
  std::vector<unsigned char> host(total, 0);
  unsigned char *base = (unsigned char *) ctx->tdnp_params.p;
  TdnpParams &tp = ctx->tdnp;
  std::memset(&tp, 0, sizeof(tp));
  for (int q = 0; q < 3; q++) {
    const DenseNet &n0 = (*nets[q])[0];
    TdnpNet &tn = tp.net[q];
    tn.nl = n0.nlayers;
    tn.act = n0.act;
    for (int l = 0; l <= n0.nlayers; l++) tn.sizes[l] = n0.sizes[l];
    for (int l = 0; l < n0.nlayers; l++) {
      const int K = n0.sizes[l], N = n0.sizes[l + 1];
      const size_t per = (size_t) K * N;
      float *W = (float *) &host[off[q][l][0]];
      float *Wt = (float *) &host[off[q][l][1]];
      float *b = (float *) &host[off[q][l][2]];
      bool hasb = false;
      for (int el = 0; el < m.nelt; el++) {
        const DenseNet &net = (*nets[q])[el];
        for (int i = 0; i < K; i++)
          for (int j = 0; j < N; j++) {
            const float w = (float) net.W[l][(size_t) i * N + j];
            W[el * per + (size_t) i * N + j] = w;
            Wt[el * per + (size_t) j * K + i] = w;
          }
        if (!net.B[l].empty()) {
          hasb = true;
          for (int j = 0; j < N; j++) b[(size_t) el * N + j] = (float) net.B[l][j];
        }
      }
      tn.W[l] = (const float *) (base + off[q][l][0]);
      tn.Wt[l] = (const float *) (base + off[q][l][1]);
      tn.b[l] = hasb ? (const float *) (base + off[q][l][2]) : nullptr;
      tn.res[l] = n0.res_layer(l) ? 1 : 0;
    }
  }
  tp.ydim = m.ydim;
  tp.sommerfeld = m.sommerfeld ? 1 : 0;
  if ((e = hipMemcpy(ctx->tdnp_params.p, host.data(), total, hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(ctx, e, "TDNP parameter upload");
  if ((e = ctx->etemp_dev.ensure(sizeof(double))) != hipSuccess) return hip_fail(ctx, e, "etemp alloc");
  if ((e = hipMemcpy(ctx->etemp_dev.p, &ctx->etemp, sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(ctx, e, "etemp upload");
  tp.etemp = (const double *) ctx->etemp_dev.p;
  return TMD_OK;
}

// Upload table and MLP parameters.
int upload_model(tmd_ctx *ctx)
{
  const Model &m = ctx->model;
  ModelGeo &g = ctx->geo;
  std::memset(&g, 0, sizeof(g));
  g.nelt = m.nelt;
  g.k = m.k;
  g.tab_n = m.tab_n;
  for (int a = 0; a < m.nelt; a++) {
    g.slots[a] = 0;
    for (int b = 0; b < m.nelt; b++) g.slots[a] |= (uint32_t) m.eltij[a][b] << (4 * b);
  }
  // the kernels hold T as the fixed multiplicity table (plus the T-symmetric
  // extras of max_moment 3); check that the model's T is exactly that
  g.tsym = m.T_sym ? 1 : 0;
  g.mm = m.max_moment;
  if (m.max_moment < 3 || m.max_moment > 5 || m.nm != m.max_moment + 1)
    return (set_err(ctx, "max_moment 3, 4 or 5 expected after padding"), TMD_ERR_MODEL);
  if (m.max_moment == 3) {
    if (m.nd != kND) return (set_err(ctx, "unexpected monomial count"), TMD_ERR_MODEL);
    for (int mm = 0; mm < kNM; mm++)
      for (int d = 0; d < kND; d++) {
        double t = kDeg[d] == mm ? (double) kMult[d] : 0.0;
        if (g.tsym && mm == 2 && d == 0) t = -1.0 / 3.0;
        if (g.tsym && mm == 3 && d >= 1 && d <= 3) t = -3.0 / 5.0;
        if (m.T[mm * m.nd + d] != t) return (set_err(ctx, "unexpected descriptor T table"), TMD_ERR_MODEL);
      }
  } else {
    // max_moment 4, 5: T[deg d][d] = deg! / (a! b! c!), no T-symmetric extras
    // (tensormd.cpp:217,229 apply them only up to max_moment 3)
    int d = 0;
    auto fact = [](int n) {
      double v = 1;
      for (int q = 2; q <= n; q++) v *= q;
      return v;
    };
    for (int n = 0; n <= m.max_moment; n++)
      for (int a = n; a >= 0; a--)
        for (int b = n - a; b >= 0; b--, d++)
          for (int mm = 0; mm < m.nm; mm++) {
            const double t = mm == n ? fact(n) / (fact(a) * fact(b) * fact(n - a - b)) : 0.0;
            if (d >= m.nd || m.T[mm * m.nd + d] != t)
              return (set_err(ctx, "unexpected descriptor T table"), TMD_ERR_MODEL);
          }
  }
  g.rc2 = m.rc * m.rc;
  g.dx = m.dx;
  g.xdx = m.xdx;
  g.dxf = (float) m.dx;
  // table as float4 [n][k]
  std::vector<float> tabf(m.tab.size());
  for (size_t i = 0; i < tabf.size(); i++) tabf[i] = (float) m.tab[i];
  hipError_t e = ctx->tab.ensure(tabf.size() * 4);
  if (e != hipSuccess) return hip_fail(ctx, e, "table alloc");
  e = hipMemcpy(ctx->tab.p, tabf.data(), tabf.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(ctx, e, "table upload");

  if (m.tdnp) return upload_tdnp(ctx);
  // MLP: per layer, all elements packed
  const DenseNet &n0 = m.mlp[0];
  const int L = n0.nlayers;
  if (L > kMaxLayers) return (set_err(ctx, "more than 8 MLP layers"), TMD_ERR_MODEL);
  size_t total = 0;
  struct Off {
    size_t W, Wt, Wb, Wtb, b, af, ab, swf, swb;
  };
  const bool fp8 = ctx->opt.mlp_precision == TMD_MLP_FP8;
  const int mode = ctx->opt.mlp_precision;
  // the fast kernel's A operands, in fragment order, for this mode only
  bool fast;
  {
    MlpParams t;
    std::memset(&t, 0, sizeof(t));
    t.nl = L;
    for (int l = 0; l <= L; l++) t.sizes[l] = n0.sizes[l];
    fast = mlp_fast_ok(t, mode);
  }
  // bytes per weight: FP32, BF16, FP8, or a BF16 hi + lo pair
  const size_t aesz = mode == TMD_MLP_FP32 ? 4 : (mode == TMD_MLP_BF16 ? 2 : (mode == TMD_MLP_FP8 ? 1 : 4));
  std::vector<Off> offs(L);
  auto al = [](size_t x) { return (x + 255) & ~(size_t) 255; };
  for (int l = 0; l < L; l++) {
    size_t per = (size_t) n0.sizes[l] * n0.sizes[l + 1];
    offs[l].W = total;
    total = al(total + per * m.nelt * 4);
    offs[l].Wt = total;
    total = al(total + per * m.nelt * 4);
    offs[l].Wb = total;
    total = al(total + per * m.nelt * 2);
    offs[l].Wtb = total;
    total = al(total + per * m.nelt * 2);
    offs[l].b = total;
    total = al(total + (size_t) n0.sizes[l + 1] * m.nelt * 4);
    if (fast && l < L - 1) {
      offs[l].af = total;
      total = al(total + per * m.nelt * aesz);
      offs[l].ab = total;
      total = al(total + per * m.nelt * aesz);
    }
    if (fp8) {  // per-output-channel scales of the e4m3 operands
      offs[l].swf = total;
      total = al(total + (size_t) n0.sizes[l + 1] * m.nelt * 4);
      offs[l].swb = total;
      total = al(total + (size_t) n0.sizes[l] * m.nelt * 4);
    }
  }
  e = ctx->params.ensure(total);
  if (e != hipSuccess) return hip_fail(ctx, e, "parameter alloc");
  std::vector<unsigned char> host(total, 0);
  MlpParams &mp = ctx->mlp;
  std::memset(&mp, 0, sizeof(mp));
  mp.nl = L;
  for (int l = 0; l <= L; l++) mp.sizes[l] = n0.sizes[l];
  mp.act = n0.act;
  unsigned char *base = (unsigned char *) ctx->params.p;
  for (int l = 0; l < L; l++) {
    int K = n0.sizes[l], N = n0.sizes[l + 1];
    size_t per = (size_t) K * N;
    float *W = (float *) &host[offs[l].W];
    float *Wt = (float *) &host[offs[l].Wt];
    uint16_t *Wb = (uint16_t *) &host[offs[l].Wb];
    uint16_t *Wtb = (uint16_t *) &host[offs[l].Wtb];
    float *b = (float *) &host[offs[l].b];
    bool hasb = false;
    for (int el = 0; el < m.nelt; el++) {
      const DenseNet &net = m.mlp[el];
      for (int i = 0; i < K; i++)
        for (int j = 0; j < N; j++) {
          float w = (float) net.W[l][(size_t) i * N + j];
          W[el * per + (size_t) i * N + j] = w;
          Wt[el * per + (size_t) j * K + i] = w;
          Wb[el * per + (size_t) i * N + j] = bf16_bits(w);
          Wtb[el * per + (size_t) j * K + i] = bf16_bits(w);
        }
      if (!net.B[l].empty()) {
        hasb = true;
        for (int j = 0; j < N; j++) b[el * N + j] = (float) net.B[l][j];
      }
      if (fast && l < L - 1 && !fp8) {
        // fragment order: [tile t][step s][lane][v]; lane = (group g) * 16 +
        // (row in tile); forward outputs n = 16 t + row, inner k = s KS + g V
        // + v (KS, V = 16, 4 for FP32, whose four MFMAs per step take the
        // values v; 32, 8 for BF16); backward swaps K and N.  Split BF16:
        // the lane's 8 high parts, then its 8 low parts.
        const bool f32 = mode == TMD_MLP_FP32;
        const bool split = mode == TMD_MLP_BF16X3;
        const int KS = f32 ? 16 : 32, V = f32 ? 4 : 8;
        const double *Wd = net.W[l].data();
        auto frag = [&](unsigned char *dst, int out, int inner, bool fwd) {
          const int nst = inner / KS;
          for (int t = 0; t < out / 16; t++)
            for (int st = 0; st < nst; st++)
              for (int lane = 0; lane < 64; lane++)
                for (int v = 0; v < V; v++) {
                  const int o = 16 * t + lane % 16, in = st * KS + (lane / 16) * V + v;
                  const double w = fwd ? Wd[(size_t) in * N + o] : Wd[(size_t) o * N + in];
                  const size_t q = (((size_t) t * nst + st) * 64 + lane) * V + v;
                  if (f32) {
                    reinterpret_cast<float *>(dst)[q] = (float) w;
                  } else if (split) {
                    const float wf = (float) w;
                    const uint16_t hi = bf16_bits(wf);
                    const float hf = bf16_float(hi);
                    const size_t q2 = (((size_t) t * nst + st) * 64 + lane) * 2 * V + v;
                    reinterpret_cast<uint16_t *>(dst)[q2] = hi;
                    reinterpret_cast<uint16_t *>(dst)[q2 + V] = bf16_bits(wf - hf);
                  } else {
                    reinterpret_cast<uint16_t *>(dst)[q] = bf16_bits((float) w);
                  }
                }
        };
        frag(&host[offs[l].af] + el * per * aesz, N, K, true);
        frag(&host[offs[l].ab] + el * per * aesz, K, N, false);
      }
      if (fp8) {
        // scale = max |w| / 440 over the output channel of each GEMM (forward:
        // column n of W; backward: row k), so that w / scale fits e4m3
        float *swf = (float *) &host[offs[l].swf] + (size_t) el * N;
        float *swb = (float *) &host[offs[l].swb] + (size_t) el * K;
        for (int j = 0; j < N; j++) swf[j] = 0.0f;
        for (int i = 0; i < K; i++) {
          float mx = 0.0f;
          for (int j = 0; j < N; j++) {
            float a = std::fabs((float) net.W[l][(size_t) i * N + j]);
            mx = std::max(mx, a);
            swf[j] = std::max(swf[j], a);
          }
          swb[i] = mx > 0.0f ? mx / 440.0f : 1.0f;
        }
        for (int j = 0; j < N; j++) swf[j] = swf[j] > 0.0f ? swf[j] / 440.0f : 1.0f;
      }
    }
    mp.Wf32[l] = (const float *) (base + offs[l].W);
    mp.Wt32[l] = (const float *) (base + offs[l].Wt);
    mp.Wb16[l] = (const uint16_t *) (base + offs[l].Wb);
    mp.Wtb[l] = (const uint16_t *) (base + offs[l].Wtb);
    mp.b[l] = hasb ? (const float *) (base + offs[l].b) : nullptr;
    if (fast && l < L - 1) {
      mp.Af[l] = base + offs[l].af;
      mp.Ab[l] = base + offs[l].ab;
    }
    if (fp8) {
      mp.swf[l] = (const float *) (base + offs[l].swf);
      mp.swb[l] = (const float *) (base + offs[l].swb);
    }
    mp.wstride[l] = (long) per;
    mp.bstride[l] = N;
    mp.res[l] = n0.res_layer(l) ? 1 : 0;
  }
  mp.wout = mp.Wf32[L - 1];
  mp.bout = mp.b[L - 1];
  mp.b[L - 1] = nullptr;
  // the fused kernel keeps one row tile's activations in LDS
  if (mlp_lds_bytes(mp, mode, 16, nullptr) > 160 * 1024)
    return (set_err(ctx, "MLP too wide for the fused kernel's LDS budget"), TMD_ERR_MODEL);
  if (fp8 && !mlp_fast_ok(mp, mode))
    return (set_err(ctx, "the FP8 MLP needs hidden and input widths that are multiples of 64"),
            TMD_ERR_MODEL);
  if (mode == TMD_MLP_BF16X3 && !mlp_fast_ok(mp, mode))
    return (set_err(ctx, "the split-BF16 MLP needs hidden and input widths that are multiples of 64"),
            TMD_ERR_MODEL);
  e = hipMemcpy(ctx->params.p, host.data(), total, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(ctx, e, "parameter upload");
  if (fp8) {
    for (int l = 0; l < L - 1; l++) {
      e = launch_quant_fp8(mp.Wf32[l], mp.swf[l], mp.swb[l], (uint8_t *) mp.Af[l], (uint8_t *) mp.Ab[l],
                           n0.sizes[l], n0.sizes[l + 1], m.nelt, ctx->stream);
      if (e != hipSuccess) return hip_fail(ctx, e, "FP8 weight quantization");
    }
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "FP8 weights");
  }
  return TMD_OK;
}

// Allocate the MLP/descriptor workspace for `nrows` rows.
int ensure_rows(tmd_ctx *ctx, Workspace &w, int nrows)
{
  const Model &m = ctx->model;
  const int gw = m.gwidth();
  hipError_t e;
  if ((e = w.G.ensure((size_t) nrows * gw * 4)) != hipSuccess) return hip_fail(ctx, e, "G alloc");
  if ((e = w.dEdG.ensure((size_t) nrows * gw * 4)) != hipSuccess) return hip_fail(ctx, e, "dEdG alloc");
  if ((e = w.E.ensure((size_t) nrows * 8)) != hipSuccess) return hip_fail(ctx, e, "E alloc");
  if ((e = w.row2atom.ensure((size_t) nrows * 4)) != hipSuccess) return hip_fail(ctx, e, "row alloc");
  if ((e = w.red.ensure(4096 * 8)) != hipSuccess) return hip_fail(ctx, e, "red alloc");
  if (m.tdnp && (e = w.sa.ensure((size_t) nrows * 16)) != hipSuccess) return hip_fail(ctx, e, "TDNP alloc");
  return TMD_OK;
}

int ensure_kept(tmd_ctx *ctx, Workspace &w, int nrows, int cap)
{
  hipError_t e = w.kl.ensure((size_t) std::max(nrows, 1) * kept_cache_words(cap) * 4);
  return e == hipSuccess ? TMD_OK : hip_fail(ctx, e, "kept-list cache alloc");
}

int ensure_atoms(tmd_ctx *ctx, Workspace &w, int natoms)
{
  hipError_t e = w.ew.ensure((size_t) std::max(natoms, 1) * 7 * sizeof(double));
  return e == hipSuccess ? TMD_OK : hip_fail(ctx, e, "eatom/watom alloc");
}

// descriptor -> MLP fwd/bwd -> force kernel -> FP64 reductions
int run_chain(tmd_ctx *ctx, Workspace &w, const Chain &c, hipStream_t s, PhaseEvents *pe)
{
  const int mode = ctx->opt.mlp_precision;  // 0 fp32, 1 bf16, 2 fp8
  hipError_t e;
  const int *row2atom = (const int *) w.row2atom.p;
  double *eatom = (double *) w.ew.p;
  double *watom = eatom + c.natoms;
  // the kept-list cache is sized before any capture; never allocate here
  if (w.kl.bytes < (size_t) w.rl.nrows * kept_cache_words(c.cap) * 4)
    return (set_err(ctx, "internal: kept-list cache not allocated for this list"), TMD_ERR_STATE);
  uint32_t *kl = (uint32_t *) w.kl.p;
  if (pe && (e = hipEventRecord(pe->ev[0], s)) != hipSuccess) return hip_fail(ctx, e, "timing event");
  Geom geo = c.geo;
  geo.g_bf16 = c.g_bf16 && !ctx->model.tdnp && ctx->geo.mm <= 3 && mlp_bf16_rows(ctx->mlp, mode, c.rtile);
  w.g_bf16 = geo.g_bf16;
  e = launch_descriptor(ctx->geo, geo, c.pos, c.nbr, c.nn, c.cap, row2atom, w.rl.nrows,
                        (const float4 *) ctx->tab.p, (float *) w.G.p, kl, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "descriptor launch");
  if (pe && (e = hipEventRecord(pe->ev[1], s)) != hipSuccess) return hip_fail(ctx, e, "timing event");
  if (ctx->model.tdnp) {
    if (w.sa.bytes < (size_t) w.rl.nrows * 16)
      return (set_err(ctx, "internal: TDNP buffer not allocated for these rows"), TMD_ERR_STATE);
    e = launch_tdnp(w.rl, ctx->tdnp, row2atom, (const float *) w.G.p, (double *) w.E.p, (double *) w.sa.p,
                    (float *) w.dEdG.p, s, c.dyn);
    if (e != hipSuccess) return hip_fail(ctx, e, "TDNP launch");
  } else {
    e = launch_mlp(mode, w.rl, ctx->mlp, (const float *) w.G.p, (double *) w.E.p, (float *) w.dEdG.p, s,
                   c.rtile, c.dyn, 0, -1, geo.g_bf16);
    if (e != hipSuccess) return hip_fail(ctx, e, "mlp launch");
  }
  if (pe && (e = hipEventRecord(pe->ev[2], s)) != hipSuccess) return hip_fail(ctx, e, "timing event");
  if (c.dyn)
    e = launch_zero_atoms(c.f, 3, (int) (c.fbytes / (3 * sizeof(double))), c.dyn, s);
  else
    e = launch_zero(c.f, c.fbytes, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "zero f");
  if (c.vatom && (e = launch_zero(c.vatom, c.vbytes, s)) != hipSuccess)
    return hip_fail(ctx, e, "zero vatom");
  if (c.zero_ew && (e = launch_zero(eatom, (size_t) c.natoms * 7 * sizeof(double), s)) != hipSuccess)
    return hip_fail(ctx, e, "zero eatom");
  if (pe && (e = hipEventRecord(pe->ev[3], s)) != hipSuccess) return hip_fail(ctx, e, "timing event");
  e = launch_force(ctx->geo, c.geo, c.pos, c.cap, row2atom, w.rl.nrows, (const float4 *) ctx->tab.p,
                   kl, (const float *) w.dEdG.p, (const double *) w.E.p, c.scale, c.f, eatom,
                   watom, c.vatom, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "force launch");
  if (pe && (e = hipEventRecord(pe->ev[4], s)) != hipSuccess) return hip_fail(ctx, e, "timing event");
  if (c.struct_off)
    e = launch_reduce_struct(eatom, watom, c.struct_off, c.nstruct, c.struct_out, s, c.dyn);
  else
    e = launch_reduce_ev(eatom, watom, c.natoms, (double *) w.red.p, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "reduce launch");
  if (pe && (e = hipEventRecord(pe->ev[5], s)) != hipSuccess) return hip_fail(ctx, e, "timing event");
  return TMD_OK;
}

// rows grouped by element (list order inside an element, tensormd.cpp:377-387),
// each group padded to kRowTile
void make_rows(int nelt, const std::vector<int> &atoms, const std::vector<int> &elem_of_atom,
               RowLayout &rl, std::vector<int> &row2atom)
{
  std::vector<int> cnt(nelt, 0);
  for (int a : atoms) cnt[elem_of_atom[a]]++;
  std::memset(&rl, 0, sizeof(rl));
  int off = 0;
  for (int e = 0; e < nelt; e++) {
    rl.row_off[e] = off;
    off += (cnt[e] + kRowTile - 1) / kRowTile * kRowTile;
  }
  for (int e = nelt; e <= kMaxElts; e++) rl.row_off[e] = off;
  rl.nrows = off;
  row2atom.assign(off, -1);
  std::vector<int> fill(nelt);
  for (int e = 0; e < nelt; e++) fill[e] = rl.row_off[e];
  for (int a : atoms) row2atom[fill[elem_of_atom[a]]++] = a;
}

}  // namespace tmd

// ===========================================================================
extern "C" {

void tmd_options_default(tmd_options *opt)
{
  if (!opt) return;
  opt->mlp_precision = TMD_MLP_FP32;
  opt->interp_dx = 0.001;
  opt->cmax_exact = 1;
  opt->device = -1;
  opt->table_device = 0;
}

const char *tmd_last_error(const tmd_ctx *ctx)
{
  if (ctx) return ctx->err.c_str();
  return g_thread_err.c_str();
}

int tmd_create(const char *npz_path, const char *const *elem_symbols, int ntypes,
               const tmd_options *opt, tmd_ctx **out)
{
  if (!out) {
    set_err(nullptr, "tmd_create: out is NULL");
    return TMD_ERR_ARG;
  }
  *out = nullptr;
  if (!npz_path || ntypes < 1 || !elem_symbols) {
    set_err(nullptr, "tmd_create: invalid arguments");
    return TMD_ERR_ARG;
  }
  std::unique_ptr<tmd_ctx> ctx(new tmd_ctx());
  if (opt)
    ctx->opt = *opt;
  else
    tmd_options_default(&ctx->opt);
  if (ctx->opt.mlp_precision < TMD_MLP_FP32 || ctx->opt.mlp_precision > TMD_MLP_BF16X3) {
    set_err(nullptr, "tmd_create: unknown mlp_precision");
    return TMD_ERR_ARG;
  }

  std::string err;
  if (!load_model(npz_path, ctx->opt.interp_dx, ctx->model, err, ctx->opt.table_device == 0)) {
    set_err(nullptr, err);
    return TMD_ERR_MODEL;
  }
  if (ctx->model.tdnp && ctx->opt.mlp_precision != TMD_MLP_FP32) {
    set_err(nullptr, "TDNP models run the FP32 path only (mlp_precision must be TMD_MLP_FP32)");
    return TMD_ERR_MODEL;
  }
  // PairTensorMD::setup_eltmap (pair_tensormd.cpp:349-370)
  ctx->ntypes = ntypes;
  ctx->type_map.assign(ntypes + 1, -1);
  int mapped = 0;
  for (int t = 1; t <= ntypes; t++) {
    const char *s = elem_symbols[t - 1];
    if (!s) continue;
    for (int e = 0; e < ctx->model.nelt; e++)
      if (ctx->model.symbols[e] == s) ctx->type_map[t] = e;
    if (ctx->type_map[t] >= 0) mapped++;
  }
  if (mapped == 0) {
    set_err(nullptr, "Incorrect args for pair coefficients: no type maps to a model element");
    return TMD_ERR_ARG;
  }
  int dev = ctx->opt.device;
  hipError_t e;
  if (dev >= 0) {
    if ((e = hipSetDevice(dev)) != hipSuccess) return hip_fail(nullptr, e, "hipSetDevice");
  } else {
    if ((e = hipGetDevice(&dev)) != hipSuccess) return hip_fail(nullptr, e, "hipGetDevice");
  }
  ctx->device = dev;
  if (ctx->opt.table_device && !build_table_device(ctx->model, ctx->opt.interp_dx, err)) {
    set_err(nullptr, err);
    return TMD_ERR_DEVICE;
  }
  if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
    return hip_fail(nullptr, e, "hipStreamCreate");
  {
    std::string serr;
    if (noalloc_scratch_setup(dev, serr) != TMD_OK) {
      set_err(nullptr, serr);
      return TMD_ERR_DEVICE;
    }
  }
  int rc = upload_model(ctx.get());
  if (rc != TMD_OK) return rc;
  *out = ctx.release();
  return TMD_OK;
}

tmd_ctx::~tmd_ctx()
{
  if (stream) hipStreamDestroy(stream);
}

void tmd_destroy(tmd_ctx *ctx)
{
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->server) {
    server_destroy(ctx->server);
    ctx->server = nullptr;
  }
  hipStreamSynchronize(ctx->stream);
  ctx->list_ws = Workspace();
  ctx->list_host.release();
  ctx->tab.release();
  ctx->params.release();
  delete ctx;
}

int tmd_cutoff(const tmd_ctx *ctx, double *rc)
{
  if (!ctx || !rc) return TMD_ERR_ARG;
  *rc = ctx->model.rc;
  return TMD_OK;
}

int tmd_masses(const tmd_ctx *ctx, double *mass_by_type)
{
  if (!ctx || !mass_by_type) return TMD_ERR_ARG;
  mass_by_type[0] = 0.0;
  for (int t = 1; t <= ctx->ntypes; t++)
    mass_by_type[t] = ctx->type_map[t] >= 0 ? ctx->model.masses[ctx->type_map[t]] : 0.0;
  return TMD_OK;
}

int tmd_type_map(const tmd_ctx *ctx, int *map)
{
  if (!ctx || !map) return TMD_ERR_ARG;
  for (int t = 0; t <= ctx->ntypes; t++) map[t] = ctx->type_map[t];
  return TMD_OK;
}

int tmd_model_info(const tmd_ctx *ctx, int *info)
{
  if (!ctx || !info) return TMD_ERR_ARG;
  info[0] = ctx->model.nelt;
  info[1] = ctx->model.k;
  info[2] = ctx->model.max_moment_file;
  info[3] = ctx->model.gwidth_file();
  info[4] = ctx->model.tdnp ? ctx->model.tdH[0].nlayers : ctx->model.mlp[0].nlayers;
  return TMD_OK;
}

int tmd_radial_table(const tmd_ctx *ctx, double *tab, long cap, int *n, double *dx)
{
  if (!ctx || !n || !dx) return TMD_ERR_ARG;
  const tmd::Model &m = ctx->model;
  *n = m.tab_n;
  *dx = m.dx;
  if (!tab) return TMD_OK;
  if (cap < (long) m.tab.size()) return TMD_ERR_CAPACITY;
  std::copy(m.tab.begin(), m.tab.end(), tab);
  return TMD_OK;
}

int tmd_compute_list(tmd_ctx *ctx, int nlocal, int nall, const double *x, const int *type,
                     int inum, const int *ilist, const int *numneigh,
                     const int *const *firstneigh, double scale, double *f, double *eatom,
                     double *vatom, double *energy, double *virial)
{
  ListArgs a;
  a.nlocal = nlocal;
  a.nall = nall;
  a.x = x;
  a.type = type;
  a.inum = inum;
  a.ilist = ilist;
  a.numneigh = numneigh;
  a.firstneigh = firstneigh;
  a.scale = scale;
  a.f = f;
  a.eatom = eatom;
  a.vatom = vatom;
  a.energy = energy;
  a.virial = virial;
  return list_compute(ctx, a, ListResources(), -1);
}

int tmd_compute_list_serial(tmd_ctx *ctx, long long list_serial, int nlocal, int nall,
                            const double *x, const int *type, int inum, const int *ilist,
                            const int *numneigh, const int *const *firstneigh, double scale,
                            double *f, double *eatom, double *vatom, double *energy, double *virial)
{
  if (list_serial < 0) {
    set_err(ctx, "tmd_compute_list_serial: list_serial must be >= 0");
    return TMD_ERR_ARG;
  }
  ListArgs a;
  a.nlocal = nlocal;
  a.nall = nall;
  a.x = x;
  a.type = type;
  a.inum = inum;
  a.ilist = ilist;
  a.numneigh = numneigh;
  a.firstneigh = firstneigh;
  a.scale = scale;
  a.f = f;
  a.eatom = eatom;
  a.vatom = vatom;
  a.energy = energy;
  a.virial = virial;
  a.who = "tmd_compute_list_serial";
  return list_compute(ctx, a, ListResources(), list_serial);
}

int tmd_descriptors_list(tmd_ctx *ctx, int nlocal, int nall, const double *x, const int *type, int inum,
                         const int *ilist, const int *numneigh, const int *const *firstneigh, double *G)
{
  if (!G) {
    set_err(ctx, "tmd_descriptors_list: G is NULL");
    return TMD_ERR_ARG;
  }
  ListArgs a;
  a.nlocal = nlocal;
  a.nall = nall;
  a.x = x;
  a.type = type;
  a.inum = inum;
  a.ilist = ilist;
  a.numneigh = numneigh;
  a.firstneigh = firstneigh;
  a.Gout = G;
  a.who = "tmd_descriptors_list";
  return list_compute(ctx, a, ListResources(), -1);
}

int tmd_set_electron_temperature(tmd_ctx *ctx, double etemp_ev)
{
  if (!ctx || !(etemp_ev >= 0.0)) {
    set_err(ctx, "tmd_set_electron_temperature: invalid arguments");
    return TMD_ERR_ARG;
  }
  ctx->etemp = etemp_ev;
  if (!ctx->model.tdnp) return TMD_OK;  // only TDNP models read it
  if (hipError_t de = hipSetDevice(ctx->device); de != hipSuccess) return hip_fail(ctx, de, "hipSetDevice");
  // stream-ordered: evaluations already queued keep the temperature they
  // were queued with
  hipError_t e = hipMemcpyAsync(ctx->etemp_dev.p, &ctx->etemp, sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? TMD_OK : hip_fail(ctx, e, "etemp upload");
}

int tmd_tdnp_results(const tmd_ctx *ctx, double *entropy, double *free_energy, double *satom, double *aatom)
{
  if (!ctx) return TMD_ERR_ARG;
  if (!ctx->model.tdnp) return TMD_ERR_MODEL;
  if (!ctx->td_valid) return TMD_ERR_STATE;
  if (entropy) *entropy = ctx->td_entropy;
  if (free_energy) *free_energy = ctx->td_free_energy;
  if (satom) std::copy(ctx->td_satom.begin(), ctx->td_satom.end(), satom);
  if (aatom) std::copy(ctx->td_aatom.begin(), ctx->td_aatom.end(), aatom);
  return TMD_OK;
}

int tmd_device_memory(size_t *free_bytes, size_t *total_bytes)
{
  if (!free_bytes || !total_bytes) return TMD_ERR_ARG;
  hipError_t e = hipMemGetInfo(free_bytes, total_bytes);
  if (e != hipSuccess) return hip_fail(nullptr, e, "tmd_device_memory");
  return TMD_OK;
}

// STREAM-like copy: bytes read = bytes written = `bytes`.
int tmd_bench_copy(size_t bytes, int iters, double *gbps)
{
  if (bytes < 16 || iters < 1 || !gbps) return TMD_ERR_ARG;
  bytes &= ~(size_t) 15;
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_zero(a, bytes, s);
  if (e == hipSuccess) e = launch_copy(a, b, bytes, s);  // warm
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int q = 0; q < iters && e == hipSuccess; q++) e = launch_copy(a, b, bytes, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) *gbps = 2.0 * (double) bytes * iters / (ms * 1e-3) / 1e9;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (s) hipStreamDestroy(s);
  if (a) hipFree(a);
  if (b) hipFree(b);
  if (e != hipSuccess) return hip_fail(nullptr, e, "tmd_bench_copy");
  return TMD_OK;
}

}  // extern "C"
