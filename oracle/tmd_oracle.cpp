// FP64 CPU oracle for the pair_style tensormd force evaluation.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
// path: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are
// the only callers.  Nothing in the product links or calls it.
//
// It is a from-scratch restatement of the reference semantics (the reference
// has no runnable CPU path: src/ML-TENSORMD/tensormd.cpp:77,389-432,562-604
// always call the GPU library).  Each function names the reference lines it
// follows.  Everything here is IEEE double; no FMA contraction is allowed in
// the geometry (the Makefile builds with -ffp-contract=off) so that the
// neighbor set is bit-identical to the device build, which evaluates the same
// expression in the same order.
//
// Parity status per function (what pins it) is listed in each header comment.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum Act { RELU = 0, SOFTPLUS = 1, TANH = 2, SQUAREPLUS = 3 };

struct Dense {
  int nlayers = 0;
  std::vector<int> sizes;  // sizes[0] = input width, sizes[l+1] = layer l output
  int act = 0;
  bool resnet = false;
  bool outbias = false;
  std::vector<std::vector<double>> W;  // W[l] row-major [in][out]
  std::vector<std::vector<double>> B;  // B[l] [out] (empty if absent)
};

struct Model {
  std::map<std::string, std::vector<double>> dkeys;
  std::map<std::string, std::vector<int>> ikeys;

  int nelt = 0;
  double rc = 0;
  int max_moment = 0;
  int nm = 0;   // max_moment + 1
  int nd = 0;   // number of monomials
  int k = 0;    // num_filters
  int fctype = 0;
  bool T_sym = false;
  bool fp32_model = false;
  std::vector<double> T;  // [nm][nd]
  std::vector<int> mono;  // [nd][3] exponents
  int eltij[8][8];
  Dense fnn;
  bool use_fnn = true;
  // analytic radial functions (use_fnn 0): kind and [parameter][filter]
  int radial_fn = 0;
  std::vector<std::vector<double>> rparams;
  std::vector<Dense> mlp;
  std::vector<double> masses;
  // Kerley table: tab[(n*4 + q)*k + j], q = 0:y 1:s 2:c1 3:c2
  int tab_n = 0;
  double dx = 0, xdx = 0;
  std::vector<double> tab;
  bool ready = false;
  std::string err;
  // TDNP (tdnp = 1): per element H, U and S networks instead of the main
  // MLP; the electron temperature (eV) and the per-atom electron entropy and
  // free energy of the last compute call (scaled), with their sums
  bool tdnp = false;
  bool sommerfeld = false;
  int ydim = 0;
  std::vector<Dense> tdH, tdU, tdS;
  double etemp = 0.0;
  std::vector<double> last_s, last_a;
  double last_S = 0.0, last_A = 0.0;
  // Reduced-precision emulation (tmdo_set_emulation): bit 0 descriptor,
  // bit 1 MLP, bit 2 force stage in FP32; bit 3 runs the emulated chain
  // even when no stage is FP32 (it must then agree with the FP64 chain).
  int emul = 0;
};

// ---------------------------------------------------------------------------
// Monomials M_d(u) in the reference order (gpu/cuda/kernel.cu:1168-1406):
// order n = 0..max_moment; inside an order, x-exponent descending, then
// y-exponent descending.  Multiplicities T[m][d] = m!/(a!b!c!) reproduce the
// table at src/ML-TENSORMD/tensormd.cpp:205-276 (pinned by a unit test that
// compares the closed form with the literal table and with the identity
// sum_d T[m][d] M_d(u)^2 = |u|^(2m)).
void build_monomials(Model &m)
{
  m.mono.clear();
  for (int n = 0; n <= m.max_moment; n++)
    for (int a = n; a >= 0; a--)
      for (int b = n - a; b >= 0; b--) {
        m.mono.push_back(a);
        m.mono.push_back(b);
        m.mono.push_back(n - a - b);
      }
  m.nd = (int) m.mono.size() / 3;
  m.nm = m.max_moment + 1;
  auto fact = [](int n) {
    double f = 1;
    for (int i = 2; i <= n; i++) f *= i;
    return f;
  };
  m.T.assign((size_t) m.nm * m.nd, 0.0);
  for (int d = 0; d < m.nd; d++) {
    int a = m.mono[3 * d], b = m.mono[3 * d + 1], c = m.mono[3 * d + 2];
    int n = a + b + c;
    m.T[n * m.nd + d] = fact(n) / (fact(a) * fact(b) * fact(c));
  }
  // tensormd.cpp:217,229-233: traceless corrections, only for max_moment <= 3
  if (m.T_sym && m.max_moment <= 3) {
    if (m.max_moment > 1) m.T[2 * m.nd + 0] = -1.0 / 3.0;
    if (m.max_moment > 2) {
      m.T[3 * m.nd + 1] = -3.0 / 5.0;
      m.T[3 * m.nd + 2] = -3.0 / 5.0;
      m.T[3 * m.nd + 3] = -3.0 / 5.0;
    }
  }
}

inline double ipow(double x, int e)
{
  double r = 1.0;
  for (int i = 0; i < e; i++) r *= x;
  return r;
}

// M_d(u) and dM_d/du_beta for all d.
void monomials(const Model &m, const double u[3], double *M, double *dMdu /*[3][nd] or null*/)
{
  for (int d = 0; d < m.nd; d++) {
    int a = m.mono[3 * d], b = m.mono[3 * d + 1], c = m.mono[3 * d + 2];
    double px = ipow(u[0], a), py = ipow(u[1], b), pz = ipow(u[2], c);
    M[d] = px * py * pz;
    if (dMdu) {
      dMdu[0 * m.nd + d] = a ? a * ipow(u[0], a - 1) * py * pz : 0.0;
      dMdu[1 * m.nd + d] = b ? b * px * ipow(u[1], b - 1) * pz : 0.0;
      dMdu[2 * m.nd + d] = c ? c * px * py * ipow(u[2], c - 1) : 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Activations.  Forward/derivative pairs follow gpu/cuda/nnp.cu:9-49 (softplus,
// squareplus on the resnet path) and activation.hpp:158-261 (relu, tanh on the
// dense path).  The derivative is always the true derivative of the forward
// map; the reference's dense-path squareplus/relu gradient (activation.hpp:
// 173-177, 209-221, applied to post-activation values by nnp.cpp:352) is a
// known defect that is not reproduced.
inline double act_fwd(int act, double z, double *dz)
{
  switch (act) {
    case SOFTPLUS: {
      // threshold = log(eps) + 2 (nnp.cu:30); linear above it
      const double thr = std::log(2.22045e-16) + 2.0;
      double e = std::exp(z);
      double h = (z < -thr) ? std::log1p(e) : z;
      *dz = 1.0 / (1.0 + 1.0 / e);
      return h;
    }
    case SQUAREPLUS: {
      double s = std::sqrt(z * z + 4.0);
      double h = 0.5 * (z + s);
      *dz = h / s;
      return h;
    }
    case TANH: {
      double h = std::tanh(z);
      *dz = 1.0 - h * h;
      return h;
    }
    default: {
      double h = z >= 0 ? z : 0.0;
      *dz = z >= 0 ? 1.0 : 0.0;
      return h;
    }
  }
}

// Dense network forward (+ optional backward).  Semantics of
// NeuralNetworkPotential::forward_resnet/forward_dense (nnp.cpp:222-303,
// gpu/cuda/nnp.cu:200-281): z_l = h_l W_l + b_l; for hidden layers
// h_{l+1} = act(z_l) + (resnet_l ? h_l : 0), resnet_l = resnet && l > 0 &&
// in == out (nnp.cu:245); the last layer is linear (+ bias if outbias).
// Backward with seed `gout` follows nnp.cu:149-198 / nnp.cpp:325-366.
struct DenseWork {
  std::vector<std::vector<double>> h, dz;  // per layer
};

void dense_forward(const Dense &net, const double *x, double *y, DenseWork &w)
{
  int L = net.nlayers;
  w.h.resize(L);
  w.dz.resize(L);
  w.h[0].assign(x, x + net.sizes[0]);
  for (int l = 0; l < L; l++) {
    int K = net.sizes[l], N = net.sizes[l + 1];
    const double *W = net.W[l].data();
    std::vector<double> z(N, 0.0);
    for (int i = 0; i < K; i++) {
      double hi = w.h[l][i];
      if (hi == 0.0) continue;
      const double *Wr = W + (size_t) i * N;
      for (int j = 0; j < N; j++) z[j] += hi * Wr[j];
    }
    if (!net.B[l].empty())
      for (int j = 0; j < N; j++) z[j] += net.B[l][j];
    if (l == L - 1) {
      for (int j = 0; j < N; j++) y[j] = z[j];
      break;
    }
    bool res = net.resnet && l > 0 && K == N;
    std::vector<double> hn(N), d(N);
    for (int j = 0; j < N; j++) {
      hn[j] = act_fwd(net.act, z[j], &d[j]);
      if (res) hn[j] += w.h[l][j];
    }
    w.h[l + 1] = hn;
    w.dz[l] = d;
  }
}

// grad wrt input; gy is d(out)/d(y) seed (length sizes[L]).
void dense_backward(const Dense &net, const DenseWork &w, const double *gy, double *gx)
{
  int L = net.nlayers;
  std::vector<double> g(gy, gy + net.sizes[L]);
  for (int l = L - 1; l >= 0; l--) {
    int K = net.sizes[l], N = net.sizes[l + 1];
    // g is dE/d(output of layer l); for hidden layers multiply by act'
    std::vector<double> gz(N);
    if (l == L - 1) {
      gz = g;
    } else {
      for (int j = 0; j < N; j++) gz[j] = g[j] * w.dz[l][j];
    }
    std::vector<double> gin(K, 0.0);
    const double *W = net.W[l].data();
    for (int i = 0; i < K; i++) {
      const double *Wr = W + (size_t) i * N;
      double s = 0.0;
      for (int j = 0; j < N; j++) s += Wr[j] * gz[j];
      gin[i] = s;
    }
    bool res = (l < L - 1) && net.resnet && l > 0 && K == N;
    if (res)
      for (int i = 0; i < K; i++) gin[i] += g[i];
    g.swap(gin);
  }
  for (int i = 0; i < net.sizes[0]; i++) gx[i] = g[i];
}

// ---------------------------------------------------------------------------
// Radial table: y_n = FNN(n dx) * fc(n dx), n = 0..N-1, N = int(rc/dx) + 1
// evaluated in the model's Scalar type (fnn.cpp:82-102); Kerley coefficients
// as BatchRation1d::setup (ration.hpp:166-225).  Pinned by the golden log
// energies of the shipped models and by node/continuity tests.
double cutoff_fc(const Model &m, double r)
{
  if (m.fctype == 0) {
    // cutoff.cpp:43-46
    double Z = std::min(r / m.rc, 1.0) * M_PI;
    return 0.5 + 0.5 * std::cos(Z);
  }
  // cutoff.cpp:47-51 (gamma = 5); note cwiseMin(pi) as in the reference
  const double gamma = 5.0;
  double Z = std::min(r / m.rc, M_PI);
  double Z4 = Z * Z * Z * Z, Z5 = Z4 * Z;
  return 1.0 + gamma * Z5 * Z - (gamma + 1.0) * Z5;
}

// Analytic radial functions of a use_fnn = 0 model, following
// Descriptor::pexp/morse/density/sf (src/ML-TENSORMD/descriptor.cpp:51-132):
//   0 PowerExp  exp(-(r / rl)^pl)
//   1 Morse     D (exp(-2 gamma (r - r0)) - 2 exp(-gamma (r - r0)))
//   2 Density   A exp(-beta (r - re) / re)
//   3 SF        exp(-eta (r - omega)^2 / rc^2)
// Pinned against the closed forms evaluated independently in numpy and, via
// the table, against the GPU path.
void radial_fn_values(const Model &m, double r, double *y)
{
  const auto &p = m.rparams;
  for (int j = 0; j < m.k; j++) {
    double v;
    if (m.radial_fn == 0) {
      v = std::exp(-std::pow(r / p[0][j], p[1][j]));
    } else if (m.radial_fn == 1) {
      double g = p[1][j] * (r - p[2][j]);
      v = p[0][j] * (std::exp(-2.0 * g) - 2.0 * std::exp(-g));
    } else if (m.radial_fn == 2) {
      v = p[0][j] * std::exp(-p[1][j] * (r - p[2][j]) / p[2][j]);
    } else {
      double d = r - p[1][j];
      v = std::exp(-p[0][j] * d * d / (m.rc * m.rc));
    }
    y[j] = v;
  }
}

// the radial functions before the cutoff: the FNN or the analytic ones
void radial_values(const Model &m, double r, double *y, DenseWork &w)
{
  if (m.use_fnn)
    dense_forward(m.fnn, &r, y, w);
  else
    radial_fn_values(m, r, y);
}

void build_table(Model &m, double delta)
{
  int n;
  std::vector<double> xs;
  if (m.fp32_model) {
    float d = (float) delta, rcf = (float) m.rc;
    n = (int) (rcf / d) + 1;
    for (int i = 0; i < n; i++) xs.push_back((double) ((float) i * d));
    m.dx = (double) d;
    m.xdx = (double) (1.0f / d);
  } else {
    n = (int) (m.rc / delta) + 1;
    for (int i = 0; i < n; i++) xs.push_back((double) i * delta);
    m.dx = delta;
    m.xdx = 1.0 / delta;
  }
  int k = m.k;
  m.tab_n = n;
  m.tab.assign((size_t) n * 4 * k, 0.0);
  DenseWork w;
  std::vector<double> y(k);
  for (int i = 0; i < n; i++) {
    double x = xs[i];
    radial_values(m, x, y.data(), w);
    double fc = cutoff_fc(m, x);
    for (int j = 0; j < k; j++) m.tab[((size_t) i * 4 + 0) * k + j] = y[j] * fc;
  }
  double xdx = m.xdx;
  for (int i = 0; i < n - 1; i++)
    for (int j = 0; j < k; j++)
      m.tab[((size_t) i * 4 + 1) * k + j] =
          (m.tab[((size_t) (i + 1) * 4) * k + j] - m.tab[((size_t) i * 4) * k + j]) * xdx;
  for (int i = 0; i < n - 1; i++) {
    for (int j = 0; j < k; j++) {
      double s = m.tab[((size_t) i * 4 + 1) * k + j];
      double c1, c2;
      if (i == 0) {
        double sp = m.tab[((size_t) (i + 1) * 4 + 1) * k + j];
        if (s * (s - (sp - s) / 2) <= 0)
          c2 = s * xdx;
        else
          c2 = (sp - s) * xdx / 2;
        c1 = 0;
      } else if (i == n - 2) {
        double sm = m.tab[((size_t) (i - 1) * 4 + 1) * k + j];
        c1 = (s - sm) * xdx / 2;
        c2 = 0;
      } else {
        double sm = m.tab[((size_t) (i - 1) * 4 + 1) * k + j];
        double sp = m.tab[((size_t) (i + 1) * 4 + 1) * k + j];
        if (i == 1 && sm * (sm - (s - sm) / 2) <= 0)
          c1 = (s - 2 * sm) * xdx;
        else
          c1 = (s - sm) * xdx / 2;
        c2 = (sp - s) * xdx / 2;
      }
      m.tab[((size_t) i * 4 + 2) * k + j] = c1;
      m.tab[((size_t) i * 4 + 3) * k + j] = c2;
    }
  }
}

// Evaluation of H_k(r), dH_k/dr: gpu/cuda/ration.cu:27-69 with
// endpoints_force_zero (ration.hpp:29).
void radial(const Model &m, double r, double *H, double *dH)
{
  int k = m.k;
  int i = (int) (r * m.xdx);
  i = std::min(m.tab_n - 2, i);
  if (i <= 0 || i == m.tab_n - 2) {
    for (int j = 0; j < k; j++) H[j] = dH[j] = 0.0;
    return;
  }
  double dx = m.dx;
  double q = r - i * dx;
  double p = dx - q;
  const double *Y = &m.tab[((size_t) i * 4 + 0) * k];
  const double *S = &m.tab[((size_t) i * 4 + 1) * k];
  const double *C1 = &m.tab[((size_t) i * 4 + 2) * k];
  const double *C2 = &m.tab[((size_t) i * 4 + 3) * k];
  for (int j = 0; j < k; j++) {
    double c1 = C1[j], c2 = C2[j], s = S[j];
    double c3 = std::fabs(c2 * p);
    double c4 = c3 + std::fabs(c1 * q);
    if (c4 > 0.0) c3 /= c4;
    c4 = c2 + c3 * (c1 - c2);
    H[j] = Y[j] + q * (s - p * c4);
    dH[j] = s + (q - p) * c4 + dx * (c4 - c2) * (1 - c3);
  }
}

// ---------------------------------------------------------------------------
// Per-atom chain: descriptor (kernel.cu:1122-1416 slots, :1028-1061 G),
// MLP (nnp), dE/dS (kernel.cu:730-817), V/U/F1/F2 (kernel.cu:869-965,
// :583-646, :76-557), scatter (kernel.cu:1639-1758).
struct Nbr {
  double d[3];  // displacement x_j - x_i (with image)
  int j;        // scatter target
  int ej;       // element of j
};

struct AtomOut {
  double E;
  double s = 0.0, a = 0.0;  // TDNP: electron entropy and free energy (unscaled)
  std::vector<double> F;  // [nnbr][3] dE_i/dDelta (scaled)
  std::vector<double> G, dEdG;
};

// TDNP atom (TDNP::forward/backward, src/ML-TENSORMD/tdnp.cpp:150-218, with
// set_element_tdnp_device and the axpy of backward_tdnp_GPU,
// gpu/cuda/nnp.cu:294-316): y = H(G) with the electron temperature T (eV)
// appended as a last column (tdnp.cpp:83-113 give H's last layer a zero
// weight column that the temperature then overwrites); u = U(y), s = S(y)
// (times T for a Sommerfeld model); the atom's energy is u, its free energy
// a = u - T s, and the gradient that drives the forces is dA/dG: the
// backward pass of H seeded with dU/dy - T dS/dy (T^2 for Sommerfeld) with
// the temperature column's gradient zeroed.  Pinned by torch autograd of
// the composed networks and by finite differences of the free energy.
void tdnp_atom(const Model &m, int ei, const std::vector<double> &G, double &u, double &s, double &a,
               std::vector<double> &g)
{
  const Dense &H = m.tdH[ei], &U = m.tdU[ei], &S = m.tdS[ei];
  DenseWork wh, wu, ws;
  std::vector<double> y(m.ydim);
  dense_forward(H, G.data(), y.data(), wh);
  const double T = m.etemp;
  y[m.ydim - 1] = T;
  dense_forward(U, y.data(), &u, wu);
  dense_forward(S, y.data(), &s, ws);
  if (m.sommerfeld) s *= T;
  a = u - T * s;
  std::vector<double> gu(m.ydim), gs(m.ydim);
  const double one = 1.0;
  dense_backward(U, wu, &one, gu.data());
  dense_backward(S, ws, &one, gs.data());
  const double sc = m.sommerfeld ? T * T : T;
  std::vector<double> gy(m.ydim - 1);
  for (int j = 0; j < m.ydim - 1; j++) gy[j] = gu[j] - sc * gs[j];
  g.assign(H.sizes[0], 0.0);
  dense_backward(H, wh, gy.data(), g.data());
}

void atom_chain_fp64(const Model &m, int ei, const std::vector<Nbr> &nb, double scale, AtomOut &out,
                     bool keep_desc)
{
  const int k = m.k, nd = m.nd, nm = m.nm, nb_slots = m.nelt;
  const double rc2 = m.rc * m.rc;
  size_t n = nb.size();
  // geometry per neighbor
  std::vector<int> keep(n, 0), slot(n, 0);
  std::vector<double> r(n), u(3 * n), M(n * nd), H(n * k), dH(n * k);
  std::vector<double> P((size_t) nb_slots * k * nd, 0.0);
  for (size_t c = 0; c < n; c++) {
    const double *dl = nb[c].d;
    double r2 = dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2];
    if (!(r2 < rc2 && r2 > 1e-10)) continue;  // kernel.cu:1163
    keep[c] = 1;
    slot[c] = m.eltij[ei][nb[c].ej];
    r[c] = std::sqrt(r2);
    double rinv = 1.0 / r[c];
    for (int a = 0; a < 3; a++) u[3 * c + a] = dl[a] * rinv;
    monomials(m, &u[3 * c], &M[c * nd], nullptr);
    radial(m, r[c], &H[c * k], &dH[c * k]);
    double *Pb = &P[(size_t) slot[c] * k * nd];
    for (int j = 0; j < k; j++)
      for (int d = 0; d < nd; d++) Pb[j * nd + d] += H[c * k + j] * M[c * nd + d];
  }
  // G[b][k][m]
  int gw = nb_slots * k * nm;
  std::vector<double> G(gw, 0.0);
  for (int b = 0; b < nb_slots; b++)
    for (int j = 0; j < k; j++) {
      const double *Pp = &P[((size_t) b * k + j) * nd];
      for (int mm = 0; mm < nm; mm++) {
        double s = 0.0;
        for (int d = 0; d < nd; d++) s += m.T[mm * nd + d] * Pp[d] * Pp[d];
        G[(b * k + j) * nm + mm] = s;
      }
      double g0 = std::sqrt(G[(b * k + j) * nm + 0]);
      G[(b * k + j) * nm + 0] = Pp[0] < 0 ? -g0 : g0;  // kernel.cu:1048,1058
    }
  // MLP
  double E;
  std::vector<double> g(gw);
  if (m.tdnp) {
    tdnp_atom(m, ei, G, E, out.s, out.a, g);
  } else {
    const Dense &net = m.mlp[ei];
    DenseWork w;
    dense_forward(net, G.data(), &E, w);
    double one = 1.0;
    dense_backward(net, w, &one, g.data());
  }
  out.E = E;
  if (keep_desc) {
    out.G = G;
    out.dEdG = g;
  }
  // dE/dS (kernel.cu:752-767): g0' = g0/(2 G0) guarded by |G0| >= 1e-8
  std::vector<double> S((size_t) nb_slots * k * nd);
  for (int b = 0; b < nb_slots; b++)
    for (int j = 0; j < k; j++) {
      double gp[8];
      for (int mm = 0; mm < nm; mm++) gp[mm] = g[(b * k + j) * nm + mm];
      double G0 = G[(b * k + j) * nm + 0];
      gp[0] = (std::fabs(G0) < 1e-8) ? 0.0 : gp[0] / G0 * 0.5;
      const double *Pp = &P[((size_t) b * k + j) * nd];
      for (int d = 0; d < nd; d++) {
        double s = 0.0;
        for (int mm = 0; mm < nm; mm++) s += gp[mm] * 2.0 * m.T[mm * nd + d];
        S[((size_t) b * k + j) * nd + d] = s * Pp[d];
      }
    }
  // per-neighbor force
  out.F.assign(3 * n, 0.0);
  std::vector<double> dMdu(3 * nd), Mc(nd);
  for (size_t c = 0; c < n; c++) {
    if (!keep[c]) continue;
    const double *Sb = &S[(size_t) slot[c] * k * nd];
    const double *uc = &u[3 * c];
    monomials(m, uc, Mc.data(), dMdu.data());
    // F1 = (sum_k U_k dH_k) u,  U_k = sum_d S[k][d] M_d
    double f1 = 0.0;
    std::vector<double> V(nd, 0.0);
    for (int j = 0; j < k; j++) {
      double Uk = 0.0;
      for (int d = 0; d < nd; d++) Uk += Sb[j * nd + d] * M[c * nd + d];
      f1 += Uk * dH[c * k + j];
      double h = H[c * k + j];
      for (int d = 0; d < nd; d++) V[d] += Sb[j * nd + d] * h;
    }
    double rinv = 1.0 / r[c];
    double F[3];
    for (int a = 0; a < 3; a++) {
      // F2_a = sum_d V_d dM_d/dDelta_a, dM/dDelta_a = sum_b dM/du_b (delta_ab - u_a u_b)/r
      double f2 = 0.0;
      for (int d = 0; d < nd; d++) {
        double dm = 0.0;
        for (int b = 0; b < 3; b++) {
          double J = ((a == b ? 1.0 : 0.0) - uc[a] * uc[b]) * rinv;
          dm += dMdu[b * nd + d] * J;
        }
        f2 += V[d] * dm;
      }
      F[a] = (f1 * uc[a] + f2) * scale;
    }
    for (int a = 0; a < 3; a++) out.F[3 * c + a] = F[a];
  }
}

// ---------------------------------------------------------------------------
// Reduced-precision emulation of the device chain (test infrastructure: it
// prices the rounding floor of an FP32 evaluation, it is not a reference).
// The algorithm is atom_chain_fp64's; the arithmetic type of each stage is a
// template parameter:
//   RD  descriptor: unit vector, monomials, table interpolation (table values
//       rounded to RD, the device's expression order), P = sum H M, G;
//   RM  MLP hidden layers forward and backward (output layer: RM products
//       summed in FP64 with the bias rounded to RM, as the device does);
//   RF  force chain: S, U, V, F1, F2 per neighbor.
// Displacements, r and the table cell stay FP64 in every stage, and with
// every type double the result equals atom_chain_fp64 to rounding.
template <class R>
inline R fmar(R a, R b, R c)
{
  return std::fma(a, b, c);
}

template <class R>
void radial_emul(const Model &m, double r, R *H, R *dH)
{
  int k = m.k;
  int i = std::min(m.tab_n - 2, (int) (r * m.xdx));
  if (i <= 0 || i == m.tab_n - 2) {
    for (int j = 0; j < k; j++) H[j] = dH[j] = R(0);
    return;
  }
  double qd = r - (double) i * m.dx;
  const R q = (R) qd, p = (R) (m.dx - qd), qp = q - p, dx = (R) m.dx;
  const double *Y = &m.tab[((size_t) i * 4 + 0) * k];
  const double *S = &m.tab[((size_t) i * 4 + 1) * k];
  const double *C1 = &m.tab[((size_t) i * 4 + 2) * k];
  const double *C2 = &m.tab[((size_t) i * 4 + 3) * k];
  for (int j = 0; j < k; j++) {
    const R y = (R) Y[j], s = (R) S[j], c1 = (R) C1[j], c2 = (R) C2[j];
    R a = std::fabs(c2) * std::fabs(p);
    R den = fmar(std::fabs(c1), std::fabs(q), a);
    R w = den > R(0) ? a / den : R(0);
    R dc = c1 - c2;
    R c = fmar(w, dc, c2);
    H[j] = fmar(q, fmar(-p, c, s), y);
    dH[j] = fmar(dx * dc, fmar(-w, w, w), fmar(qp, c, s));
  }
}

template <class R>
inline R act_emul(int act, R z, R *dz)
{
  switch (act) {
    case SOFTPLUS: {
      // the device's FP32 threshold is -(log(eps_f32) + 2); FP64 keeps eps_f64
      const R thr = sizeof(R) == 4 ? R(13.942385) : R(-(std::log(2.22045e-16) + 2.0));
      R e = std::exp(z);
      R h = z < thr ? std::log1p(e) : z;
      *dz = R(1) / (R(1) + R(1) / e);
      return h;
    }
    case SQUAREPLUS: {
      R s = std::sqrt(z * z + R(4));
      R h = z >= R(0) ? R(0.5) * (z + s) : R(2) / (s - z);
      *dz = h / s;
      return h;
    }
    case TANH: {
      R h = std::tanh(z);
      *dz = R(1) - h * h;
      return h;
    }
    default:
      *dz = z >= R(0) ? R(1) : R(0);
      return z >= R(0) ? z : R(0);
  }
}

template <class RM>
double mlp_emul(const Dense &net, const std::vector<double> &G, std::vector<double> &gx)
{
  const int L = net.nlayers;
  std::vector<std::vector<RM>> h(L), dz(L);
  h[0].resize(net.sizes[0]);
  for (int i = 0; i < net.sizes[0]; i++) h[0][i] = (RM) G[i];
  double E = 0.0;
  for (int l = 0; l < L; l++) {
    const int K = net.sizes[l], N = net.sizes[l + 1];
    const double *W = net.W[l].data();
    if (l == L - 1) {
      for (int i = 0; i < K; i++) E += (double) (h[l][i] * (RM) W[i]);
      if (!net.B[l].empty()) E += (double) (RM) net.B[l][0];
      break;
    }
    std::vector<RM> z(N, RM(0));
    for (int i = 0; i < K; i++)
      for (int j = 0; j < N; j++) z[j] = fmar(h[l][i], (RM) W[(size_t) i * N + j], z[j]);
    const bool res = net.resnet && l > 0 && K == N;
    h[l + 1].resize(N);
    dz[l].resize(N);
    for (int j = 0; j < N; j++) {
      RM zz = net.B[l].empty() ? z[j] : z[j] + (RM) net.B[l][j];
      RM v = act_emul<RM>(net.act, zz, &dz[l][j]);
      h[l + 1][j] = res ? v + h[l][j] : v;
    }
  }
  // backward, seed 1 at the output
  std::vector<RM> g(net.sizes[L - 1]);
  for (int j = 0; j < net.sizes[L - 1]; j++) g[j] = (RM) net.W[L - 1][j];
  for (int l = L - 2; l >= 0; l--) {
    const int K = net.sizes[l], N = net.sizes[l + 1];
    const double *W = net.W[l].data();
    std::vector<RM> gz(N), gin(K, RM(0));
    for (int j = 0; j < N; j++) gz[j] = g[j] * dz[l][j];
    for (int i = 0; i < K; i++) {
      RM s = RM(0);
      for (int j = 0; j < N; j++) s = fmar((RM) W[(size_t) i * N + j], gz[j], s);
      gin[i] = s;
    }
    if (net.resnet && l > 0 && K == N)
      for (int i = 0; i < K; i++) gin[i] += g[i];
    g.swap(gin);
  }
  gx.resize(net.sizes[0]);
  for (int i = 0; i < net.sizes[0]; i++) gx[i] = (double) g[i];
  return E;
}

template <class R>
void monomials_emul(const Model &m, const R u[3], R *M, R *dMdu)
{
  auto pw = [](R x, int e) {
    R r = R(1);
    for (int i = 0; i < e; i++) r *= x;
    return r;
  };
  for (int d = 0; d < m.nd; d++) {
    int a = m.mono[3 * d], b = m.mono[3 * d + 1], c = m.mono[3 * d + 2];
    R px = pw(u[0], a), py = pw(u[1], b), pz = pw(u[2], c);
    M[d] = px * py * pz;
    if (dMdu) {
      dMdu[0 * m.nd + d] = a ? R(a) * pw(u[0], a - 1) * py * pz : R(0);
      dMdu[1 * m.nd + d] = b ? R(b) * px * pw(u[1], b - 1) * pz : R(0);
      dMdu[2 * m.nd + d] = c ? R(c) * px * py * pw(u[2], c - 1) : R(0);
    }
  }
}

template <class RD, class RM, class RF>
void atom_chain_emul(const Model &m, int ei, const std::vector<Nbr> &nb, double scale,
                     AtomOut &out, bool keep_desc)
{
  const int k = m.k, nd = m.nd, nm = m.nm, nslot = m.nelt;
  const double rc2 = m.rc * m.rc;
  const size_t n = nb.size();
  std::vector<int> keep(n, 0), slot(n, 0);
  std::vector<double> r(n);
  std::vector<RD> u(3 * n), M(n * nd), H(n * k), dH(n * k);
  std::vector<RD> P((size_t) nslot * k * nd, RD(0));
  // slot-sorted kept order (the device's compaction), list order within a slot
  std::vector<size_t> order;
  for (int b = 0; b < nslot; b++)
    for (size_t c = 0; c < n; c++) {
      const double *dl = nb[c].d;
      double r2 = dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2];
      if (!(r2 < rc2 && r2 > 1e-10)) continue;
      if (m.eltij[ei][nb[c].ej] != b) continue;
      order.push_back(c);
      keep[c] = 1;
      slot[c] = b;
      r[c] = std::sqrt(r2);
    }
  for (size_t c : order) {
    const double rinv = 1.0 / r[c];
    for (int a = 0; a < 3; a++) u[3 * c + a] = (RD) (nb[c].d[a] * rinv);
    monomials_emul<RD>(m, &u[3 * c], &M[c * nd], nullptr);
    radial_emul<RD>(m, r[c], &H[c * k], &dH[c * k]);
    RD *Pb = &P[(size_t) slot[c] * k * nd];
    for (int j = 0; j < k; j++)
      for (int d = 0; d < nd; d++) Pb[j * nd + d] = fmar(H[c * k + j], M[c * nd + d], Pb[j * nd + d]);
  }
  const int gw = nslot * k * nm;
  std::vector<double> G(gw, 0.0);
  for (int b = 0; b < nslot; b++)
    for (int j = 0; j < k; j++) {
      const RD *Pp = &P[((size_t) b * k + j) * nd];
      for (int mm = 0; mm < nm; mm++) {
        RD s = RD(0);
        for (int d = 0; d < nd; d++) s = fmar((RD) m.T[mm * nd + d] * Pp[d], Pp[d], s);
        G[(b * k + j) * nm + mm] = (double) s;
      }
      G[(b * k + j) * nm + 0] = (double) Pp[0];  // sign(P0) sqrt(P0^2)
    }
  std::vector<double> g;
  if (m.tdnp)  // the TDNP networks stay FP64 in the emulation
    tdnp_atom(m, ei, G, out.E, out.s, out.a, g);
  else
    out.E = mlp_emul<RM>(m.mlp[ei], G, g);
  if (keep_desc) {
    out.G = G;
    out.dEdG = g;
  }
  std::vector<RF> S((size_t) nslot * k * nd);
  for (int b = 0; b < nslot; b++)
    for (int j = 0; j < k; j++) {
      RF gp[8];
      for (int mm = 0; mm < nm; mm++) gp[mm] = (RF) g[(b * k + j) * nm + mm];
      const RF G0 = (RF) G[(b * k + j) * nm + 0];
      gp[0] = std::fabs(G0) < RF(1e-8) ? RF(0) : gp[0] / G0 * RF(0.5);
      const RD *Pp = &P[((size_t) b * k + j) * nd];
      for (int d = 0; d < nd; d++) {
        RF s = RF(0);
        for (int mm = 0; mm < nm; mm++) s = fmar(gp[mm], RF(2) * (RF) m.T[mm * nd + d], s);
        S[((size_t) b * k + j) * nd + d] = s * (RF) Pp[d];
      }
    }
  out.F.assign(3 * n, 0.0);
  std::vector<RF> dMdu(3 * nd), Mc(nd), V(nd);
  for (size_t c : order) {
    const RF *Sb = &S[(size_t) slot[c] * k * nd];
    RF uc[3] = {(RF) u[3 * c], (RF) u[3 * c + 1], (RF) u[3 * c + 2]};
    monomials_emul<RF>(m, uc, Mc.data(), dMdu.data());
    RF f1 = RF(0);
    std::fill(V.begin(), V.end(), RF(0));
    for (int j = 0; j < k; j++) {
      RF Uk = RF(0);
      for (int d = 0; d < nd; d++) Uk = fmar(Sb[j * nd + d], Mc[d], Uk);
      f1 = fmar(Uk, (RF) dH[c * k + j], f1);
      const RF h = (RF) H[c * k + j];
      for (int d = 0; d < nd; d++) V[d] = fmar(Sb[j * nd + d], h, V[d]);
    }
    const RF rinv = (RF) (1.0 / r[c]);
    for (int a = 0; a < 3; a++) {
      RF f2 = RF(0);
      for (int d = 0; d < nd; d++) {
        RF dm = RF(0);
        for (int b = 0; b < 3; b++) {
          RF J = ((a == b ? RF(1) : RF(0)) - uc[a] * uc[b]) * rinv;
          dm = fmar(dMdu[b * nd + d], J, dm);
        }
        f2 = fmar(V[d], dm, f2);
      }
      out.F[3 * c + a] = (double) fmar(f1, uc[a], f2) * scale;
    }
  }
}

void atom_chain(const Model &m, int ei, const std::vector<Nbr> &nb, double scale, AtomOut &out,
                bool keep_desc)
{
  if (m.emul == 0) {
    atom_chain_fp64(m, ei, nb, scale, out, keep_desc);
    return;
  }
  switch (m.emul & 7) {
    case 0: atom_chain_emul<double, double, double>(m, ei, nb, scale, out, keep_desc); break;
    case 1: atom_chain_emul<float, double, double>(m, ei, nb, scale, out, keep_desc); break;
    case 2: atom_chain_emul<double, float, double>(m, ei, nb, scale, out, keep_desc); break;
    case 3: atom_chain_emul<float, float, double>(m, ei, nb, scale, out, keep_desc); break;
    case 4: atom_chain_emul<double, double, float>(m, ei, nb, scale, out, keep_desc); break;
    case 5: atom_chain_emul<float, double, float>(m, ei, nb, scale, out, keep_desc); break;
    case 6: atom_chain_emul<double, float, float>(m, ei, nb, scale, out, keep_desc); break;
    default: atom_chain_emul<float, float, float>(m, ei, nb, scale, out, keep_desc); break;
  }
}

thread_local std::string g_err;

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

const char *tmdo_last_error() { return g_err.c_str(); }

void *tmdo_new() { return new Model(); }

void tmdo_free(void *h) { delete static_cast<Model *>(h); }

// Electron temperature (eV) of a TDNP model (PairTensorMD::etemp).
int tmdo_set_etemp(void *h, double etemp)
{
  static_cast<Model *>(h)->etemp = etemp;
  return 0;
}

// TDNP results of the last compute call: per-atom electron entropy and
// free energy (scaled; zero for atoms without a row) and their sums.
int tmdo_tdnp_last(void *h, long n, double *satom, double *aatom, double *S, double *A)
{
  auto *m = static_cast<Model *>(h);
  if (!m->tdnp || (long) m->last_s.size() != n) {
    g_err = "no TDNP results of that size";
    return -1;
  }
  std::copy(m->last_s.begin(), m->last_s.end(), satom);
  std::copy(m->last_a.begin(), m->last_a.end(), aatom);
  *S = m->last_S;
  *A = m->last_A;
  return 0;
}

// Selects the reduced-precision emulation (flags as Model::emul; 0 = the
// FP64 chain).
int tmdo_set_emulation(void *h, int flags)
{
  static_cast<Model *>(h)->emul = flags;
  return 0;
}

int tmdo_set_ints(void *h, const char *key, const int *v, int n)
{
  auto *m = static_cast<Model *>(h);
  m->ikeys[key] = std::vector<int>(v, v + n);
  return 0;
}

int tmdo_set_doubles(void *h, const char *key, const double *v, long n)
{
  auto *m = static_cast<Model *>(h);
  m->dkeys[key] = std::vector<double>(v, v + n);
  return 0;
}

// Builds the model from the pushed npz keys (schema of tensormd.cpp:108-177,
// nnp.cpp:110-145, fnn.cpp:32-55) and the radial table for `interp_dx`.
int tmdo_finalize(void *h, double interp_dx)
{
  auto *m = static_cast<Model *>(h);
  auto geti = [&](const std::string &k, int def, bool req) -> int {
    auto it = m->ikeys.find(k);
    if (it == m->ikeys.end() || it->second.empty()) {
      if (req) throw std::string("missing key ") + k;
      return def;
    }
    return it->second[0];
  };
  auto getd = [&](const std::string &k) -> const std::vector<double> & {
    auto it = m->dkeys.find(k);
    if (it == m->dkeys.end()) throw std::string("missing key ") + k;
    return it->second;
  };
  try {
    m->nelt = geti("nelt", 0, true);
    if (m->nelt < 1 || m->nelt > 8) throw std::string("nelt out of range");
    m->rc = getd("rmax")[0];
    m->max_moment = geti("max_moment", 0, true);
    if (m->max_moment < 0 || m->max_moment > 5) throw std::string("max_moment out of range");
    m->fctype = geti("fctype", 0, true);
    m->T_sym = geti("is_T_symmetric", 0, false) == 1;
    m->fp32_model = geti("precision", 64, false) == 32;
    m->use_fnn = geti("use_fnn", 0, true) == 1;
    m->tdnp = geti("tdnp", 0, false) == 1;
    m->masses = getd("masses");
    build_monomials(*m);
    // element slots: tensormd.cpp:165-176
    for (int i = 0; i < m->nelt; i++) {
      int kk = 1;
      for (int j = 0; j < m->nelt; j++) m->eltij[i][j] = (i == j) ? 0 : kk++;
    }
    if (!m->use_fnn) {
      // descriptor.h:37-87: descriptor::method (default PowerExp) and one
      // array per parameter, num_filters long
      m->radial_fn = geti("descriptor::method", 0, false);
      const char *names[4][3] = {{"descriptor::rl", "descriptor::pl", nullptr},
                                 {"descriptor::D", "descriptor::gamma", "descriptor::r0"},
                                 {"descriptor::A", "descriptor::beta", "descriptor::re"},
                                 {"descriptor::eta", "descriptor::omega", nullptr}};
      if (m->radial_fn < 0 || m->radial_fn > 3) throw std::string("descriptor::method out of range");
      for (int q = 0; q < 3 && names[m->radial_fn][q]; q++) m->rparams.push_back(getd(names[m->radial_fn][q]));
      m->k = (int) m->rparams[0].size();
    }
    // FNN
    Dense &f = m->fnn;
    if (m->use_fnn) {
      f.nlayers = geti("fnn::nlayers", 0, true);
      auto fs = m->ikeys.at("fnn::layer_sizes");
      f.sizes.assign(1, 1);
      for (int v : fs) f.sizes.push_back(v);
      f.act = geti("fnn::actfn", 0, true);
      f.outbias = geti("fnn::apply_output_bias", 0, true) == 1;
      f.resnet = geti("fnn::use_resnet_dt", 0, true) == 1;
      {
        int nres = 0;
        for (int l = 1; l < f.nlayers; l++)
          if (f.sizes[l] == f.sizes[l + 1]) nres++;
        if (nres == 0) f.resnet = false;  // nnp.cpp:179-181
      }
      for (int l = 0; l < f.nlayers; l++) {
        f.W.push_back(getd("fnn::weights_0_" + std::to_string(l)));
        if (l < f.nlayers - 1 || f.outbias)
          f.B.push_back(getd("fnn::biases_0_" + std::to_string(l)));
        else
          f.B.emplace_back();
        if ((int) f.W[l].size() != f.sizes[l] * f.sizes[l + 1]) throw std::string("fnn weight shape");
      }
      m->k = geti("fnn::num_filters", 0, true);
      if (f.sizes.back() != m->k) throw std::string("fnn output width != num_filters");
    }
    // the per-element networks of one scope ("" for the main MLP, "H::",
    // "U::", "S::" for TDNP; nnp.cpp:110-205, tdnp.cpp:60-148)
    auto read_nets = [&](const std::string &scope, int nin, std::vector<Dense> &out) {
      int nl = geti(scope + "nlayers", 0, true);
      auto ls = m->ikeys.at(scope + "layer_sizes");
      int act = geti(scope + "actfn", 0, true);
      bool outb = geti(scope + "apply_output_bias", 0, true) == 1;
      bool res = geti(scope + "use_resnet_dt", 0, true) == 1;
      for (int e = 0; e < m->nelt; e++) {
        Dense d;
        d.nlayers = nl;
        d.sizes.assign(1, nin);
        for (int v : ls) d.sizes.push_back(v);
        d.act = act;
        d.outbias = outb;
        d.resnet = res;
        int nres = 0;
        for (int l = 1; l < nl; l++)
          if (d.sizes[l] == d.sizes[l + 1]) nres++;
        if (nres == 0) d.resnet = false;
        for (int l = 0; l < nl; l++) {
          d.W.push_back(getd(scope + "weights_" + std::to_string(e) + "_" + std::to_string(l)));
          if (l < nl - 1 || outb)
            d.B.push_back(getd(scope + "biases_" + std::to_string(e) + "_" + std::to_string(l)));
          else
            d.B.emplace_back();
          if ((int) d.W[l].size() != d.sizes[l] * d.sizes[l + 1])
            throw std::string(scope + "weight shape");
        }
        out.push_back(d);
      }
    };
    int nin = m->nelt * m->k * m->nm;
    if (m->tdnp) {
      m->sommerfeld = geti("tdnp::Sommerfeld", 0, false) == 1;
      read_nets("H::", nin, m->tdH);
      m->ydim = m->tdH[0].sizes.back() + 1;  // plus the temperature column
      read_nets("U::", m->ydim, m->tdU);
      read_nets("S::", m->ydim, m->tdS);
      for (int e = 0; e < m->nelt; e++)
        if (m->tdU[e].sizes.back() != 1 || m->tdS[e].sizes.back() != 1)
          throw std::string("U and S output widths must be 1");
    } else {
      read_nets("", nin, m->mlp);
      for (const Dense &d : m->mlp)
        if (d.sizes.back() != 1) throw std::string("mlp output width must be 1");
    }
    if (!(interp_dx > 0)) throw std::string("interp dx must be > 0");
    build_table(*m, interp_dx);
  } catch (const std::string &e) {
    g_err = e;
    return -1;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
  m->ready = true;
  return 0;
}

int tmdo_info(void *h, int *out /* nelt, k, nm, nd, tab_n, gwidth */)
{
  auto *m = static_cast<Model *>(h);
  if (!m->ready) return -1;
  out[0] = m->nelt;
  out[1] = m->k;
  out[2] = m->nm;
  out[3] = m->nd;
  out[4] = m->tab_n;
  out[5] = m->nelt * m->k * m->nm;
  return 0;
}

int tmdo_get_T(void *h, double *T) /* [nm][nd] */
{
  auto *m = static_cast<Model *>(h);
  std::copy(m->T.begin(), m->T.end(), T);
  return 0;
}

int tmdo_get_table(void *h, double *tab) /* [tab_n][4][k] */
{
  auto *m = static_cast<Model *>(h);
  std::copy(m->tab.begin(), m->tab.end(), tab);
  return 0;
}

// H_k(r), dH_k/dr for an array of r
int tmdo_radial(void *h, int n, const double *r, double *H, double *dH)
{
  auto *m = static_cast<Model *>(h);
  for (int i = 0; i < n; i++) radial(*m, r[i], H + (size_t) i * m->k, dH + (size_t) i * m->k);
  return 0;
}

// Radial functions (FNN or analytic) times fc(r), evaluated directly (no table)
int tmdo_fnn(void *h, int n, const double *r, double *y)
{
  auto *m = static_cast<Model *>(h);
  DenseWork w;
  for (int i = 0; i < n; i++) {
    radial_values(*m, r[i], y + (size_t) i * m->k, w);
    double fc = cutoff_fc(*m, r[i]);
    for (int j = 0; j < m->k; j++) y[(size_t) i * m->k + j] *= fc;
  }
  return 0;
}

int tmdo_monomials(void *h, int n, const double *u, double *M, double *dMdu)
{
  auto *m = static_cast<Model *>(h);
  for (int i = 0; i < n; i++)
    monomials(*m, u + 3 * i, M + (size_t) i * m->nd, dMdu ? dMdu + (size_t) i * 3 * m->nd : nullptr);
  return 0;
}

// MLP forward + input gradient for element e on nrows rows of width gwidth.
int tmdo_mlp(void *h, int e, int nrows, const double *G, double *E, double *dEdG)
{
  auto *m = static_cast<Model *>(h);
  if (m->tdnp) {
    g_err = "a TDNP model has no single MLP";
    return -1;
  }
  int gw = m->nelt * m->k * m->nm;
  std::vector<int> bad(1, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nrows; i++) {
    DenseWork w;
    dense_forward(m->mlp[e], G + (size_t) i * gw, &E[i], w);
    if (dEdG) {
      double one = 1.0;
      dense_backward(m->mlp[e], w, &one, dEdG + (size_t) i * gw);
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// LAMMPS-shaped compute (PairTensorMD::compute, pair_tensormd.cpp:141-184):
// x[nall][3], elem[nall] (0-based element index after the type map),
// full neighbor list in CSR form (neigh_off[ii] .. neigh_off[ii+1]) per ilist
// entry.  f is OVERWRITTEN for all nall atoms (kernel.cu:1766); eatom[nlocal],
// vatom[nall][6] and the fdotr virial (pair.cpp:1735-1783) are optional.
int tmdo_compute_list(void *h, int nlocal, int nall, const double *x, const int *elem, int inum,
                      const int *ilist, const long *neigh_off, const int *neigh, double scale,
                      double *f, double *eatom, double *vatom, double *energy, double *virial,
                      double *G_out /*[inum][gw] or null*/, double *dEdG_out)
{
  auto *m = static_cast<Model *>(h);
  if (!m->ready) {
    g_err = "model not finalized";
    return -1;
  }
  int gw = m->nelt * m->k * m->nm;
  std::fill(f, f + 3 * (size_t) nall, 0.0);
  if (vatom) std::fill(vatom, vatom + 6 * (size_t) nall, 0.0);
  std::vector<AtomOut> outs(inum);
#pragma omp parallel for schedule(dynamic, 16)
  for (int ii = 0; ii < inum; ii++) {
    int i = ilist[ii];
    std::vector<Nbr> nb;
    for (long p = neigh_off[ii]; p < neigh_off[ii + 1]; p++) {
      int j = neigh[p] & 0x1FFFFFFF;  // NEIGHMASK
      Nbr q;
      q.d[0] = x[3 * j + 0] - x[3 * i + 0];
      q.d[1] = x[3 * j + 1] - x[3 * i + 1];
      q.d[2] = x[3 * j + 2] - x[3 * i + 2];
      q.j = j;
      q.ej = elem[j];
      nb.push_back(q);
    }
    atom_chain(*m, elem[i], nb, scale, outs[ii], G_out != nullptr);
    if (G_out) {
      std::copy(outs[ii].G.begin(), outs[ii].G.end(), G_out + (size_t) ii * gw);
      std::copy(outs[ii].dEdG.begin(), outs[ii].dEdG.end(), dEdG_out + (size_t) ii * gw);
    }
  }
  // serial, deterministic scatter in list order
  double etot = 0.0;
  if (m->tdnp) {
    m->last_s.assign(nlocal, 0.0);
    m->last_a.assign(nlocal, 0.0);
    m->last_S = m->last_A = 0.0;
  }
  for (int ii = 0; ii < inum; ii++) {
    int i = ilist[ii];
    etot += outs[ii].E * scale;
    if (eatom) eatom[i] = outs[ii].E * scale;
    if (m->tdnp) {
      m->last_s[i] = outs[ii].s * scale;
      m->last_a[i] = outs[ii].a * scale;
      m->last_S += m->last_s[i];
      m->last_A += m->last_a[i];
    }
    long c = 0;
    for (long p = neigh_off[ii]; p < neigh_off[ii + 1]; p++, c++) {
      int j = neigh[p] & 0x1FFFFFFF;
      const double *F = &outs[ii].F[3 * c];
      if (F[0] == 0.0 && F[1] == 0.0 && F[2] == 0.0) continue;
      double d[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
      for (int a = 0; a < 3; a++) {
        f[3 * i + a] += F[a];
        f[3 * j + a] -= F[a];
      }
      if (vatom) {
        double v[6] = {-d[0] * F[0], -d[1] * F[1], -d[2] * F[2],
                       -d[0] * F[1], -d[0] * F[2], -d[1] * F[2]};
        for (int q = 0; q < 6; q++) {
          vatom[6 * i + q] += 0.5 * v[q];
          vatom[6 * j + q] += 0.5 * v[q];
        }
      }
    }
  }
  *energy = etot;
  if (virial) {
    for (int q = 0; q < 6; q++) virial[q] = 0.0;
    for (int i = 0; i < nall; i++) {
      virial[0] += f[3 * i + 0] * x[3 * i + 0];
      virial[1] += f[3 * i + 1] * x[3 * i + 1];
      virial[2] += f[3 * i + 2] * x[3 * i + 2];
      virial[3] += f[3 * i + 1] * x[3 * i + 0];
      virial[4] += f[3 * i + 2] * x[3 * i + 0];
      virial[5] += f[3 * i + 2] * x[3 * i + 1];
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Periodic cell neighbor enumeration.  cell = rows a, b, c (cartesian, 9
// doubles).  Positions may be anywhere; the image shift s (integers) is the
// lattice translation added to x_j.  Displacement, in this exact order:
//   d_a = (x_j[a] - x_i[a]) + ((s0*h00 + s1*h10) + s2*h20)   (per component)
// and the pair is kept when d.d <= cut*cut (LAMMPS npair_full_bin.cpp:99).
// Brute force over all j and all images in range; used as the bit-exact
// reference for the device neighbor build.
static inline void disp(const double *x, const double *cell, int i, int j, const int s[3],
                        double d[3])
{
  for (int a = 0; a < 3; a++) {
    double sh = (s[0] * cell[0 * 3 + a] + s[1] * cell[1 * 3 + a]) + s[2] * cell[2 * 3 + a];
    d[a] = (x[3 * j + a] - x[3 * i + a]) + sh;
  }
}

static void image_range(const double *cell, double cut, int nimg[3])
{
  // perpendicular widths: V / |b x c| etc.
  const double *A = cell, *B = cell + 3, *C = cell + 6;
  auto cross = [](const double *u, const double *v, double *w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
  };
  double bc[3], ca[3], ab[3];
  cross(B, C, bc);
  cross(C, A, ca);
  cross(A, B, ab);
  double V = std::fabs(A[0] * bc[0] + A[1] * bc[1] + A[2] * bc[2]);
  double w[3] = {V / std::sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]),
                 V / std::sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]),
                 V / std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2])};
  // positions are assumed wrapped to [0,1) fractional; +1 covers that slack
  for (int a = 0; a < 3; a++) nimg[a] = (int) std::ceil(cut / w[a]) + 1;
}

// Returns total pairs; fills up to cap entries.  Order: i ascending, then j
// ascending, then image (s0, s1, s2) ascending.
long tmdo_neighbors_periodic(int n, const double *x, const double *cell, double cut, long cap,
                             long *off /*[n+1]*/, int *nj, int *nshift /*[cap][3]*/)
{
  int nimg[3];
  image_range(cell, cut, nimg);
  double cut2 = cut * cut;
  std::vector<std::vector<int>> lists(n);
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < n; i++) {
    auto &L = lists[i];
    for (int j = 0; j < n; j++)
      for (int s0 = -nimg[0]; s0 <= nimg[0]; s0++)
        for (int s1 = -nimg[1]; s1 <= nimg[1]; s1++)
          for (int s2 = -nimg[2]; s2 <= nimg[2]; s2++) {
            if (j == i && s0 == 0 && s1 == 0 && s2 == 0) continue;
            int s[3] = {s0, s1, s2};
            double d[3];
            disp(x, cell, i, j, s, d);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 <= cut2) {
              L.push_back(j);
              L.push_back(s0);
              L.push_back(s1);
              L.push_back(s2);
            }
          }
  }
  long tot = 0;
  for (int i = 0; i < n; i++) {
    off[i] = tot;
    long cnt = (long) lists[i].size() / 4;
    for (long p = 0; p < cnt; p++) {
      if (tot + p < cap) {
        nj[tot + p] = lists[i][4 * p];
        for (int a = 0; a < 3; a++) nshift[3 * (tot + p) + a] = lists[i][4 * p + 1 + a];
      }
    }
    tot += cnt;
  }
  off[n] = tot;
  return tot;
}

// Same set as tmdo_neighbors_periodic, via a cell list (for large cells).
// Bins are on fractional coordinates; positions must be wrapped to [0,1).
long tmdo_neighbors_binned(int n, const double *x, const double *cell, double cut, long cap,
                           long *off, int *nj, int *nshift)
{
  // inverse cell for fractional coordinates
  const double *h = cell;
  double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
      h[2] * (h[3] * h[7] - h[4] * h[6]);
  double inv[9];
  inv[0] = (h[4] * h[8] - h[5] * h[7]) / det;
  inv[1] = (h[2] * h[7] - h[1] * h[8]) / det;
  inv[2] = (h[1] * h[5] - h[2] * h[4]) / det;
  inv[3] = (h[5] * h[6] - h[3] * h[8]) / det;
  inv[4] = (h[0] * h[8] - h[2] * h[6]) / det;
  inv[5] = (h[2] * h[3] - h[0] * h[5]) / det;
  inv[6] = (h[3] * h[7] - h[4] * h[6]) / det;
  inv[7] = (h[1] * h[6] - h[0] * h[7]) / det;
  inv[8] = (h[0] * h[4] - h[1] * h[3]) / det;
  // x_cart = f0*A + f1*B + f2*C -> f = x * inv (row vector times inv)
  int nb[3];
  double w[3];
  // bins per axis: width >= cut/2 each -> nb = floor(w / (cut/2)) , >= 1
  {
    const double *A = cell, *B = cell + 3, *C = cell + 6;
    auto cross = [](const double *u, const double *v, double *o) {
      o[0] = u[1] * v[2] - u[2] * v[1];
      o[1] = u[2] * v[0] - u[0] * v[2];
      o[2] = u[0] * v[1] - u[1] * v[0];
    };
    double bc[3], ca[3], ab[3];
    cross(B, C, bc);
    cross(C, A, ca);
    cross(A, B, ab);
    double V = std::fabs(A[0] * bc[0] + A[1] * bc[1] + A[2] * bc[2]);
    w[0] = V / std::sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
    w[1] = V / std::sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]);
    w[2] = V / std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
  }
  for (int a = 0; a < 3; a++) nb[a] = std::max(1, (int) std::floor(w[a] / (0.5 * cut)));
  int reach[3];
  for (int a = 0; a < 3; a++) reach[a] = (int) std::ceil(cut / (w[a] / nb[a])) + 1;
  int nbins = nb[0] * nb[1] * nb[2];
  std::vector<int> binof(n), bcount(nbins + 1, 0);
  std::vector<int> bidx(3 * n);
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) {
      double fr = x[3 * i] * inv[0 * 3 + a] + x[3 * i + 1] * inv[1 * 3 + a] + x[3 * i + 2] * inv[2 * 3 + a];
      int bi = (int) std::floor(fr * nb[a]);
      bi = std::min(std::max(bi, 0), nb[a] - 1);
      bidx[3 * i + a] = bi;
    }
    binof[i] = (bidx[3 * i] * nb[1] + bidx[3 * i + 1]) * nb[2] + bidx[3 * i + 2];
    bcount[binof[i] + 1]++;
  }
  for (int b = 0; b < nbins; b++) bcount[b + 1] += bcount[b];
  std::vector<int> sorted(n), fill(bcount.begin(), bcount.end() - 1);
  for (int i = 0; i < n; i++) sorted[fill[binof[i]]++] = i;
  double cut2 = cut * cut;
  std::vector<std::vector<int>> lists(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    std::vector<std::array<int, 4>> L;
    for (int o0 = -reach[0]; o0 <= reach[0]; o0++)
      for (int o1 = -reach[1]; o1 <= reach[1]; o1++)
        for (int o2 = -reach[2]; o2 <= reach[2]; o2++) {
          int t[3] = {bidx[3 * i] + o0, bidx[3 * i + 1] + o1, bidx[3 * i + 2] + o2};
          int s[3];
          for (int a = 0; a < 3; a++) {
            int q = t[a];
            int sh = (q >= 0) ? q / nb[a] : -((-q + nb[a] - 1) / nb[a]);
            s[a] = sh;
            t[a] = q - sh * nb[a];
          }
          int b = (t[0] * nb[1] + t[1]) * nb[2] + t[2];
          for (int p = bcount[b]; p < bcount[b + 1]; p++) {
            int j = sorted[p];
            if (j == i && s[0] == 0 && s[1] == 0 && s[2] == 0) continue;
            double d[3];
            disp(x, cell, i, j, s, d);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 <= cut2) L.push_back({j, s[0], s[1], s[2]});
          }
        }
    std::sort(L.begin(), L.end());
    auto &o = lists[i];
    for (auto &e : L)
      for (int q = 0; q < 4; q++) o.push_back(e[q]);
  }
  long tot = 0;
  for (int i = 0; i < n; i++) {
    off[i] = tot;
    long cnt = (long) lists[i].size() / 4;
    for (long p = 0; p < cnt; p++)
      if (tot + p < cap) {
        nj[tot + p] = lists[i][4 * p];
        for (int a = 0; a < 3; a++) nshift[3 * (tot + p) + a] = lists[i][4 * p + 1 + a];
      }
    tot += cnt;
  }
  off[n] = tot;
  return tot;
}

// Energy/forces/virial of a periodic structure from a neighbor list with
// image shifts (as produced above).  virial = sum over pairs of -d (x) F
// (equal to the fdotr virial over LAMMPS ghosts).
int tmdo_compute_periodic(void *h, int n, const double *x, const int *elem, const double *cell,
                          const long *off, const int *nj, const int *nshift, double scale,
                          double *f, double *energy, double *eatom, double *vatom, double *virial,
                          double *G_out, double *dEdG_out)
{
  auto *m = static_cast<Model *>(h);
  if (!m->ready) {
    g_err = "model not finalized";
    return -1;
  }
  int gw = m->nelt * m->k * m->nm;
  std::vector<AtomOut> outs(n);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; i++) {
    std::vector<Nbr> nb;
    for (long p = off[i]; p < off[i + 1]; p++) {
      Nbr q;
      disp(x, cell, i, nj[p], &nshift[3 * p], q.d);
      q.j = nj[p];
      q.ej = elem[nj[p]];
      nb.push_back(q);
    }
    atom_chain(*m, elem[i], nb, scale, outs[i], G_out != nullptr);
    if (G_out) {
      std::copy(outs[i].G.begin(), outs[i].G.end(), G_out + (size_t) i * gw);
      std::copy(outs[i].dEdG.begin(), outs[i].dEdG.end(), dEdG_out + (size_t) i * gw);
    }
  }
  std::fill(f, f + 3 * (size_t) n, 0.0);
  if (vatom) std::fill(vatom, vatom + 6 * (size_t) n, 0.0);
  double vir[6] = {0, 0, 0, 0, 0, 0};
  double etot = 0.0;
  if (m->tdnp) {
    m->last_s.assign(n, 0.0);
    m->last_a.assign(n, 0.0);
    m->last_S = m->last_A = 0.0;
  }
  for (int i = 0; i < n; i++) {
    etot += outs[i].E * scale;
    if (eatom) eatom[i] = outs[i].E * scale;
    if (m->tdnp) {
      m->last_s[i] = outs[i].s * scale;
      m->last_a[i] = outs[i].a * scale;
      m->last_S += m->last_s[i];
      m->last_A += m->last_a[i];
    }
    long c = 0;
    for (long p = off[i]; p < off[i + 1]; p++, c++) {
      const double *F = &outs[i].F[3 * c];
      int j = nj[p];
      double d[3];
      disp(x, cell, i, j, &nshift[3 * p], d);
      for (int a = 0; a < 3; a++) {
        f[3 * i + a] += F[a];
        f[3 * j + a] -= F[a];
      }
      double v[6] = {-d[0] * F[0], -d[1] * F[1], -d[2] * F[2],
                     -d[0] * F[1], -d[0] * F[2], -d[1] * F[2]};
      for (int q = 0; q < 6; q++) vir[q] += v[q];
      if (vatom)
        for (int q = 0; q < 6; q++) {
          vatom[6 * i + q] += 0.5 * v[q];
          vatom[6 * j + q] += 0.5 * v[q];
        }
    }
  }
  *energy = etot;
  if (virial)
    for (int q = 0; q < 6; q++) virial[q] = vir[q];
  return 0;
}

// Velocity-Verlet NVE (FixNVE::initial_integrate/final_integrate,
// src/fix_nve.cpp:64-130) in metal units, with a neighbor list rebuilt
// every step (so the trajectory depends only on the force field).
// Per step s = 1..nsteps the potential energy and (optionally) forces after
// the step are recorded.  mass_by_elem[nelt].
int tmdo_nve(void *h, int n, double *x, double *v, const int *elem, const double *cell,
             const double *mass_by_elem, double dt, int nsteps, double *epot /*[nsteps+1]*/,
             double *ekin /*[nsteps+1]*/, double *f_hist /*[nsteps+1][n][3] or null*/)
{
  auto *m = static_cast<Model *>(h);
  const double ftm2v = 1.0 / 1.0364269e-4;
  const double mvv2e = 1.0364269e-4;
  std::vector<double> f(3 * (size_t) n);
  auto force = [&](double *e) -> int {
    long cap = (long) n * 400;
    std::vector<long> off(n + 1);
    std::vector<int> nj(cap), ns(3 * cap);
    long tot = tmdo_neighbors_binned(n, x, cell, m->rc, cap, off.data(), nj.data(), ns.data());
    if (tot > cap) {
      g_err = "neighbor capacity";
      return -1;
    }
    return tmdo_compute_periodic(h, n, x, elem, cell, off.data(), nj.data(), ns.data(), 1.0,
                                 f.data(), e, nullptr, nullptr, nullptr, nullptr, nullptr);
  };
  auto kin = [&]() {
    double s = 0;
    for (int i = 0; i < n; i++)
      for (int a = 0; a < 3; a++) s += mass_by_elem[elem[i]] * v[3 * i + a] * v[3 * i + a];
    return 0.5 * mvv2e * s;
  };
  // wrap helper: keep fractional coordinates in [0,1)
  const double *hc = cell;
  double det = hc[0] * (hc[4] * hc[8] - hc[5] * hc[7]) - hc[1] * (hc[3] * hc[8] - hc[5] * hc[6]) +
      hc[2] * (hc[3] * hc[7] - hc[4] * hc[6]);
  double inv[9];
  inv[0] = (hc[4] * hc[8] - hc[5] * hc[7]) / det;
  inv[1] = (hc[2] * hc[7] - hc[1] * hc[8]) / det;
  inv[2] = (hc[1] * hc[5] - hc[2] * hc[4]) / det;
  inv[3] = (hc[5] * hc[6] - hc[3] * hc[8]) / det;
  inv[4] = (hc[0] * hc[8] - hc[2] * hc[6]) / det;
  inv[5] = (hc[2] * hc[3] - hc[0] * hc[5]) / det;
  inv[6] = (hc[3] * hc[7] - hc[4] * hc[6]) / det;
  inv[7] = (hc[1] * hc[6] - hc[0] * hc[7]) / det;
  inv[8] = (hc[0] * hc[4] - hc[1] * hc[3]) / det;
  auto wrap = [&]() {
    for (int i = 0; i < n; i++) {
      double fr[3];
      for (int a = 0; a < 3; a++)
        fr[a] = x[3 * i] * inv[0 * 3 + a] + x[3 * i + 1] * inv[1 * 3 + a] + x[3 * i + 2] * inv[2 * 3 + a];
      int s[3];
      bool any = false;
      for (int a = 0; a < 3; a++) {
        s[a] = (int) std::floor(fr[a]);
        if (s[a] != 0) any = true;
      }
      if (any)
        for (int a = 0; a < 3; a++)
          x[3 * i + a] -= (s[0] * hc[0 * 3 + a] + s[1] * hc[1 * 3 + a]) + s[2] * hc[2 * 3 + a];
    }
  };
  wrap();
  double e;
  if (force(&e)) return -1;
  epot[0] = e;
  ekin[0] = kin();
  if (f_hist) std::copy(f.begin(), f.end(), f_hist);
  double dtf = 0.5 * dt * ftm2v;
  for (int s = 1; s <= nsteps; s++) {
    for (int i = 0; i < n; i++) {
      double dtfm = dtf / mass_by_elem[elem[i]];
      for (int a = 0; a < 3; a++) {
        v[3 * i + a] += dtfm * f[3 * i + a];
        x[3 * i + a] += dt * v[3 * i + a];
      }
    }
    wrap();
    if (force(&e)) return -1;
    for (int i = 0; i < n; i++) {
      double dtfm = dtf / mass_by_elem[elem[i]];
      for (int a = 0; a < 3; a++) v[3 * i + a] += dtfm * f[3 * i + a];
    }
    epot[s] = e;
    ekin[s] = kin();
    if (f_hist) std::copy(f.begin(), f.end(), f_hist + (size_t) s * 3 * n);
  }
  return 0;
}

int tmdo_num_threads()
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// threads of the following parallel regions (benchmark baselines: 1 thread
// and all cores)
void tmdo_set_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void) n;
#endif
}

}  // extern "C"
