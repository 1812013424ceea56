// Model loading and the host-side precomputation.
//
// Schema: TensorMD::read_npz (src/ML-TENSORMD/tensormd.cpp:108-177),
// NeuralNetworkPotential::setup_global (src/ML-TENSORMD/nnp.cpp:110-205),
// FeaturePotential::setup_global (src/ML-TENSORMD/fnn.cpp:32-72).
// Table: FeaturePotential::setup_ration (fnn.cpp:77-105) with Cutoff
// (src/ML-TENSORMD/cutoff.cpp:39-53) and BatchRation1d::setup
// (src/ML-TENSORMD/ration.hpp:166-225).
#include "tmd_model.h"

#include <cmath>

#include "npz.h"

namespace tmd {

double softplus_threshold()
{
  return std::log(2.22045e-16) + 2.0;  // nnp.cu:30
}

void table_grid(Model &m, double delta, std::vector<double> &xs)
{
  xs.clear();
  int n;
  if (m.fp32_model) {
    // the reference computes the grid in float for precision-32 models
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
  m.tab_n = n;
}

namespace {

double act_value(int act, double z)
{
  switch (act) {
    case ACT_SOFTPLUS: {
      const double thr = softplus_threshold();
      return (z < -thr) ? std::log1p(std::exp(z)) : z;
    }
    case ACT_SQUAREPLUS: return 0.5 * (z + std::sqrt(z * z + 4.0));
    case ACT_TANH: return std::tanh(z);
    default: return z >= 0 ? z : 0.0;
  }
}

// The radial network forward in FP64 (the reference evaluates it once on the
// GPU in forward_fnn_GPU, gpu/cuda/nnp.cu:59-116).
void fnn_forward(const DenseNet &net, double r, double *y)
{
  std::vector<double> h(1, r), z;
  for (int l = 0; l < net.nlayers; l++) {
    int K = net.sizes[l], N = net.sizes[l + 1];
    z.assign(N, 0.0);
    for (int i = 0; i < K; i++)
      for (int j = 0; j < N; j++) z[j] += h[i] * net.W[l][(size_t) i * N + j];
    if (!net.B[l].empty())
      for (int j = 0; j < N; j++) z[j] += net.B[l][j];
    if (l == net.nlayers - 1) {
      for (int j = 0; j < N; j++) y[j] = z[j];
      return;
    }
    bool res = net.resnet && l > 0 && K == N;
    std::vector<double> hn(N);
    for (int j = 0; j < N; j++) hn[j] = act_value(net.act, z[j]) + (res ? h[j] : 0.0);
    h.swap(hn);
  }
}

// The analytic radial functions f_k(r) (Descriptor::pexp, morse, density,
// sf; src/ML-TENSORMD/descriptor.cpp:51-132).
void analytic_forward(const Model &m, double r, double *y)
{
  const std::vector<std::vector<double>> &p = m.rparams;
  for (int k = 0; k < m.k; k++) {
    switch (m.radial_fn) {
      case RADIAL_PEXP:  // exp(-(r / rl)^pl)
        y[k] = std::exp(-std::pow(r / p[0][k], p[1][k]));
        break;
      case RADIAL_MORSE: {  // D [exp(-2 gamma (r - r0)) - 2 exp(-gamma (r - r0))]
        const double g = p[1][k] * (r - p[2][k]);
        y[k] = p[0][k] * (std::exp(-2.0 * g) - 2.0 * std::exp(-g));
        break;
      }
      case RADIAL_DENSITY:  // A exp(-beta (r - re) / re)
        y[k] = p[0][k] * std::exp(-p[1][k] * ((r - p[2][k]) / p[2][k]));
        break;
      default: {  // exp(-eta (r - omega)^2 / rc^2)
        const double d = r - p[1][k];
        y[k] = std::exp(-p[0][k] * d * d / (m.rc * m.rc));
      }
    }
  }
}

double cutoff_fc(const Model &m, double r)
{
  if (m.fctype == 0) {
    double Z = std::min(r / m.rc, 1.0) * M_PI;
    return 0.5 + 0.5 * std::cos(Z);
  }
  const double gamma = 5.0;
  double Z = std::min(r / m.rc, M_PI);
  double Z4 = Z * Z * Z * Z, Z5 = Z4 * Z;
  return 1.0 + gamma * Z5 * Z - (gamma + 1.0) * Z5;
}

void build_T(Model &m)
{
  // multiplicities m!/(a!b!c!) in the monomial order of kernel.cu:1168-1406;
  // equals the literal table of tensormd.cpp:205-276
  std::vector<int> mono;
  for (int n = 0; n <= m.max_moment; n++)
    for (int a = n; a >= 0; a--)
      for (int b = n - a; b >= 0; b--) {
        mono.push_back(a);
        mono.push_back(b);
        mono.push_back(n - a - b);
      }
  m.nd = (int) mono.size() / 3;
  m.nm = m.max_moment + 1;
  auto fact = [](int n) {
    double f = 1;
    for (int i = 2; i <= n; i++) f *= i;
    return f;
  };
  m.T.assign((size_t) m.nm * m.nd, 0.0);
  for (int d = 0; d < m.nd; d++) {
    int a = mono[3 * d], b = mono[3 * d + 1], c = mono[3 * d + 2];
    m.T[(a + b + c) * m.nd + d] = fact(a + b + c) / (fact(a) * fact(b) * fact(c));
  }
  if (m.T_sym && m.max_moment <= 3) {
    if (m.max_moment > 1) m.T[2 * m.nd + 0] = -1.0 / 3.0;
    if (m.max_moment > 2)
      for (int d = 1; d <= 3; d++) m.T[3 * m.nd + d] = -3.0 / 5.0;
  }
}

// A max_moment < 3 model as the equivalent max_moment-3 model: the
// descriptor channels G[b][k][m] with m > max_moment get zero first-layer
// weights, so they add nothing to any activation and receive dE/dG = 0 (no
// force either); the input index (b k_num + k)(max_moment + 1) + m of the
// file (tensormd.cpp:193-196) becomes (b k_num + k) 4 + m.  T is then the
// max_moment-3 table (its extra rows only meet the zero weights).
void pad_moments(Model &m)
{
  const int nm0 = m.max_moment + 1, rows = m.nelt * m.k;
  for (DenseNet &d : m.tdnp ? m.tdH : m.mlp) {
    const int N = d.sizes[1];
    std::vector<double> W((size_t) rows * 4 * N, 0.0);
    for (int r = 0; r < rows; r++)
      for (int mo = 0; mo < nm0; mo++)
        for (int j = 0; j < N; j++)
          W[((size_t) r * 4 + mo) * N + j] = d.W[0][((size_t) r * nm0 + mo) * N + j];
    d.W[0] = std::move(W);
    d.sizes[0] = rows * 4;
  }
  m.max_moment = 3;
  build_T(m);
}

void build_table(Model &m, double delta)
{
  std::vector<double> xs;
  table_grid(m, delta, xs);
  const int n = m.tab_n, k = m.k;
  m.tab.assign((size_t) n * k * 4, 0.0);
  auto at = [&](int i, int j, int q) -> double & { return m.tab[((size_t) i * k + j) * 4 + q]; };
  std::vector<double> y(k);
  for (int i = 0; i < n; i++) {
    if (m.use_fnn)
      fnn_forward(m.fnn, xs[i], y.data());
    else
      analytic_forward(m, xs[i], y.data());
    double fc = cutoff_fc(m, xs[i]);
    for (int j = 0; j < k; j++) at(i, j, 0) = y[j] * fc;
  }
  const double xdx = m.xdx;
  for (int i = 0; i < n - 1; i++)
    for (int j = 0; j < k; j++) at(i, j, 1) = (at(i + 1, j, 0) - at(i, j, 0)) * xdx;
  for (int i = 0; i < n - 1; i++)
    for (int j = 0; j < k; j++) {
      double s = at(i, j, 1), c1, c2;
      if (i == 0) {
        double sp = at(i + 1, j, 1);
        c2 = (s * (s - (sp - s) / 2) <= 0) ? s * xdx : (sp - s) * xdx / 2;
        c1 = 0;
      } else if (i == n - 2) {
        c1 = (s - at(i - 1, j, 1)) * xdx / 2;
        c2 = 0;
      } else {
        double sm = at(i - 1, j, 1), sp = at(i + 1, j, 1);
        c1 = (i == 1 && sm * (sm - (s - sm) / 2) <= 0) ? (s - 2 * sm) * xdx : (s - sm) * xdx / 2;
        c2 = (sp - s) * xdx / 2;
      }
      at(i, j, 2) = c1;
      at(i, j, 3) = c2;
    }
}

}  // namespace

bool load_model(const std::string &path, double dx, Model &m, std::string &err, bool host_table)
{
  Npz z;
  if (!read_npz(path, z, err)) return false;
  auto has = [&](const std::string &k) { return z.find(k) != z.end(); };
  auto geti = [&](const std::string &k, int def, bool req) -> int {
    if (!has(k) || z[k].size() == 0) {
      if (req) throw std::string("model file lacks key '" + k + "'");
      return def;
    }
    return (int) z[k].get_int(0);
  };
  auto getd = [&](const std::string &k) -> std::vector<double> {
    if (!has(k)) throw std::string("model file lacks key '" + k + "'");
    return z[k].as_doubles();
  };
  try {
    m.nelt = geti("nelt", 0, true);
    if (m.nelt < 1 || m.nelt > kMaxModelElts) throw std::string("nelt must be in 1..8");
    m.rc = getd("rmax").at(0);
    if (!(m.rc > 0)) throw std::string("rmax must be > 0");
    m.max_moment = geti("max_moment", 0, true);
    if (m.max_moment < 0 || m.max_moment > 5) throw std::string("max_moment must be in 0..5");
    m.max_moment_file = m.max_moment;
    m.fctype = geti("fctype", 0, true);
    m.T_sym = geti("is_T_symmetric", 0, false) == 1;
    m.fp32_model = geti("precision", 64, false) == 32;
    m.tdnp = geti("tdnp", 0, false) == 1;
    m.use_fnn = geti("use_fnn", 0, true) == 1;
    m.masses = getd("masses");
    if ((int) m.masses.size() != m.nelt) throw std::string("masses size != nelt");
    std::vector<int> nums = z.at("numbers").as_ints();
    m.symbols.clear();
    for (int e = 0; e < m.nelt; e++) {
      // pair_tensormd.cpp:353-361: two ASCII codes per element, 0-padded
      std::string s;
      if (2 * e < (int) nums.size()) s.push_back((char) nums[2 * e]);
      if (2 * e + 1 < (int) nums.size() && nums[2 * e + 1] > 0) s.push_back((char) nums[2 * e + 1]);
      m.symbols.push_back(s);
    }
    build_T(m);
    for (int i = 0; i < m.nelt; i++) {
      int kk = 1;
      for (int j = 0; j < m.nelt; j++) m.eltij[i][j] = (i == j) ? 0 : kk++;  // tensormd.cpp:165-176
    }
    if (m.use_fnn) {
      // radial network
      DenseNet &f = m.fnn;
      f.nlayers = geti("fnn::nlayers", 0, true);
      if (f.nlayers < 2) throw std::string("fnn::nlayers must be >= 2");  // fnn.cpp:64-67
      std::vector<int> fs = z.at("fnn::layer_sizes").as_ints();
      if ((int) fs.size() != f.nlayers) throw std::string("fnn::layer_sizes length");
      f.sizes.assign(1, 1);
      for (int v : fs) f.sizes.push_back(v);
      f.act = geti("fnn::actfn", 1, true);
      if (f.act != ACT_SOFTPLUS && f.act != ACT_SQUAREPLUS)
        throw std::string("fnn::actfn must be softplus(1) or squareplus(3), the activations "
                          "the reference radial network applies (nnp.cu:89-95)");
      f.outbias = geti("fnn::apply_output_bias", 0, true) == 1;
      f.resnet = geti("fnn::use_resnet_dt", 0, true) == 1;
      for (int l = 0; l < f.nlayers; l++) {
        f.W.push_back(getd("fnn::weights_0_" + std::to_string(l)));
        if ((int) f.W[l].size() != f.sizes[l] * f.sizes[l + 1]) throw std::string("fnn weight shape");
        if (l < f.nlayers - 1 || f.outbias)
          f.B.push_back(getd("fnn::biases_0_" + std::to_string(l)));
        else
          f.B.emplace_back();
      }
      m.k = geti("fnn::num_filters", 0, true);
      if (f.sizes.back() != m.k) throw std::string("fnn output width != fnn::num_filters");
    } else {
      // analytic radial functions: descriptor::method and one parameter
      // array per parameter of the function, of length num_filters
      // (Descriptor::Descriptor, src/ML-TENSORMD/descriptor.h:37-87)
      m.radial_fn = geti("descriptor::method", RADIAL_PEXP, false);
      static const char *const keys[4][3] = {{"rl", "pl", nullptr},
                                             {"D", "gamma", "r0"},
                                             {"A", "beta", "re"},
                                             {"eta", "omega", nullptr}};
      if (m.radial_fn < 0 || m.radial_fn > 3) throw std::string("descriptor::method must be 0..3");
      m.rparams.clear();
      for (int q = 0; q < 3 && keys[m.radial_fn][q]; q++)
        m.rparams.push_back(getd(std::string("descriptor::") + keys[m.radial_fn][q]));
      m.k = (int) m.rparams[0].size();
      for (const auto &v : m.rparams)
        if ((int) v.size() != m.k) throw std::string("descriptor parameter arrays differ in length");
      for (int k = 0; k < m.k; k++) {
        const bool bad = (m.radial_fn == RADIAL_PEXP && !(m.rparams[0][k] != 0.0)) ||
                         (m.radial_fn == RADIAL_DENSITY && !(m.rparams[2][k] != 0.0));
        if (bad) throw std::string("descriptor parameter divides by zero");
      }
    }
    if (!filters_supported(m.k))
      throw std::string("num_filters " + std::to_string(m.k) +
                        " is not supported: the descriptor kernels take 16, 24, 32, 64 or 128");
    // per-element networks of one scope: "" for the main MLP
    // (NeuralNetworkPotential::setup_global, nnp.cpp:110-205); "H::", "U::",
    // "S::" for TDNP (TDNP::setup_one, tdnp.cpp:60-148)
    auto read_nets = [&](const std::string &scope, int nin, std::vector<DenseNet> &out) {
      int nl = geti(scope + "nlayers", 0, true);
      std::vector<int> ls = z.at(scope + "layer_sizes").as_ints();
      if ((int) ls.size() != nl || nl < 1) throw std::string(scope + "layer_sizes length");
      int act = geti(scope + "actfn", 0, true);
      if (act < 0 || act > 3) throw std::string(scope + "actfn must be 0..3");
      bool outb = geti(scope + "apply_output_bias", 0, true) == 1;
      bool res = geti(scope + "use_resnet_dt", 0, true) == 1;
      int nres = 0;
      for (int l = 1; l < nl; l++)
        if (ls[l - 1] == ls[l]) nres++;
      if (nres == 0) res = false;  // nnp.cpp:179-181
      for (int e = 0; e < m.nelt; e++) {
        DenseNet d;
        d.nlayers = nl;
        d.sizes.assign(1, nin);
        for (int v : ls) d.sizes.push_back(v);
        d.act = act;
        d.outbias = outb;
        d.resnet = res;
        for (int l = 0; l < nl; l++) {
          const std::string wk = scope + "weights_" + std::to_string(e) + "_" + std::to_string(l);
          d.W.push_back(getd(wk));
          if ((int) d.W[l].size() != d.sizes[l] * d.sizes[l + 1]) throw std::string(wk + " shape");
          if (l < nl - 1 || outb) {
            d.B.push_back(getd(scope + "biases_" + std::to_string(e) + "_" + std::to_string(l)));
            if ((int) d.B[l].size() < d.sizes[l + 1]) throw std::string(scope + "bias shape");
          } else {
            d.B.emplace_back();
          }
        }
        out.push_back(std::move(d));
      }
    };
    if (m.tdnp) {
      m.sommerfeld = geti("tdnp::Sommerfeld", 0, false) == 1;
      read_nets("H::", m.gwidth(), m.tdH);
      m.ydim = m.tdH[0].sizes.back() + 1;  // plus the electron-temperature column
      read_nets("U::", m.ydim, m.tdU);
      read_nets("S::", m.ydim, m.tdS);
      for (int e = 0; e < m.nelt; e++)
        if (m.tdU[e].sizes.back() != 1 || m.tdS[e].sizes.back() != 1)
          throw std::string("TDNP U and S networks must end in width 1");
    } else {
      read_nets("", m.gwidth(), m.mlp);
      if (m.mlp[0].nlayers < 2) throw std::string("layer_sizes length");
      if (m.mlp[0].sizes.back() != 1) throw std::string("the last layer must have width 1");
    }
    if (m.max_moment < 3) pad_moments(m);
    if (!(dx > 0)) throw std::string("interp dx must be > 0");
    if (m.rc / dx > 4.0e6) throw std::string("interp dx too small");
    if (host_table) {
      build_table(m, dx);
    } else {
      std::vector<double> xs;
      table_grid(m, dx, xs);
    }
  } catch (const std::string &e) {
    err = e;
    return false;
  } catch (const std::exception &e) {
    err = std::string("model file: ") + e.what();
    return false;
  }
  return true;
}

}  // namespace tmd
