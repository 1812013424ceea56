// Host-side TensorMD model: parameters read from the reference .npz schema,
// the multiplicity matrix T, the element-slot map and the radial
// interpolation table.  Everything here is built once per model on the host
// in FP64 and uploaded by the device layer.
#pragma once
#include <string>
#include <vector>

namespace tmd {

enum Act { ACT_RELU = 0, ACT_SOFTPLUS = 1, ACT_TANH = 2, ACT_SQUAREPLUS = 3 };

// Radial filter counts the descriptor and force kernels are instantiated
// for.  Every element count up to kMaxModelElts runs with each of them:
// models whose slots do not fit one register pass (k = 64 with more than 4
// elements, k = 128 with more than 2) accumulate the slots in several passes.
constexpr bool filters_supported(int k) { return k == 16 || k == 24 || k == 32 || k == 64 || k == 128; }
constexpr int kMaxModelElts = 8;

struct DenseNet {
  int nlayers = 0;
  std::vector<int> sizes;  // sizes[0] = input width, sizes[l + 1] = output of layer l
  int act = 0;
  bool resnet = false;
  bool outbias = false;
  std::vector<std::vector<double>> W;  // [l] row-major [in][out]
  std::vector<std::vector<double>> B;  // [l] [out]; empty when absent
  bool res_layer(int l) const { return resnet && l > 0 && l < nlayers - 1 && sizes[l] == sizes[l + 1]; }
};

// Analytic radial functions of a use_fnn = 0 model (Descriptor_fn,
// src/ML-TENSORMD/descriptor.h:26), one parameter set per filter.
enum RadialFn { RADIAL_PEXP = 0, RADIAL_MORSE = 1, RADIAL_DENSITY = 2, RADIAL_SF = 3 };

struct Model {
  int nelt = 0;
  double rc = 0;
  int max_moment = 0;       // as evaluated (3: smaller ones are padded)
  int max_moment_file = 0;  // as in the model file
  int nm = 0;  // max_moment + 1
  int nd = 0;  // monomials
  int k = 0;   // num_filters
  int fctype = 0;
  bool T_sym = false;
  bool fp32_model = false;
  std::vector<double> masses;
  std::vector<std::string> symbols;
  std::vector<double> T;  // [nm][nd]
  int eltij[8][8] = {};
  bool use_fnn = true;       // radial functions from the FNN, else analytic
  DenseNet fnn;
  int radial_fn = RADIAL_PEXP;             // analytic kind (use_fnn false)
  std::vector<std::vector<double>> rparams;  // [parameter][filter]
  std::vector<DenseNet> mlp;  // per element
  // TDNP models (tdnp = 1): per element H (descriptor -> ydim - 1 outputs,
  // linear last layer), U and S (ydim -> 1) instead of mlp; the last of the
  // ydim inputs of U and S is the electron temperature
  bool tdnp = false;
  bool sommerfeld = false;
  int ydim = 0;
  std::vector<DenseNet> tdH, tdU, tdS;
  // Kerley table, [n][k][4] = {y, s, c1, c2}
  int tab_n = 0;
  double dx = 0, xdx = 0;
  std::vector<double> tab;
  int gwidth() const { return nelt * k * nm; }
  // descriptor width of the model file (max_moment_file + 1 moments)
  int gwidth_file() const { return nelt * k * (max_moment_file + 1); }
};

// Reads `path`, validates the schema and builds T, slots and (host_table)
// the table for interpolation step `dx`; without host_table only the grid
// (tab_n, dx, xdx) is set and the table is left to build_table_device.
// Returns false with a message on failure.
bool load_model(const std::string &path, double dx, Model &m, std::string &err, bool host_table = true);

// The table's grid points r_i (FeaturePotential::setup_ration, fnn.cpp:81-86):
// sets m.tab_n, m.dx and m.xdx.
void table_grid(Model &m, double delta, std::vector<double> &xs);

// Below -threshold softplus(z) = log1p(exp(z)), else z (nnp.cu:30).
double softplus_threshold();

// The table [n][k][4] = {y, s, c1, c2} for step `delta`, built on the
// current device in FP64 (the radial network or analytic functions, the
// cutoff and the Kerley coefficients: the work forward_fnn_GPU,
// gpu/cuda/nnp.cu:59-116, and BatchRation1d::setup, ration.hpp:166-225, do
// at model load) and copied to m.tab.
bool build_table_device(Model &m, double delta, std::string &err);

}  // namespace tmd
