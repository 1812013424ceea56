// Internal interface between the host layer (tmd_api.cpp) and the gfx950
// kernels (tmd_device.hip).  All pointers are device pointers unless named
// h_*; every launcher enqueues on `stream` and returns hipError_t.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace tmd {

// Neighbor entries are one uint32: j in the low 23 bits, the image-shift code
// (s0+3)*49 + (s1+3)*7 + (s2+3), s in [-3, 3], in the top 9 bits.
constexpr int kNbrJBits = 23;
constexpr uint32_t kNbrJMask = (1u << kNbrJBits) - 1u;
constexpr int kMaxAtoms = (1 << kNbrJBits) - 1;
constexpr uint32_t kShiftCenter = 3u * 49u + 3u * 7u + 3u;
constexpr int kMaxElts = 8;
constexpr int kND = 20;  // monomials for max_moment 3
constexpr int kNM = 4;

struct Cell {
  double h[9];    // rows a, b, c
  double inv[9];  // inverse (fractional = x * inv)
};

// Per-model constants passed by value to the descriptor/force kernels.
struct ModelGeo {
  int nelt;
  int k;
  int tab_n;
  int eltij[kMaxElts][kMaxElts];
  float T[kNM][kND];
  double rc2;
  double dx, xdx;
  float dxf;
};

// Row layout of the MLP: rows are atoms grouped by element; each element
// group starts at a multiple of kRowTile.  row2atom[r] = -1 for padding.
constexpr int kRowTile = 64;
struct RowLayout {
  int nrows;                     // padded total
  int row_off[kMaxElts + 1];     // element e occupies [row_off[e], row_off[e+1])
};

// The per-element MLP, all elements packed: layer l of element e starts at
// W*[l] + e * wstride[l].  Forward B operands: Wf32 = W [K][N] (fp32),
// Wtb = W^T [N][K] (bf16); backward: Wt32 = W^T [N][K] (fp32), Wb16 = W
// [K][N] (bf16).  The LDS fields are filled by the launcher.
constexpr int kMaxLayers = 8;
struct MlpParams {
  int nl;                        // layers including the width-1 output layer
  int sizes[kMaxLayers + 1];     // sizes[0] = descriptor width
  int act;
  int res[kMaxLayers];           // resnet skip of hidden layer l
  const float *Wf32[kMaxLayers];
  const float *Wt32[kMaxLayers];
  const uint16_t *Wtb[kMaxLayers];
  const uint16_t *Wb16[kMaxLayers];
  const float *b[kMaxLayers];    // hidden biases (or null)
  const float *wout;             // output layer weights [K_last]
  const float *bout;             // output bias (or null)
  long wstride[kMaxLayers];
  int bstride[kMaxLayers];
  // LDS layout (elements of the mode's type)
  int offH[kMaxLayers], offD[kMaxLayers], ld[kMaxLayers + 1];
  int offGH, ldGH;
  int storeD;  // D_l kept in LDS (otherwise recomputed from H_l: tanh)
};

// ---- neighbor build ---------------------------------------------------------
struct NbrBuild {
  int n;
  int cap;            // entries per atom
  double cut2;        // (rc + skin)^2
  double trig2;       // (skin/2)^2
  int nb[3];          // bins per axis
  int reach[3];       // stencil half-width per axis
  int *bin_count;     // [nbins + 1]
  int *bin_start;     // [nbins + 1]
  int *bin_cursor;    // [nbins]
  int *bin_atoms;     // [n]
  int *atom_bin;      // [n] flat bin
  int *atom_bc;       // [n][3] bin coordinates
  double *x_build;    // [n][4]
  unsigned long long *ctl;  // [0] max d^2 bits, [1] force flag, [2] overflow, [3] build count
};

hipError_t nbr_step(const NbrBuild &nb, const Cell &cell, double4 *pos, uint32_t *nbr, int *nn,
                    hipStream_t stream);

// ---- descriptor / force -------------------------------------------------------
hipError_t launch_descriptor(const ModelGeo &g, const Cell &cell, const double4 *pos,
                             const uint32_t *nbr, const int *nn, int cap, const int *row2atom,
                             int nrows, const float4 *tab, float *G, hipStream_t stream);

// Force/virial: dEdG rows -> f (atomics, FP64, must be zeroed), per-row
// virial W[row][6] (sum over the atom's pairs of -d (x) F), optional vatom
// (FP64 atomics; must be zeroed).
hipError_t launch_force(const ModelGeo &g, const Cell &cell, const double4 *pos,
                        const uint32_t *nbr, const int *nn, int cap, const int *row2atom,
                        int nrows, const float4 *tab, const float *dEdG, double scale,
                        double *f, double *wrow, double *vatom, hipStream_t stream);

// ---- MLP ----------------------------------------------------------------------
// mode: 0 fp32, 1 bf16.  E[row] (fp32) and dEdG[row][width] (fp32).
hipError_t launch_mlp(int mode, const RowLayout &rl, const MlpParams &mp, const float *G, float *E,
                      float *dEdG, hipStream_t stream);
size_t mlp_lds_bytes(const MlpParams &mp, int mode, int R, MlpParams *out);

// ---- reductions / integration ---------------------------------------------------
hipError_t launch_zero(void *p, size_t bytes, hipStream_t stream);
// plain device-to-device copy (bandwidth calibration)
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t stream);
// out[0] = scale * sum_r E[r] over real rows; out[1..6] = sum_r wrow[r][*]
hipError_t launch_reduce_ev(const float *E, const double *wrow, const int *row2atom, int nrows,
                            double scale, double *out, hipStream_t stream);
hipError_t launch_vv_first(int n, double4 *pos, double *v, const double *f,
                           const double *dtfm_by_elem, double dt, hipStream_t stream);
hipError_t launch_vv_second(int n, const double4 *pos, double *v, const double *f,
                            const double *dtfm_by_elem, hipStream_t stream);
// kinetic energy sum(m v^2)/2 * mvv2e into out[0] (out zeroed inside)
hipError_t launch_ekin(int n, const double4 *pos, const double *v, const double *mass_by_elem,
                       double *out, hipStream_t stream);
// per-step energy history through a device counter (safe to replay in a graph)
hipError_t launch_record_step(const double *epot, const double *ekin, double *hist,
                              unsigned long long *counter, int cap, hipStream_t stream);
// copy e (per-step scalars) into history slots
hipError_t launch_copy_scalar(const double *src, double *dst, hipStream_t stream);

}  // namespace tmd
