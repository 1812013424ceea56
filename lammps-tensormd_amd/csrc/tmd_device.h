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

// One periodic cell (a batch of structures has one per structure).  The
// neighbor-build fields are ignored by the descriptor/force kernels.
struct CellInfo {
  double h[9];    // rows a, b, c
  double inv[9];  // inverse (fractional = x * inv)
  int nb[3];      // bins per axis
  int reach[3];   // stencil half-width per axis
  int bin_base;   // first global bin of this cell
  int pad;
};

// Monomials of max_moment 3 in the order degree 0, 1, 2, 3 and, within a
// degree, x-power then y-power descending (tensormd.cpp:205-276): degree and
// multinomial multiplicity of monomial d.  T[m][d] = kMult[d] if kDeg[d] == m;
// a T-symmetric model adds T[2][0] = -1/3 and T[3][1..3] = -3/5.
constexpr int kDeg[kND] = {0, 1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3};
constexpr int kMult[kND] = {1, 1, 1, 1, 1, 2, 2, 1, 2, 1, 1, 3, 3, 3, 6, 3, 1, 3, 3, 1};

// Per-model constants passed by value to the descriptor/force kernels.
// slots[ei] packs the slot of neighbor element ej in bits 4ej..4ej+3
// (eltij of tensormd.cpp:165-176).
struct ModelGeo {
  int nelt;
  int k;
  int tab_n;
  int tsym;  // 1: T-symmetric descriptor
  int mm;    // max_moment as evaluated: 3 (0-2 are padded to 3), 4 or 5
  uint32_t slots[kMaxElts];
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

// Sizes of one batch of structures, kept in device memory so that one
// captured graph (launched with grids sized for the capacities) serves
// every batch up to those capacities.  Kernels given a non-null BatchDims
// read their extents from it; workgroups beyond them exit at once.
struct BatchDims {
  int natoms;
  int nbins;
  int nstruct;
  int pad;
  RowLayout rl;
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
  // fast path: the MFMA A operands of hidden layer l in the context's
  // precision (FP32, BF16, a BF16 hi + lo pair, or FP8 e4m3 divided by the
  // per-output-channel scale), stored in fragment order: tile t of 16 output channels, step s
  // of KS inner values, lane, then the lane's values, so that a wave loads
  // one fragment as one contiguous block.  Af: forward (outputs N, inner K),
  // Ab: backward (outputs K, inner N).  Element e at byte offset
  // e * wstride * (4, 2 or 1).
  const void *Af[kMaxLayers];
  const void *Ab[kMaxLayers];
  const float *swf[kMaxLayers];     // forward per-output-channel scales [N]
  const float *swb[kMaxLayers];     // backward per-output-channel scales [K]
  const float *b[kMaxLayers];    // hidden biases (or null)
  const float *wout;             // output layer weights [K_last]
  const float *bout;             // output bias (or null)
  long wstride[kMaxLayers];
  int bstride[kMaxLayers];
  // LDS layout (elements of the mode's type)
  int offH[kMaxLayers], offD[kMaxLayers], ld[kMaxLayers + 1];
  int offGH, ldGH;
  int offRS;   // FP8: per-row scale and inverse scale [2][R] (floats, byte offset)
  int offQ, ldq;  // FP8: the next GEMM's B operand rows, quantized [R][ldq] (bytes)
  int storeD;  // D_l kept in LDS (otherwise recomputed from H_l: tanh)
};

// ---- neighbor build ---------------------------------------------------------
struct NbrBuild {
  int n;
  int cap;            // entries per atom
  double cut2;        // (rc + skin)^2
  double trig2;       // (skin/2)^2
  int nbins;          // total bins over all cells
  const CellInfo *cells;   // [ncells] (device)
  const int *atom_struct;  // [n] cell of each atom, or null (one cell)
  const BatchDims *dyn;    // device extents (n, nbins) when non-null
  int *bin_count;     // [nbins + 1]
  int *bin_start;     // [nbins + 1]
  int *bin_cursor;    // [nbins]
  int *bin_atoms;     // [n]
  int *atom_bin;      // [n] flat bin
  int *atom_bc;       // [n][3] bin coordinates
  double *x_build;    // [n][4]
  unsigned long long *ctl;  // [0] max d^2 bits, [1] force flag, [2] overflow, [3] build count
};

hipError_t nbr_step(const NbrBuild &nb, double4 *pos, uint32_t *nbr, int *nn, hipStream_t stream);
// force the next nbr_step to rebuild (device-side flag; graph-safe)
hipError_t nbr_force_rebuild(const NbrBuild &nb, hipStream_t stream);

// ---- descriptor / force -------------------------------------------------------
// Geometry of an evaluation: cells[atom_struct[i]] (atom_struct null: cells[0]).
struct Geom {
  const CellInfo *cells;
  const int *atom_struct;
  const BatchDims *dyn;  // rows beyond dyn->rl.nrows are skipped when non-null
  // the order in which the descriptor and force kernels take the rows: a
  // permutation of [0, nrows) (launch position p runs row order[p]); null:
  // row order.  Rows are element blocks, so atoms of one element run
  // together; the standalone system orders them by atom index instead,
  // which for spatially ordered atoms keeps neighboring atoms' work together
  const int *order = nullptr;
  // the descriptor kernel stores G as BF16 (round to nearest even, the BF16
  // MLP's own operand rounding) instead of FP32; max_moment <= 3 kernels only
  bool g_bf16 = false;
};
// The descriptor kernel's compaction of each row's list, reused by the force
// kernel of the same evaluation: kl[row][kKeptHdr + cap], kl[0] = nk,
// kl[1 + b] = first entry of slot b (b = 0..nelt), then the nk kept entries
// (within rc, sorted by slot).
constexpr int kKeptHdr = 16;
__host__ __device__ inline size_t kept_cache_words(int cap) { return (size_t) kKeptHdr + (size_t) cap; }

// Descriptor rows G[row] and the kept list of every row.
hipError_t launch_descriptor(const ModelGeo &g, const Geom &geo, const double4 *pos,
                             const uint32_t *nbr, const int *nn, int cap, const int *row2atom,
                             int nrows, const float4 *tab, float *G, uint32_t *kl,
                             hipStream_t stream, int row_begin = 0);

// Force/virial over the descriptor's kept lists: dEdG rows -> f (atomics,
// FP64, must be zeroed); per atom: eatom[i] = scale * E[row] and
// watom[i][6] = sum over the atom's pairs of -d (x) F (both written only for
// atoms that own a row); optional vatom (FP64 atomics; must be zeroed).
hipError_t launch_force(const ModelGeo &g, const Geom &geo, const double4 *pos, int cap,
                        const int *row2atom, int nrows, const float4 *tab, const uint32_t *kl,
                        const float *dEdG, const double *E, double scale, double *f, double *eatom,
                        double *watom, double *vatom, hipStream_t stream, int row_begin = 0);

// The same two kernels for max_moment 4 and 5 models (g.mm 4 or 5): 35 or 56
// monomials (tmd_desc_hm.hip); launch_descriptor/launch_force call them.
hipError_t launch_descriptor_hm(const ModelGeo &g, const Geom &geo, const double4 *pos,
                                const uint32_t *nbr, const int *nn, int cap, const int *row2atom,
                                int nrows, const float4 *tab, float *G, uint32_t *kl,
                                hipStream_t stream, int row_begin = 0);
hipError_t launch_force_hm(const ModelGeo &g, const Geom &geo, const double4 *pos, int cap,
                           const int *row2atom, int nrows, const float4 *tab, const uint32_t *kl,
                           const float *dEdG, const double *E, double scale, double *f,
                           double *eatom, double *watom, double *vatom, hipStream_t stream,
                           int row_begin = 0);

// ---- MLP ----------------------------------------------------------------------
// mode: 0 fp32, 1 bf16, 2 fp8 (e4m3 operands: weights with per-output-channel
// scales, activations and gradients with per-row scales; FP32 accumulation),
// 3 split bf16 (hi + lo operands, three BF16 products each).
// E[row] (FP64 output sum) and dEdG[row][width] (fp32).  rtile: rows per
// workgroup (16, 32 or 64; 0 = default).
// With dyn, rl gives the capacity (grid) and dyn->rl the live layout.
// Rows [row_begin, row_end) only (row_end < 0: to rl.nrows); both bounds
// are multiples of 64 (element blocks start on such rows).
// g_bf16: G holds BF16 rows (the descriptor kernel's Geom::g_bf16 output);
// only where mlp_bf16_rows() holds.
hipError_t launch_mlp(int mode, const RowLayout &rl, const MlpParams &mp, const float *G, double *E,
                      float *dEdG, hipStream_t stream, int rtile = 0,
                      const BatchDims *dyn = nullptr, int row_begin = 0, int row_end = -1,
                      bool g_bf16 = false);
// the MLP launch takes BF16 descriptor rows: BF16 mode, tanh, a compile-time
// layer shape, 32-row tiles
bool mlp_bf16_rows(const MlpParams &mp, int mode, int rtile);
size_t mlp_lds_bytes(const MlpParams &mp, int mode, int R, MlpParams *out);
// the fast (aligned) MLP kernel applies: required by the FP8 mode
bool mlp_fast_ok(const MlpParams &mp, int mode);
// FP8 A operands (MlpParams::Af/Ab layout) of one layer's weights for nelt
// elements: W [nelt][K][N] FP32, per-output-channel scales swf [nelt][N]
// (forward) and swb [nelt][K] (backward)
hipError_t launch_quant_fp8(const float *W, const float *swf, const float *swb, uint8_t *Af,
                            uint8_t *Ab, int K, int N, int nelt, hipStream_t stream);

// ---- TDNP: temperature-dependent models (src/ML-TENSORMD/tdnp.cpp) ------------------
// One network of every element: layer l of element e is W[l] + e K N
// ([K][N]), Wt[l] + e N K ([N][K]) and b[l] + e N (null without bias).
struct TdnpNet {
  int nl;
  int act;
  int sizes[kMaxLayers + 1];
  int res[kMaxLayers];
  const float *W[kMaxLayers];
  const float *Wt[kMaxLayers];
  const float *b[kMaxLayers];
};
// net[0] = H (descriptor -> ydim - 1, linear last layer), net[1] = U and
// net[2] = S (ydim -> 1); etemp: device scalar (eV), read at launch time.
struct TdnpParams {
  TdnpNet net[3];
  int ydim;
  int sommerfeld;
  const double *etemp;
};
// Per row: E[row] = u (FP64), sa[row] = {s, a} (FP64, a = u - T s), and
// dEdG[row] = dA/dG.  Padding rows get zeros.
hipError_t launch_tdnp(const RowLayout &rl, const TdnpParams &tp, const int *row2atom, const float *G,
                       double *E, double *sa, float *dEdG, hipStream_t stream,
                       const BatchDims *dyn = nullptr);

// ---- batches of structures (serving) ---------------------------------------------
// Capacities of one batch slot (grids of its captured graph).
struct BatchCaps {
  int atoms, rows, bins, structs;
  int nelt;  // model elements (valid element indices in pos.w)
};
// Device arrays of a batch slot.
struct BatchArrays {
  BatchDims *dyn;
  int *struct_off;    // [structs + 1]
  CellInfo *cells;    // [structs]
  int *atom_struct;   // [atoms]
  int *row2atom;      // [rows]
  double4 *pos;       // [atoms]
  const double *struct_out;  // [structs][7]
  const double *f;           // [atoms][3]
  const double *eatom;       // [atoms]
  unsigned long long *ctl;   // neighbor control word
};
// Dense host staging of a batch: BatchDims | struct_off[ns+1] | cells[ns] |
// atom_struct[n] | row2atom[nrows] | pos[n].  Byte offsets from the start:
// off[1] cells, off[2] atom_struct, off[3] row2atom, off[4] pos, off[5] and
// off[6] the total size (struct_off follows the header directly).
__host__ __device__ inline void batch_stage_offsets(int ns, int n, int nrows, size_t *off)
{
  auto up = [](size_t v) { return (v + 255) & ~(size_t) 255; };
  off[0] = 0;
  off[1] = up(sizeof(BatchDims) + 4 * (size_t) (ns + 1));
  off[2] = up(off[1] + sizeof(CellInfo) * (size_t) ns);
  off[3] = up(off[2] + 4 * (size_t) n);
  off[4] = up(off[3] + 4 * (size_t) nrows);
  off[5] = off[4] + sizeof(double4) * (size_t) n;
  off[6] = off[5];
}
// Dense host output of a batch: 8 header doubles (list overflow count,
// atoms, structures, error flag) | struct_out[ns][7] | f[n][3] | eatom[n].
inline size_t batch_out_bytes(int ns, int n) { return 8 * (8 + 7 * (size_t) ns + 4 * (size_t) n); }
// Copy a staged batch (host memory, read by the kernels) into the slot's
// device arrays, clamping every count and index to the capacities.
hipError_t launch_batch_unpack(const void *stage, const BatchCaps &caps, const BatchArrays &a,
                               hipStream_t stream);
// Write a finished batch's results to host memory `out` and clear the
// list-overflow word.
hipError_t launch_batch_pack(void *out, const BatchCaps &caps, const BatchArrays &a,
                             hipStream_t stream);

// ---- reductions / integration ---------------------------------------------------
hipError_t launch_zero(void *p, size_t bytes, hipStream_t stream);
// plain device-to-device copy (bandwidth calibration)
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t stream);
// out[0] = sum_i eatom[i]; out[1..6] = sum_i watom[i][*], i < n (fixed order)
hipError_t launch_reduce_ev(const double *eatom, const double *watom, int n, double *out,
                            hipStream_t stream);
// per structure s: out[7 s + 0] = sum eatom, out[7 s + 1..6] = sum watom over
// atoms [off[s], off[s+1])
// (with dyn: nstruct is the capacity, dyn->nstruct the live count)
hipError_t launch_reduce_struct(const double *eatom, const double *watom, const int *off,
                                int nstruct, double *out, hipStream_t stream,
                                const BatchDims *dyn = nullptr);
// p[0 .. per_atom * dyn->natoms) = 0 (grid sized for cap_atoms)
hipError_t launch_zero_atoms(double *p, int per_atom, int cap_atoms, const BatchDims *dyn,
                             hipStream_t stream);
hipError_t launch_vv_first(int n, double4 *pos, double *v, const double *f,
                           const double *dtfm_by_elem, double dt, hipStream_t stream);
hipError_t launch_vv_second(int n, const double4 *pos, double *v, const double *f,
                            const double *dtfm_by_elem, hipStream_t stream);
// kinetic energy sum(m v^2)/2 * mvv2e into out[0] (out zeroed inside)
hipError_t launch_ekin(int n, const double4 *pos, const double *v, const double *mass_by_elem,
                       double *out, hipStream_t stream);
// per-step energy history through a device counter (safe to replay in a graph)
// *acc += sum of the kept-neighbor counts of the real rows (row2atom >= 0)
// in the kept-list cache kl (words per row)
hipError_t launch_sum_kept(const uint32_t *kl, size_t words, const int *row2atom, int nrows,
                           unsigned long long *acc, hipStream_t s);
hipError_t launch_record_step(const double *epot, const double *ekin, double *hist,
                              unsigned long long *counter, int cap, hipStream_t stream);
// copy e (per-step scalars) into history slots
hipError_t launch_copy_scalar(const double *src, double *dst, hipStream_t stream);

}  // namespace tmd
