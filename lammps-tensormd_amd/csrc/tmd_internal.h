// Host-side state behind the opaque C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/tensormd_amd.h"
#include "tmd_device.h"
#include "tmd_model.h"

namespace tmd {

// Device allocation that only grows, or (view()) a fixed window of memory
// the buffer does not own: ensure() on a view never allocates and fails with
// hipErrorOutOfMemory when the window is too small.
struct DeviceBuffer {
  void *p = nullptr;
  size_t bytes = 0;
  bool owned = true;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : p(o.p), bytes(o.bytes), owned(o.owned)
  {
    o.p = nullptr;
    o.bytes = 0;
    o.owned = true;
  }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept
  {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      owned = o.owned;
      o.p = nullptr;
      o.bytes = 0;
      o.owned = true;
    }
    return *this;
  }
  ~DeviceBuffer();
  void release();
  hipError_t ensure(size_t nbytes);
  void set_view(void *base, size_t nbytes)
  {
    release();
    p = base;
    bytes = nbytes;
    owned = false;
  }
};

// Page-locked host allocation (hipHostMalloc) that only grows.
struct PinnedBuffer {
  void *p = nullptr;
  size_t bytes = 0;
  PinnedBuffer() = default;
  PinnedBuffer(const PinnedBuffer &) = delete;
  PinnedBuffer &operator=(const PinnedBuffer &) = delete;
  ~PinnedBuffer();
  void release();
  hipError_t ensure(size_t nbytes);
};

// Buffers of one evaluation: positions, list, MLP activations, outputs.
// ew holds eatom[n] followed by watom[n][6]; kl holds the descriptor
// kernel's kept list of every row, reused by the force kernel.
struct Workspace {
  RowLayout rl{};
  DeviceBuffer pos, nbr, nn, f, vatom, ew;
  DeviceBuffer G, dEdG, E, row2atom, red, cells;
  DeviceBuffer kl;
  DeviceBuffer sa;  // TDNP models: per row electron entropy and free energy (FP64 pairs)
  bool g_bf16 = false;  // G holds BF16 rows (the last chain's choice)
};

struct Server;
void server_destroy(Server *s);

}  // namespace tmd

struct tmd_ctx {
  std::string err;
  tmd_options opt{};
  tmd::Model model;
  int device = 0;
  hipStream_t stream = nullptr;
  int ntypes = 0;
  std::vector<int> type_map;
  tmd::ModelGeo geo{};
  tmd::DeviceBuffer tab, params;
  tmd::MlpParams mlp{};
  // TDNP models: network parameters, the electron temperature (eV; a device
  // copy that launches read, so captured graphs see updates) and the
  // entropy and free energy of the last list-mode evaluation (scaled)
  tmd::DeviceBuffer tdnp_params, etemp_dev;
  tmd::TdnpParams tdnp{};
  double etemp = 0.0;
  std::vector<double> td_satom, td_aatom;
  double td_entropy = 0.0, td_free_energy = 0.0;
  bool td_valid = false;
  tmd::Workspace list_ws;
  // list mode: pinned host staging, and what list_ws holds (the serial of
  // the last tmd_compute_list_serial call and the list's shape)
  tmd::PinnedBuffer list_host;
  bool list_cached = false;
  long long list_serial = 0;
  int list_nlocal = 0, list_nall = 0, list_inum = 0, list_cap = 0;
  tmd::Server *server = nullptr;  // request loop (tmd_serve_start)
  tmd_ctx() = default;
  tmd_ctx(const tmd_ctx &) = delete;
  tmd_ctx &operator=(const tmd_ctx &) = delete;
  ~tmd_ctx();  // destroys the stream (also when tmd_create fails part way)
};

namespace tmd {
// Optional events recorded between the phases of one evaluation:
// [0] start, [1] after descriptor, [2] after MLP, [3] after the force-buffer
// clears (just before the force kernel), [4] after force, [5] after reductions
struct PhaseEvents {
  hipEvent_t ev[6];
};
// Records an error message for the calling thread and, on a caller's
// thread, in ctx->err (tmd_last_error).
void set_err(tmd_ctx *ctx, const std::string &m);
// Called first on a thread the library starts itself: set_err then leaves
// every ctx->err alone on that thread.
void mark_library_thread();
// The calling thread's last error message.
const std::string &thread_error();
int hip_fail(tmd_ctx *ctx, hipError_t e, const char *what);
int ensure_rows(tmd_ctx *ctx, Workspace &w, int nrows);
// the kept-list cache for nrows rows of a list with `cap` entries per atom
// (allocate before any graph capture of run_chain)
int ensure_kept(tmd_ctx *ctx, Workspace &w, int nrows, int cap);
// Inputs and outputs of one descriptor -> MLP -> force -> reduction chain.
struct Chain {
  Geom geo{};
  const double4 *pos = nullptr;
  const uint32_t *nbr = nullptr;
  const int *nn = nullptr;
  int cap = 0;
  double scale = 1.0;
  double *f = nullptr;       // zeroed, then accumulated
  size_t fbytes = 0;
  double *vatom = nullptr;   // optional; zeroed, then accumulated
  size_t vbytes = 0;
  int natoms = 0;            // eatom/watom entries (in w.ew)
  bool zero_ew = false;      // some atoms own no row: clear eatom/watom first
  const int *struct_off = nullptr;  // per-structure sums into struct_out[7 s ..]
  int nstruct = 0;
  double *struct_out = nullptr;     // (null struct_off: one sum into w.red[0..6])
  int rtile = 0;
  const BatchDims *dyn = nullptr;   // device extents of a batch (capture-once graphs)
  // BF16 descriptor rows where the MLP launch takes them (mlp_bf16_rows):
  // the BF16 MLP rounds its layer-0 operands to BF16 (nearest even) anyway,
  // so storing them rounded gives the same results with half the G traffic
  // (128k-atom step: MLP 0.246 -> 0.204 ms, descriptor 0.436 -> 0.426 ms);
  // FP32 rows elsewhere
  bool g_bf16 = false;
};
int ensure_atoms(tmd_ctx *ctx, Workspace &w, int natoms);
int run_chain(tmd_ctx *ctx, Workspace &w, const Chain &c, hipStream_t s, PhaseEvents *pe = nullptr);
// ---- list mode (Pair::compute over a LAMMPS full list) ----
struct ListArgs {
  int nlocal = 0, nall = 0;
  const double *x = nullptr;
  const int *type = nullptr;
  int inum = 0;
  const int *ilist = nullptr;
  const int *numneigh = nullptr;
  const int *const *firstneigh = nullptr;
  double scale = 1.0;
  double *f = nullptr, *eatom = nullptr, *vatom = nullptr, *energy = nullptr, *virial = nullptr;
  double *Gout = nullptr;  // set: descriptor rows only ([nlocal][file width])
  const char *who = "tmd_compute_list";
};
// Where one list-mode call's buffers live.  fixed = false: the context's
// growable buffers (list_ws, list_host).  fixed = true: device buffers are
// carved from [dev_base, dev_base + dev_bytes) and host staging from
// [host, host + host_bytes); nothing is allocated, and a call that needs
// more is refused with TMD_ERR_CAPACITY.
struct ListResources {
  bool fixed = false;
  void *dev_base = nullptr;
  size_t dev_bytes = 0;
  unsigned char *host = nullptr;
  size_t host_bytes = 0;
};
// serial >= 0 (tmd_compute_list_serial): the list is reused from the
// previous call with the same serial and shape; serial < 0: always packed.
int list_compute(tmd_ctx *ctx, const ListArgs &a, const ListResources &r, long long serial);
// The fixed device scratch of the no-allocation entry point on `device`
// (tmd_noalloc.hip): looked up once per device, at context creation.
int noalloc_scratch_setup(int device, std::string &err);

bool invert3(const double *h, double *inv);
double cell_volume(const double *h);
int setup_cell(const double *h, double cut, CellInfo &c, std::string &err);
void make_rows(int nelt, const std::vector<int> &atoms, const std::vector<int> &elem_of_atom,
               RowLayout &rl, std::vector<int> &row2atom);
}  // namespace tmd
