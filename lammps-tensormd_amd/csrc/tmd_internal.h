// Host-side state behind the opaque C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/tensormd_amd.h"
#include "tmd_device.h"
#include "tmd_model.h"

namespace tmd {

// Owning device allocation that only grows.
struct DeviceBuffer {
  void *p = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : p(o.p), bytes(o.bytes)
  {
    o.p = nullptr;
    o.bytes = 0;
  }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept
  {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DeviceBuffer();
  void release();
  hipError_t ensure(size_t nbytes);
};

// Buffers of one evaluation: positions, list, MLP activations, outputs.
struct Workspace {
  RowLayout rl{};
  DeviceBuffer pos, nbr, nn, f, vatom;
  DeviceBuffer G, dEdG, E, wrow, row2atom, red;
};

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
  tmd::Workspace list_ws;
};

namespace tmd {
// Optional events recorded between the phases of one evaluation:
// [0] start, [1] after descriptor, [2] after MLP, [3] after force, [4] after reductions
struct PhaseEvents {
  hipEvent_t ev[5];
};
void set_err(tmd_ctx *ctx, const std::string &m);
int hip_fail(tmd_ctx *ctx, hipError_t e, const char *what);
int ensure_rows(tmd_ctx *ctx, Workspace &w, int nrows);
int run_chain(tmd_ctx *ctx, Workspace &w, const Cell &cell, const double4 *pos, const uint32_t *nbr,
              const int *nn, int cap, double scale, double *f, size_t fbytes, double *vatom,
              size_t vbytes, hipStream_t s, PhaseEvents *pe = nullptr);
void make_rows(int nelt, const std::vector<int> &atoms, const std::vector<int> &elem_of_atom,
               RowLayout &rl, std::vector<int> &row2atom);
}  // namespace tmd
