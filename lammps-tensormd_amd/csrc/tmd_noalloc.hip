// tmd_compute_list_noalloc: the list-mode force evaluation with no device
// (or host) allocation in the call.  Every device buffer of the call is a
// window of `scratch`, a fixed file-scope device array; the host staging is
// a fixed static array.  For callers that cannot allocate: code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// `scratch` is a device symbol: its host-side name is not a device address.
// noalloc_scratch_setup() looks its address up once per device (with that
// device current) at context creation, and only that address is handed to
// kernels and copies.  Each device holds its own copy of the array until
// the process ends.
//
// The array keeps the last call's contents; every buffer that must start at
// zero (forces, per-atom virials, per-atom energies of atoms without a row)
// is cleared by a kernel inside the call.  Calls are serialized by a
// process-wide lock (the host staging is shared by all devices); without it
// two calls on one device would race on the scratch.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "tmd_internal.h"

namespace {

// 64 MiB per device: a 2,048-atom C1-shaped call (5.8k positions with
// ghosts) uses about 4 MiB; a 4-element k = 24 model fits about 10,000 local
// atoms.
constexpr size_t kScratchBytes = 64ull << 20;
constexpr size_t kHostStageBytes = 32ull << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

alignas(256) unsigned char g_host_stage[kHostStageBytes];

std::mutex g_mu;  // guards g_scratch and serializes the entry point
void *g_scratch[kMaxDevices] = {};

}  // namespace

namespace tmd {

int noalloc_scratch_setup(int device, std::string &err)
{
  if (device < 0 || device >= kMaxDevices) {
    err = "device ordinal out of range for the no-allocation scratch table";
    return TMD_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return TMD_OK;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e == hipSuccess && cur != device) e = hipErrorInvalidDevice;  // the caller made it current
  void *ptr = nullptr;
  if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    err = std::string("no-allocation scratch lookup: ") + hipGetErrorString(e);
    return TMD_ERR_DEVICE;
  }
  g_scratch[device] = ptr;
  return TMD_OK;
}

}  // namespace tmd

extern "C" int tmd_compute_list_noalloc(tmd_ctx *ctx, int nlocal, int nall, const double *x,
                                        const int *type, int inum, const int *ilist,
                                        const int *numneigh, const int *const *firstneigh,
                                        double scale, double *f, double *eatom, double *vatom,
                                        double *energy, double *virial)
{
  using namespace tmd;
  if (!ctx) {
    set_err(nullptr, "tmd_compute_list_noalloc: ctx is NULL");
    return TMD_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  void *ptr = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch[ctx->device] : nullptr;
  if (!ptr) {  // tmd_create looks the address up; a context always has it
    set_err(ctx, "tmd_compute_list_noalloc: no scratch address for this device");
    return TMD_ERR_STATE;
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
  a.who = "tmd_compute_list_noalloc";
  ListResources r;
  r.fixed = true;
  r.dev_base = ptr;
  r.dev_bytes = kScratchBytes;
  r.host = g_host_stage;
  r.host_bytes = kHostStageBytes;
  return list_compute(ctx, a, r, -1);
}
