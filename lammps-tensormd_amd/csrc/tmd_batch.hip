// Staging kernels of the batched single-point path.  One kernel copies a
// packed batch (pinned host memory, read directly by the GPU) into the
// slot's fixed device arrays and one writes the results back, so a batch is
// a single graph launch with no memcpy nodes and no host round trips.
//
// Every count and index read from the staging area is clamped to the slot
// capacities before any kernel uses it; a clamp sets an error flag that the
// host reads back with the results.
#include "tmd_device.h"

namespace tmd {

namespace {

__global__ void k_batch_header(const unsigned char *stage, BatchCaps caps, BatchArrays a)
{
  if (threadIdx.x != 0) return;
  BatchDims d = *reinterpret_cast<const BatchDims *>(stage);
  int err = 0;
  auto clampi = [&err](int v, int hi) {
    if (v < 0 || v > hi) {
      err = 1;
      return v < 0 ? 0 : hi;
    }
    return v;
  };
  d.natoms = clampi(d.natoms, caps.atoms);
  d.nstruct = clampi(d.nstruct, caps.structs);
  d.nbins = clampi(d.nbins, caps.bins);
  d.rl.nrows = clampi(d.rl.nrows, caps.rows);
  // element row groups: non-decreasing, empty beyond the model's elements
  int prev = 0;
  for (int e = 0; e <= kMaxElts; e++) {
    int v = clampi(d.rl.row_off[e], d.rl.nrows);
    if (v < prev || (e >= caps.nelt && v != d.rl.nrows && e > 0)) {
      err = 1;
      v = e >= caps.nelt ? d.rl.nrows : prev;
    }
    d.rl.row_off[e] = v;
    prev = v;
  }
  if (d.rl.row_off[0] != 0) err = 1;
  d.rl.row_off[0] = 0;
  d.pad = err;
  *a.dyn = d;
  a.ctl[1] = 1ull;  // every batch builds its neighbor list
}

__global__ void k_batch_unpack(const unsigned char *stage, BatchCaps caps, BatchArrays a)
{
  const BatchDims d = *a.dyn;
  const int ns = d.nstruct, n = d.natoms, nr = d.rl.nrows;
  size_t off[7];
  batch_stage_offsets(ns, n, nr, off);
  const unsigned char *body = stage + sizeof(BatchDims);
  const int *soff = reinterpret_cast<const int *>(body);
  const CellInfo *cells = reinterpret_cast<const CellInfo *>(stage + off[1]);
  const int *as = reinterpret_cast<const int *>(stage + off[2]);
  const int *r2a = reinterpret_cast<const int *>(stage + off[3]);
  const double4 *pos = reinterpret_cast<const double4 *>(stage + off[4]);
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  int bad = 0;
  for (int q = tid; q <= ns; q += stride) {
    int v = soff[q];
    if (v < 0 || v > n) bad = 1;
    a.struct_off[q] = min(max(v, 0), n);
  }
  for (int q = tid; q < ns; q += stride) {
    CellInfo c = cells[q];
    long nb = 1;
    for (int k = 0; k < 3; k++) {
      if (c.nb[k] < 1 || c.nb[k] > 1024 || c.reach[k] < 0 || c.reach[k] > 64) {
        bad = 1;
        c.nb[k] = 1;
        c.reach[k] = 0;
      }
      nb *= c.nb[k];
    }
    if (c.bin_base < 0 || c.bin_base + nb > d.nbins) {
      bad = 1;
      c.nb[0] = c.nb[1] = c.nb[2] = 1;
      c.reach[0] = c.reach[1] = c.reach[2] = 0;
      c.bin_base = 0;
    }
    a.cells[q] = c;
  }
  for (int q = tid; q < n; q += stride) {
    int s = as[q];
    if (s < 0 || s >= ns) {
      bad = 1;
      s = 0;
    }
    a.atom_struct[q] = s;
  }
  for (int q = tid; q < nr; q += stride) {
    int i = r2a[q];
    if (i >= n) {
      bad = 1;
      i = -1;
    }
    a.row2atom[q] = i < 0 ? -1 : i;
  }
  for (int q = tid; q < n; q += stride) {
    double4 p = pos[q];
    if (!(p.w >= 0.0 && p.w < (double) caps.nelt)) {
      bad = 1;
      p.w = 0.0;
    }
    a.pos[q] = p;
  }
  if (bad) atomicOr(&a.dyn->pad, 1);
}

__global__ void k_batch_pack(double *out, BatchArrays a)
{
  const BatchDims d = *a.dyn;
  const int ns = d.nstruct, n = d.natoms;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  if (tid == 0) {
    out[0] = (double) a.ctl[2];
    a.ctl[2] = 0ull;
    out[1] = (double) n;
    out[2] = (double) ns;
    out[3] = (double) d.pad;
    for (int q = 4; q < 8; q++) out[q] = 0.0;
  }
  double *so = out + 8;
  double *fo = so + 7 * (size_t) ns;
  double *eo = fo + 3 * (size_t) n;
  for (int q = tid; q < 7 * ns; q += stride) so[q] = a.struct_out[q];
  for (int q = tid; q < 3 * n; q += stride) fo[q] = a.f[q];
  for (int q = tid; q < n; q += stride) eo[q] = a.eatom[q];
}

int grid_for(long work)
{
  long b = (work + 255) / 256;
  return (int) (b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

}  // namespace

hipError_t launch_batch_unpack(const void *stage, const BatchCaps &caps, const BatchArrays &a,
                               hipStream_t stream)
{
  const unsigned char *st = static_cast<const unsigned char *>(stage);
  k_batch_header<<<1, 64, 0, stream>>>(st, caps, a);
  long work = caps.atoms > caps.rows ? caps.atoms : caps.rows;
  k_batch_unpack<<<grid_for(work), 256, 0, stream>>>(st, caps, a);
  return hipGetLastError();
}

hipError_t launch_batch_pack(void *out, const BatchCaps &caps, const BatchArrays &a,
                             hipStream_t stream)
{
  k_batch_pack<<<grid_for(3L * caps.atoms), 256, 0, stream>>>(static_cast<double *>(out), a);
  return hipGetLastError();
}

}  // namespace tmd
