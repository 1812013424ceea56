// Device neighbor build for periodic cells (any triclinic cell).
//
// Replaces LAMMPS NPairFullBin::build (src/npair_full_bin.cpp:68-126) with
// binning (src/nbin_standard.cpp:82-160) and the displacement trigger of
// Neighbor::decide/check_distance (src/neighbor.cpp:2293-2360, triggersq at
// :344).  The full list keeps every (j, image) with |d|^2 <= (rc+skin)^2,
// i != j or a non-zero image (npair_full_bin.cpp:99).  Bins are cells of the
// fractional lattice; a stencil of (2*reach+1)^3 bins reaches every
// candidate, including several periodic images when the cell is small.
//
// The rebuild decision is taken on the device: every kernel here reads the
// control word and exits when no rebuild is due, so a step never waits for
// the host.
//
// One launch serves a batch of independent structures: each atom carries
// its cell (atom_struct), each cell owns a contiguous range of global bins,
// and the stencil of an atom only visits its own cell's bins, so atoms of
// different structures never see each other.
#include "tmd_device.h"

namespace tmd {

namespace {

__device__ inline void displacement(const double4 &xi, const double4 &xj, int s0, int s1, int s2,
                                    const CellInfo &c, double d[3])
{
#pragma clang fp contract(off)
  // Same expression, same order as the oracle: (xj - xi) + ((s0 h0 + s1 h1) + s2 h2)
  double sh0 = ((double) s0 * c.h[0] + (double) s1 * c.h[3]) + (double) s2 * c.h[6];
  double sh1 = ((double) s0 * c.h[1] + (double) s1 * c.h[4]) + (double) s2 * c.h[7];
  double sh2 = ((double) s0 * c.h[2] + (double) s1 * c.h[5]) + (double) s2 * c.h[8];
  d[0] = (xj.x - xi.x) + sh0;
  d[1] = (xj.y - xi.y) + sh1;
  d[2] = (xj.z - xi.z) + sh2;
}

__device__ inline double dist2(const double d[3])
{
#pragma clang fp contract(off)
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

__device__ inline const CellInfo &cell_of(const NbrBuild &nb, int i)
{
  return nb.cells[nb.atom_struct ? nb.atom_struct[i] : 0];
}

__device__ inline int live_atoms(const NbrBuild &nb) { return nb.dyn ? nb.dyn->natoms : nb.n; }
__device__ inline int live_bins(const NbrBuild &nb, int nbins) { return nb.dyn ? nb.dyn->nbins : nbins; }

__device__ inline bool rebuild_due(const unsigned long long *ctl, double trig2)
{
  double m = __longlong_as_double((long long) ctl[0]);
  return ctl[1] != 0ull || m > trig2;
}

// max |x - x_build|^2 over atoms -> ctl[0] (as ordered bits of a positive double)
__global__ void k_nb_check(NbrBuild nb, const double4 *pos, const double *xb, unsigned long long *ctl)
{
  const int n = live_atoms(nb);
  if ((int) (blockIdx.x * blockDim.x) >= n) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < n) {
    double4 p = pos[i];
    double dx = p.x - xb[4 * i], dy = p.y - xb[4 * i + 1], dz = p.z - xb[4 * i + 2];
    d2 = dx * dx + dy * dy + dz * dz;
  }
  for (int o = 32; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o));
  __shared__ double wmax[16];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < (int) (blockDim.x >> 6); w++) m = fmax(m, wmax[w]);
    if (m > 0.0) atomicMax(&ctl[0], (unsigned long long) __double_as_longlong(m));
  }
}

// wrap into the cell, remember build positions, bin the atoms
__global__ void k_nb_bin(NbrBuild nb, double4 *pos)
{
  if (!rebuild_due(nb.ctl, nb.trig2)) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= live_atoms(nb)) return;
  const CellInfo &c = cell_of(nb, i);
  double4 p = pos[i];
  double f[3];
  f[0] = p.x * c.inv[0] + p.y * c.inv[3] + p.z * c.inv[6];
  f[1] = p.x * c.inv[1] + p.y * c.inv[4] + p.z * c.inv[7];
  f[2] = p.x * c.inv[2] + p.y * c.inv[5] + p.z * c.inv[8];
  double s[3] = {floor(f[0]), floor(f[1]), floor(f[2])};
  if (s[0] != 0.0 || s[1] != 0.0 || s[2] != 0.0) {
    p.x -= (s[0] * c.h[0] + s[1] * c.h[3]) + s[2] * c.h[6];
    p.y -= (s[0] * c.h[1] + s[1] * c.h[4]) + s[2] * c.h[7];
    p.z -= (s[0] * c.h[2] + s[1] * c.h[5]) + s[2] * c.h[8];
    pos[i] = p;
    for (int a = 0; a < 3; a++) f[a] -= s[a];
  }
  nb.x_build[4 * i + 0] = p.x;
  nb.x_build[4 * i + 1] = p.y;
  nb.x_build[4 * i + 2] = p.z;
  int b[3];
  for (int a = 0; a < 3; a++) {
    int q = (int) floor(f[a] * c.nb[a]);
    b[a] = min(max(q, 0), c.nb[a] - 1);
    nb.atom_bc[3 * i + a] = b[a];
  }
  int flat = c.bin_base + (b[0] * c.nb[1] + b[1]) * c.nb[2] + b[2];
  nb.atom_bin[i] = flat;
  atomicAdd(&nb.bin_count[flat], 1);
}

// exclusive scan of bin counts (single workgroup of 1024 threads)
__global__ void k_nb_scan(NbrBuild nb, int nbins)
{
  if (!rebuild_due(nb.ctl, nb.trig2)) return;
  nbins = live_bins(nb, nbins);
  __shared__ int part[1024];
  int t = threadIdx.x;
  int per = (nbins + 1023) / 1024;
  int lo = t * per, hi = min(lo + per, nbins);
  int s = 0;
  for (int b = lo; b < hi; b++) s += nb.bin_count[b];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = (t >= o) ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = (t == 0) ? 0 : part[t - 1];
  for (int b = lo; b < hi; b++) {
    nb.bin_start[b] = run;
    nb.bin_cursor[b] = run;
    run += nb.bin_count[b];
  }
  if (t == 1023) nb.bin_start[nbins] = part[1023];
}

__global__ void k_nb_fill(NbrBuild nb)
{
  if (!rebuild_due(nb.ctl, nb.trig2)) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= live_atoms(nb)) return;
  int slot = atomicAdd(&nb.bin_cursor[nb.atom_bin[i]], 1);
  nb.bin_atoms[slot] = i;
}

// make the order inside each bin deterministic (ascending atom index)
__global__ void k_nb_sortbins(NbrBuild nb, int nbins)
{
  if (!rebuild_due(nb.ctl, nb.trig2)) return;
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= live_bins(nb, nbins)) return;
  int lo = nb.bin_start[b], hi = nb.bin_start[b + 1];
  for (int p = lo + 1; p < hi; p++) {
    int v = nb.bin_atoms[p];
    int q = p - 1;
    while (q >= lo && nb.bin_atoms[q] > v) {
      nb.bin_atoms[q + 1] = nb.bin_atoms[q];
      q--;
    }
    nb.bin_atoms[q + 1] = v;
  }
}

// One wavefront per atom, one lane per stencil bin (rounds of 64 bins in
// the stencil order o0, o1, o2): each lane counts the kept candidates of its
// bin, a prefix sum over the lanes gives each bin its output offset, and the
// lanes write their entries there.  The list order is the serial walk's
// (stencil order, then ascending atom index within a bin), with ~2 rounds of
// a few candidates per lane instead of one thread walking ~400 candidates.
__global__ __launch_bounds__(256) void k_nb_list(NbrBuild nb, const double4 *pos, uint32_t *nbr, int *nn)
{
  if (!rebuild_due(nb.ctl, nb.trig2)) return;
  const int lane = threadIdx.x & 63;
  const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (i >= live_atoms(nb)) return;
  const CellInfo &c = cell_of(nb, i);
  const double4 xi = pos[i];
  const int bc0 = nb.atom_bc[3 * i], bc1 = nb.atom_bc[3 * i + 1], bc2 = nb.atom_bc[3 * i + 2];
  const int n0 = c.nb[0], n1 = c.nb[1], n2 = c.nb[2];
  const int w1 = 2 * c.reach[1] + 1, w2 = 2 * c.reach[2] + 1;
  const int nst = (2 * c.reach[0] + 1) * w1 * w2;
  uint32_t *out = nbr + (size_t) i * nb.cap;
  int base = 0;  // entries of the earlier rounds (wave-uniform)
  for (int r0 = 0; r0 < nst; r0 += 64) {
    const int sidx = r0 + lane;
    int lo = 0, hi = 0, s0 = 0, s1 = 0, s2 = 0;
    if (sidx < nst) {
      const int o0 = sidx / (w1 * w2) - c.reach[0];
      const int o1 = (sidx / w2) % w1 - c.reach[1];
      const int o2 = sidx % w2 - c.reach[2];
      const int q0 = bc0 + o0, q1 = bc1 + o1, q2 = bc2 + o2;
      s0 = (q0 >= 0) ? q0 / n0 : -((-q0 + n0 - 1) / n0);
      s1 = (q1 >= 0) ? q1 / n1 : -((-q1 + n1 - 1) / n1);
      s2 = (q2 >= 0) ? q2 / n2 : -((-q2 + n2 - 1) / n2);
      const int t0 = q0 - s0 * n0, t1 = q1 - s1 * n1, t2 = q2 - s2 * n2;
      const int b = c.bin_base + (t0 * n1 + t1) * n2 + t2;
      lo = nb.bin_start[b];
      hi = nb.bin_start[b + 1];
    }
    const bool self_image = s0 == 0 && s1 == 0 && s2 == 0;
    auto kept = [&](int j) {
      if (j == i && self_image) return false;
      double d[3];
      displacement(xi, pos[j], s0, s1, s2, c, d);
      return dist2(d) <= nb.cut2;
    };
    int cnt = 0;
    for (int p = lo; p < hi; p++) cnt += kept(nb.bin_atoms[p]) ? 1 : 0;
    // exclusive prefix sum of the lane counts
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o);
      if (lane >= o) incl += v;
    }
    int k = base + incl - cnt;
    const uint32_t code = (uint32_t) ((s0 + 3) * 49 + (s1 + 3) * 7 + (s2 + 3)) << kNbrJBits;
    for (int p = lo; p < hi; p++) {
      const int j = nb.bin_atoms[p];
      if (kept(j)) {
        if (k < nb.cap) out[k] = (uint32_t) j | code;
        k++;
      }
    }
    base += __shfl(incl, 63);
  }
  if (lane == 0) {
    nn[i] = min(base, nb.cap);
    if (base > nb.cap) atomicMax(&nb.ctl[2], (unsigned long long) base);
  }
}

__global__ void k_nb_done(NbrBuild nb, int nbins)
{
  // runs as one workgroup: clear counts for the next build, reset the control word
  bool due = rebuild_due(nb.ctl, nb.trig2);
  nbins = live_bins(nb, nbins);
  __syncthreads();
  if (due) {
    for (int b = threadIdx.x; b <= nbins; b += blockDim.x) nb.bin_count[b] = 0;
    if (threadIdx.x == 0) nb.ctl[3] += 1ull;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    nb.ctl[0] = 0ull;
    nb.ctl[1] = 0ull;
  }
}

__global__ void k_nb_force(unsigned long long *ctl)
{
  if (threadIdx.x == 0) ctl[1] = 1ull;
}

}  // namespace

hipError_t nbr_force_rebuild(const NbrBuild &nb, hipStream_t stream)
{
  k_nb_force<<<1, 64, 0, stream>>>(nb.ctl);
  return hipGetLastError();
}

hipError_t nbr_step(const NbrBuild &nb, double4 *pos, uint32_t *nbr, int *nn, hipStream_t stream)
{
  if (nb.n <= 0) return hipSuccess;
  const int nbins = nb.nbins;
  int blocks = (nb.n + 255) / 256;
  k_nb_check<<<(nb.n + 1023) / 1024, 1024, 0, stream>>>(nb, pos, nb.x_build, nb.ctl);
  k_nb_bin<<<blocks, 256, 0, stream>>>(nb, pos);
  k_nb_scan<<<1, 1024, 0, stream>>>(nb, nbins);
  k_nb_fill<<<blocks, 256, 0, stream>>>(nb);
  k_nb_sortbins<<<(nbins + 255) / 256, 256, 0, stream>>>(nb, nbins);
  k_nb_list<<<(nb.n + 3) / 4, 256, 0, stream>>>(nb, pos, nbr, nn);
  k_nb_done<<<1, 1024, 0, stream>>>(nb, nbins);
  return hipGetLastError();
}

}  // namespace tmd
