// Small device kernels: buffer clearing (a kernel, so it is safe inside
// captured graphs), fixed-order FP64 reductions of energy/virial and kinetic
// energy, and the two velocity-Verlet half steps (src/fix_nve.cpp:64-130).
#include "tmd_device.h"

#include <algorithm>

namespace tmd {

namespace {

constexpr int kRedParts = 240;

__global__ void k_zero(uint4 *p, size_t n16, unsigned char *tail, size_t ntail)
{
  size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for (size_t q = i; q < n16; q += stride) p[q] = make_uint4(0, 0, 0, 0);
  if (i < ntail) tail[i] = 0;
}

// Each block writes one partial; the final kernel adds partials in order,
// so the result does not depend on scheduling.
// sums of eatom[i], watom[i][0..5] over atoms [lo, hi) by one 256-thread
// workgroup, in a fixed order; thread 0..6 of the result hold the 7 sums
__device__ inline void block_sum7(const double *eatom, const double *watom, int lo, int hi,
                                  int stride_lo, int stride, double (*sh)[256], double *res)
{
  int t = threadIdx.x;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = lo + stride_lo + t; i < hi; i += stride) {
    v[0] += eatom[i];
    for (int q = 0; q < 6; q++) v[1 + q] += watom[(size_t) i * 6 + q];
  }
  for (int q = 0; q < 7; q++) sh[q][t] = v[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int q = 0; q < 7; q++) sh[q][t] += sh[q][t + o];
    __syncthreads();
  }
  if (t < 7) res[t] = sh[t][0];
}

__global__ void k_reduce_ev_part(const double *eatom, const double *watom, int n, double *part)
{
  __shared__ double sh[7][256];
  block_sum7(eatom, watom, 0, n, blockIdx.x * 256, gridDim.x * 256, sh,
             part + (size_t) blockIdx.x * 7);
}

// one wave per quantity; fixed shuffle tree (deterministic)
__global__ void k_reduce_ev_final(const double *part, int nparts, double *out)
{
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (q >= 7) return;
  double s = 0.0;
  for (int p = lane; p < nparts; p += 64) s += part[(size_t) p * 7 + q];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[q] = s;
}

// one workgroup per structure
__global__ void k_reduce_struct(const double *eatom, const double *watom, const int *off,
                                int nstruct, double *out, const BatchDims *dyn)
{
  __shared__ double sh[7][256];
  const int st = blockIdx.x;
  if (st >= nstruct || (dyn && st >= dyn->nstruct)) return;
  block_sum7(eatom, watom, off[st], off[st + 1], 0, 256, sh, out + (size_t) st * 7);
}

__global__ void k_vv_first(int n, double4 *pos, double *v, const double *f, const double *dtfm,
                           double dt)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 p = pos[i];
  double c = dtfm[(int) p.w];
  double vx = v[3 * i] + c * f[3 * i];
  double vy = v[3 * i + 1] + c * f[3 * i + 1];
  double vz = v[3 * i + 2] + c * f[3 * i + 2];
  v[3 * i] = vx;
  v[3 * i + 1] = vy;
  v[3 * i + 2] = vz;
  p.x += dt * vx;
  p.y += dt * vy;
  p.z += dt * vz;
  pos[i] = p;
}

__global__ void k_vv_second(int n, const double4 *pos, double *v, const double *f, const double *dtfm)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double c = dtfm[(int) pos[i].w];
  v[3 * i] += c * f[3 * i];
  v[3 * i + 1] += c * f[3 * i + 1];
  v[3 * i + 2] += c * f[3 * i + 2];
}

__global__ void k_ekin_part(int n, const double4 *pos, const double *v, const double *mass,
                            double *part)
{
  __shared__ double sh[256];
  int t = threadIdx.x;
  double s = 0.0;
  for (int i = blockIdx.x * 256 + t; i < n; i += gridDim.x * 256) {
    double m = mass[(int) pos[i].w];
    s += m * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  }
  sh[t] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_ekin_final(const double *part, int nparts, double *out)
{
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int p = lane; p < nparts; p += 64) s += part[p];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (threadIdx.x == 0) out[0] = 0.5 * 1.0364269e-4 * s;  // mvv2e, metal units
}

__global__ void k_zero_atoms(double *p, int per_atom, const BatchDims *dyn)
{
  size_t n = (size_t) per_atom * dyn->natoms;
  size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_copy_scalar(const double *src, double *dst)
{
  if (threadIdx.x == 0) dst[0] = src[0];
}

// hist[2*n] = epot, hist[2*n+1] = ekin, n = counter++ (graph-replay safe)
__global__ void k_record_step(const double *epot, const double *ekin, double *hist,
                              unsigned long long *counter, int cap)
{
  if (threadIdx.x == 0) {
    unsigned long long n = counter[0];
    if (n < (unsigned long long) cap) {
      hist[2 * n] = epot[0];
      hist[2 * n + 1] = ekin ? ekin[0] : 0.0;
    }
    counter[0] = n + 1;
  }
}

__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
  size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for (size_t q = i; q < n16; q += stride) dst[q] = src[q];
}

}  // namespace

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t s)
{
  size_t n16 = bytes / 16;
  k_copy<<<8192, 256, 0, s>>>((const uint4 *) src, (uint4 *) dst, n16);
  return hipGetLastError();
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t s)
{
  if (bytes == 0) return hipSuccess;
  size_t n16 = bytes / 16, tail = bytes - n16 * 16;
  int blocks = (int) std::min<size_t>(4096, (n16 + 255) / 256 + 1);
  k_zero<<<blocks, 256, 0, s>>>((uint4 *) p, n16, (unsigned char *) p + n16 * 16, tail);
  return hipGetLastError();
}

hipError_t launch_reduce_ev(const double *eatom, const double *watom, int n, double *out,
                            hipStream_t s)
{
  double *part = out + 8;  // out holds 8 + kRedParts*7 doubles
  k_reduce_ev_part<<<kRedParts, 256, 0, s>>>(eatom, watom, n, part);
  k_reduce_ev_final<<<1, 448, 0, s>>>(part, kRedParts, out);
  return hipGetLastError();
}

hipError_t launch_reduce_struct(const double *eatom, const double *watom, const int *off,
                                int nstruct, double *out, hipStream_t s, const BatchDims *dyn)
{
  if (nstruct <= 0) return hipSuccess;
  k_reduce_struct<<<nstruct, 256, 0, s>>>(eatom, watom, off, nstruct, out, dyn);
  return hipGetLastError();
}

hipError_t launch_zero_atoms(double *p, int per_atom, int cap_atoms, const BatchDims *dyn,
                             hipStream_t s)
{
  size_t n = (size_t) per_atom * cap_atoms;
  if (n == 0) return hipSuccess;
  k_zero_atoms<<<(unsigned) ((n + 255) / 256), 256, 0, s>>>(p, per_atom, dyn);
  return hipGetLastError();
}

hipError_t launch_vv_first(int n, double4 *pos, double *v, const double *f, const double *dtfm,
                           double dt, hipStream_t s)
{
  k_vv_first<<<(n + 255) / 256, 256, 0, s>>>(n, pos, v, f, dtfm, dt);
  return hipGetLastError();
}

hipError_t launch_vv_second(int n, const double4 *pos, double *v, const double *f,
                            const double *dtfm, hipStream_t s)
{
  k_vv_second<<<(n + 255) / 256, 256, 0, s>>>(n, pos, v, f, dtfm);
  return hipGetLastError();
}

hipError_t launch_ekin(int n, const double4 *pos, const double *v, const double *mass, double *out,
                       hipStream_t s)
{
  double *part = out + 8;
  k_ekin_part<<<kRedParts, 256, 0, s>>>(n, pos, v, mass, part);
  k_ekin_final<<<1, 64, 0, s>>>(part, kRedParts, out);
  return hipGetLastError();
}

// Sum of the kept-neighbor counts over the real rows (the first word of
// each row of the descriptor's kept-list cache), added to *acc: one atomic
// per wave.
__global__ void k_sum_kept(const uint32_t *kl, size_t words, const int *row2atom, int nrows,
                           unsigned long long *acc)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (r < nrows && row2atom[r] >= 0) v = kl[(size_t) r * words];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(acc, v);
}

hipError_t launch_sum_kept(const uint32_t *kl, size_t words, const int *row2atom, int nrows,
                           unsigned long long *acc, hipStream_t s)
{
  if (nrows <= 0) return hipSuccess;
  k_sum_kept<<<(nrows + 255) / 256, 256, 0, s>>>(kl, words, row2atom, nrows, acc);
  return hipGetLastError();
}

hipError_t launch_record_step(const double *epot, const double *ekin, double *hist,
                              unsigned long long *counter, int cap, hipStream_t s)
{
  k_record_step<<<1, 64, 0, s>>>(epot, ekin, hist, counter, cap);
  return hipGetLastError();
}

hipError_t launch_copy_scalar(const double *src, double *dst, hipStream_t s)
{
  k_copy_scalar<<<1, 64, 0, s>>>(src, dst);
  return hipGetLastError();
}

}  // namespace tmd
