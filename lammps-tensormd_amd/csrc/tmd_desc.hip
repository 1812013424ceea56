// Fused descriptor and force/virial kernels (gfx950, wave64).
//
// One wavefront owns one atom.  Its neighbor list is walked in chunks of 64
// entries: each lane takes one entry and forms the FP64 displacement; the
// kept neighbors (1e-10 < r^2 < rc^2) are sorted by element slot with
// ballots and staged in LDS with their monomials M_d(u) and radial values
// H_k(r) (Kerley interpolation of the table).  The lanes then accumulate
// P[b][k][d] = sum_c H_k(r_c) M_d(u_c) as register tiles of KT x 4 entries.
//
// This replaces, for one atom at a time and without materialising anything
// per neighbor in HBM, the reference chain
//   setup_tensors_device (gpu/cuda/kernel.cu:1122-1416),
//   ration_compute_device (gpu/cuda/ration.cu:13-71),
//   the P GEMM + kernel_1_fused_device (kernel.cu:1098-1120, :1028-1061),
// and, in the force kernel,
//   kernel2_device (kernel.cu:730-817), the V GEMM (:989-1006),
//   kernel_fused_F2_device (:583-646), kernel4_device (:869-965) and
//   sum_forces_device (:1714-1758).
#include "tmd_device.h"

namespace tmd {

namespace {

constexpr int kMS = 24;  // LDS stride of a monomial row (16-byte aligned)

template <int K>
struct Tiling {
  static constexpr int KT = (K <= 32) ? 4 : ((K <= 64) ? 8 : 16);  // k per lane
  static constexpr int LPS = (K / KT) * 5;                           // lanes per slot
  static constexpr int NG = (64 / LPS) > 0 ? (64 / LPS) : 1;         // slot groups per wave
  static constexpr int HS = K + 1;                                   // odd stride for per-lane rows
};

__device__ inline void displacement(const double4 &xi, const double4 &xj, uint32_t code,
                                    const Cell &c, double d[3])
{
#pragma clang fp contract(off)
  int s0 = (int) (code / 49u) - 3, s1 = (int) ((code / 7u) % 7u) - 3, s2 = (int) (code % 7u) - 3;
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

__device__ inline void monomials20(float x, float y, float z, float *M)
{
  float xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
  M[0] = 1.0f;
  M[1] = x;
  M[2] = y;
  M[3] = z;
  M[4] = xx;
  M[5] = xy;
  M[6] = xz;
  M[7] = yy;
  M[8] = yz;
  M[9] = zz;
  M[10] = x * xx;
  M[11] = x * xy;
  M[12] = x * xz;
  M[13] = x * yy;
  M[14] = x * yz;
  M[15] = x * zz;
  M[16] = y * yy;
  M[17] = y * yz;
  M[18] = y * zz;
  M[19] = z * zz;
}

// Kerley rational interpolation of one filter (gpu/cuda/ration.cu:59-68);
// the caller handles the zeroed end cells.
__device__ inline void kerley(float4 t, float q, float p, float dx, float &h, float &dh)
{
  float c1 = t.z, c2 = t.w, s = t.y;
  float c3 = fabsf(c2 * p);
  float c4 = c3 + fabsf(c1 * q);
  if (c4 > 0.0f) c3 = c3 / c4;
  c4 = c2 + c3 * (c1 - c2);
  h = t.x + q * (s - p * c4);
  dh = s + (q - p) * c4 + dx * (c4 - c2) * (1.0f - c3);
}

// Per-chunk staging shared by both kernels.
struct ChunkLDS {
  float *M;       // [64][kMS]
  float *H;       // [64][HS]
  float *dH;      // [64][HS] (force kernel only)
  float4 *rec;    // [64]: ux, uy, uz, r
  float4 *drec;   // [64]: dx, dy, dz (fp32 of the fp64 displacement), j as bits
  int *tcell;     // [64]
  float2 *qp;     // [64]
  int *slot;      // [64]
  int *sstart;    // [kMaxElts + 1]
};

// Geometry + slot sort of one raw chunk.  Returns the number of kept
// neighbors staged; sstart[b] .. sstart[b+1] is slot b's range.
template <bool FORCE>
__device__ inline int stage_chunk(const ModelGeo &g, const Cell &cell, const double4 &xi, int ei,
                                  const uint32_t *list, int base, int nraw, const double4 *pos,
                                  ChunkLDS &L)
{
  const int lane = threadIdx.x & 63;
  int idx = base + lane;
  bool keep = false;
  int slot = -1, j = 0, tcell = 0;
  double d[3] = {0, 0, 0}, r = 0.0, rinv = 0.0;
  if (idx < nraw) {
    uint32_t e = list[idx];
    j = (int) (e & kNbrJMask);
    double4 xj = pos[j];
    displacement(xi, xj, e >> kNbrJBits, cell, d);
    double r2 = dist2(d);
    if (r2 < g.rc2 && r2 > 1e-10) {  // kernel.cu:1163
      keep = true;
      slot = g.eltij[ei][(int) xj.w];
      r = sqrt(r2);
      rinv = 1.0 / r;
    }
  }
  // stable partition by slot: position = start(slot) + rank among equal slots
  int start = 0, mypos = -1;
  for (int b = 0; b < g.nelt; b++) {
    unsigned long long m = __ballot(slot == b);
    if (lane == 0) L.sstart[b] = start;
    if (slot == b) mypos = start + (int) __builtin_amdgcn_mbcnt_hi((unsigned) (m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned) m, 0u));
    start += __popcll(m);
  }
  if (lane == 0) L.sstart[g.nelt] = start;
  if (keep) {
    float ux = (float) (d[0] * rinv), uy = (float) (d[1] * rinv), uz = (float) (d[2] * rinv);
    float M[kND];
    monomials20(ux, uy, uz, M);
    float *Mr = L.M + mypos * kMS;
#pragma unroll
    for (int q = 0; q < kND; q++) Mr[q] = M[q];
    L.rec[mypos] = make_float4(ux, uy, uz, (float) r);
    // table cell and offsets in FP64 (ration.cu:29-32)
    int n = (int) (r * g.xdx);
    n = min(g.tab_n - 2, n);
    double q = r - (double) n * g.dx;
    double p = g.dx - q;
    L.tcell[mypos] = n;
    L.qp[mypos] = make_float2((float) q, (float) p);
    L.slot[mypos] = slot;
    if (FORCE) L.drec[mypos] = make_float4((float) d[0], (float) d[1], (float) d[2], __int_as_float(j));
  }
  return start;
}

// H (and dH) for the staged neighbors: lanes over (neighbor, filter)
template <int K, bool FORCE>
__device__ inline void stage_radial(const ModelGeo &g, const float4 *tab, int nk, ChunkLDS &L)
{
  const int lane = threadIdx.x & 63;
  constexpr int HS = FORCE ? Tiling<K>::HS : K;
  for (int e = lane; e < nk * K; e += 64) {
    int c = e / K;
    int kk = e - c * K;
    int n = L.tcell[c];
    float h = 0.0f, dh = 0.0f;
    if (n > 0 && n < g.tab_n - 2) {  // endpoint cells are zero (ration.hpp:29)
      float2 qp = L.qp[c];
      kerley(tab[(size_t) n * K + kk], qp.x, qp.y, g.dxf, h, dh);
    }
    L.H[c * HS + kk] = h;
    if (FORCE) L.dH[c * HS + kk] = dh;
  }
}

// P[b][k][d] += sum over the chunk's slot-b neighbors of H_k M_d.
template <int K, bool FORCE>
__device__ inline void accumulate_P(const ModelGeo &g, ChunkLDS &L, float *sP)
{
  using T = Tiling<K>;
  constexpr int HS = FORCE ? T::HS : K;
  const int lane = threadIdx.x & 63;
  const int grp = lane / T::LPS;
  const int w = lane - grp * T::LPS;
  const int kq = w / 5, dq = w - (w / 5) * 5;
  if (grp >= T::NG) return;
  for (int b = grp; b < g.nelt; b += T::NG) {
    int c0 = L.sstart[b], c1 = L.sstart[b + 1];
    if (c0 == c1) continue;
    float acc[T::KT][4];
#pragma unroll
    for (int a = 0; a < T::KT; a++)
#pragma unroll
      for (int q = 0; q < 4; q++) acc[a][q] = 0.0f;
    for (int c = c0; c < c1; c++) {
      float4 m = *reinterpret_cast<const float4 *>(L.M + c * kMS + dq * 4);
      const float *hr = L.H + c * HS + kq * T::KT;
#pragma unroll
      for (int a = 0; a < T::KT; a++) {
        float h = hr[a];
        acc[a][0] += h * m.x;
        acc[a][1] += h * m.y;
        acc[a][2] += h * m.z;
        acc[a][3] += h * m.w;
      }
    }
    float *P = sP + ((size_t) b * K + kq * T::KT) * kND + dq * 4;
#pragma unroll
    for (int a = 0; a < T::KT; a++)
#pragma unroll
      for (int q = 0; q < 4; q++) P[a * kND + q] += acc[a][q];
  }
}

template <int K, bool FORCE>
__device__ inline void build_P(const ModelGeo &g, const Cell &cell, const double4 &xi, int ei,
                               const uint32_t *list, int nraw, const double4 *pos,
                               const float4 *tab, ChunkLDS &L, float *sP)
{
  const int lane = threadIdx.x & 63;
  for (int e = lane; e < g.nelt * K * kND; e += 64) sP[e] = 0.0f;
  for (int base = 0; base < nraw; base += 64) {
    __syncthreads();
    int nk = stage_chunk<false>(g, cell, xi, ei, list, base, nraw, pos, L);
    __syncthreads();
    stage_radial<K, FORCE>(g, tab, nk, L);
    __syncthreads();
    accumulate_P<K, FORCE>(g, L, sP);
  }
  __syncthreads();
}

template <int K>
size_t lds_bytes(int nelt, bool force)
{
  constexpr int HS = Tiling<K>::HS;
  size_t b = 0;
  b += (size_t) nelt * K * kND * 4;  // P / S
  b += 64 * kMS * 4;                 // M
  b += (size_t) 64 * (force ? HS : K) * 4 * (force ? 2 : 1);
  b += 64 * 16 * (force ? 2 : 1);    // rec (+ drec)
  b += 64 * 4 + 64 * 8 + 64 * 4 + (kMaxElts + 1) * 4;
  return b + 64;
}

__device__ inline ChunkLDS carve(char *base, int nelt, int K, int HS, bool force, float **sP)
{
  ChunkLDS L;
  char *p = base;
  *sP = (float *) p;
  p += (size_t) nelt * K * kND * 4;
  L.M = (float *) p;
  p += 64 * kMS * 4;
  L.H = (float *) p;
  p += (size_t) 64 * HS * 4;
  if (force) {
    L.dH = (float *) p;
    p += (size_t) 64 * HS * 4;
  } else {
    L.dH = nullptr;
  }
  L.rec = (float4 *) p;
  p += 64 * 16;
  if (force) {
    L.drec = (float4 *) p;
    p += 64 * 16;
  } else {
    L.drec = nullptr;
  }
  L.tcell = (int *) p;
  p += 64 * 4;
  L.qp = (float2 *) p;
  p += 64 * 8;
  L.slot = (int *) p;
  p += 64 * 4;
  L.sstart = (int *) p;
  return L;
}

template <int K>
__global__ __launch_bounds__(64) void k_descriptor(ModelGeo g, Cell cell, const double4 *pos,
                                                   const uint32_t *nbr, const int *nn, int cap,
                                                   const int *row2atom, int nrows,
                                                   const float4 *tab, float *G)
{
  extern __shared__ __align__(16) char smem[];
  const int row = blockIdx.x;
  if (row >= nrows) return;
  const int lane = threadIdx.x & 63;
  const int gw = g.nelt * K * kNM;
  float *Grow = G + (size_t) row * gw;
  const int i = row2atom[row];
  if (i < 0) {  // padding row: zero descriptor
    for (int e = lane; e < gw; e += 64) Grow[e] = 0.0f;
    return;
  }
  float *sP;
  ChunkLDS L = carve(smem, g.nelt, K, K, false, &sP);
  const double4 xi = pos[i];
  const int ei = (int) xi.w;
  build_P<K, false>(g, cell, xi, ei, nbr + (size_t) i * cap, nn[i], pos, tab, L, sP);
  // G[b][k][m] = sum_d T[m][d] P^2 (m >= 1); G[b][k][0] = P0 (= sign(P0) sqrt(P0^2))
  for (int e = lane; e < g.nelt * K; e += 64) {
    const float *P = sP + (size_t) e * kND;
    float Gv[kNM] = {0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < kND; d++) {
      float p2 = P[d] * P[d];
#pragma unroll
      for (int m = 1; m < kNM; m++) Gv[m] += g.T[m][d] * p2;
    }
    Gv[0] = P[0];
    *reinterpret_cast<float4 *>(Grow + (size_t) e * kNM) = make_float4(Gv[0], Gv[1], Gv[2], Gv[3]);
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_force(ModelGeo g, Cell cell, const double4 *pos,
                                              const uint32_t *nbr, const int *nn, int cap,
                                              const int *row2atom, int nrows, const float4 *tab,
                                              const float *dEdG, double scale, double *f,
                                              double *wrow, double *vatom)
{
  using T = Tiling<K>;
  constexpr int HS = T::HS;
  extern __shared__ __align__(16) char smem[];
  const int row = blockIdx.x;
  if (row >= nrows) return;
  const int lane = threadIdx.x & 63;
  const int i = row2atom[row];
  if (i < 0) {
    if (lane < 6) wrow[(size_t) row * 6 + lane] = 0.0;
    return;
  }
  float *sP;
  ChunkLDS L = carve(smem, g.nelt, K, HS, true, &sP);
  const double4 xi = pos[i];
  const int ei = (int) xi.w;
  const uint32_t *list = nbr + (size_t) i * cap;
  const int nraw = nn[i];
  build_P<K, true>(g, cell, xi, ei, list, nraw, pos, tab, L, sP);

  // S[b][k][d] = P[b][k][d] * sum_m 2 T[m][d] g'_m  (kernel.cu:752-767), in place
  const float *grow = dEdG + (size_t) row * g.nelt * K * kNM;
  for (int e = lane; e < g.nelt * K; e += 64) {
    float *P = sP + (size_t) e * kND;
    float4 gv = *reinterpret_cast<const float4 *>(grow + (size_t) e * kNM);
    float G0 = P[0];
    float g0 = (fabsf(G0) < 1e-8f) ? 0.0f : gv.x / G0 * 0.5f;
    float gm[kNM] = {g0, gv.y, gv.z, gv.w};
#pragma unroll
    for (int d = 0; d < kND; d++) {
      float s = 0.0f;
#pragma unroll
      for (int m = 0; m < kNM; m++) s += gm[m] * 2.0f * g.T[m][d];
      P[d] = s * P[d];
    }
  }
  __syncthreads();

  double fi[3] = {0, 0, 0};
  double wi[6] = {0, 0, 0, 0, 0, 0};
  const float fscale = 1.0f;
  for (int base = 0; base < nraw; base += 64) {
    __syncthreads();
    int nk = stage_chunk<true>(g, cell, xi, ei, list, base, nraw, pos, L);
    __syncthreads();
    stage_radial<K, true>(g, tab, nk, L);
    __syncthreads();
    if (lane < nk) {
      const int c = lane;
      const int b = L.slot[c];
      float M[kND];
      const float *Mr = L.M + c * kMS;
#pragma unroll
      for (int q = 0; q < kND; q++) M[q] = Mr[q];
      float V[kND];
#pragma unroll
      for (int q = 0; q < kND; q++) V[q] = 0.0f;
      float f1 = 0.0f;
      const float *S = sP + (size_t) b * K * kND;
      const float *Hr = L.H + c * HS;
      const float *dHr = L.dH + c * HS;
      for (int kk = 0; kk < K; kk++) {
        float h = Hr[kk], dh = dHr[kk];
        const float4 *Sk = reinterpret_cast<const float4 *>(S + kk * kND);
        float U = 0.0f;
#pragma unroll
        for (int q4 = 0; q4 < kND / 4; q4++) {
          float4 s = Sk[q4];
          U += s.x * M[4 * q4] + s.y * M[4 * q4 + 1] + s.z * M[4 * q4 + 2] + s.w * M[4 * q4 + 3];
          V[4 * q4] += s.x * h;
          V[4 * q4 + 1] += s.y * h;
          V[4 * q4 + 2] += s.z * h;
          V[4 * q4 + 3] += s.w * h;
        }
        f1 += U * dh;
      }
      float4 rc = L.rec[c];
      float x = rc.x, y = rc.y, z = rc.z;
      // g_beta = sum_d V_d dM_d/du_beta for the 20 monomials
      float gx = V[1] + 2.f * x * V[4] + y * V[5] + z * V[6] + 3.f * x * x * V[10] +
          2.f * x * y * V[11] + 2.f * x * z * V[12] + y * y * V[13] + y * z * V[14] + z * z * V[15];
      float gy = V[2] + x * V[5] + 2.f * y * V[7] + z * V[8] + x * x * V[11] + 2.f * x * y * V[13] +
          x * z * V[14] + 3.f * y * y * V[16] + 2.f * y * z * V[17] + z * z * V[18];
      float gz = V[3] + x * V[6] + y * V[8] + 2.f * z * V[9] + x * x * V[12] + x * y * V[14] +
          2.f * x * z * V[15] + y * y * V[17] + 2.f * y * z * V[18] + 3.f * z * z * V[19];
      float gu = gx * x + gy * y + gz * z;
      float rinv = 1.0f / rc.w;
      // dE_i/dDelta = F1 + F2 (kernel.cu:559-571 and :583-646)
      double F0 = (double) (f1 * x + (gx - gu * x) * rinv) * scale;
      double F1v = (double) (f1 * y + (gy - gu * y) * rinv) * scale;
      double F2v = (double) (f1 * z + (gz - gu * z) * rinv) * scale;
      (void) fscale;
      float4 dr = L.drec[c];
      int j = __float_as_int(dr.w);
      // f_i += F, f_j -= F (kernel.cu:1726-1730)
      fi[0] += F0;
      fi[1] += F1v;
      fi[2] += F2v;
      atomicAdd(&f[3 * (size_t) j + 0], -F0);
      atomicAdd(&f[3 * (size_t) j + 1], -F1v);
      atomicAdd(&f[3 * (size_t) j + 2], -F2v);
      double dx = dr.x, dy = dr.y, dz = dr.z;
      double v[6] = {-dx * F0, -dy * F1v, -dz * F2v, -dx * F1v, -dx * F2v, -dy * F2v};
#pragma unroll
      for (int q = 0; q < 6; q++) wi[q] += v[q];
      if (vatom) {
#pragma unroll
        for (int q = 0; q < 6; q++) atomicAdd(&vatom[6 * (size_t) j + q], 0.5 * v[q]);
      }
    }
  }
  // wave reductions
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; a++) fi[a] += __shfl_xor(fi[a], o);
#pragma unroll
    for (int q = 0; q < 6; q++) wi[q] += __shfl_xor(wi[q], o);
  }
  if (lane == 0) {
    atomicAdd(&f[3 * (size_t) i + 0], fi[0]);
    atomicAdd(&f[3 * (size_t) i + 1], fi[1]);
    atomicAdd(&f[3 * (size_t) i + 2], fi[2]);
  }
  if (lane < 6) {
    double v = wi[lane];
    wrow[(size_t) row * 6 + lane] = v;
    if (vatom) atomicAdd(&vatom[6 * (size_t) i + lane], 0.5 * v);
  }
}

template <int K>
hipError_t desc_impl(const ModelGeo &g, const Cell &cell, const double4 *pos, const uint32_t *nbr,
                     const int *nn, int cap, const int *row2atom, int nrows, const float4 *tab,
                     float *G, hipStream_t s)
{
  size_t lds = lds_bytes<K>(g.nelt, false);
  if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
  if (lds > 64 * 1024)
    hipFuncSetAttribute((const void *) k_descriptor<K>, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
  k_descriptor<K><<<nrows, 64, lds, s>>>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, G);
  return hipGetLastError();
}

template <int K>
hipError_t force_impl(const ModelGeo &g, const Cell &cell, const double4 *pos, const uint32_t *nbr,
                      const int *nn, int cap, const int *row2atom, int nrows, const float4 *tab,
                      const float *dEdG, double scale, double *f, double *wrow, double *vatom,
                      hipStream_t s)
{
  size_t lds = lds_bytes<K>(g.nelt, true);
  if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
  if (lds > 64 * 1024)
    hipFuncSetAttribute((const void *) k_force<K>, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
  k_force<K><<<nrows, 64, lds, s>>>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, dEdG, scale,
                                    f, wrow, vatom);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_descriptor(const ModelGeo &g, const Cell &cell, const double4 *pos,
                             const uint32_t *nbr, const int *nn, int cap, const int *row2atom,
                             int nrows, const float4 *tab, float *G, hipStream_t stream)
{
  if (nrows <= 0) return hipSuccess;
  switch (g.k) {
    case 16: return desc_impl<16>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, G, stream);
    case 24: return desc_impl<24>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, G, stream);
    case 32: return desc_impl<32>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, G, stream);
    case 64: return desc_impl<64>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, G, stream);
    case 128: return desc_impl<128>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, G, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_force(const ModelGeo &g, const Cell &cell, const double4 *pos,
                        const uint32_t *nbr, const int *nn, int cap, const int *row2atom,
                        int nrows, const float4 *tab, const float *dEdG, double scale, double *f,
                        double *wrow, double *vatom, hipStream_t stream)
{
  if (nrows <= 0) return hipSuccess;
  switch (g.k) {
    case 16: return force_impl<16>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, dEdG, scale, f, wrow, vatom, stream);
    case 24: return force_impl<24>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, dEdG, scale, f, wrow, vatom, stream);
    case 32: return force_impl<32>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, dEdG, scale, f, wrow, vatom, stream);
    case 64: return force_impl<64>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, dEdG, scale, f, wrow, vatom, stream);
    case 128: return force_impl<128>(g, cell, pos, nbr, nn, cap, row2atom, nrows, tab, dEdG, scale, f, wrow, vatom, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace tmd
