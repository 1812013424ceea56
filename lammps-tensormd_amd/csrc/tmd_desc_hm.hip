// Descriptor and force kernels for max_moment 4 and 5 models (35 and 56
// monomials; the reference's setup_tensors_device and calculate_dM_g/_h,
// gpu/cuda/kernel.cu:1233-1406 and :553-556, and the T rows of
// src/ML-TENSORMD/tensormd.cpp:234-273).
//
// The same algorithm as tmd_desc.hip (one wavefront per atom, the kept list
// compacted and sorted by slot, P = sum_c H M per slot, G = sum_d T P^2,
// S = P (2 sum_m T g'), per-neighbor chain rule), written for any moment
// count instead of register-tiled for 20 monomials:
//  * slots are processed one at a time; P of the slot is accumulated in
//    blocks of (k, d) pairs, 64 x PB per block (PB accumulators per lane),
//    from 32-neighbor chunks staged in LDS (monomials and H), and stored
//    row-major [k][d] in LDS;
//  * G and S are formed from the LDS copy of P;
//  * the chain rule runs one lane per neighbor of the slot, with the
//    monomials, V = S^T H and the gradient of the monomials in registers.
// These models are not shipped with the reference; the kernels favour
// generality over the register tiling of the max_moment-3 kernels.
#include "tmd_desc_common.h"

namespace tmd {

namespace {

// Monomials u_x^a u_y^b u_z^c of degree 0..MM in the reference order
// (degree, then x-power and y-power descending) with their degree and
// multinomial multiplicity deg! / (a! b! c!) = T[deg][d].
template <int MM>
struct Moments {
  static constexpr int ND = (MM + 1) * (MM + 2) * (MM + 3) / 6;
  static constexpr int NM = MM + 1;
  struct Table {
    int a[ND], b[ND], c[ND], deg[ND];
    float mult[ND];
  };
  static constexpr Table make()
  {
    Table t{};
    int d = 0;
    for (int n = 0; n <= MM; n++)
      for (int a = n; a >= 0; a--)
        for (int b = n - a; b >= 0; b--) {
          const int c = n - a - b;
          long fn = 1, fa = 1, fb = 1, fc = 1;
          for (int q = 2; q <= n; q++) fn *= q;
          for (int q = 2; q <= a; q++) fa *= q;
          for (int q = 2; q <= b; q++) fb *= q;
          for (int q = 2; q <= c; q++) fc *= q;
          t.a[d] = a;
          t.b[d] = b;
          t.c[d] = c;
          t.deg[d] = n;
          t.mult[d] = (float) (fn / (fa * fb * fc));
          d++;
        }
    return t;
  }
  static constexpr Table tab = make();
  // first monomial of degree m
  static constexpr int first(int m) { return m * (m + 1) * (m + 2) / 6; }
};

constexpr int kHmChunk = 32;  // neighbors staged at a time for P

template <int K, int MM>
struct HmLayout {
  static constexpr int ND = Moments<MM>::ND;
  static constexpr int PAIRS = K * ND;
  static constexpr int PB = (PAIRS + 63) / 64 < 28 ? (PAIRS + 63) / 64 : 28;  // accumulators per lane
  static constexpr int BLK = 64 * PB;
  static constexpr int NBLK = (PAIRS + BLK - 1) / BLK;
  // LDS: kept[cap] | M[chunk][ND] | H[chunk][K] | P or S [K][ND]
  __host__ __device__ static size_t bytes(int cap, size_t *off)
  {
    auto up = [](size_t v) { return (v + 15) & ~(size_t) 15; };
    off[0] = 0;
    off[1] = up((size_t) cap * 4);
    off[2] = off[1] + up((size_t) kHmChunk * ND * 4);
    off[3] = off[2] + up((size_t) kHmChunk * K * 4);
    return off[3] + up((size_t) PAIRS * 4);
  }
};

template <int MM>
__device__ inline void monomials_hm(float x, float y, float z, float *M)
{
  using Mo = Moments<MM>;
  float px[MM + 1], py[MM + 1], pz[MM + 1];
  px[0] = py[0] = pz[0] = 1.0f;
#pragma unroll
  for (int e = 1; e <= MM; e++) {
    px[e] = px[e - 1] * x;
    py[e] = py[e - 1] * y;
    pz[e] = pz[e - 1] * z;
  }
#pragma unroll
  for (int d = 0; d < Mo::ND; d++) M[d] = px[Mo::tab.a[d]] * py[Mo::tab.b[d]] * pz[Mo::tab.c[d]];
}

// Stage kept neighbors [base, min(base + 32, end)) of the slot: two lanes per
// neighbor, both form the geometry; lane 0 of the pair writes the monomials,
// each writes H of half of the filters (zero outside the live table cells).
template <int K, int MM>
__device__ inline void stage_hm(const ModelGeo &g, const Box &cell, const double4 &xi, const double4 *pos,
                                const float4 *tab, const uint32_t *kept, int base, int end, float *Ms,
                                float *Hs)
{
  constexpr int ND = Moments<MM>::ND;
  const int lane = threadIdx.x & 63;
  const int c = lane >> 1, part = lane & 1;
  if (base + c < end) {
    NbrGeo ng;
    nbr_geometry(g, cell, xi, pos, kept[base + c], ng);
    if (part == 0) {
      float M[ND];
      monomials_hm<MM>(ng.ux, ng.uy, ng.uz, M);
#pragma unroll
      for (int d = 0; d < ND; d++) Ms[c * ND + d] = M[d];
    }
    float *Hc = Hs + c * K + part * (K / 2);
    if (table_cell_live(g, ng.n)) {
      const float4 *row = tab + (size_t) ng.n * K + part * (K / 2);
      for (int k = 0; k < K / 2; k++) Hc[k] = kerley_h(row[k], ng.q, ng.pq);
    } else {
      for (int k = 0; k < K / 2; k++) Hc[k] = 0.0f;
    }
  }
  __syncthreads();
}

// P[k][d] of one slot (kept entries [s0, s1)) into LDS Pl, row-major.
template <int K, int MM>
__device__ inline void build_P_hm(const ModelGeo &g, const Box &cell, const double4 &xi, const double4 *pos,
                                  const float4 *tab, const uint32_t *kept, int s0, int s1, float *Ms,
                                  float *Hs, float *Pl)
{
  using Lo = HmLayout<K, MM>;
  constexpr int ND = Lo::ND;
  const int lane = threadIdx.x & 63;
  for (int blk = 0; blk < Lo::NBLK; blk++) {
    int hk[Lo::PB], md[Lo::PB];
    float acc[Lo::PB];
#pragma unroll
    for (int j = 0; j < Lo::PB; j++) {
      const int p = min(blk * Lo::BLK + j * 64 + lane, Lo::PAIRS - 1);
      hk[j] = p / ND;
      md[j] = p - (p / ND) * ND;
      acc[j] = 0.0f;
    }
    for (int base = s0; base < s1; base += kHmChunk) {
      stage_hm<K, MM>(g, cell, xi, pos, tab, kept, base, s1, Ms, Hs);
      const int nc = min(kHmChunk, s1 - base);
      for (int c = 0; c < nc; c++) {
#pragma unroll
        for (int j = 0; j < Lo::PB; j++) acc[j] = fmaf(Hs[c * K + hk[j]], Ms[c * ND + md[j]], acc[j]);
      }
      __syncthreads();  // the next chunk overwrites the staging area
    }
#pragma unroll
    for (int j = 0; j < Lo::PB; j++) {
      const int p = blk * Lo::BLK + j * 64 + lane;
      if (p < Lo::PAIRS) Pl[p] = acc[j];
    }
  }
  __syncthreads();
}

template <int K, int MM>
__global__ __launch_bounds__(64) void k_descriptor_hm(ModelGeo g, Geom geo, const double4 *pos,
                                                      const uint32_t *nbr, const int *nn, int cap,
                                                      const int *row2atom, int rbase, int nrows,
                                                      const float4 *tab, float *G, uint32_t *kl)
{
  using Mo = Moments<MM>;
  using Lo = HmLayout<K, MM>;
  constexpr int ND = Mo::ND, NM = Mo::NM;
  extern __shared__ __align__(16) unsigned char smem[];
  const int row = geo.order ? geo.order[rbase + (int) blockIdx.x] : rbase + (int) blockIdx.x;
  if (row >= nrows || (geo.dyn && row >= geo.dyn->rl.nrows)) return;
  const int lane = threadIdx.x & 63;
  const int gw = g.nelt * K * NM;
  float *Grow = G + (size_t) row * gw;
  const int i = row2atom[row];
  if (i < 0) {  // padding row
    for (int e = lane; e < gw; e += 64) Grow[e] = 0.0f;
    return;
  }
  size_t off[4];
  Lo::bytes(cap, off);
  uint32_t *kept = reinterpret_cast<uint32_t *>(smem + off[0]);
  float *Ms = reinterpret_cast<float *>(smem + off[1]);
  float *Hs = reinterpret_cast<float *>(smem + off[2]);
  float *Pl = reinterpret_cast<float *>(smem + off[3]);
  int start[kMaxElts + 1];
  const double4 xi = pos[i];
  const Box cell = load_box(geo, i);
  const int nk = compact(g, cell, xi, (int) xi.w, nbr + (size_t) i * cap, nn[i], pos, kept, start);
  __syncthreads();
  {  // the row's kept list with its slot starts, for the force kernel
    uint32_t *kr = kl + (size_t) row * kept_cache_words(cap);
    if (lane == 0) {
      kr[0] = (uint32_t) nk;
      for (int b = 0; b <= g.nelt; b++) kr[1 + b] = (uint32_t) start[b];
    }
    for (int e = lane; e < nk; e += 64) kr[kKeptHdr + e] = kept[e];
  }
  for (int b = 0; b < g.nelt; b++) {
    build_P_hm<K, MM>(g, cell, xi, pos, tab, kept, start[b], start[b + 1], Ms, Hs, Pl);
    // G[b][k][m] = sum over monomials d of degree m of T[m][d] P[k][d]^2,
    // G[b][k][0] = P[k][0]
    for (int e = lane; e < K * NM; e += 64) {
      const int k = e / NM, m = e - (e / NM) * NM;
      const float *Pk = Pl + k * ND;
      float v;
      if (m == 0) {
        v = Pk[0];
      } else {
        v = 0.0f;
        for (int d = Mo::first(m); d < Mo::first(m + 1); d++) v = fmaf(Mo::tab.mult[d] * Pk[d], Pk[d], v);
      }
      Grow[((size_t) b * K + k) * NM + m] = v;
    }
    __syncthreads();  // Pl is rebuilt for the next slot
  }
}

template <int K, int MM>
__global__ __launch_bounds__(64) void k_force_hm(ModelGeo g, Geom geo, const double4 *pos, int cap,
                                                 const int *row2atom, int rbase, int nrows, const float4 *tab,
                                                 const uint32_t *kl, const float *dEdG, const double *E,
                                                 double scale, double *f, double *eatom, double *watom,
                                                 double *vatom)
{
  using Mo = Moments<MM>;
  using Lo = HmLayout<K, MM>;
  constexpr int ND = Mo::ND, NM = Mo::NM;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double facc[9];
  const int row = geo.order ? geo.order[rbase + (int) blockIdx.x] : rbase + (int) blockIdx.x;
  if (row >= nrows || (geo.dyn && row >= geo.dyn->rl.nrows)) return;
  const int lane = threadIdx.x & 63;
  const int i = row2atom[row];
  if (i < 0) return;  // padding row
  size_t off[4];
  Lo::bytes(0, off);  // the kept list stays in HBM
  float *Ms = reinterpret_cast<float *>(smem + off[1]);
  float *Hs = reinterpret_cast<float *>(smem + off[2]);
  float *Pl = reinterpret_cast<float *>(smem + off[3]);
  const double4 xi = pos[i];
  const Box cell = load_box(geo, i);
  const uint32_t *kr = kl + (size_t) row * kept_cache_words(cap);
  int start[kMaxElts + 1];
  for (int b = 0; b <= g.nelt; b++) start[b] = (int) kr[1 + b];
  const uint32_t *kept = kr + kKeptHdr;
  const float *grow = dEdG + (size_t) row * g.nelt * K * NM;
  if (lane < 9) facc[lane] = 0.0;
  for (int b = 0; b < g.nelt; b++) {
    build_P_hm<K, MM>(g, cell, xi, pos, tab, kept, start[b], start[b + 1], Ms, Hs, Pl);
    // S[k][d] = P[k][d] * 2 T[deg d][d] g'[k][deg d], g'_0 = g_0 / (2 G_0)
    // (0 where |G_0| < 1e-8, kernel.cu:752-767)
    for (int k = lane; k < K; k += 64) {
      float *Pk = Pl + k * ND;
      const float *gk = grow + ((size_t) b * K + k) * NM;
      float w[NM];
      const float G0 = Pk[0];
      w[0] = fabsf(G0) < 1e-8f ? 0.0f : gk[0] / G0;
#pragma unroll
      for (int m = 1; m < NM; m++) w[m] = 2.0f * gk[m];
#pragma unroll
      for (int d = 0; d < ND; d++) Pk[d] *= Mo::tab.mult[d] * w[Mo::tab.deg[d]];
    }
    __syncthreads();
    // chain rule, one lane per neighbor of the slot
    for (int base = start[b]; base < start[b + 1]; base += 64) {
      double cf[9];
#pragma unroll
      for (int q = 0; q < 9; q++) cf[q] = 0.0;
      if (base + lane < start[b + 1]) {
        NbrGeo ng;
        nbr_geometry(g, cell, xi, pos, kept[base + lane], ng);
        if (table_cell_live(g, ng.n)) {
          const float x = ng.ux, y = ng.uy, z = ng.uz;
          float M[ND], V[ND];
          monomials_hm<MM>(x, y, z, M);
#pragma unroll
          for (int d = 0; d < ND; d++) V[d] = 0.0f;
          float f1 = 0.0f;
          const float qp = ng.q - ng.pq;
          const float4 *trow = tab + (size_t) ng.n * K;
          for (int k = 0; k < K; k++) {
            float h, dh;
            kerley(trow[k], ng.q, ng.pq, qp, g.dxf, h, dh);
            const float *Sk = Pl + k * ND;
            float U = 0.0f;
#pragma unroll
            for (int d = 0; d < ND; d++) {
              const float s = Sk[d];
              U = fmaf(s, M[d], U);
              V[d] = fmaf(s, h, V[d]);
            }
            f1 = fmaf(U, dh, f1);
          }
          // g_beta = sum_d V_d dM_d/du_beta
          float px[MM + 1], py[MM + 1], pz[MM + 1];
          px[0] = py[0] = pz[0] = 1.0f;
#pragma unroll
          for (int e = 1; e <= MM; e++) {
            px[e] = px[e - 1] * x;
            py[e] = py[e - 1] * y;
            pz[e] = pz[e - 1] * z;
          }
          float gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
          for (int d = 1; d < ND; d++) {
            constexpr auto &t = Mo::tab;
            const int a = t.a[d], bb = t.b[d], c = t.c[d];
            if (a > 0) gx = fmaf(V[d], (float) a * px[a > 0 ? a - 1 : 0] * py[bb] * pz[c], gx);
            if (bb > 0) gy = fmaf(V[d], (float) bb * px[a] * py[bb > 0 ? bb - 1 : 0] * pz[c], gy);
            if (c > 0) gz = fmaf(V[d], (float) c * px[a] * py[bb] * pz[c > 0 ? c - 1 : 0], gz);
          }
          const float gu = gx * x + gy * y + gz * z;
          const float rinv = __builtin_amdgcn_rcpf((float) ng.r);
          const double F0 = (double) (f1 * x + (gx - gu * x) * rinv) * scale;
          const double F1v = (double) (f1 * y + (gy - gu * y) * rinv) * scale;
          const double F2v = (double) (f1 * z + (gz - gu * z) * rinv) * scale;
          const int j = ng.j;
          atomicAdd(&f[3 * (size_t) j + 0], -F0);
          atomicAdd(&f[3 * (size_t) j + 1], -F1v);
          atomicAdd(&f[3 * (size_t) j + 2], -F2v);
          const double dx = ng.d[0], dy = ng.d[1], dz = ng.d[2];
          const double v[6] = {-dx * F0, -dy * F1v, -dz * F2v, -dx * F1v, -dx * F2v, -dy * F2v};
          if (vatom) {
#pragma unroll
            for (int q = 0; q < 6; q++) atomicAdd(&vatom[6 * (size_t) j + q], 0.5 * v[q]);
          }
          cf[0] = F0;
          cf[1] = F1v;
          cf[2] = F2v;
#pragma unroll
          for (int q = 0; q < 6; q++) cf[3 + q] = v[q];
        }
      }
#pragma unroll
      for (int q = 0; q < 9; q++) cf[q] = wave_sum_d(cf[q]);
      if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 9; q++) facc[q] += cf[q];
      }
    }
    __syncthreads();  // Pl is rebuilt for the next slot
  }
  if (lane == 0) {
    atomicAdd(&f[3 * (size_t) i + 0], facc[0]);
    atomicAdd(&f[3 * (size_t) i + 1], facc[1]);
    atomicAdd(&f[3 * (size_t) i + 2], facc[2]);
  }
  if (lane < 6) {
    const double v = facc[3 + lane];
    watom[6 * (size_t) i + lane] = v;
    if (vatom) atomicAdd(&vatom[6 * (size_t) i + lane], 0.5 * v);
  }
  if (lane == 6) eatom[i] = scale * E[row];
}

template <int K, int MM>
hipError_t desc_hm(const ModelGeo &g, const Geom &geo, const double4 *pos, const uint32_t *nbr, const int *nn,
                   int cap, const int *row2atom, int rbase, int nrows, const float4 *tab, float *G, uint32_t *kl,
                   hipStream_t s)
{
  size_t off[4];
  const size_t lds = HmLayout<K, MM>::bytes(cap, off);
  if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
  if (hipError_t e = hipFuncSetAttribute((const void *) k_descriptor_hm<K, MM>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
      e != hipSuccess)
    return e;
  k_descriptor_hm<K, MM><<<nrows - rbase, 64, lds, s>>>(g, geo, pos, nbr, nn, cap, row2atom, rbase, nrows, tab, G,
                                                        kl);
  return hipGetLastError();
}

template <int K, int MM>
hipError_t force_hm(const ModelGeo &g, const Geom &geo, const double4 *pos, int cap, const int *row2atom,
                    int rbase, int nrows, const float4 *tab, const uint32_t *kl, const float *dEdG, const double *E,
                    double scale, double *f, double *eatom, double *watom, double *vatom, hipStream_t s)
{
  size_t off[4];
  const size_t lds = HmLayout<K, MM>::bytes(0, off);
  if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
  if (hipError_t e = hipFuncSetAttribute((const void *) k_force_hm<K, MM>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
      e != hipSuccess)
    return e;
  k_force_hm<K, MM><<<nrows - rbase, 64, lds, s>>>(g, geo, pos, cap, row2atom, rbase, nrows, tab, kl, dEdG, E,
                                                   scale, f, eatom, watom, vatom);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_descriptor_hm(const ModelGeo &g, const Geom &geo, const double4 *pos, const uint32_t *nbr,
                                const int *nn, int cap, const int *row2atom, int nrows, const float4 *tab,
                                float *G, uint32_t *kl, hipStream_t stream, int row_begin)
{
  if (nrows <= row_begin) return hipSuccess;
  if (g.nelt < 1 || g.nelt > kMaxElts) return hipErrorInvalidValue;
#define TMD_HM_CASE(K)                                                                                        \
  case K:                                                                                                     \
    return g.mm == 4 ? desc_hm<K, 4>(g, geo, pos, nbr, nn, cap, row2atom, row_begin, nrows, tab, G, kl, stream) \
                     : desc_hm<K, 5>(g, geo, pos, nbr, nn, cap, row2atom, row_begin, nrows, tab, G, kl, stream);
  if (g.mm != 4 && g.mm != 5) return hipErrorInvalidValue;
  switch (g.k) {
    TMD_HM_CASE(16)
    TMD_HM_CASE(24)
    TMD_HM_CASE(32)
    TMD_HM_CASE(64)
    TMD_HM_CASE(128)
    default: return hipErrorInvalidValue;
  }
#undef TMD_HM_CASE
}

hipError_t launch_force_hm(const ModelGeo &g, const Geom &geo, const double4 *pos, int cap, const int *row2atom,
                           int nrows, const float4 *tab, const uint32_t *kl, const float *dEdG, const double *E,
                           double scale, double *f, double *eatom, double *watom, double *vatom,
                           hipStream_t stream, int row_begin)
{
  if (nrows <= row_begin) return hipSuccess;
  if (g.nelt < 1 || g.nelt > kMaxElts) return hipErrorInvalidValue;
#define TMD_HM_CASE(K)                                                                                     \
  case K:                                                                                                  \
    return g.mm == 4 ? force_hm<K, 4>(g, geo, pos, cap, row2atom, row_begin, nrows, tab, kl, dEdG, E, scale, \
                                      f, eatom, watom, vatom, stream)                                      \
                     : force_hm<K, 5>(g, geo, pos, cap, row2atom, row_begin, nrows, tab, kl, dEdG, E, scale, \
                                      f, eatom, watom, vatom, stream);
  if (g.mm != 4 && g.mm != 5) return hipErrorInvalidValue;
  switch (g.k) {
    TMD_HM_CASE(16)
    TMD_HM_CASE(24)
    TMD_HM_CASE(32)
    TMD_HM_CASE(64)
    TMD_HM_CASE(128)
    default: return hipErrorInvalidValue;
  }
#undef TMD_HM_CASE
}

}  // namespace tmd
