// Fused descriptor and force/virial kernels (gfx950, wave64).
//
// One wavefront owns one atom.
//  1. Sweep the atom's neighbor list (cutoff rc+skin) in chunks of 64: each
//     lane forms one FP64 displacement, keeps 1e-10 < r^2 < rc^2
//     (gpu/cuda/kernel.cu:1163) and finds the element slot b
//     (tensormd.cpp:165-176).  Kept entries are compacted into LDS sorted
//     by slot (ballot prefix counts), so later passes run on full lanes.
//  2. For each chunk of 32 kept neighbors (two lanes per neighbor; 4 or 8
//     for a short last chunk): geometry u = d/r (FP32 from FP64),
//     monomials M_d(u), the Kerley-interpolated radial values H_k(r)
//     (gpu/cuda/ration.cu:13-71) with lanes over (neighbor, filter), then
//     P[b][k][d] += H_k M_d accumulated in registers: each lane owns a
//     KT x DT block of (k, d) for every slot.
//  3. Descriptor: G[b][k][m] = sum_d T[m][d] P^2 (m >= 1), G[b][k][0] = P0
//     (= sign(P0) sqrt(P0^2), kernel.cu:1028-1061).
//     Force: S = P * (2 sum_m T[m][d] g'_m), g'_0 = g_0/(2 G_0) guarded at
//     1e-8 (kernel.cu:730-817); per neighbor U_k = S M, V_d = S H,
//     F = scale (sum_k U_k H'_k) u + scale V . dM/dd (kernel.cu:559-646,
//     :869-965); f_i += F, f_j -= F, virial -d (x) F (kernel.cu:1714-1758).
//
// Nothing per neighbor is stored in HBM: the reference materialises R, M,
// H, dH/dr, P, dE/dS, V, F1, F2 per padded slot (kernel.cu:1545-1582).
#include "tmd_desc_common.h"

namespace tmd {

namespace {

constexpr int kMS = 24;  // LDS stride of a monomial row (16-byte aligned)

constexpr int kChunkRows = 32;  // neighbors staged at a time

template <int K>
struct Tiling {
  // each lane owns KT filters x DT monomials of every slot
  static constexpr int KT = K == 16 ? 1 : (K <= 32 ? 2 : (K == 64 ? 4 : 8));
  static constexpr int DT = (K == 24) ? 4 : 5;
  static constexpr int NDG = kND / DT;
  static constexpr int LANES = (K / KT) * NDG;  // 60 or 64
  // H and dH: [half][kChunk][HS] with the filters of a half in a row of odd
  // stride HS = K/2 + 1: the 64 lanes of a neighbor pass (32 rows x 2 halves)
  // hit 64 distinct banks
  static constexpr int KH = K / 2;
  static constexpr int HS = KH + 1;
  static constexpr int HALF = kChunkRows * HS;  // floats per half
  // S rows of one slot, padded so that slots b and b + 2 start on different banks
  static constexpr int SSTR = K * kND + 4;
  // LDS offset of filter k of row c
  __host__ __device__ static constexpr int hidx(int c, int k)
  {
    return (k >= KH ? HALF + k - KH : k) + c * HS;
  }
};

// T[m][d] of a max_moment-3 descriptor (kDeg, kMult, plus the T-symmetric
// extras T[2][0] = -1/3, T[3][1..3] = -3/5) for a lane-dependent d.
__device__ inline float t_entry(int m, int d, int tsym)
{
  constexpr unsigned long long kMultBits = [] {
    unsigned long long b = 0;
    for (int i = 0; i < kND; i++) b |= (unsigned long long) kMult[i] << (3 * i);
    return b;
  }();
  const int deg = d >= 10 ? 3 : (d >= 4 ? 2 : (d >= 1 ? 1 : 0));
  float t = deg == m ? (float) ((kMultBits >> (3 * d)) & 7ull) : 0.0f;
  if (tsym && m == 2 && d == 0) t = (float) (-1.0 / 3.0);
  if (tsym && m == 3 && d >= 1 && d <= 3) t = (float) (-3.0 / 5.0);
  return t;
}

// True when monomials [d0, d0 + n) include one of degree m.
__host__ __device__ constexpr bool group_has_degree(int d0, int n, int m)
{
  for (int d = d0; d < d0 + n && d < kND; d++)
    if (kDeg[d] == m) return true;
  return false;
}

constexpr int kChunk = kChunkRows;

struct LDSMap {
  uint32_t *kept;  // [cap] compacted, slot-sorted neighbor entries
  float *M;        // [kChunk][kMS] monomials of the staged chunk
  float *H;        // [2 halves][kChunk][HS] radial functions of the staged chunk
};

// descriptor LDS: kept list, monomials, radial functions
template <int K>
__host__ __device__ inline size_t lds_layout(int cap, size_t *offs)
{
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t r = o;
    o += (bytes + 15) & ~(size_t) 15;
    return r;
  };
  offs[0] = take((size_t) cap * 4);
  offs[1] = take((size_t) kChunk * kMS * 4);
  offs[2] = take((size_t) 2 * Tiling<K>::HALF * 4);
  return o;
}

// force LDS: the staging area (monomials, radial functions), then P and S
// of every slot [nelt][SSTR] at offs[0]: over the staging area when P is
// built in one pass, after it when it takes several (the staging area is
// still in use while the first passes' P is stored).  The kept list stays
// in HBM.
template <int K>
__host__ __device__ inline size_t force_lds_layout(int nelt, bool multipass, size_t *offs)
{
  const size_t mreg = (size_t) kChunk * kMS * 4;
  const size_t hreg = (size_t) 2 * Tiling<K>::HALF * 4;
  const size_t preg = (size_t) nelt * Tiling<K>::SSTR * 4;
  offs[1] = 0;
  offs[2] = mreg;
  if (multipass) {
    offs[0] = (mreg + hreg + 15) & ~(size_t) 15;
    return offs[0] + preg;
  }
  offs[0] = 0;
  return mreg + hreg > preg ? mreg + hreg : preg;
}

template <int K>
__device__ inline LDSMap carve(unsigned char *base, int cap)
{
  size_t offs[3];
  lds_layout<K>(cap, offs);
  LDSMap L;
  L.kept = (uint32_t *) (base + offs[0]);
  L.M = (float *) (base + offs[1]);
  L.H = (float *) (base + offs[2]);
  return L;
}

// Lanes per neighbor in the chain-rule pass: one lane per neighbor (a chunk
// of 64) while more than 48 remain, then as lanes_per_nbr.  One wide chunk
// runs the per-neighbor set-up and the force scatter once instead of twice.
__device__ inline int chain_lanes(int rem) { return rem > 48 ? 1 : lanes_per_nbr(rem); }

// table loads in flight per lane: divides K/2, K/4 and K/8
template <int K>
constexpr int kTabBatch = ((K / 8) % 4 == 0) ? 4 : K / 8;

// Phase 2 staging of kept neighbors [base, base + 64/Lp) for the P
// accumulation: Lp adjacent lanes per neighbor.  Each computes the
// neighbor's geometry and interpolates K/Lp of the filters from the table
// row of the neighbor's cell (immediate-offset loads, no per-element index
// arithmetic).  Writes the monomials (M) and H to LDS.
template <int K>
__device__ inline void stage(const ModelGeo &g, const Box &cell, const double4 &xi,
                             const double4 *pos, const float4 *tab, int base, int nk, int Lp,
                             LDSMap &L)
{
  static_assert(kChunk == 32, "at most 32 neighbors staged at a time");
  using T = Tiling<K>;
  const int lane = threadIdx.x & 63;
  const int c = lane / Lp, part = lane - c * Lp;
  const int KL = K / Lp;
  if (base + c < nk) {
    NbrGeo ng;
    nbr_geometry(g, cell, xi, pos, L.kept[base + c], ng);
    if (part == 0) {
      float M[kND];
      monomials20(ng.ux, ng.uy, ng.uz, M);
      float4 *Mr = reinterpret_cast<float4 *>(L.M + c * kMS);
#pragma unroll
      for (int q4 = 0; q4 < kND / 4; q4++)
        Mr[q4] = make_float4(M[4 * q4], M[4 * q4 + 1], M[4 * q4 + 2], M[4 * q4 + 3]);
    }
    // the lane's filters lie in one half of the H layout (K/Lp divides K/2)
    float *Hc = L.H + T::hidx(c, part * KL);
    constexpr int B = kTabBatch<K>;
    if (table_cell_live(g, ng.n)) {
      const float4 *row = tab + (size_t) ng.n * K + part * KL;
      // six rows per wait where they divide the lane's filters: the staging
      // is latency-bound, and each wait covering more loads cuts the
      // descriptor and the P rebuild by 7-8 % at K = 24 (twelve rows raise
      // the descriptor to 143 VGPRs and lose)
      constexpr int B6 = (K / 2) % 6 == 0 ? 6 : B;
      if (KL % B6 == 0) {
        for (int k0 = 0; k0 < KL; k0 += B6) {
          float4 t[B6];
#pragma unroll
          for (int u = 0; u < B6; u++) t[u] = row[k0 + u];
#pragma unroll
          for (int u = 0; u < B6; u++) Hc[k0 + u] = kerley_h(t[u], ng.q, ng.pq);
        }
      } else {
        for (int k0 = 0; k0 < KL; k0 += B) {
          float4 t[B];
#pragma unroll
          for (int u = 0; u < B; u++) t[u] = row[k0 + u];
#pragma unroll
          for (int u = 0; u < B; u++) Hc[k0 + u] = kerley_h(t[u], ng.q, ng.pq);
        }
      }
    } else {
      for (int k = 0; k < KL; k++) Hc[k] = 0.0f;
    }
  }
  __syncthreads();
}

// Phase 2 accumulation over the staged chunk into the lane's register block.
// The pass covers nsl slots whose kept entries start at start[0..nsl]; slot
// b of the pass accumulates into acc[b].
template <int K, int NS>
__device__ inline void accumulate(const LDSMap &L, const int *start, int nsl, int base,
                                  float (*acc)[Tiling<K>::KT][Tiling<K>::DT])
{
  using T = Tiling<K>;
  const int lane = threadIdx.x & 63;
  if (lane >= T::LANES) return;
  const int kq = lane / T::NDG, dq = lane - (lane / T::NDG) * T::NDG;
  const float *Hb = L.H + T::hidx(0, kq * T::KT);  // KT divides K/2: a lane's filters share a half
  const float *Mb = L.M + dq * T::DT;
#pragma unroll
  for (int b = 0; b < NS; b++) {
    if (b >= nsl) break;
    int c0 = max(start[b] - base, 0), c1 = min(start[b + 1] - base, kChunk);
#pragma unroll 4
    for (int c = c0; c < c1; c++) {
      float h[T::KT], m[T::DT];
#pragma unroll
      for (int a = 0; a < T::KT; a++) h[a] = Hb[c * T::HS + a];
#pragma unroll
      for (int q = 0; q < T::DT; q++) m[q] = Mb[c * kMS + q];
#pragma unroll
      for (int a = 0; a < T::KT; a++)
#pragma unroll
        for (int q = 0; q < T::DT; q++) acc[b][a][q] += h[a] * m[q];
    }
  }
}

// Phase 2: P of the nsl slots of one pass in the lane's register block, over
// the kept entries L.kept[start[0], start[nsl]) (the list is slot-sorted, so
// a pass stages only its own slots' neighbors).
template <int K, int NS>
__device__ inline void accumulate_P(const ModelGeo &g, const Box &cell, const double4 &xi,
                                    const double4 *pos, const float4 *tab, LDSMap &L,
                                    const int *start, int nsl,
                                    float (*acc)[Tiling<K>::KT][Tiling<K>::DT])
{
  using T = Tiling<K>;
#pragma unroll
  for (int b = 0; b < NS; b++)
#pragma unroll
    for (int a = 0; a < T::KT; a++)
#pragma unroll
      for (int q = 0; q < T::DT; q++) acc[b][a][q] = 0.0f;
  const int end = start[nsl];
  for (int base = start[0]; base < end;) {
    const int Lp = lanes_per_nbr(end - base);
    stage<K>(g, cell, xi, pos, tab, base, end, Lp, L);
    accumulate<K, NS>(L, start, nsl, base, acc);
    __syncthreads();
    base += 64 / Lp;
  }
}

typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));

// G[b0 + b][k][m] = sum_d T[m][d] P^2 (m >= 1) for the nsl slots of a pass:
// partial sums over the lane's d block, then a sum over the NDG lanes that
// share k; G[b][k][0] = P0 (= sign(P0) sqrt(P0^2)), held by the dq == 0 lane.
// gbf16: the row is stored as BF16 (round to nearest even).
template <int K, int NS>
__device__ inline void write_G(const ModelGeo &g, float (*acc)[Tiling<K>::KT][Tiling<K>::DT],
                               float *Grow, int b0, int nsl, bool gbf16)
{
  using T = Tiling<K>;
  const int lane = threadIdx.x & 63;
  const int kq = lane / T::NDG, dq = lane - (lane / T::NDG) * T::NDG;
  const int gbase = kq * T::NDG;
  float tc[kNM][T::DT];
#pragma unroll
  for (int m = 1; m < kNM; m++)
#pragma unroll
    for (int q = 0; q < T::DT; q++) tc[m][q] = t_entry(m, dq * T::DT + q, g.tsym);
#pragma unroll
  for (int b = 0; b < NS; b++) {
    if (b >= nsl) break;
#pragma unroll
    for (int a = 0; a < T::KT; a++) {
      float part[kNM] = {0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < T::DT; q++) {
        float p2 = acc[b][a][q] * acc[b][a][q];
#pragma unroll
        for (int m = 1; m < kNM; m++) part[m] += tc[m][q] * p2;
      }
      // sum over the NDG lanes of the filter: only the d groups that hold a
      // monomial of degree m add to G_m (the others' partial sums are exact
      // zeros; the T-symmetric extras are in the d group 0, the lane itself)
      float tot[kNM] = {0, part[1], part[2], part[3]};
#pragma unroll
      for (int t = 1; t < T::NDG; t++)
#pragma unroll
        for (int m = 1; m < kNM; m++)
          if (group_has_degree(t * T::DT, T::DT, m)) tot[m] += __shfl(part[m], gbase + t);
      if (lane < T::LANES && dq == 0) {
        const int k = kq * T::KT + a;
        const size_t o = ((size_t) (b0 + b) * K + k) * kNM;
        if (gbf16) {
          const f32x4v xv = {acc[b][a][0], tot[1], tot[2], tot[3]};
          *reinterpret_cast<bf16x4v *>(reinterpret_cast<uint16_t *>(Grow) + o) = __builtin_convertvector(xv, bf16x4v);
        } else {
          *reinterpret_cast<float4 *>(Grow + o) = make_float4(acc[b][a][0], tot[1], tot[2], tot[3]);
        }
      }
    }
  }
}

// NS: slots accumulated per register pass.  MP: the model has more slots
// than NS, and the kernel runs ceil(nelt / NS) passes over the slot-sorted
// kept list (each neighbor is still staged once).
template <int K, int NS, bool MP>
__global__ __launch_bounds__(64) void k_descriptor(ModelGeo g, Geom geo, const double4 *pos,
                                                   const uint32_t *nbr, const int *nn, int cap,
                                                   const int *row2atom, int rbase, int nrows,
                                                   const float4 *tab, float *G, uint32_t *kl)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int row = geo.order ? geo.order[rbase + (int) blockIdx.x] : rbase + (int) blockIdx.x;
  if (row >= nrows || (geo.dyn && row >= geo.dyn->rl.nrows)) return;
  const int lane = threadIdx.x & 63;
  const int gw = g.nelt * K * kNM;
  // BF16 rows: gw 2-byte values per row (gw is a multiple of 4)
  float *Grow = geo.g_bf16 ? reinterpret_cast<float *>(reinterpret_cast<uint16_t *>(G) + (size_t) row * gw)
                           : G + (size_t) row * gw;
  const int i = row2atom[row];
  if (i < 0) {  // padding row
    const int nw = geo.g_bf16 ? gw / 2 : gw;
    for (int e = lane; e < nw; e += 64) Grow[e] = 0.0f;
    return;
  }
  using T = Tiling<K>;
  LDSMap L = carve<K>(smem, cap);
  int start[kMaxElts + 1];
  const double4 xi = pos[i];
  const Box cell = load_box(geo, i);
  // phase 1: compaction
  const int nk = compact(g, cell, xi, (int) xi.w, nbr + (size_t) i * cap, nn[i], pos, L.kept, start);
  __syncthreads();
  // the row's kept list with its slot starts, for the force kernel
  auto write_kept = [&]() {
    uint32_t *kr = kl + (size_t) row * kept_cache_words(cap);
    if (lane == 0) {
      kr[0] = (uint32_t) nk;
      for (int b = 0; b <= g.nelt; b++) kr[1 + b] = (uint32_t) start[b];
    }
    for (int e = lane; e < nk; e += 64) kr[kKeptHdr + e] = L.kept[e];
  };
  if constexpr (!MP) {
    float acc[NS][T::KT][T::DT];
    accumulate_P<K, NS>(g, cell, xi, pos, tab, L, start, g.nelt, acc);
    write_kept();
    write_G<K, NS>(g, acc, Grow, 0, g.nelt, geo.g_bf16);
  } else {
    __shared__ int s_start[kMaxElts + 1];
    if (lane == 0)
      for (int b = 0; b <= g.nelt; b++) s_start[b] = start[b];
    write_kept();
    __syncthreads();
    for (int b0 = 0; b0 < g.nelt; b0 += NS) {
      const int nsl = min(NS, g.nelt - b0);
      float acc[NS][T::KT][T::DT];
      accumulate_P<K, NS>(g, cell, xi, pos, tab, L, s_start + b0, nsl, acc);
      write_G<K, NS>(g, acc, Grow, b0, nsl, geo.g_bf16);
    }
  }
}

// One chunk of the chain-rule pass: LP adjacent lanes per neighbor (neighbors
// [base, base + 64/LP)), each over K/LP of the filters.  A lane recomputes
// its neighbor's geometry and interpolates H and dH/dr of its filters in
// registers, so the only LDS traffic is the broadcast read of S.  Adds the
// neighbor forces to f_j (FP64 atomics) and to the atom's fi, wi.
template <int K, int LP>
__device__ inline void force_chunk(const ModelGeo &g, const Box &cell, const double4 &xi,
                                   const double4 *pos, const float4 *tab, const float *PS,
                                   const uint32_t *kept, const int *start, int base, int nk,
                                   double scale, double *f, double *vatom, double *facc)
{
  using T = Tiling<K>;
  const int lane = threadIdx.x & 63;
  const int c = lane / LP, sub = lane - c * LP;
  constexpr int KL = K / LP;
  {
    const bool live = base + c < nk;
    int b = 0;
    while (b + 1 < g.nelt && base + c >= start[b + 1]) b++;
    NbrGeo ng;
    ng.ux = ng.uy = ng.uz = 0.0f;
    ng.r = 1.0;
    float V[kND];
#pragma unroll
    for (int q = 0; q < kND; q++) V[q] = 0.0f;
    float f1 = 0.0f;
    if (live) {
      nbr_geometry(g, cell, xi, pos, kept[base + c], ng);
      if (table_cell_live(g, ng.n)) {
        float M[kND];
        monomials20(ng.ux, ng.uy, ng.uz, M);
        const float qp = ng.q - ng.pq;
        const float4 *trow = tab + (size_t) ng.n * K + sub * KL;
        const float *S = PS + (size_t) b * T::SSTR + sub * KL * kND;
        // batches of B filters: load the table rows, interpolate H and dH/dr,
        // then U_k = sum_d S_kd M_d (M_0 = 1), V_d += S_kd H_k (V_0 is not
        // needed: dM_0/du = 0), f1 += U_k dH_k.  No software pipelining: four
        // waves per SIMD hide the row latency, and a prefetch ring costs
        // registers and copies (measured slower).  Six rows per batch where
        // they divide the lane's filters (K = 24: 3 % faster than 3 or 8 rows;
        // 12 rows spill), else the staging batch.
        constexpr int B = KL % 6 == 0 ? 6 : kTabBatch<K>;
#pragma unroll 1
        for (int k0 = 0; k0 < KL; k0 += B) {
          float4 t[B];
#pragma unroll
          for (int u = 0; u < B; u++) t[u] = trow[k0 + u];
#pragma unroll
          for (int u = 0; u < B; u++) {
            float h, dh;
            kerley(t[u], ng.q, ng.pq, qp, g.dxf, h, dh);
            const float4 *Sk = reinterpret_cast<const float4 *>(S + (k0 + u) * kND);
            float sv[kND];
#pragma unroll
            for (int q4 = 0; q4 < kND / 4; q4++) {
              float4 v4 = Sk[q4];
              sv[4 * q4] = v4.x;
              sv[4 * q4 + 1] = v4.y;
              sv[4 * q4 + 2] = v4.z;
              sv[4 * q4 + 3] = v4.w;
            }
            float U = sv[0];
#pragma unroll
            for (int q = 1; q < kND; q++) U = fmaf(sv[q], M[q], U);
#pragma unroll
            for (int q = 1; q < kND; q++) V[q] = fmaf(sv[q], h, V[q]);
            f1 = fmaf(U, dh, f1);
          }
        }
      }
    }
    // combine the Lp partial sums into the neighbor's first lane
    if constexpr (LP >= 2) {
#pragma unroll
      for (int q = 1; q < kND; q++) V[q] += dpp_f<kDppXor1>(V[q]);
      f1 += dpp_f<kDppXor1>(f1);
    }
    if constexpr (LP >= 4) {
#pragma unroll
      for (int q = 1; q < kND; q++) V[q] += dpp_f<kDppXor2>(V[q]);
      f1 += dpp_f<kDppXor2>(f1);
    }
    if constexpr (LP == 8) {
#pragma unroll
      for (int q = 1; q < kND; q++) V[q] += __shfl_xor(V[q], 4);
      f1 += __shfl_xor(f1, 4);
    }
    double cf[9];
#pragma unroll
    for (int q = 0; q < 9; q++) cf[q] = 0.0;
    if (live && sub == 0) {
      const float x = ng.ux, y = ng.uy, z = ng.uz;
      // g_beta = sum_d V_d dM_d/du_beta
      float gx = V[1] + 2.f * x * V[4] + y * V[5] + z * V[6] + 3.f * x * x * V[10] +
          2.f * x * y * V[11] + 2.f * x * z * V[12] + y * y * V[13] + y * z * V[14] + z * z * V[15];
      float gy = V[2] + x * V[5] + 2.f * y * V[7] + z * V[8] + x * x * V[11] + 2.f * x * y * V[13] +
          x * z * V[14] + 3.f * y * y * V[16] + 2.f * y * z * V[17] + z * z * V[18];
      float gz = V[3] + x * V[6] + y * V[8] + 2.f * z * V[9] + x * x * V[12] + x * y * V[14] +
          2.f * x * z * V[15] + y * y * V[17] + 2.f * y * z * V[18] + 3.f * z * z * V[19];
      float gu = gx * x + gy * y + gz * z;
      // 1-ulp hardware reciprocal (a division is a ten-instruction FP64 sequence)
      float rinv = __builtin_amdgcn_rcpf((float) ng.r);
      // dE_i/dd = F1 + F2; dM/dd_a = sum_b dM/du_b (delta_ab - u_a u_b)/r
      double F0 = (double) (f1 * x + (gx - gu * x) * rinv) * scale;
      double F1v = (double) (f1 * y + (gy - gu * y) * rinv) * scale;
      double F2v = (double) (f1 * z + (gz - gu * z) * rinv) * scale;
      const int j = ng.j;
      atomicAdd(&f[3 * (size_t) j + 0], -F0);
      atomicAdd(&f[3 * (size_t) j + 1], -F1v);
      atomicAdd(&f[3 * (size_t) j + 2], -F2v);
      const double dx = ng.d[0], dy = ng.d[1], dz = ng.d[2];
      double v[6] = {-dx * F0, -dy * F1v, -dz * F2v, -dx * F1v, -dx * F2v, -dy * F2v};
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
    // the chunk's own-force and virial sums, in a fixed lane order, into the
    // atom's LDS accumulators (nothing FP64 is live through the filter loop);
    // only each neighbor's first lane holds values
#pragma unroll
    for (int q = 0; q < 9; q++) cf[q] = wave_sum_d_lp<LP>(cf[q]);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < 9; q++) facc[q] += cf[q];
    }
  }
}

// 4 waves per SIMD (at most 128 VGPRs): the chain-rule pass hides its table
// and LDS latency with more waves; the few spilled values live outside its loops
// Waves per SIMD asked of the force kernel: 4 (at most 128 VGPRs, a few
// values spilled outside the loops) where the P block of a lane is small;
// the wider variants keep their registers and run fewer waves.
template <int K, int NS>
constexpr int kForceWaves = (K <= 24 && NS * Tiling<K>::KT * Tiling<K>::DT <= 40) ? 4 : 1;

// P of the nsl slots of a pass -> LDS [b0 + b][k][d]
template <int K, int NS>
__device__ inline void store_P(float (*acc)[Tiling<K>::KT][Tiling<K>::DT], float *PS, int b0, int nsl)
{
  using T = Tiling<K>;
  const int lane = threadIdx.x & 63;
  if (lane < T::LANES) {
    const int kq = lane / T::NDG, dq = lane - (lane / T::NDG) * T::NDG;
#pragma unroll
    for (int b = 0; b < NS; b++) {
      if (b >= nsl) break;
#pragma unroll
      for (int a = 0; a < T::KT; a++)
#pragma unroll
        for (int q = 0; q < T::DT; q++)
          PS[(size_t) (b0 + b) * T::SSTR + (kq * T::KT + a) * kND + dq * T::DT + q] = acc[b][a][q];
    }
  }
}

// NS, MP: as for k_descriptor
template <int K, int NS, bool MP>
__global__ __launch_bounds__(64, (kForceWaves<K, NS>)) void k_force(ModelGeo g, Geom geo, const double4 *pos, int cap,
                                              const int *row2atom, int rbase, int nrows, const float4 *tab,
                                              const uint32_t *kl, const float *dEdG,
                                              const double *E, double scale, double *f,
                                              double *eatom, double *watom, double *vatom)
{
  using T = Tiling<K>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int row = geo.order ? geo.order[rbase + (int) blockIdx.x] : rbase + (int) blockIdx.x;
  if (row >= nrows || (geo.dyn && row >= geo.dyn->rl.nrows)) return;
  const int lane = threadIdx.x & 63;
  const int i = row2atom[row];
  if (i < 0) return;  // padding row
  size_t offs[3];
  force_lds_layout<K>(g.nelt, MP, offs);
  LDSMap L;
  L.M = reinterpret_cast<float *>(smem + offs[1]);
  L.H = reinterpret_cast<float *>(smem + offs[2]);
  float *PS = reinterpret_cast<float *>(smem + offs[0]);
  const double4 xi = pos[i];
  const Box cell = load_box(geo, i);
  // the descriptor kernel's kept list of this row (slot-sorted, within rc)
  const uint32_t *kr = kl + (size_t) row * kept_cache_words(cap);
  const int nk = (int) kr[0];
  int start[kMaxElts + 1];
  for (int b = 0; b <= g.nelt; b++) start[b] = (int) kr[1 + b];
  const uint32_t *kept = kr + kKeptHdr;
  L.kept = const_cast<uint32_t *>(kept);
  if constexpr (!MP) {
    float acc[NS][T::KT][T::DT];
    accumulate_P<K, NS>(g, cell, xi, pos, tab, L, start, g.nelt, acc);
    // P -> LDS over the staging area (accumulate_P ends with a barrier)
    store_P<K, NS>(acc, PS, 0, g.nelt);
  } else {
    __shared__ int s_start[kMaxElts + 1];
    if (lane == 0)
      for (int b = 0; b <= g.nelt; b++) s_start[b] = start[b];
    __syncthreads();
    for (int b0 = 0; b0 < g.nelt; b0 += NS) {
      const int nsl = min(NS, g.nelt - b0);
      float acc[NS][T::KT][T::DT];
      accumulate_P<K, NS>(g, cell, xi, pos, tab, L, s_start + b0, nsl, acc);
      store_P<K, NS>(acc, PS, b0, nsl);
    }
  }
  __syncthreads();

  // S[b][k][d] = P * sum_m 2 T[m][d] g'_m, in place: T[m][d] is the
  // multiplicity of monomial d in its degree m, plus the T-symmetric extras
  const float *grow = dEdG + (size_t) row * g.nelt * K * kNM;
  for (int e = lane; e < g.nelt * K; e += 64) {
    float *P = PS + (size_t) (e / K) * T::SSTR + (e % K) * kND;
    float4 gv = *reinterpret_cast<const float4 *>(grow + (size_t) e * kNM);
    float G0 = P[0];
    float g0 = (fabsf(G0) < 1e-8f) ? 0.0f : gv.x * __builtin_amdgcn_rcpf(G0) * 0.5f;
    float w[kNM] = {2.0f * g0, 2.0f * gv.y, 2.0f * gv.z, 2.0f * gv.w};
    float w0 = w[0], w1 = w[1];
    if (g.tsym) {
      w0 = fmaf(w[2], (float) (-1.0 / 3.0), w0);
      w1 = fmaf(w[3], (float) (-3.0 / 5.0), w1);
    }
    P[0] *= w0;
#pragma unroll
    for (int d = 1; d < kND; d++) {
      const float wd = kDeg[d] == 1 ? w1 : (kMult[d] == 1 ? w[kDeg[d]] : (float) kMult[d] * w[kDeg[d]]);
      P[d] *= wd;
    }
  }
  __syncthreads();

  // Phase 3: the chain rule, chunk by chunk (1 lane per neighbor, or 2 / 4 / 8
  // for a short last chunk)
  // own force (3) and virial (6) of the atom, summed chunk by chunk
  __shared__ double facc[9];
  if (lane < 9) facc[lane] = 0.0;
  for (int base = 0; base < nk;) {
    const int Lp = chain_lanes(nk - base);
    if (Lp == 1)
      force_chunk<K, 1>(g, cell, xi, pos, tab, PS, kept, start, base, nk, scale, f, vatom, facc);
    else if (Lp == 2)
      force_chunk<K, 2>(g, cell, xi, pos, tab, PS, kept, start, base, nk, scale, f, vatom, facc);
    else if (Lp == 4)
      force_chunk<K, 4>(g, cell, xi, pos, tab, PS, kept, start, base, nk, scale, f, vatom, facc);
    else
      force_chunk<K, 8>(g, cell, xi, pos, tab, PS, kept, start, base, nk, scale, f, vatom, facc);
    base += 64 / Lp;
  }
  __syncthreads();
  double fi[3], wi[6];
#pragma unroll
  for (int a = 0; a < 3; a++) fi[a] = facc[a];
#pragma unroll
  for (int q = 0; q < 6; q++) wi[q] = facc[3 + q];

  if (lane == 0) {
    atomicAdd(&f[3 * (size_t) i + 0], fi[0]);
    atomicAdd(&f[3 * (size_t) i + 1], fi[1]);
    atomicAdd(&f[3 * (size_t) i + 2], fi[2]);
  }
  if (lane < 6) {
    double v = wi[lane];
    watom[6 * (size_t) i + lane] = v;
    if (vatom) atomicAdd(&vatom[6 * (size_t) i + lane], 0.5 * v);
  }
  if (lane == 6) eatom[i] = scale * E[row];
}

template <int K, int NS, bool MP>
hipError_t desc_impl(const ModelGeo &g, const Geom &geo, const double4 *pos, const uint32_t *nbr,
                     const int *nn, int cap, const int *row2atom, int rbase, int nrows, const float4 *tab,
                     float *G, uint32_t *kl, hipStream_t s)
{
  size_t offs[3];
  size_t lds = lds_layout<K>(cap, offs);
  if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
  if (hipError_t ae = hipFuncSetAttribute((const void *) k_descriptor<K, NS, MP>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
      ae != hipSuccess)
    return ae;
  k_descriptor<K, NS, MP><<<nrows - rbase, 64, lds, s>>>(g, geo, pos, nbr, nn, cap, row2atom, rbase, nrows, tab,
                                                         G, kl);
  return hipGetLastError();
}

template <int K, int NS, bool MP>
hipError_t force_impl(const ModelGeo &g, const Geom &geo, const double4 *pos, int cap,
                      const int *row2atom, int rbase, int nrows, const float4 *tab, const uint32_t *kl,
                      const float *dEdG, const double *E, double scale, double *f, double *eatom,
                      double *watom, double *vatom, hipStream_t s)
{
  size_t offs[3];
  const size_t lds = force_lds_layout<K>(g.nelt, MP, offs);
  if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
  if (hipError_t ae = hipFuncSetAttribute((const void *) k_force<K, NS, MP>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
      ae != hipSuccess)
    return ae;
  k_force<K, NS, MP><<<nrows - rbase, 64, lds, s>>>(g, geo, pos, cap, row2atom, rbase, nrows, tab, kl, dEdG, E,
                                                    scale, f, eatom, watom, vatom);
  return hipGetLastError();
}

// Slots per register pass: the smallest of 1, 2, 4, 8 that holds nelt
// within a budget of 128 accumulators per lane; models with more slots than
// the largest such count (k = 64: 4, k = 128: 2) take several passes.
template <int K, bool FORCE>
hipError_t dispatch(const ModelGeo &g, const Geom &geo, const double4 *pos, const uint32_t *nbr,
                    const int *nn, int cap, const int *row2atom, int rbase, int nrows, const float4 *tab,
                    float *G, uint32_t *kl, const float *dEdG, const double *E, double scale,
                    double *f, double *eatom, double *watom, double *vatom, hipStream_t s)
{
  constexpr int per = Tiling<K>::KT * Tiling<K>::DT;
  constexpr int kMaxNS = 8 * per <= 128 ? 8 : (4 * per <= 128 ? 4 : (2 * per <= 128 ? 2 : 1));
#define TMD_LAUNCH(NS, MP)                                                                          \
  return FORCE ? force_impl<K, NS, MP>(g, geo, pos, cap, row2atom, rbase, nrows, tab, kl, dEdG, E,   \
                                       scale, f, eatom, watom, vatom, s)                            \
               : desc_impl<K, NS, MP>(g, geo, pos, nbr, nn, cap, row2atom, rbase, nrows, tab, G, kl, s);
  static_assert(kMaxNS >= 2, "every filter count holds two slots per pass");
  if (g.nelt < 1 || g.nelt > kMaxElts) return hipErrorInvalidValue;
  if (g.nelt <= 1) TMD_LAUNCH(1, false)
  if (g.nelt <= 2) TMD_LAUNCH(2, false)
  if constexpr (kMaxNS >= 8) {
    if (g.nelt <= 4) TMD_LAUNCH(4, false)
    TMD_LAUNCH(8, false)
  } else if constexpr (kMaxNS == 4) {
    if (g.nelt <= 4) TMD_LAUNCH(4, false)
    TMD_LAUNCH(4, true)
  } else {
    TMD_LAUNCH(2, true)
  }
#undef TMD_LAUNCH
}

}  // namespace

hipError_t launch_descriptor(const ModelGeo &g, const Geom &geo, const double4 *pos,
                             const uint32_t *nbr, const int *nn, int cap, const int *row2atom,
                             int nrows, const float4 *tab, float *G, uint32_t *kl,
                             hipStream_t stream, int row_begin)
{
  if (nrows <= row_begin) return hipSuccess;
  if (geo.g_bf16 && g.mm > 3) return hipErrorInvalidValue;  // FP32 rows only there
  if (geo.order && row_begin != 0) return hipErrorInvalidValue;  // an order covers whole launches
  if (g.mm > 3)
    return launch_descriptor_hm(g, geo, pos, nbr, nn, cap, row2atom, nrows, tab, G, kl, stream, row_begin);
#define TMD_K_CASE(K)                                                                                    \
  case K:                                                                                                \
    return dispatch<K, false>(g, geo, pos, nbr, nn, cap, row2atom, row_begin, nrows, tab, G, kl, nullptr, \
                              nullptr, 0, nullptr, nullptr, nullptr, nullptr, stream);
  switch (g.k) {
    TMD_K_CASE(16)
    TMD_K_CASE(24)
    TMD_K_CASE(32)
    TMD_K_CASE(64)
    TMD_K_CASE(128)
    default: return hipErrorInvalidValue;
  }
#undef TMD_K_CASE
}

hipError_t launch_force(const ModelGeo &g, const Geom &geo, const double4 *pos, int cap,
                        const int *row2atom, int nrows, const float4 *tab, const uint32_t *kl,
                        const float *dEdG, const double *E, double scale, double *f, double *eatom,
                        double *watom, double *vatom, hipStream_t stream, int row_begin)
{
  if (nrows <= row_begin) return hipSuccess;
  if (geo.order && row_begin != 0) return hipErrorInvalidValue;  // an order covers whole launches
  if (g.mm > 3)
    return launch_force_hm(g, geo, pos, cap, row2atom, nrows, tab, kl, dEdG, E, scale, f, eatom, watom, vatom,
                           stream, row_begin);
#define TMD_K_CASE(K)                                                                                       \
  case K:                                                                                                   \
    return dispatch<K, true>(g, geo, pos, nullptr, nullptr, cap, row2atom, row_begin, nrows, tab, nullptr,  \
                             const_cast<uint32_t *>(kl), dEdG, E, scale, f, eatom, watom, vatom, stream);
  switch (g.k) {
    TMD_K_CASE(16)
    TMD_K_CASE(24)
    TMD_K_CASE(32)
    TMD_K_CASE(64)
    TMD_K_CASE(128)
    default: return hipErrorInvalidValue;
  }
#undef TMD_K_CASE
}

}  // namespace tmd
