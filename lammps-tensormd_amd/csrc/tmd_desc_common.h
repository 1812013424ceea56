// Device helpers shared by the descriptor and force kernels
// (tmd_desc.hip: max_moment 0-3; tmd_desc_hm.hip: max_moment 4-5):
// periodic displacements, the Kerley interpolation, wave reductions, the
// slot-sorted compaction of an atom's list and per-neighbor geometry.
#pragma once
#include "tmd_device.h"

namespace tmd {

namespace {

// Lattice rows of the cell an atom belongs to (uniform per workgroup).
struct Box {
  double h[9];
};

__device__ inline Box load_box(const Geom &geo, int i)
{
  const CellInfo &c = geo.cells[geo.atom_struct ? geo.atom_struct[i] : 0];
  Box b;
#pragma unroll
  for (int q = 0; q < 9; q++) b.h[q] = c.h[q];
  return b;
}

__device__ inline void displacement(const double4 &xi, const double4 &xj, uint32_t code,
                                    const Box &c, double d[3])
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

// Kerley rational interpolation of one filter (gpu/cuda/ration.cu:59-68),
// table entry t = (y, s, c1, c2), q = r - n dx, p = dx - q, qp = q - p:
//   w = |c2 p| / (|c2 p| + |c1 q|)  (0 when both vanish)
//   c = c2 + w (c1 - c2),  h = y + q (s - p c),
//   dh/dr = s + (q - p) c + dx w (1 - w) (c1 - c2)
// (the last term is the reference's dx (c - c2)(1 - w)).  The max() stands
// for its "denominator > 0" test: a zero denominator has a zero numerator.
// The caller zeroes the two end cells.
__device__ inline float kerley_w(float4 t, float q, float p)
{
  float a = fabsf(t.w) * fabsf(p);
  float den = fmaf(fabsf(t.z), fabsf(q), a);
  return a * __builtin_amdgcn_rcpf(fmaxf(den, 1.17549435e-38f));  // 1-ulp reciprocal
}

__device__ inline float kerley_h(float4 t, float q, float p)
{
  float c = fmaf(kerley_w(t, q, p), t.z - t.w, t.w);
  return fmaf(q, fmaf(-p, c, t.y), t.x);
}

__device__ inline void kerley(float4 t, float q, float p, float qp, float dx, float &h, float &dh)
{
  float w = kerley_w(t, q, p);
  float dc = t.z - t.w;
  float c = fmaf(w, dc, t.w);
  h = fmaf(q, fmaf(-p, c, t.y), t.x);
  dh = fmaf(dx * dc, fmaf(-w, w, w), fmaf(qp, c, t.y));
}

// Cross-lane moves within a row of 16 lanes (DPP: no LDS traffic).
// quad_perm [1,0,3,2] and [2,3,0,1] swap lanes 2i <-> 2i+1 and i <-> i^2;
// row_ror:4 and row_ror:8 rotate a row of 16 lanes.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppRor4 = 0x124, kDppRor8 = 0x128;

template <int CTRL>
__device__ inline float dpp_f(float v)
{
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ inline double dpp_d(double v)
{
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_mov_dpp((int) (b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_mov_dpp((int) (b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long) hi << 32) | (unsigned int) lo);
}

// Sum over the 64 lanes, the result in every lane (fixed order: the same
// lanes combine in the same order on every call).
__device__ inline double wave_sum_d(double v)
{
  v += dpp_d<kDppXor1>(v);
  v += dpp_d<kDppXor2>(v);
  v += dpp_d<kDppRor4>(v);
  v += dpp_d<kDppRor8>(v);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

// The same sum when only lanes that are multiples of LP (1, 2, 4 or 8) hold
// non-zero values; the total is exact in lane 0 (the first steps, which
// would only add zeros there, are skipped).
template <int LP>
__device__ inline double wave_sum_d_lp(double v)
{
  if constexpr (LP < 2) v += dpp_d<kDppXor1>(v);
  if constexpr (LP < 4) v += dpp_d<kDppXor2>(v);
  if constexpr (LP < 8) v += dpp_d<kDppRor4>(v);
  v += dpp_d<kDppRor8>(v);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

__device__ inline unsigned lane_rank(unsigned long long m)
{
  return __builtin_amdgcn_mbcnt_hi((unsigned) (m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned) m, 0u));
}

// Phase 1: compact the kept entries (slot-sorted).  start[b] .. start[b+1].
__device__ inline int kept_slot(const ModelGeo &g, const Box &cell, const double4 &xi, int ei,
                                uint32_t e, const double4 &xj)
{
  double d[3];
  displacement(xi, xj, e >> kNbrJBits, cell, d);
  double r2 = dist2(d);
  return (r2 < g.rc2 && r2 > 1e-10) ? (int) ((g.slots[ei] >> (4 * (int) xj.w)) & 15u)
                                    : -1;  // kernel.cu:1163
}

// Lists of at most 64 NCH entries: every entry's slot stays in registers.
// All loads are unconditional, with indices clamped into the row (and to
// atom 0 past its end), and issued before any is used: exec-masked loads
// make the compiler wait for each one at its branch join.
template <int NCH>
__device__ inline int compact_short(const ModelGeo &g, const Box &cell, const double4 &xi, int ei,
                                    const uint32_t *list, int nraw, const double4 *pos,
                                    uint32_t *kept, int *start)
{
  const int lane = threadIdx.x & 63;
  const int last = max(nraw - 1, 0);
  uint32_t ent[NCH];
  double4 xj[NCH];
  int slot[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    const int idx = ch * 64 + lane;
    const uint32_t e = list[min(idx, last)];
    ent[ch] = idx < nraw ? e : 0u;
  }
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) xj[ch] = pos[ent[ch] & kNbrJMask];
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    const int sl = kept_slot(g, cell, xi, ei, ent[ch], xj[ch]);
    slot[ch] = ch * 64 + lane < nraw ? sl : -1;
  }
  int s = 0;
  for (int b = 0; b < g.nelt; b++) {
    start[b] = s;
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) s += __popcll(__ballot(slot[ch] == b));
  }
  start[g.nelt] = s;
  for (int b = 0; b < g.nelt; b++) {
    int run = start[b];
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) {
      unsigned long long m = __ballot(slot[ch] == b);
      if (slot[ch] == b) kept[run + (int) lane_rank(m)] = ent[ch];
      run += __popcll(m);
    }
  }
  return s;
}

__device__ inline int compact(const ModelGeo &g, const Box &cell, const double4 &xi, int ei,
                              const uint32_t *list, int nraw, const double4 *pos, uint32_t *kept,
                              int *start)
{
  const int lane = threadIdx.x & 63;
  if (nraw <= 128) return compact_short<2>(g, cell, xi, ei, list, nraw, pos, kept, start);
  if (nraw <= 192) return compact_short<3>(g, cell, xi, ei, list, nraw, pos, kept, start);
  if (nraw <= 256) return compact_short<4>(g, cell, xi, ei, list, nraw, pos, kept, start);
  // long lists: two sweeps
  int cnt[kMaxElts];
  for (int b = 0; b < kMaxElts; b++) cnt[b] = 0;
  for (int base = 0; base < nraw; base += 64) {
    int idx = base + lane, slot = -1;
    if (idx < nraw) {
      uint32_t e = list[idx];
      slot = kept_slot(g, cell, xi, ei, e, pos[e & kNbrJMask]);
    }
    for (int b = 0; b < g.nelt; b++) cnt[b] += __popcll(__ballot(slot == b));
  }
  int s = 0;
  for (int b = 0; b < g.nelt; b++) {
    start[b] = s;
    s += cnt[b];
  }
  start[g.nelt] = s;
  int run[kMaxElts];
  for (int b = 0; b < kMaxElts; b++) run[b] = start[b];
  for (int base = 0; base < nraw; base += 64) {
    int idx = base + lane, slot = -1;
    uint32_t e = 0;
    if (idx < nraw) {
      e = list[idx];
      slot = kept_slot(g, cell, xi, ei, e, pos[e & kNbrJMask]);
    }
    for (int b = 0; b < g.nelt; b++) {
      unsigned long long m = __ballot(slot == b);
      if (slot == b) kept[run[b] + (int) lane_rank(m)] = e;
      run[b] += __popcll(m);
    }
  }
  return s;
}

// 1/sqrt(r2) in FP64: hardware estimate plus one Newton step (relative
// error ~1e-15, far below the FP32 arithmetic that uses it)
__device__ inline double rsqrt_fp64(double r2)
{
  double y = __builtin_amdgcn_rsq(r2);
  return y * (1.5 - 0.5 * r2 * y * y);
}

// Geometry of one kept neighbor: FP64 displacement, unit vector and r, and
// the radial-table cell n with the offsets q, dx - q (ration.cu:29-32).
struct NbrGeo {
  double d[3];
  double r;
  float ux, uy, uz;
  float q, pq;
  int n, j;
};

__device__ inline void nbr_geometry(const ModelGeo &g, const Box &cell, const double4 &xi,
                                    const double4 *pos, uint32_t e, NbrGeo &o)
{
  o.j = (int) (e & kNbrJMask);
  displacement(xi, pos[o.j], e >> kNbrJBits, cell, o.d);
  double r2 = o.d[0] * o.d[0] + o.d[1] * o.d[1] + o.d[2] * o.d[2];
  double rinv = rsqrt_fp64(r2);
  o.r = r2 * rinv;
  o.ux = (float) (o.d[0] * rinv);
  o.uy = (float) (o.d[1] * rinv);
  o.uz = (float) (o.d[2] * rinv);
  o.n = min(g.tab_n - 2, (int) (o.r * g.xdx));
  double qd = o.r - (double) o.n * g.dx;
  o.q = (float) qd;
  o.pq = (float) (g.dx - qd);
}

// the end cells of the table are zero (ration.hpp:29)
__device__ inline bool table_cell_live(const ModelGeo &g, int n) { return n > 0 && n < g.tab_n - 2; }

// Lanes per neighbor for a chunk when `rem` kept neighbors remain: 2 (a
// chunk of 32 neighbors), 4 (16) or 8 (8).  A short last chunk spreads each
// neighbor's filters over more lanes instead of leaving lanes idle.
__device__ inline int lanes_per_nbr(int rem) { return rem > 16 ? 2 : (rem > 8 ? 4 : 8); }

}  // namespace

}  // namespace tmd
