// Device-resident periodic systems: the LAMMPS step loop around the pair
// style (Verlet::run, src/verlet.cpp:246-350; FixNVE, src/fix_nve.cpp:64-130;
// Neighbor::decide, src/neighbor.cpp:2293-2360) with positions, velocities,
// forces and the neighbor list kept on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/tensormd_amd.h"
#include "tmd_device.h"
#include "tmd_internal.h"

using namespace tmd;

struct tmd_system {
  tmd_ctx *ctx = nullptr;
  int n = 0;
  double skin = 0;
  CellInfo cell{};
  NbrBuild nb{};
  int nbins = 0;
  Workspace w;
  std::vector<int> row2atom;
  DeviceBuffer vel, dtfm, mass, bins, atominfo, xbuild, ctl, hist;
  DeviceBuffer snap;  // x, v, f at the start of the current run segment
  DeviceBuffer order;  // rows in atom order (the descriptor and force kernels' launch order)
  bool have_forces = false;
  double dt_cached = -1;
  int nbuilds_seen = 0;
  // one captured MD step (velocity Verlet + neighbor check + forces + history)
  hipGraphExec_t step_graph = nullptr;
  bool graph_hist = false;
  bool use_graph = true;
  // sampled kernel timing inside tmd_md_run (tmd_system_set_kernel_timing):
  // every time_every-th step is launched kernel by kernel with events around
  // the descriptor, MLP and force kernels
  int time_every = 0;
  std::vector<PhaseEvents> tev;  // one set per sampled step of a run segment
  DeviceBuffer tkept;            // kept-neighbor count summed over a segment's samples
  double tsum[3] = {0, 0, 0};    // descriptor, MLP, force kernel ms summed over the samples
  double tkept_sum = 0;          // kept neighbors summed over the samples (all atoms)
  long long tsamples = 0;
  // a step of tmd_md_step_async whose list-capacity check is still open:
  // the control words are copied into pinned memory behind it, and the
  // event marks the copy; the next call on the system resolves it
  bool pending = false;
  double pending_dt = 0;
  hipEvent_t pend_ev = nullptr;
  PinnedBuffer pend_ctl;
};

namespace tmd {

bool invert3(const double *h, double *inv)
{
  double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
      h[2] * (h[3] * h[7] - h[4] * h[6]);
  if (!(std::fabs(det) > 0)) return false;
  inv[0] = (h[4] * h[8] - h[5] * h[7]) / det;
  inv[1] = (h[2] * h[7] - h[1] * h[8]) / det;
  inv[2] = (h[1] * h[5] - h[2] * h[4]) / det;
  inv[3] = (h[5] * h[6] - h[3] * h[8]) / det;
  inv[4] = (h[0] * h[8] - h[2] * h[6]) / det;
  inv[5] = (h[2] * h[3] - h[0] * h[5]) / det;
  inv[6] = (h[3] * h[7] - h[4] * h[6]) / det;
  inv[7] = (h[1] * h[6] - h[0] * h[7]) / det;
  inv[8] = (h[0] * h[4] - h[1] * h[3]) / det;
  return true;
}

// perpendicular widths of the cell
void widths(const double *h, double w[3])
{
  const double *A = h, *B = h + 3, *C = h + 6;
  auto cross = [](const double *u, const double *v, double *o) {
    o[0] = u[1] * v[2] - u[2] * v[1];
    o[1] = u[2] * v[0] - u[0] * v[2];
    o[2] = u[0] * v[1] - u[1] * v[0];
  };
  double bc[3], ca[3], ab[3];
  cross(B, C, bc);
  cross(C, A, ca);
  cross(A, B, ab);
  double V = std::fabs(A[0] * bc[0] + A[1] * bc[1] + A[2] * bc[2]);
  w[0] = V / std::sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
  w[1] = V / std::sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]);
  w[2] = V / std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
}

// Bins and stencil reach of one periodic cell for the list cutoff `cut`;
// bins are cells of the fractional lattice at least cut/2 wide.
int setup_cell(const double *h, double cut, CellInfo &c, std::string &err)
{
  std::memset(&c, 0, sizeof(c));
  std::memcpy(c.h, h, 9 * sizeof(double));
  if (!invert3(c.h, c.inv)) {
    err = "singular cell";
    return TMD_ERR_ARG;
  }
  double w[3];
  widths(c.h, w);
  long nbins = 1;
  for (int a = 0; a < 3; a++) {
    c.nb[a] = std::max(1, (int) std::min(1024.0, std::floor(w[a] / (0.5 * cut))));
    double alpha = cut * c.nb[a] / w[a];
    c.reach[a] = (int) std::ceil(alpha + 1e-6);
    // image shifts are encoded in [-3, 3]: |s| <= ceil(cut / w) + 1
    if (std::ceil(cut / w[a]) + 1 > 3) {
      err = "cell narrower than (rc+skin)/2; replicate it first";
      return TMD_ERR_ARG;
    }
    nbins *= c.nb[a];
  }
  if (nbins > (1 << 26)) {
    err = "too many bins";
    return TMD_ERR_CAPACITY;
  }
  return TMD_OK;
}

double cell_volume(const double *h)
{
  return std::fabs(h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                   h[2] * (h[3] * h[7] - h[4] * h[6]));
}

}  // namespace tmd

namespace {

int sys_fail(tmd_system *s, int code, const std::string &m)
{
  set_err(s ? s->ctx : nullptr, m);
  return code;
}

int alloc_list(tmd_system *s, int cap)
{
  hipError_t e;
  s->nb.cap = cap;
  if ((e = s->w.nbr.ensure((size_t) s->n * cap * 4)) != hipSuccess) return hip_fail(s->ctx, e, "neighbor list alloc");
  return ensure_kept(s->ctx, s->w, s->w.rl.nrows, cap);
}

// One force evaluation at the current positions (neighbor list maintained).
Chain sys_chain(tmd_system *s)
{
  Chain c;
  c.geo.cells = (const CellInfo *) s->w.cells.p;
  c.geo.order = (const int *) s->order.p;
  c.pos = (const double4 *) s->w.pos.p;
  c.nbr = (const uint32_t *) s->w.nbr.p;
  c.nn = (const int *) s->w.nn.p;
  c.cap = s->nb.cap;
  c.f = (double *) s->w.f.p;
  c.fbytes = sizeof(double) * 3 * (size_t) s->n;
  c.natoms = s->n;
  c.g_bf16 = true;
  return c;
}

int sys_forces(tmd_system *s, PhaseEvents *pe = nullptr)
{
  hipStream_t st = s->ctx->stream;
  hipError_t e = nbr_step(s->nb, (double4 *) s->w.pos.p, (uint32_t *) s->w.nbr.p, (int *) s->w.nn.p, st);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "neighbor build");
  int rc = run_chain(s->ctx, s->w, sys_chain(s), st, pe);
  if (rc != TMD_OK || !pe) return rc;
  // the kept-neighbor counts of this evaluation, for the sampled work per launch
  e = launch_sum_kept((const uint32_t *) s->w.kl.p, kept_cache_words(s->nb.cap),
                      (const int *) s->w.row2atom.p, s->w.rl.nrows, (unsigned long long *) s->tkept.p, st);
  return e == hipSuccess ? TMD_OK : hip_fail(s->ctx, e, "kept count");
}

// the control words read after the queued work: an overflow grows the list
// and asks for a rebuild
int apply_overflow(tmd_system *s, const unsigned long long *ctl, bool *overflow)
{
  s->nbuilds_seen = (int) ctl[3];
  *overflow = ctl[2] != 0;
  if (*overflow) {
    int need = (int) ctl[2];
    int cap = ((need + need / 4 + 16) + 7) & ~7;
    if (s->step_graph) {
      hipGraphExecDestroy(s->step_graph);
      s->step_graph = nullptr;
    }
    int rc = alloc_list(s, cap);
    if (rc != TMD_OK) return rc;
    unsigned long long reset[3] = {0ull, 1ull, 0ull};
    if (hipError_t ce = hipMemcpy(s->ctl.p, reset, sizeof(reset), hipMemcpyHostToDevice);
        ce != hipSuccess)
      return hip_fail(s->ctx, ce, "list control reset");
  }
  return TMD_OK;
}

// overflow check after the queued work (waits for it)
int check_overflow(tmd_system *s, bool *overflow)
{
  unsigned long long ctl[4];
  hipError_t e = hipMemcpyAsync(ctl, s->ctl.p, sizeof(ctl), hipMemcpyDeviceToHost, s->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "device step");
  return apply_overflow(s, ctl, overflow);
}

int md_run_impl(tmd_system *s, int nsteps, double dt, double *epot, double *ekin, bool defer);

// Closes the open step of tmd_md_step_async: waits for it and, if its list
// overflowed, restores the snapshot taken before it and runs it again with
// the grown list.
int resolve_pending(tmd_system *s)
{
  if (!s->pending) return TMD_OK;
  s->pending = false;
  if (hipError_t e = hipEventSynchronize(s->pend_ev); e != hipSuccess) return hip_fail(s->ctx, e, "device step");
  unsigned long long ctl[4];
  std::memcpy(ctl, s->pend_ctl.p, sizeof(ctl));
  bool over = false;
  int rc = apply_overflow(s, ctl, &over);
  if (rc != TMD_OK || !over) return rc;
  const size_t xb = sizeof(double4) * (size_t) s->n, vb = 24 * (size_t) s->n;
  unsigned char *snap = (unsigned char *) s->snap.p;
  hipStream_t st = s->ctx->stream;
  hipError_t e = hipMemcpyAsync(s->w.pos.p, snap, xb, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->vel.p, snap + xb, vb, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->w.f.p, snap + xb + vb, vb, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "snapshot restore");
  return md_run_impl(s, 1, s->pending_dt, nullptr, nullptr, false);
}

}  // namespace

extern "C" {

int tmd_system_create(tmd_ctx *ctx, int natoms, const double *x, const int *type,
                      const double *cellv, double skin, tmd_system **out)
{
  if (!ctx || !out || !x || !type || !cellv || natoms < 1 || !(skin >= 0)) {
    set_err(ctx, "tmd_system_create: invalid arguments");
    return TMD_ERR_ARG;
  }
  *out = nullptr;
  if (natoms > kMaxAtoms) {
    set_err(ctx, "tmd_system_create: more atoms than the 2^23-1 limit of the neighbor encoding");
    return TMD_ERR_CAPACITY;
  }
  if (hipError_t de = hipSetDevice(ctx->device); de != hipSuccess)
    return hip_fail(ctx, de, "hipSetDevice");
  std::unique_ptr<tmd_system> s(new tmd_system());
  s->ctx = ctx;
  s->n = natoms;
  s->skin = skin;
  const Model &m = ctx->model;
  double cut = m.rc + skin;
  {
    std::string err;
    int rc = setup_cell(cellv, cut, s->cell, err);
    if (rc != TMD_OK) {
      set_err(ctx, "tmd_system_create: " + err);
      return rc;
    }
  }
  s->nbins = s->cell.nb[0] * s->cell.nb[1] * s->cell.nb[2];
  s->nb.nbins = s->nbins;
  s->nb.n = natoms;
  s->nb.cut2 = cut * cut;
  s->nb.trig2 = 0.25 * skin * skin;
  // expected neighbors within rc+skin, with margin for disorder
  double V = cell_volume(s->cell.h);
  double expect = 4.0 / 3.0 * M_PI * cut * cut * cut * natoms / V;
  // thermal motion widens the local-density spread: ~1.5x the mean plus a
  // floor; an overflow during a run is still caught and the run rolled back
  int cap = (int) std::ceil(expect * 1.5 + 16.0);
  cap = (cap + 7) & ~7;

  std::vector<double4> pos(natoms);
  std::vector<int> elem(natoms);
  std::vector<int> atoms(natoms);
  for (int i = 0; i < natoms; i++) {
    int t = type[i];
    int e = (t >= 1 && t <= ctx->ntypes) ? ctx->type_map[t] : -1;
    if (e < 0) {
      set_err(ctx, "tmd_system_create: atom type not mapped to a model element");
      return TMD_ERR_ARG;
    }
    elem[i] = e;
    atoms[i] = i;
    pos[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], (double) e);
  }
  make_rows(m.nelt, atoms, elem, s->w.rl, s->row2atom);
  int rc = ensure_rows(ctx, s->w, s->w.rl.nrows);
  if (rc != TMD_OK) return rc;
  {
    // the descriptor and force kernels take the rows in atom order (the
    // element blocks interleaved), padding rows last: for atoms given in
    // spatial order, consecutive workgroups gather the same neighbors
    // (1.2% on the 128k-atom step)
    std::vector<int> ord;
    ord.reserve(s->w.rl.nrows);
    for (int r = 0; r < s->w.rl.nrows; r++)
      if (s->row2atom[r] >= 0) ord.push_back(r);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return s->row2atom[a] < s->row2atom[b]; });
    for (int r = 0; r < s->w.rl.nrows; r++)
      if (s->row2atom[r] < 0) ord.push_back(r);
    hipError_t oe = s->order.ensure(4 * std::max<size_t>(ord.size(), 1));
    if (oe == hipSuccess) oe = hipMemcpy(s->order.p, ord.data(), 4 * ord.size(), hipMemcpyHostToDevice);
    if (oe != hipSuccess) return hip_fail(ctx, oe, "row order upload");
  }
  hipError_t e;
  auto chk = [&](hipError_t er, const char *what) -> int {
    if (er != hipSuccess) return hip_fail(ctx, er, what);
    return TMD_OK;
  };
  if ((rc = ensure_atoms(ctx, s->w, natoms))) return rc;
  if ((rc = chk(s->w.cells.ensure(sizeof(CellInfo)), "cell alloc"))) return rc;
  if ((rc = chk(s->w.pos.ensure(sizeof(double4) * natoms), "pos alloc"))) return rc;
  if ((rc = chk(s->w.nn.ensure(4 * (size_t) natoms), "nn alloc"))) return rc;
  if ((rc = chk(s->w.f.ensure(sizeof(double) * 3 * (size_t) natoms), "f alloc"))) return rc;
  if ((rc = chk(s->vel.ensure(sizeof(double) * 3 * (size_t) natoms), "v alloc"))) return rc;
  if ((rc = chk(s->dtfm.ensure(sizeof(double) * kMaxElts), "dtfm alloc"))) return rc;
  if ((rc = chk(s->mass.ensure(sizeof(double) * kMaxElts), "mass alloc"))) return rc;
  if ((rc = chk(s->bins.ensure(sizeof(int) * (3 * (size_t) (s->nbins + 1) + 1)), "bins alloc"))) return rc;
  if ((rc = chk(s->atominfo.ensure(sizeof(int) * 5 * (size_t) natoms), "atom bins alloc"))) return rc;
  if ((rc = chk(s->xbuild.ensure(sizeof(double) * 4 * (size_t) natoms), "xbuild alloc"))) return rc;
  if ((rc = chk(s->ctl.ensure(sizeof(unsigned long long) * 8), "ctl alloc"))) return rc;
  if ((rc = chk(s->hist.ensure(sizeof(double) * (2 * 4096 + 4096 + 64)), "hist alloc"))) return rc;
  if ((rc = alloc_list(s.get(), cap))) return rc;
  int *bins = (int *) s->bins.p;
  s->nb.bin_count = bins;
  s->nb.bin_start = bins + (s->nbins + 1);
  s->nb.bin_cursor = bins + 2 * (s->nbins + 1);
  int *ai = (int *) s->atominfo.p;
  s->nb.bin_atoms = ai;
  s->nb.atom_bin = ai + natoms;
  s->nb.atom_bc = ai + 2 * (size_t) natoms;
  s->nb.x_build = (double *) s->xbuild.p;
  s->nb.ctl = (unsigned long long *) s->ctl.p;
  s->nb.cells = (const CellInfo *) s->w.cells.p;
  s->nb.atom_struct = nullptr;
  hipStream_t st = ctx->stream;
  std::vector<double> masses(kMaxElts, 1.0);
  for (int el = 0; el < m.nelt; el++) masses[el] = m.masses[el];
  unsigned long long ctl0[8] = {0ull, 1ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull};  // force first build
  if ((e = hipMemcpyAsync(s->w.pos.p, pos.data(), sizeof(double4) * natoms, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(s->w.row2atom.p, s->row2atom.data(), 4 * s->row2atom.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(s->mass.p, masses.data(), sizeof(double) * kMaxElts, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(s->ctl.p, ctl0, sizeof(ctl0), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(s->w.cells.p, &s->cell, sizeof(CellInfo), hipMemcpyHostToDevice, st)) != hipSuccess)
    return hip_fail(ctx, e, "system upload");
  if ((e = launch_zero(s->vel.p, sizeof(double) * 3 * (size_t) natoms, st)) != hipSuccess) return hip_fail(ctx, e, "zero v");
  if ((e = launch_zero(s->bins.p, s->bins.bytes, st)) != hipSuccess) return hip_fail(ctx, e, "zero bins");
  if ((e = launch_zero(s->w.nn.p, 4 * (size_t) natoms, st)) != hipSuccess) return hip_fail(ctx, e, "zero nn");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(ctx, e, "system create");
  *out = s.release();
  return TMD_OK;
}

void tmd_system_destroy(tmd_system *s)
{
  if (!s) return;
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->ctx->stream);
  if (s->step_graph) hipGraphExecDestroy(s->step_graph);
  if (s->pend_ev) hipEventDestroy(s->pend_ev);
  for (auto &pe : s->tev)
    for (auto &ev : pe.ev) hipEventDestroy(ev);
  delete s;
}

int tmd_system_set_positions(tmd_system *s, const double *x)
{
  if (!s || !x) return TMD_ERR_ARG;
  if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess)
    return hip_fail(s->ctx, de, "hipSetDevice");
  if (int rc = resolve_pending(s); rc != TMD_OK) return rc;
  std::vector<double4> pos(s->n);
  hipStream_t st = s->ctx->stream;
  hipError_t e = hipMemcpyAsync(pos.data(), s->w.pos.p, sizeof(double4) * s->n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "set positions");
  for (int i = 0; i < s->n; i++) {
    pos[i].x = x[3 * i];
    pos[i].y = x[3 * i + 1];
    pos[i].z = x[3 * i + 2];
  }
  unsigned long long force_build = 1ull;
  e = hipMemcpyAsync(s->w.pos.p, pos.data(), sizeof(double4) * s->n, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync((unsigned long long *) s->ctl.p + 1, &force_build, 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "set positions");
  s->have_forces = false;
  return TMD_OK;
}

int tmd_system_set_velocities(tmd_system *s, const double *v)
{
  if (!s || !v) return TMD_ERR_ARG;
  if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess)
    return hip_fail(s->ctx, de, "hipSetDevice");
  if (int rc = resolve_pending(s); rc != TMD_OK) return rc;
  hipError_t e = hipMemcpy(s->vel.p, v, sizeof(double) * 3 * (size_t) s->n, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "set velocities");
  return TMD_OK;
}

int tmd_system_compute(tmd_system *s)
{
  if (!s) return TMD_ERR_ARG;
  if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess)
    return hip_fail(s->ctx, de, "hipSetDevice");
  if (int rc = resolve_pending(s); rc != TMD_OK) return rc;
  for (int attempt = 0; attempt < 4; attempt++) {
    int rc = sys_forces(s);
    if (rc != TMD_OK) return rc;
    bool over = false;
    rc = check_overflow(s, &over);
    if (rc != TMD_OK) return rc;
    if (!over) {
      s->have_forces = true;
      return TMD_OK;
    }
  }
  return sys_fail(s, TMD_ERR_CAPACITY, "neighbor list capacity could not be satisfied");
}

int tmd_system_get(tmd_system *s, double *x, double *v, double *f, double *eatom, double *energy,
                   double *virial)
{
  if (!s) return TMD_ERR_ARG;
  if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess)
    return hip_fail(s->ctx, de, "hipSetDevice");
  if (int rc = resolve_pending(s); rc != TMD_OK) return rc;
  hipStream_t st = s->ctx->stream;
  std::vector<double4> pos;
  double red[7];
  hipError_t e = hipSuccess;
  if (x) {
    pos.resize(s->n);
    e = hipMemcpyAsync(pos.data(), s->w.pos.p, sizeof(double4) * s->n, hipMemcpyDeviceToHost, st);
  }
  if (e == hipSuccess && v) e = hipMemcpyAsync(v, s->vel.p, sizeof(double) * 3 * (size_t) s->n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && f) e = hipMemcpyAsync(f, s->w.f.p, sizeof(double) * 3 * (size_t) s->n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && (energy || virial)) e = hipMemcpyAsync(red, s->w.red.p, sizeof(red), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && eatom)
    e = hipMemcpyAsync(eatom, s->w.ew.p, sizeof(double) * (size_t) s->n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "tmd_system_get");
  if (x)
    for (int i = 0; i < s->n; i++) {
      x[3 * i] = pos[i].x;
      x[3 * i + 1] = pos[i].y;
      x[3 * i + 2] = pos[i].z;
    }
  if (energy) *energy = red[0];
  if (virial)
    for (int q = 0; q < 6; q++) virial[q] = red[1 + q];
  return TMD_OK;
}

int tmd_md_run(tmd_system *s, int nsteps, double dt, double *epot, double *ekin)
{
  if (!s || nsteps < 0 || !(dt > 0)) return TMD_ERR_ARG;
  if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess)
    return hip_fail(s->ctx, de, "hipSetDevice");
  if (int rc = resolve_pending(s); rc != TMD_OK) return rc;
  return md_run_impl(s, nsteps, dt, epot, ekin, false);
}

int tmd_md_step_async(tmd_system *s, double dt)
{
  if (!s || !(dt > 0)) return TMD_ERR_ARG;
  if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess)
    return hip_fail(s->ctx, de, "hipSetDevice");
  if (int rc = resolve_pending(s); rc != TMD_OK) return rc;
  // sampled kernel timing needs the synchronous loop
  return md_run_impl(s, 1, dt, nullptr, nullptr, s->time_every == 0);
}

}  // extern "C"

namespace {

// nsteps of velocity Verlet in run segments; defer (one step, no history):
// the step's capacity check is left open (resolve_pending) instead of
// waiting for it
int md_run_impl(tmd_system *s, int nsteps, double dt, double *epot, double *ekin, bool defer)
{
  hipStream_t st = s->ctx->stream;
  const Model &m = s->ctx->model;
  hipError_t e;
  if (dt != s->dt_cached) {
    const double ftm2v = 1.0 / 1.0364269e-4;  // metal units
    std::vector<double> dtfm(kMaxElts, 0.0);
    for (int el = 0; el < m.nelt; el++) dtfm[el] = 0.5 * dt * ftm2v / m.masses[el];
    if ((e = hipMemcpy(s->dtfm.p, dtfm.data(), sizeof(double) * kMaxElts, hipMemcpyHostToDevice)) != hipSuccess)
      return hip_fail(s->ctx, e, "dtfm upload");
    s->dt_cached = dt;
    if (s->step_graph) {  // dt is baked into the captured launch
      hipGraphExecDestroy(s->step_graph);
      s->step_graph = nullptr;
    }
  }
  if (!s->have_forces) {
    int rc = tmd_system_compute(s);
    if (rc != TMD_OK) return rc;
  }
  double *hist = (double *) s->hist.p;
  const bool want_hist = epot || ekin;
  // layout of the history buffer: [0, 2*kHist) pairs, then scratch
  const int kHist = 4096;
  double *ekin_out = hist + 2 * kHist;
  unsigned long long *counter = (unsigned long long *) (hist + 2 * kHist + 4096);
  auto enqueue_step = [&](PhaseEvents *pe) -> int {
    hipError_t er;
    if ((er = launch_vv_first(s->n, (double4 *) s->w.pos.p, (double *) s->vel.p, (const double *) s->w.f.p,
                              (const double *) s->dtfm.p, dt, st)) != hipSuccess)
      return hip_fail(s->ctx, er, "vv first");
    int rc = sys_forces(s, pe);
    if (rc != TMD_OK) return rc;
    if ((er = launch_vv_second(s->n, (const double4 *) s->w.pos.p, (double *) s->vel.p,
                               (const double *) s->w.f.p, (const double *) s->dtfm.p, st)) != hipSuccess)
      return hip_fail(s->ctx, er, "vv second");
    if (want_hist) {
      if ((er = launch_ekin(s->n, (const double4 *) s->w.pos.p, (const double *) s->vel.p,
                            (const double *) s->mass.p, ekin_out, st)) != hipSuccess)
        return hip_fail(s->ctx, er, "ekin");
      if ((er = launch_record_step((const double *) s->w.red.p, ekin_out, hist, counter, kHist, st)) != hipSuccess)
        return hip_fail(s->ctx, er, "history");
    }
    return TMD_OK;
  };
  if (s->use_graph && (!s->step_graph || s->graph_hist != want_hist)) {
    if (s->step_graph) {
      hipGraphExecDestroy(s->step_graph);
      s->step_graph = nullptr;
    }
    hipGraph_t graph;
    if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hip_fail(s->ctx, e, "graph capture");
    int rc = enqueue_step(nullptr);
    e = hipStreamEndCapture(st, &graph);
    if (rc != TMD_OK) return rc;
    if (e != hipSuccess) return hip_fail(s->ctx, e, "graph capture");
    e = hipGraphInstantiate(&s->step_graph, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
      s->step_graph = nullptr;
      return hip_fail(s->ctx, e, "graph instantiate");
    }
    s->graph_hist = want_hist;
  }
  // Run in segments.  Each starts from a device snapshot of x, v and f; a
  // neighbor-list overflow inside a segment (a truncated list) grows the
  // list and reruns the segment from the snapshot, so no step ever uses a
  // truncated list.
  const int kSeg = 200;
  const int every = s->time_every;
  if (every > 0) {
    const size_t need = (size_t) (kSeg + every - 1) / every;
    while (s->tev.size() < need) {
      PhaseEvents pe;
      for (auto &ev : pe.ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return hip_fail(s->ctx, e, "timing events");
      s->tev.push_back(pe);
    }
    if ((e = s->tkept.ensure(sizeof(unsigned long long))) != hipSuccess)
      return hip_fail(s->ctx, e, "timing counter");
  }
  const size_t xb = sizeof(double4) * (size_t) s->n, vb = 24 * (size_t) s->n;
  if ((e = s->snap.ensure(xb + 2 * vb)) != hipSuccess) return hip_fail(s->ctx, e, "snapshot alloc");
  unsigned char *snap = (unsigned char *) s->snap.p;
  auto copy3 = [&](bool save) -> hipError_t {
    hipError_t er;
    void *a[3] = {s->w.pos.p, s->vel.p, s->w.f.p};
    void *b[3] = {snap, snap + xb, snap + xb + vb};
    size_t n[3] = {xb, vb, vb};
    for (int q = 0; q < 3; q++)
      if ((er = hipMemcpyAsync(save ? b[q] : a[q], save ? a[q] : b[q], n[q], hipMemcpyDeviceToDevice, st)) !=
          hipSuccess)
        return er;
    return hipSuccess;
  };
  for (int done = 0; done < nsteps;) {
    int nrun = std::min(kSeg, nsteps - done);
    if ((e = copy3(true)) != hipSuccess) return hip_fail(s->ctx, e, "snapshot");
    bool over = true;
    if (defer) {
      // one step behind its snapshot; the control words follow it into
      // pinned memory, read by resolve_pending
      if (!s->pend_ev && (e = hipEventCreateWithFlags(&s->pend_ev, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(s->ctx, e, "step event");
      if ((e = s->pend_ctl.ensure(4 * sizeof(unsigned long long))) != hipSuccess)
        return hip_fail(s->ctx, e, "step control staging");
      if (s->use_graph)
        e = hipGraphLaunch(s->step_graph, st);
      else if (int rc = enqueue_step(nullptr); rc != TMD_OK)
        return rc;
      if (e == hipSuccess)
        e = hipMemcpyAsync(s->pend_ctl.p, s->ctl.p, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipEventRecord(s->pend_ev, st);
      if (e != hipSuccess) return hip_fail(s->ctx, e, "step launch");
      s->pending = true;
      s->pending_dt = dt;
      return TMD_OK;
    }
    for (int attempt = 0; attempt < 4 && over; attempt++) {
      if (want_hist) {
        // stream-ordered clear (this loop is never captured into a graph)
        if ((e = hipMemsetAsync(counter, 0, sizeof(unsigned long long), st)) != hipSuccess)
          return hip_fail(s->ctx, e, "history reset");
      }
      if (every > 0) {
        if ((e = hipMemsetAsync(s->tkept.p, 0, sizeof(unsigned long long), st)) != hipSuccess)
          return hip_fail(s->ctx, e, "timing reset");
      }
      int nsamp = 0;
      for (int q = 0; q < nrun; q++) {
        if (every > 0 && (done + q) % every == 0) {  // a sampled step, kernel by kernel
          int rc = enqueue_step(&s->tev[(size_t) nsamp++]);
          if (rc != TMD_OK) return rc;
        } else if (s->use_graph) {
          if ((e = hipGraphLaunch(s->step_graph, st)) != hipSuccess) return hip_fail(s->ctx, e, "graph launch");
        } else {
          int rc = enqueue_step(nullptr);
          if (rc != TMD_OK) return rc;
        }
      }
      int rc = check_overflow(s, &over);  // grows the list and requests a rebuild
      if (rc != TMD_OK) return rc;
      if (!over && nsamp > 0) {  // the segment stands: keep its samples (the stream is idle)
        hipError_t te = hipSuccess;
        auto el = [&](hipEvent_t a, hipEvent_t b) {
          float t = 0.f;
          if (te == hipSuccess) te = hipEventElapsedTime(&t, a, b);
          return (double) t;
        };
        for (int j = 0; j < nsamp; j++) {
          const PhaseEvents &pe = s->tev[(size_t) j];
          s->tsum[0] += el(pe.ev[0], pe.ev[1]);
          s->tsum[1] += el(pe.ev[1], pe.ev[2]);
          s->tsum[2] += el(pe.ev[3], pe.ev[4]);
        }
        if (te != hipSuccess) return hip_fail(s->ctx, te, "kernel timing");
        unsigned long long kept = 0;
        if ((e = hipMemcpy(&kept, s->tkept.p, sizeof(kept), hipMemcpyDeviceToHost)) != hipSuccess)
          return hip_fail(s->ctx, e, "timing counter copy");
        s->tkept_sum += (double) kept;
        s->tsamples += nsamp;
      }
      if (over) {
        if ((e = copy3(false)) != hipSuccess) return hip_fail(s->ctx, e, "snapshot restore");
        if (s->use_graph && !s->step_graph) {  // the list grew: capture again
          hipGraph_t graph;
          if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return hip_fail(s->ctx, e, "graph capture");
          rc = enqueue_step(nullptr);
          e = hipStreamEndCapture(st, &graph);
          if (rc != TMD_OK) return rc;
          if (e != hipSuccess) return hip_fail(s->ctx, e, "graph capture");
          e = hipGraphInstantiate(&s->step_graph, graph, nullptr, nullptr, 0);
          hipGraphDestroy(graph);
          if (e != hipSuccess) {
            s->step_graph = nullptr;
            return hip_fail(s->ctx, e, "graph instantiate");
          }
        }
      }
    }
    if (over) return sys_fail(s, TMD_ERR_CAPACITY, "neighbor list overflow persists after growing the list");
    if (want_hist) {
      std::vector<double> h(2 * (size_t) nrun);
      if ((e = hipMemcpy(h.data(), hist, sizeof(double) * 2 * nrun, hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(s->ctx, e, "history copy");
      for (int q = 0; q < nrun; q++) {
        if (epot) epot[done + q] = h[2 * q];
        if (ekin) ekin[done + q] = h[2 * q + 1];
      }
    }
    done += nrun;
  }
  return TMD_OK;
}

}  // namespace

extern "C" {

int tmd_system_set_kernel_timing(tmd_system *s, int every)
{
  if (!s || every < 0) return TMD_ERR_ARG;
  s->time_every = every;
  s->tsum[0] = s->tsum[1] = s->tsum[2] = 0;
  s->tkept_sum = 0;
  s->tsamples = 0;
  return TMD_OK;
}

int tmd_system_kernel_timing(const tmd_system *s, double *ms, double *kept, long long *nsamples)
{
  if (!s || !ms) return TMD_ERR_ARG;
  const double n = s->tsamples > 0 ? (double) s->tsamples : 1.0;
  for (int k = 0; k < 3; k++) ms[k] = s->tsum[k] / n;
  if (kept) kept[0] = s->tkept_sum / (n * (double) s->n);
  if (nsamples) nsamples[0] = s->tsamples;
  return TMD_OK;
}

int tmd_system_set_graph(tmd_system *s, int enable)
{
  if (!s) return TMD_ERR_ARG;
  s->use_graph = enable != 0;
  return TMD_OK;
}


int tmd_md_profile(tmd_system *s, int nsteps, double dt, double *ms)
{
  if (!s || nsteps < 1 || nsteps > 100000 || !(dt > 0) || !ms) return TMD_ERR_ARG;
  int rc = tmd_md_run(s, 1, dt, nullptr, nullptr);  // forces, dt constants, warm
  if (rc != TMD_OK) return rc;
  hipStream_t st = s->ctx->stream;
  const int NE = 10;
  std::vector<hipEvent_t> ev((size_t) nsteps * NE, nullptr);
  struct EventsGuard {  // destroyed on every return path
    std::vector<hipEvent_t> &v;
    ~EventsGuard()
    {
      for (auto &x : v)
        if (x) hipEventDestroy(x);
    }
  } guard{ev};
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) return sys_fail(s, TMD_ERR_DEVICE, "hipEventCreate");
  hipError_t e = hipSuccess;
  for (int q = 0; q < nsteps && e == hipSuccess; q++) {
    hipEvent_t *E = &ev[(size_t) q * NE];
    PhaseEvents pe;
    for (int k = 0; k < 6; k++) pe.ev[k] = E[3 + k];
    e = hipEventRecord(E[0], st);
    if (e == hipSuccess)
      e = launch_vv_first(s->n, (double4 *) s->w.pos.p, (double *) s->vel.p, (const double *) s->w.f.p,
                          (const double *) s->dtfm.p, dt, st);
    if (e == hipSuccess) e = hipEventRecord(E[1], st);
    if (e == hipSuccess)
      e = nbr_step(s->nb, (double4 *) s->w.pos.p, (uint32_t *) s->w.nbr.p, (int *) s->w.nn.p, st);
    if (e == hipSuccess) e = hipEventRecord(E[2], st);
    if (e == hipSuccess) {
      rc = run_chain(s->ctx, s->w, sys_chain(s), st, &pe);
      if (rc != TMD_OK) return rc;
    }
    if (e == hipSuccess)
      e = launch_vv_second(s->n, (const double4 *) s->w.pos.p, (double *) s->vel.p,
                           (const double *) s->w.f.p, (const double *) s->dtfm.p, st);
    if (e == hipSuccess) e = hipEventRecord(E[9], st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "profile");
  // phases: neighbor(1-2), descriptor(3-4), mlp(4-5), force kernel(6-7),
  // reductions + force-buffer clears(7-8, 5-6), integrate(0-1, 8-9)
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  auto el = [&](hipEvent_t a, hipEvent_t b) {
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    return (double) t;
  };
  for (int q = 0; q < nsteps; q++) {
    hipEvent_t *E = &ev[(size_t) q * NE];
    acc[0] += el(E[1], E[2]);
    acc[1] += el(E[3], E[4]);
    acc[2] += el(E[4], E[5]);
    acc[3] += el(E[6], E[7]);
    acc[4] += el(E[7], E[8]) + el(E[5], E[6]);
    acc[5] += el(E[0], E[1]) + el(E[8], E[9]);
    acc[6] += el(E[0], E[9]);
  }
  if (e != hipSuccess) return hip_fail(s->ctx, e, "profile timing");
  for (int k = 0; k < 7; k++) ms[k] = acc[k] / nsteps;
  bool over = false;
  rc = check_overflow(s, &over);
  if (rc != TMD_OK) return rc;
  if (over) return sys_fail(s, TMD_ERR_CAPACITY, "neighbor list overflow during profiling");
  return TMD_OK;
}

int tmd_system_nbuilds(const tmd_system *s) { return s ? s->nbuilds_seen : 0; }

int tmd_system_neighbors(tmd_system *s, long cap, long *off, int *j, int *shift, long *total)
{
  if (!s || !off || !total) return TMD_ERR_ARG;
  if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess)
    return hip_fail(s->ctx, de, "hipSetDevice");
  if (int rc = resolve_pending(s); rc != TMD_OK) return rc;
  std::vector<int> nn(s->n);
  std::vector<uint32_t> list((size_t) s->n * s->nb.cap);
  hipError_t e = hipMemcpy(nn.data(), s->w.nn.p, 4 * (size_t) s->n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(list.data(), s->w.nbr.p, 4 * list.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "neighbor export");
  long t = 0;
  for (int i = 0; i < s->n; i++) {
    off[i] = t;
    for (int p = 0; p < nn[i]; p++, t++) {
      if (t < cap && j && shift) {
        uint32_t v = list[(size_t) i * s->nb.cap + p];
        uint32_t c = v >> kNbrJBits;
        j[t] = (int) (v & kNbrJMask);
        shift[3 * t] = (int) (c / 49) - 3;
        shift[3 * t + 1] = (int) ((c / 7) % 7) - 3;
        shift[3 * t + 2] = (int) (c % 7) - 3;
      }
    }
  }
  off[s->n] = t;
  *total = t;
  return TMD_OK;
}

int tmd_system_descriptors(tmd_system *s, float *G, float *dEdG)
{
  if (!s) return TMD_ERR_ARG;
  if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess)
    return hip_fail(s->ctx, de, "hipSetDevice");
  if (int rc = resolve_pending(s); rc != TMD_OK) return rc;
  int gw = s->ctx->model.gwidth();
  int nr = s->w.rl.nrows;
  std::vector<float> g((size_t) nr * gw), dg((size_t) nr * gw);
  hipError_t e;
  if (s->w.g_bf16) {  // BF16 rows: widen (exact)
    std::vector<uint16_t> gb(g.size());
    e = hipMemcpy(gb.data(), s->w.G.p, 2 * gb.size(), hipMemcpyDeviceToHost);
    for (size_t q = 0; q < gb.size(); q++) {
      const uint32_t bits = (uint32_t) gb[q] << 16;
      std::memcpy(&g[q], &bits, 4);
    }
  } else {
    e = hipMemcpy(g.data(), s->w.G.p, 4 * g.size(), hipMemcpyDeviceToHost);
  }
  if (e == hipSuccess) e = hipMemcpy(dg.data(), s->w.dEdG.p, 4 * dg.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(s->ctx, e, "descriptor export");
  for (int r = 0; r < nr; r++) {
    int a = s->row2atom[r];
    if (a < 0) continue;
    // in the model file's layout [b][k][max_moment_file + 1] (a model with
    // fewer moments is evaluated with its missing channels padded)
    const int nmf = s->ctx->model.max_moment_file + 1, nm = s->ctx->model.nm;
    const int gwf = s->ctx->model.gwidth_file();
    for (int q = 0; q < gwf; q++) {
      const size_t src = (size_t) r * gw + (size_t) (q / nmf) * nm + q % nmf;
      if (G) G[(size_t) a * gwf + q] = g[src];
      if (dEdG) dEdG[(size_t) a * gwf + q] = dg[src];
    }
  }
  return TMD_OK;
}

}  // extern "C"
