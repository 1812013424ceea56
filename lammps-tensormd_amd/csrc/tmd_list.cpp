// List mode: Pair::compute over a LAMMPS full neighbor list
// (PairTensorMD::compute, src/ML-TENSORMD/pair_tensormd.cpp:141-184, with the
// host pack of setup_tensors_GPU, gpu/cuda/kernel.cu:1455-1596, and the
// force download of sum_forces_GPU, kernel.cu:1760-1789).
//
// One call: pack the element-tagged positions and (unless the caller's list
// serial says the list is unchanged) the padded list and the element-sorted
// rows into host staging, upload, run descriptor -> MLP -> force ->
// reductions, download f, eatom, vatom and the sums.  The staging and device
// buffers are either the context's growable ones (pinned host staging,
// device buffers grown on demand) or a fixed area given by the caller
// (the no-allocation entry point).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "tmd_internal.h"

namespace tmd {

namespace {

size_t al256(size_t v) { return (v + 255) & ~(size_t) 255; }

// Sizes of one call, fixed before anything is staged.
struct ListPlan {
  int cap = 8;    // list row stride (max numneigh over ilist, rounded up to 8)
  int nrows = 0;  // MLP rows: per element, ceil(count / kRowTile) * kRowTile
  int cnt[kMaxElts] = {};
};

// Host staging: zero cell | pos[nall] | list[nlocal][cap] | nn[nlocal] |
// row2atom[nrows] | f[nall][3] (with stage_f: the forces come back into
// pinned memory at the full DMA rate and are copied out by host threads,
// instead of a copy into the caller's pageable array)
struct HostStage {
  CellInfo *cell;
  double4 *pos;
  uint32_t *list;
  int *nn;
  int *row2atom;
  double *f;  // null without stage_f
};

size_t host_stage_bytes(const ListArgs &a, const ListPlan &p, size_t *off, bool stage_f)
{
  off[0] = 0;
  off[1] = al256(sizeof(CellInfo));
  off[2] = off[1] + al256(sizeof(double4) * (size_t) std::max(a.nall, 1));
  off[3] = off[2] + al256(4 * (size_t) std::max(a.nlocal, 1) * p.cap);
  off[4] = off[3] + al256(4 * (size_t) std::max(a.nlocal, 1));
  off[5] = off[4] + al256(4 * (size_t) std::max(p.nrows, 1));
  return off[5] + (stage_f ? al256(3 * sizeof(double) * (size_t) std::max(a.nall, 1)) : 0);
}

HostStage carve_host(unsigned char *base, const size_t *off, bool stage_f)
{
  HostStage h;
  h.cell = (CellInfo *) (base + off[0]);
  h.pos = (double4 *) (base + off[1]);
  h.list = (uint32_t *) (base + off[2]);
  h.nn = (int *) (base + off[3]);
  h.row2atom = (int *) (base + off[4]);
  h.f = stage_f ? (double *) (base + off[5]) : nullptr;
  return h;
}

// Device buffers of one call in a fixed area, in the order of
// carve_device; returns the bytes needed.
size_t device_bytes(const ListArgs &a, const ListPlan &p, int gw, size_t *off)
{
  const size_t nall = (size_t) std::max(a.nall, 1), nloc = (size_t) std::max(a.nlocal, 1);
  const size_t rows = (size_t) std::max(p.nrows, 1);
  size_t sz[13] = {
      sizeof(double4) * nall,                     // pos
      4 * nloc * p.cap,                           // nbr
      4 * nloc,                                   // nn
      4 * rows,                                   // row2atom
      sizeof(CellInfo),                           // cells
      4 * rows * gw,                              // G
      4 * rows * gw,                              // dEdG
      8 * rows,                                   // E
      4096 * 8,                                   // red
      4 * rows * kept_cache_words(p.cap),         // kl
      7 * 8 * nloc,                               // ew
      3 * 8 * nall,                               // f
      a.vatom ? 6 * 8 * nall : 0,                 // vatom
  };
  size_t o = 0;
  for (int q = 0; q < 13; q++) {
    off[q] = o;
    o += al256(sz[q]);
  }
  off[13] = o;
  return o;
}

void carve_device(Workspace &w, unsigned char *base, const size_t *off)
{
  DeviceBuffer *b[13] = {&w.pos, &w.nbr, &w.nn, &w.row2atom, &w.cells, &w.G, &w.dEdG,
                         &w.E, &w.red, &w.kl, &w.ew, &w.f, &w.vatom};
  for (int q = 0; q < 13; q++) b[q]->set_view(base + off[q], off[q + 1] - off[q]);
}

// Host worker threads kept for the process (up to 15, started at the first
// call that has work for them), so that a list-mode call does not pay for
// creating threads.  run(n, job) calls job(0 .. n-1) on the workers and the
// calling thread and returns when all are done; concurrent callers take
// turns.
class HostPool {
 public:
  static HostPool &get()
  {
    static HostPool pool;
    return pool;
  }
  int size() const { return nworkers + 1; }
  void run(int n, const std::function<void(int)> &job)
  {
    std::lock_guard<std::mutex> turn(run_mu);
    {
      std::lock_guard<std::mutex> g(mu);
      cur = &job;
      njobs = n;
      next = 0;
      finished = 0;
      ++gen;
    }
    cv.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu);
    done_cv.wait(lk, [&] { return finished == njobs; });
    cur = nullptr;
  }
  ~HostPool()
  {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv.notify_all();
    for (auto &t : workers) t.join();
  }

 private:
  HostPool()
  {
    nworkers = (int) std::min<unsigned>(15u, std::max(1u, std::thread::hardware_concurrency()) - 1);
    for (int i = 0; i < nworkers; i++) workers.emplace_back([this] { loop(); });
  }
  // takes jobs of the current run until none is left
  void work()
  {
    for (;;) {
      int k;
      const std::function<void(int)> *f;
      {
        std::lock_guard<std::mutex> g(mu);
        if (!cur || next >= njobs) return;
        k = next++;
        f = cur;
      }
      (*f)(k);
      std::lock_guard<std::mutex> g(mu);
      if (++finished == njobs) done_cv.notify_all();
    }
  }
  void loop()
  {
    unsigned long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || gen != seen; });
        if (stop) return;
        seen = gen;
      }
      work();
    }
  }
  std::mutex run_mu, mu;
  std::condition_variable cv, done_cv;
  std::vector<std::thread> workers;
  const std::function<void(int)> *cur = nullptr;
  int nworkers = 0, njobs = 0, next = 0, finished = 0;
  unsigned long gen = 0;
  bool stop = false;
};

// fn(lo, hi) over [0, n) in up to 16 pieces on the host pool when there is
// enough work (at least min_per_thread items a piece)
template <class F>
void parallel_range(int n, int min_per_thread, bool threads, F fn)
{
  int nt = threads ? std::min(16, HostPool::get().size()) : 1;
  nt = std::min(nt, n / std::max(min_per_thread, 1));
  if (nt <= 1) {
    fn(0, n);
    return;
  }
  const int per = (n + nt - 1) / nt;
  const std::function<void(int)> job = [&](int t) {
    const int lo = t * per, hi = std::min(n, lo + per);
    if (lo < hi) fn(lo, hi);
  };
  HostPool::get().run(nt, job);
}

// Waits for the stream on every return path, so that no copy from the host
// staging is still in flight when the staging is reused or the call returns.
struct StreamDrain {
  hipStream_t s;
  bool done = false;
  ~StreamDrain()
  {
    if (!done) hipStreamSynchronize(s);
  }
};

}  // namespace

int list_compute(tmd_ctx *ctx, const ListArgs &a, const ListResources &r, long long serial)
{
  const char *who = a.who;
  if (!ctx) {
    set_err(nullptr, std::string(who) + ": ctx is NULL");
    return TMD_ERR_ARG;
  }
  if (a.nall < a.nlocal || a.nlocal < 0 || a.inum < 0 || a.inum > a.nlocal || !a.x || !a.type ||
      (!a.Gout && (!a.f || !a.energy)) || (a.inum > 0 && (!a.ilist || !a.numneigh || !a.firstneigh))) {
    set_err(ctx, std::string(who) + ": invalid arguments");
    return TMD_ERR_ARG;
  }
  if (a.nall > kMaxAtoms) {
    set_err(ctx, std::string(who) + ": nall exceeds the 2^23-1 atom limit of the neighbor encoding");
    return TMD_ERR_CAPACITY;
  }
  if (hipError_t de = hipSetDevice(ctx->device); de != hipSuccess) return hip_fail(ctx, de, "hipSetDevice");
  const Model &model = ctx->model;
  const int nelt = model.nelt;
  const bool threads = !r.fixed;
  // the list in list_ws is that of the previous serial call with this serial
  const bool reuse = !r.fixed && serial >= 0 && ctx->list_cached && ctx->list_serial == serial &&
                     ctx->list_nlocal == a.nlocal && ctx->list_nall == a.nall && ctx->list_inum == a.inum;
  if (!r.fixed) ctx->list_cached = false;  // list_ws is rewritten below unless reused

  // ---- plan: list stride and rows ----
  ListPlan plan;
  if (reuse) {
    plan.cap = ctx->list_cap;
    plan.nrows = ctx->list_ws.rl.nrows;
  } else {
    int cap = 8;
    for (int ii = 0; ii < a.inum; ii++) {
      const int i = a.ilist[ii];
      if (i < 0 || i >= a.nlocal) {
        set_err(ctx, std::string(who) + ": ilist entry out of range");
        return TMD_ERR_ARG;
      }
      const int t = a.type[i];
      const int e = (t >= 1 && t <= ctx->ntypes) ? ctx->type_map[t] : -1;
      if (e < 0) {
        set_err(ctx, std::string(who) + ": atom of a type not mapped to a model element");
        return TMD_ERR_ARG;
      }
      plan.cnt[e]++;
      cap = std::max(cap, a.numneigh[i]);
    }
    plan.cap = (cap + 7) & ~7;
    for (int e = 0; e < nelt; e++) plan.nrows += (plan.cnt[e] + kRowTile - 1) / kRowTile * kRowTile;
  }

  // ---- host staging ----
  size_t hoff[6];
  const bool stage_f = !r.fixed && !a.Gout;
  const size_t hbytes = host_stage_bytes(a, plan, hoff, stage_f);
  unsigned char *hbase;
  if (r.fixed) {
    if (hbytes > r.host_bytes) {
      set_err(ctx, std::string(who) + ": the call needs " + std::to_string(hbytes) +
                       " bytes of host staging, more than the fixed " + std::to_string(r.host_bytes));
      return TMD_ERR_CAPACITY;
    }
    hbase = r.host;
  } else {
    hipError_t e = ctx->list_host.ensure(hbytes);
    if (e != hipSuccess) return hip_fail(ctx, e, "pinned staging alloc");
    hbase = (unsigned char *) ctx->list_host.p;
  }
  HostStage hs = carve_host(hbase, hoff, stage_f);

  // positions with the element in w; every position's type must be mapped
  std::atomic<int> bad_type{0};
  parallel_range(a.nall, 32768, threads, [&](int lo, int hi) {
    for (int i = lo; i < hi; i++) {
      const int t = a.type[i];
      const int e = (t >= 1 && t <= ctx->ntypes) ? ctx->type_map[t] : -1;
      if (e < 0) {
        bad_type.store(1, std::memory_order_relaxed);
        continue;
      }
      hs.pos[i] = make_double4(a.x[3 * i], a.x[3 * i + 1], a.x[3 * i + 2], (double) e);
    }
  });
  if (bad_type.load()) {
    set_err(ctx, std::string(who) + ": atom of a type not mapped to a model element");
    return TMD_ERR_ARG;
  }
  RowLayout rows = ctx->list_ws.rl;  // the cached layout when the list is reused
  if (!reuse) {
    // padded list rows of the atoms in ilist; entries as j | center image code
    std::memset(hs.nn, 0, 4 * (size_t) std::max(a.nlocal, 1));
    std::atomic<int> bad_j{0};
    parallel_range(a.inum, 2048, threads, [&](int lo, int hi) {
      for (int ii = lo; ii < hi; ii++) {
        const int i = a.ilist[ii];
        const int n = a.numneigh[i];
        const int *src = a.firstneigh[i];
        uint32_t *dst = hs.list + (size_t) i * plan.cap;
        hs.nn[i] = n;
        for (int p = 0; p < n; p++) {
          const int j = src[p] & 0x1FFFFFFF;  // NEIGHMASK
          if (j < 0 || j >= a.nall) {
            bad_j.store(1, std::memory_order_relaxed);
            break;
          }
          dst[p] = (uint32_t) j | (kShiftCenter << kNbrJBits);
        }
      }
    });
    if (bad_j.load()) {
      set_err(ctx, std::string(who) + ": neighbor index out of range");
      return TMD_ERR_ARG;
    }
    // rows grouped by element, list order inside an element
    // (tensormd.cpp:377-387), each group padded to kRowTile
    std::memset(&rows, 0, sizeof(rows));
    int off = 0, fill[kMaxElts];
    for (int e = 0; e < nelt; e++) {
      rows.row_off[e] = off;
      fill[e] = off;
      off += (plan.cnt[e] + kRowTile - 1) / kRowTile * kRowTile;
    }
    for (int e = nelt; e <= kMaxElts; e++) rows.row_off[e] = off;
    rows.nrows = off;
    for (int q = 0; q < off; q++) hs.row2atom[q] = -1;
    for (int ii = 0; ii < a.inum; ii++) {
      const int i = a.ilist[ii];
      hs.row2atom[fill[ctx->type_map[a.type[i]]]++] = i;
    }
  }
  const int gwf = model.gwidth_file();
  if (a.Gout) std::memset(a.Gout, 0, sizeof(double) * (size_t) a.nlocal * gwf);
  if (plan.nrows == 0) {
    if (a.Gout) return TMD_OK;
    std::memset(a.f, 0, sizeof(double) * 3 * (size_t) a.nall);
    if (a.vatom) std::memset(a.vatom, 0, sizeof(double) * 6 * (size_t) a.nall);
    if (a.eatom) std::memset(a.eatom, 0, sizeof(double) * (size_t) a.nlocal);
    *a.energy = 0.0;
    if (a.virial) std::memset(a.virial, 0, 6 * sizeof(double));
    return TMD_OK;
  }

  // ---- device buffers ----
  Workspace fixed_ws;
  Workspace &w = r.fixed ? fixed_ws : ctx->list_ws;
  if (r.fixed) {
    size_t doff[14];
    const size_t need = device_bytes(a, plan, model.gwidth(), doff);
    if (need > r.dev_bytes) {
      set_err(ctx, std::string(who) + ": the call needs " + std::to_string(need) +
                       " bytes of device scratch, more than the fixed " + std::to_string(r.dev_bytes));
      return TMD_ERR_CAPACITY;
    }
    carve_device(w, (unsigned char *) r.dev_base, doff);
  } else {
    int rc = ensure_rows(ctx, w, plan.nrows);
    if (rc != TMD_OK) return rc;
    if ((rc = ensure_kept(ctx, w, plan.nrows, plan.cap)) != TMD_OK) return rc;
    if ((rc = ensure_atoms(ctx, w, a.nlocal)) != TMD_OK) return rc;
    hipError_t e;
    if ((e = w.pos.ensure(sizeof(double4) * std::max(a.nall, 1))) != hipSuccess) return hip_fail(ctx, e, "pos alloc");
    if ((e = w.nbr.ensure(4 * (size_t) std::max(a.nlocal, 1) * plan.cap)) != hipSuccess)
      return hip_fail(ctx, e, "nbr alloc");
    if ((e = w.nn.ensure(4 * (size_t) std::max(a.nlocal, 1))) != hipSuccess) return hip_fail(ctx, e, "nn alloc");
    if ((e = w.f.ensure(sizeof(double) * 3 * (size_t) a.nall)) != hipSuccess) return hip_fail(ctx, e, "f alloc");
    if (a.vatom && (e = w.vatom.ensure(sizeof(double) * 6 * (size_t) a.nall)) != hipSuccess)
      return hip_fail(ctx, e, "vatom alloc");
    if ((e = w.cells.ensure(sizeof(CellInfo))) != hipSuccess) return hip_fail(ctx, e, "cell alloc");
  }

  w.rl = rows;

  // ---- upload ----
  hipStream_t s = ctx->stream;
  StreamDrain drain{s};
  hipError_t e;
  // list mode: positions are already images (ghosts), the one cell is zero
  std::memset(hs.cell, 0, sizeof(CellInfo));
  if ((e = hipMemcpyAsync(w.pos.p, hs.pos, sizeof(double4) * a.nall, hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(ctx, e, "position upload");
  if (!reuse &&
      ((e = hipMemcpyAsync(w.nbr.p, hs.list, 4 * (size_t) std::max(a.nlocal, 1) * plan.cap, hipMemcpyHostToDevice,
                           s)) != hipSuccess ||
       (e = hipMemcpyAsync(w.nn.p, hs.nn, 4 * (size_t) std::max(a.nlocal, 1), hipMemcpyHostToDevice, s)) !=
           hipSuccess ||
       (e = hipMemcpyAsync(w.row2atom.p, hs.row2atom, 4 * (size_t) plan.nrows, hipMemcpyHostToDevice, s)) !=
           hipSuccess ||
       (e = hipMemcpyAsync(w.cells.p, hs.cell, sizeof(CellInfo), hipMemcpyHostToDevice, s)) != hipSuccess))
    return hip_fail(ctx, e, "list upload");

  if (a.Gout) {
    Geom geo;
    std::memset(&geo, 0, sizeof(geo));
    geo.cells = (const CellInfo *) w.cells.p;
    e = launch_descriptor(ctx->geo, geo, (const double4 *) w.pos.p, (const uint32_t *) w.nbr.p,
                          (const int *) w.nn.p, plan.cap, (const int *) w.row2atom.p, w.rl.nrows,
                          (const float4 *) ctx->tab.p, (float *) w.G.p, (uint32_t *) w.kl.p, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "descriptor launch");
    const int gw = model.gwidth(), nm = model.nm, nmf = model.max_moment_file + 1;
    std::vector<float> g((size_t) w.rl.nrows * gw);
    if ((e = hipMemcpyAsync(g.data(), w.G.p, 4 * g.size(), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return hip_fail(ctx, e, "descriptor download");
    drain.done = true;
    // the model file's layout [b][k][max_moment_file + 1]
    for (int rr = 0; rr < w.rl.nrows; rr++) {
      const int at = hs.row2atom[rr];
      if (at < 0) continue;
      for (int q = 0; q < gwf; q++)
        a.Gout[(size_t) at * gwf + q] = g[(size_t) rr * gw + (size_t) (q / nmf) * nm + q % nmf];
    }
    return TMD_OK;
  }

  Chain c;
  c.geo.cells = (const CellInfo *) w.cells.p;
  c.pos = (const double4 *) w.pos.p;
  c.nbr = (const uint32_t *) w.nbr.p;
  c.nn = (const int *) w.nn.p;
  c.cap = plan.cap;
  c.scale = a.scale;
  c.f = (double *) w.f.p;
  c.fbytes = sizeof(double) * 3 * (size_t) a.nall;
  c.vatom = a.vatom ? (double *) w.vatom.p : nullptr;
  c.vbytes = sizeof(double) * 6 * (size_t) a.nall;
  c.natoms = a.nlocal;
  c.zero_ew = a.inum < a.nlocal;
  c.g_bf16 = true;
  int rc = run_chain(ctx, w, c, s);
  if (rc != TMD_OK) return rc;
  // the sums land in the host staging's cell slot (pinned or fixed host
  // memory), read after the stream is drained
  double *red = reinterpret_cast<double *>(hs.cell);
  if ((e = hipMemcpyAsync(hs.f ? hs.f : a.f, w.f.p, sizeof(double) * 3 * (size_t) a.nall, hipMemcpyDeviceToHost,
                          s)) != hipSuccess ||
      (a.vatom && (e = hipMemcpyAsync(a.vatom, w.vatom.p, sizeof(double) * 6 * (size_t) a.nall,
                                      hipMemcpyDeviceToHost, s)) != hipSuccess) ||
      (e = hipMemcpyAsync(red, w.red.p, 7 * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (a.eatom && (e = hipMemcpyAsync(a.eatom, w.ew.p, sizeof(double) * (size_t) a.nlocal, hipMemcpyDeviceToHost,
                                      s)) != hipSuccess))
    return hip_fail(ctx, e, "result download");
  // TDNP: the per-row electron entropy and free energy (the context's
  // allocating paths keep them for tmd_tdnp_results)
  std::vector<double> sa;
  const bool td = model.tdnp && !r.fixed;
  if (td) {
    sa.resize(2 * (size_t) w.rl.nrows);
    if ((e = hipMemcpyAsync(sa.data(), w.sa.p, 8 * sa.size(), hipMemcpyDeviceToHost, s)) != hipSuccess)
      return hip_fail(ctx, e, "TDNP download");
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(ctx, e, who);
  drain.done = true;
  if (hs.f) {
    const size_t nf = 3 * (size_t) a.nall;
    parallel_range((int) ((nf + 4095) / 4096), 16, threads, [&](int lo, int hi) {
      const size_t b = (size_t) lo * 4096, e2 = std::min(nf, (size_t) hi * 4096);
      std::memcpy(a.f + b, hs.f + b, sizeof(double) * (e2 - b));
    });
  }
  if (td) {
    ctx->td_satom.assign(a.nlocal, 0.0);
    ctx->td_aatom.assign(a.nlocal, 0.0);
    double S = 0.0, A = 0.0;
    for (int rr = 0; rr < w.rl.nrows; rr++) {  // atoms in row order (fixed)
      const int at = hs.row2atom[rr];
      if (at < 0) continue;
      ctx->td_satom[at] = a.scale * sa[2 * (size_t) rr];
      ctx->td_aatom[at] = a.scale * sa[2 * (size_t) rr + 1];
      S += ctx->td_satom[at];
      A += ctx->td_aatom[at];
    }
    ctx->td_entropy = S;
    ctx->td_free_energy = A;
    ctx->td_valid = true;
  }
  *a.energy = red[0];
  if (a.virial)
    for (int q = 0; q < 6; q++) a.virial[q] = red[1 + q];
  if (!r.fixed && serial >= 0) {
    ctx->list_cached = true;
    ctx->list_serial = serial;
    ctx->list_nlocal = a.nlocal;
    ctx->list_nall = a.nall;
    ctx->list_inum = a.inum;
    ctx->list_cap = plan.cap;
  }
  return TMD_OK;
}

}  // namespace tmd
