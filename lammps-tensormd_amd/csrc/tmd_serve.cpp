// Request loop for batched single-point energy/force/virial evaluations of
// small periodic structures (continuous batching).
//
// Each request is what one PairTensorMD::compute (src/ML-TENSORMD/
// pair_tensormd.cpp:141-184) does for a small periodic cell; here many of
// them share one launch.  A worker thread owns the device side:
//   - slots: each has its own stream, device arrays sized for a capacity
//     (atoms, rows, bins, structures) fixed at start from the free device
//     memory, and pinned host staging/output areas that the batch kernels
//     read and write directly;
//   - batching: whenever a slot is free, every queued request that still
//     fits its capacities is packed (FIFO) into one batch, so requests that
//     arrive while others run join the next batch, and batches on different
//     slots overlap on the device;
//   - launch: the batch is one captured HIP graph per capacity bucket
//     (grids sized for the bucket, live sizes read from a device header), so
//     a batch costs one graph launch plus two events;
//   - tuning: on a bucket's first use the MLP row tile (16/32/64 rows per
//     workgroup) is chosen by timing each candidate on a full bucket.
// Structures are kept apart on the device by a per-atom structure index:
// each structure has its own cell and bin range, and sums are per
// structure in a fixed order, so a structure's energy and virial are the
// same alone and in any batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/tensormd_amd.h"
#include "tmd_device.h"
#include "tmd_internal.h"

using namespace tmd;

namespace tmd {

namespace {

double now_s()
{
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

constexpr int kMaxBuckets = 8;

struct Request {
  tmd_ticket id = 0;
  int natoms = 0;
  std::vector<double4> pos;  // element index in .w
  CellInfo cell{};           // bin_base set when batched
  int nbins = 1;
  int listcap = 8;           // neighbor entries per atom this structure needs (estimate)
  double *f_out = nullptr;
  double *eatom_out = nullptr;
  int status = TMD_PENDING;
  double energy = 0.0;
  double virial[6] = {0, 0, 0, 0, 0, 0};
  double t_submit = 0, t_start = 0, t_done = 0;
  int batch_structs = 0, batch_atoms = 0;
};

struct Bucket {
  BatchCaps caps{};
  hipGraphExec_t exec = nullptr;
  int rtile = 0;  // 0 = not tuned
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  Workspace w;  // pos, nbr, nn, f, ew, G, dEdG, E, row2atom, cells
  DeviceBuffer dyn, soff, astruct, sout, bins, atominfo, xbuild, ctl;
  unsigned char *stage = nullptr;  // pinned host (device-visible)
  double *out = nullptr;           // pinned host (device-visible)
  unsigned char *stage_dev = nullptr;
  double *out_dev = nullptr;
  int listcap = 0;
  Bucket buckets[kMaxBuckets];
  int nbuckets = 0;
  std::vector<Request *> batch;
  bool busy = false;
  int batch_atoms = 0, batch_structs = 0;
  double t_launch = 0;
};

}  // namespace

struct Server {
  tmd_ctx *ctx = nullptr;
  tmd_serve_options opt{};
  BatchCaps full{};
  double cut = 0;
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv;       // work for the worker
  std::condition_variable done_cv;  // a request finished
  std::deque<Request *> queue;
  std::unordered_map<tmd_ticket, Request *> reqs;
  bool stopping = false;
  std::thread worker;
  tmd_ticket next_id = 1;
  std::string fatal;  // set when the worker stops on a device error
  // statistics (worker thread; read under mu)
  long batches = 0, structs = 0, atoms = 0, captures = 0, reruns = 0;
  double gpu_ms = 0;
};

namespace {

int serve_fail(tmd_ctx *ctx, int code, const std::string &m)
{
  set_err(ctx, m);
  return code;
}

// Device bytes per atom of one slot (rows ~ atoms).
size_t bytes_per_atom(const tmd_ctx *ctx, int listcap)
{
  const size_t gw = (size_t) ctx->model.gwidth();
  return 8 * gw + 8 /*E*/ + 4 /*row2atom*/ + 4 * (size_t) listcap + 4 /*nn*/ + 32 /*pos*/ +
      4 * kept_cache_words(listcap) /*kept lists*/ +
      24 /*f*/ + 56 /*eatom, watom*/ + 4 /*atom_struct*/ + 12 /*bins*/ + 20 /*atom bins*/ +
      32 /*x_build*/ + 48 /*staging, output*/;
}

BatchArrays arrays_of(Slot &sl)
{
  BatchArrays a;
  a.dyn = (BatchDims *) sl.dyn.p;
  a.struct_off = (int *) sl.soff.p;
  a.cells = (CellInfo *) sl.w.cells.p;
  a.atom_struct = (int *) sl.astruct.p;
  a.row2atom = (int *) sl.w.row2atom.p;
  a.pos = (double4 *) sl.w.pos.p;
  a.struct_out = (const double *) sl.sout.p;
  a.f = (const double *) sl.w.f.p;
  a.eatom = (const double *) sl.w.ew.p;
  a.ctl = (unsigned long long *) sl.ctl.p;
  return a;
}

void drop_graphs(Slot &sl)
{
  for (int b = 0; b < sl.nbuckets; b++)
    if (sl.buckets[b].exec) {
      hipGraphExecDestroy(sl.buckets[b].exec);
      sl.buckets[b].exec = nullptr;
    }
}

int alloc_list(Server *S, Slot &sl, int listcap)
{
  hipError_t e = sl.w.nbr.ensure((size_t) S->full.atoms * listcap * 4);
  if (e != hipSuccess) return hip_fail(S->ctx, e, "serve: neighbor list alloc");
  int rc = ensure_kept(S->ctx, sl.w, S->full.rows, listcap);
  if (rc != TMD_OK) return rc;
  sl.listcap = listcap;
  drop_graphs(sl);
  return TMD_OK;
}

int create_slot(Server *S, Slot &sl, int listcap)
{
  tmd_ctx *ctx = S->ctx;
  const BatchCaps &c = S->full;
  hipError_t e;
  auto chk = [&](hipError_t er, const char *what) -> int {
    return er == hipSuccess ? TMD_OK : hip_fail(ctx, er, what);
  };
  int rc;
  if ((rc = chk(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "serve: stream"))) return rc;
  if ((rc = chk(hipEventCreate(&sl.ev_start), "serve: event"))) return rc;
  if ((rc = chk(hipEventCreate(&sl.ev_done), "serve: event"))) return rc;
  if ((rc = ensure_rows(ctx, sl.w, c.rows))) return rc;
  if ((rc = ensure_atoms(ctx, sl.w, c.atoms))) return rc;
  if ((rc = chk(sl.w.pos.ensure(sizeof(double4) * c.atoms), "serve: pos"))) return rc;
  if ((rc = chk(sl.w.nn.ensure(4 * (size_t) c.atoms), "serve: nn"))) return rc;
  if ((rc = chk(sl.w.f.ensure(24 * (size_t) c.atoms), "serve: f"))) return rc;
  if ((rc = chk(sl.w.cells.ensure(sizeof(CellInfo) * c.structs), "serve: cells"))) return rc;
  if ((rc = chk(sl.dyn.ensure(sizeof(BatchDims)), "serve: header"))) return rc;
  if ((rc = chk(sl.soff.ensure(4 * (size_t) (c.structs + 1)), "serve: offsets"))) return rc;
  if ((rc = chk(sl.astruct.ensure(4 * (size_t) c.atoms), "serve: atom structure"))) return rc;
  if ((rc = chk(sl.sout.ensure(56 * (size_t) c.structs), "serve: sums"))) return rc;
  if ((rc = chk(sl.bins.ensure(sizeof(int) * (3 * (size_t) (c.bins + 1) + 1)), "serve: bins"))) return rc;
  if ((rc = chk(sl.atominfo.ensure(sizeof(int) * 5 * (size_t) c.atoms), "serve: atom bins"))) return rc;
  if ((rc = chk(sl.xbuild.ensure(sizeof(double) * 4 * (size_t) c.atoms), "serve: x_build"))) return rc;
  if ((rc = chk(sl.ctl.ensure(sizeof(unsigned long long) * 8), "serve: ctl"))) return rc;
  if ((rc = alloc_list(S, sl, listcap))) return rc;
  size_t off[7];
  batch_stage_offsets(c.structs, c.atoms, c.rows, off);
  if ((rc = chk(hipHostMalloc((void **) &sl.stage, off[6], hipHostMallocCoherent), "serve: staging"))) return rc;
  if ((rc = chk(hipHostMalloc((void **) &sl.out, batch_out_bytes(c.structs, c.atoms), hipHostMallocCoherent),
                "serve: output")))
    return rc;
  if ((rc = chk(hipHostGetDevicePointer((void **) &sl.stage_dev, sl.stage, 0), "serve: staging map"))) return rc;
  if ((rc = chk(hipHostGetDevicePointer((void **) &sl.out_dev, sl.out, 0), "serve: output map"))) return rc;
  // clear everything a kernel may read before it is written
  hipStream_t st = sl.stream;
  if ((e = launch_zero(sl.bins.p, sl.bins.bytes, st)) != hipSuccess ||
      (e = launch_zero(sl.ctl.p, sl.ctl.bytes, st)) != hipSuccess ||
      (e = launch_zero(sl.xbuild.p, sl.xbuild.bytes, st)) != hipSuccess ||
      (e = launch_zero(sl.w.G.p, sl.w.G.bytes, st)) != hipSuccess ||
      (e = launch_zero(sl.w.nn.p, sl.w.nn.bytes, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return hip_fail(ctx, e, "serve: slot init");
  // capacity buckets: full, /4, /16, ... (at least 1024 atoms)
  sl.nbuckets = 0;
  for (int a = c.atoms; sl.nbuckets < kMaxBuckets; a /= 4) {
    Bucket &b = sl.buckets[sl.nbuckets++];
    b.caps.atoms = a;
    b.caps.rows = (a + c.nelt * kRowTile + kRowTile - 1) / kRowTile * kRowTile;
    b.caps.rows = std::min(b.caps.rows, c.rows);
    b.caps.bins = std::min(a, c.bins);
    b.caps.structs = std::min(c.structs, a);
    b.caps.nelt = c.nelt;
    if (a / 4 < 1024) break;
  }
  return TMD_OK;
}

void destroy_slot(Slot &sl)
{
  drop_graphs(sl);
  if (sl.stream) hipStreamSynchronize(sl.stream);
  if (sl.stage) hipHostFree(sl.stage);
  if (sl.out) hipHostFree(sl.out);
  if (sl.ev_start) hipEventDestroy(sl.ev_start);
  if (sl.ev_done) hipEventDestroy(sl.ev_done);
  if (sl.stream) hipStreamDestroy(sl.stream);
  sl.stage = nullptr;
  sl.out = nullptr;
  sl.stream = nullptr;
}

// The kernels of one batch on the slot's stream (captured or eager).
int enqueue_batch(Server *S, Slot &sl, const Bucket &b, int rtile)
{
  tmd_ctx *ctx = S->ctx;
  hipStream_t st = sl.stream;
  BatchArrays a = arrays_of(sl);
  hipError_t e = launch_batch_unpack(sl.stage_dev, b.caps, a, st);
  if (e != hipSuccess) return hip_fail(ctx, e, "serve: unpack");
  NbrBuild nb{};
  nb.n = b.caps.atoms;
  nb.cap = sl.listcap;
  nb.cut2 = S->cut * S->cut;
  nb.trig2 = 0.0;
  nb.nbins = b.caps.bins;
  nb.cells = a.cells;
  nb.atom_struct = a.atom_struct;
  nb.dyn = a.dyn;
  int *bins = (int *) sl.bins.p;
  nb.bin_count = bins;
  nb.bin_start = bins + (S->full.bins + 1);
  nb.bin_cursor = bins + 2 * (S->full.bins + 1);
  int *ai = (int *) sl.atominfo.p;
  nb.bin_atoms = ai;
  nb.atom_bin = ai + S->full.atoms;
  nb.atom_bc = ai + 2 * (size_t) S->full.atoms;
  nb.x_build = (double *) sl.xbuild.p;
  nb.ctl = a.ctl;
  e = nbr_step(nb, a.pos, (uint32_t *) sl.w.nbr.p, (int *) sl.w.nn.p, st);
  if (e != hipSuccess) return hip_fail(ctx, e, "serve: neighbor build");
  sl.w.rl.nrows = b.caps.rows;  // grid extent; live layout is on the device
  Chain c;
  c.geo.cells = a.cells;
  c.geo.atom_struct = a.atom_struct;
  c.geo.dyn = a.dyn;
  c.pos = a.pos;
  c.nbr = (const uint32_t *) sl.w.nbr.p;
  c.nn = (const int *) sl.w.nn.p;
  c.cap = sl.listcap;
  c.f = (double *) sl.w.f.p;
  c.fbytes = 24 * (size_t) b.caps.atoms;
  c.natoms = S->full.atoms;  // eatom/watom layout of the slot
  c.struct_off = a.struct_off;
  c.nstruct = b.caps.structs;
  c.struct_out = (double *) sl.sout.p;
  c.rtile = rtile;
  c.dyn = a.dyn;
  c.g_bf16 = true;
  int rc = run_chain(ctx, sl.w, c, st);
  if (rc != TMD_OK) return rc;
  e = launch_batch_pack(sl.out_dev, b.caps, a, st);
  if (e != hipSuccess) return hip_fail(ctx, e, "serve: pack");
  return TMD_OK;
}

// Pick the MLP row tile of a bucket by timing each candidate over a full
// bucket of rows (the values in G do not matter for the timing).
int tune_bucket(Server *S, Slot &sl, Bucket &b)
{
  tmd_ctx *ctx = S->ctx;
  const int mode = ctx->opt.mlp_precision;  // 0 fp32, 1 bf16, 2 fp8
  RowLayout rl{};
  rl.nrows = b.caps.rows;
  const int nelt = ctx->model.nelt;
  int per = (b.caps.rows / nelt) / kRowTile * kRowTile;
  for (int e = 0; e <= kMaxElts; e++) rl.row_off[e] = std::min(e, nelt) * per;
  rl.row_off[nelt] = b.caps.rows;
  for (int e = nelt + 1; e <= kMaxElts; e++) rl.row_off[e] = b.caps.rows;
  const int cand[3] = {16, 32, 64};
  float best = 1e30f;
  int pick = 0;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return serve_fail(ctx, TMD_ERR_DEVICE, "serve: event");
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return serve_fail(ctx, TMD_ERR_DEVICE, "serve: event");
  }
  for (int q = 0; q < 3; q++) {
    MlpParams m;
    if (mlp_lds_bytes(ctx->mlp, mode, cand[q], &m) > 160 * 1024) continue;
    hipError_t e = launch_mlp(mode, rl, ctx->mlp, (const float *) sl.w.G.p, (double *) sl.w.E.p,
                              (float *) sl.w.dEdG.p, sl.stream, cand[q]);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventRecord(e0, sl.stream);
    for (int r = 0; r < 3 && e == hipSuccess; r++)
      e = launch_mlp(mode, rl, ctx->mlp, (const float *) sl.w.G.p, (double *) sl.w.E.p,
                     (float *) sl.w.dEdG.p, sl.stream, cand[q]);
    if (e == hipSuccess) e = hipEventRecord(e1, sl.stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) {
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      return hip_fail(ctx, e, "serve: tile tuning");
    }
    if (ms < best) {
      best = ms;
      pick = cand[q];
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  b.rtile = pick ? pick : 16;
  return TMD_OK;
}

int launch_batch(Server *S, Slot &sl)
{
  // smallest bucket that holds the batch
  int bi = 0;
  for (int q = 0; q < sl.nbuckets; q++) {
    const BatchCaps &c = sl.buckets[q].caps;
    const BatchDims *d = (const BatchDims *) sl.stage;
    if (d->natoms <= c.atoms && d->rl.nrows <= c.rows && d->nbins <= c.bins && d->nstruct <= c.structs)
      bi = q;
  }
  Bucket &b = sl.buckets[bi];
  int rc;
  if (S->opt.autotune && b.rtile == 0 && (rc = tune_bucket(S, sl, b)) != TMD_OK) return rc;
  const int rtile = S->opt.autotune ? b.rtile : 0;
  hipError_t e;
  if (S->opt.use_graph && !b.exec) {
    hipGraph_t g;
    if ((e = hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hip_fail(S->ctx, e, "serve: capture");
    rc = enqueue_batch(S, sl, b, rtile);
    e = hipStreamEndCapture(sl.stream, &g);
    if (rc != TMD_OK) return rc;
    if (e != hipSuccess) return hip_fail(S->ctx, e, "serve: capture");
    e = hipGraphInstantiate(&b.exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      b.exec = nullptr;
      return hip_fail(S->ctx, e, "serve: instantiate");
    }
    S->captures++;
  }
  if ((e = hipEventRecord(sl.ev_start, sl.stream)) != hipSuccess) return hip_fail(S->ctx, e, "serve: event");
  if (S->opt.use_graph) {
    if ((e = hipGraphLaunch(b.exec, sl.stream)) != hipSuccess) return hip_fail(S->ctx, e, "serve: graph launch");
  } else if ((rc = enqueue_batch(S, sl, b, rtile)) != TMD_OK) {
    return rc;
  }
  if ((e = hipEventRecord(sl.ev_done, sl.stream)) != hipSuccess) return hip_fail(S->ctx, e, "serve: event");
  sl.busy = true;
  sl.t_launch = now_s();
  return TMD_OK;
}

// Pack sl.batch into the slot's staging area.
void stage_batch(Server *S, Slot &sl)
{
  const int nelt = S->ctx->model.nelt;
  int ns = (int) sl.batch.size(), n = 0, nbins = 0;
  std::vector<int> cnt(nelt, 0);
  for (Request *r : sl.batch) {
    n += r->natoms;
    nbins += r->nbins;
    for (int i = 0; i < r->natoms; i++) cnt[(int) r->pos[i].w]++;
  }
  BatchDims d{};
  d.natoms = n;
  d.nbins = nbins;
  d.nstruct = ns;
  int off = 0;
  for (int e = 0; e < nelt; e++) {
    d.rl.row_off[e] = off;
    off += (cnt[e] + kRowTile - 1) / kRowTile * kRowTile;
  }
  for (int e = nelt; e <= kMaxElts; e++) d.rl.row_off[e] = off;
  d.rl.nrows = off;
  size_t o[7];
  batch_stage_offsets(ns, n, d.rl.nrows, o);
  unsigned char *st = sl.stage;
  std::memcpy(st, &d, sizeof(d));
  int *soff = (int *) (st + sizeof(BatchDims));
  CellInfo *cells = (CellInfo *) (st + o[1]);
  int *as = (int *) (st + o[2]);
  int *r2a = (int *) (st + o[3]);
  double4 *pos = (double4 *) (st + o[4]);
  std::vector<int> cursor(nelt);
  for (int e = 0; e < nelt; e++) cursor[e] = d.rl.row_off[e];
  for (int r = 0; r < d.rl.nrows; r++) r2a[r] = -1;
  int a0 = 0, b0 = 0;
  for (int s = 0; s < ns; s++) {
    Request *r = sl.batch[s];
    soff[s] = a0;
    CellInfo c = r->cell;
    c.bin_base = b0;
    cells[s] = c;
    std::memcpy(pos + a0, r->pos.data(), sizeof(double4) * r->natoms);
    for (int i = 0; i < r->natoms; i++) {
      as[a0 + i] = s;
      r2a[cursor[(int) r->pos[i].w]++] = a0 + i;
    }
    a0 += r->natoms;
    b0 += r->nbins;
  }
  soff[ns] = a0;
  sl.batch_atoms = n;
  sl.batch_structs = ns;
}

// Results of a finished batch -> requests.  Returns false when the batch
// must run again (neighbor list overflow: the list was grown).
bool finish_batch(Server *S, Slot &sl, bool *rerun)
{
  *rerun = false;
  const double *out = sl.out;
  float ms = 0.f;  // the batch's device time feeds the statistics only
  if (hipEventElapsedTime(&ms, sl.ev_start, sl.ev_done) != hipSuccess) ms = 0.f;
  const long overflow = (long) out[0];
  const int err = (int) out[3];
  if (overflow > 0 && !err) {
    int need = (int) overflow;
    int cap = ((need + need / 4 + 16) + 7) & ~7;
    if (alloc_list(S, sl, cap) == TMD_OK) {
      *rerun = true;
      std::lock_guard<std::mutex> lk(S->mu);
      S->reruns++;
      return false;
    }
  }
  const int ns = sl.batch_structs, n = sl.batch_atoms;
  const double *so = out + 8;
  const double *fo = so + 7 * (size_t) ns;
  const double *eo = fo + 3 * (size_t) n;
  const double t = now_s();
  int a0 = 0;
  for (int s = 0; s < ns; s++) {
    Request *r = sl.batch[s];
    if (!err && overflow == 0) {
      r->energy = so[7 * s];
      for (int q = 0; q < 6; q++) r->virial[q] = so[7 * s + 1 + q];
      if (r->f_out) std::memcpy(r->f_out, fo + 3 * (size_t) a0, 24 * (size_t) r->natoms);
      if (r->eatom_out) std::memcpy(r->eatom_out, eo + a0, 8 * (size_t) r->natoms);
    }
    a0 += r->natoms;
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Request *r : sl.batch) {
      r->status = (err || overflow) ? TMD_ERR_CAPACITY : TMD_OK;
      r->t_start = sl.t_launch;
      r->t_done = t;
      r->batch_structs = ns;
      r->batch_atoms = n;
      r->pos.clear();
      r->pos.shrink_to_fit();
    }
    S->batches++;
    S->structs += ns;
    S->atoms += n;
    S->gpu_ms += ms;
  }
  S->done_cv.notify_all();
  sl.batch.clear();
  sl.busy = false;
  return true;
}

void fail_batch(Server *S, Slot &sl, int code)
{
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Request *r : sl.batch) {
      r->status = code;
      r->t_done = now_s();
    }
  }
  S->done_cv.notify_all();
  sl.batch.clear();
  sl.busy = false;
}

// Take queued requests that fit the slot (caller holds S->mu).
void take_batch(Server *S, Slot &sl)
{
  long atoms = 0, bins = 0;
  const BatchCaps &c = S->full;
  while (!S->queue.empty() && (int) sl.batch.size() < c.structs) {
    Request *r = S->queue.front();
    if (atoms + r->natoms > c.atoms || bins + r->nbins > c.bins) break;
    atoms += r->natoms;
    bins += r->nbins;
    sl.batch.push_back(r);
    S->queue.pop_front();
  }
}

void worker_main(Server *S)
{
  mark_library_thread();  // errors go to S->fatal, never to ctx->err
  hipSetDevice(S->ctx->device);
  for (;;) {
    bool progressed = false, any_busy = false;
    for (auto &slp : S->slots) {
      Slot &sl = *slp;
      if (!sl.busy) continue;
      hipError_t q = hipEventQuery(sl.ev_done);
      if (q == hipErrorNotReady) {
        any_busy = true;
        continue;
      }
      progressed = true;
      if (q != hipSuccess) {
        {
          std::lock_guard<std::mutex> lk(S->mu);
          S->fatal = std::string("serve: batch failed: ") + hipGetErrorString(q);
        }
        fail_batch(S, sl, TMD_ERR_DEVICE);
        continue;
      }
      bool rerun = false;
      finish_batch(S, sl, &rerun);
      if (rerun) {
        stage_batch(S, sl);
        if (launch_batch(S, sl) != TMD_OK) fail_batch(S, sl, TMD_ERR_DEVICE);
        else any_busy = true;
      }
    }
    for (auto &slp : S->slots) {
      Slot &sl = *slp;
      if (sl.busy) continue;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->fatal.empty()) {
          // refuse further work after a device error
          while (!S->queue.empty()) {
            Request *r = S->queue.front();
            S->queue.pop_front();
            r->status = TMD_ERR_DEVICE;
          }
          S->done_cv.notify_all();
          break;
        }
        take_batch(S, sl);
      }
      if (sl.batch.empty()) break;
      int need = 0;
      for (Request *r : sl.batch) need = std::max(need, r->listcap);
      if (need > sl.listcap && alloc_list(S, sl, need) != TMD_OK) {
        fail_batch(S, sl, TMD_ERR_NOMEM);
        continue;
      }
      stage_batch(S, sl);
      if (launch_batch(S, sl) != TMD_OK) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->fatal = "serve: launch failed: " + thread_error();
      }
      if (!sl.busy) fail_batch(S, sl, TMD_ERR_DEVICE);
      progressed = true;
      any_busy = any_busy || sl.busy;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->stopping && S->queue.empty() && !any_busy) {
      bool idle = true;
      for (auto &slp : S->slots) idle = idle && !slp->busy;
      if (idle) break;
    }
    if (!progressed) {
      if (any_busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      } else {
        S->cv.wait_for(lk, std::chrono::milliseconds(5),
                       [S] { return !S->queue.empty() || S->stopping; });
      }
    }
  }
}

}  // namespace

void server_destroy(Server *S)
{
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
  }
  S->cv.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (auto &sl : S->slots) destroy_slot(*sl);
  for (auto &kv : S->reqs) delete kv.second;
  delete S;
}

}  // namespace tmd

extern "C" {

double tmd_clock(void) { return tmd::now_s(); }

void tmd_serve_options_default(tmd_serve_options *o)
{
  if (!o) return;
  o->slots = 2;
  o->max_batch_atoms = 65536;
  o->max_batch_structs = 512;
  o->hbm_fraction = 0.25;
  o->use_graph = 1;
  o->autotune = 1;
  o->max_queue = 65536;
}

int tmd_serve_start(tmd_ctx *ctx, const tmd_serve_options *opt)
{
  if (!ctx) return TMD_ERR_ARG;
  if (ctx->server) return serve_fail(ctx, TMD_ERR_STATE, "tmd_serve_start: already serving");
  tmd_serve_options o;
  tmd_serve_options_default(&o);
  if (opt) o = *opt;
  if (o.slots < 1 || o.slots > 8 || o.max_batch_atoms < 16 || o.max_batch_atoms > kMaxAtoms ||
      o.max_batch_structs < 1 || !(o.hbm_fraction > 0 && o.hbm_fraction <= 0.9) || o.max_queue < 1)
    return serve_fail(ctx, TMD_ERR_ARG, "tmd_serve_start: invalid options");
  if (hipError_t de = hipSetDevice(ctx->device); de != hipSuccess)
    return hip_fail(ctx, de, "hipSetDevice");
  std::unique_ptr<Server> S(new Server());
  S->ctx = ctx;
  S->opt = o;
  S->cut = ctx->model.rc;
  const int nelt = ctx->model.nelt;
  // admission against free device memory
  size_t free_b = 0, total_b = 0;
  hipError_t e = hipMemGetInfo(&free_b, &total_b);
  if (e != hipSuccess) return hip_fail(ctx, e, "tmd_serve_start: hipMemGetInfo");
  const int listcap0 = 160;
  const double budget = o.hbm_fraction * (double) free_b / o.slots;
  long atoms = o.max_batch_atoms;
  const double per = (double) bytes_per_atom(ctx, listcap0);
  while (atoms > 1024 && per * (atoms + nelt * kRowTile) + 1e6 > budget) atoms /= 2;
  if (per * (atoms + nelt * kRowTile) > budget)
    return serve_fail(ctx, TMD_ERR_NOMEM, "tmd_serve_start: not enough free device memory");
  S->full.atoms = (int) atoms;
  S->full.rows = (int) ((atoms + nelt * kRowTile + kRowTile - 1) / kRowTile * kRowTile);
  S->full.bins = (int) atoms;
  S->full.structs = (int) std::min<long>(o.max_batch_structs, atoms);
  S->full.nelt = nelt;
  for (int q = 0; q < o.slots; q++) {
    S->slots.emplace_back(new Slot());
    int rc = create_slot(S.get(), *S->slots.back(), listcap0);
    if (rc != TMD_OK) {
      for (auto &sl : S->slots) destroy_slot(*sl);
      return rc;
    }
  }
  Server *raw = S.release();
  raw->worker = std::thread(worker_main, raw);
  ctx->server = raw;
  return TMD_OK;
}

int tmd_serve_submit(tmd_ctx *ctx, const tmd_structure *s, tmd_ticket *ticket)
{
  if (!ctx || !s || !ticket || !s->x || !s->type || s->natoms < 1)
    return serve_fail(ctx, TMD_ERR_ARG, "tmd_serve_submit: invalid arguments");
  Server *S = ctx->server;
  if (!S) return serve_fail(ctx, TMD_ERR_STATE, "tmd_serve_submit: not serving");
  if (s->natoms > S->full.atoms)
    return serve_fail(ctx, TMD_ERR_NOMEM, "tmd_serve_submit: structure larger than a batch slot");
  std::unique_ptr<Request> r(new Request());
  std::string err;
  int rc = setup_cell(s->cell, S->cut, r->cell, err);
  if (rc != TMD_OK) return serve_fail(ctx, rc, "tmd_serve_submit: " + err);
  // at most one bin per atom: fewer, larger bins never change the list
  double w[3];
  {
    const double *h = r->cell.h;
    const double V = cell_volume(h);
    auto nrm = [](double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); };
    w[0] = V / nrm(h[4] * h[8] - h[5] * h[7], h[5] * h[6] - h[3] * h[8], h[3] * h[7] - h[4] * h[6]);
    w[1] = V / nrm(h[7] * h[2] - h[8] * h[1], h[8] * h[0] - h[6] * h[2], h[6] * h[1] - h[7] * h[0]);
    w[2] = V / nrm(h[1] * h[5] - h[2] * h[4], h[2] * h[3] - h[0] * h[5], h[0] * h[4] - h[1] * h[3]);
  }
  while ((long) r->cell.nb[0] * r->cell.nb[1] * r->cell.nb[2] > s->natoms) {
    int k = 0;
    for (int q = 1; q < 3; q++)
      if (r->cell.nb[q] > r->cell.nb[k]) k = q;
    r->cell.nb[k] = (r->cell.nb[k] + 1) / 2;
    r->cell.reach[k] = (int) std::ceil(S->cut * r->cell.nb[k] / w[k] + 1e-6);
  }
  r->nbins = r->cell.nb[0] * r->cell.nb[1] * r->cell.nb[2];
  const double V = cell_volume(r->cell.h);
  const double expect = 4.0 / 3.0 * M_PI * S->cut * S->cut * S->cut * s->natoms / V;
  r->listcap = ((int) std::ceil(std::max(expect * 1.35, expect + 32.0)) + 7) & ~7;
  r->natoms = s->natoms;
  r->pos.resize(s->natoms);
  for (int i = 0; i < s->natoms; i++) {
    int t = s->type[i];
    int el = (t >= 1 && t <= ctx->ntypes) ? ctx->type_map[t] : -1;
    if (el < 0) return serve_fail(ctx, TMD_ERR_ARG, "tmd_serve_submit: atom type not mapped to a model element");
    r->pos[i] = make_double4(s->x[3 * i], s->x[3 * i + 1], s->x[3 * i + 2], (double) el);
  }
  r->f_out = s->f;
  r->eatom_out = s->eatom;
  r->t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->fatal.empty()) return serve_fail(ctx, TMD_ERR_DEVICE, S->fatal);
    if (S->stopping) return serve_fail(ctx, TMD_ERR_STATE, "tmd_serve_submit: stopping");
    if ((int) S->queue.size() >= S->opt.max_queue)
      return serve_fail(ctx, TMD_ERR_CAPACITY, "tmd_serve_submit: queue full");
    r->id = S->next_id++;
    *ticket = r->id;
    S->reqs[r->id] = r.get();
    S->queue.push_back(r.release());
  }
  S->cv.notify_one();
  return TMD_OK;
}

static void fill_result(const tmd::Request *r, tmd_result *res)
{
  if (!res) return;
  res->status = r->status;
  res->energy = r->energy;
  for (int q = 0; q < 6; q++) res->virial[q] = r->virial[q];
  res->t_submit = r->t_submit;
  res->t_start = r->t_start;
  res->t_done = r->t_done;
  res->batch_structs = r->batch_structs;
  res->batch_atoms = r->batch_atoms;
}

int tmd_serve_poll(tmd_ctx *ctx, tmd_ticket ticket, tmd_result *res)
{
  if (!ctx || !ctx->server) return TMD_ERR_STATE;
  Server *S = ctx->server;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return TMD_ERR_ARG;
  if (it->second->status == TMD_PENDING) return TMD_PENDING;
  fill_result(it->second, res);
  delete it->second;
  S->reqs.erase(it);
  return TMD_OK;
}

int tmd_serve_wait(tmd_ctx *ctx, tmd_ticket ticket, tmd_result *res)
{
  if (!ctx || !ctx->server) return TMD_ERR_STATE;
  Server *S = ctx->server;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return TMD_ERR_ARG;
  Request *r = it->second;
  S->done_cv.wait(lk, [r] { return r->status != TMD_PENDING; });
  fill_result(r, res);
  delete r;
  S->reqs.erase(ticket);
  return TMD_OK;
}

int tmd_serve_stop(tmd_ctx *ctx)
{
  if (!ctx || !ctx->server) return TMD_ERR_STATE;
  std::string fatal;
  {
    std::lock_guard<std::mutex> lk(ctx->server->mu);
    fatal = ctx->server->fatal;
  }
  server_destroy(ctx->server);
  ctx->server = nullptr;
  if (!fatal.empty()) return serve_fail(ctx, TMD_ERR_DEVICE, fatal);
  return TMD_OK;
}

int tmd_serve_replay(tmd_ctx *ctx, int n, const tmd_structure *s, const double *arrival,
                     tmd_result *res)
{
  if (!ctx || n < 0 || (n > 0 && (!s || !arrival || !res))) return TMD_ERR_ARG;
  if (!ctx->server) return serve_fail(ctx, TMD_ERR_STATE, "tmd_serve_replay: not serving");
  std::vector<tmd_ticket> tk(n, 0);
  const double t0 = now_s();
  for (int k = 0; k < n; k++) {
    const double due = t0 + arrival[k];
    for (double t = now_s(); t < due; t = now_s())
      if (due - t > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
    int rc;
    while ((rc = tmd_serve_submit(ctx, &s[k], &tk[k])) == TMD_ERR_CAPACITY)
      std::this_thread::sleep_for(std::chrono::microseconds(100));  // queue full: back off
    if (rc != TMD_OK) {
      res[k] = tmd_result{};
      res[k].status = rc;
      tk[k] = 0;
    }
  }
  for (int k = 0; k < n; k++)
    if (tk[k]) {
      int rc = tmd_serve_wait(ctx, tk[k], &res[k]);
      if (rc != TMD_OK) res[k].status = rc;
    }
  return TMD_OK;
}

int tmd_serve_stats(tmd_ctx *ctx, double *st)
{
  if (!ctx || !ctx->server || !st) return TMD_ERR_ARG;
  Server *S = ctx->server;
  for (int q = 0; q < 24; q++) st[q] = 0.0;
  std::lock_guard<std::mutex> lk(S->mu);
  st[0] = (double) S->batches;
  st[1] = (double) S->structs;
  st[2] = (double) S->atoms;
  st[3] = (double) S->captures;
  st[4] = (double) S->reruns;
  st[5] = S->gpu_ms;
  st[6] = (double) S->full.atoms;
  st[7] = (double) S->slots.size();
  if (!S->slots.empty()) {
    const Slot &sl = *S->slots[0];
    for (int b = 0; b < sl.nbuckets && b < 8; b++) {
      st[8 + b] = (double) sl.buckets[b].rtile;
      st[16 + b] = (double) sl.buckets[b].caps.atoms;
    }
  }
  return TMD_OK;
}

}  // extern "C"
