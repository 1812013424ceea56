/*
 * tensormd_amd.h -- C ABI of the MI355X (gfx950) implementation of LAMMPS
 * `pair_style tensormd`.
 *
 * Plain C types only: caller-owned host buffers, library-owned device
 * buffers, integer status codes (TMD_OK == 0), never exit().  One context per
 * model; a context and everything created from it must be used from one host
 * thread at a time.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   tmd_create          PairTensorMD::coeff + read_potential_file +
 *                       TensorMD::read_npz/setup_global/setup_interpolation
 *                       (src/ML-TENSORMD/pair_tensormd.cpp:214-317, :376-399;
 *                        src/ML-TENSORMD/tensormd.cpp:108-280, :437-446)
 *   tmd_cutoff          PairTensorMD::init_one (pair_tensormd.cpp:336-341)
 *   tmd_masses          the atom->set_mass loop in coeff (pair_tensormd.cpp:302-312)
 *   tmd_type_map        PairTensorMD::setup_eltmap (pair_tensormd.cpp:349-370)
 *   tmd_compute_list    PairTensorMD::compute (pair_tensormd.cpp:141-184) with
 *                       TensorMD::calc_tensor_density/run (tensormd.cpp:368-433,
 *                       :450-615) and Pair::virial_fdotr_compute
 *                       (src/pair.cpp:1735-1783)
 *   tmd_compute_list_serial  the same, with the list kept on the device
 *                       between LAMMPS neighbor builds (the reference uploads
 *                       it every step, gpu/cuda/kernel.cu:1500-1521)
 *   tmd_set_electron_temperature, tmd_tdnp_results  the TDNP state of
 *                       PairTensorMD (etemp, eentropy, free_energy, satom,
 *                       aatom; pair_tensormd.cpp:123-137, :171-178) and
 *                       TDNP::forward/backward (src/ML-TENSORMD/tdnp.cpp:150-218)
 *   tmd_descriptors_list ComputeTensorMDDensity::compute_peratom
 *                       (src/ML-TENSORMD/compute_tensormd_density.cpp:129-205)
 *   tmd_system_*,       the LAMMPS step loop around the pair style:
 *   tmd_md_run,         Verlet::run (src/verlet.cpp:246-350), NPairFullBin::build
 *   tmd_md_step_async   (src/npair_full_bin.cpp:68-126), Neighbor::decide
 *                       (src/neighbor.cpp:2293-2360), FixNVE
 *                       (src/fix_nve.cpp:64-130), kept on the device
 *   tmd_radial_table    the table of FeaturePotential::setup_ration
 *                       (src/ML-TENSORMD/fnn.cpp:77-105), built on the host
 *                       or, with tmd_options.table_device, on the device as
 *                       forward_fnn_GPU does (gpu/cuda/nnp.cu:59-116)
 *   tmd_destroy         PairTensorMD::~PairTensorMD (pair_tensormd.cpp:68-83)
 *   tmd_serve_*         no reference counterpart: batched single-point
 *                       requests (each one a PairTensorMD::compute of a
 *                       small periodic structure) with continuous batching
 */
#ifndef TENSORMD_AMD_H
#define TENSORMD_AMD_H

#include <stddef.h> /* size_t */

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TMD_OK 0
#define TMD_ERR_ARG (-1)      /* invalid argument / input (LAMMPS: error->all) */
#define TMD_ERR_MODEL (-2)    /* unreadable or unsupported model file */
#define TMD_ERR_DEVICE (-3)   /* HIP runtime error */
#define TMD_ERR_NOMEM (-4)    /* device allocation failed or over budget */
#define TMD_ERR_CAPACITY (-5) /* a fixed capacity was exceeded */
#define TMD_ERR_STATE (-6)    /* call not valid in the current state */
#define TMD_PENDING 1         /* tmd_serve_poll: request not finished yet */

/* ---- MLP arithmetic ------------------------------------------------------ */
/* Descriptors and forces are FP32 with FP64 displacements and FP64
 * energy/virial sums in every mode; the mode selects the MLP GEMMs. */
#define TMD_MLP_FP32 0 /* plain FP32 path */
#define TMD_MLP_BF16 1 /* BF16 operands, FP32 accumulation */
#define TMD_MLP_FP8 2  /* FP8 e4m3 operands, per-output-channel FP32 scales */
#define TMD_MLP_BF16X3 3 /* split BF16: operands as hi + lo BF16 pairs, three
                            BF16 MFMA products per product, FP32 accumulation
                            (about 16 significant bits) */

typedef struct tmd_ctx tmd_ctx;
typedef struct tmd_system tmd_system;

typedef struct tmd_options {
  int mlp_precision; /* TMD_MLP_* */
  double interp_dx;  /* pair_coeff "interp dx"; must be > 0 (the only radial
                        path the reference wires, tensormd.cpp:419-423) */
  int cmax_exact;    /* pair_coeff "cmax exact|default" (1|0).  Accepted for
                        compatibility; padding does not exist here, so it
                        has no effect on results. */
  int device;        /* HIP device ordinal; -1 = current device */
  int table_device;  /* 1: build the radial table on the device at load, as
                        forward_fnn_GPU (gpu/cuda/nnp.cu:59-116) does; 0
                        (default): on the host.  Both in FP64; they differ
                        only in the rounding of exp/log1p/tanh/cos/pow. */
} tmd_options;

/* Fill `opt` with defaults: FP32, interp_dx 0.001, cmax exact, device -1. */
void tmd_options_default(tmd_options *opt);

/* Load a TensorMD model (.npz, reference schema) and map LAMMPS atom types
 * 1..ntypes to model elements by symbol: elem_symbols[t-1] is the symbol
 * given for type t on the pair_coeff line ("NULL" leaves a type unmapped).
 * Builds the radial interpolation table and uploads all parameters. */
int tmd_create(const char *npz_path, const char *const *elem_symbols, int ntypes,
               const tmd_options *opt, tmd_ctx **out);
void tmd_destroy(tmd_ctx *ctx);

/* Last error message of `ctx` (or of the calling thread if ctx is NULL). */
const char *tmd_last_error(const tmd_ctx *ctx);

int tmd_cutoff(const tmd_ctx *ctx, double *rc);
/* mass_by_type[t] for t = 1..ntypes (index 0 unused); 0 for unmapped types */
int tmd_masses(const tmd_ctx *ctx, double *mass_by_type);
/* map[t] = model element index for type t (0-based), -1 if unmapped */
int tmd_type_map(const tmd_ctx *ctx, int *map);
/* nelt, num_filters, max_moment, descriptor width, mlp layer count, as in the
 * model file (max_moment 0..5; a model with fewer than 4 moments is evaluated
 * with zero first-layer weights on the moment channels it lacks) */
int tmd_model_info(const tmd_ctx *ctx, int *info5);
/* The FP64 radial table the context uploaded (FeaturePotential::setup_ration,
 * fnn.cpp:77-105): n grid points x num_filters x {y, dy/dr slope, c1, c2}
 * (the Kerley form of ration.hpp:166-225), step dx.  Writes *n and *dx and,
 * when tab is non-NULL and cap >= n * num_filters * 4, the table. */
int tmd_radial_table(const tmd_ctx *ctx, double *tab, long cap, int *n, double *dx);

/* ---- Pair::compute drop-in ------------------------------------------------
 * x[nall][3] (local + ghost), type[nall] (1-based), full neighbor list
 * (ilist[inum], numneigh[i], firstneigh[i][*]; the top 3 bits of an entry
 * are LAMMPS special-bond flags and are masked off).
 * Outputs: f[nall][3] is OVERWRITTEN (reference kernel.cu:1766,1783);
 * eatom[nlocal] and vatom[nall][6] are written if non-NULL (vatom is
 * zeroed first; order xx,yy,zz,xy,xz,yz); *energy = sum of scaled atomic
 * energies; virial[6] (if non-NULL) = sum_i x_i (x) f_i over nall atoms
 * (the fdotr virial).  Energies and virials are FP64 sums. */
int tmd_compute_list(tmd_ctx *ctx, int nlocal, int nall, const double *x, const int *type,
                     int inum, const int *ilist, const int *numneigh,
                     const int *const *firstneigh, double scale, double *f, double *eatom,
                     double *vatom, double *energy, double *virial);

/* tmd_compute_list for a caller that knows when its neighbor list was
 * rebuilt (LAMMPS: neighbor->lastcall, the step of the last build; the
 * reference re-packs and re-uploads the whole list every step,
 * gpu/cuda/kernel.cu:1500-1521).  When list_serial (>= 0) equals that of
 * the context's previous tmd_compute_list_serial call, with the same nlocal,
 * nall and inum and no other list-mode call in between, the list (ilist,
 * numneigh, firstneigh and the types of the atoms in ilist) is taken as
 * unchanged and stays on the device: only x is packed and uploaded.  Any
 * other serial packs and uploads everything.  Outputs as tmd_compute_list. */
int tmd_compute_list_serial(tmd_ctx *ctx, long long list_serial, int nlocal, int nall,
                            const double *x, const int *type, int inum, const int *ilist,
                            const int *numneigh, const int *const *firstneigh, double scale,
                            double *f, double *eatom, double *vatom, double *energy,
                            double *virial);

/* tmd_compute_list with no allocation in the call.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as tmd_compute_list; its device buffers are
 * windows of a fixed 64 MiB device array (one per device, held until the
 * process ends) and its host staging a fixed 32 MiB static array.  A call
 * that needs more (about 10,000 local atoms of a 4-element k = 24 model)
 * returns TMD_ERR_CAPACITY.  Calls are serialized by a process-wide lock;
 * the arrays keep the last call's contents. */
int tmd_compute_list_noalloc(tmd_ctx *ctx, int nlocal, int nall, const double *x,
                             const int *type, int inum, const int *ilist, const int *numneigh,
                             const int *const *firstneigh, double scale, double *f,
                             double *eatom, double *vatom, double *energy, double *virial);

/* ---- TDNP: temperature-dependent models (tdnp = 1) ----------------------------
 * The electron temperature in eV (PairTensorMD::etemp: the pair_coeff
 * keyword "etemp T" gives T / 11604.522060401008 eV, pair_tensormd.cpp:33,
 * 245-247; fix tensormd/etemp resets it every N steps,
 * fix_tensormd_etemp.cpp:80-88).  0 until set.  A TDNP model's atomic energy
 * is the electron internal energy u and its forces are -dA/dx with the free
 * energy A = U - T S (tdnp.cpp:150-218).  Accepted for any model; only TDNP
 * models read it. */
int tmd_set_electron_temperature(tmd_ctx *ctx, double etemp_ev);
/* The electron entropy S and free energy A (scaled sums; PairTensorMD::
 * eentropy and free_energy, read by compute tensormd/tdpe) and, where
 * non-NULL, the per-atom satom/aatom[nlocal] (zero for atoms not in ilist)
 * of the context's last tmd_compute_list or tmd_compute_list_serial call.
 * TMD_ERR_MODEL for a model without TDNP, TMD_ERR_STATE before such a call. */
int tmd_tdnp_results(const tmd_ctx *ctx, double *entropy, double *free_energy, double *satom,
                     double *aatom);

/* `compute tensormd/density`: the descriptor row of every atom in ilist,
 * G[nlocal][width] in the model file's width and [b][k][m] layout (rows of
 * atoms not in ilist are zero), from the same inputs as tmd_compute_list.
 * Computed in FP32 on the device, returned as double. */
int tmd_descriptors_list(tmd_ctx *ctx, int nlocal, int nall, const double *x, const int *type,
                         int inum, const int *ilist, const int *numneigh,
                         const int *const *firstneigh, double *G);

/* ---- device-resident periodic system ----------------------------------------
 * cell = rows a, b, c (Cartesian, 9 doubles); any triclinic cell.  Positions
 * and velocities stay on the device; the neighbor list (cutoff rc + skin)
 * is rebuilt on the device when an atom has moved more than skin/2 since the
 * last build (LAMMPS "neigh_modify every 1 delay 0 check yes"). */
int tmd_system_create(tmd_ctx *ctx, int natoms, const double *x, const int *type,
                      const double *cell, double skin, tmd_system **out);
void tmd_system_destroy(tmd_system *sys);
int tmd_system_set_positions(tmd_system *sys, const double *x);
int tmd_system_set_velocities(tmd_system *sys, const double *v);
/* Evaluate energy/forces/virial at the current positions. */
int tmd_system_compute(tmd_system *sys);
/* Copy results to the host; any pointer may be NULL.  virial = sum over
 * pairs of -d (x) F (equals the fdotr virial over periodic images). */
int tmd_system_get(tmd_system *sys, double *x, double *v, double *f, double *eatom,
                   double *energy, double *virial);
/* nsteps of velocity-Verlet NVE (metal units) with device neighbor
 * rebuilds; epot/ekin[nsteps] receive per-step energies if non-NULL
 * (read back once at the end).  */
int tmd_md_run(tmd_system *sys, int nsteps, double dt, double *epot, double *ekin);
/* One step of tmd_md_run enqueued without waiting for the device: the
 * per-step call of a host loop (fix nve/tensormd/amd, in place of the
 * initial/final_integrate pair of FixNVE, fix_nve.cpp:64-130, around
 * Verlet's force call), so that the host's own work overlaps the step.  The
 * step's neighbor-list capacity check, and any device error, are taken up
 * by the next call on sys; a step whose list overflowed is then run again
 * from the state before it with a larger list, as tmd_md_run does. */
int tmd_md_step_async(tmd_system *sys, double dt);
/* Run nsteps MD steps with HIP events recorded on the library stream
 * between phases; ms[0..6] = mean ms per step spent in: neighbor
 * maintenance, descriptor kernel, MLP kernel, force kernel, FP64
 * reductions (with the force-buffer clears), integration, and the whole
 * step. */
int tmd_md_profile(tmd_system *sys, int nsteps, double dt, double *ms);
/* Sampled kernel timing inside tmd_md_run.  No counterpart in the reference
 * (benchmark instrumentation: kernel durations measured over a timed run).
 * every > 0: every every-th step of the following tmd_md_run calls is
 * launched kernel by kernel with HIP events on the library stream around
 * the descriptor, MLP and force kernels, the other steps replay the
 * captured graph; every = 0 turns it off.  Clears the samples. */
int tmd_system_set_kernel_timing(tmd_system *sys, int every);
/* ms[0..2] = mean ms of the descriptor, MLP and force kernels over the
 * sampled steps; *kept = mean kept neighbors (within rc) per atom at those
 * steps (NULL allowed); *nsamples = number of sampled steps (NULL allowed).
 * Steps of a run segment that is rerun after a list overflow are not
 * counted. */
int tmd_system_kernel_timing(const tmd_system *sys, double *ms, double *kept, long long *nsamples);
/* Replay each MD step as one captured HIP graph (default 1) or launch its
 * kernels one by one (0); results are the same either way. */
int tmd_system_set_graph(tmd_system *sys, int enable);
/* Number of neighbor-list builds done so far (up to the last step whose
 * check was taken up; see tmd_md_step_async). */
int tmd_system_nbuilds(const tmd_system *sys);
/* Export the current full neighbor list (cutoff rc + skin): off[natoms+1],
 * and up to cap entries of (j, image shift s0 s1 s2).  Returns the total
 * entry count in *total.  Order within an atom is unspecified. */
int tmd_system_neighbors(tmd_system *sys, long cap, long *off, int *j, int *shift, long *total);
/* Descriptor rows G[natoms][width] and dE/dG of the last compute (FP32,
 * atom order, width and [b][k][m] layout of the model file), for
 * kernel-level parity tests.  Where the BF16 MLP takes its descriptor rows
 * as BF16 (tanh networks of the benchmark layer shapes), G is those rows,
 * rounded to BF16 (nearest even) and widened exactly. */
int tmd_system_descriptors(tmd_system *sys, float *G, float *dEdG);

/* ---- serving: batched single-point requests --------------------------------
 * A request is one periodic structure (any triclinic cell at least rc/2
 * wide in every direction); its energy, virial and forces are those of a
 * tmd_system with skin 0.  Requests are packed into batches (continuous
 * batching): a library thread launches a batch on a free slot (its own
 * stream and device buffers) as soon as one is free, taking every queued
 * request that fits the slot's capacities, so requests that arrive while
 * others run are admitted as slots free up and batches on different slots
 * overlap on the device.  Slot capacities are fixed at tmd_serve_start from
 * the free device memory.  A batch runs as one captured HIP graph per
 * capacity bucket (sizes are read on the device), and the MLP row tile of
 * each bucket is chosen by timing the candidates on first use.
 *
 * Energies and virials of a structure do not depend on what else is in the
 * batch (fixed-order sums); forces agree to FP64 rounding (atomics). */
typedef long long tmd_ticket;

typedef struct tmd_serve_options {
  int slots;             /* batches in flight (default 2) */
  int max_batch_atoms;   /* atoms per batch (default 65536; lowered to fit) */
  int max_batch_structs; /* structures per batch (default 512) */
  double hbm_fraction;   /* share of free device memory for the slots (0.25) */
  int use_graph;         /* 1: one captured graph per capacity bucket (default) */
  int autotune;          /* 1: time the MLP row tiles per bucket (default) */
  int max_queue;         /* queued requests before submit refuses (65536) */
} tmd_serve_options;

typedef struct tmd_structure {
  int natoms;
  const double *x;   /* [natoms][3], copied at submit */
  const int *type;   /* [natoms], 1-based LAMMPS types */
  double cell[9];    /* rows a, b, c */
  double *f;         /* out [natoms][3] or NULL; written before the request
                        completes, must stay valid until then */
  double *eatom;     /* out [natoms] or NULL (same rule) */
} tmd_structure;

typedef struct tmd_result {
  int status;          /* TMD_OK or an error code for this request */
  double energy;       /* sum of atomic energies (FP64) */
  double virial[6];    /* sum over pairs of -d (x) F: xx yy zz xy xz yz */
  double t_submit;     /* tmd_clock() at submit */
  double t_start;      /* when its batch was launched */
  double t_done;       /* when its results were written */
  int batch_structs;   /* size of the batch it ran in */
  int batch_atoms;
} tmd_result;

void tmd_serve_options_default(tmd_serve_options *opt);
/* Start the request loop of `ctx` (one per context). */
int tmd_serve_start(tmd_ctx *ctx, const tmd_serve_options *opt);
/* Queue one structure; *ticket identifies it.  Refuses (TMD_ERR_CAPACITY)
 * when max_queue requests are waiting and rejects structures larger than a
 * slot (TMD_ERR_NOMEM) or with invalid types or cells (TMD_ERR_ARG). */
int tmd_serve_submit(tmd_ctx *ctx, const tmd_structure *s, tmd_ticket *ticket);
/* TMD_PENDING while running; TMD_OK once done (fills *res and releases the
 * ticket); TMD_ERR_ARG for an unknown ticket. */
int tmd_serve_poll(tmd_ctx *ctx, tmd_ticket ticket, tmd_result *res);
/* Block until the request is done, then as tmd_serve_poll. */
int tmd_serve_wait(tmd_ctx *ctx, tmd_ticket ticket, tmd_result *res);
/* Finish every queued request, then stop the loop. */
int tmd_serve_stop(tmd_ctx *ctx);
/* st[0] batches, st[1] structures, st[2] atoms, st[3] graph captures,
 * st[4] batches re-run after a neighbor-list overflow, st[5] summed device
 * time of the batches (ms), st[6] slot capacity in atoms, st[7] slots,
 * st[8..15] MLP row tile chosen per bucket (largest bucket first, 0 = not
 * tuned yet), st[16..23] bucket capacities in atoms. */
int tmd_serve_stats(tmd_ctx *ctx, double *st24);
/* Load generator: submit s[k] at arrival[k] seconds after the call (from
 * the calling thread; a full queue delays the submit), then wait for every
 * request; res[k] receives each result (status carries submit errors). */
int tmd_serve_replay(tmd_ctx *ctx, int n, const tmd_structure *s, const double *arrival,
                     tmd_result *res);
/* Monotonic clock in seconds (the clock of tmd_result times). */
double tmd_clock(void);

/* ---- measurement ----------------------------------------------------------
 * Device-to-device copy of `bytes` (read + write) repeated `iters` times on
 * the current device; *gbps = achieved read+write GB/s.  Re-measures the
 * HBM roofline and calibrates memory counters.  No reference counterpart. */
int tmd_bench_copy(size_t bytes, int iters, double *gbps);
/* Free and total device memory (bytes) of the current device, for
 * footprint reports.  No reference counterpart. */
int tmd_device_memory(size_t *free_bytes, size_t *total_bytes);

#ifdef __cplusplus
}
#endif
#endif
