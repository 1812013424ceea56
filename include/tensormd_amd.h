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
 *   tmd_system_*,       the LAMMPS step loop around the pair style:
 *   tmd_md_run          Verlet::run (src/verlet.cpp:246-350), NPairFullBin::build
 *                       (src/npair_full_bin.cpp:68-126), Neighbor::decide
 *                       (src/neighbor.cpp:2293-2360), FixNVE
 *                       (src/fix_nve.cpp:64-130), kept on the device
 *   tmd_destroy         PairTensorMD::~PairTensorMD (pair_tensormd.cpp:68-83)
 */
#ifndef TENSORMD_AMD_H
#define TENSORMD_AMD_H

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

/* ---- MLP arithmetic ------------------------------------------------------ */
/* Descriptors and forces are FP32 with FP64 displacements and FP64
 * energy/virial sums in every mode; the mode selects the MLP GEMMs. */
#define TMD_MLP_FP32 0 /* plain FP32 path */
#define TMD_MLP_BF16 1 /* BF16 operands, FP32 accumulation */
#define TMD_MLP_FP8 2  /* FP8 e4m3 operands, per-output-channel FP32 scales */

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
/* nelt, num_filters, max_moment, descriptor width, mlp layer count */
int tmd_model_info(const tmd_ctx *ctx, int *info5);

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
/* Run nsteps MD steps with HIP events recorded on the library stream
 * between phases; ms[0..6] = mean ms per step spent in: neighbor
 * maintenance, descriptor kernel, MLP GEMMs, force kernel, FP64
 * reductions, integration, and the whole step. */
int tmd_md_profile(tmd_system *sys, int nsteps, double dt, double *ms);
/* Replay each MD step as one captured HIP graph (default 1) or launch its
 * kernels one by one (0); results are the same either way. */
int tmd_system_set_graph(tmd_system *sys, int enable);
/* Number of neighbor-list builds done so far. */
int tmd_system_nbuilds(const tmd_system *sys);
/* Export the current full neighbor list (cutoff rc + skin): off[natoms+1],
 * and up to cap entries of (j, image shift s0 s1 s2).  Returns the total
 * entry count in *total.  Order within an atom is unspecified. */
int tmd_system_neighbors(tmd_system *sys, long cap, long *off, int *j, int *shift, long *total);
/* Descriptor rows G[natoms][width] and dE/dG of the last compute (FP32,
 * atom order), for kernel-level parity tests. */
int tmd_system_descriptors(tmd_system *sys, float *G, float *dEdG);

/* ---- measurement ----------------------------------------------------------
 * Device-to-device copy of `bytes` (read + write) repeated `iters` times on
 * the current device; *gbps = achieved read+write GB/s.  Re-measures the
 * HBM roofline and calibrates memory counters.  No reference counterpart. */
int tmd_bench_copy(size_t bytes, int iters, double *gbps);

#ifdef __cplusplus
}
#endif
#endif
