/* ----------------------------------------------------------------------
   fix nve/tensormd/amd: the velocity-Verlet step of fix nve
   (src/fix_nve.cpp:64-130 of the reference tree) together with the pair
   evaluation, run as one device step per LAMMPS step.

   setup():             a tmd_system (the pair style's context) from the
                        host x, v and types; the pair style switches to
                        reporting device results
   initial_integrate(): one device step enqueued (tmd_md_step_async):
                        v += dtf f/m, x += dt v, the device neighbor check
                        (and rebuild), forces, energy and virial at the new
                        positions, v += dtf f/m
   pair compute:        PairTensorMDAMD::compute asks pair_results() for the
                        energy, virial and per-atom energies of the step
   end_of_step():       on the steps LAMMPS' Output writes (output->next),
                        x, v and f go back to the host
   post_run():          x, v and f go back to the host; the system is freed

   Host x stays what it was at the last write-back between those steps, so
   LAMMPS' own displacement check sees no motion and its host neighbor list
   is not rebuilt; after a write-back the next check may rebuild it once.
   Written-back positions keep the host image flags: x = device position +
   (image at setup - current image) box vectors, which is what LAMMPS holds
   between its own rebuilds, and the ghosts are refreshed by forward_comm.
------------------------------------------------------------------------- */

#include "fix_nve_tensormd_amd.h"

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "force.h"
#include "modify.h"
#include "neighbor.h"
#include "output.h"
#include "pair_tensormd_amd.h"
#include "update.h"
#include "utils.h"

#include "tensormd_amd.h"

#include <cmath>
#include <cstring>

using namespace LAMMPS_NS;
using namespace FixConst;

FixNVETensorMDAMD::FixNVETensorMDAMD(LAMMPS *lmp, int narg, char **arg) :
    Fix(lmp, narg, arg), pair(nullptr), sys(nullptr), sys_ctx(nullptr), dtv(0.0), natoms(0),
    maxtag(0)
{
  if (narg != 3) error->all(FLERR, "Illegal fix nve/tensormd/amd command");
  if (igroup != 0) error->all(FLERR, "fix nve/tensormd/amd integrates every atom: use group all");
  time_integrate = 1;
  for (double &c : cell) c = 0.0;
}

FixNVETensorMDAMD::~FixNVETensorMDAMD()
{
  // a run that ended without post_run (an error) may leave the system; it
  // belongs to the pair style's context, so it is freed only while that
  // context is still the pair style's
  if (sys && force->pair == pair && pair->context() == sys_ctx) release();
}

int FixNVETensorMDAMD::setmask()
{
  return INITIAL_INTEGRATE | END_OF_STEP | POST_RUN;
}

void FixNVETensorMDAMD::device_fail(int code, const char *where)
{
  error->one(FLERR, "fix nve/tensormd/amd {}: {} (status {})", where, tmd_last_error(sys_ctx), code);
}

void FixNVETensorMDAMD::init()
{
  pair = dynamic_cast<PairTensorMDAMD *>(force->pair);
  if (!pair || !pair->context()) error->all(FLERR, "fix nve/tensormd/amd needs pair_style tensormd/amd with its pair_coeff");
  if (comm->nprocs != 1) error->all(FLERR, "fix nve/tensormd/amd runs on one MPI rank");
  if (!domain->xperiodic || !domain->yperiodic || !domain->zperiodic)
    error->all(FLERR, "fix nve/tensormd/amd needs a box periodic in x, y and z");
  if (strcmp(update->unit_style, "metal") != 0) error->all(FLERR, "fix nve/tensormd/amd needs units metal");
  if (!atom->tag_enable) error->all(FLERR, "fix nve/tensormd/amd needs atom IDs");
  if (atom->rmass) error->all(FLERR, "fix nve/tensormd/amd uses per-type masses, not per-atom masses");
  if (utils::strmatch(update->integrate_style, "^respa")) error->all(FLERR, "fix nve/tensormd/amd does not support run_style respa");
  if (pair->tdpe_flag)
    error->all(FLERR, "fix nve/tensormd/amd: the TDNP entropy and free energy (pair_coeff etemp) are computed on the host path only");
  // the device step uses the model's masses
  const int nt = atom->ntypes;
  std::vector<double> mass(nt + 1, 0.0);
  std::vector<int> tmap(nt + 1, -1);
  if (tmd_masses(pair->context(), mass.data()) != TMD_OK || tmd_type_map(pair->context(), tmap.data()) != TMD_OK)
    error->all(FLERR, "fix nve/tensormd/amd: {}", tmd_last_error(pair->context()));
  for (int t = 1; t <= nt; t++)
    if (tmap[t] >= 0 && std::fabs(atom->mass[t] - mass[t]) > 1e-12 * mass[t])
      error->all(FLERR, "fix nve/tensormd/amd: the mass of type {} ({}) differs from the model's ({})", t,
                 atom->mass[t], mass[t]);
  // every other fix that acts during a step would see host copies that
  // are not updated between output steps
  const int during_step = INITIAL_INTEGRATE | POST_INTEGRATE | PRE_EXCHANGE | PRE_NEIGHBOR | POST_NEIGHBOR |
      PRE_FORCE | PRE_REVERSE | POST_FORCE | FINAL_INTEGRATE | END_OF_STEP;
  for (int i = 0; i < modify->nfix; i++) {
    Fix *f = modify->fix[i];
    if (f == this) continue;
    if (f->time_integrate || (modify->fmask[i] & during_step))
      error->all(FLERR, "fix nve/tensormd/amd keeps x, v and f on the device during a step; fix {} ({}) acts on them",
                 f->id, f->style);
  }
  dtv = update->dt;
}

void FixNVETensorMDAMD::reset_dt()
{
  dtv = update->dt;
}

void FixNVETensorMDAMD::setup(int /*vflag*/)
{
  if (sys) release();
  natoms = atom->nlocal;
  if ((bigint) natoms != atom->natoms) error->all(FLERR, "fix nve/tensormd/amd: every atom must be on this rank");
  if (natoms < 1) error->all(FLERR, "fix nve/tensormd/amd: no atoms");
  // cell rows a = (xprd, 0, 0), b = (xy, yprd, 0), c = (xz, yz, zprd)
  const double c9[9] = {domain->xprd, 0.0, 0.0, domain->xy, domain->yprd, 0.0, domain->xz, domain->yz, domain->zprd};
  std::memcpy(cell, c9, sizeof(cell));
  sys_ctx = pair->context();
  int rc = tmd_system_create(sys_ctx, natoms, &atom->x[0][0], atom->type, cell, neighbor->skin, &sys);
  if (rc != TMD_OK) {
    sys = nullptr;
    device_fail(rc, "setup");
  }
  if ((rc = tmd_system_set_velocities(sys, &atom->v[0][0])) != TMD_OK) device_fail(rc, "setup");
  dev_type.assign(atom->type, atom->type + natoms);
  dev_tag.assign(atom->tag, atom->tag + natoms);
  dev_image.assign(atom->image, atom->image + natoms);
  maxtag = 0;
  for (tagint t : dev_tag) maxtag = t > maxtag ? t : maxtag;
  pair->resident = this;
}

void FixNVETensorMDAMD::release()
{
  if (pair && pair->resident == this) pair->resident = nullptr;
  if (sys) tmd_system_destroy(sys);
  sys = nullptr;
}

void FixNVETensorMDAMD::initial_integrate(int /*vflag*/)
{
  pair->send_etemp();  // a temperature changed through extract("etemp")
  // enqueued without waiting: LAMMPS' host work of the step overlaps it,
  // and the next call on the system (pair_results, a write-back or the
  // next step) takes up its list check
  const int rc = tmd_md_step_async(sys, dtv);
  if (rc != TMD_OK) device_fail(rc, "step");
}

void FixNVETensorMDAMD::end_of_step()
{
  if (update->ntimestep == output->next) sync_to_host();
}

void FixNVETensorMDAMD::post_run()
{
  if (!sys) return;
  sync_to_host();
  release();
}

// host index of each device atom, by atom ID (LAMMPS may have sorted the
// host arrays at one of its neighbor rebuilds)
std::vector<int> FixNVETensorMDAMD::host_index()
{
  if (atom->nlocal != natoms) error->one(FLERR, "fix nve/tensormd/amd: the number of atoms changed during the run");
  std::vector<int> at(maxtag + 1, -1);
  for (int i = 0; i < natoms; i++) {
    const tagint t = atom->tag[i];
    if (t < 0 || t > maxtag) error->one(FLERR, "fix nve/tensormd/amd: atom IDs changed during the run");
    at[t] = i;
  }
  std::vector<int> h(natoms);
  for (int k = 0; k < natoms; k++) {
    h[k] = at[dev_tag[k]];
    if (h[k] < 0 || atom->type[h[k]] != dev_type[k])
      error->one(FLERR, "fix nve/tensormd/amd: atom IDs or types changed during the run");
  }
  return h;
}

void FixNVETensorMDAMD::sync_to_host()
{
  const std::vector<int> h = host_index();
  stage.resize(9 * (size_t) natoms);
  double *xs = stage.data(), *vs = xs + 3 * (size_t) natoms, *fs = vs + 3 * (size_t) natoms;
  const int rc = tmd_system_get(sys, xs, vs, fs, nullptr, nullptr, nullptr);
  if (rc != TMD_OK) device_fail(rc, "write-back");
  double **x = atom->x, **v = atom->v, **f = atom->f;
  for (int k = 0; k < natoms; k++) {
    const int i = h[k];
    double p[3] = {xs[3 * k], xs[3 * k + 1], xs[3 * k + 2]};
    domain->unmap(p, dev_image[k]);       // unwrapped position
    domain->unmap_inv(p, atom->image[i]);  // in the host's current image
    for (int d = 0; d < 3; d++) {
      x[i][d] = p[d];
      v[i][d] = vs[3 * k + d];
      f[i][d] = fs[3 * k + d];
    }
  }
  comm->forward_comm();  // ghosts of the written-back positions
}

void FixNVETensorMDAMD::pair_results(int want_energy, int want_virial, double *eatom, double &energy,
                                     double *virial)
{
  if (!want_energy && !want_virial && !eatom) return;
  double e = 0.0, vir[6];
  if (eatom) stage.resize((size_t) natoms);
  const int rc = tmd_system_get(sys, nullptr, nullptr, nullptr, eatom ? stage.data() : nullptr, &e, vir);
  if (rc != TMD_OK) device_fail(rc, "results");
  if (want_energy) energy = e;
  if (want_virial)
    for (int q = 0; q < 6; q++) virial[q] = vir[q];
  if (eatom) {
    const std::vector<int> h = host_index();
    for (int k = 0; k < natoms; k++) eatom[h[k]] = stage[k];
  }
}

double FixNVETensorMDAMD::memory_usage()
{
  return (double) stage.capacity() * sizeof(double) +
      (double) natoms * (sizeof(int) + sizeof(tagint) + sizeof(imageint));
}
