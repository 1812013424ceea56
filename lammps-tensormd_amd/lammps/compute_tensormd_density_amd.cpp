/* compute tensormd/density/amd: see compute_tensormd_density_amd.h. */

#include "compute_tensormd_density_amd.h"

#include "atom.h"
#include "error.h"
#include "force.h"
#include "memory.h"
#include "neigh_list.h"
#include "neighbor.h"
#include "update.h"

#include "pair_tensormd_amd.h"
#include "tensormd_amd.h"

using namespace LAMMPS_NS;

ComputeTensorMDDensityAMD::ComputeTensorMDDensityAMD(LAMMPS *lmp, int narg, char **arg) :
    Compute(lmp, narg, arg), list(nullptr), pair(nullptr), nmax(0), G(nullptr)
{
  if (narg != 3) error->all(FLERR, "Illegal compute tensormd/density/amd command");
  pair = current_pair();
  int info[5];
  if (tmd_model_info(pair->context(), info) != TMD_OK)
    error->all(FLERR, tmd_last_error(pair->context()));
  size_peratom_cols = info[3];
  peratom_flag = 1;
  timeflag = 1;
}

ComputeTensorMDDensityAMD::~ComputeTensorMDDensityAMD()
{
  memory->destroy(G);
}

// the pair style whose model this compute evaluates; it must be
// tensormd/amd with its pair_coeff already read
PairTensorMDAMD *ComputeTensorMDDensityAMD::current_pair()
{
  auto *p = dynamic_cast<PairTensorMDAMD *>(force->pair);
  if (!p || !p->context())
    error->all(FLERR, "compute tensormd/density/amd needs pair_style tensormd/amd and its pair_coeff first");
  return p;
}

void ComputeTensorMDDensityAMD::init()
{
  pair = current_pair();
  neighbor->add_request(this, NeighConst::REQ_FULL | NeighConst::REQ_OCCASIONAL);
}

void ComputeTensorMDDensityAMD::init_list(int /*id*/, NeighList *ptr)
{
  list = ptr;
}

void ComputeTensorMDDensityAMD::compute_peratom()
{
  invoked_peratom = update->ntimestep;
  if (atom->nmax > nmax) {
    memory->destroy(G);
    nmax = atom->nmax;
    memory->create(G, nmax, size_peratom_cols, "tensormd/density/amd:G");
    array_atom = G;
  }
  const int nlocal = atom->nlocal;
  if (nlocal == 0) return;
  neighbor->build_one(list);
  const int nall = nlocal + atom->nghost;
  const int st = tmd_descriptors_list(pair->context(), nlocal, nall, &atom->x[0][0], atom->type,
                                      list->inum, list->ilist, list->numneigh, list->firstneigh,
                                      &G[0][0]);
  if (st != TMD_OK) error->one(FLERR, tmd_last_error(pair->context()));
}

double ComputeTensorMDDensityAMD::memory_usage()
{
  return (double) nmax * size_peratom_cols * sizeof(double);
}
