/* compute tensormd/density/amd: see compute_tensormd_density_amd.h. */

#include "compute_tensormd_density_amd.h"

#include "atom.h"
#include "error.h"
#include "force.h"
#include "memory.h"
#include "neigh_list.h"
#include "neighbor.h"
#include "update.h"
#include "utils.h"

#include "pair_tensormd_amd.h"
#include "tensormd_amd.h"

using namespace LAMMPS_NS;

ComputeTensorMDDensityAMD::ComputeTensorMDDensityAMD(LAMMPS *lmp, int narg, char **arg) :
    Compute(lmp, narg, arg), list(nullptr), pair(nullptr), ctx(nullptr), nmax(0), G(nullptr)
{
  if (narg != 3) error->all(FLERR, "Illegal compute tensormd/density/amd command");
  pair = current_pair();
  load_model();
  int info[5];
  if (tmd_model_info(ctx, info) != TMD_OK) error->all(FLERR, tmd_last_error(ctx));
  size_peratom_cols = info[3];
  peratom_flag = 1;
  timeflag = 1;
}

ComputeTensorMDDensityAMD::~ComputeTensorMDDensityAMD()
{
  if (ctx) tmd_destroy(ctx);
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

// (Re)load the pair style's model file with the reference compute's table
// step, dx = 1e-4 (compute_tensormd_density.cpp:106-124), and the pair's
// element map (pair_coeff arguments: file, one symbol per type, keywords).
void ComputeTensorMDDensityAMD::load_model()
{
  const std::vector<std::string> &args = pair->coeff_args();
  if (ctx && args == loaded) return;
  const int n = atom->ntypes;
  if ((int) args.size() < 1 + n) error->all(FLERR, "compute tensormd/density/amd: pair_coeff arguments missing");
  const std::string file = utils::get_potential_file_path(args[0]);
  std::vector<const char *> sym(n);
  for (int t = 0; t < n; t++) sym[t] = args[1 + t] == "NULL" ? nullptr : args[1 + t].c_str();
  tmd_options opt;
  tmd_options_default(&opt);
  opt.interp_dx = 1e-4;
  if (ctx) {
    tmd_destroy(ctx);
    ctx = nullptr;
  }
  if (tmd_create(file.c_str(), sym.data(), n, &opt, &ctx) != TMD_OK)
    error->all(FLERR, "compute tensormd/density/amd: cannot load {}: {}", file, tmd_last_error(nullptr));
  loaded = args;
}

// A pair_coeff issued after this compute was created may load another
// model: reload it, and refuse a descriptor width other than the one the
// per-atom array was declared with (fixes and dumps hold that width).
void ComputeTensorMDDensityAMD::init()
{
  pair = current_pair();
  load_model();
  int info[5];
  if (tmd_model_info(ctx, info) != TMD_OK) error->all(FLERR, tmd_last_error(ctx));
  if (info[3] != size_peratom_cols)
    error->all(FLERR, "compute tensormd/density/amd: the pair style's model has descriptor width {}, "
               "this compute was created for {}; create the compute again", info[3], size_peratom_cols);
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
  const int st = tmd_descriptors_list(ctx, nlocal, nall, &atom->x[0][0], atom->type,
                                      list->inum, list->ilist, list->numneigh, list->firstneigh,
                                      &G[0][0]);
  if (st != TMD_OK) error->one(FLERR, tmd_last_error(ctx));
}

double ComputeTensorMDDensityAMD::memory_usage()
{
  return (double) nmax * size_peratom_cols * sizeof(double);
}
