/* ----------------------------------------------------------------------
   compute tensormd/tdpe/amd (compute_tensormd_tdpe.cpp:28-76 of the
   reference tree): a global vector of 3, the pair style's electron internal
   energy (eng_vdwl), free energy and entropy of the last force evaluation,
   summed over MPI ranks.
------------------------------------------------------------------------- */

#include "compute_tensormd_tdpe_amd.h"

#include "error.h"
#include "force.h"
#include "pair_tensormd_amd.h"
#include "update.h"

using namespace LAMMPS_NS;

ComputeTensorMDTDPEAMD::ComputeTensorMDTDPEAMD(LAMMPS *lmp, int narg, char **arg) :
    Compute(lmp, narg, arg), pair(nullptr)
{
  if (narg < 3) error->all(FLERR, "Illegal compute tensormd/tdpe/amd command");
  pair = dynamic_cast<PairTensorMDAMD *>(force->pair);
  if (!pair)
    error->all(FLERR, "pair_style tensormd/amd must be defined before compute tensormd/tdpe/amd");
  if (pair->tdpe_flag == 0)
    error->all(FLERR, "compute tensormd/tdpe/amd requires a temperature-dependent model (etemp keyword)");
  vector_flag = 1;
  size_vector = 3;
  extvector = 0;
  peatomflag = 1;
  timeflag = 1;
  vector = new double[size_vector];
}

ComputeTensorMDTDPEAMD::~ComputeTensorMDTDPEAMD()
{
  delete[] vector;
}

void ComputeTensorMDTDPEAMD::init()
{
  pair = dynamic_cast<PairTensorMDAMD *>(force->pair);
  if (!pair) error->all(FLERR, "compute tensormd/tdpe/amd requires pair_style tensormd/amd");
}

void ComputeTensorMDTDPEAMD::compute_vector()
{
  invoked_vector = update->ntimestep;
  double values[3] = {0.0, 0.0, 0.0};
  if (pair && pair->compute_flag && pair->tdpe_flag) {
    values[0] = pair->eng_vdwl;
    values[1] = pair->free_energy;
    values[2] = pair->eentropy;
  }
  MPI_Allreduce(values, vector, size_vector, MPI_DOUBLE, MPI_SUM, world);
}
