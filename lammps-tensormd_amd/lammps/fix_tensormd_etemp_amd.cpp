/* ----------------------------------------------------------------------
   fix tensormd/etemp/amd (fix_tensormd_etemp.cpp:31-94 of the reference
   tree): at the end of every N-th step the pair style's electron temperature
   becomes the kinetic temperature of compute thermo_temp, in eV (the same
   11604.51812 K/eV as the reference fix); compute_scalar reports it in K.
   The pair style passes a changed temperature to the device before its
   next force evaluation.
------------------------------------------------------------------------- */

#include "fix_tensormd_etemp_amd.h"

#include "compute.h"
#include "error.h"
#include "force.h"
#include "modify.h"
#include "pair_tensormd_amd.h"
#include "update.h"
#include "utils.h"

using namespace LAMMPS_NS;
using namespace FixConst;

static constexpr double kEvToKelvin = 11604.51812;

FixTensorMDElectronTemperatureAMD::FixTensorMDElectronTemperatureAMD(LAMMPS *lmp, int narg, char **arg) :
    Fix(lmp, narg, arg), pair(nullptr), kelvin(0.0)
{
  if (narg != 4) error->all(FLERR, "Illegal fix tensormd/etemp/amd command");
  nevery = utils::inumeric(FLERR, arg[3], false, lmp);
  if (nevery < 0) error->all(FLERR, "Illegal fix tensormd/etemp/amd command");
  scalar_flag = 1;
}

int FixTensorMDElectronTemperatureAMD::setmask()
{
  return END_OF_STEP;
}

void FixTensorMDElectronTemperatureAMD::init()
{
  pair = dynamic_cast<PairTensorMDAMD *>(force->pair);
  if (!pair) error->all(FLERR, "pair_style must be tensormd/amd for fix tensormd/etemp/amd");
  kelvin = pair->etemp * kEvToKelvin;
}

void FixTensorMDElectronTemperatureAMD::end_of_step()
{
  if (nevery == 0 || update->ntimestep % nevery) return;
  modify->clearstep_compute();
  Compute *c_temp = modify->get_compute_by_id("thermo_temp");
  if (!c_temp) error->all(FLERR, "fix tensormd/etemp/amd: compute thermo_temp does not exist");
  kelvin = c_temp->compute_scalar();
  pair->etemp = kelvin / kEvToKelvin;
  modify->addstep_compute(update->ntimestep + nevery);
}

double FixTensorMDElectronTemperatureAMD::compute_scalar()
{
  return kelvin;
}
