/* Registers pair_style tensormd/amd, compute tensormd/density/amd,
   compute tensormd/tdpe/amd, fix tensormd/etemp/amd and
   fix nve/tensormd/amd with a
   LAMMPS binary built with PKG_PLUGIN (src/lammpsplugin.h:17-39,
   src/PLUGIN/plugin.cpp:98-122 of the reference tree):
   `plugin load tensormd_amd_plugin.so`. */

#include "lammpsplugin.h"
#include "version.h"

#include "compute_tensormd_density_amd.h"
#include "compute_tensormd_tdpe_amd.h"
#include "fix_nve_tensormd_amd.h"
#include "fix_tensormd_etemp_amd.h"
#include "pair_tensormd_amd.h"

using namespace LAMMPS_NS;

static Pair *tensormd_amd_creator(LAMMPS *lmp)
{
  return new PairTensorMDAMD(lmp);
}

static Compute *tensormd_density_amd_creator(LAMMPS *lmp, int narg, char **arg)
{
  return new ComputeTensorMDDensityAMD(lmp, narg, arg);
}

static Compute *tensormd_tdpe_amd_creator(LAMMPS *lmp, int narg, char **arg)
{
  return new ComputeTensorMDTDPEAMD(lmp, narg, arg);
}

static Fix *tensormd_etemp_amd_creator(LAMMPS *lmp, int narg, char **arg)
{
  return new FixTensorMDElectronTemperatureAMD(lmp, narg, arg);
}

static Fix *nve_tensormd_amd_creator(LAMMPS *lmp, int narg, char **arg)
{
  return new FixNVETensorMDAMD(lmp, narg, arg);
}

extern "C" void lammpsplugin_init(void *lmp, void *handle, void *regfunc)
{
  lammpsplugin_t plugin;
  auto reg = (lammpsplugin_regfunc) regfunc;
  plugin.version = LAMMPS_VERSION;
  plugin.style = "pair";
  plugin.name = "tensormd/amd";
  plugin.info = "TensorMD pair style on MI355X (libtensormd_amd)";
  plugin.author = "tensormd-mi355x";
  plugin.creator.v1 = (lammpsplugin_factory1 *) &tensormd_amd_creator;
  plugin.handle = handle;
  (*reg)(&plugin, lmp);

  plugin.style = "compute";
  plugin.name = "tensormd/density/amd";
  plugin.info = "TensorMD per-atom descriptors on MI355X (libtensormd_amd)";
  plugin.creator.v2 = (lammpsplugin_factory2 *) &tensormd_density_amd_creator;
  (*reg)(&plugin, lmp);

  plugin.name = "tensormd/tdpe/amd";
  plugin.info = "TensorMD electron energy, free energy and entropy (TDNP models) on MI355X";
  plugin.creator.v2 = (lammpsplugin_factory2 *) &tensormd_tdpe_amd_creator;
  (*reg)(&plugin, lmp);

  plugin.style = "fix";
  plugin.name = "tensormd/etemp/amd";
  plugin.info = "electron temperature of pair_style tensormd/amd from the thermo temperature";
  plugin.creator.v2 = (lammpsplugin_factory2 *) &tensormd_etemp_amd_creator;
  (*reg)(&plugin, lmp);

  plugin.name = "nve/tensormd/amd";
  plugin.info = "velocity-Verlet NVE with x, v and f kept on the MI355X (pair_style tensormd/amd)";
  plugin.creator.v2 = (lammpsplugin_factory2 *) &nve_tensormd_amd_creator;
  (*reg)(&plugin, lmp);
}
