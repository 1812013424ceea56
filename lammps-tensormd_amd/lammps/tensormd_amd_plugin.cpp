/* Registers pair_style tensormd/amd with a LAMMPS binary built with
   PKG_PLUGIN (src/lammpsplugin.h:17-39, src/PLUGIN/plugin.cpp:98-122 of the
   reference tree): `plugin load tensormd_amd_plugin.so`. */

#include "lammpsplugin.h"
#include "version.h"

#include "pair_tensormd_amd.h"

using namespace LAMMPS_NS;

static Pair *tensormd_amd_creator(LAMMPS *lmp)
{
  return new PairTensorMDAMD(lmp);
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
}
