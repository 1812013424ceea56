/* -*- c++ -*- ----------------------------------------------------------
   fix tensormd/etemp/amd: every N steps, set the electron temperature of
   pair_style tensormd/amd to the thermo temperature (two-temperature
   equilibrium); the counterpart of fix tensormd/etemp
   (src/ML-TENSORMD/fix_tensormd_etemp.{h,cpp} of the reference tree).
     fix ID group tensormd/etemp/amd N       (N = 0: never update)
------------------------------------------------------------------------- */

#ifdef FIX_CLASS
// clang-format off
FixStyle(tensormd/etemp/amd, FixTensorMDElectronTemperatureAMD);
// clang-format on
#else

#ifndef LMP_FIX_TENSORMD_ETEMP_AMD_H
#define LMP_FIX_TENSORMD_ETEMP_AMD_H

#include "fix.h"

namespace LAMMPS_NS {

class PairTensorMDAMD;

class FixTensorMDElectronTemperatureAMD : public Fix {
 public:
  FixTensorMDElectronTemperatureAMD(class LAMMPS *, int, char **);
  int setmask() override;
  void init() override;
  void end_of_step() override;
  double compute_scalar() override;

 private:
  PairTensorMDAMD *pair;
  double kelvin;
};

}    // namespace LAMMPS_NS

#endif
#endif
