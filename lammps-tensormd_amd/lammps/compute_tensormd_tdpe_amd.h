/* -*- c++ -*- ----------------------------------------------------------
   compute tensormd/tdpe/amd: the electron internal energy, free energy and
   entropy of a temperature-dependent (TDNP) model evaluated by
   pair_style tensormd/amd; the counterpart of compute tensormd/tdpe
   (src/ML-TENSORMD/compute_tensormd_tdpe.{h,cpp} of the reference tree).
     compute ID group tensormd/tdpe/amd      -> c_ID[1..3] = U, A, S
------------------------------------------------------------------------- */

#ifdef COMPUTE_CLASS
// clang-format off
ComputeStyle(tensormd/tdpe/amd, ComputeTensorMDTDPEAMD);
// clang-format on
#else

#ifndef LMP_COMPUTE_TENSORMD_TDPE_AMD_H
#define LMP_COMPUTE_TENSORMD_TDPE_AMD_H

#include "compute.h"

namespace LAMMPS_NS {

class PairTensorMDAMD;

class ComputeTensorMDTDPEAMD : public Compute {
 public:
  ComputeTensorMDTDPEAMD(class LAMMPS *, int, char **);
  ~ComputeTensorMDTDPEAMD() override;
  void init() override;
  void compute_vector() override;

 private:
  PairTensorMDAMD *pair;
};

}    // namespace LAMMPS_NS

#endif
#endif
