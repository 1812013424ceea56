/* -*- c++ -*- ----------------------------------------------------------
   compute tensormd/density/amd: per-atom descriptor rows of the model
   loaded by pair_style tensormd/amd, from libtensormd_amd (MI355X).

   Same output as compute tensormd/density
   (src/ML-TENSORMD/compute_tensormd_density.{h,cpp} of the reference tree):
     compute ID group tensormd/density/amd
   gives a per-atom array with the model file's descriptor width, in the
   file's [element][filter][moment] column order.  Like the reference
   compute it loads the pair style's model file itself, with the radial
   table at interp dx = 1e-4 whatever the pair style uses
   (compute_tensormd_density.cpp:106-124); the descriptors are computed in
   FP32 on the device.
------------------------------------------------------------------------- */

#ifdef COMPUTE_CLASS
// clang-format off
ComputeStyle(tensormd/density/amd, ComputeTensorMDDensityAMD);
// clang-format on
#else

#ifndef LMP_COMPUTE_TENSORMD_DENSITY_AMD_H
#define LMP_COMPUTE_TENSORMD_DENSITY_AMD_H

#include "compute.h"

#include <string>
#include <vector>

struct tmd_ctx;

namespace LAMMPS_NS {

class PairTensorMDAMD;

class ComputeTensorMDDensityAMD : public Compute {
 public:
  ComputeTensorMDDensityAMD(class LAMMPS *, int, char **);
  ~ComputeTensorMDDensityAMD() override;
  void init() override;
  void init_list(int, class NeighList *) override;
  void compute_peratom() override;
  double memory_usage() override;

 private:
  class NeighList *list;
  PairTensorMDAMD *pair;
  ::tmd_ctx *ctx;                    // this compute's own context (dx = 1e-4)
  std::vector<std::string> loaded;   // the pair_coeff arguments it was built from
  int nmax;
  double **G;

  PairTensorMDAMD *current_pair();
  void load_model();
};

}    // namespace LAMMPS_NS

#endif
#endif
