/* -*- c++ -*- ----------------------------------------------------------
   compute tensormd/density/amd: per-atom descriptor rows of the model
   loaded by pair_style tensormd/amd, from libtensormd_amd (MI355X).

   Same output as compute tensormd/density
   (src/ML-TENSORMD/compute_tensormd_density.{h,cpp} of the reference tree):
     compute ID group tensormd/density/amd
   gives a per-atom array with the model file's descriptor width, in the
   file's [element][filter][moment] column order.  It uses the context of
   the current pair style instead of loading the model a second time.
------------------------------------------------------------------------- */

#ifdef COMPUTE_CLASS
// clang-format off
ComputeStyle(tensormd/density/amd, ComputeTensorMDDensityAMD);
// clang-format on
#else

#ifndef LMP_COMPUTE_TENSORMD_DENSITY_AMD_H
#define LMP_COMPUTE_TENSORMD_DENSITY_AMD_H

#include "compute.h"

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
  int nmax;
  double **G;

  PairTensorMDAMD *current_pair();
};

}    // namespace LAMMPS_NS

#endif
#endif
