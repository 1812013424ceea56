/* -*- c++ -*- ----------------------------------------------------------
   pair_style tensormd/amd: LAMMPS adaptor of libtensormd_amd (MI355X).

   Same user interface and results as pair_style tensormd
   (src/ML-TENSORMD/pair_tensormd.h:34-83 of the reference tree):
     pair_style tensormd/amd
     pair_coeff * * model.npz El1 El2 ... [interp dx] [etemp T] [cmax exact|default]
   With `suffix amd` (or -sf amd) an input that says `pair_style tensormd`
   selects this class.  Extra keywords: `precision fp32|bf16|bf16x3|fp8`
   (MLP GEMMs; default fp32) and `table host|device` (where the radial
   table is built at load; default host).
------------------------------------------------------------------------- */

#ifdef PAIR_CLASS
// clang-format off
PairStyle(tensormd/amd, PairTensorMDAMD);
// clang-format on
#else

#ifndef LMP_PAIR_TENSORMD_AMD_H
#define LMP_PAIR_TENSORMD_AMD_H

#include "pair.h"

#include <string>
#include <vector>

struct tmd_ctx;

namespace LAMMPS_NS {

class FixNVETensorMDAMD;

class PairTensorMDAMD : public Pair {
 public:
  explicit PairTensorMDAMD(class LAMMPS *);
  ~PairTensorMDAMD() override;
  void compute(int, int) override;
  void settings(int, char **) override;
  void coeff(int, char **) override;
  void init_style() override;
  double init_one(int, int) override;
  void *extract(const char *, int &) override;
  double memory_usage() override;

  void ev_setup(int, int, int alloc = 1) override;

  // TDNP state, as PairTensorMD (pair_tensormd.h): electron temperature in
  // eV (pair_coeff "etemp T", fix tensormd/etemp/amd), tdpe_flag set by the
  // etemp keyword, and the electron entropy and free energy of the last
  // compute (sums and, with per-atom energies, per atom)
  double etemp;
  int tdpe_flag;
  double eentropy, free_energy;
  double *satom, *aatom;
  tmd_ctx *context() const { return ctx; }
  // the pair_coeff arguments after "* *" (model file, symbols, keywords),
  // for compute tensormd/density/amd, which loads the model itself
  const std::vector<std::string> &coeff_args() const { return kwargs; }
  // pass a changed etemp to the device
  void send_etemp();
  // set by fix nve/tensormd/amd for the length of a run: positions and
  // forces are on the device, compute() reports the device results
  FixNVETensorMDAMD *resident;

 protected:
  tmd_ctx *ctx;
  double cutmax;
  double **scale;  // fix adapt scaling (one factor for all types)
  std::vector<std::string> kwargs;
  double etemp_sent;  // the temperature the context holds
  int maxsatom;

  void allocate();
  void fail(int code, const char *where);
};

}    // namespace LAMMPS_NS

#endif
#endif
