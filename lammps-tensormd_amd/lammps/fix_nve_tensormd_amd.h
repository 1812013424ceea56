/* -*- c++ -*- ----------------------------------------------------------
   fix nve/tensormd/amd: velocity-Verlet NVE with x, v and f kept on the
   MI355X between steps, for pair_style tensormd/amd.  The counterpart of
   fix nve (src/fix_nve.cpp:64-130 of the reference tree) with the pair
   style's device context doing the whole step: both half kicks, the drift,
   the neighbor list (rc + the LAMMPS skin, rebuilt on the device when an
   atom has moved more than skin/2) and the force evaluation.
     fix ID all nve/tensormd/amd
   Host x, v and f are written back at output steps (thermo, dump, restart:
   the steps LAMMPS' Output schedules) and at the end of every run; the
   pair style reports the device energy and virial on the steps that ask
   for them.  One MPI rank, a periodic box, per-type masses equal to the
   model's, and no other fix that acts during a step.
------------------------------------------------------------------------- */

#ifdef FIX_CLASS
// clang-format off
FixStyle(nve/tensormd/amd, FixNVETensorMDAMD);
// clang-format on
#else

#ifndef LMP_FIX_NVE_TENSORMD_AMD_H
#define LMP_FIX_NVE_TENSORMD_AMD_H

#include "fix.h"

#include <vector>

struct tmd_ctx;
struct tmd_system;

namespace LAMMPS_NS {

class PairTensorMDAMD;

class FixNVETensorMDAMD : public Fix {
 public:
  FixNVETensorMDAMD(class LAMMPS *, int, char **);
  ~FixNVETensorMDAMD() override;
  int setmask() override;
  void init() override;
  void setup(int) override;
  void initial_integrate(int) override;
  void end_of_step() override;
  void post_run() override;
  void reset_dt() override;
  double memory_usage() override;

  // PairTensorMDAMD::compute while the run is on the device: energy, global
  // virial (xx yy zz xy xz yz) and, with eatom non-NULL, per-atom energies
  // of the positions of the current step
  void pair_results(int want_energy, int want_virial, double *eatom, double &energy, double *virial);

 private:
  PairTensorMDAMD *pair;
  tmd_system *sys;
  tmd_ctx *sys_ctx;                 // the pair style's context that made sys
  double dtv;
  int natoms;                       // atoms on the device
  std::vector<int> dev_type;        // LAMMPS type of each device atom
  std::vector<tagint> dev_tag;      // atom ID of each device atom
  std::vector<imageint> dev_image;  // image flags at the start of the run
  tagint maxtag;
  std::vector<double> stage;        // x, v, f or eatom in device order
  double cell[9];                   // cell rows a, b, c of the device system

  void release();
  std::vector<int> host_index();
  void sync_to_host();
  void device_fail(int code, const char *where);
};

}    // namespace LAMMPS_NS

#endif
#endif
