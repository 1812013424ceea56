/* ----------------------------------------------------------------------
   pair_style tensormd/amd: LAMMPS adaptor of libtensormd_amd (MI355X).

   Behaviour follows PairTensorMD (src/ML-TENSORMD/pair_tensormd.cpp of the
   reference tree); each method cites what it replaces.  All model and
   force work is done by the C ABI in tensormd_amd.h.
------------------------------------------------------------------------- */

#include "pair_tensormd_amd.h"

#include "atom.h"
#include "comm.h"
#include "error.h"
#include "fix_nve_tensormd_amd.h"
#include "force.h"
#include "memory.h"
#include "neigh_list.h"
#include "neighbor.h"
#include "utils.h"

#include "tensormd_amd.h"

#include <cstring>
#include <string>
#include <vector>

using namespace LAMMPS_NS;

/* flags as in pair_tensormd.cpp:37-61 */
PairTensorMDAMD::PairTensorMDAMD(LAMMPS *lmp) : Pair(lmp)
{
  single_enable = 0;
  restartinfo = 0;
  one_coeff = 1;
  manybody_flag = 1;
  centroidstressflag = CENTROID_NOTAVAIL;
  ctx = nullptr;
  cutmax = 0.0;
  etemp = 0.0;
  etemp_sent = 0.0;
  tdpe_flag = 0;
  eentropy = free_energy = 0.0;
  satom = aatom = nullptr;
  maxsatom = 0;
  scale = nullptr;
  resident = nullptr;
}

PairTensorMDAMD::~PairTensorMDAMD()
{
  if (ctx) tmd_destroy(ctx);
  memory->destroy(satom);
  memory->destroy(aatom);
  if (allocated) {
    memory->destroy(setflag);
    memory->destroy(cutsq);
    memory->destroy(scale);
    delete[] map;
  }
}

/* pair_tensormd.cpp:123-137: per-atom electron entropy and free energy
   arrays beside eatom when the etemp keyword was given */
void PairTensorMDAMD::ev_setup(int eflag, int vflag, int alloc)
{
  Pair::ev_setup(eflag, vflag, alloc);
  if (tdpe_flag && alloc && maxsatom < maxeatom) {
    memory->destroy(satom);
    memory->destroy(aatom);
    maxsatom = maxeatom;
    memory->create(satom, maxsatom, "pair:tensormd:tdnp:satom");
    memory->create(aatom, maxsatom, "pair:tensormd:tdnp:aatom");
  }
}

void PairTensorMDAMD::fail(int code, const char *where)
{
  error->all(FLERR, "{}: {} (status {})", where, tmd_last_error(ctx), code);
}

/* pair_tensormd.cpp:141-184: f is overwritten for all local + ghost atoms,
   eatom/vatom are written, eng_vdwl = sum of scaled atomic energies, and the
   global virial is the fdotr sum.  The list is identified by the step of
   the last neighbor build (neighbor->lastcall): between builds only the
   positions go to the device (the reference repacks and uploads the whole
   list every step, gpu/cuda/kernel.cu:1500-1521). */
void PairTensorMDAMD::send_etemp()
{
  if (etemp != etemp_sent) {  // fix tensormd/etemp/amd or extract("etemp") changed it
    int trc = tmd_set_electron_temperature(ctx, etemp);
    if (trc != TMD_OK) fail(trc, "pair tensormd/amd electron temperature");
    etemp_sent = etemp;
  }
}

void PairTensorMDAMD::compute(int eflag, int vflag)
{
  ev_init(eflag, vflag);
  if (resident) {
    // fix nve/tensormd/amd evaluated this step's forces on the device; f
    // reaches the host at output steps, energy and virial here
    if (vflag_atom || cvflag_atom)
      error->all(FLERR, "pair tensormd/amd: per-atom virials are not available while fix nve/tensormd/amd runs");
    if (scale[1][1] != 1.0)
      error->all(FLERR, "pair tensormd/amd: fix adapt scaling is not available while fix nve/tensormd/amd runs");
    resident->pair_results(eflag_global, vflag_global, eflag_atom ? eatom : nullptr, eng_vdwl, virial);
    return;
  }
  send_etemp();
  const int nlocal = atom->nlocal;
  const int nall = nlocal + atom->nghost;
  double energy = 0.0;
  const double scale_factor = scale[1][1];
  double *eat = eflag_atom ? eatom : nullptr;
  double *vat = vflag_atom ? &vatom[0][0] : nullptr;
  int rc;
  if (neighbor->lastcall >= 0)
    rc = tmd_compute_list_serial(ctx, (long long) neighbor->lastcall, nlocal, nall, &atom->x[0][0],
                                 atom->type, list->inum, list->ilist, list->numneigh,
                                 list->firstneigh, scale_factor, &atom->f[0][0], eat, vat, &energy,
                                 nullptr);
  else
    rc = tmd_compute_list(ctx, nlocal, nall, &atom->x[0][0], atom->type, list->inum, list->ilist,
                          list->numneigh, list->firstneigh, scale_factor, &atom->f[0][0], eat, vat,
                          &energy, nullptr);
  if (rc != TMD_OK) fail(rc, "pair tensormd/amd compute");
  if (eflag) eng_vdwl = energy;
  // TDNP (tensormd.cpp:471-485): the energy is the electron internal energy;
  // entropy and free energy beside it (zero for a model without TDNP)
  if (tdpe_flag) {
    const bool per_atom = eflag_atom && satom && aatom;
    rc = tmd_tdnp_results(ctx, &eentropy, &free_energy, per_atom ? satom : nullptr,
                          per_atom ? aatom : nullptr);
    if (rc == TMD_ERR_MODEL) {
      eentropy = free_energy = 0.0;
    } else if (rc != TMD_OK) {
      fail(rc, "pair tensormd/amd TDNP results");
    }
  }
  if (vflag_fdotr) virial_fdotr_compute();
}

void PairTensorMDAMD::allocate()
{
  allocated = 1;
  const int n = atom->ntypes;
  memory->create(setflag, n + 1, n + 1, "pair:setflag");
  memory->create(cutsq, n + 1, n + 1, "pair:cutsq");
  memory->create(scale, n + 1, n + 1, "pair:scale");
  map = new int[n + 1];
  for (int i = 0; i <= n; i++) map[i] = -1;
}

/* pair_tensormd.cpp:205-208 */
void PairTensorMDAMD::settings(int narg, char ** /*arg*/)
{
  if (narg != 0) error->all(FLERR, "Illegal pair_style command");
}

/* pair_tensormd.cpp:214-317 */
void PairTensorMDAMD::coeff(int narg, char **arg)
{
  if (!allocated) allocate();
  if (narg < 4) error->all(FLERR, "Incorrect args for pair coefficients");
  if (strcmp(arg[0], "*") != 0 || strcmp(arg[1], "*") != 0)
    error->all(FLERR, "Incorrect args for pair coefficients");
  const std::string file = utils::get_potential_file_path(arg[2]);
  if (file.empty()) error->all(FLERR, "Cannot open potential file {}", arg[2]);

  tmd_options opt;
  tmd_options_default(&opt);
  std::vector<std::string> symbols;
  bool read_elt = true;
  for (int i = 3; i < narg;) {
    const std::string key = arg[i];
    auto value = [&]() -> std::string {
      if (i + 1 >= narg) error->all(FLERR, "pair_coeff: missing value after {}", key);
      return arg[i + 1];
    };
    if (key == "interp") {
      opt.interp_dx = utils::numeric(FLERR, value(), false, lmp);
      i += 2;
      read_elt = false;
    } else if (key == "etemp") {
      etemp = utils::numeric(FLERR, value(), false, lmp) / 11604.522060401008;  // K -> eV
      tdpe_flag = 1;
      i += 2;
      read_elt = false;
    } else if (key == "cmax") {
      const std::string v = value();
      if (v == "exact") opt.cmax_exact = 1;
      else if (v == "default") opt.cmax_exact = 0;
      else error->all(FLERR, "Unknown option for cmax");
      i += 2;
      read_elt = false;
    } else if (key == "table") {  // where the radial table is built at load
      const std::string v = value();
      if (v == "host") opt.table_device = 0;
      else if (v == "device") opt.table_device = 1;
      else error->all(FLERR, "Unknown option for table");
      i += 2;
      read_elt = false;
    } else if (key == "precision") {
      const std::string v = value();
      if (v == "fp32") opt.mlp_precision = TMD_MLP_FP32;
      else if (v == "bf16") opt.mlp_precision = TMD_MLP_BF16;
      else if (v == "fp8") opt.mlp_precision = TMD_MLP_FP8;
      else if (v == "bf16x3") opt.mlp_precision = TMD_MLP_BF16X3;
      else error->all(FLERR, "Unknown option for precision");
      i += 2;
      read_elt = false;
    } else if (read_elt) {
      symbols.push_back(key);
      i++;
    } else {
      error->all(FLERR, "Unknown pair_coeff keyword {}", key);
    }
  }
  const int n = atom->ntypes;
  if ((int) symbols.size() != n)
    error->all(FLERR, "pair_coeff: give one element symbol (or NULL) per atom type");
  std::vector<const char *> sym(n);
  for (int t = 0; t < n; t++) sym[t] = symbols[t] == "NULL" ? nullptr : symbols[t].c_str();
  if (ctx) {
    tmd_destroy(ctx);
    ctx = nullptr;
  }
  int rc = tmd_create(file.c_str(), sym.data(), n, &opt, &ctx);
  if (rc != TMD_OK) {
    error->all(FLERR, "pair_coeff: cannot load {}: {}", file, tmd_last_error(nullptr));
  }
  if ((rc = tmd_cutoff(ctx, &cutmax)) != TMD_OK) fail(rc, "pair_coeff");
  if ((rc = tmd_set_electron_temperature(ctx, etemp)) != TMD_OK) fail(rc, "pair_coeff");
  etemp_sent = etemp;
  std::vector<int> tmap(n + 1, -1);
  std::vector<double> mass(n + 1, 0.0);
  if ((rc = tmd_type_map(ctx, tmap.data())) != TMD_OK) fail(rc, "pair_coeff");
  if ((rc = tmd_masses(ctx, mass.data())) != TMD_OK) fail(rc, "pair_coeff");
  int count = 0;
  for (int i = 1; i <= n; i++) {
    map[i] = tmap[i];
    for (int j = i; j <= n; j++) {
      setflag[i][j] = 0;
      if (tmap[i] >= 0 && tmap[j] >= 0) {
        setflag[i][j] = 1;
        count++;
      }
      scale[i][j] = 1.0;
    }
    if (tmap[i] >= 0) atom->set_mass(FLERR, i, mass[i]);
  }
  if (count == 0) error->all(FLERR, "Incorrect args for pair coefficients");
  kwargs.assign(arg + 2, arg + narg);
}

/* pair_tensormd.cpp:323-330 */
void PairTensorMDAMD::init_style()
{
  if (force->newton_pair == 0) error->all(FLERR, "Pair style tensormd/amd requires newton pair on");
  neighbor->add_request(this, NeighConst::REQ_FULL);
}

/* pair_tensormd.cpp:336-341 */
double PairTensorMDAMD::init_one(int i, int j)
{
  if (setflag[i][j] == 0) scale[i][j] = 1.0;
  scale[j][i] = scale[i][j];
  return cutmax;
}

/* pair_tensormd.cpp:415-425 */
void *PairTensorMDAMD::extract(const char *str, int &dim)
{
  if (strcmp(str, "scale") == 0) {
    dim = 2;
    return (void *) scale;
  }
  if (strcmp(str, "etemp") == 0) {
    dim = 0;
    return &etemp;
  }
  return nullptr;
}

/* the reference reports 0 (tensormd.cpp:619-623); device memory is not counted */
double PairTensorMDAMD::memory_usage()
{
  return 0.0;
}
