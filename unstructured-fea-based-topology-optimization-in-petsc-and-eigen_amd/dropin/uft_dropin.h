// Drop-in C++ host classes with the reference's names and call signatures
// (OptPara.h, InputMeshData.h, SolveFem.h, PDEFilter.h, MMA.h, Write2Vtu.h),
// implemented over the C-ABI of include/uft.h.  The reference's main.cpp
// compiles against this header with only its #include lines replaced by
// #include "uft_dropin.h" (PetscInitialize, PetscPrintf, MPI_Wtime, ... are
// provided below for one process on one GPU).
//
// PETSc's Vec becomes a handle to a device-resident vector of the context;
// the mesh matrices that the reference passes around (elementEigen,
// eleXCoord, ...) become MeshRef handles, because the mesh lives on the
// device once InputMeshData has loaded it.
#pragma once
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <memory>
#include <vector>
#include <time.h>

#include "uft.h"

typedef int PetscErrorCode;
typedef double PetscScalar;
typedef int PetscInt;
typedef bool PetscBool;
#define PETSC_TRUE true
#define PETSC_FALSE false
#define CHKERRQ(e)                                                                                   \
  do {                                                                                               \
    if (e) {                                                                                         \
      std::fprintf(stderr, "uft error %d at %s:%d: %s\n", (int)(e), __FILE__, __LINE__, UftLastError()); \
      return (e);                                                                                    \
    }                                                                                                \
  } while (0)

// ------------------------------------------------------------------ runtime
// Options database of "-key value" pairs (the PETSc options the reference reads:
// OptPara.cpp:92-104, SolveFem.cpp:11, -ksp_rtol for both solvers) plus the
// mesh source: -mesh_dir DIR (reference text files, with -nn -ne -nfixed -nload)
// or -box NX,NY,NZ [-jitter J -seed S -load_kind K] for the synthetic cantilever;
// -npe 8|4 selects hex8 or tet4.
struct UftRuntime {
  std::map<std::string, std::string> opt;
  uft_ctx *ctx = nullptr;
  int device = 0;
  static UftRuntime &get() {
    static UftRuntime r;
    return r;
  }
  bool has(const char *k) const { return opt.count(k) > 0; }
  double real(const char *k, double d) const { return has(k) ? std::atof(opt.at(k).c_str()) : d; }
  int integer(const char *k, int d) const { return has(k) ? std::atoi(opt.at(k).c_str()) : d; }
  std::string str(const char *k, const char *d) const { return has(k) ? opt.at(k) : std::string(d); }
};

inline const char *UftLastError() { return uft_last_error(UftRuntime::get().ctx); }

inline PetscErrorCode UftInitialize(int *argc, char ***argv) {
  UftRuntime &R = UftRuntime::get();
  for (int i = 1; i < *argc; ++i) {
    if ((*argv)[i][0] != '-') continue;
    std::string k = (*argv)[i];
    std::string v = (i + 1 < *argc && ((*argv)[i + 1][0] != '-' || std::isdigit((unsigned char)(*argv)[i + 1][1]) ||
                                       (*argv)[i + 1][1] == '.'))
                        ? (*argv)[++i]
                        : "1";
    R.opt[k] = v;
  }
  R.device = R.integer("-device", 0);
  return uft_create(&R.ctx, R.device);
}
inline PetscErrorCode UftFinalize() {
  UftRuntime &R = UftRuntime::get();
  uft_destroy(R.ctx);
  R.ctx = nullptr;
  return 0;
}

// ------------------------------------------------------------------ PETSc / MPI names of the driver
// The calls the reference's main.cpp makes outside the five classes
// (main.cpp:38-44,112,132,135), on one process and one GPU: the driver then
// compiles with only its #include lines replaced by this header.
typedef int PetscMPIInt;
typedef int MPI_Comm;
#define MPI_COMM_WORLD 0
#define PETSC_COMM_WORLD 0
#define PETSC_NULL nullptr
namespace Eigen {}  // the driver's "using namespace Eigen;" (main.cpp:12)
inline PetscErrorCode PetscInitialize(int *argc, char ***argv, const char * /*file*/, const char * /*help*/) {
  return UftInitialize(argc, argv);
}
inline PetscErrorCode PetscFinalize() { return UftFinalize(); }
inline int MPI_Comm_rank(MPI_Comm, int *rank) {
  *rank = 0;
  return 0;
}
inline int MPI_Comm_size(MPI_Comm, int *size) {
  *size = 1;
  return 0;
}
inline double MPI_Wtime() {
  timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}
inline PetscErrorCode PetscPrintf(MPI_Comm, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vprintf(fmt, ap);
  va_end(ap);
  return 0;
}

// handle to a device vector of the context (replaces PETSc Vec)
// A PETSc Vec stand-in: a context vector id, or (for the MMA restart
// vectors, which the reference keeps as user-owned Vecs) a host buffer made
// by VecCreateHost.
struct Vec {
  int id = -1;
  std::shared_ptr<std::vector<double>> host;
};
inline Vec VecCreateHost(long long n) {
  Vec v;
  v.host = std::make_shared<std::vector<double>>(n, 0.0);
  return v;
}
struct MeshRef {};  // stands in for the reference's Eigen mesh matrices and raw BC arrays

// ------------------------------------------------------------------ OptPara
class OptPara {
 public:
  OptPara() : OptPara(1) {}
  explicit OptPara(PetscInt nconstraints) {
    m = nconstraints;
    SetUp();
  }
  Vec X{UFT_X}, Xold{UFT_XOLD}, XFiled{UFT_XFILED}, XPhys{UFT_XPHYS}, dfdx{UFT_DFDX};
  Vec *dgdx = nullptr;
  Vec xmin{UFT_XMIN}, xmax{UFT_XMAX}, Vol{UFT_VOL};
  PetscScalar fx = 0, Xmin, Xmax, Emin, Emax, volfrac, penal, movlim, rmin;
  PetscScalar *gx = nullptr;
  PetscInt maxItr, m;
  PetscScalar x0, beta0, betamax, fscale = 1.0;
  PetscBool proj;
  uft_params params;
  ~OptPara() {
    delete[] gx;
    delete[] dgdx;
  }
  PetscErrorCode SetMMA(PetscInt *itr, class MMA **mma);

 private:
  // defaults and options (OptPara::SetUp / SetUpOpt, OptPara.cpp:61-120)
  void SetUp() {
    UftRuntime &R = UftRuntime::get();
    uft_default_params(&params);
    maxItr = R.integer("-maxItr", params.maxItr);
    rmin = R.real("-rmin", params.rmin);
    penal = R.real("-penal", params.penal);
    Emin = R.real("-Emin", params.Emin);
    Emax = R.real("-Emax", params.Emax);
    movlim = R.real("-movlim", params.movlim);
    Xmin = R.real("-Xmin", params.Xmin);
    Xmax = R.real("-Xmax", params.Xmax);
    volfrac = R.real("-volfrac", params.volfrac);
    x0 = R.real("-x0", params.x0);
    beta0 = R.real("-beta0", params.beta0);
    betamax = R.real("-betamax", params.betamax);
    proj = R.integer("-proj", 0) != 0;
    params.Emin = Emin; params.Emax = Emax; params.penal = penal; params.volfrac = volfrac; params.rmin = rmin;
    params.movlim = movlim; params.Xmin = Xmin; params.Xmax = Xmax; params.x0 = x0; params.beta0 = beta0;
    params.betamax = betamax; params.proj = proj; params.m = m; params.maxItr = maxItr;
    params.nu = R.real("-nu", params.nu);
    params.rtol_el = R.real("-ksp_rtol", params.rtol_el);  // unprefixed: both KSPs (SolveFem.cpp:135, PDEFilter.cpp:135)
    params.rtol_f = R.real("-ksp_rtol", params.rtol_f);
    // -mixed 1: FP32 preconditioner blocks; -mixed 2: FP16 smoother blocks
    {
      const int mx = R.integer("-mixed", 0);
      params.precision = mx == 2 ? UFT_MIXED_H : (mx ? UFT_MIXED : UFT_FP64);
    }
    params.cheb_degree = R.integer("-cheb_degree", params.cheb_degree);
    params.cycle_type = R.integer("-cycle_type", params.cycle_type);
    params.esteig_steps = R.integer("-esteig_steps", params.esteig_steps);
    std::printf("################### Optimization settings ####################\n");
    std::printf("# -rmin: %f\n# -volfrac: %f\n# -penal: %f\n# -Emin/-Emax: %e - %e \n", rmin, volfrac, penal, Emin, Emax);
    std::printf("# -Xmin/-Xmax: %e - %e \n# -maxItr: %i\n# -movlim: %f\n# -x0: %f\n", Xmin, Xmax, maxItr, movlim, x0);
    std::printf("# -beta0: %f\n# -betamax: %f\n# -proj: %i  (0/1)\n", beta0, betamax, (int)proj);
    std::printf("##############################################################\n");
    gx = new PetscScalar[m]();
    dgdx = new Vec[m];
    for (int j = 0; j < m; ++j) dgdx[j].id = UFT_DGDX0 + j;
    uft_params_set(R.ctx, &params);
  }
};

// ------------------------------------------------------------------ InputMeshData
class InputMeshData {
 public:
  // InputMeshData::GetNodeAndElementData + GetBcs (InputMeshData.cpp:33-285)
  explicit InputMeshData(Vec /*X*/) {
    UftRuntime &R = UftRuntime::get();
    const int npe = R.integer("-npe", 8);
    int rc = 0;
    const std::string cache = R.str("-mesh_cache", "");
    bool from_cache = false;
    if (!cache.empty()) {
      std::FILE *f = std::fopen(cache.c_str(), "rb");
      if (f) {
        std::fclose(f);
        rc = uft_cache_load(R.ctx, cache.c_str());
        from_cache = true;
      }
    }
    if (from_cache) {
      // mesh and BCs came from the binary cache
    } else if (R.has("-mesh_dir")) {
      const std::string d = R.str("-mesh_dir", ".");
      const long long nn = R.integer("-nn", 410451), ne = R.integer("-ne", 392000);
      rc = uft_mesh_load_text(R.ctx, npe, nn, ne, (d + "/vertex.txt").c_str(), (d + "/element.txt").c_str());
      if (!rc) rc = uft_bc_load_text(R.ctx, d.c_str(), R.integer("-nfixed", 8733), R.integer("-nlc", 1),
                                     R.integer("-nload", 41));
    } else {
      int nx = 40, ny = 20, nz = 20;
      std::sscanf(R.str("-box", "40,20,20").c_str(), "%d,%d,%d", &nx, &ny, &nz);
      const long long nn = uft_box_nodes(nx, ny, nz), ne = uft_box_elems(npe, nx, ny, nz);
      std::vector<double> xyz(nn * 3);
      std::vector<int32_t> el(ne * npe);
      rc = uft_gen_box(npe, nx, ny, nz, 2.0, 1.0, 1.0, R.real("-jitter", 0.0), (uint64_t)R.real("-seed", 0),
                       xyz.data(), el.data());
      std::vector<int32_t> fixed(3 * nn), ld(nn + 9);
      std::vector<double> lv(nn + 9);
      int nf = 0, nl = 0;
      if (!rc) rc = uft_box_bcs(nx, ny, nz, R.integer("-load_kind", 0), fixed.data(), &nf, ld.data(), lv.data(), &nl);
      if (!rc) rc = uft_mesh_set(R.ctx, npe, nn, ne, el.data(), xyz.data());
      if (!rc) rc = uft_bc_set(R.ctx, nf, fixed.data(), 1, nl, ld.data(), lv.data());
    }
    if (!rc && !cache.empty() && !from_cache) rc = uft_cache_save(R.ctx, cache.c_str());
    if (rc) {
      std::fprintf(stderr, "InputMeshData: %s\n", UftLastError());
      std::exit(1);
    }
  }
  MeshRef elementEigen, eleXCoord, eleYCoord, eleZCoord, vertex, element;
  int **loadDofs = nullptr;
  double **load = nullptr;
  int *fixedDofs = nullptr;
  double *vertex_loc = nullptr;
};

// ------------------------------------------------------------------ SolveFem
class SolveFem {
 public:
  // SolveFem ctor (SolveFem.cpp:6-15): DOF map, volumes; here the whole
  // device setup (element operators, filter, hierarchy) happens once.
  SolveFem(MeshRef, MeshRef, MeshRef, MeshRef, Vec /*Vol*/, Vec /*X*/) {
    UftRuntime &R = UftRuntime::get();
    if (uft_setup(R.ctx)) {
      std::fprintf(stderr, "SolveFem: %s\n", UftLastError());
      std::exit(1);
    }
  }
  Vec U{UFT_U};
  Vec VMS{UFT_VMS};
  PetscScalar nu = 0.3;
  // SolveFem::ComputeStress (stress/ComputeStress.cpp:1-113): von Mises
  // stress of U into VMS
  PetscErrorCode ComputeStress(Vec XPhys, PetscScalar Emax, MeshRef, MeshRef, MeshRef, MeshRef) {
    uft_ctx *c = UftRuntime::get().ctx;
    uft_params p;
    uft_params_get(c, &p);
    if (p.Emax != Emax) {
      p.Emax = Emax;
      PetscErrorCode e = uft_params_set(c, &p);
      if (e) return e;
    }
    return uft_stress(c, XPhys.id, VMS.id);
  }
  // SolveFem::ComputeObjectiveAndSensitivities (SolveFem.h:39-40)
  PetscErrorCode ComputeObjectiveAndSensitivities(PetscScalar *fx, PetscScalar *gx, Vec dfdx, PetscInt m, Vec *dgdx,
                                                  Vec XPhys, PetscScalar Emin, PetscScalar Emax, PetscScalar penal,
                                                  PetscScalar volfrac, MeshRef, MeshRef, MeshRef, Vec /*Vol*/, int **,
                                                  double **, int *, double *) {
    uft_ctx *c = UftRuntime::get().ctx;
    uft_params p;
    uft_params_get(c, &p);
    if (p.Emin != Emin || p.Emax != Emax || p.penal != penal || p.volfrac != volfrac) {
      p.Emin = Emin; p.Emax = Emax; p.penal = penal; p.volfrac = volfrac;
      PetscErrorCode e = uft_params_set(c, &p);
      if (e) return e;
    }
    std::vector<int> ids(m);
    for (int j = 0; j < m; ++j) ids[j] = dgdx[j].id;
    PetscErrorCode e = uft_fem_obj_sens(c, fx, gx, dfdx.id, m, ids.data(), XPhys.id);
    int it = 0, itf = 0, outer = 0;
    double rr = 0;
    uft_solve_stats(c, &it, &rr, &itf, &outer);
    std::printf("Norm of error %e, Iterations %d\n", rr, it);
    return e == UFT_ERR_NOCONV ? 0 : e;  // the reference accepts an unconverged solve
  }
};

// ------------------------------------------------------------------ PDEFilter
class PDEFilter {
 public:
  PDEFilter(MeshRef, MeshRef, MeshRef, MeshRef, PetscScalar /*rmin*/, Vec /*X*/) {}
  PetscErrorCode Filter(Vec X, Vec XFiled, PetscScalar Xmin) {
    uft_ctx *c = UftRuntime::get().ctx;
    set_xmin(c, Xmin);
    return uft_filter(c, X.id, XFiled.id);
  }
  PetscErrorCode ComputeFilteredSens(Vec dfdx, Vec *dgdx, PetscInt m) {
    std::vector<int> ids(m);
    for (int j = 0; j < m; ++j) ids[j] = dgdx[j].id;
    return uft_filter_sens(UftRuntime::get().ctx, dfdx.id, m, ids.data());
  }
  PetscErrorCode HeavisideFilter(Vec XFiled, Vec XPhys, PetscScalar x0, PetscScalar beta0, PetscScalar betamax,
                                 PetscInt itr, PetscScalar Xmin, PetscBool proj) {
    uft_ctx *c = UftRuntime::get().ctx;
    uft_params p;
    uft_params_get(c, &p);
    p.x0 = x0; p.beta0 = beta0; p.betamax = betamax; p.Xmin = Xmin; p.proj = proj;
    PetscErrorCode e = uft_params_set(c, &p);
    return e ? e : uft_heaviside(c, XFiled.id, XPhys.id, itr);
  }
  PetscErrorCode ComputeProjectedSens(Vec XFiled, Vec dfdx, Vec *dgdx, PetscInt m, PetscScalar /*x0*/) {
    std::vector<int> ids(m);
    for (int j = 0; j < m; ++j) ids[j] = dgdx[j].id;
    return uft_projected_sens(UftRuntime::get().ctx, XFiled.id, dfdx.id, m, ids.data());
  }

 private:
  static void set_xmin(uft_ctx *c, double Xmin) {
    uft_params p;
    uft_params_get(c, &p);
    if (p.Xmin != Xmin) {
      p.Xmin = Xmin;
      uft_params_set(c, &p);
    }
  }
};

// ------------------------------------------------------------------ MMA
class MMA {
 public:
  // MMA(n, m, x[, a, c, d]) (MMA.h:10-13): a, c, d default to 0, 1000, 0
  MMA(PetscInt n, PetscInt m, Vec /*x*/) : n_(n), m_(m) {}
  MMA(PetscInt n, PetscInt m, Vec /*x*/, PetscScalar *a, PetscScalar *c, PetscScalar *d) : n_(n), m_(m) {
    if (a && c && d) uft_mma_set_penalization(UftRuntime::get().ctx, a, c, d);
  }
  // restart constructors (MMA.h:14-18, MMA.cpp:5-176): state from host Vecs
  MMA(PetscInt n, PetscInt m, PetscInt itr, Vec xo1, Vec xo2, Vec U, Vec L) : n_(n), m_(m) {
    uft_ctx *c = UftRuntime::get().ctx;
    if (!(xo1.host && xo2.host && U.host && L.host)) {
      std::fprintf(stderr, "MMA restart: xo1, xo2, U, L must be host Vecs (VecCreateHost)\n");
      return;
    }
    uft_mma_state_set(c, (int)itr, xo1.host->data(), xo2.host->data(), U.host->data(), L.host->data());
    if (itr < 3) std::printf("NOT A LEGAL RESTART POINT (k<3): EXPECT BREAKDOWN\n");
  }
  MMA(PetscInt n, PetscInt m, PetscInt itr, Vec xo1, Vec xo2, Vec U, Vec L, PetscScalar *a, PetscScalar *c,
      PetscScalar *d)
      : MMA(n, m, itr, xo1, xo2, U, L) {
    if (a && c && d) uft_mma_set_penalization(UftRuntime::get().ctx, a, c, d);
  }
  PetscErrorCode SetOuterMovelimit(PetscScalar Xmin, PetscScalar Xmax, PetscScalar movlim, Vec x, Vec xmin, Vec xmax) {
    uft_ctx *c = UftRuntime::get().ctx;
    uft_params p;
    uft_params_get(c, &p);
    if (p.Xmin != Xmin || p.Xmax != Xmax || p.movlim != movlim) {
      p.Xmin = Xmin; p.Xmax = Xmax; p.movlim = movlim;
      PetscErrorCode e = uft_params_set(c, &p);
      if (e) return e;
    }
    return uft_mma_movelimit(c, x.id, xmin.id, xmax.id);
  }
  PetscErrorCode Update(Vec xval, Vec dfdx, PetscScalar *gx, Vec *dgdx, Vec xmin, Vec xmax) {
    std::vector<int> ids(m_);
    for (int j = 0; j < m_; ++j) ids[j] = dgdx[j].id;
    return uft_mma_update(UftRuntime::get().ctx, xval.id, dfdx.id, gx, ids.data(), xmin.id, xmax.id);
  }
  // MMA::Restart (MMA.cpp:302-342): copies the restart state out; the
  // iteration counter is returned by RestartIteration
  PetscErrorCode Restart(Vec xo1, Vec xo2, Vec U, Vec L) {
    if (!(xo1.host && xo2.host && U.host && L.host)) return UFT_ERR_ARG;
    return uft_mma_state_get(UftRuntime::get().ctx, &k_, xo1.host->data(), xo2.host->data(), U.host->data(),
                             L.host->data());
  }
  PetscInt RestartIteration() const { return k_; }
  PetscErrorCode SetAsymptotes(PetscScalar init, PetscScalar decrease, PetscScalar increase) {
    return uft_mma_set_asymptotes(UftRuntime::get().ctx, init, decrease, increase);
  }
  PetscErrorCode ConstraintModification(PetscBool conMod) {
    return uft_mma_constraint_modification(UftRuntime::get().ctx, conMod ? 1 : 0);
  }
  PetscErrorCode SetRobustAsymptotesType(PetscInt val) {
    if (uft_mma_set_robust_asymptotes(UftRuntime::get().ctx, (int)val))
      std::printf("ERROR in MMA.cc/h: RobustAsymptotesType cannot be set to: %d \n", (int)val);
    return 0;
  }
  PetscErrorCode KKTresidual(Vec x, Vec dfdx, PetscScalar *gx, Vec *dgdx, Vec xmin, Vec xmax, PetscScalar *norm2,
                             PetscScalar *normInf) {
    std::vector<int> ids(m_);
    for (int j = 0; j < m_; ++j) ids[j] = dgdx[j].id;
    return uft_mma_kkt_residual(UftRuntime::get().ctx, x.id, dfdx.id, gx, ids.data(), xmin.id, xmax.id, norm2,
                                normInf);
  }
  PetscScalar DesignChange(Vec x, Vec xold) {
    double ch = 0.0;
    uft_mma_design_change(UftRuntime::get().ctx, x.id, xold.id, &ch);
    return ch;
  }

 private:
  PetscInt n_, m_;
  int k_ = 0;
};

inline PetscErrorCode OptPara::SetMMA(PetscInt * /*itr*/, MMA **mma) {
  // a = 0, c = 1000, d = 0 (OptPara.cpp:166-174) are the context's MMA constants
  *mma = new MMA((PetscInt)uft_size(UftRuntime::get().ctx, UFT_SIZE_NE), m, X, nullptr, nullptr, nullptr);
  return 0;
}

// ------------------------------------------------------------------ Write2Vtu
// Write2Vtu::WriteResults (Write2Vtu.cpp:5-173) through uft_write_vtu: the
// reference's file names (output_initial.vtu for itr 0, output_final.vtu for
// -1, output_<itr>.vtu otherwise, Write2Vtu.cpp:35-46) in the -vtu_dir
// directory, 6 digits like the reference (-vtu_digits to change); nothing is
// written without -vtu_dir.
class Write2Vtu {
 public:
  PetscErrorCode WriteResults(MeshRef, MeshRef, Vec XPhys, Vec U, PetscInt itr) {
    UftRuntime &R = UftRuntime::get();
    const std::string dir = R.str("-vtu_dir", "");
    if (dir.empty()) return 0;
    char name[64];
    if (itr == 0)
      std::snprintf(name, sizeof(name), "output_initial.vtu");
    else if (itr == -1)
      std::snprintf(name, sizeof(name), "output_final.vtu");
    else
      std::snprintf(name, sizeof(name), "output_%d.vtu", (int)itr);
    const PetscErrorCode e =
        uft_write_vtu(R.ctx, (dir + "/" + name).c_str(), XPhys.id, U.id, R.integer("-vtu_digits", 6));
    if (e) std::fprintf(stderr, "Write2Vtu: %s\n", UftLastError());
    return e;
  }
};
