// Drop-in driver: the design loop of the reference's main.cpp:35-137 written
// against the drop-in classes.  The object construction order, the loop
// condition, the per-iteration call sequence and the log line are the
// reference's; the initialisation call and the timer are the only changes.
//
//   bin/uft_opt -box 140,70,40 -jitter 0.15 -seed 392000 -Emax 1 -volfrac 0.3 \
//               -rmin 0.0214285714 -maxItr 50 -ksp_rtol 1e-8
#include <chrono>

#include "uft_dropin.h"

static double wtime() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv) {
  PetscErrorCode ierr = UftInitialize(&argc, &argv);
  if (ierr) {
    std::fprintf(stderr, "UftInitialize failed (%d): no HIP device?\n", ierr);
    return ierr;
  }
  const double t_ini = wtime();

  OptPara *opt = new OptPara();
  InputMeshData *mesh = new InputMeshData(opt->XPhys);
  SolveFem *fem = new SolveFem(mesh->elementEigen, mesh->eleXCoord, mesh->eleYCoord, mesh->eleZCoord, opt->Vol,
                               opt->XPhys);
  PDEFilter *pde = new PDEFilter(mesh->elementEigen, mesh->eleXCoord, mesh->eleYCoord, mesh->eleZCoord, opt->rmin,
                                 opt->XPhys);
  MMA *mma;
  PetscInt itr = 0;
  opt->SetMMA(&itr, &mma);
  Write2Vtu *vtu = new Write2Vtu();
  ierr = pde->Filter(opt->X, opt->XFiled, opt->Xmin);
  CHKERRQ(ierr);
  ierr = pde->HeavisideFilter(opt->XFiled, opt->XPhys, opt->x0, opt->beta0, opt->betamax, itr, opt->Xmin, opt->proj);
  CHKERRQ(ierr);
  PetscScalar ch = 1.0;

  while (itr < opt->maxItr && ch > 0.01) {
    itr++;
    const double t1 = wtime();
    ierr = fem->ComputeObjectiveAndSensitivities(&(opt->fx), opt->gx, opt->dfdx, opt->m, opt->dgdx, opt->XPhys,
                                                 opt->Emin, opt->Emax, opt->penal, opt->volfrac, mesh->eleXCoord,
                                                 mesh->eleYCoord, mesh->eleZCoord, opt->Vol, mesh->loadDofs,
                                                 mesh->load, mesh->fixedDofs, mesh->vertex_loc);
    CHKERRQ(ierr);
    if (itr == 1) vtu->WriteResults(mesh->vertex, mesh->element, opt->XPhys, fem->U, 0);
    if (opt->proj) {
      ierr = pde->ComputeProjectedSens(opt->XFiled, opt->dfdx, opt->dgdx, opt->m, opt->x0);
      CHKERRQ(ierr);
    }
    ierr = pde->ComputeFilteredSens(opt->dfdx, opt->dgdx, opt->m);
    CHKERRQ(ierr);
    ierr = mma->SetOuterMovelimit(opt->Xmin, opt->Xmax, opt->movlim, opt->X, opt->xmin, opt->xmax);
    CHKERRQ(ierr);
    ierr = mma->Update(opt->X, opt->dfdx, opt->gx, opt->dgdx, opt->xmin, opt->xmax);
    CHKERRQ(ierr);
    ch = mma->DesignChange(opt->X, opt->Xold);
    ierr = pde->Filter(opt->X, opt->XFiled, opt->Xmin);
    CHKERRQ(ierr);
    ierr = pde->HeavisideFilter(opt->XFiled, opt->XPhys, opt->x0, opt->beta0, opt->betamax, itr, opt->Xmin, opt->proj);
    CHKERRQ(ierr);
    const double t2 = wtime();
    std::printf("It.: %i, fx: %f, gx[0]: %f, ch.: %f, time: %f\n", itr, opt->fx, opt->gx[0], ch, t2 - t1);
    if (itr % 20 == 0) vtu->WriteResults(mesh->vertex, mesh->element, opt->XPhys, fem->U, itr);
  }
  vtu->WriteResults(mesh->vertex, mesh->element, opt->XPhys, fem->U, -1);

  delete opt;
  delete mesh;
  delete fem;
  delete pde;
  delete mma;
  delete vtu;
  std::printf("total time: %f hours\n", (wtime() - t_ini) / 3600);
  return UftFinalize();
}
