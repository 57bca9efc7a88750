/*
 * uft.h -- C-ABI of the MI355X SIMP topology-optimization core.
 *
 * One context (uft_ctx) holds one design problem on one GPU: the mesh, the
 * boundary conditions, the parameters, every design vector of the reference's
 * OptPara (device resident), the solver hierarchy and the MMA state.  Entry
 * points mirror the calls of the reference's design loop (main.cpp:64-119);
 * each cites the reference interface it replaces.  Plain pointers and sizes
 * only; no exceptions cross this boundary.  Every function returns an int
 * status (UFT_OK == 0); uft_last_error() describes the last failure.
 *
 * Threading: a context is used from one host thread at a time; calls are
 * ordered on the context's own HIP stream.  Contexts are independent.
 *
 * Vectors are addressed by id.  Element vectors (length ne) replace the PETSc
 * Vecs of OptPara (OptPara.h:28-35); UFT_U is the displacement (length 3*nn,
 * node-interleaved DOF 3n+c, SolveFem.h:42); UFT_RX the filter's node field.
 */
#ifndef UFT_H
#define UFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct uft_ctx uft_ctx;

enum uft_status {
  UFT_OK = 0,
  UFT_ERR_ARG = 1,      /* bad argument / size / id */
  UFT_ERR_IO = 2,       /* missing or short mesh/BC file */
  UFT_ERR_NOCONV = 3,   /* Krylov solve did not reach its tolerance */
  UFT_ERR_DETJ = 4,     /* non-positive Jacobian determinant in an element */
  UFT_ERR_OOM = 5,      /* device or host allocation failed */
  UFT_ERR_HIP = 6,      /* HIP runtime error */
  UFT_ERR_STATE = 7     /* call out of order (e.g. before uft_setup) */
};

enum uft_vec_id {
  UFT_X = 0, UFT_XFILED = 1, UFT_XPHYS = 2, UFT_XOLD = 3, UFT_DFDX = 4,
  UFT_XMIN = 5, UFT_XMAX = 6, UFT_VOL = 7, UFT_U = 8, UFT_RX = 9,
  UFT_VMS = 10, /* von Mises stress per element (SolveFem::VMS, stress/ComputeStress.cpp:6) */
  UFT_DGDX0 = 16 /* UFT_DGDX0 + j for constraint j < m */
};

enum uft_size_id { UFT_SIZE_NE = 0, UFT_SIZE_NN = 1, UFT_SIZE_NDOF = 2, UFT_SIZE_NPE = 3 };

/* UFT_FP64: every operator product in FP64.  UFT_MIXED: FP64 flexible outer
 * Krylov loop (FP64 residuals, inner products, iterate; the outer product
 * K p with FP32 blocks), multigrid preconditioner in FP32 with FP32 fine
 * and coarse operators; the coarsest level's dense inverse is computed in
 * FP64 and applied as its FP32 copy.
 * UFT_MIXED_H: as UFT_MIXED with the smoother's operators stored low: with
 * fine_assembled = 1 (default) the fine level as assembled 3x3 node blocks
 * of D^-1 A in int8 with an FP32 scale per block (one right-hand side) or
 * in FP16 (several load cases), level 1's blocks in FP16 scaled to unit
 * diagonal, the smoothed prolongator's cycle copies in FP16, the coarsest
 * inverse applied from a row-scaled FP16 copy (one design, one load case);
 * with fine_assembled = 0 FP16 element blocks scaled per element.  FP32
 * arithmetic throughout the cycle. */
enum uft_precision { UFT_FP64 = 0, UFT_MIXED = 1, UFT_MIXED_H = 2 };

/* Parameters.  Defaults (uft_default_params) are the reference's
 * (OptPara.cpp:66-78, SolveFem.cpp:8-11,134, PDEFilter.cpp:117-120). */
typedef struct uft_params {
  double Emin, Emax, penal, volfrac, rmin, nu, movlim, Xmin, Xmax, x0, beta0, betamax;
  int proj, m, maxItr;
  double rtol_el;      /* elasticity solve: ||F-KU|| <= rtol_el ||F||   (-ksp_rtol) */
  double rtol_f;       /* filter CG: ||D^-1 r|| <= rtol_f ||D^-1 b||    (-ksp_rtol) */
  int maxit_el, maxit_f;
  int precision;       /* enum uft_precision */
  int amg_levels;      /* most levels incl. the fine one (reference: PCGAMGSetNlevels 4); coarsening also stops
                          at coarse_max */
  int cheb_degree;     /* Chebyshev-Jacobi smoother degree per level */
  int coarse_max;      /* stop coarsening at or below this many coarse DOFs (default 4096: one design, with one
                          right-hand side or several load cases, keeps three levels and solves the coarsest
                          exactly with its dense inverse; design batches stop at 600 per design at most) */
  int cycle_type;      /* coarse-grid cycle: 0 V, 1 W, 2 K-cycle on every coarse level, 3 K-cycle on level 1 only */
  double cheb_lo, cheb_hi; /* Chebyshev interval [cheb_lo, cheb_hi] * lambda_max estimate
                              (-mg_levels_ksp_chebyshev_esteig 0,lo,0,hi; PCGAMG: 0.1, 1.1) */
  int esteig_steps;    /* Lanczos steps of each level's lambda_max(D^-1 A) estimate */
  int load_cases;      /* with several load cases: 0 superposed into one right-hand side (the reference's
                          INSERT, SolveFem.cpp:277-283); 1 separate right-hand sides (at most 8) solved
                          together, compliance and sensitivities summed over the cases */
  int graph;           /* 1: each Krylov iteration of the elasticity solve after the first is captured once
                          in a hipGraph and replayed (re-captured when a single design's hierarchy is
                          refreshed); 0 (default): launched kernel by kernel.  Same results either way.
                          While the device timers are on (uft_profile_enable) the solve launches kernel by
                          kernel. */
  int cheb_degree_coarse; /* Chebyshev degree on multigrid levels 2 and below (their applies are bound by
                             launch latency, not by bandwidth); 0 (default): cheb_degree */
  int fine_assembled;  /* UFT_MIXED_H, and UFT_MIXED with one load case (FP32 blocks for its smoother and
                          outer product): 1 (default) the fine operators are assembled at each hierarchy
                          refresh into 3x3 node blocks,
                          the smoother's rows divided by the diagonal in int8 with an FP32 scale per block
                          (FP16 with several load cases) and FP32 for the outer product (one
                          launch per apply, no element-local round trip); with 5-8 load cases the
                          eight-right-hand-side products run on node tiles whose neighbour entries are
                          staged in LDS; 2: the same without the node tiles; 0: the element blocks.
                          Fixed at setup. */
  int coarse_dense_max; /* the coarsest multigrid level is solved with its dense inverse (recomputed at every
                           hierarchy refresh), so coarsening goes on until a level has at most this many
                           unknowns (default 4096, at most 6144; design batches 3072 per design at most).
                           Fixed at setup. */
  int prolong_smooth;  /* 1 (default): smoothed aggregation on the fine level as PCGAMG does (SolveFem.cpp:129-136):
                          at every hierarchy refresh the tentative prolongator is smoothed once with the current
                          operator, P = (I - omega D^-1 A) P_tent with omega = 4 / (3 lambda_max(D^-1 A)), and the
                          level-1 operator is P^T A P (per design in a batch); used with the assembled node pattern
                          (fine_assembled = 1) and at least three levels, otherwise the tentative prolongator.
                          0: the tentative prolongator.  Fixed at setup. */
  int coarse_lag;      /* one design with a coarsest level of 512 unknowns or more: 1 (default) the cycle applies
                          the dense inverse of the previous hierarchy refresh's coarsest operator while the current
                          one is inverted on a stream of its own, beside the solve (the first refresh inverts its
                          own operator before the solve); 0: every refresh inverts its own coarsest operator
                          before the solve.  Either way the outer iteration stops on the same residual test, so
                          the solution is the same to the tolerance.  Fixed at setup. */
} uft_params;

/* --- context ------------------------------------------------------------ */
/* Replaces PetscInitialize + OptPara() (main.cpp:41,50). */
int uft_create(uft_ctx **ctx, int device);
void uft_destroy(uft_ctx *ctx);
const char *uft_last_error(const uft_ctx *ctx);
void uft_default_params(uft_params *p);
/* OptPara::SetUp / SetUpOpt defaults and option overrides (OptPara.cpp:61-120). */
int uft_params_set(uft_ctx *ctx, const uft_params *p);
int uft_params_get(const uft_ctx *ctx, uft_params *p);

/* --- mesh and boundary conditions -------------------------------------- */
/* InputMeshData::GetNodeAndElementData (InputMeshData.cpp:33-226): elem is
 * ne x npe 0-based node ids (npe 8 = hex8 diagonal numbering, 4 = tet4),
 * xyz is nn x 3. */
int uft_mesh_set(uft_ctx *ctx, int npe, int64_t nn, int64_t ne, const int32_t *elem, const double *xyz);
/* Same, reading the reference text formats vertex.txt / element.txt
 * (InputMeshData.cpp:100-138): whitespace-separated, no terminator needed. */
int uft_mesh_load_text(uft_ctx *ctx, int npe, int64_t nn, int64_t ne, const char *vertex_path,
                       const char *element_path);
/* InputMeshData::GetBcs (InputMeshData.cpp:228-285): fixed DOFs, and nlc load
 * cases of nload (dof, value) pairs each, INSERTed into one RHS (SolveFem.cpp:277-283). */
int uft_bc_set(uft_ctx *ctx, int nfixed, const int32_t *fixed, int nlc, int nload,
               const int32_t *load_dofs, const double *load_vals);
/* Same, reading dir/fixeddofs.txt, dir/loaddof_%d.txt, dir/load_%d.txt
 * (InputMeshData.cpp:238,252,271). */
int uft_bc_load_text(uft_ctx *ctx, const char *dir, int nfixed, int nlc, int nload);
/* The two text readers above without a context or a GPU (host arrays only;
 * errors in uft_last_error(NULL)): xyz receives nn x 3 doubles and elem
 * ne x npe ids (InputMeshData.cpp:100-149); fixed receives nfixed DOFs,
 * load_dofs / load_vals nlc x nload entries, case-major
 * (InputMeshData.cpp:228-285). */
int uft_mesh_read_text(int npe, int64_t nn, int64_t ne, const char *vertex_path, const char *element_path,
                       double *xyz, int32_t *elem);
int uft_bc_read_text(const char *dir, int nfixed, int nlc, int nload, int32_t *fixed, int32_t *load_dofs,
                     double *load_vals);
/* Inspection of the node tiles that the eight-right-hand-side products use
 * (no counterpart in the reference; host only, no context or GPU): the tiles
 * of the mesh's node pattern (at most 256 nodes each, halos of at most 840
 * nodes).  Sizes into ntile, nhalo (total halo entries) and hmax; with node
 * (ntile x 256: the node at each tile position, -1 for padding), hoff
 * (ntile + 1) and hnode (nhalo) given, the tiles themselves (call once for
 * the sizes, then with arrays of those sizes).  UFT_ERR_ARG when the
 * pattern cannot be tiled. */
int uft_node_tiles_host(int npe, int64_t nn, int64_t ne, const int32_t *elem, const double *xyz, int64_t *ntile,
                        int64_t *nhalo, int *hmax, int32_t *node, int32_t *hoff, int32_t *hnode);
/* Binary cache of the mesh and BCs set on ctx (no counterpart in the
 * reference, which re-parses the text files on every start,
 * InputMeshData.cpp:100-149,228-285): uft_cache_save writes them with a
 * checksum; uft_cache_load reads a cache back (UFT_ERR_IO on a bad header,
 * a truncated file or a checksum mismatch) as uft_mesh_set + uft_bc_set. */
int uft_cache_save(uft_ctx *ctx, const char *path);
int uft_cache_load(uft_ctx *ctx, const char *path);
/* Builds the per-element operators, volumes, filter, solver hierarchy and MMA
 * state: SolveFem ctor (SolveFem.cpp:6-15), PDEFilter ctor
 * (PDEFilter.cpp:36-140), OptPara::SetMMA (OptPara.cpp:158-177); sets
 * X = XFiled = XPhys = Xold = volfrac (OptPara.cpp:150-153). */
int uft_setup(uft_ctx *ctx);
/* Host copy of the mesh (elem ne x npe, xyz nn x 3), for output such as the
 * VTU writer (Write2Vtu.cpp:5-173 gathers the same data on rank 0). */
int uft_mesh_get(const uft_ctx *ctx, int32_t *elem, double *xyz);
/* Write2Vtu::WriteResults (Write2Vtu.cpp:5-173): ASCII VTU of the context's
 * mesh with the displacement u_id (point data ux, uy, uz) and the element
 * field xphys_id (cell data xphys), copied from the device; hex8 corners in
 * VTK order 0,1,3,2,4,5,7,6 (Write2Vtu.cpp:89-96), cell type 12 (hex8) or
 * 10 (tet4); numbers in scientific notation with `digits` digits after the
 * point (the reference writes 6; 17 round-trips FP64).  UFT_ERR_IO when the
 * file cannot be written. */
int uft_write_vtu(uft_ctx *ctx, const char *path, int xphys_id, int u_id, int digits);
/* The same file from host arrays, without a context or a GPU (errors in
 * uft_last_error(NULL)). */
int uft_vtu_write_host(const char *path, int npe, int64_t nn, int64_t ne, const double *xyz, const int32_t *elem,
                       const double *u, const double *xphys, int digits);
int64_t uft_size(const uft_ctx *ctx, int which);
double uft_total_volume(const uft_ctx *ctx);

/* --- vectors (PETSc Vec replacement) ------------------------------------ */
int uft_vec_get(uft_ctx *ctx, int id, double *host, int64_t n);
int uft_vec_set(uft_ctx *ctx, int id, const double *host, int64_t n);
int uft_vec_fill(uft_ctx *ctx, int id, double value);
int uft_vec_copy(uft_ctx *ctx, int src, int dst);
/* Device address of a vector (stream-ordered on the context stream). */
void *uft_vec_device(uft_ctx *ctx, int id);

/* --- the design iteration (main.cpp:72-107) ------------------------------ */
/* SolveFem::ComputeObjectiveAndSensitivities (SolveFem.h:39-40,
 * SolveFem.cpp:159-265): solves K(xphys) U = F, then fx = sum E(x) u^T KE u,
 * gx[0] = sum x V / TV - volfrac, dfdx, dgdx[0] = V/TV. */
int uft_fem_obj_sens(uft_ctx *ctx, double *fx, double *gx, int dfdx_id, int m, const int *dgdx_ids,
                     int xphys_id);
/* PDEFilter::Filter (PDEFilter.h:22, PDEFilter.cpp:142-165). */
int uft_filter(uft_ctx *ctx, int x_id, int xf_id);
/* PDEFilter::ComputeFilteredSens (PDEFilter.h:24, PDEFilter.cpp:167-183). */
int uft_filter_sens(uft_ctx *ctx, int dfdx_id, int m, const int *dgdx_ids);
/* PDEFilter::HeavisideFilter (PDEFilter.h:26, PDEFilter.cpp:242-293). */
int uft_heaviside(uft_ctx *ctx, int xf_id, int xp_id, int itr);
/* PDEFilter::ComputeProjectedSens (PDEFilter.h:28, PDEFilter.cpp:295-341). */
int uft_projected_sens(uft_ctx *ctx, int xf_id, int dfdx_id, int m, const int *dgdx_ids);
/* MMA::SetOuterMovelimit (MMA.h:43-44, MMA.cpp:370-389). */
int uft_mma_movelimit(uft_ctx *ctx, int x_id, int xmin_id, int xmax_id);
/* MMA::Update (MMA.h:23, MMA.cpp:483-502): x is overwritten in place. */
int uft_mma_update(uft_ctx *ctx, int x_id, int dfdx_id, const double *gx, const int *dgdx_ids,
                   int xmin_id, int xmax_id);
/* MMA::DesignChange (MMA.h:52, MMA.cpp:391-410): ch = max|x - xold|; xold := x. */
int uft_mma_design_change(uft_ctx *ctx, int x_id, int xold_id, double *ch);
/* Whole loop body of main.cpp:72-107 on the context's own vectors. */
int uft_design_iteration(uft_ctx *ctx, int itr, double *fx, double *gx, double *ch);

/* Allocation-free design change.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP stream/graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * result as uft_mma_design_change on caller-owned DEVICE arrays, without a
 * context: *ch = max_i |x[i] - xold[i]| (0 for n = 0), then xold := x.
 * x, xold (n doubles) and ch (one double) are device pointers on GPU
 * `device`; stream is a hipStream_t of that device or NULL (its null
 * stream); the result is stream-ordered (synchronise before reading ch).
 * The block partials live in a 512 KiB static device array per GPU, held
 * until the process ends, so n is limited to 2^24 (UFT_ERR_ARG beyond, with
 * the limit in uft_last_error(NULL)).  The call makes no hipMalloc or hipFree
 * and creates no stream or event.  All calls on one GPU share its array:
 * calls that overlap on one GPU race. */
int uft_mma_design_change_noalloc(int device, const double *x, double *xold, int64_t n, double *ch, void *stream);

/* --- diagnostics and checks ---------------------------------------------- */
/* Iterations and true relative residual of the last elasticity solve
 * (the "Norm of error %e, Iterations %d" line, SolveFem.cpp:142-152), and the
 * iteration count of the last filter solve. */
int uft_solve_stats(const uft_ctx *ctx, int *it_el, double *relres_el, int *it_f, int *outer_el);
/* y = K(xphys) u with Dirichlet rows/columns eliminated and unit diagonal
 * (SolveFem.cpp:297-333, matrix-free), host buffers of length 3*nn. */
int uft_apply_K(uft_ctx *ctx, int xphys_id, const double *u, double *y);
/* Unscaled element stiffness KE of element e as a full nd x nd row-major
 * matrix (SolveFem.cpp:339-408 / tet4/Tet4Isoparametric.cpp). */
int uft_element_ke(uft_ctx *ctx, int64_t e, double *ke);
/* Helmholtz element matrix KF (npe x npe) and T (npe) (PDEFilter.cpp:185-240). */
int uft_element_kf(uft_ctx *ctx, int64_t e, double *kf, double *t);
/* Shape of the solver hierarchy (no counterpart in the reference, whose GAMG
 * hierarchy is internal to PETSc): level count (at most 8), unknowns per level
 * and, per level, the element count (level 0) or 6x6 block count. */
int uft_solver_info(const uft_ctx *ctx, int *nlev, int64_t *n_per_level, int64_t *blocks_per_level);
/* The assembled fine operators (fine_assembled; no counterpart in the
 * reference): slots of their sliced-ELL storage (3x3 node blocks, padding
 * included), nonzero node blocks, bytes per stored column index (2 or 4) and
 * bytes per stored smoother block (12: int8 entries with an FP32 block scale,
 * one right-hand side; 18: FP16, the multi-load-case form; the mixed outer
 * product's blocks are FP32; 72: the FP64 form's FP64 blocks, which its
 * solves apply when the hierarchy is smoothed aggregation); all 0 when the
 * fine level uses element blocks. */
int uft_fine_operator_info(const uft_ctx *ctx, int64_t *slots, int64_t *blocks, int *col_bytes, int *value_bytes);
/* Node tiles of the eight-right-hand-side fine products (fine_assembled = 1
 * with 5-8 load cases; no counterpart in the reference): number of tiles (of
 * at most 256 nodes, one workgroup each), slots of their tile-ordered
 * sliced-ELL storage and the largest halo (neighbour nodes of one tile,
 * staged in LDS, at most 840); all 0 when the node-ordered kernels run. */
int uft_fine_tiles_info(const uft_ctx *ctx, int64_t *tiles, int64_t *slots, int *max_halo);
/* --- multi-load-case form (load_cases = 1; no counterpart in the reference,
 * which superposes its cases) --------------------------------------------- */
/* Number of load cases; per case the Krylov iterations and final relative
 * residual of the last solve (arrays of up to 8).  0 in the other forms. */
int uft_load_case_stats(const uft_ctx *ctx, int *iters, double *relres);
/* Displacement of load case k after the last solve (host, 3*nn doubles);
 * UFT_U holds case 0. */
int uft_load_case_u(uft_ctx *ctx, int k, double *host);
/* y_k = K(XPhys) u_k for every load case through the multi-load-case
 * operator kernels (host arrays ncase x 3nn, case-major): kind 0 the FP64
 * kernel of the outer loop, kind 1 the preconditioner's (FP32 vectors;
 * BF16 matrix-core element products in the mixed forms). */
int uft_apply_K_cases(uft_ctx *ctx, int kind, const double *u, double *y);
/* Diagnostic (no counterpart in the reference): y = A_l x with the operator
 * of multigrid level l as of the last hierarchy refresh (level 0 = the fine
 * stiffness of XPhys), in the stored precision kind 0 = FP64, 1 = FP32,
 * 2 = the UFT_MIXED_H smoother's low-precision blocks (int8 with a block scale or
 * FP16); x, y of length n_l. */
int uft_level_apply(uft_ctx *ctx, int level, int kind, const double *x, double *y);
/* Measurement (no counterpart in the reference): average duration in
 * microseconds of reps back-to-back launches of a fine operator apply as
 * the solve issues it, on the operators of the last hierarchy refresh and
 * the solver's own work vectors (their contents are overwritten), timed with
 * HIP events on the context's stream: kind 0 the FP64 stiffness apply
 * (the assembled FP64 blocks when the solves use them, else element and
 * node phase), 1 the mixed forms' FP32 outer product, 2 the smoother's
 * operator with its Chebyshev-step epilogue (UFT_MIXED_H, or UFT_MIXED on
 * assembled FP32 blocks), 3 the FP64 element phase alone. Kind 0 needs the
 * FP64 form, kinds 1 and 2 a mixed form; one design.  With several load cases (load_cases = 1) kinds
 * 0-2 apply to all right-hand sides at once (kind 3 is not available). */
int uft_time_fine_operator(uft_ctx *ctx, int kind, int reps, double *us);

/* --- MMA settings, restart and KKT residual (MMA.h:8-52) ------------------ */
/* MMA::SetAsymptotes (MMA.cpp:347-355): asymptote initialisation,
 * decrease and increase factors (defaults 0.5, 0.7, 1.2). */
int uft_mma_set_asymptotes(uft_ctx *ctx, double init, double decrease, double increase);
/* MMA::SetRobustAsymptotesType (MMA.cpp:357-368): 0 (default) or 1; any
 * other value selects 0 and returns UFT_ERR_ARG. */
int uft_mma_set_robust_asymptotes(uft_ctx *ctx, int type);
/* MMA::ConstraintModification (MMA.h:33-36): 1 adds the objective's
 * regularisation to the constraint approximations (default 0). */
int uft_mma_constraint_modification(uft_ctx *ctx, int on);
/* User-defined subproblem penalisation a, c, d (m each) of
 * MMA(n, m, x, a, c, d) (MMA.h:13; defaults a = 0, c = 1000, d = 0). */
int uft_mma_set_penalization(uft_ctx *ctx, const double *a, const double *c, const double *d);
/* MMA::Restart (MMA.cpp:302-342): the state a restart needs -- iteration
 * counter k and xo1, xo2, U, L (host buffers of ne doubles). */
int uft_mma_state_get(uft_ctx *ctx, int *k, double *xo1, double *xo2, double *U, double *L);
/* The restart constructors MMA(n, m, k, xo1, xo2, U, L[, a, c, d])
 * (MMA.cpp:5-176): load that state.  k < 3 is accepted with the reference's
 * warning in uft_last_error ("not a legal restart point"). */
int uft_mma_state_set(uft_ctx *ctx, int k, const double *xo1, const double *xo2, const double *U, const double *L);
/* MMA::KKTresidual (MMA.cpp:412-480) of the last uft_mma_update's
 * multipliers: 2-norm and inf-norm. */
int uft_mma_kkt_residual(uft_ctx *ctx, int x_id, int dfdx_id, const double *gx, const int *dgdx_ids, int xmin_id,
                         int xmax_id, double *norm2, double *norminf);
/* SolveFem::ComputeStress (stress/ComputeStress.cpp:1-113): von Mises
 * stress of the current displacement U at each element's centroid,
 * sqrt(xphys) * Emax * D * B(0) * u_e, into vector vms_id (UFT_VMS).  For
 * tet4 the constant strain is used (the reference's routine is hex8 only). */
int uft_stress(uft_ctx *ctx, int xphys_id, int vms_id);

/* Multigrid inspection (no counterpart in the reference): the 6x6-block
 * operator of level >= 1 as refreshed for the last solve (rp: units+1, ci and
 * val: blocks and 36*blocks, row-major blocks), and the transfer of level l to
 * l+1 (parent: units of level l; P: units * bs * 6 with bs = 3 on level 0, 6
 * above; the operator of level l+1 is P^T A_l P). */
int uft_amg_operator_get(uft_ctx *ctx, int level, int32_t *rp, int32_t *ci, double *val);
int uft_amg_transfer_get(uft_ctx *ctx, int level, int32_t *parent, double *P);
/* With prolong_smooth: the smoothed prolongator of level 0 as of the last
 * refresh (the transfer the cycle uses; uft_amg_transfer_get then returns the
 * tentative one it was smoothed from): *nent entries, row pointers pptr
 * (nodes + 1) and aggregate ids pagg per entry, P with a row-major 3 x 6
 * block per entry, and omega.  *nent = 0 without it.  Any of pptr, pagg, P,
 * omega may be NULL (call with them NULL first to size the arrays). */
int uft_amg_prolongator_get(uft_ctx *ctx, int64_t *nent, int32_t *pptr, int32_t *pagg, double *P, double *omega);
/* The dense inverse of the coarsest operator used by the coarse solve
 * (row-major n x n, n = unknowns of the last level). */
int uft_amg_coarse_inverse_get(uft_ctx *ctx, double *inv);
/* Wall time in ms of the phases of the last uft_design_iteration (the
 * reference prints only the whole iteration, main.cpp:112): 0 solver
 * hierarchy refresh, 1 elasticity Krylov solve, 2 sensitivities, 3 sensitivity
 * filter, 4 MMA update + design change, 5 density filter + projection,
 * 6 whole iteration.  ms must hold 7 doubles. */
int uft_phase_times(const uft_ctx *ctx, double *ms);
/* Autotuner (no counterpart in the reference): times every element-phase
 * kernel variant of the fine stiffness apply (workgroup size, streaming
 * loads, occupancy floor) on the device and keeps the fastest for this mesh;
 * uft_setup runs it once.  All variants give bit-identical results.
 * variant[0]/[1] receive the FP64/FP32 choice; us_per_variant (nmax entries)
 * the FP64 timings in microseconds. */
int uft_autotune(uft_ctx *ctx, int reps, int *variant, double *us_per_variant, int nmax);
/* Smoother-schedule autotuner (no counterpart in the reference, whose
 * multigrid schedule is fixed by options): times one elasticity solve at the
 * current XPhys for each Chebyshev degree 1..3 and each cycle type (V, W,
 * K-cycle; V and W for design batches), keeps the fastest schedule in the
 * context's parameters and returns it; ms (12 doubles, [degree - 1][cycle])
 * receives the times, an unconverged schedule timing as 1e300.  The solve
 * leaves UFT_U as the last candidate's solution. */
int uft_autotune_smoother(uft_ctx *ctx, int *degree, int *cycle_type, double *ms);
/* Synchronise the context stream. */
int uft_sync(uft_ctx *ctx);
/* Device timers: HIP events recorded on the context stream around the
 * launches of the timed kernels while enabled (no counterpart in the
 * reference, whose only instrumentation is MPI_Wtime around the loop body,
 * main.cpp:74,109).  on = k > 0 times one in every k fine applies (1: all;
 * a sample keeps the events' own cost out of the throughput); 0 disables.
 * uft_profile_read returns the number of launches timed and their summed
 * duration in ms for kind 0 = element phase of the FP64 fine stiffness
 * apply, 1 = its node phase, 2/3 = the same for the FP32 apply of the mixed
 * form. */
int uft_profile_enable(uft_ctx *ctx, int on);
int uft_profile_read(uft_ctx *ctx, int kind, int64_t *count, double *total_ms);
/* Solve traffic since the last uft_profile_enable (no counterpart in the
 * reference): the summed algorithmic bytes of every operator apply made
 * inside the elasticity solves -- fine level (element blocks: Ne (S w +
 * 4 npe + 8), S unique entries of w bytes; assembled node blocks: blocks x
 * (block bytes + column index)), coarse levels (blocks x (36 w + 4)) and the
 * coarsest dense inverse (entries x w), each plus its input and output
 * vectors once; transfers and vector updates are not counted, the hierarchy
 * refresh is excluded -- the summed wall time of those solves, the number of
 * applies and of solves.  bytes / ms is a lower bound of the solve's HBM
 * rate. */
int uft_solve_traffic(const uft_ctx *ctx, double *bytes, double *ms, int64_t *applies, int64_t *solves);

/* --- workload generator (synthetic cantilever) --------------------------- */
int64_t uft_box_nodes(int nx, int ny, int nz);
int64_t uft_box_elems(int npe, int nx, int ny, int nz);
/* Box [0,lx]x[0,ly]x[0,lz], node id i+(nx+1)(j+(ny+1)k), interior nodes
 * jittered by (2u-1)*jitter*h per axis from a SplitMix64 stream. */
int uft_gen_box(int npe, int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                uint64_t seed, double *xyz, int32_t *elem);
/* Cantilever BCs: x=0 face clamped; load_kind 0 = -1/(nz+1) on the y-DOFs of
 * the bottom edge of x=lx, 1 = -1/9 on a 3x3 node patch at mid-height,
 * 2 = eight load cases of nine entries each (load_dofs/load_vals need 72):
 * 3x3 patches of -1/9 on x=lx centred at (j, k) = (1 + floor(u (ny-1)),
 * 1 + floor(u (nz-1))), u from the SplitMix64 stream of seed 8. */
int uft_box_bcs(int nx, int ny, int nz, int load_kind, int32_t *fixed, int *nfixed, int32_t *load_dofs,
                double *load_vals, int *nload);

/* --- independent designs in one context ---------------------------------
 * No counterpart in the reference, which optimises one design per run; for
 * callers that optimise many designs (parameter sweeps, ensembles).  The mesh
 * given to uft_mesh_set is the concatenation of ndesign independent meshes:
 * design d owns elements [elem_off[d], elem_off[d+1]) and nodes
 * [node_off[d], node_off[d+1]), and its elements use only its own nodes;
 * uft_bc_set takes the fixed and load DOFs in the concatenated numbering.
 * volfrac[d] and rmin[d] replace those parameters for design d; all other
 * parameters are shared.  Call after uft_mesh_set and before uft_setup.
 * Every kernel launch then covers all designs, while each design keeps its
 * own Krylov scalars and stopping tests, filter solve, volume constraint,
 * sensitivities and MMA dual problem, so each design follows the iterates of
 * a context that holds it alone (to rounding, and to the solver tolerance:
 * its multigrid levels are formed per design).  uft_design_iteration then
 * advances every design by one iteration and returns fx = the sum of the
 * compliances, gx[0] = the largest volume-constraint value and ch = the
 * largest design change; uft_batch_results returns them per design.
 * Restrictions: m = 1, superposed load cases, at most 64 designs, each
 * design's coarsest multigrid level at most 3072 unknowns; the K-cycle
 * (cycle_type 2, 3) runs as the W-cycle. */
int uft_batch_set(uft_ctx *ctx, int ndesign, const int64_t *elem_off, const int64_t *node_off,
                  const double *volfrac, const double *rmin);
/* Per design results of the last iteration (any pointer may be NULL; arrays
 * of ndesign): compliance, volume constraint, design change, elasticity
 * Krylov iterations and final relative residual, filter CG iterations.
 * Returns ndesign (0 without a batch). */
int uft_batch_results(const uft_ctx *ctx, double *fx, double *gx, double *ch, int *it_el,
                      double *relres_el, int *it_f);

#ifdef __cplusplus
}
#endif
#endif
