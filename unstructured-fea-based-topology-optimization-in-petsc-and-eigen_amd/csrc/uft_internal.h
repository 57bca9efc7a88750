// Internal declarations shared by the host orchestration (uft_api.cpp,
// uft_amg_setup.cpp, uft_mesh.cpp) and the gfx950 kernels (*.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "uft.h"

#define UFT_MAXM 8
#define UFT_MAXLEV 8
#define UFT_RED_BLOCK 256
#define UFT_TAIL_MAX 3072  // coarsest unknowns handled by the fused coarse tail (LDS)
#define UFT_SWEEP_MIN 512  // coarsest levels from this size are inverted by the tiled sweep (kernels_dense.hip)
#define UFT_ELL_SLICE 64      // nodes per slice of the assembled fine operator (one wavefront)
// node tiles of the multi-load-case fine products: UFT_TILE nodes per
// workgroup, whose neighbours' vector entries (at most UFT_TILE_HCAP nodes,
// 3 x 8 FP32 values each, 80,640 B of LDS) are staged once per tile
#define UFT_TILE 256
#define UFT_TILE_HCAP 840

// Per-row epilogue of a fused operator apply t = A u (see kernels_cycle.hip):
//  STORE     y = t
//  STEP      r -= t; dout = c1 u + c2 D^-1 r; x += dout (+ u when addprev)
//  RES       r = base - t
//  RESFIRST  r = base - t; dout = it D^-1 r
//  RESFIRSTX r = kr = base - t; dout = kd = it D^-1 r
//  DOTS2     y = t; partials of u.t, u.w1
//  DOTS3     partials of u.w1, u.t, u.w2
//  STOREDOT  y = t; partials of u.t
enum { ROW_STORE = 0, ROW_STEP, ROW_RES, ROW_RESFIRST, ROW_RESFIRSTX, ROW_DOTS2, ROW_DOTS3, ROW_STOREDOT };
template <typename T>
struct RowOp {
  int mode = ROW_STORE;
  int addprev = 0;
  T *y = nullptr, *r = nullptr, *dout = nullptr, *x = nullptr, *kr = nullptr, *kd = nullptr;
  const T *base = nullptr, *dinv = nullptr, *w1 = nullptr, *w2 = nullptr;
  double c1 = 0.0, c2 = 0.0, it = 0.0;
  double *part = nullptr;
};

// Segment grid of a design batch: blocks of at most 256 consecutive rows,
// none spanning two designs, so that block partial sums are per design and a
// block's design is uniform.  span == nullptr: the plain grid of nb(n) blocks
// over all rows.
struct SegG {
  const int64_t *span = nullptr;  // per block: first row, end row
  const int32_t *des = nullptr;   // per block: design
  int nb = 0;
};
#define UFT_MAXSEG 64   // designs per batch
#define UFT_BSC 96      // per-design device scalars

// Slots of a batch design's device scalars (sc + d * UFT_BSC): elasticity
// flexible CG, filter CG, Lanczos, then general reductions from B_RED.
enum {
  B_RZ = 0, B_PQ = 1, B_ALPHA = 2, B_RR = 3, B_RZN = 4, B_RZO = 5, B_BETA = 6, B_FN = 7, B_TOL = 8, B_ACT = 9,
  B_IT = 10, B_TR = 11, B_RT = 12, B_NEED = 13,
  B_FB = 16, B_FZZ = 17, B_FBETA = 18, B_FBO = 19, B_FPQ = 20, B_FAL = 21, B_FPC = 22, B_FTOL = 23, B_FACT = 24,
  B_FIT = 25, B_FSN = 26, B_FDIV = 27, B_FZZN = 28, B_FZRN = 29,
  B_LRZ = 32, B_LPQ = 33, B_LAL = 34, B_LRR = 35, B_LRZN = 36, B_LBE = 37,
  B_RED = 40,
  B_KC = 44,  // K-cycle scalars of levels 1 and 2, eight each (the single-design K-cycle's slots)
  // the MMA dual interior-point state (m = 1): lambda, mu, epsilon, residual,
  // inner steps of the level, all steps, finished, stepping, y, z, b, tolerance
  B_MLAM = 64, B_MMU = 65, B_MEPS = 66, B_MERR = 67, B_MLOOP = 68, B_MSTEPS = 69, B_MDONE = 70, B_MSTEP = 71,
  B_MY = 72, B_MZ = 73, B_MB = 74, B_MTOL = 75
};
// batch-wide flags of the device-controlled loops (three doubles each):
// elasticity CG: 0 some design asks for its FP64 residual, 1 some design is
// active, 2 FP64 residual passes so far; MMA dual: 0 some design steps, 1 some
// stepping design had a negative lambda
enum { BF_NEED = 0, BF_ACT = 1, BF_NRES = 2, MF_STEP = 0, MF_NEG = 1 };
#define UFT_BATCH_KLEV 2  // coarse levels that run the K-cycle in a design batch (the W-cycle below)

// E(x) = Emin + x^penal * (Emax - Emin)   (SolveFem.cpp:304)
struct SimpCoef {
  double Emin, dE, penal;
};

// --------------------------------------------------------------------------
// Multigrid hierarchy (host-built aggregates, device-applied V-cycle)
// --------------------------------------------------------------------------
struct AmgLevel {
  // operator of this level (level 0 is the matrix-free fine operator)
  int64_t n = 0;        // scalar unknowns
  int64_t nunit = 0;    // nodes (level 0) or aggregates (level >= 1)
  int bs = 3;           // unknowns per unit: 3 on level 0, 6 above
  int64_t nblk = 0;     // BSR blocks (level >= 1)
  int32_t *d_rp = nullptr, *d_ci = nullptr;  // BSR pattern (level >= 1)
  int32_t *d_brow = nullptr;                 // block -> block row (level >= 1)
  double *d_val = nullptr;                   // nblk * 36
  float *d_valf = nullptr;
  uint16_t *d_valh = nullptr;                // FP16 D^-1/2 A D^-1/2 (UFT_MIXED_H, level 1)
  float *d_dh = nullptr;                     // D^1/2 per scalar row (with d_valh)
  double *d_dinv = nullptr;
  float *d_dinvf = nullptr;
  double lmax = 2.0;
  // transfer to the next coarser level: each unit u of this level belongs to
  // parent[u] on the next level with a bs x 6 block P_u.
  double *d_P = nullptr;
  float *d_Pf = nullptr;
  int32_t *d_parent = nullptr;
  int32_t *d_cptr = nullptr, *d_child = nullptr;  // next-level unit -> children
  // level 0, smoothed prolongator P = (I - omega D^-1 A) P_tent of the
  // current operator (rebuilt at each refresh; see uft_amg_build_host)
  bool sa = false;
  int64_t sa_np = 0, sa_nv = 0;      // prolongator entries, entries of A P (3x6 blocks each)
  int32_t *d_sa_pptr = nullptr, *d_sa_pagg = nullptr, *d_sa_pnode = nullptr;
  int32_t *d_sa_vptr = nullptr, *d_sa_vagg = nullptr, *d_sa_vnode = nullptr;
  int32_t *d_sa_sptr = nullptr, *d_sa_snode = nullptr, *d_sa_sidx = nullptr;
  double *d_sa_V = nullptr;          // A P, sa_nv x 18 (FP64 form)
  float *d_sa_Vf = nullptr;          // the same in the mixed forms
  int32_t *d_sa_ublk = nullptr;      // the next level's upper blocks (I <= J), in block order
  int64_t sa_nub = 0;
  double sa_omega = 0.0;       // omega of the next refresh
  double sa_omega_used = 0.0;  // omega of the last one
  // contribution lists built once at setup (fixed order): per prolongator
  // entry the slots of its row whose column lies in its aggregate; per entry
  // of A P the (slot, prolongator entry) pairs; per upper block of level 1 the
  // (prolongator entry, A P entry) pairs of its support nodes
  int32_t *d_sa_wptr = nullptr, *d_sa_wslot = nullptr;
  int32_t *d_sa_vlptr = nullptr, *d_sa_vlslot = nullptr, *d_sa_vlent = nullptr;
  int32_t *d_sa_gptr = nullptr, *d_sa_gp = nullptr, *d_sa_gv = nullptr;
  // the cycle's copies of P: sliced ELL by node (slice of 64 nodes, entry
  // slot s = poff + 64 q + n for entry q of lane n; FP64/FP32: value m at
  // 18 (s - n) + 64 m + n; FP16: three planes, values 0-7 at 8 s, 8-15 at
  // 8 (sa_ne_ell + s), 16-17 at 16 sa_ne_ell + 2 s; aggregate -1 on padding)
  // for the prolongation, support order (18 values per entry) for the
  // restriction
  int64_t sa_ne_ell = 0;
  int64_t *d_sa_poff = nullptr;      // per slice, first entry slot (nslice + 1)
  int32_t *d_sa_pw = nullptr;        // per slice, entry slots per node
  int32_t *d_sa_eagg = nullptr;      // per entry slot, aggregate or -1
  int32_t *d_sa_spos = nullptr;      // per entry, its position in the support order
  double *d_sa_Pe = nullptr, *d_sa_Ps = nullptr;  // FP64 cycle
  float *d_sa_Pef = nullptr, *d_sa_Psf = nullptr; // FP32 cycle
  uint16_t *d_sa_Peh = nullptr, *d_sa_Psh = nullptr;  // FP16 copies (UFT_MIXED_H)
  // deterministic Galerkin contribution lists for the next level's blocks
  int64_t ngl = 0;
  int32_t *d_glptr = nullptr;  // per next-level block
  int32_t *d_gl = nullptr;     // level 0: element of each entry ; else block index
  int32_t *d_gtr = nullptr;    // level 0: block -> its transposed block
  double *d_G = nullptr;       // level 0: 6x6 density-free entry matrices (ngl * 36)
  // work vectors (length n)
  double *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr, *t = nullptr;
  float *xf = nullptr, *bf = nullptr, *rf = nullptr, *df = nullptr, *tf = nullptr;
  // coarse levels: W/K-cycle work vectors
  double *kc = nullptr, *kv = nullptr, *kd = nullptr, *kr = nullptr;
  float *kcf = nullptr, *kvf = nullptr, *kdf = nullptr, *krf = nullptr;
};

struct AmgHier {
  int nlev = 0;  // levels incl. fine; last level solved densely
  AmgLevel L[UFT_MAXLEV];
  int64_t ncoarse = 0;
  double *d_cmat = nullptr;  // dense coarsest matrix / its inverse (ncoarse^2)
  double *d_cinv = nullptr;
  float *d_cinvf = nullptr;
  double *d_cwork = nullptr;
  // lagged coarsest inverse (coarse_lag): the cycle applies the inverse of
  // the previous refresh's coarsest operator while the current one is being
  // inverted into the spare buffers on side stream 2
  double *d_cinv2 = nullptr;
  float *d_cinvf2 = nullptr;
  // UFT_MIXED_H (one design, one load case): the cycle's copy of the inverse
  // in FP16 scaled by row (entries / row max) with the FP32 row scales, and
  // the spare pair of the lagged inverse
  uint16_t *d_cinvh = nullptr, *d_cinvh2 = nullptr;
  float *d_cinvs = nullptr, *d_cinvs2 = nullptr;
  bool lag_spare = false;    // the spare buffers hold the inverse of the last refresh (swap at the next)
  bool lag_pending = false;  // a spare inverse is being computed on side[2] (event ev_inv)
};

// Multi-load-case form: nr right-hand sides solved together, vectors stored
// interleaved [unknown][nr] (kernels_mrhs.hip).
struct MrLevelBuf {
  void *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr, *t = nullptr;
  void *kc = nullptr, *kv = nullptr, *kd = nullptr, *kr = nullptr;
};
struct MrState {
  int nr = 0;      // right-hand sides per block (2, 4 or 8); 0: single right-hand side
  int ncase = 0;   // load cases in use (<= nr; the rest are zero)
  double *F = nullptr, *U = nullptr;                                // ndof x nr
  double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *zo = nullptr;
  float *rf = nullptr, *zf = nullptr;                               // mixed: FP32 copies for the cycle
  float *pf = nullptr, *qf = nullptr;                               // mixed: FP32 outer product
  MrLevelBuf L[UFT_MAXLEV];                                         // cycle work (FP64 or FP32)
  void *Ye = nullptr;                                               // nd x ne x nr element-local scratch
  double *sc = nullptr;                                             // per right-hand-side scalars
  double relres[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int it[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Rows of one vector space of a design batch and its segment grid.
struct SegSpace {
  int64_t nrows = 0;
  std::vector<int64_t> off;     // nseg + 1 row offsets
  std::vector<int32_t> dfirst;  // nseg + 1: first block of each design
  int64_t *d_span = nullptr;
  int32_t *d_des = nullptr;
  int32_t *d_dfirst = nullptr;
  int nb = 0;
  SegG g() const {
    SegG r;
    r.span = d_span;
    r.des = d_des;
    r.nb = nb;
    return r;
  }
};

// Independent designs in one context (uft_batch_set).
struct BatchState {
  int nseg = 0;                       // 0: one design
  std::vector<int64_t> eoff, noff;    // element / node offsets
  std::vector<double> volfrac, rmin, TV;
  SegSpace E, N, D;                   // elements, nodes, DOFs
  SegSpace R[UFT_MAXLEV];             // rows of multigrid level l >= 1
  std::vector<int64_t> uoff[UFT_MAXLEV];  // unit offsets per level (host)
  std::vector<int64_t> coff;          // coarsest dense blocks: offsets of the n_d x n_d blocks
  int32_t *d_cu = nullptr;            // coarsest unit offsets (nseg + 1)
  int64_t *d_coff = nullptr;
  int32_t *d_cdes = nullptr;          // coarsest units: design
  double *sc = nullptr;               // nseg x UFT_BSC device scalars
  double *h = nullptr;                // pinned host mirror of sc
  double *d_vf = nullptr;             // per design volfrac
  double *d_invTV = nullptr;          // per design 1 / total volume
  double *d_lam = nullptr;            // per design MMA dual variable
  double *d_gt = nullptr;             // per design gx and the dual solve's tolerance
  int32_t *d_edes = nullptr, *d_ndes = nullptr;  // design of every element / node
  double *d_fac = nullptr;            // per design Chebyshev scaling 1 / lambda_max
  double *d_log = nullptr;            // Lanczos log: nseg x 64 steps x 2
  // per design results of the last iteration
  std::vector<double> fx, gx, ch, relres;
  std::vector<int> it, itf;
  // per design MMA dual state (m = 1)
  std::vector<double> lam, mu, b;
};

// host-side description produced by uft_amg_setup.cpp
struct AmgHostLevel {
  int64_t nunit_next = 0;
  std::vector<int32_t> parent;       // per unit of this level
  std::vector<double> P;             // per unit: bs*6 (row-major bs x 6)
  std::vector<int32_t> cptr, child;  // next-level unit -> children of this level
  // pattern of the next level operator
  std::vector<int32_t> rp, ci;
  // Galerkin lists for the next level blocks (level 0: per upper block, the
  // elements of its entries, the entries' block and the transposed block)
  std::vector<int32_t> glptr, gl, gblk, tr;
  // level 0 with a smoothed prolongator (sa): per node the aggregates of its
  // prolongator row (its own and its neighbours', sorted), per node the
  // aggregates of its row of A P (the union over its neighbours' rows), and
  // per aggregate its support: the (node, prolongator entry) pairs naming it,
  // in node order
  bool sa = false;
  std::vector<int32_t> sa_pptr, sa_pagg, sa_vptr, sa_vagg, sa_sptr, sa_snode, sa_sidx;
};

// --------------------------------------------------------------------------
// Context
// --------------------------------------------------------------------------
struct uft_ctx {
  std::vector<void *> allocs;  // every device allocation, freed by uft_destroy
  int device = 0;
  hipStream_t s = nullptr;
  std::string err;
  uft_params P{};
  bool have_mesh = false, have_bc = false, is_setup = false;

  // mesh (host copies)
  int npe = 0, nd = 0;
  int64_t nn = 0, ne = 0, ndof = 0;
  std::vector<int32_t> elem;  // ne x npe
  std::vector<double> xyz;    // nn x 3
  std::vector<int32_t> fixed;
  int nlc = 0, nload = 0;
  std::vector<int32_t> ldofs;
  std::vector<double> lvals;

  // device mesh / element operators
  int32_t *d_elem = nullptr;   // [npe][ne]
  double *d_xyz = nullptr;     // nn x 3
  double *d_KE = nullptr;      // packed upper triangle, pair-interleaved [np/2][ne][2]
  float *d_KEf = nullptr;      // FP32 copy for the mixed form [np/2][ne][2]
  uint4 *d_KEh = nullptr;      // FP16 smoother copy, KE/s_e in groups of eight [ng][ne]
  uint32_t *d_KEb = nullptr;   // BF16 full rows for the multi-load-case matrix-core products [ne][nd][ndp]
  float *d_KEs = nullptr;      // per element scale s_e = max |KE_e|
  double *d_KF = nullptr;      // Helmholtz packed [npf/2][ne][2]
  double *d_T = nullptr;       // [npe][ne]
  double *d_detmin = nullptr;  // per element minimum det J
  uint8_t *d_fixm = nullptr;   // per node: bit c set if DOF 3n+c is fixed
  int32_t *d_iptr = nullptr, *d_inc = nullptr;  // node -> (e*npe + local)
  double *d_F = nullptr;
  double *d_Ye = nullptr;      // element-local scratch nd x ne
  double *d_Ee = nullptr;      // per element E(x_e) of the last hierarchy refresh
  float *d_Yef = nullptr;
  double TV = 0.0;
  // assembled node operators (fine_assembled) on one sliced-ELL node pattern:
  // the filter's Helmholtz operator (FP64) and, in UFT_MIXED_H, the
  // smoother's FP16 3x3 blocks (rows scaled by the diagonal) and the outer
  // product's FP32 blocks
  struct {
    int64_t nslice = 0, nslot = 0, nnzb = 0;
    int wmax = 0;
    int64_t *d_off = nullptr;    // per slice: first slot (nslice + 1)
    int32_t *d_width = nullptr;  // per slice: slots per node
    int32_t *d_col = nullptr;    // per slot: column node
    int16_t *d_cd = nullptr;     // per slot: column node - row node, when every one fits 16 bits
    int32_t *d_cptr = nullptr;   // per slot: first element contribution (nslot + 1)
    int32_t *d_clist = nullptr;  // element contributions (e * npe + a) * npe + b
    uint4 *d_v8 = nullptr;       // per slot: block entries 0..7 / diag(row), FP16 (multi-load-case form)
    uint16_t *d_v1 = nullptr;    // per slot: entry 8
    uint2 *d_q8 = nullptr;       // per slot: entries 0..7 / diag(row) as int8 times the block scale
    uint32_t *d_qs = nullptr;    // per slot: FP32 block scale, entry 8 (int8) in its low byte
    float *d_dg = nullptr;       // per scalar row: the diagonal (the row scale)
    double *d_hv = nullptr;      // per slot: the filter's Helmholtz entry (FP64, assembled at setup)
    float4 *d_w8 = nullptr;      // per slot: FP32 block entries 0..7 (the outer operator)
    float *d_w1 = nullptr;       // per slot: FP32 entry 8
    double2 *d_e8 = nullptr;     // FP64 form (one load case): FP64 entries 2q, 2q+1 of slot k at [q nslot + k]
    double *d_e1 = nullptr;      // ... and entry 8 (the fine operator of the solves)
    // node tiles (several load cases): the same blocks in tile order, see
    // NodeTiles; the assembly writes them beside the node-ordered ones
    int64_t ntile = 0, tslot = 0;
    int thmax = 0;
    int32_t *d_tnode = nullptr, *d_thoff = nullptr, *d_thnode = nullptr, *d_twidth = nullptr, *d_kmap = nullptr;
    int64_t *d_toff = nullptr;
    uint16_t *d_tlc = nullptr;
    uint4 *d_tv8 = nullptr;
    uint16_t *d_tv1 = nullptr;
    float4 *d_tw8 = nullptr;
    float *d_tw1 = nullptr;
  } fell;

  // vectors
  std::vector<double *> vec;   // by id
  std::vector<int64_t> vlen;

  // elasticity solver work (ndof)
  double *d_r = nullptr, *d_z = nullptr, *d_p = nullptr, *d_q = nullptr;
  double *d_w = nullptr;       // mixed form: FP64 correction accumulator
  float *d_rf = nullptr, *d_zf = nullptr, *d_pf = nullptr, *d_qf = nullptr, *d_uf = nullptr;
  AmgHier amg;
  bool amg_built = false;
  MrState mr;

  // filter work (nn)
  double *d_fr = nullptr, *d_fz = nullptr, *d_fp = nullptr, *d_fq = nullptr, *d_fdinv = nullptr;
  int f_pred[UFT_MAXM + 2] = {};  // last iteration counts of the filter solves (queue sizes)
  double *d_TX = nullptr;

  // reductions
  double *d_part = nullptr;    // partials
  int64_t part_cap = 0;        // doubles in d_part
  double *d_scal = nullptr;    // device scalars
  double *h_scal = nullptr;    // pinned host mirror

  // heaviside state (PDEFilter.h:50-51)
  double beta = 1.0, btemp = 1.0;

  // MMA state (MMA.h:76-109)
  int mk = 0;
  double asyminit = 0.5, asymdec = 0.7, asyminc = 1.2;
  int mma_robust = 0, mma_conmod = 0;  // RobustAsymptotesType, constraintModification
  double a[UFT_MAXM], c[UFT_MAXM], dd[UFT_MAXM], y[UFT_MAXM], z = 0.0;
  double lam[UFT_MAXM], mu[UFT_MAXM], sv[2 * UFT_MAXM], b[UFT_MAXM], grad[UFT_MAXM],
      Hess[UFT_MAXM * UFT_MAXM];
  double *d_L = nullptr, *d_U = nullptr, *d_alpha = nullptr, *d_beta = nullptr, *d_p0 = nullptr,
         *d_q0 = nullptr, *d_pij = nullptr, *d_qij = nullptr, *d_xo1 = nullptr, *d_xo2 = nullptr;

  const double *cur_xphys = nullptr;  // density of the operator being solved
  double fnorm = 0.0;
  double *d_lanczos = nullptr;       // CG scalar log for eigenvalue estimates

  // fine operator 3x3 node blocks on the node pattern for the smoothed
  // prolongator: FP64 in the FP64 form, FP32 in the mixed forms
  double *d_a64 = nullptr;
  float *d_a32 = nullptr;

  // uft_solve_traffic: algorithmic bytes of the operator applies made inside
  // the elasticity solves (hierarchy refresh excluded) and the solves' time
  bool count_on = false;
  double solve_bytes = 0.0, solve_ms = 0.0;
  int64_t solve_applies = 0, solve_count = 0;

  // device timers (uft_profile_*): event triples around fine applies
  bool prof = false;
  int prof_every = 1;   // time one in every prof_every fine applies of each kind
  int64_t prof_tick[3] = {0, 0, 0};
  std::vector<hipEvent_t> evpool;
  // the hierarchy refresh's concurrent branches (fine blocks + fine-level
  // eigenvalue estimate, coarse-level estimates, coarsest inverse) and their
  // fork / join events; the coarse estimates' block partials
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_inv = nullptr;  // the lagged coarsest inverse done on side[2]
  hipEvent_t ev_side[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double *d_part2 = nullptr;
  std::vector<int> evkind;  // per triple: 0 FP64 apply, 1 FP32 apply
  size_t evused = 0;
  double prof_ms[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int64_t prof_n[6] = {0, 0, 0, 0, 0, 0};

  // wall time of the phases of the last design iteration (ms): 0 hierarchy
  // refresh, 1 elasticity Krylov solve, 2 sensitivities, 3 sensitivity
  // filter, 4 MMA, 5 density filter + projection, 6 whole iteration
  double t_phase[8] = {0, 0, 0, 0, 0, 0, 0, 0};

  // element-phase kernel variants chosen by uft_autotune
  int elem_variant = 0, elem_variant_f = 0, elem_variant_h = 0;
  double tune_us[16] = {0}, tune_us_f[16] = {0}, tune_us_h[16] = {0};

  BatchState bt;

  // hipGraph of one Krylov iteration of the elasticity solve (uft_params.graph)
  hipGraph_t ig_graph = nullptr;
  hipGraphExec_t ig_exec = nullptr;
  const void *ig_key[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t ig_refresh = -1;  // hierarchy refresh the graph was captured after
  int64_t refreshes = 0;

  // stats
  int it_el = 0, it_f = 0, outer_el = 0, dip_steps = 0, reliable = 0;
  int el_prev_its = 0;  // Krylov iterations of the last device-controlled elasticity solve (round sizes)
  double relres_el = 0.0;
  double fx = 0.0, gx[UFT_MAXM];
};

// --------------------------------------------------------------------------
// Kernel launch wrappers (kernels_*.hip)
// --------------------------------------------------------------------------
namespace uftk {
int red_blocks(int64_t n);
// element operators
void gen_ke(hipStream_t s, int npe, int64_t ne, const int32_t *elem, const double *xyz, double nu,
            double *KE, float *KEf, double *detmin);
void gen_kf(hipStream_t s, int npe, int64_t ne, const int32_t *elem, const double *xyz, double rmin,
            double *KF, double *T, double *vol, double *part, SegG sg = SegG(), const double *rmind = nullptr);
// fine operator apply: y = K_D(x) u   (two-phase element/node)
// dotw != nullptr: also y.dotw, reduced in the same launch into *dot_out
void apply_K(hipStream_t s, const uft_ctx *C, const double *xphys, const double *u, double *y,
             const double *dotw, double *dot_out, hipEvent_t *ev = nullptr);
void apply_Kf(hipStream_t s, const uft_ctx *C, const double *xphys, const float *u, float *y,
              hipEvent_t *ev = nullptr);
int elem_variant_count();
void elem_phase(hipStream_t s, const uft_ctx *C, int variant, const double *xphys, const double *u, double *Ye);
void elem_phase_f(hipStream_t s, const uft_ctx *C, int variant, const double *xphys, const float *u, float *Ye);
void elem_phase_h(hipStream_t s, const uft_ctx *C, int variant, const double *xphys, const float *u, float *Ye);
void ke_to_half(hipStream_t s, const uft_ctx *C);
int ke_half_groups(int npe);
void diag_K(hipStream_t s, const uft_ctx *C, const double *xphys, double *dinv, float *dinvf);
void sens(hipStream_t s, const uft_ctx *C, const double *xphys, const double *u, double *dfdx,
          double *dgdx0, double *part, SegG sg = SegG(), const double *invTVd = nullptr);
// Helmholtz filter
// y = KF x; with dot_out also x.y (reduced in fixed order into *dot_out)
void apply_KF(hipStream_t s, const uft_ctx *C, const double *x, double *y, double *dot_out = nullptr);
void diag_KF(hipStream_t s, const uft_ctx *C, double *dinv);
void tf_mult(hipStream_t s, const uft_ctx *C, const double *xe, double *TX);
void tfx_t(hipStream_t s, const uft_ctx *C, const double *RX, double *xe, int clamp, double Xmin);
// reductions / vector ops
void final_sum(hipStream_t s, const double *part, int nb, int K, double *out);
void final_max(hipStream_t s, const double *part, int nb, double *out);
void dot(hipStream_t s, int64_t n, const double *a, const double *b, double *part, double *out);
void dot2(hipStream_t s, int64_t n, const double *a, const double *b, const double *c, double *part,
          double *out);  // out[0]=a.b, out[1]=c.c
void scalar_div(hipStream_t s, double *sc, int num, int den, int dst);
void fill(hipStream_t s, int64_t n, double v, double *x);
void fillf(hipStream_t s, int64_t n, float v, float *x);
void copy(hipStream_t s, int64_t n, const double *a, double *b);
void d2f(hipStream_t s, int64_t n, const double *a, float *b);
void f2d_add(hipStream_t s, int64_t n, const float *a, double *b);
void scale(hipStream_t s, int64_t n, const double *a, const double *dinv, double *z);
void cg_update(hipStream_t s, int64_t n, const double *sc, int ialpha, double *u, const double *p,
               double *r, const double *q, double *part, double *rr_out);
void xpby(hipStream_t s, int64_t n, const double *z, const double *sc, int ibeta, double *p);
// the same, also writing pf = (float)p
void xpby_lo(hipStream_t s, int64_t n, const double *z, const double *sc, int ibeta, double *p, float *pf);
void axpy_inplace(hipStream_t s, int64_t n, double a, const double *x, double *y);
void sub(hipStream_t s, int64_t n, const double *a, const double *b, double *c);
void set_dofs(hipStream_t s, int n, const int32_t *idx, const double *val, double *x);
// templated float variants for the mixed form
// u += a p ; r -= a q ; rf = (float) r with a = sc[ia] / sc[ib]; r.r into *out
void cg_update_lo(hipStream_t s, int64_t n, const double *sc, int ia, int ib, double *u, const double *p, double *r,
                  const float *q, float *rf, double *part, double *out);
// the same, also forming the next fine cycle's first Chebyshev step on rf
// (zero guess): cr = rf, cd = cx = it dinvf rf
void cg_update_lo_cheb(hipStream_t s, int64_t n, const double *sc, int ia, int ib, double *u, const double *p,
                       double *r, const float *q, float *rf, double *part, double *out, const float *dinvf, double it,
                       float *cr, float *cd, float *cx, const double *stop = nullptr);
// device control of the single-design flexible CG of the mixed forms (ctl:
// stop, reliable-update request, iterations, reliable updates, last FP64
// residual norm, tolerance, iteration limit): initial state; the test after
// an iteration's update; the reliable update itself (FP64 K U through the
// element phase, r = F - K U, its FP32 copy and the next cycle's first
// Chebyshev step, r.r into rr_out and the stopping test), every launch of
// which returns at once unless ctl asked for it
void ksp_ctl_init(hipStream_t s, double *ctl, double rtrue, double tol, int maxit);
void ksp_ctl_iter(hipStream_t s, const double *rr, double *ctl);
void reliable_update(hipStream_t s, const uft_ctx *C, const double *xphys, const double *U, double *r, float *rf,
                     const float *dinvf, double it, float *cr, float *cd, float *cx, double *rr_out, double *ctl);
// zo = z ; z = (double) zf ; r.z, r.zo into out[0..1]
void z_from_f32(hipStream_t s, int64_t n, const float *zf, const double *r, double *z, double *zo, double *part,
                double *out);
void cg_updatef(hipStream_t s, int64_t n, const double *sc, int ialpha, float *u, const float *p,
                float *r, const float *q, double *part, double *rr_out);
void xpbyf(hipStream_t s, int64_t n, const float *z, const double *sc, int ibeta, float *p);
void dotf(hipStream_t s, int64_t n, const float *a, const float *b, double *part, double *out);
// multigrid
void cheb_first(hipStream_t s, int64_t n, const double *b, const double *dinv, double inv_theta,
                double *r, double *d, double *x, int x_is_zero);
void cheb_step(hipStream_t s, int64_t n, const double *t, const double *dinv, double c1, double c2,
               double *r, double *d, double *x);
void cheb_firstf(hipStream_t s, int64_t n, const float *b, const float *dinv, double inv_theta, float *r,
                 float *d, float *x, int x_is_zero);
void cheb_stepf(hipStream_t s, int64_t n, const float *t, const float *dinv, double c1, double c2,
                float *r, float *d, float *x);
void sub_inplace(hipStream_t s, int64_t n, const double *t, double *r);
void sub_inplacef(hipStream_t s, int64_t n, const float *t, float *r);
void subf(hipStream_t s, int64_t n, const float *a, const float *b, float *c);
void bsr6_mv(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val,
             const double *x, double *y);
void bsr6_mvf(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const float *val,
              const float *x, float *y);
void bsr6_dinv(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, double *val,
               double *dinv, float *dinvf);
void restrict_(hipStream_t s, int bs, int64_t ncoarse_units, const int32_t *cptr, const int32_t *child,
               const double *P, const double *r, double *bc);
void prolong_add(hipStream_t s, int bs, int64_t nunits, const int32_t *parent, const double *P,
                 const double *xc, double *x);
void restrictf(hipStream_t s, int bs, int64_t ncoarse_units, const int32_t *cptr, const int32_t *child,
               const float *P, const float *r, float *bc);
void prolong_addf(hipStream_t s, int bs, int64_t nunits, const int32_t *parent, const float *P,
                  const float *xc, float *x);
// Ee[e] = E(x_e) (SIMP modulus)
void simp_moduli(hipStream_t s, const uft_ctx *C, const double *xphys, double *Ee);
// level 0 -> 1: G (ngl x 36) once per mesh; then val = sum_k E(x_e(k)) G_k per block
// (galerkin_fine writes E(x_e) to Ee first, or with xphys null reads it there)
void galerkin_pre(hipStream_t s, const uft_ctx *C, int64_t ngl, const int32_t *gblk, const int32_t *gl,
                  const int32_t *brow, const int32_t *ci, const int32_t *parent, const double *Q, double *G);
void galerkin_fine(hipStream_t s, const uft_ctx *C, const double *xphys, int64_t nblk, const int32_t *glptr,
                   const int32_t *gl, const int32_t *gtr, const int32_t *brow, const int32_t *ci,
                   const double *G, double *Ee, double *val);
void galerkin_coarse(hipStream_t s, int64_t nblk, const int32_t *glptr, const int32_t *gl,
                     const int32_t *brow, const int32_t *ci, const double *P, const double *val,
                     double *valc);
void bsr_to_dense(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val,
                  int64_t n, double *dense);
void dense_spd_inverse(hipStream_t s, int64_t n, double *A, double *Ainv, float *Ainvf);
void dense_mv(hipStream_t s, int64_t n, const double *A, const double *x, double *y);
// blocked symmetric Gauss-Jordan sweep on 64x64 tiles (kernels_dense.hip):
// the inverse of the SPD row-major n x n matrix A into Ainv (FP32 copy into
// Ainvf when given), S = dense_sweep_scratch(n) doubles of scratch
int64_t dense_sweep_scratch(int64_t n);
void dense_spd_sweep(hipStream_t s, int64_t n, const double *A, double *S, double *Ainv, float *Ainvf);
void dense_spd_sweep_bsr(hipStream_t s, int64_t n, int64_t nblk, const int32_t *brow, const int32_t *ci,
                         const double *val, double *S, double *Ainv, float *Ainvf);
// y = A x, row-major FP32 n x n, n a multiple of 4 and 16-byte aligned operands
void dense_mvf4(hipStream_t s, int64_t n, const float *A, const float *x, float *y);
// row-scaled FP16 copy of an FP32 matrix (entries / row max, FP32 row scales)
// and its product (n a multiple of 8)
void dense_to_half_rows(hipStream_t s, int64_t n, const float *A, uint16_t *H, float *sc);
void dense_mvh(hipStream_t s, int64_t n, const uint16_t *H, const float *sc, const float *x, float *y);
void dense_mvf(hipStream_t s, int64_t n, const float *A, const float *x, float *y);
// multi-load-case kernels (kernels_mrhs.hip)
void mr_cheb_first(hipStream_t s, int nr, int64_t n, const double *b, const double *dinv, double it, double *r,
                   double *d, double *x);
void mr_cheb_first(hipStream_t s, int nr, int64_t n, const float *b, const float *dinv, double it, float *r, float *d,
                   float *x);
void mr_elem_phase(hipStream_t s, const uft_ctx *C, int nr, const double *xphys, const double *u, double *Ye);
// FP32 vectors: BF16 matrix-core products (8 cases, mixed forms) or, with blocks32, the FP32 element blocks
void mr_elem_phase(hipStream_t s, const uft_ctx *C, int nr, const double *xphys, const float *u, float *Ye,
                   bool blocks32 = false);
void ke_to_bf16(hipStream_t s, const uft_ctx *C);
int ke_bf16_rows_words(int npe);
void mr_gather_op(hipStream_t s, const uft_ctx *C, int nr, const double *u, const double *Ye, const RowOp<double> &o);
void mr_gather_op(hipStream_t s, const uft_ctx *C, int nr, const float *u, const float *Ye, const RowOp<float> &o);
// the assembled fine operator (FP16 smoother blocks, or FP32 with fp32) on nr
// interleaved right-hand sides, with the row epilogue (mr_gather_blocks blocks)
void mr_ell_op(hipStream_t s, const uft_ctx *C, int nr, const float *u, const RowOp<float> &o, bool fp32);
int mr_ell_blocks(const uft_ctx *C, const float *u, int mode, bool fp32);
int mr_gather_blocks(int64_t nn, int nr);
void mr_bsr_op(hipStream_t s, int nr, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val,
               const double *u, const RowOp<double> &o);
void mr_bsr_op(hipStream_t s, int nr, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const float *val,
               const float *u, const RowOp<float> &o);
// the level's FP16 blocks of D^-1/2 A D^-1/2 with dh = D^1/2 (UFT_MIXED_H)
void mr_bsr_op_h(hipStream_t s, int nr, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const uint16_t *valh,
                 const float *dh, const float *u, const RowOp<float> &o);
// the FP64 form's load cases on the assembled FP64 node blocks (fell.d_e8)
void mr_ell_op64(hipStream_t s, const uft_ctx *C, int nr, const double *u, const RowOp<double> &o);
int mr_bsr_blocks(int64_t nb_rows);
void mr_restrict_first(hipStream_t s, int nr, int bs, int64_t nc, const int32_t *cptr, const int32_t *child,
                       const double *P, const double *r, double *bc, const double *dinv, double it, double *rc,
                       double *dc, double *xc);
void mr_restrict_first(hipStream_t s, int nr, int bs, int64_t nc, const int32_t *cptr, const int32_t *child,
                       const float *P, const float *r, float *bc, const float *dinv, double it, float *rc, float *dc,
                       float *xc);
void mr_prolong_comb(hipStream_t s, int nr, int bs, int64_t nu, const int32_t *parent, const double *P,
                     const double *sc, int ia, int ib, const double *xa, const double *xb, double *x);
void mr_prolong_comb(hipStream_t s, int nr, int bs, int64_t nu, const int32_t *parent, const float *P, const double *sc,
                     int ia, int ib, const float *xa, const float *xb, float *x);
void mr_lincomb_first(hipStream_t s, int nr, int64_t n, const double *sc, const double *b, const double *v, double *kr,
                      double *r, const double *dinv, double it, double *d, double *kd);
void mr_lincomb_first(hipStream_t s, int nr, int64_t n, const double *sc, const float *b, const float *v, float *kr,
                      float *r, const float *dinv, double it, float *d, float *kd);
// (part: room for 3 x 8 x 128 more doubles after the nb partial rows)
void mr_final_kc1(hipStream_t s, int nr, double *part, int nb, double *sc);
void mr_final_kc2(hipStream_t s, int nr, double *part, int nb, double *sc);
// xc = Ainv P^T r on the coarsest level (bc: restriction scratch of the coarsest size x nr)
void mr_tail_solve(hipStream_t s, int nr, int bs, int nc_units, const int32_t *cptr, const int32_t *child,
                   const double *P, const double *Ainv, double *bc, const double *r, double *xc);
void mr_tail_solve(hipStream_t s, int nr, int bs, int nc_units, const int32_t *cptr, const int32_t *child,
                   const float *P, const float *Ainv, float *bc, const float *r, float *xc);
void mr_dot(hipStream_t s, int nr, int64_t n, const double *a, const double *b, double *part, double *out);
void mr_dot_pr(hipStream_t s, int nr, int64_t n, const double *r, const double *z, const double *zo, double *part,
               double *out);
void mr_cg_update(hipStream_t s, int nr, int64_t n, const double *alpha, double *x, const double *p, double *r,
                  const double *q, double *part, double *rr_out);
void mr_xpby(hipStream_t s, int nr, int64_t n, const double *z, const double *beta, double *p);
void mr_cg_update_lo(hipStream_t s, int nr, int64_t n, const double *alpha, double *x, const double *p, double *r,
                     const float *q, float *rf, double *part, double *rr_out);
// r_k = res_k (and rf_k) for the right-hand sides k in mask; r.r of every one into rr_out
void mr_take_res(hipStream_t s, int nr, int64_t n, unsigned mask, const double *res, double *r, float *rf,
                 double *part, double *rr_out);
void mr_xpby_lo(hipStream_t s, int nr, int64_t n, const double *z, const double *beta, double *p, float *pf);
void mr_alpha(hipStream_t s, int nr, unsigned mask, const double *rz, const double *pq, double *alpha);
void mr_beta(hipStream_t s, int nr, unsigned mask, const double *rzn, double *rzprev, double *beta);
void mr_sens(hipStream_t s, const uft_ctx *C, int nr, const double *xphys, const double *u, double *dfdx,
             double *part);
int mr_sens_blocks(int64_t ne, int nr);
void mr_extract(hipStream_t s, int64_t n, int nr, int k, const double *src, double *dst);
void mr_set_dofs(hipStream_t s, int n, int nr, int k, const int32_t *idx, const double *val, double *F);
// fused cycle kernels (kernels_cycle.hip)
void bsr6_op(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val, const double *u,
             const RowOp<double> &op);
void bsr6_op(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const float *val, const float *u,
             const RowOp<float> &op);
int bsr6_op_blocks(int64_t nb_rows);
void bsr6_op_h(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const uint16_t *valh,
               const float *dh, const float *u, const RowOp<float> &op);
int bsr6_op_h_blocks(int64_t nb_rows);
void bsr6_to_half(hipStream_t s, int64_t nblk, int64_t nb_rows, const int32_t *brow, const int32_t *ci,
                  const double *val, const double *dinv, uint16_t *valh, float *dh);
void gather_op(hipStream_t s, const uft_ctx *C, const double *u, const double *Ye, const RowOp<double> &op);
void gather_op(hipStream_t s, const uft_ctx *C, const float *u, const float *Ye, const RowOp<float> &op);
// assembled FP16 fine operator: assembly from the per-element moduli E(x_e)
// (rows scaled by the level-0 diagonal 1 / dinv) and the apply with the row
// epilogue (block partials over gather_op_blocks(nn) blocks)
void fine_ell_fill(hipStream_t s, const uft_ctx *C, const double *Ee, const double *dinv, double *a64 = nullptr,
                   float *a32 = nullptr);
// smoothed prolongator of level 0 and the level-1 operator P^T A P from the
// fine blocks (d_a64 or d_a32), the fine diagonal and L[0].sa_omega
// (kernels_sa.hip)
void sa_refresh(hipStream_t s, uft_ctx *C);
// setup: the contribution lists (which 0: prolongator slots, 1: A P pairs,
// 2: Galerkin pairs); count writes the list length per row, fill the lists
// at the row offsets ptr
void sa_lists(hipStream_t s, const uft_ctx *C, int which, const int32_t *ptr, int32_t *cnt_or_a, int32_t *b);
void sa_restrict_first(hipStream_t s, const AmgLevel &L0, int64_t nc, const double *r, double *bc, const double *dinv,
                       double it, double *rc, double *dc, double *xc);
void sa_restrict_first(hipStream_t s, const AmgLevel &L0, int64_t nc, const float *r, float *bc, const float *dinv,
                       double it, float *rc, float *dc, float *xc);
void sa_prolong_comb(hipStream_t s, const AmgLevel &L0, const double *sc, int ia, int ib, const double *xa,
                     const double *xb, double *x, const int32_t *ndes = nullptr);
void sa_prolong_comb(hipStream_t s, const AmgLevel &L0, const double *sc, int ia, int ib, const float *xa,
                     const float *xb, float *x, const int32_t *ndes = nullptr);
// the same transfers for nr interleaved right-hand sides (the multi-load-case
// form; per right-hand-side coefficients at sc[8 q + ia])
void mr_sa_restrict_first(hipStream_t s, int nr, const AmgLevel &L0, int64_t nc, const double *r, double *bc,
                          const double *dinv, double it, double *rc, double *dc, double *xc);
void mr_sa_restrict_first(hipStream_t s, int nr, const AmgLevel &L0, int64_t nc, const float *r, float *bc,
                          const float *dinv, double it, float *rc, float *dc, float *xc);
// (ncu: units of level 1; with xb the combination goes through tmp, a free
// level-1 vector of nr right-hand sides)
void mr_sa_prolong_comb(hipStream_t s, int nr, const AmgLevel &L0, int64_t ncu, const double *sc, int ia, int ib,
                        const double *xa, const double *xb, double *x, double *tmp);
void mr_sa_prolong_comb(hipStream_t s, int nr, const AmgLevel &L0, int64_t ncu, const double *sc, int ia, int ib,
                        const float *xa, const float *xb, float *x, float *tmp);
void fine_ell_op(hipStream_t s, const uft_ctx *C, const float *u, const RowOp<float> &op);
// the same pattern with FP32 entries (unscaled): the mixed forms' outer operator
void fine_ell_op32(hipStream_t s, const uft_ctx *C, const float *u, const RowOp<float> &op);
// the FP64 form's fine operator from the assembled FP64 node blocks of the
// last hierarchy refresh (fell.d_e8 / d_e1)
void fine_ell_op64(hipStream_t s, const uft_ctx *C, const double *u, const RowOp<double> &op);
// the filter's Helmholtz operator on the node pattern: assembled once (KF is
// density-free), then y = KF x (with dot_out: x.y into it via d_part)
void helm_ell_fill(hipStream_t s, const uft_ctx *C);
void helm_ell_apply(hipStream_t s, const uft_ctx *C, const double *x, double *y, double *dot_out);
// filter CG step on the assembled pattern: p = z + (sc[ib] / sc[ibo]) p (z
// when first), y = KF p, *dot_out = p.y; nothing once *stop is set
void helm_ell_cg(hipStream_t s, const uft_ctx *C, const double *z, const double *sc, int ib, int ibo, int first,
                 const double *stop, double *p, double *y, double *dot_out);
int gather_op_blocks(int64_t nn);
void restrict_first(hipStream_t s, int bs, int64_t nc, const int32_t *cptr, const int32_t *child, const double *P,
                    const double *r, double *bc, const double *dinv, double it, double *rc, double *dc, double *xc);
void restrict_first(hipStream_t s, int bs, int64_t nc, const int32_t *cptr, const int32_t *child, const float *P,
                    const float *r, float *bc, const float *dinv, double it, float *rc, float *dc, float *xc);
void prolong_comb(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const double *P, const double *sc, int ia,
                  int ib, const double *xa, const double *xb, double *x);
void prolong_comb(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const float *P, const double *sc, int ia,
                  int ib, const float *xa, const float *xb, float *x);
void lincomb_first(hipStream_t s, int64_t n, const double *sc, const double *b, const double *v, double *kr, double *r,
                   const double *dinv, double it, double *d, double *kd);
void lincomb_first(hipStream_t s, int64_t n, const double *sc, const float *b, const float *v, float *kr, float *r,
                   const float *dinv, double it, float *d, float *kd);
void final_kc1(hipStream_t s, const double *part, int nb, double *sc);
void final_kc2(hipStream_t s, const double *part, int nb, double *sc);
// final_kc1 + lincomb_first in one launch (nb partial rows summed by every block)
void lincomb_first_kc(hipStream_t s, int64_t n, const double *part, int nb, double *sc, const double *b,
                      const double *v, double *kr, double *r, const double *dinv, double it, double *d, double *kd);
void lincomb_first_kc(hipStream_t s, int64_t n, const double *part, int nb, double *sc, const float *b,
                      const float *v, float *kr, float *r, const float *dinv, double it, float *d, float *kd);
// final_kc2 + the K-cycle prolongation x += P (c5 xa + c6 xb) in one launch
void prolong_kc(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const double *P, const double *part, int nb,
                double *sc, const double *xa, const double *xb, double *x);
void prolong_kc(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const float *P, const double *part, int nb,
                double *sc, const float *xa, const float *xb, float *x);
// xc = Ainv P^T r on the coarsest level (one workgroup; ncoarse <= UFT_TAIL_MAX)
// With a design batch (nseg > 1): one workgroup per design, units
// [cu[d], cu[d+1]) and the inverse block at Ainv + coff[d].
void tail_solve(hipStream_t s, int nc_units, const int32_t *cptr, const int32_t *child, const double *P,
                const double *Ainv, const double *r, double *xc, int nseg = 1, const int32_t *cu = nullptr,
                const int64_t *coff = nullptr);
void tail_solve(hipStream_t s, int nc_units, const int32_t *cptr, const int32_t *child, const float *P,
                const double *Ainv, const float *r, float *xc, int nseg = 1, const int32_t *cu = nullptr,
                const int64_t *coff = nullptr);
// One design: the same (FP64 inverse in the FP64 cycle, its FP32 copy in the
// FP32 cycle), then x_u += P_u xc_parent(u) over the nu_up units of the level
// above in the same workgroup (the prolongation of coarse_direct).
void tail_solve_prolong(hipStream_t s, int nc_units, const int32_t *cptr, const int32_t *child, const double *P,
                        const double *Ainv, const double *r, double *xc, int64_t nu_up, const int32_t *parent,
                        double *x);
void tail_solve_prolong(hipStream_t s, int nc_units, const int32_t *cptr, const int32_t *child, const float *P,
                        const float *Ainv, const float *r, float *xc, int64_t nu_up, const int32_t *parent,
                        float *x);
void fixed_unit_diag(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, double *val);
// MMA
void mma_movelimit(hipStream_t s, int64_t n, const double *x, double Xmin, double Xmax, double movlim,
                   double *xmin, double *xmax);
void mma_gensub(hipStream_t s, uft_ctx *C, int64_t n, int m, int k, const double *x, const double *xmin,
                const double *xmax, const double *dfdx, const double *const *dgdx, double *part, SegG sg = SegG());
// lamd (device, per design of the segment grid) replaces lam when set (m = 1)
void mma_dual_pass(hipStream_t s, uft_ctx *C, int64_t n, int m, const double *lam, double *x, double *part,
                   SegG sg = SegG(), const double *lamd = nullptr, const double *gate = nullptr);
// out[0] = sum of squares, out[1] = max abs of the KKT residual terms
void mma_kkt(hipStream_t s, int64_t n, int m, const double *lam, const double *x, const double *dfdx,
             const double *const *dgdx, const double *xmin, const double *xmax, double *part, double *out);
void stress(hipStream_t s, const uft_ctx *C, const double *xphys, const double *u, double *vms);
void design_change(hipStream_t s, int64_t n, const double *x, double *xold, double *part, SegG sg = SegG());
void heaviside(hipStream_t s, int64_t n, const double *xf, double *xp, double beta, double x0, double Xmin);
void projected_sens(hipStream_t s, int64_t n, const double *xf, double beta, double x0, double *v);
// Krylov helpers (kernels_solve.hip)
void fcg_init(hipStream_t s, int64_t n, const double *b, const double *q, const double *dinv, double *r, double *z,
              double *part, double *out);
void fcg_update(hipStream_t s, int64_t n, const double *sc, int ia, double *x, const double *p, double *r,
                const double *q, const double *dinv, double *z, double *part, double *out);
void fcg_p(hipStream_t s, int64_t n, const double *z, const double *sc, int ib, int ibo, int first, double *p);
void fcg_p_ctl(hipStream_t s, int64_t n, const double *z, const double *sc, int ib, int ibo, int first,
               const double *stop, double *p);
// one Lanczos (Jacobi-CG) step after q = A p and p.q in sc[ipq]: r -= alpha q,
// z = D^-1 r, rz' and beta into sc, (alpha, beta) into log2
void lanczos_step(hipStream_t s, int64_t n, double *sc, int irz, int ipq, int ibeta, double *r, const double *q,
                  const double *dinv, double *z, double *part, double *log2);
void lanczos_step(hipStream_t s, int64_t n, double *sc, int irz, int ipq, int ibeta, double *r, const float *q,
                  const double *dinv, double *z, double *part, double *log2);
// x += a p ; r -= a q ; z = D^-1 r with a = sc[ia] / sc[ib]; then sc[izz] = ||z||^2,
// sc[ibo] = sc[ibeta], sc[ibeta] = z.r
void fcg_update_ab(hipStream_t s, int64_t n, double *sc, int ia, int ib, double *x, const double *p, double *r,
                   const double *q, const double *dinv, double *z, double *part, int izz, int ibeta, int ibo);
// the filter CG's device-side control in sc[ic .. ic + 4] (state, tolerance,
// divergence bound, iteration count, ||z||): fcg_start after fcg_init's sums
// in sc[i0], sc[i0 + 1]; fcg_update_ctl is fcg_update_ab that counts the
// iteration, applies the tests and does nothing once the state is set
void fcg_start(hipStream_t s, double *sc, int i0, int ic, double rtol);
void fcg_update_ctl(hipStream_t s, int64_t n, double *sc, int ia, int ib, double *x, const double *p, double *r,
                    const double *q, const double *dinv, double *z, double *part, int izz, int ibeta, int ibo,
                    int ic);
void sc_div(hipStream_t s, double *sc, int a, int b, int dst);
void sc_beta(hipStream_t s, double *sc, int inew, int iold, int ibeta, double *log2);
void sc_copy(hipStream_t s, const double *src, double *dst);
void f2d(hipStream_t s, int64_t n, const float *a, double *b);
void fill_hash(hipStream_t s, int64_t n, double *x);
void dot_pr(hipStream_t s, int64_t n, const double *r, const double *z, const double *zo, double *part, double *out);
void sc_beta_flex(hipStream_t s, double *sc, int irz, int irzo, int iprev, int ibeta);
// final sums of (r.z, r.zo) into sc[irz], sc[irz + 1], then beta = (r.z - r.zo) /
// sc[iprev] into sc[ibeta] and sc[iprev] = r.z, in one launch; with the
// partials' kernel: dot_pr_beta (FP64 z), z_from_f32_beta (z from the FP32 cycle)
void final_beta_flex(hipStream_t s, const double *part, int nb, double *sc, int irz, int iprev, int ibeta);
void dot_pr_beta(hipStream_t s, int64_t n, const double *r, const double *z, const double *zo, double *part,
                 double *sc, int irz, int iprev, int ibeta);
void z_from_f32_beta(hipStream_t s, int64_t n, const float *zf, const double *r, double *z, double *zo, double *part,
                     double *sc, int irz, int iprev, int ibeta);
// the same writing the new z to znew (the caller swaps z and zo), no copy
void z_from_f32_swap_beta(hipStream_t s, int64_t n, const float *zf, const double *r, const double *zold,
                          double *znew, double *part, double *sc, int irz, int iprev, int ibeta);
void lincomb(hipStream_t s, int64_t n, const double *sc, int ia, const double *a, int ib, const double *b, double *y);
void lincombf(hipStream_t s, int64_t n, const double *sc, int ia, const float *a, int ib, const float *b, float *y);
void axpy_inplace_t(hipStream_t s, int64_t n, const double *d, double *x);
void axpy_inplace_t(hipStream_t s, int64_t n, const float *d, float *x);
void kcycle_coef1(hipStream_t s, double *sc);
void kcycle_coef2(hipStream_t s, double *sc);
// design batches (kernels_batch.hip): per-design scalars at sc + d * UFT_BSC
void seg_final(hipStream_t s, const double *part, const int32_t *dfirst, int nseg, int K, double *sc, int base,
               const double *gate = nullptr);
void seg_final_max(hipStream_t s, const double *part, const int32_t *dfirst, int nseg, double *sc, int base);
void seg_dot(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, const double *a, const double *b,
             double *part, double *sc, int base);
// design-batch K-cycle (kernels_batch.hip): K = 2 or 3 per-design inner
// products a_k . b_k into sc[d * UFT_BSC + base + k]; the coefficients of the
// two inner steps at base; the second start and the prolongation with the
// coefficients of each row's design (segment grid of the rows)
void seg_dots(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, int K, const double *const *a,
              const double *const *b, double *part, double *sc, int base);
void seg_dots(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, int K, const float *const *a,
              const float *const *b, double *part, double *sc, int base);
void bsc_kc1(hipStream_t s, int nseg, double *sc, int base);
void bsc_kc2(hipStream_t s, int nseg, double *sc, int base);
void seg_lincomb_first(hipStream_t s, SegG g, int64_t n, const double *sc, int ic, const double *b, const double *v,
                       double *kr, double *r, const double *dinv, double it, double *d, double *kd);
void seg_lincomb_first(hipStream_t s, SegG g, int64_t n, const double *sc, int ic, const float *b, const float *v,
                       float *kr, float *r, const float *dinv, double it, float *d, float *kd);
void seg_prolong_kc(hipStream_t s, SegG g, int bs, int64_t n, const int32_t *parent, const double *P, const double *sc,
                    int i5, const double *xa, const double *xb, double *x);
void seg_prolong_kc(hipStream_t s, SegG g, int bs, int64_t n, const int32_t *parent, const float *P, const double *sc,
                    int i5, const float *xa, const float *xb, float *x);
void seg_dot_pr(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, const double *r, const double *z,
                const double *zo, double *part, double *sc, int base);
void seg_cg_update(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, double *sc, int ia, double *x,
                   const double *p, double *r, const double *q, double *part, int base);
void seg_xpby(hipStream_t s, SegG g, int64_t n, const double *z, const double *sc, int ib, double *p);
void seg_fill(hipStream_t s, SegG g, int64_t n, const double *vals, double *x);
void seg_scale_dinv(hipStream_t s, SegG g, int64_t n, const double *fac, double *dinv, float *dinvf);
void seg_fcg_init(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, const double *b, const double *q,
                  const double *dinv, double *r, double *z, double *part, double *sc, int base);
void seg_fcg_update(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, double *sc, int ia, double *x,
                    const double *p, double *r, const double *q, const double *dinv, double *z, double *part, int base);
void seg_fcg_p(hipStream_t s, SegG g, int64_t n, const double *z, const double *sc, int ic, int first, double *p);
// per-design scalar recurrences (one thread per design)
void bsc_el_init(hipStream_t s, int nseg, double *sc, double rtol);
void bsc_el_alpha(hipStream_t s, int nseg, double *sc);
void bsc_el_check(hipStream_t s, int nseg, double *sc, double *flags);
void bsc_el_beta(hipStream_t s, int nseg, double *sc);
void bsc_el_check_lo(hipStream_t s, int nseg, double *sc, double *flags);
void bsc_el_true(hipStream_t s, int nseg, double *sc, double *flags);
// MMA dual interior point of a batch on the device (m = 1): start (after the
// GenSub sums in B_RED; gt: per design gx and the tolerance), the step
// decision, the Newton step with its line search, the dual residual
void bsc_mma_init(hipStream_t s, int nseg, double *sc, const double *gt, double a, double c, double *lamd);
void bsc_mma_decide(hipStream_t s, int nseg, double *sc, double a, double c, double *lamd, double *flags);
void bsc_mma_step(hipStream_t s, int nseg, double *sc, double a, double c, double *lamd);
void bsc_mma_resid(hipStream_t s, int nseg, double *sc, double a);
// q = K U in FP64 when *gate != 0 (element and node phase; every workgroup
// returns at once otherwise); with dgate (a design batch's scalars + slot)
// only the elements and nodes of the designs whose dgate entry is nonzero
void apply_K_gated(hipStream_t s, const uft_ctx *C, const double *xphys, const double *U, double *q, const double *gate,
                   const double *dgate = nullptr);
void seg_true_res(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, double *sc, const double *F,
                  const double *q, double *r, double *part, int base, const double *gate = nullptr);
void seg_dot_lo(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, const double *a, const float *b,
                double *part, double *sc, int base);
void seg_cg_update_lo(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, double *sc, int ia, double *x,
                      const double *p, double *r, const float *q, double *part, int base);
void bsc_f_init(hipStream_t s, int nseg, double *sc, double rtol);
void bsc_f_alpha(hipStream_t s, int nseg, double *sc);
void bsc_f_after(hipStream_t s, int nseg, double *sc);
void bsc_lanczos_alpha(hipStream_t s, int nseg, double *sc, double *log, int step);
void bsc_lanczos_beta(hipStream_t s, int nseg, double *sc, double *log, int step);
// coarsest level of a batch: block-diagonal dense matrix and its inverse
void seg_bsr_to_dense(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val,
                      const int32_t *udes, const int32_t *cu, const int64_t *coff, double *dense);
void seg_dense_inverse(hipStream_t s, int nseg, const int32_t *cu, const int64_t *coff, double *A, double *Ainv);
}  // namespace uftk

// host helpers (uft_api.cpp)
int uft_set_error(uft_ctx *C, int code, const char *fmt, ...);
#define UFT_HIP(C, call)                                                                       \
  do {                                                                                         \
    hipError_t _e = (call);                                                                    \
    if (_e != hipSuccess)                                                                      \
      return uft_set_error((C), UFT_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                           __FILE__, __LINE__);                                                \
  } while (0)

// amg setup (uft_amg_setup.cpp)
int uft_amg_build_host(uft_ctx *C, std::vector<AmgHostLevel> &H, std::vector<int64_t> &nunits, bool sa);
// mesh helpers (uft_mesh.cpp)
void uft_node_incidence(int64_t nn, int64_t ne, int npe, const int32_t *elem, std::vector<int32_t> &ptr,
                        std::vector<int32_t> &inc);
// Sliced-ELL pattern of the assembled fine operator (3x3 node blocks): slices
// of UFT_ELL_SLICE consecutive nodes, slot s of node n at
// off[n / slice] + s * slice + n % slice, the row's neighbours ascending,
// padding slots pointing at the node itself; slot i's element contributions
// are clist[cptr[i] .. cptr[i + 1]), each (e * npe + a) * npe + b: element e
// holds the row's node as local node a and the column's node as local node b.
// False when a node has more than 255 neighbours or the codes overflow int32.
bool uft_fine_ell_pattern(int64_t nn, int npe, const int32_t *elem, const std::vector<int32_t> &iptr,
                          const std::vector<int32_t> &inc, std::vector<int64_t> &off, std::vector<int32_t> &width,
                          std::vector<int32_t> &col, std::vector<int32_t> &cptr, std::vector<int32_t> &clist);
// Node tiles of the sliced-ELL pattern for the multi-load-case products:
// recursive coordinate bisection of the nodes (cut across the longest extent,
// the lower part a multiple of `tile` nodes) into tiles of at most `tile`
// nodes whose neighbour sets (halos) hold at most `hcap` nodes (a tile over
// the cap is halved again).  Tile t's nodes, ascending, sit at positions
// t tile + i (node[] = -1 past its count); the positions form slices of
// UFT_ELL_SLICE with the node-ordered pattern's slot order per node.
struct NodeTiles {
  int64_t ntile = 0, nslot = 0;
  int hmax = 0;
  std::vector<int32_t> node;          // [ntile * tile] node at each position, -1 for padding
  std::vector<int32_t> hoff, hnode;   // tile t's halo: hnode[hoff[t] .. hoff[t + 1]), ascending
  std::vector<int64_t> off;           // per slice of positions: first slot (nslice + 1)
  std::vector<int32_t> width;         // per slice of positions: slots per position
  std::vector<uint16_t> lc;           // per tiled slot: index of the column node in the tile's halo
  std::vector<int32_t> kmap;          // per node-ordered slot: its tiled slot, -1 for padding
};
bool uft_node_tiles(int64_t nn, const double *xyz, const std::vector<int64_t> &off, const std::vector<int32_t> &width,
                    const std::vector<int32_t> &col, int tile, int hcap, NodeTiles &T);
