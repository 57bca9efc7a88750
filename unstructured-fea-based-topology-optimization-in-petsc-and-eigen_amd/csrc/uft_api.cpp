// C-ABI implementation: context lifetime, mesh/BC ingest, setup, and the
// host orchestration of the device design iteration (main.cpp:64-119).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "uft_internal.h"

// ---------------------------------------------------------------- errors
// Messages of calls that take no context (uft_mma_design_change_noalloc)
// are kept per host thread and read with uft_last_error(NULL).
static thread_local std::string g_free_err;

int uft_set_error(uft_ctx *C, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (C) C->err = buf;
  else g_free_err = buf;
  return code;
}

namespace {

// device scalar slots
enum {
  S_RZ = 0, S_PQ = 1, S_ALPHA = 2, S_RR = 3, S_RZNEW = 4, S_BETA = 5,
  S_F0 = 8,  // filter: 8 ||Db||^2, 9 ||z||^2, 10 z.r (beta), 11 beta_old, 12 pq, 13 alpha
  S_MISC = 16,
  S_RED = 32,  // general reduction outputs (up to 48 values)
  S_K = 96,    // K-cycle scalars, 8 per level
  S_FD = 200,  // filter CG control: state, tolerance, divergence bound, iterations, ||z||
  S_CTL = 208, // elasticity CG control of the mixed forms (uftk::ksp_ctl_init): 7 values
  S_BFL = 216, // design batches: elasticity CG flags (BF_*), 3 values
  S_MFL = 220  // design batches: MMA dual flags (MF_*), 2 values
};

int dalloc(uft_ctx *C, void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return uft_set_error(C, UFT_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  C->allocs.push_back(*p);
  return UFT_OK;
}
template <typename T>
int dalloc_n(uft_ctx *C, T **p, int64_t n) {
  return dalloc(C, (void **)p, (size_t)std::max<int64_t>(n, 1) * sizeof(T));
}
template <typename T>
int upload(uft_ctx *C, T **p, const std::vector<T> &h) {
  int rc = dalloc_n(C, p, (int64_t)h.size());
  if (rc) return rc;
  if (!h.empty()) UFT_HIP(C, hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, C->s));
  return UFT_OK;
}

// segment grid of a vector space whose rows are grouped by design (off:
// nseg + 1 row offsets): blocks of at most 256 rows inside one design
int build_space(uft_ctx *C, SegSpace &S, const std::vector<int64_t> &off) {
  const int nseg = (int)off.size() - 1;
  S.off = off;
  S.nrows = off[nseg];
  std::vector<int64_t> span;
  std::vector<int32_t> des;
  S.dfirst.assign(nseg + 1, 0);
  for (int d = 0; d < nseg; ++d) {
    S.dfirst[d] = (int32_t)des.size();
    for (int64_t r = off[d]; r < off[d + 1]; r += 256) {
      span.push_back(r);
      span.push_back(std::min(off[d + 1], r + 256));
      des.push_back(d);
    }
  }
  S.dfirst[nseg] = (int32_t)des.size();
  S.nb = (int)des.size();
  if (S.nb == 0) return uft_set_error(C, UFT_ERR_ARG, "design batch: empty vector space");
  int rc;
  if ((rc = upload(C, &S.d_span, span))) return rc;
  if ((rc = upload(C, &S.d_des, des))) return rc;
  return upload(C, &S.d_dfirst, S.dfirst);
}

void free_all(uft_ctx *C) {
  if (C->s) (void)hipStreamSynchronize(C->s);
  for (hipStream_t t : C->side)  // a lagged coarsest inverse may still run
    if (t) (void)hipStreamSynchronize(t);
  for (void *p : C->allocs) (void)hipFree(p);
  C->allocs.clear();
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int sync(uft_ctx *C) {
  UFT_HIP(C, hipStreamSynchronize(C->s));
  UFT_HIP(C, hipGetLastError());
  return UFT_OK;
}

// copy n device scalars to the pinned mirror and wait
int fetch(uft_ctx *C, const double *dsrc, int n, double *hdst) {
  UFT_HIP(C, hipMemcpyAsync(hdst, dsrc, n * sizeof(double), hipMemcpyDeviceToHost, C->s));
  return sync(C);
}

bool valid_vec(const uft_ctx *C, int id) {
  return id >= 0 && id < (int)C->vec.size() && C->vec[id] != nullptr;
}

// ---------------------------------------------------------------- operators
// fine stiffness apply, with an event triple recorded around its two phases
// while the device timers are enabled
hipEvent_t *prof_triple(uft_ctx *C, int kind) {
  if (!C->prof) return nullptr;
  if (C->prof_tick[kind]++ % C->prof_every != 0) return nullptr;
  while (C->evpool.size() < C->evused + 3) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    C->evpool.push_back(e);
  }
  hipEvent_t *ev = &C->evpool[C->evused];
  C->evkind.resize(C->evpool.size() / 3 + 1);
  C->evkind[C->evused / 3] = kind;
  C->evused += 3;
  return ev;
}

bool is_mixed(const uft_ctx *C) { return C->P.precision != UFT_FP64; }

// ---- algorithmic bytes of one operator apply: the stored operator once
// (Ne (S w + 4 npe + 8) for element blocks with S unique entries of w bytes,
// node ids and density; blocks x (block bytes + column index) for assembled
// blocks) plus the input and output vectors once.
double fine_bytes64(const uft_ctx *C) {
  const int S = C->npe == 8 ? 300 : 78;
  return (double)C->ne * (S * 8 + 4 * C->npe + 8) + 2.0 * (double)C->ndof * 8;
}
double fine_bytes_lo(const uft_ctx *C, int w) {
  const int S = C->npe == 8 ? 300 : 78;
  return (double)C->ne * (S * w + 4 * C->npe + 8 + (w == 2 ? 4 : 0)) + 2.0 * (double)C->ndof * 4;
}
// the FP64 form's assembled operator: FP64 blocks, FP64 u and output
double ell_bytes64(const uft_ctx *C) {
  const int cb = C->fell.d_cd ? 2 : 4;
  return (double)C->fell.nnzb * (72 + cb) + (double)C->ndof * 8 * 2 + (double)C->nn;
}
double ell_bytes(const uft_ctx *C, int value_bytes, bool scaled) {
  const int cb = C->fell.d_cd ? 2 : 4;
  return (double)C->fell.nnzb * (value_bytes + cb) + (double)C->ndof * 4 * (scaled ? 3 : 2) + (double)C->nn;
}
double coarse_bytes(const uft_ctx *C, int l, int w) {
  const AmgLevel &L = C->amg.L[l];
  return (double)L.nblk * (36.0 * w + 4) + 2.0 * (double)L.n * (w == 8 ? 8 : 4) + (w == 2 ? 4.0 * L.n : 0.0);
}
// the same for nr interleaved right-hand sides: the stored operator once,
// each vector nr times (w bytes per vector entry)
double mr_fine_bytes(const uft_ctx *C, int kw, int w, int nr) {
  const int S = C->npe == 8 ? 300 : 78;
  return (double)C->ne * (S * kw + 4 * C->npe + 8) + 2.0 * nr * (double)C->ndof * w;
}
double mr_ell_bytes(const uft_ctx *C, int value_bytes, bool scaled, int nr) {
  const int cb = C->fell.d_cd ? 2 : 4;
  return (double)C->fell.nnzb * (value_bytes + cb) + (double)C->ndof * 4 * (2 * nr + (scaled ? 1 : 0)) +
         (double)C->nn;
}
double mr_coarse_bytes(const uft_ctx *C, int l, int w, int nr) {
  const AmgLevel &L = C->amg.L[l];
  return (double)L.nblk * (36.0 * w + 4) + 2.0 * nr * (double)L.n * w;
}
inline void count_apply(uft_ctx *C, double bytes) {
  if (!C->count_on) return;
  C->solve_bytes += bytes;
  C->solve_applies++;
}
// the assembled smoother blocks (int8 for one right-hand side, FP16 for the
// multi-load-case form) exist
// assembled fine blocks that the refresh fills: the low-precision smoother
// blocks of UFT_MIXED_H and the FP32 blocks (UFT_MIXED_H's outer product;
// UFT_MIXED's smoother and outer product)
bool fine_blocks(const uft_ctx *C) { return C->fell.d_q8 || C->fell.d_v8 || C->fell.d_w8; }
bool smoother_lo(const uft_ctx *C) { return C->fell.d_q8 || C->fell.d_v8; }

void graph_release(uft_ctx *C) {
  if (C->ig_exec) (void)hipGraphExecDestroy(C->ig_exec);
  if (C->ig_graph) (void)hipGraphDestroy(C->ig_graph);
  C->ig_exec = nullptr;
  C->ig_graph = nullptr;
  C->ig_refresh = -1;
}

// Graph replays are used unless the device timers are on: event records
// captured into a graph cannot be read back as elapsed times, so a timed
// solve launches kernel by kernel (the same kernels and arguments).
bool use_graph(const uft_ctx *C) { return C->P.graph && !C->prof; }

// Launch `body` (a fixed sequence of kernel launches on C->s whose arguments
// depend only on the key pointers and the hierarchy refresh) through a
// hipGraph: captured on first use, replayed afterwards.  The body contains no
// memset or copy nodes (vector fills and copies are kernels).
template <typename F>
int graph_launch(uft_ctx *C, const void *k0, const void *k1, const void *k2, F &&body) {
  const void *key[4] = {k0, k1, k2,
                        (const void *)(intptr_t)(256 * C->P.cheb_degree_coarse + 16 * C->P.cheb_degree +
                                                 C->P.cycle_type + 1)};
  // a batch folds its Chebyshev scaling into D^-1, so its cycle's launch
  // arguments do not change with a refresh; a single design's do
  bool same = C->ig_exec && (C->bt.nseg || C->ig_refresh == C->refreshes);
  for (int k = 0; k < 4; ++k) same = same && key[k] == C->ig_key[k];
  if (!same) {
    graph_release(C);
    hipError_t e = hipStreamBeginCapture(C->s, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      body();
      e = hipStreamEndCapture(C->s, &C->ig_graph);
    }
    if (e != hipSuccess) {
      graph_release(C);
      return uft_set_error(C, UFT_ERR_HIP, "Krylov iteration capture failed: %s", hipGetErrorString(e));
    }
    UFT_HIP(C, hipGraphInstantiate(&C->ig_exec, C->ig_graph, nullptr, nullptr, 0));
    for (int k = 0; k < 4; ++k) C->ig_key[k] = key[k];
    C->ig_refresh = C->refreshes;
  }
  UFT_HIP(C, hipGraphLaunch(C->ig_exec, C->s));
  return UFT_OK;
}


void fine_apply(uft_ctx *C, const double *xphys, const double *x, double *y, const double *dotw, double *part) {
  if (C->fell.d_e8 && xphys == C->cur_xphys && (!dotw || dotw == x)) {
    // the assembled FP64 blocks of this density's hierarchy refresh
    count_apply(C, ell_bytes64(C));
    hipEvent_t *ev = prof_triple(C, 0);
    if (ev) (void)hipEventRecord(ev[0], C->s);
    RowOp<double> op;
    op.mode = dotw ? ROW_STOREDOT : ROW_STORE;
    op.y = y;
    op.part = C->d_part;
    uftk::fine_ell_op64(C->s, C, x, op);
    if (ev) (void)hipEventRecord(ev[1], C->s);
    if (ev) (void)hipEventRecord(ev[2], C->s);
    if (dotw) uftk::final_sum(C->s, C->d_part, uftk::gather_op_blocks(C->nn), 1, part);
    return;
  }
  count_apply(C, fine_bytes64(C));
  uftk::apply_K(C->s, C, xphys, x, y, dotw, part, prof_triple(C, 0));
}

// Outer operator product of the mixed forms with the FP32 element blocks:
// qf = K(x) p in FP32 (p rounded to FP32 into d_pf unless have_pf), and p.q
// into sc[slot] (slot >= 0)
void fine_apply_lo(uft_ctx *C, const double *xphys, const double *p, int slot, bool have_pf = false) {
  hipStream_t s = C->s;
  if (!have_pf) uftk::d2f(s, C->ndof, p, C->d_pf);
  hipEvent_t *ev = prof_triple(C, 2);
  RowOp<float> op;
  op.mode = slot >= 0 ? ROW_STOREDOT : ROW_STORE;
  op.y = C->d_qf;
  op.part = C->d_part;
  count_apply(C, C->fell.d_w8 ? ell_bytes(C, 36, false) : fine_bytes_lo(C, 4));
  if (ev) (void)hipEventRecord(ev[0], s);
  if (C->fell.d_w8) {  // assembled FP32 blocks of the current hierarchy refresh: one launch
    uftk::fine_ell_op32(s, C, C->d_pf, op);
    if (ev) (void)hipEventRecord(ev[1], s);
  } else {
    uftk::elem_phase_f(s, C, C->elem_variant_f, xphys, C->d_pf, C->d_Yef);
    if (ev) (void)hipEventRecord(ev[1], s);
    uftk::gather_op(s, C, C->d_pf, C->d_Yef, op);
  }
  if (ev) (void)hipEventRecord(ev[2], s);
  if (slot >= 0) uftk::final_sum(s, C->d_part, uftk::gather_op_blocks(C->nn), 1, C->d_scal + slot);
}

// y = A_l x on multigrid level l
void level_apply(uft_ctx *C, int l, const double *x, double *y) {
  AmgLevel &Lv = C->amg.L[l];
  if (l == 0) {
    fine_apply(C, C->cur_xphys, x, y, nullptr, nullptr);
  } else {
    count_apply(C, coarse_bytes(C, l, 8));
    uftk::bsr6_mv(C->s, Lv.nunit, Lv.d_rp, Lv.d_ci, Lv.d_val, x, y);
  }
}

// Chebyshev interval [lo, hi] * lambda_max(D^-1 A); PCGAMG's default is
// [0.1, 1.1] (-mg_levels_ksp_chebyshev_esteig 0,0.1,0,1.1).
struct ChebCoef {
  double inv_theta, c1[16], c2[16];
};
// lmax: the level's estimate already scaled by cheb_hi; emin = lo/hi of it.
ChebCoef cheb_coef(double lmax, int deg, double lo_over_hi) {
  ChebCoef cc{};
  const double emax = lmax, emin = lo_over_hi * lmax;
  const double theta = 0.5 * (emax + emin), delta = 0.5 * (emax - emin);
  const double sigma = theta / delta;
  cc.inv_theta = 1.0 / theta;
  double rho = 1.0 / sigma;
  for (int k = 1; k < deg && k < 16; ++k) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    cc.c1[k] = rn * rho;
    cc.c2[k] = 2.0 * rn / delta;
    rho = rn;
  }
  return cc;
}

// ---------------------------------------------------------------- cycles
// Overloads so one cycle template serves FP64 and FP32 vectors.
void k_cheb_first(hipStream_t s, int64_t n, const double *b, const double *di, double it, double *r, double *d,
                  double *x, int z) { uftk::cheb_first(s, n, b, di, it, r, d, x, z); }
void k_cheb_first(hipStream_t s, int64_t n, const float *b, const float *di, double it, float *r, float *d, float *x,
                  int z) { uftk::cheb_firstf(s, n, b, di, it, r, d, x, z); }
void k_restrict(hipStream_t s, const AmgLevel &Lv, int64_t nc, const double *r, double *bc) {
  uftk::restrict_(s, Lv.bs, nc, Lv.d_cptr, Lv.d_child, Lv.d_P, r, bc);
}
void k_restrict(hipStream_t s, const AmgLevel &Lv, int64_t nc, const float *r, float *bc) {
  uftk::restrictf(s, Lv.bs, nc, Lv.d_cptr, Lv.d_child, Lv.d_Pf, r, bc);
}
template <typename T>
struct LV {
  T *x, *b, *r, *d, *t, *kc, *kv, *kd, *kr;
  const T *dinv;
};
LV<double> lv(AmgLevel &L, double *) { return {L.x, L.b, L.r, L.d, L.t, L.kc, L.kv, L.kd, L.kr, L.d_dinv}; }
LV<float> lv(AmgLevel &L, float *) { return {L.xf, L.bf, L.rf, L.df, L.tf, L.kcf, L.kvf, L.kdf, L.krf, L.d_dinvf}; }
const double *tr_P(const AmgLevel &L, double *) { return L.d_P; }
const float *tr_P(const AmgLevel &L, float *) { return L.d_Pf; }

// ---- fused multigrid cycle
// Operator apply with a per-row epilogue on level l (l = 0: matrix-free fine
// operator of the current density; above: 6x6 block-sparse).
void lev_op(uft_ctx *C, int l, const double *u, const RowOp<double> &op) {
  count_apply(C, l == 0 ? (C->fell.d_e8 ? ell_bytes64(C) : fine_bytes64(C)) : coarse_bytes(C, l, 8));
  if (l == 0) {
    hipEvent_t *ev = prof_triple(C, 0);
    if (ev) (void)hipEventRecord(ev[0], C->s);
    if (C->fell.d_e8) {  // the assembled FP64 blocks of the current hierarchy refresh: one launch
      uftk::fine_ell_op64(C->s, C, u, op);
      if (ev) (void)hipEventRecord(ev[1], C->s);
      if (ev) (void)hipEventRecord(ev[2], C->s);
      return;
    }
    uftk::elem_phase(C->s, C, C->elem_variant, C->cur_xphys, u, C->d_Ye);
    if (ev) (void)hipEventRecord(ev[1], C->s);
    uftk::gather_op(C->s, C, u, C->d_Ye, op);
    if (ev) (void)hipEventRecord(ev[2], C->s);
  } else {
    const AmgLevel &Lv = C->amg.L[l];
    uftk::bsr6_op(C->s, Lv.nunit, Lv.d_rp, Lv.d_ci, Lv.d_val, u, op);
  }
}
void lev_op(uft_ctx *C, int l, const float *u, const RowOp<float> &op) {
  if (l == 0)
    count_apply(C, smoother_lo(C) ? ell_bytes(C, C->fell.d_q8 ? 12 : 18, true)
                                  : (C->fell.d_w8 ? ell_bytes(C, 36, false) : fine_bytes_lo(C, C->d_KEh ? 2 : 4)));
  else
    count_apply(C, coarse_bytes(C, l, C->amg.L[l].d_valh ? 2 : 4));
  if (l == 0) {
    hipEvent_t *ev = prof_triple(C, 1);
    if (ev) (void)hipEventRecord(ev[0], C->s);
    if (fine_blocks(C)) {  // assembled blocks: one launch (timed as the element phase)
      if (smoother_lo(C))
        uftk::fine_ell_op(C->s, C, u, op);
      else  // UFT_MIXED: the FP32 blocks
        uftk::fine_ell_op32(C->s, C, u, op);
      if (ev) (void)hipEventRecord(ev[1], C->s);
      if (ev) (void)hipEventRecord(ev[2], C->s);
      return;
    }
    if (C->d_KEh)
      uftk::elem_phase_h(C->s, C, C->elem_variant_h, C->cur_xphys, u, C->d_Yef);
    else
      uftk::elem_phase_f(C->s, C, C->elem_variant_f, C->cur_xphys, u, C->d_Yef);
    if (ev) (void)hipEventRecord(ev[1], C->s);
    uftk::gather_op(C->s, C, u, C->d_Yef, op);
    if (ev) (void)hipEventRecord(ev[2], C->s);
  } else {
    const AmgLevel &Lv = C->amg.L[l];
    if (Lv.d_valh)
      uftk::bsr6_op_h(C->s, Lv.nunit, Lv.d_rp, Lv.d_ci, Lv.d_valh, Lv.d_dh, u, op);
    else
      uftk::bsr6_op(C->s, Lv.nunit, Lv.d_rp, Lv.d_ci, Lv.d_valf, u, op);
  }
}
// blocks of the level-l operator launch (its block partials), per vector type
int lev_op_blocks(const uft_ctx *C, int l, bool f32) {
  if (l == 0) return uftk::gather_op_blocks(C->nn);
  const AmgLevel &Lv = C->amg.L[l];
  return f32 && Lv.d_valh ? uftk::bsr6_op_h_blocks(Lv.nunit) : uftk::bsr6_op_blocks(Lv.nunit);
}

// How a coarse level l >= 1 is solved inside the cycle of level l-1:
// 0 one cycle (V), 1 two stationary cycles (W), 2 a two-step K-cycle.
// Chebyshev degree of level l: cheb_degree, or cheb_degree_coarse (when set)
// on levels 2 and below, whose applies are launch-latency bound
int level_degree(const uft_ctx *C, int l) {
  const int d = (l >= 2 && C->P.cheb_degree_coarse > 0) ? C->P.cheb_degree_coarse : C->P.cheb_degree;
  return std::max(1, d);
}

int level_kind(const uft_ctx *C, int l) {
  const int ct = C->P.cycle_type;
  if (ct == 0 || (ct == 3 && l > 1)) return 0;
  // design batches: the K-cycle with per-design inner products on the first
  // coarse levels (their scalars have device slots), the W-cycle below
  if (C->bt.nseg) return ct >= 2 && l <= UFT_BATCH_KLEV ? 2 : 1;
  return ct == 1 ? 1 : 2;
}

template <typename T>
void cycle_body(uft_ctx *C, int l, const T *b, T *x);
const SegSpace &level_space(const uft_ctx *C, int l);

// Solve on coarse level l (1 <= l < coarsest) whose first smoothing step from
// a zero guess has been done by the restriction; the result is left where the
// prolongation reads it (prolong_from).
template <typename T>
void coarse_body(uft_ctx *C, int l) {
  AmgLevel &Lv = C->amg.L[l];
  LV<T> v = lv(Lv, (T *)nullptr);
  const int kind = level_kind(C, l);
  const double it = cheb_coef(Lv.lmax, level_degree(C, l), C->P.cheb_lo / C->P.cheb_hi).inv_theta;
  if (kind == 0) {
    cycle_body<T>(C, l, v.b, v.x);
  } else if (kind == 1) {
    cycle_body<T>(C, l, v.b, v.x);
    RowOp<T> op;
    op.mode = ROW_RESFIRSTX;
    op.base = v.b;
    op.r = v.r;
    op.kr = v.kr;
    op.dinv = v.dinv;
    op.it = it;
    op.dout = v.d;
    op.kd = v.kd;
    lev_op(C, l, v.x, op);
    cycle_body<T>(C, l, v.kr, v.kd);
  } else if (C->bt.nseg) {  // design batch: the same K-cycle with each design's own coefficients
    BatchState &B = C->bt;
    const SegSpace &R = level_space(C, l);
    const SegG g = R.g();
    const int base = B_KC + 8 * (l - 1);
    cycle_body<T>(C, l, v.b, v.kc);
    RowOp<T> op;
    op.mode = ROW_STORE;  // kv = A kc
    op.y = v.kv;
    lev_op(C, l, v.kc, op);
    const T *a2[2] = {v.kc, v.kc}, *b2[2] = {v.kv, v.b};  // rho1 = c.v, alpha1 = c.b
    uftk::seg_dots(C->s, g, R.d_dfirst, B.nseg, Lv.n, 2, a2, b2, C->d_part, B.sc, base);
    uftk::bsc_kc1(C->s, B.nseg, B.sc, base);
    uftk::seg_lincomb_first(C->s, g, Lv.n, B.sc, base + 7, v.b, v.kv, v.kr, v.r, v.dinv, it, v.d, v.kd);
    cycle_body<T>(C, l, v.kr, v.kd);
    RowOp<T> o2;
    o2.mode = ROW_STORE;  // t = A kd
    o2.y = v.t;
    lev_op(C, l, v.kd, o2);
    const T *a3[3] = {v.kd, v.kd, v.kd}, *b3[3] = {v.kv, v.t, v.kr};  // gamma = d.v, beta = d.Ad, alpha2 = d.r
    uftk::seg_dots(C->s, g, R.d_dfirst, B.nseg, Lv.n, 3, a3, b3, C->d_part, B.sc, base + 2);
    uftk::bsc_kc2(C->s, B.nseg, B.sc, base);
  } else {
    double *sc = C->d_scal + S_K + 8 * l;
    cycle_body<T>(C, l, v.b, v.kc);
    RowOp<T> op;
    op.mode = ROW_DOTS2;
    op.y = v.kv;
    op.w1 = v.b;
    op.part = C->d_part;
    lev_op(C, l, v.kc, op);
    const int nb = lev_op_blocks(C, l, std::is_same<T, float>::value);
    if (nb <= 1024) {  // few partial rows: every block of the combination sums them itself
      uftk::lincomb_first_kc(C->s, Lv.n, C->d_part, nb, sc, v.b, v.kv, v.kr, v.r, v.dinv, it, v.d, v.kd);
    } else {
      uftk::final_kc1(C->s, C->d_part, nb, sc);
      uftk::lincomb_first(C->s, Lv.n, sc, v.b, v.kv, v.kr, v.r, v.dinv, it, v.d, v.kd);
    }
    cycle_body<T>(C, l, v.kr, v.kd);
    RowOp<T> o3;
    o3.mode = ROW_DOTS3;
    o3.w1 = v.kv;
    o3.w2 = v.kr;
    o3.part = C->d_part;
    lev_op(C, l, v.kd, o3);
    // level >= 2: the final sums are taken by the prolongation into level
    // l - 1 (prolong_from), which follows directly
    if (l == 1) uftk::final_kc2(C->s, C->d_part, nb, sc);
  }
}

// x_l += P (result of coarse level l+1, 1 <= l+1 < coarsest)
template <typename T>
void prolong_from(uft_ctx *C, int l, T *x) {
  AmgLevel &Lv = C->amg.L[l], &Ln = C->amg.L[l + 1];
  LV<T> vn = lv(Ln, (T *)nullptr);
  const T *P = tr_P(Lv, (T *)nullptr);
  if (Lv.sa) {  // level 0's smoothed prolongator, row by row
    switch (level_kind(C, l + 1)) {
      case 0: uftk::sa_prolong_comb(C->s, Lv, nullptr, -1, -1, vn.x, (const T *)nullptr, x); break;
      case 1: uftk::sa_prolong_comb(C->s, Lv, nullptr, -1, -1, vn.x, vn.kd, x); break;
      default:
        if (C->bt.nseg)  // each node with its design's K-cycle coefficients
          uftk::sa_prolong_comb(C->s, Lv, C->bt.sc, B_KC + 8 * l + 5, B_KC + 8 * l + 6, vn.kc, vn.kd, x,
                                C->bt.d_ndes);
        else
          uftk::sa_prolong_comb(C->s, Lv, C->d_scal + S_K + 8 * (l + 1), 5, 6, vn.kc, vn.kd, x);
        break;
    }
    return;
  }
  switch (level_kind(C, l + 1)) {
    case 0:
      uftk::prolong_comb(C->s, Lv.bs, Lv.nunit, Lv.d_parent, P, nullptr, -1, -1, vn.x, (const T *)nullptr, x);
      break;
    case 1:  // x + kd of the two stationary cycles
      uftk::prolong_comb(C->s, Lv.bs, Lv.nunit, Lv.d_parent, P, nullptr, -1, -1, vn.x, vn.kd, x);
      break;
    default:  // K-cycle: c5 kc + c6 kd
      if (C->bt.nseg)  // each row with its design's coefficients
        uftk::seg_prolong_kc(C->s, level_space(C, l).g(), Lv.bs, Lv.n, Lv.d_parent, P, C->bt.sc, B_KC + 8 * l + 5,
                             vn.kc, vn.kd, x);
      else if (l >= 1)  // the level's final K-cycle sums in the same launch (few partials per block)
        uftk::prolong_kc(C->s, Lv.bs, Lv.nunit, Lv.d_parent, P, C->d_part,
                         lev_op_blocks(C, l + 1, std::is_same<T, float>::value), C->d_scal + S_K + 8 * (l + 1),
                         vn.kc, vn.kd, x);
      else
        uftk::prolong_comb(C->s, Lv.bs, Lv.nunit, Lv.d_parent, P, C->d_scal + S_K + 8 * (l + 1), 5, 6, vn.kc, vn.kd,
                           x);
      break;
  }
}

// Coarse-grid correction of level l when level l+1 is the coarsest (dense
// FP64 inverse): x += P Ainv P^T r.
template <typename T>
void coarse_direct(uft_ctx *C, int l, const T *r, T *x) {
  AmgHier &H = C->amg;
  AmgLevel &Lv = H.L[l], &Ln = H.L[l + 1];
  LV<T> vn = lv(Ln, (T *)nullptr);
  const T *P = tr_P(Lv, (T *)nullptr);
  {  // the dense inverse once (block diagonal in a batch: every design's block)
    double ent = (double)H.ncoarse * H.ncoarse;
    if (C->bt.nseg) {
      ent = 0.0;
      for (int d = 0; d < C->bt.nseg; ++d) {
        const double nd = (double)(C->bt.uoff[H.nlev - 1][d + 1] - C->bt.uoff[H.nlev - 1][d]) * 6.0;
        ent += nd * nd;
      }
    }
    const bool half = std::is_same<T, float>::value && H.d_cinvh;  // row-scaled FP16 copy plus row scales
    count_apply(C, half ? ent * 2 + 4.0 * H.ncoarse + 2.0 * (double)H.ncoarse * sizeof(T)
                        : ent * (C->bt.nseg || std::is_same<T, double>::value ? 8 : 4) +
                              2.0 * (double)H.ncoarse * sizeof(T));
  }
  if (C->bt.nseg) {  // block-diagonal coarsest level: one workgroup per design
    uftk::tail_solve(C->s, (int)Ln.nunit, Lv.d_cptr, Lv.d_child, P, H.d_cinv, r, vn.x, C->bt.nseg, C->bt.d_cu,
                     C->bt.d_coff);
    uftk::prolong_comb(C->s, Lv.bs, Lv.nunit, Lv.d_parent, P, nullptr, -1, -1, vn.x, (const T *)nullptr, x);
    return;
  }
  if (Lv.bs == 6 && H.ncoarse <= UFT_TAIL_MAX) {  // restriction, dense solve and prolongation: one launch
    if constexpr (std::is_same<T, float>::value)  // FP32 inverse inside the FP32 cycle
      uftk::tail_solve_prolong(C->s, (int)Ln.nunit, Lv.d_cptr, Lv.d_child, P, H.d_cinvf, r, vn.x, Lv.nunit,
                               Lv.d_parent, x);
    else
      uftk::tail_solve_prolong(C->s, (int)Ln.nunit, Lv.d_cptr, Lv.d_child, P, H.d_cinv, r, vn.x, Lv.nunit,
                               Lv.d_parent, x);
    return;
  }
  k_restrict(C->s, Lv, Ln.nunit, r, vn.b);
  if constexpr (std::is_same<T, double>::value) {
    uftk::dense_mv(C->s, H.ncoarse, H.d_cinv, vn.b, vn.x);
  } else if (H.d_cinvh && ((uintptr_t)vn.b % 16) == 0) {
    uftk::dense_mvh(C->s, H.ncoarse, H.d_cinvh, H.d_cinvs, vn.b, vn.x);  // the row-scaled FP16 copy
  } else if (H.d_cinvf && H.ncoarse % 4 == 0 && ((uintptr_t)vn.b % 16) == 0 && ((uintptr_t)vn.x % 16) == 0) {
    uftk::dense_mvf4(C->s, H.ncoarse, H.d_cinvf, vn.b, vn.x);  // the FP32 copy of the inverse
  } else if (H.d_cinvf) {
    uftk::dense_mvf(C->s, H.ncoarse, H.d_cinvf, vn.b, vn.x);
  } else {
    uftk::f2d(C->s, Ln.n, vn.b, Ln.kc);
    uftk::dense_mv(C->s, H.ncoarse, H.d_cinv, Ln.kc, Ln.kv);
    uftk::d2f(C->s, Ln.n, Ln.kv, vn.x);
  }
  uftk::prolong_comb(C->s, Lv.bs, Lv.nunit, Lv.d_parent, P, nullptr, -1, -1, vn.x, (const T *)nullptr, x);
}

// One cycle on level l for right-hand side b, after its first Chebyshev
// step from a zero guess (r = b, d = it D^-1 b, x = d) is in place: the
// remaining pre-smoothing steps, the coarse-grid correction, and the same
// polynomial as post-smoothing (a symmetric cycle).  Result in x.
template <typename T>
void cycle_body(uft_ctx *C, int l, const T *b, T *x) {
  AmgHier &H = C->amg;
  AmgLevel &Lv = H.L[l];
  LV<T> v = lv(Lv, (T *)nullptr);
  const int deg = level_degree(C, l);
  const ChebCoef cc = cheb_coef(Lv.lmax, deg, C->P.cheb_lo / C->P.cheb_hi);
  T *dc = v.d, *dn = v.t;
  RowOp<T> st;
  st.mode = ROW_STEP;
  st.r = v.r;
  st.dinv = v.dinv;
  st.x = x;
  for (int k = 1; k < deg; ++k) {
    st.c1 = cc.c1[k];
    st.c2 = cc.c2[k];
    st.dout = dn;
    st.addprev = 0;
    lev_op(C, l, dc, st);
    std::swap(dc, dn);
  }
  RowOp<T> rs;
  rs.mode = ROW_RES;  // r = b - A x
  rs.base = v.r;
  rs.r = v.r;
  lev_op(C, l, dc, rs);
  if (l + 1 == H.nlev - 1) {
    coarse_direct<T>(C, l, v.r, x);
  } else {
    AmgLevel &Ln = H.L[l + 1];
    LV<T> vn = lv(Ln, (T *)nullptr);
    const double itn = cheb_coef(Ln.lmax, level_degree(C, l + 1), C->P.cheb_lo / C->P.cheb_hi).inv_theta;
    T *xn = level_kind(C, l + 1) == 2 ? vn.kc : vn.x;
    if (Lv.sa)  // smoothed prolongator of level 0: through the aggregates' supports
      uftk::sa_restrict_first(C->s, Lv, Ln.nunit, v.r, vn.b, vn.dinv, itn, vn.r, vn.d, xn);
    else
      uftk::restrict_first(C->s, Lv.bs, Ln.nunit, Lv.d_cptr, Lv.d_child, tr_P(Lv, (T *)nullptr), v.r, vn.b,
                           vn.dinv, itn, vn.r, vn.d, xn);
    coarse_body<T>(C, l + 1);
    prolong_from<T>(C, l, x);
  }
  RowOp<T> rf;
  rf.mode = ROW_RESFIRST;  // r = b - A x ; d = it D^-1 r
  rf.base = b;
  rf.r = v.r;
  rf.dinv = v.dinv;
  rf.it = cc.inv_theta;
  rf.dout = v.d;
  lev_op(C, l, x, rf);
  dc = v.d;
  dn = v.t;
  if (deg == 1) {
    uftk::axpy_inplace_t(C->s, Lv.n, v.d, x);
    return;
  }
  for (int k = 1; k < deg; ++k) {
    st.c1 = cc.c1[k];
    st.c2 = cc.c2[k];
    st.dout = dn;
    st.addprev = k == 1;
    lev_op(C, l, dc, st);
    std::swap(dc, dn);
  }
}

// the fine cycle's first Chebyshev coefficient
double fine_first_coef(uft_ctx *C) {
  return cheb_coef(C->amg.L[0].lmax, level_degree(C, 0), C->P.cheb_lo / C->P.cheb_hi).inv_theta;
}
// x = B_0 b: one fine-level cycle (the preconditioner of the elasticity solve);
// first_done: the first Chebyshev step (r = b, d = x = it D^-1 b) is already
// in place
template <typename T>
void fine_cycle(uft_ctx *C, const T *b, T *x, bool first_done = false) {
  AmgLevel &L0 = C->amg.L[0];
  LV<T> v = lv(L0, (T *)nullptr);
  if (!first_done) k_cheb_first(C->s, L0.n, b, v.dinv, fine_first_coef(C), v.r, v.d, x, 1);
  cycle_body<T>(C, 0, b, x);
}
void vcycle(uft_ctx *C, int, const double *b, double *x) { fine_cycle<double>(C, b, x); }
void vcycle_f(uft_ctx *C, int, const float *b, float *x) { fine_cycle<float>(C, b, x); }

// ---- multi-load-case cycle: the same fused cycle on interleaved
// [unknown][nr] vectors, per right-hand-side K-cycle scalars at
// mr.sc + 128 + 64 l + 8 r.
template <typename T>
struct MLV {
  T *x, *b, *r, *d, *t, *kc, *kv, *kd, *kr;
  const T *dinv;
};
template <typename T>
MLV<T> mlv(uft_ctx *C, int l) {
  MrLevelBuf &B = C->mr.L[l];
  AmgLevel &Lv = C->amg.L[l];
  const T *dinv;
  if constexpr (std::is_same<T, double>::value) dinv = Lv.d_dinv; else dinv = Lv.d_dinvf;
  return {(T *)B.x, (T *)B.b, (T *)B.r, (T *)B.d, (T *)B.t, (T *)B.kc, (T *)B.kv, (T *)B.kd, (T *)B.kr, dinv};
}
double *mr_ksc(uft_ctx *C, int l) { return C->mr.sc + 128 + 64 * l; }

template <typename T>
void mr_lev_op(uft_ctx *C, int l, const T *u, const RowOp<T> &op) {
  const int nr = C->mr.nr;
  constexpr bool f64 = std::is_same<T, double>::value;
  if (l == 0 && f64 && C->fell.d_e8)  // FP64 node blocks, FP64 vectors
    count_apply(C, (double)C->fell.nnzb * (72 + (C->fell.d_cd ? 2 : 4)) + 2.0 * nr * (double)C->ndof * 8 +
                       (double)C->nn);
  else if (l == 0)
    count_apply(C, !f64 && C->fell.d_v8 ? mr_ell_bytes(C, 18, true, nr) : mr_fine_bytes(C, f64 ? 8 : 2, f64 ? 8 : 4, nr));
  else if (!f64 && C->amg.L[l].d_valh)  // FP16 blocks, FP32 row scales and vectors
    count_apply(C, (double)C->amg.L[l].nblk * (72.0 + 4) + 2.0 * nr * (double)C->amg.L[l].n * 4 +
                       4.0 * (double)C->amg.L[l].n);
  else
    count_apply(C, mr_coarse_bytes(C, l, f64 ? 8 : 4, nr));
  if (l == 0) {
    hipEvent_t *ev = prof_triple(C, std::is_same<T, double>::value ? 0 : 1);
    if (ev) (void)hipEventRecord(ev[0], C->s);
    if constexpr (std::is_same<T, float>::value) {
      if (C->fell.d_v8) {  // assembled FP16 blocks: one launch for all right-hand sides
        uftk::mr_ell_op(C->s, C, nr, u, op, false);
        if (ev) (void)hipEventRecord(ev[1], C->s);
        if (ev) (void)hipEventRecord(ev[2], C->s);
        return;
      }
    }
    if constexpr (std::is_same<T, double>::value) {
      if (C->fell.d_e8) {  // the FP64 form: the assembled FP64 node blocks of the current refresh
        uftk::mr_ell_op64(C->s, C, nr, u, op);
        if (ev) (void)hipEventRecord(ev[1], C->s);
        if (ev) (void)hipEventRecord(ev[2], C->s);
        return;
      }
    }
    uftk::mr_elem_phase(C->s, C, nr, C->cur_xphys, u, (T *)C->mr.Ye);
    if (ev) (void)hipEventRecord(ev[1], C->s);
    uftk::mr_gather_op(C->s, C, nr, u, (const T *)C->mr.Ye, op);
    if (ev) (void)hipEventRecord(ev[2], C->s);
  } else {
    const AmgLevel &Lv = C->amg.L[l];
    if constexpr (std::is_same<T, float>::value) {
      if (Lv.d_valh) {  // UFT_MIXED_H: the level's FP16 smoother blocks
        uftk::mr_bsr_op_h(C->s, nr, Lv.nunit, Lv.d_rp, Lv.d_ci, Lv.d_valh, Lv.d_dh, u, op);
        return;
      }
    }
    const T *val;
    if constexpr (std::is_same<T, double>::value) val = Lv.d_val; else val = Lv.d_valf;
    uftk::mr_bsr_op(C->s, nr, Lv.nunit, Lv.d_rp, Lv.d_ci, val, u, op);
  }
}
int mr_lev_blocks(const uft_ctx *C, int l) {
  return l == 0 ? uftk::mr_gather_blocks(C->nn, C->mr.nr) : uftk::mr_bsr_blocks(C->amg.L[l].nunit);
}
// mixed forms' outer product: q = K p with the FP32 element blocks (the
// cycle's fine operator uses the BF16 matrix-core form), p.q per right-hand
// side into the block partials
void mr_outer_apply_lo(uft_ctx *C, const float *pf, float *qf) {
  const int nr = C->mr.nr;
  count_apply(C, C->fell.d_w8 ? mr_ell_bytes(C, 36, false, nr) : mr_fine_bytes(C, 4, 4, nr));
  RowOp<float> op;
  op.mode = ROW_STOREDOT;
  op.y = qf;
  op.part = C->d_part;
  hipEvent_t *ev = prof_triple(C, 2);
  if (ev) (void)hipEventRecord(ev[0], C->s);
  if (C->fell.d_w8) {  // assembled FP32 blocks: one launch for all right-hand sides
    uftk::mr_ell_op(C->s, C, nr, pf, op, true);
    if (ev) (void)hipEventRecord(ev[1], C->s);
  } else {
    uftk::mr_elem_phase(C->s, C, nr, C->cur_xphys, pf, (float *)C->mr.Ye, true);
    if (ev) (void)hipEventRecord(ev[1], C->s);
    uftk::mr_gather_op(C->s, C, nr, pf, (const float *)C->mr.Ye, op);
  }
  if (ev) (void)hipEventRecord(ev[2], C->s);
}

template <typename T>
void mr_cycle_body(uft_ctx *C, int l, const T *b, T *x);

template <typename T>
void mr_coarse_body(uft_ctx *C, int l) {
  AmgLevel &Lv = C->amg.L[l];
  MLV<T> v = mlv<T>(C, l);
  const int nr = C->mr.nr;
  const int kind = level_kind(C, l);
  const double it = cheb_coef(Lv.lmax, level_degree(C, l), C->P.cheb_lo / C->P.cheb_hi).inv_theta;
  if (kind == 0) {
    mr_cycle_body<T>(C, l, v.b, v.x);
  } else if (kind == 1) {
    mr_cycle_body<T>(C, l, v.b, v.x);
    RowOp<T> op;
    op.mode = ROW_RESFIRSTX;
    op.base = v.b;
    op.r = v.r;
    op.kr = v.kr;
    op.dinv = v.dinv;
    op.it = it;
    op.dout = v.d;
    op.kd = v.kd;
    mr_lev_op(C, l, (const T *)v.x, op);
    mr_cycle_body<T>(C, l, v.kr, v.kd);
  } else {
    double *sc = mr_ksc(C, l);
    mr_cycle_body<T>(C, l, v.b, v.kc);
    RowOp<T> op;
    op.mode = ROW_DOTS2;
    op.y = v.kv;
    op.w1 = v.b;
    op.part = C->d_part;
    mr_lev_op(C, l, (const T *)v.kc, op);
    uftk::mr_final_kc1(C->s, nr, C->d_part, mr_lev_blocks(C, l), sc);
    uftk::mr_lincomb_first(C->s, nr, Lv.n, sc, v.b, v.kv, v.kr, v.r, v.dinv, it, v.d, v.kd);
    mr_cycle_body<T>(C, l, v.kr, v.kd);
    RowOp<T> o3;
    o3.mode = ROW_DOTS3;
    o3.w1 = v.kv;
    o3.w2 = v.kr;
    o3.part = C->d_part;
    mr_lev_op(C, l, (const T *)v.kd, o3);
    uftk::mr_final_kc2(C->s, nr, C->d_part, mr_lev_blocks(C, l), sc);
  }
}

template <typename T>
void mr_prolong_from(uft_ctx *C, int l, T *x) {
  AmgLevel &Lv = C->amg.L[l], &Ln = C->amg.L[l + 1];
  MLV<T> vn = mlv<T>(C, l + 1);
  const int nr = C->mr.nr;
  if (Lv.sa) {  // level 0's smoothed prolongator, node by node
    switch (level_kind(C, l + 1)) {
      case 0:
        uftk::mr_sa_prolong_comb(C->s, nr, Lv, Ln.nunit, nullptr, -1, -1, vn.x, (const T *)nullptr, x, vn.t);
        break;
      case 1: uftk::mr_sa_prolong_comb(C->s, nr, Lv, Ln.nunit, nullptr, -1, -1, vn.x, vn.kd, x, vn.t); break;
      default: uftk::mr_sa_prolong_comb(C->s, nr, Lv, Ln.nunit, mr_ksc(C, l + 1), 5, 6, vn.kc, vn.kd, x, vn.t); break;
    }
    return;
  }
  const T *P = tr_P(Lv, (T *)nullptr);
  switch (level_kind(C, l + 1)) {
    case 0:
      uftk::mr_prolong_comb(C->s, nr, Lv.bs, Lv.nunit, Lv.d_parent, P, nullptr, -1, -1, vn.x, (const T *)nullptr, x);
      break;
    case 1:
      uftk::mr_prolong_comb(C->s, nr, Lv.bs, Lv.nunit, Lv.d_parent, P, nullptr, -1, -1, vn.x, vn.kd, x);
      break;
    default:
      uftk::mr_prolong_comb(C->s, nr, Lv.bs, Lv.nunit, Lv.d_parent, P, mr_ksc(C, l + 1), 5, 6, vn.kc, vn.kd, x);
      break;
  }
}

template <typename T>
void mr_cycle_body(uft_ctx *C, int l, const T *b, T *x) {
  AmgHier &H = C->amg;
  AmgLevel &Lv = H.L[l];
  MLV<T> v = mlv<T>(C, l);
  const int nr = C->mr.nr;
  const int deg = level_degree(C, l);
  const double lohi = C->P.cheb_lo / C->P.cheb_hi;
  const ChebCoef cc = cheb_coef(Lv.lmax, deg, lohi);
  T *dc = v.d, *dn = v.t;
  RowOp<T> st;
  st.mode = ROW_STEP;
  st.r = v.r;
  st.dinv = v.dinv;
  st.x = x;
  for (int k = 1; k < deg; ++k) {
    st.c1 = cc.c1[k];
    st.c2 = cc.c2[k];
    st.dout = dn;
    st.addprev = 0;
    mr_lev_op(C, l, (const T *)dc, st);
    std::swap(dc, dn);
  }
  RowOp<T> rs;
  rs.mode = ROW_RES;
  rs.base = v.r;
  rs.r = v.r;
  mr_lev_op(C, l, (const T *)dc, rs);
  if (l + 1 == H.nlev - 1) {
    AmgLevel &Ln = H.L[l + 1];
    MLV<T> vn = mlv<T>(C, l + 1);
    const T *P = tr_P(Lv, (T *)nullptr);
    count_apply(C, (double)H.ncoarse * H.ncoarse * sizeof(T) + 2.0 * nr * (double)H.ncoarse * sizeof(T));
    const T *cinv;
    if constexpr (std::is_same<T, double>::value) cinv = H.d_cinv; else cinv = H.d_cinvf;
    uftk::mr_tail_solve(C->s, nr, Lv.bs, (int)Ln.nunit, Lv.d_cptr, Lv.d_child, P, cinv, vn.b, v.r, vn.x);
    uftk::mr_prolong_comb(C->s, nr, Lv.bs, Lv.nunit, Lv.d_parent, P, nullptr, -1, -1, vn.x, (const T *)nullptr, x);
  } else {
    AmgLevel &Ln = H.L[l + 1];
    MLV<T> vn = mlv<T>(C, l + 1);
    const double itn = cheb_coef(Ln.lmax, level_degree(C, l + 1), lohi).inv_theta;
    T *xn = level_kind(C, l + 1) == 2 ? vn.kc : vn.x;
    if (Lv.sa)  // smoothed prolongator of level 0: through the aggregates' supports
      uftk::mr_sa_restrict_first(C->s, nr, Lv, Ln.nunit, v.r, vn.b, vn.dinv, itn, vn.r, vn.d, xn);
    else
      uftk::mr_restrict_first(C->s, nr, Lv.bs, Ln.nunit, Lv.d_cptr, Lv.d_child, tr_P(Lv, (T *)nullptr), v.r, vn.b,
                              vn.dinv, itn, vn.r, vn.d, xn);
    mr_coarse_body<T>(C, l + 1);
    mr_prolong_from<T>(C, l, x);
  }
  RowOp<T> rf;
  rf.mode = ROW_RESFIRST;
  rf.base = b;
  rf.r = v.r;
  rf.dinv = v.dinv;
  rf.it = cc.inv_theta;
  rf.dout = v.d;
  mr_lev_op(C, l, (const T *)x, rf);
  dc = v.d;
  dn = v.t;
  if (deg == 1) {
    uftk::axpy_inplace_t(C->s, Lv.n * nr, v.d, x);
    return;
  }
  for (int k = 1; k < deg; ++k) {
    st.c1 = cc.c1[k];
    st.c2 = cc.c2[k];
    st.dout = dn;
    st.addprev = k == 1;
    mr_lev_op(C, l, (const T *)dc, st);
    std::swap(dc, dn);
  }
}

template <typename T>
void mr_fine_cycle(uft_ctx *C, const T *b, T *x) {
  AmgLevel &L0 = C->amg.L[0];
  MLV<T> v = mlv<T>(C, 0);
  const double it =
      cheb_coef(L0.lmax, level_degree(C, 0), C->P.cheb_lo / C->P.cheb_hi).inv_theta;
  uftk::mr_cheb_first(C->s, C->mr.nr, L0.n, b, v.dinv, it, v.r, v.d, x);
  mr_cycle_body<T>(C, 0, b, x);
}

// largest eigenvalue of a symmetric tridiagonal matrix (bisection, Sturm counts)
// Operator of the eigenvalue estimates: the one the level's smoother applies
// (in the mixed forms the fine level's FP32 or FP16 element blocks), on FP64
// vectors
void lanczos_apply(uft_ctx *C, int l, const double *x, double *y) {
  AmgLevel &Lv = C->amg.L[l];
  if (l == 0 && is_mixed(C)) {
    uftk::d2f(C->s, Lv.n, x, Lv.xf);
    RowOp<float> op;
    op.mode = ROW_STORE;
    op.y = Lv.tf;
    lev_op(C, 0, Lv.xf, op);
    uftk::f2d(C->s, Lv.n, Lv.tf, y);
    return;
  }
  level_apply(C, l, x, y);
}

double tridiag_lmax(const std::vector<double> &d, const std::vector<double> &e) {
  const int k = (int)d.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  auto count_below = [&](double x) {  // number of eigenvalues < x
    int c = 0;
    double q = d[0] - x;
    if (q < 0) ++c;
    for (int i = 1; i < k; ++i) {
      const double qq = (q == 0.0) ? 1e-300 : q;
      q = d[i] - x - e[i - 1] * e[i - 1] / qq;
      if (q < 0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-12 * std::max(1.0, std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) >= k) hi = mid; else lo = mid;
  }
  return hi;
}

// lambda_max(D^-1 A_l) by k steps of Jacobi-PCG (Lanczos from the CG
// coefficients); the scalars stay on the device until the end.
// The steps for level l, queued on
// C->s: scalars in d_scal[sb .. sb + 4], block partials in part, (alpha,
// beta, -) per step in log.
void lanczos_enqueue(uft_ctx *C, int l, int ksteps, double *part, int sb, double *log) {
  AmgLevel &Lv = C->amg.L[l];
  hipStream_t s = C->s;
  const int64_t n = Lv.n;
  double *x = Lv.x, *r = Lv.r, *z = Lv.d, *p = Lv.b, *q = Lv.t;
  uftk::fill_hash(s, n, r);  // fixed pseudo-random start: not close to a near-null (rigid) mode
  uftk::scale(s, n, r, Lv.d_dinv, z);
  uftk::copy(s, n, z, p);
  uftk::dot(s, n, r, z, part, C->d_scal + sb);  // rz in sb
  // the fine level of the mixed forms: the smoother's operator on FP32 vectors
  const bool lo = l == 0 && is_mixed(C);
  (void)x;
  if (lo) uftk::d2f(s, n, p, Lv.xf);
  for (int j = 0; j < ksteps; ++j) {
    // q = A p and p.q (sb + 1) in the operator's epilogue; one update
    // launch (alpha, r, z = D^-1 r, r.z); alpha, beta logged; p = z + beta p
    if (lo) {
      RowOp<float> op;
      op.mode = ROW_STOREDOT;
      op.y = Lv.tf;
      op.part = part;
      lev_op(C, 0, Lv.xf, op);
      uftk::final_sum(s, part, lev_op_blocks(C, 0, true), 1, C->d_scal + sb + 1);
      uftk::lanczos_step(s, n, C->d_scal, sb, sb + 1, sb + 4, r, Lv.tf, Lv.d_dinv, z, part,
                         log + 3 * j);
      uftk::xpby_lo(s, n, z, C->d_scal, sb + 4, p, Lv.xf);
    } else {
      RowOp<double> op;
      op.mode = ROW_STOREDOT;
      op.y = q;
      op.part = part;
      lev_op(C, l, p, op);
      uftk::final_sum(s, part, lev_op_blocks(C, l, false), 1, C->d_scal + sb + 1);
      uftk::lanczos_step(s, n, C->d_scal, sb, sb + 1, sb + 4, r, q, Lv.d_dinv, z, part,
                         log + 3 * j);
      uftk::xpby(s, n, z, C->d_scal, sb + 4, p);
    }
  }
}

// lambda_max of level l from the host copy hl of its Lanczos log: the
// tridiagonal matrix's largest eigenvalue times cheb_hi
void lanczos_finish(uft_ctx *C, int l, int ksteps, const double *hl) {
  AmgLevel &Lv = C->amg.L[l];
  std::vector<double> d, e;
  for (int j = 0; j < ksteps; ++j) {
    const double al = hl[3 * j], be = hl[3 * j + 1];
    if (!(al > 0.0) || !std::isfinite(al)) break;
    double dj = 1.0 / al;
    if (j > 0) dj += hl[3 * (j - 1) + 1] / hl[3 * (j - 1)];
    d.push_back(dj);
    if (!(be > 0.0) || !std::isfinite(be)) break;
    e.push_back(std::sqrt(be) / al);
  }
  if (d.empty()) {
    Lv.lmax = 2.0 * C->P.cheb_hi;
    return;
  }
  e.resize(d.size());
  Lv.lmax = C->P.cheb_hi * tridiag_lmax(d, e);
}

// both, with one host round trip
int estimate_lmax(uft_ctx *C, int l, int ksteps) {
  lanczos_enqueue(C, l, ksteps, C->d_part, S_MISC, C->d_lanczos);
  std::vector<double> hl(3 * ksteps);
  UFT_HIP(C, hipMemcpyAsync(hl.data(), C->d_lanczos, hl.size() * sizeof(double), hipMemcpyDeviceToHost, C->s));
  int rc = sync(C);
  if (rc) return rc;
  lanczos_finish(C, l, ksteps, hl.data());
  return UFT_OK;
}

// lambda_max from a Lanczos log (alpha_j, beta_j of k Jacobi-CG steps); 0 when
// the log is unusable
double lanczos_lmax(const double *hl, int ksteps) {
  std::vector<double> d, e;
  for (int j = 0; j < ksteps; ++j) {
    const double al = hl[2 * j], be = hl[2 * j + 1];
    if (!(al > 0.0) || !std::isfinite(al)) break;
    double dj = 1.0 / al;
    if (j > 0) dj += hl[2 * (j - 1) + 1] / hl[2 * (j - 1)];
    d.push_back(dj);
    if (!(be > 0.0) || !std::isfinite(be)) break;
    e.push_back(std::sqrt(be) / al);
  }
  if (d.empty()) return 0.0;
  e.resize(d.size());
  return tridiag_lmax(d, e);
}

const SegSpace &level_space(const uft_ctx *C, int l) { return l == 0 ? C->bt.D : C->bt.R[l]; }

// Design batches: lambda_max(D^-1 A_l) of every design by the same Jacobi-CG
// Lanczos steps with per-design scalars; each design's Chebyshev scaling is
// then folded into its rows of D^-1 (D^-1 / lambda_max), so the cycle runs
// with lambda_max = 1 on every level and needs no per-design coefficients.
int estimate_lmax_batch(uft_ctx *C, int l, int ksteps) {
  AmgLevel &Lv = C->amg.L[l];
  BatchState &B = C->bt;
  const SegSpace &S = level_space(C, l);
  const SegG g = S.g();
  hipStream_t s = C->s;
  const int ns = B.nseg;
  const int64_t n = Lv.n;
  double *x = Lv.x, *r = Lv.r, *z = Lv.d, *p = Lv.b, *q = Lv.t;
  uftk::fill_hash(s, n, r);  // fixed pseudo-random start: not close to a near-null (rigid) mode
  uftk::scale(s, n, r, Lv.d_dinv, z);
  uftk::copy(s, n, z, p);
  uftk::fill(s, n, 0.0, x);
  uftk::seg_dot(s, g, S.d_dfirst, ns, n, r, z, C->d_part, B.sc, B_LRZ);
  for (int j = 0; j < ksteps; ++j) {
    lanczos_apply(C, l, p, q);
    uftk::seg_dot(s, g, S.d_dfirst, ns, n, p, q, C->d_part, B.sc, B_LPQ);
    uftk::bsc_lanczos_alpha(s, ns, B.sc, B.d_log, j);
    uftk::seg_cg_update(s, g, S.d_dfirst, ns, n, B.sc, B_LAL, x, p, r, q, C->d_part, B_LRR);
    uftk::scale(s, n, r, Lv.d_dinv, z);
    uftk::seg_dot(s, g, S.d_dfirst, ns, n, r, z, C->d_part, B.sc, B_LRZN);
    uftk::bsc_lanczos_beta(s, ns, B.sc, B.d_log, j);
    uftk::seg_xpby(s, g, n, z, B.sc, B_LBE, p);
  }
  std::vector<double> hl((size_t)ns * 128);
  int rc = fetch(C, B.d_log, ns * 128, hl.data());
  if (rc) return rc;
  double *fac = B.h + (size_t)ns * UFT_BSC;
  for (int d = 0; d < ns; ++d) {
    const double lm = lanczos_lmax(&hl[(size_t)d * 128], ksteps);
    fac[d] = 1.0 / (lm > 0.0 ? C->P.cheb_hi * lm : 2.0 * C->P.cheb_hi);
  }
  UFT_HIP(C, hipMemcpyAsync(B.d_fac, fac, ns * sizeof(double), hipMemcpyHostToDevice, s));
  uftk::seg_scale_dinv(s, g, n, B.d_fac, Lv.d_dinv, Lv.d_dinvf);
  Lv.lmax = 1.0;
  return sync(C);  // the pinned staging of fac is reused by the next level
}

// Rebuild the density-dependent parts of the hierarchy for xphys.
// operator applies of the refresh (eigenvalue estimates) are not solve traffic
struct CountOff {
  uft_ctx *C;
  bool was;
  explicit CountOff(uft_ctx *c) : C(c), was(c->count_on) { c->count_on = false; }
  ~CountOff() { C->count_on = was; }
};
int amg_refresh_impl(uft_ctx *C, const double *xphys);
int amg_refresh(uft_ctx *C, const double *xphys) {
  CountOff off(C);
  return amg_refresh_impl(C, xphys);
}
// The dense inverse of the coarsest level (and its FP32 copy in the mixed
// forms): the tiled sweep straight from the block-sparse operator for large
// levels, the small blocked Gauss-Jordan on the dense matrix otherwise.
// Returns true when the inverse is lagged (coarse_lag) and this operator's
// inverse is still to be computed beside the solve (lag_inverse).
bool coarse_inverse(uft_ctx *C, hipStream_t s) {
  AmgHier &H = C->amg;
  AmgLevel &Lc = H.L[H.nlev - 1];
  if (H.ncoarse >= UFT_SWEEP_MIN) {
    if (H.d_cinv2 && H.lag_spare) return true;
    // this operator's inverse before the solve (also the first refresh when lagged)
    uftk::dense_spd_sweep_bsr(s, H.ncoarse, Lc.nblk, Lc.d_brow, Lc.d_ci, Lc.d_val, H.d_cmat, H.d_cinv, H.d_cinvf);
    if (H.d_cinvh) uftk::dense_to_half_rows(s, H.ncoarse, H.d_cinvf, H.d_cinvh, H.d_cinvs);
    H.lag_spare = H.d_cinv2 != nullptr;
    return false;
  }
  uftk::bsr_to_dense(s, Lc.nunit, Lc.d_rp, Lc.d_ci, Lc.d_val, H.ncoarse, H.d_cmat);
  uftk::dense_spd_inverse(s, H.ncoarse, H.d_cmat, H.d_cinv, H.d_cinvf);
  if (H.d_cinvh) uftk::dense_to_half_rows(s, H.ncoarse, H.d_cinvf, H.d_cinvh, H.d_cinvs);
  return false;
}
// The lagged inverse of the current coarsest operator into the spare buffers
// on side stream 2, queued after everything on that stream (the operator;
// the last uses of the spare buffers are ordered before by the refresh's
// fork); nothing joins it into the main stream until the next refresh
// (amg_refresh_impl), so it runs beside the solve.
void lag_inverse(uft_ctx *C) {
  AmgHier &H = C->amg;
  AmgLevel &Lc = H.L[H.nlev - 1];
  uftk::dense_spd_sweep_bsr(C->side[2], H.ncoarse, Lc.nblk, Lc.d_brow, Lc.d_ci, Lc.d_val, H.d_cmat, H.d_cinv2,
                            H.d_cinvf2);
  if (H.d_cinvh2) uftk::dense_to_half_rows(C->side[2], H.ncoarse, H.d_cinvf2, H.d_cinvh2, H.d_cinvs2);
  (void)hipEventRecord(C->ev_inv, C->side[2]);
  H.lag_pending = true;
}

// Levels >= 1 after the level-1 operator: its post-processing (unit diagonal
// of empty rows, block-diagonal inverse, lower-precision copies), the coarser
// Galerkin products, the coarse levels' eigenvalue estimates and the coarsest
// inverse (on the side streams when par, one host round trip for all).
int refresh_coarse(uft_ctx *C, bool par, int ks, int first_est) {
  AmgHier &H = C->amg;
  hipStream_t s = C->s;
  for (int l = 1; l < H.nlev; ++l) {
    AmgLevel &Lv = H.L[l - 1], &Ln = H.L[l];
    if (l >= 2) uftk::galerkin_coarse(s, Ln.nblk, Lv.d_glptr, Lv.d_gl, Lv.d_brow, Lv.d_ci, Lv.d_P, Lv.d_val, Ln.d_val);
    uftk::fixed_unit_diag(s, Ln.nunit, Ln.d_rp, Ln.d_ci, Ln.d_val);
    uftk::bsr6_dinv(s, Ln.nunit, Ln.d_rp, Ln.d_ci, Ln.d_val, Ln.d_dinv, Ln.d_dinvf);
    if (Ln.d_valf) uftk::d2f(s, Ln.nblk * 36, Ln.d_val, Ln.d_valf);
    if (Ln.d_valh)
      uftk::bsr6_to_half(s, Ln.nblk, Ln.nunit, Ln.d_brow, Ln.d_ci, Ln.d_val, Ln.d_dinv, Ln.d_valh, Ln.d_dh);
  }
  if (!par) {
    if (coarse_inverse(C, s)) {
      UFT_HIP(C, hipEventRecord(C->ev_side[4], s));
      UFT_HIP(C, hipStreamWaitEvent(C->side[2], C->ev_side[4], 0));
      lag_inverse(C);
    }
    for (int l = first_est; l + 1 < H.nlev; ++l) {
      int rc = estimate_lmax(C, l, ks);
      if (rc) return rc;
    }
    return UFT_OK;
  }
  UFT_HIP(C, hipEventRecord(C->ev_side[4], s));
  for (int i = 1; i < 3; ++i) UFT_HIP(C, hipStreamWaitEvent(C->side[i], C->ev_side[4], 0));
  C->s = C->side[1];
  for (int l = 1; l + 1 < H.nlev; ++l) lanczos_enqueue(C, l, ks, C->d_part2, S_MISC + 8, C->d_lanczos + 192 * l);
  C->s = C->side[2];
  const bool lag = coarse_inverse(C, C->s);
  C->s = s;
  for (int i = 0; i < 3; ++i) {
    UFT_HIP(C, hipEventRecord(C->ev_side[i], C->side[i]));
    UFT_HIP(C, hipStreamWaitEvent(s, C->ev_side[i], 0));
  }
  if (lag) lag_inverse(C);  // after the join: the main stream does not wait for it
  std::vector<double> hl((size_t)192 * (H.nlev - 1));
  UFT_HIP(C, hipMemcpyAsync(hl.data(), C->d_lanczos, hl.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  if (int rc = sync(C)) return rc;
  for (int l = first_est; l + 1 < H.nlev; ++l) lanczos_finish(C, l, ks, hl.data() + 192 * l);
  return UFT_OK;
}

int amg_refresh_impl(uft_ctx *C, const double *xphys) {
  AmgHier &H = C->amg;
  hipStream_t s = C->s;
  if (H.lag_pending) {  // the previous operator's inverse (computed beside the last solve) becomes the cycle's
    UFT_HIP(C, hipStreamWaitEvent(s, C->ev_inv, 0));
    std::swap(H.d_cinv, H.d_cinv2);
    std::swap(H.d_cinvf, H.d_cinvf2);
    std::swap(H.d_cinvh, H.d_cinvh2);
    std::swap(H.d_cinvs, H.d_cinvs2);
    H.lag_pending = false;
  }
  C->cur_xphys = xphys;
  C->refreshes++;
  uftk::diag_K(s, C, xphys, H.L[0].d_dinv, H.L[0].d_dinvf);
  const bool par = C->side[0] && !C->bt.nseg;
  const int ks = std::min(64, std::max(2, C->P.esteig_steps));
  if (H.L[0].sa && C->bt.nseg) {
    // design batch with smoothed aggregation: the fine blocks, the fine
    // level's per-design lambda_max (folded into D^-1, so omega = 4 cheb_hi / 3
    // on the folded D^-1 is each design's 4 / (3 lambda_max)), the smoothed
    // prolongator and P^T A P, then the batch's coarse levels
    BatchState &B = C->bt;
    AmgLevel &L0 = H.L[0];
    uftk::simp_moduli(s, C, xphys, C->d_Ee);
    uftk::fine_ell_fill(s, C, C->d_Ee, L0.d_dinv, C->d_a64, C->d_a32);
    if (int rc = estimate_lmax_batch(C, 0, ks)) return rc;
    L0.sa_omega = 4.0 * C->P.cheb_hi / 3.0;
    uftk::sa_refresh(s, C);
    for (int l = 1; l < H.nlev; ++l) {
      AmgLevel &Lv = H.L[l - 1], &Ln = H.L[l];
      if (l >= 2)
        uftk::galerkin_coarse(s, Ln.nblk, Lv.d_glptr, Lv.d_gl, Lv.d_brow, Lv.d_ci, Lv.d_P, Lv.d_val, Ln.d_val);
      uftk::fixed_unit_diag(s, Ln.nunit, Ln.d_rp, Ln.d_ci, Ln.d_val);
      uftk::bsr6_dinv(s, Ln.nunit, Ln.d_rp, Ln.d_ci, Ln.d_val, Ln.d_dinv, Ln.d_dinvf);
      if (Ln.d_valf) uftk::d2f(s, Ln.nblk * 36, Ln.d_val, Ln.d_valf);
      if (Ln.d_valh)
        uftk::bsr6_to_half(s, Ln.nblk, Ln.nunit, Ln.d_brow, Ln.d_ci, Ln.d_val, Ln.d_dinv, Ln.d_valh, Ln.d_dh);
    }
    AmgLevel &Lc = H.L[H.nlev - 1];
    uftk::seg_bsr_to_dense(s, Lc.nunit, Lc.d_rp, Lc.d_ci, Lc.d_val, B.d_cdes, B.d_cu, B.d_coff, H.d_cmat);
    uftk::seg_dense_inverse(s, B.nseg, B.d_cu, B.d_coff, H.d_cmat, H.d_cinv);
    for (int l = 1; l + 1 < H.nlev; ++l) {
      int rc = estimate_lmax_batch(C, l, ks);
      if (rc) return rc;
    }
    return UFT_OK;
  }
  if (H.L[0].sa) {
    // smoothed aggregation: the fine blocks (FP64, and the mixed forms'
    // smoother and outer blocks), the fine level's lambda_max (omega needs
    // it), the smoothed prolongator and P^T A P, then the coarser levels
    // After the first refresh omega comes from the previous refresh's
    // estimate (lambda_max moves little between design iterations), so the
    // fine level's estimate runs on a side stream beside the products.
    AmgLevel &L0 = H.L[0];
    uftk::simp_moduli(s, C, xphys, C->d_Ee);
    uftk::fine_ell_fill(s, C, C->d_Ee, L0.d_dinv, C->d_a64, C->d_a32);
    const double hi = C->P.cheb_hi;
    if (par && L0.sa_omega > 0.0) {
      UFT_HIP(C, hipEventRecord(C->ev_side[3], s));
      UFT_HIP(C, hipStreamWaitEvent(C->side[0], C->ev_side[3], 0));
      C->s = C->side[0];
      lanczos_enqueue(C, 0, ks, C->d_part, S_MISC, C->d_lanczos);
      C->s = s;
      uftk::sa_refresh(s, C);
      if (int rc = refresh_coarse(C, par, ks, 0)) return rc;
    } else {
      if (int rc = estimate_lmax(C, 0, ks)) return rc;
      L0.sa_omega = 4.0 / (3.0 * (L0.lmax / hi));
      uftk::sa_refresh(s, C);
      if (int rc = refresh_coarse(C, par, ks, 1)) return rc;
    }
    L0.sa_omega = 4.0 / (3.0 * (L0.lmax / hi));  // for the next refresh
    return UFT_OK;
  }
  // Concurrent branches (each on its own stream; each does what the
  // sequential order does, so the results are the same): the fine blocks and
  // the fine level's eigenvalue estimate once E(x_e) and the fine diagonal
  // exist; after the Galerkin products the coarse levels' estimates (own
  // partials and scalar slots) and the coarsest inverse.
  if (par) {
    uftk::simp_moduli(s, C, xphys, C->d_Ee);
    UFT_HIP(C, hipEventRecord(C->ev_side[3], s));
    UFT_HIP(C, hipStreamWaitEvent(C->side[0], C->ev_side[3], 0));
    C->s = C->side[0];
    if (fine_blocks(C)) uftk::fine_ell_fill(C->s, C, C->d_Ee, H.L[0].d_dinv);
    if (H.nlev > 1) lanczos_enqueue(C, 0, ks, C->d_part, S_MISC, C->d_lanczos);
    C->s = s;
  }
  if (H.nlev > 1) {
    AmgLevel &Lv = H.L[0], &Ln = H.L[1];
    uftk::galerkin_fine(s, C, par ? nullptr : xphys, Ln.nblk, Lv.d_glptr, Lv.d_gl, Lv.d_gtr, Ln.d_brow, Ln.d_ci,
                        Lv.d_G, C->d_Ee, Ln.d_val);
  }
  if (C->bt.nseg) {
    BatchState &B = C->bt;
    for (int l = 1; l < H.nlev; ++l) {
      AmgLevel &Lv = H.L[l - 1], &Ln = H.L[l];
      if (l >= 2)
        uftk::galerkin_coarse(s, Ln.nblk, Lv.d_glptr, Lv.d_gl, Lv.d_brow, Lv.d_ci, Lv.d_P, Lv.d_val, Ln.d_val);
      uftk::fixed_unit_diag(s, Ln.nunit, Ln.d_rp, Ln.d_ci, Ln.d_val);
      uftk::bsr6_dinv(s, Ln.nunit, Ln.d_rp, Ln.d_ci, Ln.d_val, Ln.d_dinv, Ln.d_dinvf);
      if (Ln.d_valf) uftk::d2f(s, Ln.nblk * 36, Ln.d_val, Ln.d_valf);
      if (Ln.d_valh)
        uftk::bsr6_to_half(s, Ln.nblk, Ln.nunit, Ln.d_brow, Ln.d_ci, Ln.d_val, Ln.d_dinv, Ln.d_valh, Ln.d_dh);
    }
    // the level 0 -> 1 Galerkin product has written E(x_e) to d_Ee
    if (fine_blocks(C)) uftk::fine_ell_fill(s, C, C->d_Ee, H.L[0].d_dinv);
    AmgLevel &Lc = H.L[H.nlev - 1];
    uftk::seg_bsr_to_dense(s, Lc.nunit, Lc.d_rp, Lc.d_ci, Lc.d_val, B.d_cdes, B.d_cu, B.d_coff, H.d_cmat);
    uftk::seg_dense_inverse(s, B.nseg, B.d_cu, B.d_coff, H.d_cmat, H.d_cinv);
    for (int l = 0; l + 1 < H.nlev; ++l) {
      int rc = estimate_lmax_batch(C, l, std::min(64, std::max(2, C->P.esteig_steps)));
      if (rc) return rc;
    }
    return UFT_OK;
  }
  // the level 0 -> 1 Galerkin product has written E(x_e) to d_Ee
  if (fine_blocks(C) && !par) uftk::fine_ell_fill(s, C, C->d_Ee, H.L[0].d_dinv);
  return refresh_coarse(C, par, ks, 0);
}

// AMG-preconditioned CG on K(x) U = F from U = 0; converged when
// ||F - K U|| <= rtol ||F|| (the reference's FGMRES uses right
// preconditioning, i.e. the unpreconditioned residual, SolveFem.cpp:126-134).
int solve_el(uft_ctx *C, const double *xphys, double *U) {
  hipStream_t s = C->s;
  const int64_t n = C->ndof;
  double *r = C->d_r, *z = C->d_z, *p = C->d_p, *q = C->d_q, *zo = C->d_w;
  const double t0 = now_ms();
  int rc = amg_refresh(C, xphys);
  if (rc) return rc;
  C->t_phase[0] = now_ms() - t0;
  uftk::fill(s, n, 0.0, U);
  uftk::copy(s, n, C->d_F, r);
  uftk::dot(s, n, r, r, C->d_part, C->d_scal + S_RR);
  rc = fetch(C, C->d_scal + S_RR, 1, C->h_scal);
  if (rc) return rc;
  const double bn = std::sqrt(C->h_scal[0]);
  C->fnorm = bn;
  C->it_el = 0;
  C->outer_el = 1;
  if (bn == 0.0) {
    C->relres_el = 0.0;
    return UFT_OK;
  }
  vcycle(C, 0, r, z);
  uftk::dot(s, n, r, z, C->d_part, C->d_scal + S_RZ);
  uftk::copy(s, n, z, p);
  const double tol = C->P.rtol_el * bn;
  int it = 0;
  bool conv = false;
  auto head = [&]() {
    fine_apply(C, xphys, p, q, p, C->d_scal + S_PQ);
    uftk::sc_div(s, C->d_scal, S_RZ, S_PQ, S_ALPHA);
    uftk::cg_update(s, n, C->d_scal, S_ALPHA, U, p, r, q, C->d_part, C->d_scal + S_RR);
  };
  // flexible CG (Polak-Ribiere beta): the K-cycle preconditioner contains
  // inner products, so it is not a fixed linear operator
  auto next = [&]() {
    uftk::copy(s, n, z, zo);  // the previous preconditioned residual
    vcycle(C, 0, r, z);
    uftk::dot_pr_beta(s, n, r, z, zo, C->d_part, C->d_scal, S_RZNEW, S_RZ, S_BETA + 2);
    uftk::xpby(s, n, z, C->d_scal, S_BETA + 2, p);
    head();
  };
  // true residual F - K U into q; returns its norm in *tn
  auto true_residual = [&](double *tn) -> int {
    fine_apply(C, xphys, U, q, nullptr, nullptr);
    uftk::sub(s, n, C->d_F, q, q);
    uftk::dot(s, n, q, q, C->d_part, C->d_scal + S_MISC);
    const int e = fetch(C, C->d_scal + S_MISC, 1, C->h_scal);
    *tn = std::sqrt(C->h_scal[0]);
    return e;
  };
  static const char tag = 0;
  bool restart = false;
  double tn = bn;
  while (it < C->P.maxit_el) {
    if (it == 0) {
      head();
    } else if (restart) {
      // the recursive residual drifted from the true one: CG again from r = F - K U
      uftk::copy(s, n, q, r);
      vcycle(C, 0, r, z);
      uftk::dot(s, n, r, z, C->d_part, C->d_scal + S_RZ);
      uftk::copy(s, n, z, p);
      head();
      restart = false;
    } else if (!use_graph(C)) {
      next();
    } else if ((rc = graph_launch(C, xphys, U, &tag, next))) {
      return rc;
    }
    ++it;
    rc = fetch(C, C->d_scal + S_RR, 1, C->h_scal);
    if (rc) return rc;
    const double rn = std::sqrt(C->h_scal[0]);
    if (!std::isfinite(rn)) return uft_set_error(C, UFT_ERR_NOCONV, "elasticity CG diverged at iteration %d", it);
    if (rn <= tol) {
      // stop on the true residual (the reference's FGMRES is right-preconditioned, so its
      // residual is the unpreconditioned true one, SolveFem.cpp:126-152)
      if ((rc = true_residual(&tn))) return rc;
      if (tn <= tol) {
        conv = true;
        break;
      }
      restart = true;
    }
  }
  C->it_el = it;
  // true residual (SolveFem.cpp:142-152)
  if (!conv && (rc = true_residual(&tn))) return rc;
  C->relres_el = tn / bn;
  if (!conv)
    return uft_set_error(C, UFT_ERR_NOCONV, "elasticity solve: %d iterations, relative residual %.3e > %.3e", it,
                         C->relres_el, C->P.rtol_el);
  return UFT_OK;
}

// Mixed form: FP64 flexible CG (Polak-Ribiere beta, which tolerates a
// preconditioner that is not exactly linear) whose residual, inner products
// and iterate stay FP64; the preconditioner is one V-cycle on FP32 vectors
// with FP32 element blocks and FP32 coarse operators, coarsest solve FP64.
int solve_el_mixed(uft_ctx *C, const double *xphys, double *U) {
  hipStream_t s = C->s;
  const int64_t n = C->ndof;
  double *r = C->d_r, *z = C->d_z, *p = C->d_p, *q = C->d_q, *zo = C->d_w;
  float *rf = C->d_rf, *zf = C->d_zf;
  const double t0 = now_ms();
  int rc = amg_refresh(C, xphys);
  if (rc) return rc;
  C->t_phase[0] = now_ms() - t0;
  uftk::fill(s, n, 0.0, U);
  uftk::copy(s, n, C->d_F, r);
  uftk::dot(s, n, r, r, C->d_part, C->d_scal + S_RR);
  if ((rc = fetch(C, C->d_scal + S_RR, 1, C->h_scal))) return rc;
  const double bn = std::sqrt(C->h_scal[0]);
  C->fnorm = bn;
  C->it_el = 0;
  C->outer_el = 1;
  if (bn == 0.0) {
    C->relres_el = 0.0;
    return UFT_OK;
  }
  auto precond = [&]() {
    uftk::d2f(s, n, r, rf);
    vcycle_f(C, 0, rf, zf);
    uftk::f2d(s, n, zf, z);
  };
  precond();
  uftk::dot(s, n, r, z, C->d_part, C->d_scal + S_RZ);
  uftk::copy(s, n, z, p);
  const double tol = C->P.rtol_el * bn;
  // The outer operator product runs with the FP32 blocks; the FP64 residual
  // F - K U replaces the recursively updated one whenever the latter has
  // fallen 100-fold below the last FP64 residual or below the tolerance, and
  // the solve stops only on an FP64 residual (reliable updates: the
  // recursion's drift stays 1e-5 of the current residual, far below the
  // tolerance).  The update also forms the next cycle's first Chebyshev step
  // on rf, and so does a reliable update.
  AmgLevel &L0 = C->amg.L[0];
  uftk::d2f(s, n, p, C->d_pf);  // the first direction's FP32 copy
  // Device control: the stopping test and the reliable-update decision run
  // on the device after every iteration (uftk::ksp_ctl_iter), the reliable
  // update's launches are queued every iteration and return at once unless
  // asked for, and the update that changes U and r returns at once after the
  // solve has stopped.  The host queues iterations in rounds (as many as the
  // previous solve took, less one, then two at a time) and reads the control
  // scalars once per round, so the device never waits for the host between
  // iterations; the iterations queued after the stop leave U and r alone.
  // With graph = 1 one iteration is captured and replayed.
  double *ctl = C->d_scal + S_CTL;
  uftk::ksp_ctl_init(s, ctl, bn, tol, C->P.maxit_el);
  C->reliable = 0;
  const double itc = fine_first_coef(C);
  auto control = [&]() {
    uftk::ksp_ctl_iter(s, C->d_scal + S_RR, ctl);
    uftk::reliable_update(s, C, xphys, U, r, rf, L0.d_dinvf, itc, L0.rf, L0.df, zf, C->d_scal + S_RR, ctl);
  };
  auto head_dev = [&]() {  // q = K p (FP32, d_pf = p); alpha; U, r, rf and the next cycle's first step
    fine_apply_lo(C, xphys, p, S_PQ, true);
    uftk::cg_update_lo_cheb(s, n, C->d_scal, S_RZ, S_PQ, U, p, r, C->d_qf, rf, C->d_part, C->d_scal + S_RR,
                            L0.d_dinvf, itc, L0.rf, L0.df, zf, ctl);
  };
  const bool graph = use_graph(C);
  auto next_dev = [&]() {  // z = B rf; zo = previous z; beta; p; then the update and the control
    fine_cycle<float>(C, rf, zf, true);
    if (graph) {  // a captured iteration must not depend on the host's swap of z and zo
      uftk::z_from_f32_beta(s, n, zf, r, z, zo, C->d_part, C->d_scal, S_RZNEW, S_RZ, S_BETA + 2);
    } else {
      uftk::z_from_f32_swap_beta(s, n, zf, r, z, zo, C->d_part, C->d_scal, S_RZNEW, S_RZ, S_BETA + 2);
      std::swap(z, zo);
    }
    uftk::xpby_lo(s, n, z, C->d_scal, S_BETA + 2, p, C->d_pf);
    head_dev();
    control();
  };
  static const char tag = 0;
  // traffic counters after each queued iteration (the iterations queued after
  // the stop are taken back; reliable updates are counted as they ran)
  std::vector<std::pair<double, int64_t>> cnt;
  cnt.reserve(256);
  int queued = 0, round = C->el_prev_its > 1 ? C->el_prev_its - 1 : 8;
  double stop = 0.0, its = 0.0, rel = 0.0;
  while (true) {
    for (int k = 0; k < round && queued < C->P.maxit_el; ++k, ++queued) {
      if (queued == 0) {
        head_dev();
        control();
      } else if (!graph) {
        next_dev();
      } else if ((rc = graph_launch(C, xphys, U, &tag, next_dev))) {
        return rc;
      }
      cnt.emplace_back(C->solve_bytes, C->solve_applies);
    }
    if ((rc = fetch(C, ctl, 5, C->h_scal))) return rc;
    stop = C->h_scal[0];
    its = C->h_scal[2];
    rel = C->h_scal[3];
    if (stop != 0.0 || queued >= C->P.maxit_el) break;
    round = 2;
  }
  const int it = (int)its;
  C->it_el = it;
  C->el_prev_its = it;
  C->reliable = (int)rel;
  C->outer_el = C->reliable;  // FP64 residual evaluations
  if (C->count_on && it >= 1 && it <= (int)cnt.size()) {
    C->solve_bytes = cnt[it - 1].first + rel * fine_bytes64(C);
    C->solve_applies = cnt[it - 1].second + (int64_t)rel;
  }
  if (stop == 2.0) return uft_set_error(C, UFT_ERR_NOCONV, "mixed CG diverged at iteration %d", it);
  if (stop == 1.0) {  // the last reliable update is the true residual (SolveFem.cpp:142-152)
    C->relres_el = C->h_scal[4] / bn;
    return UFT_OK;
  }
  fine_apply(C, xphys, U, q, nullptr, nullptr);
  uftk::sub(s, n, C->d_F, q, q);
  uftk::dot(s, n, q, q, C->d_part, C->d_scal + S_MISC);
  if ((rc = fetch(C, C->d_scal + S_MISC, 1, C->h_scal))) return rc;
  C->relres_el = std::sqrt(C->h_scal[0]) / bn;
  return uft_set_error(C, UFT_ERR_NOCONV, "mixed elasticity solve: %d iterations, relative residual %.3e", it,
                       C->relres_el);
}

// Multi-load-case solve: one flexible CG per right-hand side, all advanced
// together (FP64 iterate, residual and inner products; per-RHS scalars); the
// cycle runs in FP64 (UFT_FP64) or FP32 with matrix-core element products
// (mixed forms).  A right-hand side stops updating once
// ||F_k - K U_k|| <= rtol ||F_k||, checked on the FP64 residual (which then
// replaces the recursive one) when the recursive one reaches the tolerance;
// the loop ends when all have.  The mixed forms' outer product runs in FP32,
// with per-case reliable updates as in the single-case mixed solve: a case's
// residual is also replaced by its FP64 residual once the recursive one has
// fallen 100-fold below the last FP64 one.
enum { MS_RZ = 0, MS_PQ = 8, MS_ALPHA = 16, MS_RR = 24, MS_RZN = 32, MS_BETA = 48, MS_FN = 56, MS_TR = 64 };
int solve_el_block(uft_ctx *C, const double *xphys) {
  hipStream_t s = C->s;
  MrState &M = C->mr;
  const int nr = M.nr;
  const int64_t n = C->ndof;
  double *sc = M.sc;
  const double t0 = now_ms();
  int rc = amg_refresh(C, xphys);
  if (rc) return rc;
  C->t_phase[0] = now_ms() - t0;
  uftk::fill(s, n * nr, 0.0, M.U);
  uftk::copy(s, n * nr, M.F, M.r);
  uftk::mr_dot(s, nr, n, M.F, M.F, C->d_part, sc + MS_FN);
  if ((rc = fetch(C, sc + MS_FN, nr, C->h_scal))) return rc;
  double fn[8], tol[8];
  unsigned mask = 0;
  for (int k = 0; k < nr; ++k) {
    fn[k] = std::sqrt(C->h_scal[k]);
    tol[k] = C->P.rtol_el * fn[k];
    M.it[k] = 0;
    M.relres[k] = 0.0;
    if (fn[k] > 0.0) mask |= 1u << k;
  }
  const bool f32 = is_mixed(C);
  // z = B r; in the mixed forms the cycle reads the FP32 copy M.rf that the
  // update (or the reliable update) wrote
  auto precond = [&](double *z) {
    if (f32) {
      mr_fine_cycle<float>(C, M.rf, M.zf);
      uftk::f2d(s, n * nr, M.zf, z);
    } else {
      mr_fine_cycle<double>(C, M.r, z);
    }
  };
  double *z = M.z, *zo = M.zo;
  int it = 0;
  if (mask) {
    if (f32) uftk::d2f(s, n * nr, M.r, M.rf);
    precond(z);
    uftk::mr_dot(s, nr, n, M.r, z, C->d_part, sc + MS_RZ);
    uftk::copy(s, n * nr, z, M.p);
    if (f32) uftk::d2f(s, n * nr, M.p, M.pf);
  }
  RowOp<double> pq;
  pq.mode = ROW_STOREDOT;
  pq.y = M.q;
  pq.part = C->d_part;
  RowOp<double> tr;  // FP64 residuals F - K U into M.q
  tr.mode = ROW_RES;
  tr.base = M.F;
  tr.r = M.q;
  double rtrue[8];   // mixed: last FP64 residual norm per case
  bool have_true = false;
  for (int k = 0; k < nr; ++k) rtrue[k] = fn[k];
  C->reliable = 0;
  while (mask && it < C->P.maxit_el) {
    if (f32) {  // q = K p (FP32 blocks), p.q per right-hand side
      mr_outer_apply_lo(C, M.pf, M.qf);
    } else {    // q = K p and p.q per right-hand side
      mr_lev_op<double>(C, 0, M.p, pq);
    }
    uftk::final_sum(s, C->d_part, f32 && C->fell.d_w8 ? uftk::mr_ell_blocks(C, M.pf, ROW_STOREDOT, true) : mr_lev_blocks(C, 0),
                    nr, sc + MS_PQ);
    uftk::mr_alpha(s, nr, mask, sc + MS_RZ, sc + MS_PQ, sc + MS_ALPHA);
    if (f32)
      uftk::mr_cg_update_lo(s, nr, n, sc + MS_ALPHA, M.U, M.p, M.r, M.qf, M.rf, C->d_part, sc + MS_RR);
    else
      uftk::mr_cg_update(s, nr, n, sc + MS_ALPHA, M.U, M.p, M.r, M.q, C->d_part, sc + MS_RR);
    ++it;
    if ((rc = fetch(C, sc + MS_RR, nr, C->h_scal))) return rc;
    unsigned need = 0;
    for (int k = 0; k < nr; ++k) {
      if (!((mask >> k) & 1u)) continue;
      const double rn = std::sqrt(C->h_scal[k]);
      if (!std::isfinite(rn)) return uft_set_error(C, UFT_ERR_NOCONV, "load case %d diverged at iteration %d", k, it);
      M.it[k] = it;
      // a case stops on its FP64 residual F_k - K U_k; the FP32 outer
      // product also replaces the recursive residual after 100-fold drops
      if (rn <= tol[k] || (f32 && rn < 1e-2 * rtrue[k])) need |= 1u << k;
    }
    // the FP64 pass computes every case's residual at once: when one case
    // asks for it, every active case takes its FP64 residual (a reliable
    // update is valid at any iteration), which keeps the cases' updates in
    // step instead of one pass per case
    if (need && f32) need |= mask;
    if (need) {  // reliable update of the cases in need
      mr_lev_op<double>(C, 0, M.U, tr);
      uftk::mr_take_res(s, nr, n, need, M.q, M.r, M.rf, C->d_part, sc + MS_RR);
      if ((rc = fetch(C, sc + MS_RR, nr, C->h_scal))) return rc;
      C->reliable++;
      for (int k = 0; k < nr; ++k) {
        if (!((need >> k) & 1u)) continue;
        rtrue[k] = std::sqrt(C->h_scal[k]);
        if (rtrue[k] <= tol[k]) mask &= ~(1u << k);
      }
      have_true = true;
    }
    if (!mask) break;
    std::swap(z, zo);
    precond(z);
    uftk::mr_dot_pr(s, nr, n, M.r, z, zo, C->d_part, sc + MS_RZN);
    uftk::mr_beta(s, nr, mask, sc + MS_RZN, sc + MS_RZ, sc + MS_BETA);
    if (f32)
      uftk::mr_xpby_lo(s, nr, n, z, sc + MS_BETA, M.p, M.pf);
    else
      uftk::mr_xpby(s, nr, n, z, sc + MS_BETA, M.p);
  }
  C->it_el = it;
  C->outer_el = f32 ? C->reliable : 1;
  double worst = 0.0;
  if (!mask && have_true) {
    // every case stopped on its FP64 residual, U unchanged since
    for (int k = 0; k < nr; ++k) {
      M.relres[k] = fn[k] > 0.0 ? rtrue[k] / fn[k] : 0.0;
      worst = std::max(worst, M.relres[k]);
    }
  } else {
    // true residuals F - K U per right-hand side (SolveFem.cpp:142-152)
    mr_lev_op<double>(C, 0, M.U, tr);
    uftk::mr_dot(s, nr, n, M.q, M.q, C->d_part, sc + MS_TR);
    if ((rc = fetch(C, sc + MS_TR, nr, C->h_scal))) return rc;
    for (int k = 0; k < nr; ++k) {
      M.relres[k] = fn[k] > 0.0 ? std::sqrt(C->h_scal[k]) / fn[k] : 0.0;
      worst = std::max(worst, M.relres[k]);
    }
  }
  C->relres_el = worst;
  C->fnorm = fn[0];
  // case 0 is the context's U (for the VTU writer and uft_vec_get)
  uftk::mr_extract(s, n, nr, 0, M.U, C->vec[UFT_U]);
  if (mask)
    return uft_set_error(C, UFT_ERR_NOCONV, "load cases: %d iterations, worst relative residual %.3e", it, worst);
  return UFT_OK;
}

// fetch every design's scalars; true when some design has flag `slot` set
int fetch_batch(uft_ctx *C, int slot, bool *any) {
  BatchState &B = C->bt;
  int rc = fetch(C, B.sc, B.nseg * UFT_BSC, B.h);
  if (rc) return rc;
  bool a = false;
  for (int d = 0; d < B.nseg; ++d) a |= B.h[(size_t)d * UFT_BSC + slot] != 0.0;
  if (any) *any = a;
  return UFT_OK;
}

// Design batches: one flexible CG per design (the single-design loop of
// solve_el / solve_el_mixed with per-design scalars, stopping tests on the
// device), all advanced together; every operator apply and every cycle
// covers all designs.  A design stops updating once
// ||F_d - K U_d|| <= rtol ||F_d||; the loop ends when all have.
int solve_el_batch(uft_ctx *C, const double *xphys, double *U) {
  hipStream_t s = C->s;
  BatchState &B = C->bt;
  const int ns = B.nseg;
  const SegG g = B.D.g();
  const int32_t *df = B.D.d_dfirst;
  const int64_t n = C->ndof;
  double *r = C->d_r, *z = C->d_z, *p = C->d_p, *q = C->d_q, *zo = C->d_w;
  const double t0 = now_ms();
  int rc = amg_refresh(C, xphys);
  if (rc) return rc;
  C->t_phase[0] = now_ms() - t0;
  uftk::fill(s, n, 0.0, U);
  uftk::copy(s, n, C->d_F, r);
  uftk::seg_dot(s, g, df, ns, n, r, r, C->d_part, B.sc, B_RR);
  uftk::bsc_el_init(s, ns, B.sc, C->P.rtol_el);
  C->reliable = 0;
  const bool f32 = is_mixed(C);
  auto precond = [&]() {
    if (f32) {
      uftk::d2f(s, n, r, C->d_rf);
      vcycle_f(C, 0, C->d_rf, C->d_zf);
      uftk::f2d(s, n, C->d_zf, z);
    } else {
      vcycle(C, 0, r, z);
    }
  };
  bool any = false;
  if ((rc = fetch_batch(C, B_ACT, &any))) return rc;
  int it = 0;
  double *flags = C->d_scal + S_BFL;
  uftk::fill(s, 3, 0.0, flags);
  if (any) {
    precond();
    uftk::seg_dot(s, g, df, ns, n, r, z, C->d_part, B.sc, B_RZ);
    uftk::copy(s, n, z, p);
  }
  // mixed forms: the outer product with the FP32 element blocks and FP64
  // residuals at reliable updates (as solve_el_mixed), the update decisions
  // per design on the device; one FP64 residual pass serves all designs
  auto head = [&]() {
    if (f32) {
      fine_apply_lo(C, xphys, p, -1);
      uftk::seg_dot_lo(s, g, df, ns, n, p, C->d_qf, C->d_part, B.sc, B_PQ);
      uftk::bsc_el_alpha(s, ns, B.sc);
      uftk::seg_cg_update_lo(s, g, df, ns, n, B.sc, B_ALPHA, U, p, r, C->d_qf, C->d_part, B_RR);
      uftk::bsc_el_check_lo(s, ns, B.sc, flags);
    } else {
      fine_apply(C, xphys, p, q, nullptr, nullptr);
      uftk::seg_dot(s, g, df, ns, n, p, q, C->d_part, B.sc, B_PQ);
      uftk::bsc_el_alpha(s, ns, B.sc);
      uftk::seg_cg_update(s, g, df, ns, n, B.sc, B_ALPHA, U, p, r, q, C->d_part, B_RR);
      uftk::bsc_el_check(s, ns, B.sc, flags);
    }
  };
  auto next = [&]() {
    uftk::copy(s, n, z, zo);  // the previous preconditioned residual
    precond();
    uftk::seg_dot_pr(s, g, df, ns, n, r, z, zo, C->d_part, B.sc, B_RZN);
    uftk::bsc_el_beta(s, ns, B.sc);
    uftk::seg_xpby(s, g, n, z, B.sc, B_BETA, p);
    head();
  };
  // FP64 residual of the designs that asked for it (FP64 form: the stopping
  // test; mixed forms: a reliable update), launched after every iteration and
  // gated on the device (flags[BF_NEED]): an idle pass returns at once
  auto true_pass = [&]() {
    uftk::apply_K_gated(s, C, xphys, U, q, flags + BF_NEED, B.sc + B_NEED);
    uftk::seg_true_res(s, g, df, ns, n, B.sc, C->d_F, q, r, C->d_part, B_RR, flags + BF_NEED);
    uftk::bsc_el_true(s, ns, B.sc, flags);
  };
  // Device control: every test runs on the per-design scalars, so the host
  // queues the iterations in rounds (as many as the last solve took, less
  // one, then two at a time) and reads the batch flags once per round; an
  // iteration queued after every design has stopped changes no design's U or
  // r (their step lengths are zero).
  static const char tag = 0;
  int round = C->el_prev_its > 1 ? C->el_prev_its - 1 : 8;
  while (any && it < C->P.maxit_el) {
    const int k = std::min(round, C->P.maxit_el - it);
    for (int j = 0; j < k; ++j) {
      if (it == 0)
        head();
      else if (!use_graph(C))
        next();
      else if ((rc = graph_launch(C, xphys, U, &tag, next)))
        return rc;
      ++it;
      true_pass();
    }
    if ((rc = fetch(C, flags, 3, C->h_scal))) return rc;
    any = C->h_scal[BF_ACT] != 0.0;
    C->reliable = (int)C->h_scal[BF_NRES];
    round = 2;
  }
  C->outer_el = f32 ? C->reliable : 1;
  // true residuals per design (SolveFem.cpp:142-152)
  fine_apply(C, xphys, U, q, nullptr, nullptr);
  uftk::sub(s, n, C->d_F, q, q);
  uftk::seg_dot(s, g, df, ns, n, q, q, C->d_part, B.sc, B_TR);
  if ((rc = fetch_batch(C, B_ACT, nullptr))) return rc;
  double worst = 0.0;
  int bad = -1;
  C->it_el = 0;
  for (int d = 0; d < ns; ++d) {
    const double *S = B.h + (size_t)d * UFT_BSC;
    const double fn = S[B_FN];
    B.relres[d] = fn > 0.0 ? std::sqrt(S[B_TR]) / fn : 0.0;
    B.it[d] = (int)S[B_IT];
    C->it_el = std::max(C->it_el, B.it[d]);
    // as for one design: a design fails only when its stopping test never
    // passed (it ran to maxit_el with its residual above the tolerance)
    if (B.it[d] >= C->P.maxit_el && !(B.relres[d] <= C->P.rtol_el) && bad < 0) bad = d;
    if (!(B.relres[d] <= worst)) worst = B.relres[d];
  }
  C->el_prev_its = C->it_el;
  C->relres_el = worst;
  C->fnorm = B.h[B_FN];
  if (bad >= 0)
    return uft_set_error(C, UFT_ERR_NOCONV, "design %d: %d iterations, relative residual %.3e > %.3e", bad, B.it[bad],
                         B.relres[bad], C->P.rtol_el);
  return UFT_OK;
}

// Design batches: the filter's Jacobi-CG (solve_filter) per design
// A finished design's scalars freeze its iterate (alpha 0, its iteration
// count kept), so the host queues the iterations in rounds as in
// solve_filter (the last call's largest count, then two at a time) and reads
// the designs' states once per round: the same iterates and counts as a read
// after every iteration.
int solve_filter_batch(uft_ctx *C, const double *b, double *x, int slot) {
  hipStream_t s = C->s;
  BatchState &B = C->bt;
  const int ns = B.nseg;
  const SegG g = B.N.g();
  const int32_t *df = B.N.d_dfirst;
  const int64_t n = C->nn;
  double *r = C->d_fr, *z = C->d_fz, *p = C->d_fp, *q = C->d_fq;
  uftk::apply_KF(s, C, x, q);
  uftk::seg_fcg_init(s, g, df, ns, n, b, q, C->d_fdinv, r, z, C->d_part, B.sc, B_FB);
  uftk::bsc_f_init(s, ns, B.sc, C->P.rtol_f);
  bool any = false;
  int rc = fetch_batch(C, B_FACT, &any);
  if (rc) return rc;
  int it = 0;
  int *pred = &C->f_pred[std::min(std::max(slot, 0), UFT_MAXM + 1)];
  int round = *pred > 0 ? *pred : 8;
  while (any && it < C->P.maxit_f) {
    const int upto = std::min(C->P.maxit_f, it + round);
    for (; it < upto; ++it) {
      uftk::seg_fcg_p(s, g, n, z, B.sc, B_FPC, it == 0, p);
      uftk::apply_KF(s, C, p, q);
      uftk::seg_dot(s, g, df, ns, n, p, q, C->d_part, B.sc, B_FPQ);
      uftk::bsc_f_alpha(s, ns, B.sc);
      uftk::seg_fcg_update(s, g, df, ns, n, B.sc, B_FAL, x, p, r, q, C->d_fdinv, z, C->d_part, B_FZZN);
      uftk::bsc_f_after(s, ns, B.sc);
    }
    if ((rc = fetch_batch(C, B_FACT, &any))) return rc;
    round = 2;
  }
  int itmax = 0;
  for (int d = 0; d < ns; ++d) {
    const double *S = B.h + (size_t)d * UFT_BSC;
    B.itf[d] = (int)S[B_FIT];
    itmax = std::max(itmax, B.itf[d]);
    if (S[B_FDIV] != 0.0) return uft_set_error(C, UFT_ERR_NOCONV, "design %d: filter CG diverged", d);
    if (S[B_FACT] != 0.0)
      return uft_set_error(C, UFT_ERR_NOCONV, "design %d: filter CG, %d iterations without convergence", d, it);
  }
  C->it_f = itmax;
  *pred = itmax;
  return UFT_OK;
}

// Jacobi-CG on KF RX = TX with the reference's KSPCG settings
// (PDEFilter.cpp:117-137): warm start from the previous RX, preconditioned
// residual norm relative to ||D^-1 b||, atol 1e-50, dtol 1e5.
// With the filter operator assembled the tests run on the device
// (fcg_start / fcg_update_ctl) and the host queues a solve's iterations in
// rounds: as many as the same solve (slot: 0 the density filter, 1 + j the
// j-th sensitivity) took last time, then two at a time; a queued iteration
// after the stop does nothing.  The iterates, the iteration count and the
// tests are those of the loop below, which tests on the host after every
// iteration.
int solve_filter(uft_ctx *C, const double *b, double *x, int slot) {
  if (C->bt.nseg) return solve_filter_batch(C, b, x, slot);
  hipStream_t s = C->s;
  const int64_t n = C->nn;
  double *r = C->d_fr, *z = C->d_fz, *p = C->d_fp, *q = C->d_fq;
  double *sc = C->d_scal;
  uftk::apply_KF(s, C, x, q);
  uftk::fcg_init(s, n, b, q, C->d_fdinv, r, z, C->d_part, sc + S_F0);  // 8: |Db|^2 9: |z|^2 10: z.r
  if (C->fell.d_hv) {
    uftk::fcg_start(s, sc, S_F0, S_FD, C->P.rtol_f);
    const int maxit = C->P.maxit_f;
    int *pred = &C->f_pred[std::min(std::max(slot, 0), UFT_MAXM + 1)];
    int queued = 0, round = *pred > 0 ? *pred : 8;
    C->it_f = 0;
    for (;;) {
      const int upto = std::min(maxit, queued + round);
      for (; queued < upto; ++queued) {
        uftk::helm_ell_cg(s, C, z, sc, S_F0 + 2, S_F0 + 3, queued == 0, sc + S_FD, p, q, sc + S_F0 + 4);
        uftk::fcg_update_ctl(s, n, sc, S_F0 + 2, S_F0 + 4, x, p, r, q, C->d_fdinv, z, C->d_part, S_F0 + 6,
                             S_F0 + 2, S_F0 + 3, S_FD);
      }
      const int rc = fetch(C, sc + S_FD, 5, C->h_scal);
      if (rc) return rc;
      if (C->h_scal[0] != 0.0 || queued >= maxit) break;
      round = 2;
    }
    const int it = (int)C->h_scal[3];
    const double state = C->h_scal[0];
    C->it_f = it;
    *pred = it;
    if (state == 2.0) return uft_set_error(C, UFT_ERR_NOCONV, "filter CG diverged at iteration %d", it);
    if (state == 0.0) return uft_set_error(C, UFT_ERR_NOCONV, "filter CG: %d iterations without convergence", it);
    return UFT_OK;
  }
  int rc = fetch(C, sc + S_F0, 2, C->h_scal);
  if (rc) return rc;
  double snorm = std::sqrt(C->h_scal[0]);
  double dp = std::sqrt(C->h_scal[1]);
  if (snorm == 0.0) snorm = dp;
  const double ttol = std::max(C->P.rtol_f * snorm, 1e-50);
  int it = 0;
  C->it_f = 0;
  if (dp <= ttol) return UFT_OK;
  while (it < C->P.maxit_f) {
    uftk::fcg_p(s, n, z, sc, S_F0 + 2, S_F0 + 3, it == 0, p);
    uftk::apply_KF(s, C, p, q, sc + S_F0 + 4);  // q = KF p and p.q
    // alpha = beta / p.q; x, r, z updated; ||z||^2; beta_old = beta; beta = z.r
    uftk::fcg_update_ab(s, n, sc, S_F0 + 2, S_F0 + 4, x, p, r, q, C->d_fdinv, z, C->d_part, S_F0 + 6, S_F0 + 2,
                        S_F0 + 3);
    ++it;
    rc = fetch(C, sc + S_F0 + 6, 1, C->h_scal);
    if (rc) return rc;
    dp = std::sqrt(C->h_scal[0]);
    if (dp <= ttol) break;
    if (!(dp < 1e5 * snorm)) {
      C->it_f = it;
      return uft_set_error(C, UFT_ERR_NOCONV, "filter CG diverged at iteration %d", it);
    }
  }
  C->it_f = it;
  if (it >= C->P.maxit_f && dp > ttol)
    return uft_set_error(C, UFT_ERR_NOCONV, "filter CG: %d iterations without convergence", it);
  return UFT_OK;
}

// Design batches: GenSub and the dual interior-point solve (SolveDIP,
// MMA.cpp:639-934) of every design, m = 1.  Each design runs its own copy of
// the single-design control flow (epsilon levels, at most 100 Newton steps
// per level, residual carried across levels, tolerance 1e-9 sqrt(1 + n_d))
// on its device scalars (uftk::bsc_mma_*, one thread per design); the
// designs advance in lockstep, one fused dual pass per step giving every
// design's x(lambda_d), constraint sum and Hessian sum.  A design that has
// finished keeps its lambda, so later passes rewrite its x unchanged.  The
// host queues the steps in rounds (as many as the last update took, then
// four at a time) and reads two flags per round: a step queued after every
// design has finished launches nothing but its decision kernel's check.
int mma_update_batch(uft_ctx *C, double *x, const double *dfdx, const double *dgdx, const double *xmin,
                     const double *xmax) {
  hipStream_t s = C->s;
  BatchState &B = C->bt;
  const int ns = B.nseg;
  const int64_t n = C->ne;
  const SegG g = B.E.g();
  int rc;
  C->mk++;
  const double *dg[1] = {dgdx};
  uftk::mma_gensub(s, C, n, 1, C->mk, x, xmin, xmax, dfdx, dg, C->d_part, g);
  uftk::seg_final(s, C->d_part, B.E.d_dfirst, ns, 1, B.sc, B_RED);
  std::swap(C->d_xo1, C->d_xo2);
  uftk::copy(s, n, x, C->d_xo1);

  const double a = C->a[0], c = C->c[0];
  double *gt = B.h + (size_t)ns * UFT_BSC;  // pinned staging: gx and the tolerance per design
  for (int d = 0; d < ns; ++d) {
    gt[2 * d] = B.gx[d];
    gt[2 * d + 1] = 1.0e-9 * std::sqrt((double)(1 + B.eoff[d + 1] - B.eoff[d]));
  }
  UFT_HIP(C, hipMemcpyAsync(B.d_gt, gt, 2 * ns * sizeof(double), hipMemcpyHostToDevice, s));
  double *flags = C->d_scal + S_MFL;
  auto pass = [&](const double *gate) {
    uftk::mma_dual_pass(s, C, n, 1, nullptr, x, C->d_part, g, B.d_lam, gate);
    uftk::seg_final(s, C->d_part, B.E.d_dfirst, ns, 2, B.sc, B_RED, gate);
  };
  uftk::bsc_mma_init(s, ns, B.sc, B.d_gt, a, c, B.d_lam);
  pass(nullptr);
  int round = C->dip_steps > 0 ? C->dip_steps + 1 : 16, total = 0;
  while (true) {
    for (int j = 0; j < round; ++j) {
      uftk::bsc_mma_decide(s, ns, B.sc, a, c, B.d_lam, flags);
      pass(flags + MF_NEG);  // the sums at a clamped lambda
      uftk::bsc_mma_step(s, ns, B.sc, a, c, B.d_lam);
      pass(flags + MF_STEP);
      uftk::bsc_mma_resid(s, ns, B.sc, a);
    }
    total += round;
    if ((rc = fetch(C, flags, 2, C->h_scal))) return rc;
    if (C->h_scal[MF_STEP] == 0.0) break;
    if (total > 20000) return uft_set_error(C, UFT_ERR_NOCONV, "MMA dual solve: %d steps without finishing", total);
    round = 4;
  }
  if ((rc = fetch(C, B.sc, ns * UFT_BSC, B.h))) return rc;
  int steps = 0;
  for (int d = 0; d < ns; ++d) {
    const double *S = B.h + (size_t)d * UFT_BSC;
    B.lam[d] = S[B_MLAM];
    B.mu[d] = S[B_MMU];
    B.b[d] = S[B_MB];
    steps = std::max(steps, (int)S[B_MSTEPS]);
  }
  C->dip_steps = steps;
  return UFT_OK;
}

int check_setup(uft_ctx *C) {
  if (!C) return UFT_ERR_ARG;
  if (!C->is_setup) return uft_set_error(C, UFT_ERR_STATE, "uft_setup has not been called");
  return UFT_OK;
}

}  // namespace

extern "C" int uft_autotune(uft_ctx *C, int reps, int *variant, double *us_per_variant, int nmax);

// ==========================================================================
extern "C" {

const char *uft_last_error(const uft_ctx *C) {
  if (C) return C->err.c_str();
  return g_free_err.empty() ? "null context" : g_free_err.c_str();
}

void uft_default_params(uft_params *p) {
  std::memset(p, 0, sizeof(*p));
  // OptPara::SetUp (OptPara.cpp:61-84)
  p->maxItr = 200; p->rmin = 0.01; p->penal = 3.0; p->Emin = 1.0e-9; p->Emax = 1.0e6; p->movlim = 0.2;
  p->Xmin = 0.01; p->Xmax = 1.0; p->volfrac = 0.12; p->x0 = 0.5; p->beta0 = 1.0; p->betamax = 8.0; p->proj = 0;
  p->m = 1;
  p->nu = 0.3;                          // SolveFem.cpp:8
  p->rtol_el = 1.0e-5;                  // SolveFem.cpp:134
  p->rtol_f = 1.0e-5;                   // PDEFilter.cpp:117
  p->maxit_el = 10000;
  p->maxit_f = 10000;                   // PDEFilter.cpp:120
  p->precision = UFT_FP64;
  p->amg_levels = 4;                    // SolveFem.cpp:131
  p->cheb_degree = 2;
  p->coarse_max = 4096;
  p->cycle_type = 2;
  p->cheb_lo = 0.1;
  p->cheb_hi = 1.1;
  p->esteig_steps = 10;
  p->load_cases = 0;
  p->graph = 0;
  p->cheb_degree_coarse = 0;
  p->fine_assembled = 1;
  p->coarse_dense_max = 4096;
  p->prolong_smooth = 1;
  p->coarse_lag = 1;
}

int uft_create(uft_ctx **out, int device) {
  if (!out) return UFT_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return UFT_ERR_HIP;
  if (device < 0 || device >= ndev) return UFT_ERR_ARG;
  uft_ctx *C = new uft_ctx();
  C->device = device;
  uft_default_params(&C->P);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&C->s, hipStreamNonBlocking) != hipSuccess) {
    delete C;
    return UFT_ERR_HIP;
  }
  if (hipHostMalloc((void **)&C->h_scal, 256 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    (void)hipStreamDestroy(C->s);
    delete C;
    return UFT_ERR_HIP;
  }
  *out = C;
  return UFT_OK;
}

void uft_destroy(uft_ctx *C) {
  if (!C) return;
  (void)hipSetDevice(C->device);
  graph_release(C);
  free_all(C);
  if (C->h_scal) (void)hipHostFree(C->h_scal);
  if (C->bt.h) (void)hipHostFree(C->bt.h);
  for (hipEvent_t e : C->ev_side)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t t : C->side)
    if (t) (void)hipStreamDestroy(t);
  if (C->ev_inv) (void)hipEventDestroy(C->ev_inv);
  if (C->s) (void)hipStreamDestroy(C->s);
  delete C;
}

int uft_params_set(uft_ctx *C, const uft_params *p) {
  if (!C || !p) return UFT_ERR_ARG;
  if (p->m < 1 || p->m > UFT_MAXM) return uft_set_error(C, UFT_ERR_ARG, "m must be in [1, %d]", UFT_MAXM);
  if (!(p->rtol_el > 0) || !(p->rtol_f > 0) || p->maxit_el < 1 || p->maxit_f < 1)
    return uft_set_error(C, UFT_ERR_ARG, "bad solver tolerances");
  if (!(p->cheb_lo >= 0.0) || !(p->cheb_hi > p->cheb_lo) || p->esteig_steps < 2 || p->esteig_steps > 64)
    return uft_set_error(C, UFT_ERR_ARG, "bad Chebyshev interval or estimate steps");
  if (p->load_cases != 0 && p->load_cases != 1)
    return uft_set_error(C, UFT_ERR_ARG, "load_cases must be 0 (superposed) or 1 (separate)");
  if (p->graph != 0 && p->graph != 1) return uft_set_error(C, UFT_ERR_ARG, "graph must be 0 or 1");
  if (p->fine_assembled < 0 || p->fine_assembled > 2)
    return uft_set_error(C, UFT_ERR_ARG, "fine_assembled must be 0, 1 or 2");
  if (p->cheb_degree < 1 || p->cheb_degree > 16 || p->cheb_degree_coarse < 0 || p->cheb_degree_coarse > 16 ||
      p->cycle_type < 0 || p->cycle_type > 3)
    return uft_set_error(C, UFT_ERR_ARG, "Chebyshev degrees must be in [1, 16] (coarse: 0 or [1, 16]), cycle_type in [0, 3]");
  if (p->prolong_smooth != 0 && p->prolong_smooth != 1)
    return uft_set_error(C, UFT_ERR_ARG, "prolong_smooth must be 0 or 1");
  if (p->coarse_lag != 0 && p->coarse_lag != 1) return uft_set_error(C, UFT_ERR_ARG, "coarse_lag must be 0 or 1");
  if (p->coarse_dense_max < 6 || p->coarse_dense_max > 6144)
    return uft_set_error(C, UFT_ERR_ARG, "coarse_dense_max must be in [6, 6144]");
  if (p->precision < UFT_FP64 || p->precision > UFT_MIXED_H)
    return uft_set_error(C, UFT_ERR_ARG, "precision must be UFT_FP64, UFT_MIXED or UFT_MIXED_H");
  if (C->is_setup && (p->m != C->P.m || p->rmin != C->P.rmin || p->nu != C->P.nu || p->precision != C->P.precision ||
                      p->load_cases != C->P.load_cases || p->fine_assembled != C->P.fine_assembled ||
                      p->amg_levels != C->P.amg_levels || p->coarse_max != C->P.coarse_max ||
                      p->coarse_dense_max != C->P.coarse_dense_max || p->prolong_smooth != C->P.prolong_smooth ||
                      p->coarse_lag != C->P.coarse_lag))
    return uft_set_error(C, UFT_ERR_STATE,
                         "m, rmin, nu, precision, fine_assembled and the hierarchy shape are fixed at setup");
  C->P = *p;
  return UFT_OK;
}
int uft_params_get(const uft_ctx *C, uft_params *p) {
  if (!C || !p) return UFT_ERR_ARG;
  *p = C->P;
  return UFT_OK;
}

int uft_mesh_set(uft_ctx *C, int npe, int64_t nn, int64_t ne, const int32_t *elem, const double *xyz) {
  if (!C) return UFT_ERR_ARG;
  if (C->is_setup) return uft_set_error(C, UFT_ERR_STATE, "mesh already set up");
  if ((npe != 8 && npe != 4) || nn <= 0 || ne <= 0 || !elem || !xyz)
    return uft_set_error(C, UFT_ERR_ARG, "uft_mesh_set: bad arguments");
  if (nn >= (int64_t)INT32_MAX / 3 || ne * npe * 36 >= (int64_t)INT32_MAX * 8)
    return uft_set_error(C, UFT_ERR_ARG, "mesh too large");
  for (int64_t i = 0; i < ne * npe; ++i)
    if (elem[i] < 0 || elem[i] >= nn)
      return uft_set_error(C, UFT_ERR_ARG, "element %lld references node %d outside [0,%lld)", (long long)(i / npe),
                           elem[i], (long long)nn);
  C->npe = npe;
  C->nd = 3 * npe;
  C->nn = nn;
  C->ne = ne;
  C->ndof = 3 * nn;
  C->elem.assign(elem, elem + ne * npe);
  C->xyz.assign(xyz, xyz + nn * 3);
  C->have_mesh = true;
  return UFT_OK;
}

int uft_bc_set(uft_ctx *C, int nfixed, const int32_t *fixed, int nlc, int nload, const int32_t *ld,
               const double *lv) {
  if (!C) return UFT_ERR_ARG;
  if (!C->have_mesh) return uft_set_error(C, UFT_ERR_STATE, "uft_bc_set before uft_mesh_set");
  if (C->is_setup) return uft_set_error(C, UFT_ERR_STATE, "boundary conditions are fixed after setup");
  if (nfixed < 0 || nlc < 1 || nload < 0 || (nfixed > 0 && !fixed) || (nload > 0 && (!ld || !lv)))
    return uft_set_error(C, UFT_ERR_ARG, "uft_bc_set: bad arguments");
  for (int i = 0; i < nfixed; ++i)
    if (fixed[i] < 0 || fixed[i] >= C->ndof) return uft_set_error(C, UFT_ERR_ARG, "fixed DOF %d out of range", fixed[i]);
  for (int64_t i = 0; i < (int64_t)nlc * nload; ++i)
    if (ld[i] < 0 || ld[i] >= C->ndof) return uft_set_error(C, UFT_ERR_ARG, "load DOF %d out of range", ld[i]);
  C->fixed.assign(fixed, fixed + nfixed);
  C->nlc = nlc;
  C->nload = nload;
  C->ldofs.assign(ld, ld + (int64_t)nlc * nload);
  C->lvals.assign(lv, lv + (int64_t)nlc * nload);
  C->have_bc = true;
  return UFT_OK;
}

// Smoothed prolongator, once at setup: the refresh's contribution lists
// (counted, then filled, by scanning the patterns on the device) and the
// cycle's two layouts of P (node-sliced for the prolongation, support order
// for the restriction).
int sa_list(uft_ctx *C, int which, int64_t rows, int32_t **ptr, int32_t **a, int32_t **b) {
  hipStream_t s = C->s;
  int32_t *cnt = nullptr;
  int rc = dalloc_n(C, &cnt, std::max<int64_t>(rows, 1));
  if (rc) return rc;
  uftk::sa_lists(s, C, which, nullptr, cnt, nullptr);
  std::vector<int32_t> hc(rows);
  if (rows) UFT_HIP(C, hipMemcpyAsync(hc.data(), cnt, rows * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if ((rc = sync(C))) return rc;
  std::vector<int32_t> hp(rows + 1, 0);
  int64_t tot = 0;
  for (int64_t r = 0; r < rows; ++r) {
    tot += hc[r];
    if (tot > INT32_MAX) return uft_set_error(C, UFT_ERR_ARG, "smoothed prolongator lists exceed 2^31 entries");
    hp[r + 1] = (int32_t)tot;
  }
  if ((rc = upload(C, ptr, hp)) || (rc = dalloc_n(C, a, std::max<int64_t>(tot, 1)))) return rc;
  if (b && (rc = dalloc_n(C, b, std::max<int64_t>(tot, 1)))) return rc;
  uftk::sa_lists(s, C, which, *ptr, *a, b ? *b : nullptr);
  return sync(C);
}

int sa_setup_lists(uft_ctx *C, const AmgHostLevel &h) {
  AmgLevel &L0 = C->amg.L[0];
  int rc;
  // support position of every entry (the A P lists name entries by it)
  std::vector<int32_t> spos(L0.sa_np, 0);
  for (size_t q = 0; q < h.sa_sidx.size(); ++q) spos[h.sa_sidx[q]] = (int32_t)q;
  if ((rc = upload(C, &L0.d_sa_spos, spos))) return rc;
  if ((rc = sa_list(C, 0, L0.sa_np, &L0.d_sa_wptr, &L0.d_sa_wslot, nullptr)) ||
      (rc = sa_list(C, 1, L0.sa_nv, &L0.d_sa_vlptr, &L0.d_sa_vlslot, &L0.d_sa_vlent)) ||
      (rc = sa_list(C, 2, L0.sa_nub, &L0.d_sa_gptr, &L0.d_sa_gp, &L0.d_sa_gv)))
    return rc;
  const int64_t nn = L0.nunit, nsl = (nn + UFT_ELL_SLICE - 1) / UFT_ELL_SLICE;
  std::vector<int64_t> poff(nsl + 1, 0);
  std::vector<int32_t> pw(nsl, 0);
  for (int64_t w = 0; w < nsl; ++w) {
    int mx = 0;
    for (int64_t n = w * UFT_ELL_SLICE; n < std::min(nn, (w + 1) * UFT_ELL_SLICE); ++n)
      mx = std::max(mx, h.sa_pptr[n + 1] - h.sa_pptr[n]);
    pw[w] = mx;
    poff[w + 1] = poff[w] + (int64_t)mx * UFT_ELL_SLICE;
  }
  std::vector<int32_t> eagg(poff[nsl], -1);
  for (int64_t n = 0; n < nn; ++n) {
    const int64_t w = n / UFT_ELL_SLICE, lane = n % UFT_ELL_SLICE;
    for (int32_t t = h.sa_pptr[n]; t < h.sa_pptr[n + 1]; ++t)
      eagg[poff[w] + (int64_t)(t - h.sa_pptr[n]) * UFT_ELL_SLICE + lane] = h.sa_pagg[t];
  }
  L0.sa_ne_ell = poff[nsl];
  if ((rc = upload(C, &L0.d_sa_poff, poff)) || (rc = upload(C, &L0.d_sa_pw, pw)) ||
      (rc = upload(C, &L0.d_sa_eagg, eagg)))
    return rc;
  if (is_mixed(C)) {
    // (padding slots are never read: the prolongation stops at aggregate -1);
    // FP16 copies with the smoother's low-precision blocks (UFT_MIXED_H)
    if (C->P.precision == UFT_MIXED_H) {
      if ((rc = dalloc_n(C, &L0.d_sa_Peh, 18 * L0.sa_ne_ell)) || (rc = dalloc_n(C, &L0.d_sa_Psh, 18 * L0.sa_np)))
        return rc;
    } else if ((rc = dalloc_n(C, &L0.d_sa_Pef, 18 * L0.sa_ne_ell)) ||
               (rc = dalloc_n(C, &L0.d_sa_Psf, 18 * L0.sa_np))) {
      return rc;
    }
  } else {
    if ((rc = dalloc_n(C, &L0.d_sa_Pe, 18 * L0.sa_ne_ell)) || (rc = dalloc_n(C, &L0.d_sa_Ps, 18 * L0.sa_np)))
      return rc;
  }
  return sync(C);
}

int uft_setup(uft_ctx *C) {
  if (!C) return UFT_ERR_ARG;
  if (!C->have_mesh || !C->have_bc) return uft_set_error(C, UFT_ERR_STATE, "mesh and BCs must be set before setup");
  if (C->is_setup) return uft_set_error(C, UFT_ERR_STATE, "already set up");
  UFT_HIP(C, hipSetDevice(C->device));
  const int npe = C->npe, nd = C->nd, m = C->P.m;
  const int64_t ne = C->ne, nn = C->nn, ndof = C->ndof;
  const int np = nd * (nd + 1) / 2, npf = npe * (npe + 1) / 2;
  hipStream_t s = C->s;
  int rc;
  // ---- mesh on device (element ids structure-of-arrays)
  {
    std::vector<int32_t> soa((size_t)ne * npe);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < npe; ++a) soa[(size_t)a * ne + e] = C->elem[e * npe + a];
    if ((rc = upload(C, &C->d_elem, soa))) return rc;
    if ((rc = upload(C, &C->d_xyz, C->xyz))) return rc;
    std::vector<uint8_t> fm(nn, 0);
    for (int32_t d : C->fixed) fm[d / 3] |= (uint8_t)(1u << (d % 3));
    if ((rc = upload(C, &C->d_fixm, fm))) return rc;
    std::vector<int32_t> iptr, inc;
    uft_node_incidence(nn, ne, npe, C->elem.data(), iptr, inc);
    if ((rc = upload(C, &C->d_iptr, iptr))) return rc;
    if ((rc = upload(C, &C->d_inc, inc))) return rc;
    // assembled node operators on the node-neighbour pattern: the filter's
    // Helmholtz operator in every form, the elasticity operators in
    // UFT_MIXED_H (also for separate load cases, and a design batch's designs
    // share no node, so its concatenated mesh assembles the same way)
    const bool elast = C->P.precision == UFT_MIXED_H;
    if (C->P.fine_assembled) {
      std::vector<int64_t> off;
      std::vector<int32_t> width, col, cptr, clist;
      if (uft_fine_ell_pattern(nn, npe, C->elem.data(), iptr, inc, off, width, col, cptr, clist)) {
        auto &F = C->fell;
        F.nslice = (int64_t)width.size();
        F.nslot = off.back();
        F.nnzb = 0;
        for (int64_t i = 0; i < F.nslot; ++i) F.nnzb += cptr[i + 1] > cptr[i];
        F.wmax = *std::max_element(width.begin(), width.end());
        // 16-bit column offsets when the ordering keeps every neighbour within 32767 nodes
        std::vector<int16_t> cd(col.size(), 0);
        bool fits = true;
        for (int64_t n = 0; n < nn && fits; ++n) {
          const int64_t w = n / UFT_ELL_SLICE, lane = n % UFT_ELL_SLICE;
          for (int q = 0; q < width[w]; ++q) {
            const int64_t k = off[w] + (int64_t)q * UFT_ELL_SLICE + lane, dlt = (int64_t)col[k] - n;
            if (dlt < INT16_MIN || dlt > INT16_MAX) {
              fits = false;
              break;
            }
            cd[k] = (int16_t)dlt;
          }
        }
        if (fits && (rc = upload(C, &F.d_cd, cd))) return rc;
        if ((rc = upload(C, &F.d_off, off)) || (rc = upload(C, &F.d_width, width)) || (rc = upload(C, &F.d_col, col)) ||
            (rc = upload(C, &F.d_cptr, cptr)) || (rc = upload(C, &F.d_clist, clist)) ||
            (rc = dalloc_n(C, &F.d_hv, F.nslot)))
          return rc;
        // smoother blocks: int8 with a block scale for one right-hand side,
        // FP16 for the multi-load-case form (its eight lanes per node share
        // each block load, so its blocks are not the bound)
        const bool mr = C->P.load_cases == 1 && C->nlc > 1;
        // UFT_MIXED (one right-hand side): the FP32 blocks serve the smoother
        // and the outer product
        if (C->P.precision == UFT_MIXED && !mr &&
            ((rc = dalloc_n(C, &F.d_w8, 2 * F.nslot)) || (rc = dalloc_n(C, &F.d_w1, F.nslot))))
          return rc;
        if (elast && ((rc = mr ? dalloc_n(C, &F.d_v8, F.nslot) : dalloc_n(C, &F.d_q8, F.nslot)) ||
                      (rc = mr ? dalloc_n(C, &F.d_v1, F.nslot) : dalloc_n(C, &F.d_qs, F.nslot)) ||
                      (rc = dalloc_n(C, &F.d_dg, ndof)) || (rc = dalloc_n(C, &F.d_w8, 2 * F.nslot)) ||
                      (rc = dalloc_n(C, &F.d_w1, F.nslot))))
          return rc;
        // node tiles for the products on eight right-hand sides (5-8 load cases;
        // fine_assembled = 2: none)
        NodeTiles T;
        if (elast && mr && C->nlc > 4 && C->P.fine_assembled == 1 &&
            uft_node_tiles(nn, C->xyz.data(), off, width, col, UFT_TILE, UFT_TILE_HCAP, T)) {
          F.ntile = T.ntile;
          F.tslot = T.nslot;
          F.thmax = T.hmax;
          if ((rc = upload(C, &F.d_tnode, T.node)) || (rc = upload(C, &F.d_thoff, T.hoff)) ||
              (rc = upload(C, &F.d_thnode, T.hnode)) || (rc = upload(C, &F.d_toff, T.off)) ||
              (rc = upload(C, &F.d_twidth, T.width)) || (rc = upload(C, &F.d_tlc, T.lc)) ||
              (rc = upload(C, &F.d_kmap, T.kmap)) || (rc = dalloc_n(C, &F.d_tv8, T.nslot)) ||
              (rc = dalloc_n(C, &F.d_tv1, T.nslot)) || (rc = dalloc_n(C, &F.d_tw8, 2 * T.nslot)) ||
              (rc = dalloc_n(C, &F.d_tw1, T.nslot)))
            return rc;
          // padding slots of the tiled copies stay zero (the assembly writes the real ones)
          UFT_HIP(C, hipMemsetAsync(F.d_tv8, 0, (size_t)T.nslot * sizeof(uint4), s));
          UFT_HIP(C, hipMemsetAsync(F.d_tv1, 0, (size_t)T.nslot * sizeof(uint16_t), s));
          UFT_HIP(C, hipMemsetAsync(F.d_tw8, 0, (size_t)T.nslot * 2 * sizeof(float4), s));
          UFT_HIP(C, hipMemsetAsync(F.d_tw1, 0, (size_t)T.nslot * sizeof(float), s));
        }
        UFT_HIP(C, hipStreamSynchronize(s));  // the host arrays end with this scope
      }
    }
    UFT_HIP(C, hipStreamSynchronize(s));
  }
  // ---- design batch: segment grids of the element, node and DOF spaces
  BatchState &B = C->bt;
  if (B.nseg) {
    if (m != 1) return uft_set_error(C, UFT_ERR_ARG, "design batches take one constraint (m = 1)");
    if (C->P.load_cases == 1 && C->nlc > 1)
      return uft_set_error(C, UFT_ERR_ARG, "design batches take superposed load cases (load_cases = 0)");
    std::vector<int64_t> doff(B.noff.size());
    for (size_t d = 0; d < doff.size(); ++d) doff[d] = 3 * B.noff[d];
    if ((rc = build_space(C, B.E, B.eoff)) || (rc = build_space(C, B.N, B.noff)) || (rc = build_space(C, B.D, doff)))
      return rc;
    if ((rc = dalloc_n(C, &B.sc, (int64_t)B.nseg * UFT_BSC))) return rc;
    uftk::fill(s, (int64_t)B.nseg * UFT_BSC, 0.0, B.sc);
    if ((rc = upload(C, &B.d_vf, B.volfrac))) return rc;
    double **bv[] = {&B.d_invTV, &B.d_lam, &B.d_fac};
    for (double **p : bv)
      if ((rc = dalloc_n(C, p, (int64_t)B.nseg))) return rc;
    if ((rc = dalloc_n(C, &B.d_gt, (int64_t)B.nseg * 2))) return rc;
    {  // design of every element and node
      std::vector<int32_t> ed((size_t)ne), nd((size_t)nn);
      for (int d = 0; d < B.nseg; ++d) {
        for (int64_t e = B.eoff[d]; e < B.eoff[d + 1]; ++e) ed[e] = d;
        for (int64_t v = B.noff[d]; v < B.noff[d + 1]; ++v) nd[v] = d;
      }
      if ((rc = upload(C, &B.d_edes, ed)) || (rc = upload(C, &B.d_ndes, nd))) return rc;
    }
    if ((rc = dalloc_n(C, &B.d_log, (int64_t)B.nseg * 128))) return rc;
    uftk::fill(s, (int64_t)B.nseg * 128, 0.0, B.d_log);
  }
  // ---- vectors
  C->vec.assign(UFT_DGDX0 + m, nullptr);
  C->vlen.assign(UFT_DGDX0 + m, 0);
  const int elem_ids[] = {UFT_X, UFT_XFILED, UFT_XPHYS, UFT_XOLD, UFT_DFDX, UFT_XMIN, UFT_XMAX, UFT_VOL, UFT_VMS};
  for (int id : elem_ids) {
    if ((rc = dalloc_n(C, &C->vec[id], ne))) return rc;
    C->vlen[id] = ne;
  }
  for (int j = 0; j < m; ++j) {
    if ((rc = dalloc_n(C, &C->vec[UFT_DGDX0 + j], ne))) return rc;
    C->vlen[UFT_DGDX0 + j] = ne;
    uftk::fill(s, ne, 0.0, C->vec[UFT_DGDX0 + j]);
  }
  if ((rc = dalloc_n(C, &C->vec[UFT_U], ndof))) return rc;
  C->vlen[UFT_U] = ndof;
  if ((rc = dalloc_n(C, &C->vec[UFT_RX], nn))) return rc;
  C->vlen[UFT_RX] = nn;
  // ---- element operators
  if ((rc = dalloc_n(C, &C->d_KE, (int64_t)np * ne))) return rc;
  if (is_mixed(C)) {
    if ((rc = dalloc_n(C, &C->d_KEf, (int64_t)np * ne))) return rc;
    if ((rc = dalloc_n(C, &C->d_Yef, (int64_t)nd * ne))) return rc;
  }
  if ((rc = dalloc_n(C, &C->d_KF, (int64_t)npf * ne))) return rc;
  if ((rc = dalloc_n(C, &C->d_T, (int64_t)npe * ne))) return rc;
  if ((rc = dalloc_n(C, &C->d_detmin, ne))) return rc;
  if ((rc = dalloc_n(C, &C->d_Ye, (int64_t)nd * ne))) return rc;
  const int64_t maxn = std::max(ndof, ne);
  // block partials of the deterministic two-launch reductions; the coarse
  // operators' fused K-cycle products write up to 3 per block of 256/32 rows
  // (a level has at most 6 rows per fine node)
  C->part_cap = (maxn / 256 + 2) * 48 + 3 * (2 * ndof * 32 / 256 + 2);
  if ((rc = dalloc_n(C, &C->d_part, C->part_cap))) return rc;
  if ((rc = dalloc_n(C, &C->d_scal, 256))) return rc;
  if ((rc = dalloc_n(C, &C->d_lanczos, 3 * 64 * UFT_MAXLEV))) return rc;
  if (!C->bt.nseg) {  // the hierarchy refresh's concurrent branches (default-priority streams)
    for (hipStream_t &t : C->side)
      if (!t && hipStreamCreateWithFlags(&t, hipStreamNonBlocking) != hipSuccess)
        return uft_set_error(C, UFT_ERR_HIP, "hipStreamCreateWithFlags failed");
    for (hipEvent_t &e : C->ev_side)
      if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
        return uft_set_error(C, UFT_ERR_HIP, "hipEventCreateWithFlags failed");
    if ((rc = dalloc_n(C, &C->d_part2, C->part_cap))) return rc;
  }
  uftk::fill(s, 256, 0.0, C->d_scal);
  uftk::gen_ke(s, npe, ne, C->d_elem, C->d_xyz, C->P.nu, C->d_KE, C->d_KEf, C->d_detmin);
  if (C->P.precision == UFT_MIXED_H) {
    if ((rc = dalloc_n(C, &C->d_KEh, (int64_t)uftk::ke_half_groups(npe) * ne))) return rc;
    if ((rc = dalloc_n(C, &C->d_KEs, ne))) return rc;
    uftk::ke_to_half(s, C);
  }
  if (B.nseg) {  // per design filter length and total volume
    std::vector<double> r2(B.nseg), inv(B.nseg);
    for (int d = 0; d < B.nseg; ++d) {
      const double r = 1.0 / (2.0 * std::sqrt(3.0)) * B.rmin[d];
      r2[d] = r * r;
    }
    double *d_r2;
    if ((rc = upload(C, &d_r2, r2))) return rc;
    uftk::gen_kf(s, npe, ne, C->d_elem, C->d_xyz, C->P.rmin, C->d_KF, C->d_T, C->vec[UFT_VOL], C->d_part, B.E.g(),
                 d_r2);
    uftk::seg_final(s, C->d_part, B.E.d_dfirst, B.nseg, 1, B.sc, B_RED);
    if ((rc = fetch(C, B.sc, B.nseg * UFT_BSC, B.h))) return rc;
    B.TV.assign(B.nseg, 0.0);
    double tv = 0.0;
    for (int d = 0; d < B.nseg; ++d) {
      B.TV[d] = B.h[(size_t)d * UFT_BSC + B_RED];
      inv[d] = 1.0 / B.TV[d];
      tv += B.TV[d];
    }
    UFT_HIP(C, hipMemcpyAsync(B.d_invTV, inv.data(), B.nseg * sizeof(double), hipMemcpyHostToDevice, s));
    uftk::fill(s, 1, tv, C->d_scal + S_MISC);
  } else {
    uftk::gen_kf(s, npe, ne, C->d_elem, C->d_xyz, C->P.rmin, C->d_KF, C->d_T, C->vec[UFT_VOL], C->d_part);
    uftk::final_sum(s, C->d_part, uftk::red_blocks(ne), 1, C->d_scal + S_MISC);
  }
  if (C->fell.d_hv) uftk::helm_ell_fill(s, C);  // the filter operator on the node pattern (density-free)
  {
    std::vector<double> dm(ne);
    UFT_HIP(C, hipMemcpyAsync(dm.data(), C->d_detmin, ne * sizeof(double), hipMemcpyDeviceToHost, s));
    if ((rc = fetch(C, C->d_scal + S_MISC, 1, C->h_scal))) return rc;
    C->TV = C->h_scal[0];
    for (int64_t e = 0; e < ne; ++e)
      if (!(dm[e] > 0.0))
        return uft_set_error(C, UFT_ERR_DETJ, "element %lld has non-positive Jacobian determinant %.6e", (long long)e,
                             dm[e]);
  }
  // ---- work vectors
  double **dv[] = {&C->d_r, &C->d_z, &C->d_p, &C->d_q, &C->d_F, &C->d_w};
  for (double **p : dv)
    if ((rc = dalloc_n(C, p, ndof))) return rc;
  double **fv[] = {&C->d_fr, &C->d_fz, &C->d_fp, &C->d_fq, &C->d_fdinv, &C->d_TX};
  for (double **p : fv)
    if ((rc = dalloc_n(C, p, nn))) return rc;
  if (is_mixed(C)) {
    float **ff[] = {&C->d_rf, &C->d_zf, &C->d_pf, &C->d_qf, &C->d_uf};
    for (float **p : ff)
      if ((rc = dalloc_n(C, p, ndof))) return rc;
  }
  // ---- load vector: INSERT per case, then zero the fixed DOFs (SolveFem.cpp:277-333)
  {
    uftk::fill(s, ndof, 0.0, C->d_F);
    int32_t *dld;
    double *dlv;
    if ((rc = upload(C, &dld, C->ldofs))) return rc;
    if ((rc = upload(C, &dlv, C->lvals))) return rc;
    for (int lc = 0; lc < C->nlc; ++lc)
      uftk::set_dofs(s, C->nload, dld + (size_t)lc * C->nload, dlv + (size_t)lc * C->nload, C->d_F);
    int32_t *dfx;
    if ((rc = upload(C, &dfx, C->fixed))) return rc;
    uftk::set_dofs(s, (int)C->fixed.size(), dfx, nullptr, C->d_F);
  }
  // ---- filter diagonal (KF is density independent)
  uftk::diag_KF(s, C, C->d_fdinv);
  // ---- multigrid hierarchy
  {
    std::vector<AmgHostLevel> H;
    std::vector<int64_t> nun;
    // smoothed aggregation on the fine level: one design, one right-hand side,
    // the assembled node pattern (its FP64 blocks are the operator that smooths
    // the prolongator) and at least three levels (the level-1 correction is not
    // the dense coarsest solve)
    const bool want_sa = C->P.prolong_smooth && C->fell.nslot > 0;
    if ((rc = uft_amg_build_host(C, H, nun, want_sa))) return rc;
    if (want_sa && nun.size() < 3) {
      H.clear();
      nun.clear();
      if ((rc = uft_amg_build_host(C, H, nun, false))) return rc;
    }
    AmgHier &A = C->amg;
    A.nlev = (int)nun.size();
    for (int l = 0; l < A.nlev; ++l) {
      AmgLevel &Lv = A.L[l];
      Lv.nunit = nun[l];
      Lv.bs = l == 0 ? 3 : 6;
      Lv.n = Lv.nunit * Lv.bs;
      double **wv[] = {&Lv.x, &Lv.b, &Lv.r, &Lv.d, &Lv.t};
      for (double **p : wv)
        if ((rc = dalloc_n(C, p, Lv.n))) return rc;
      if ((rc = dalloc_n(C, &Lv.d_dinv, Lv.n))) return rc;
      if (l > 0) {
        double **kv[] = {&Lv.kc, &Lv.kv, &Lv.kd, &Lv.kr};
        for (double **p : kv)
          if ((rc = dalloc_n(C, p, Lv.n))) return rc;
      }
      if (is_mixed(C)) {
        float **wf[] = {&Lv.xf, &Lv.bf, &Lv.rf, &Lv.df, &Lv.tf};
        for (float **p : wf)
          if ((rc = dalloc_n(C, p, Lv.n))) return rc;
        if ((rc = dalloc_n(C, &Lv.d_dinvf, Lv.n))) return rc;
        if (l > 0) {
          float **kf[] = {&Lv.kcf, &Lv.kvf, &Lv.kdf, &Lv.krf};
          for (float **p : kf)
            if ((rc = dalloc_n(C, p, Lv.n))) return rc;
        }
      }
    }
    for (int l = 0; l + 1 < A.nlev; ++l) {
      AmgHostLevel &h = H[l];
      AmgLevel &Lv = A.L[l], &Ln = A.L[l + 1];
      if ((rc = upload(C, &Lv.d_P, h.P))) return rc;
      if (is_mixed(C)) {
        std::vector<float> pf(h.P.begin(), h.P.end());
        if ((rc = upload(C, &Lv.d_Pf, pf))) return rc;
      }
      if ((rc = upload(C, &Lv.d_parent, h.parent))) return rc;
      if ((rc = upload(C, &Lv.d_cptr, h.cptr))) return rc;
      if ((rc = upload(C, &Lv.d_child, h.child))) return rc;
      if ((rc = upload(C, &Lv.d_glptr, h.glptr))) return rc;
      if ((rc = upload(C, &Lv.d_gl, h.gl))) return rc;
      Lv.ngl = (int64_t)h.gl.size();
      Ln.nblk = (int64_t)h.ci.size();
      if ((rc = upload(C, &Ln.d_rp, h.rp))) return rc;
      if ((rc = upload(C, &Ln.d_ci, h.ci))) return rc;
      std::vector<int32_t> brow(Ln.nblk);
      for (int64_t A2 = 0; A2 < Ln.nunit; ++A2)
        for (int32_t k = h.rp[A2]; k < h.rp[A2 + 1]; ++k) brow[k] = (int32_t)A2;
      if ((rc = upload(C, &Ln.d_brow, brow))) return rc;
      if ((rc = dalloc_n(C, &Ln.d_val, Ln.nblk * 36))) return rc;
      if (l == 0 && h.sa) {  // smoothed prolongator: patterns, supports, work arrays
        Lv.sa = true;
        Lv.sa_np = (int64_t)h.sa_pagg.size();
        Lv.sa_nv = (int64_t)h.sa_vagg.size();
        std::vector<int32_t> pnode(Lv.sa_np), vnode(Lv.sa_nv), ublk;
        for (int64_t i = 0; i < Lv.nunit; ++i) {
          for (int32_t k = h.sa_pptr[i]; k < h.sa_pptr[i + 1]; ++k) pnode[k] = (int32_t)i;
          for (int32_t k = h.sa_vptr[i]; k < h.sa_vptr[i + 1]; ++k) vnode[k] = (int32_t)i;
        }
        for (int64_t A2 = 0; A2 < Ln.nunit; ++A2)
          for (int32_t k = h.rp[A2]; k < h.rp[A2 + 1]; ++k)
            if (h.ci[k] >= A2) ublk.push_back(k);
        Lv.sa_nub = (int64_t)ublk.size();
        if ((rc = upload(C, &Lv.d_gtr, h.tr)) || (rc = upload(C, &Lv.d_sa_pptr, h.sa_pptr)) ||
            (rc = upload(C, &Lv.d_sa_pagg, h.sa_pagg)) || (rc = upload(C, &Lv.d_sa_pnode, pnode)) ||
            (rc = upload(C, &Lv.d_sa_vptr, h.sa_vptr)) || (rc = upload(C, &Lv.d_sa_vagg, h.sa_vagg)) ||
            (rc = upload(C, &Lv.d_sa_vnode, vnode)) || (rc = upload(C, &Lv.d_sa_sptr, h.sa_sptr)) ||
            (rc = upload(C, &Lv.d_sa_snode, h.sa_snode)) || (rc = upload(C, &Lv.d_sa_sidx, h.sa_sidx)) ||
            (rc = upload(C, &Lv.d_sa_ublk, ublk)) || (rc = dalloc_n(C, &C->d_Ee, C->ne)))
          return rc;
        if (is_mixed(C)) {
          if ((rc = dalloc_n(C, &Lv.d_sa_Vf, 18 * Lv.sa_nv)) || (rc = dalloc_n(C, &C->d_a32, 9 * C->fell.nslot)))
            return rc;
        } else if ((rc = dalloc_n(C, &Lv.d_sa_V, 18 * Lv.sa_nv)) || (rc = dalloc_n(C, &C->d_a64, 9 * C->fell.nslot))) {
          return rc;
        } else {
          // the same FP64 blocks also serve the solves' fine operator (one
          // design, a design batch or load cases; written by the refresh's assembly)
          if ((rc = dalloc_n(C, &C->fell.d_e8, 4 * C->fell.nslot)) || (rc = dalloc_n(C, &C->fell.d_e1, C->fell.nslot)))
            return rc;
        }
        if ((rc = sa_setup_lists(C, h))) return rc;
      } else if (l == 0) {  // density-free Galerkin entries of the fine level
        int32_t *d_gblk;
        if ((rc = upload(C, &Lv.d_gtr, h.tr))) return rc;
        if ((rc = upload(C, &d_gblk, h.gblk))) return rc;
        if ((rc = dalloc_n(C, &Lv.d_G, Lv.ngl * 36))) return rc;
        if ((rc = dalloc_n(C, &C->d_Ee, C->ne))) return rc;
        uftk::galerkin_pre(s, C, Lv.ngl, d_gblk, Lv.d_gl, Ln.d_brow, Ln.d_ci, Lv.d_parent, Lv.d_P, Lv.d_G);
      }
      if (is_mixed(C))
        if ((rc = dalloc_n(C, &Ln.d_valf, Ln.nblk * 36))) return rc;
      if (C->P.precision == UFT_MIXED_H && l + 2 < A.nlev) {  // coarse levels' smoother blocks in FP16 (not the coarsest)
        if ((rc = dalloc_n(C, &Ln.d_valh, Ln.nblk * 36))) return rc;
        if ((rc = dalloc_n(C, &Ln.d_dh, Ln.n))) return rc;
      }
      UFT_HIP(C, hipStreamSynchronize(s));  // host vectors of H go out of scope level by level
    }
    A.ncoarse = A.L[A.nlev - 1].n;
    if (B.nseg) {
      // rows of every coarse level, and the block-diagonal coarsest level
      for (int l = 1; l < A.nlev; ++l) {
        std::vector<int64_t> roff(B.uoff[l].size());
        for (size_t d = 0; d < roff.size(); ++d) roff[d] = 6 * B.uoff[l][d];
        if ((rc = build_space(C, B.R[l], roff))) return rc;
      }
      const std::vector<int64_t> &cu = B.uoff[A.nlev - 1];
      std::vector<int32_t> cu32(cu.begin(), cu.end()), cdes(A.L[A.nlev - 1].nunit);
      B.coff.assign(B.nseg + 1, 0);
      for (int d = 0; d < B.nseg; ++d) {
        const int64_t nd6 = 6 * (cu[d + 1] - cu[d]);
        B.coff[d + 1] = B.coff[d] + nd6 * nd6;
        for (int64_t u = cu[d]; u < cu[d + 1]; ++u) cdes[u] = d;
      }
      if ((rc = upload(C, &B.d_cu, cu32)) || (rc = upload(C, &B.d_coff, B.coff)) || (rc = upload(C, &B.d_cdes, cdes)))
        return rc;
      if ((rc = dalloc_n(C, &A.d_cmat, B.coff[B.nseg]))) return rc;
      if ((rc = dalloc_n(C, &A.d_cinv, B.coff[B.nseg]))) return rc;
    } else {
      // matrix plus the inverse's scratch (the small-size Gauss-Jordan: at most n^2/16 + n^2/2 + 2n + 1
      // doubles of panels; the tiled sweep: uftk::dense_sweep_scratch)
      const int64_t nc = A.ncoarse;
      const int64_t need = nc >= UFT_SWEEP_MIN ? uftk::dense_sweep_scratch(nc) : 2 * nc * nc + 2 * nc + 1;
      if ((rc = dalloc_n(C, &A.d_cmat, need))) return rc;
      if ((rc = dalloc_n(C, &A.d_cinv, A.ncoarse * A.ncoarse))) return rc;
      if (is_mixed(C))
        if ((rc = dalloc_n(C, &A.d_cinvf, A.ncoarse * A.ncoarse))) return rc;
      // the int8-smoother form's cycle applies the inverse from a row-scaled
      // FP16 copy (one design, one load case; the load cases and batches
      // keep FP32 and FP64)
      const bool half = C->P.precision == UFT_MIXED_H && !(C->P.load_cases == 1 && C->nlc > 1) && !C->bt.nseg &&
                        nc % 8 == 0 && nc > UFT_TAIL_MAX;
      if (half && ((rc = dalloc_n(C, &A.d_cinvh, nc * nc)) || (rc = dalloc_n(C, &A.d_cinvs, nc)))) return rc;
      if (C->P.coarse_lag && nc >= UFT_SWEEP_MIN) {
        if ((rc = dalloc_n(C, &A.d_cinv2, nc * nc))) return rc;
        if (is_mixed(C))
          if ((rc = dalloc_n(C, &A.d_cinvf2, nc * nc))) return rc;
        if (half && ((rc = dalloc_n(C, &A.d_cinvh2, nc * nc)) || (rc = dalloc_n(C, &A.d_cinvs2, nc)))) return rc;
        if (!C->ev_inv && hipEventCreateWithFlags(&C->ev_inv, hipEventDisableTiming) != hipSuccess)
          return uft_set_error(C, UFT_ERR_HIP, "hipEventCreateWithFlags failed");
      }
      A.lag_spare = A.lag_pending = false;
    }
    UFT_HIP(C, hipStreamSynchronize(s));
  }
  // ---- multi-load-case form: interleaved right-hand sides and work vectors
  if (C->P.load_cases == 1 && C->nlc > 1) {
    if (C->nlc > 8) return uft_set_error(C, UFT_ERR_ARG, "the multi-load-case form takes at most 8 load cases");
    MrState &M = C->mr;
    M.ncase = C->nlc;
    M.nr = C->nlc <= 2 ? 2 : (C->nlc <= 4 ? 4 : 8);
    const int nr = M.nr;
    const size_t ew = is_mixed(C) ? sizeof(float) : sizeof(double);
    double **mv[] = {&M.F, &M.U, &M.r, &M.z, &M.p, &M.q, &M.zo};
    for (double **p : mv)
      if ((rc = dalloc_n(C, p, ndof * nr))) return rc;
    if (is_mixed(C)) {
      if ((rc = dalloc_n(C, &M.rf, ndof * nr))) return rc;
      if ((rc = dalloc_n(C, &M.zf, ndof * nr))) return rc;
      if ((rc = dalloc_n(C, &M.pf, ndof * nr))) return rc;
      if ((rc = dalloc_n(C, &M.qf, ndof * nr))) return rc;
      if ((rc = dalloc_n(C, &C->d_KEb, (int64_t)uftk::ke_bf16_rows_words(npe) * ne))) return rc;
      uftk::ke_to_bf16(s, C);
    }
    if ((rc = dalloc(C, &M.Ye, (size_t)nd * ne * nr * sizeof(double)))) return rc;
    if ((rc = dalloc_n(C, &M.sc, 1024))) return rc;
    uftk::fill(s, 1024, 0.0, M.sc);
    int64_t need = 0;
    for (int l = 0; l < C->amg.nlev; ++l) {
      const AmgLevel &Lv = C->amg.L[l];
      MrLevelBuf &B = M.L[l];
      void **bv[] = {&B.x, &B.b, &B.r, &B.d, &B.t};
      for (void **p : bv)
        if ((rc = dalloc(C, p, (size_t)Lv.n * nr * ew))) return rc;
      if (l > 0) {
        void **kv[] = {&B.kc, &B.kv, &B.kd, &B.kr};
        for (void **p : kv)
          if ((rc = dalloc(C, p, (size_t)Lv.n * nr * ew))) return rc;
        need = std::max<int64_t>(need, (int64_t)uftk::mr_bsr_blocks(Lv.nunit) * 3 * nr);
      }
    }
    need = std::max<int64_t>(need, (int64_t)uftk::mr_gather_blocks(nn, nr) * nr);
    need += 3 * 8 * 128;  // the two-stage final sums of the K-cycle's partial rows
    need = std::max<int64_t>(need, (int64_t)((ndof * nr + 255) / 256) * 2 * nr);
    if (need > C->part_cap) {  // the block partials outgrow the single-vector buffer
      if ((rc = dalloc_n(C, &C->d_part, need + 64))) return rc;
      C->part_cap = need + 64;
    }
    // loads INSERTed case by case, then the fixed DOFs zeroed (SolveFem.cpp:277-333)
    uftk::fill(s, ndof * nr, 0.0, M.F);
    int32_t *dld, *dfx;
    double *dlv;
    if ((rc = upload(C, &dld, C->ldofs))) return rc;
    if ((rc = upload(C, &dlv, C->lvals))) return rc;
    if ((rc = upload(C, &dfx, C->fixed))) return rc;
    for (int k = 0; k < C->nlc; ++k) {
      uftk::mr_set_dofs(s, C->nload, nr, k, dld + (size_t)k * C->nload, dlv + (size_t)k * C->nload, M.F);
      uftk::mr_set_dofs(s, (int)C->fixed.size(), nr, k, dfx, nullptr, M.F);
    }
    UFT_HIP(C, hipStreamSynchronize(s));
  }
  // ---- design state (OptPara.cpp:150-153) and MMA (OptPara.cpp:158-177, MMA.cpp:178-224)
  for (int id : {UFT_X, UFT_XFILED, UFT_XPHYS, UFT_XOLD}) {
    if (B.nseg)
      uftk::seg_fill(s, B.E.g(), ne, B.d_vf, C->vec[id]);
    else
      uftk::fill(s, ne, C->P.volfrac, C->vec[id]);
  }
  for (int id : {UFT_DFDX, UFT_XMIN, UFT_XMAX}) uftk::fill(s, ne, 0.0, C->vec[id]);
  uftk::fill(s, ndof, 0.0, C->vec[UFT_U]);
  uftk::fill(s, nn, 0.0, C->vec[UFT_RX]);
  double **mv[] = {&C->d_L, &C->d_U, &C->d_alpha, &C->d_beta, &C->d_p0, &C->d_q0, &C->d_xo1, &C->d_xo2};
  for (double **p : mv) {
    if ((rc = dalloc_n(C, p, ne))) return rc;
    uftk::fill(s, ne, 0.0, *p);
  }
  if ((rc = dalloc_n(C, &C->d_pij, (int64_t)m * ne))) return rc;
  if ((rc = dalloc_n(C, &C->d_qij, (int64_t)m * ne))) return rc;
  C->mk = 0;
  for (int j = 0; j < m; ++j) {
    C->a[j] = 0.0;
    C->c[j] = 1000.0;
    C->dd[j] = 0.0;
  }
  C->beta = C->P.beta0;
  C->btemp = C->P.beta0;
  if ((rc = sync(C))) return rc;
  C->is_setup = true;
  return uft_autotune(C, 3, nullptr, nullptr, 0);
}

int uft_mesh_get(const uft_ctx *C, int32_t *elem, double *xyz) {
  if (!C || !elem || !xyz || !C->have_mesh) return UFT_ERR_ARG;
  std::memcpy(elem, C->elem.data(), C->elem.size() * sizeof(int32_t));
  std::memcpy(xyz, C->xyz.data(), C->xyz.size() * sizeof(double));
  return UFT_OK;
}

int64_t uft_size(const uft_ctx *C, int which) {
  if (!C) return -1;
  switch (which) {
    case UFT_SIZE_NE: return C->ne;
    case UFT_SIZE_NN: return C->nn;
    case UFT_SIZE_NDOF: return C->ndof;
    case UFT_SIZE_NPE: return C->npe;
  }
  return -1;
}
double uft_total_volume(const uft_ctx *C) { return C ? C->TV : 0.0; }

int uft_vec_get(uft_ctx *C, int id, double *host, int64_t n) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, id) || !host || n != C->vlen[id]) return uft_set_error(C, UFT_ERR_ARG, "uft_vec_get: bad id/size");
  UFT_HIP(C, hipMemcpyAsync(host, C->vec[id], n * sizeof(double), hipMemcpyDeviceToHost, C->s));
  return sync(C);
}
int uft_vec_set(uft_ctx *C, int id, const double *host, int64_t n) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, id) || !host || n != C->vlen[id]) return uft_set_error(C, UFT_ERR_ARG, "uft_vec_set: bad id/size");
  UFT_HIP(C, hipMemcpyAsync(C->vec[id], host, n * sizeof(double), hipMemcpyHostToDevice, C->s));
  return sync(C);
}
int uft_vec_fill(uft_ctx *C, int id, double v) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, id)) return uft_set_error(C, UFT_ERR_ARG, "uft_vec_fill: bad id");
  uftk::fill(C->s, C->vlen[id], v, C->vec[id]);
  return UFT_OK;
}
int uft_vec_copy(uft_ctx *C, int src, int dst) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, src) || !valid_vec(C, dst) || C->vlen[src] != C->vlen[dst])
    return uft_set_error(C, UFT_ERR_ARG, "uft_vec_copy: bad ids");
  uftk::copy(C->s, C->vlen[src], C->vec[src], C->vec[dst]);
  return UFT_OK;
}
void *uft_vec_device(uft_ctx *C, int id) {
  if (!C || !C->is_setup || !valid_vec(C, id)) return nullptr;
  return C->vec[id];
}

int uft_fem_obj_sens(uft_ctx *C, double *fx, double *gx, int dfdx_id, int m, const int *dgdx_ids, int xphys_id) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!fx || !gx || !valid_vec(C, dfdx_id) || !valid_vec(C, xphys_id) || m < 1 || m > C->P.m || !dgdx_ids)
    return uft_set_error(C, UFT_ERR_ARG, "uft_fem_obj_sens: bad arguments");
  for (int j = 0; j < m; ++j)
    if (!valid_vec(C, dgdx_ids[j]) || C->vlen[dgdx_ids[j]] != C->ne)
      return uft_set_error(C, UFT_ERR_ARG, "uft_fem_obj_sens: bad dgdx id");
  const double *xp = C->vec[xphys_id];
  double *U = C->vec[UFT_U];
  const double t0 = now_ms();
  C->count_on = true;
  const int src = C->bt.nseg  ? solve_el_batch(C, xp, U)
                  : C->mr.nr  ? solve_el_block(C, xp)
                  : is_mixed(C) ? solve_el_mixed(C, xp, U)
                                : solve_el(C, xp, U);
  C->count_on = false;
  const double t1 = now_ms();
  C->t_phase[1] = t1 - t0 - C->t_phase[0];
  C->solve_ms += C->t_phase[1];
  C->solve_count++;
  if (src != UFT_OK && src != UFT_ERR_NOCONV) return src;
  if (C->bt.nseg) {  // per design: compliance, volume constraint (SolveFem.cpp:208-249)
    BatchState &B = C->bt;
    uftk::sens(C->s, C, xp, U, C->vec[dfdx_id], C->vec[dgdx_ids[0]], C->d_part, B.E.g(), B.d_invTV);
    uftk::seg_final(C->s, C->d_part, B.E.d_dfirst, B.nseg, 2, B.sc, B_RED);
    if ((rc = fetch(C, B.sc, B.nseg * UFT_BSC, B.h))) return rc;
    fx[0] = 0.0;
    for (int j = 0; j < m; ++j) gx[j] = 0.0;
    double gmax = -1e300;
    for (int d = 0; d < B.nseg; ++d) {
      const double *S = B.h + (size_t)d * UFT_BSC;
      B.fx[d] = S[B_RED];
      B.gx[d] = S[B_RED + 1] / B.TV[d] - B.volfrac[d];
      fx[0] += B.fx[d];
      gmax = std::max(gmax, B.gx[d]);
    }
    gx[0] = gmax;
    C->fx = fx[0];
    C->gx[0] = gx[0];
    C->t_phase[2] = now_ms() - t1;
    return src;
  }
  // SolveFem.cpp:208-249: only constraint 0 is filled by the reference
  uftk::sens(C->s, C, xp, U, C->vec[dfdx_id], C->vec[dgdx_ids[0]], C->d_part);
  uftk::final_sum(C->s, C->d_part, uftk::red_blocks(C->ne), 2, C->d_scal + S_RED);
  if (C->mr.nr) {  // summed over the load cases
    uftk::mr_sens(C->s, C, C->mr.nr, xp, C->mr.U, C->vec[dfdx_id], C->d_part);
    uftk::final_sum(C->s, C->d_part, uftk::mr_sens_blocks(C->ne, C->mr.nr), 1, C->d_scal + S_RED);
  }
  rc = fetch(C, C->d_scal + S_RED, 2, C->h_scal);
  if (rc) return rc;
  fx[0] = C->h_scal[0];
  for (int j = 0; j < m; ++j) gx[j] = 0.0;
  gx[0] = C->h_scal[1] / C->TV - C->P.volfrac;
  C->fx = fx[0];
  C->gx[0] = gx[0];
  C->t_phase[2] = now_ms() - t1;
  return src;
}

int uft_filter(uft_ctx *C, int x_id, int xf_id) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, x_id) || !valid_vec(C, xf_id) || C->vlen[x_id] != C->ne || C->vlen[xf_id] != C->ne)
    return uft_set_error(C, UFT_ERR_ARG, "uft_filter: bad ids");
  uftk::tf_mult(C->s, C, C->vec[x_id], C->d_TX);
  rc = solve_filter(C, C->d_TX, C->vec[UFT_RX], 0);
  uftk::tfx_t(C->s, C, C->vec[UFT_RX], C->vec[xf_id], 1, C->P.Xmin);
  return rc;
}

int uft_filter_sens(uft_ctx *C, int dfdx_id, int m, const int *dgdx_ids) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, dfdx_id) || m < 0 || (m > 0 && !dgdx_ids)) return uft_set_error(C, UFT_ERR_ARG, "uft_filter_sens: bad ids");
  int out = UFT_OK;
  std::vector<int> ids{dfdx_id};
  for (int j = 0; j < m; ++j) ids.push_back(dgdx_ids[j]);
  for (size_t k = 0; k < ids.size(); ++k) {
    const int id = ids[k];
    if (!valid_vec(C, id) || C->vlen[id] != C->ne) return uft_set_error(C, UFT_ERR_ARG, "uft_filter_sens: bad id");
    uftk::tf_mult(C->s, C, C->vec[id], C->d_TX);
    rc = solve_filter(C, C->d_TX, C->vec[UFT_RX], 1 + (int)k);
    if (rc && rc != UFT_ERR_NOCONV) return rc;
    if (rc) out = rc;
    uftk::tfx_t(C->s, C, C->vec[UFT_RX], C->vec[id], 0, 0.0);
  }
  return out;
}

int uft_heaviside(uft_ctx *C, int xf_id, int xp_id, int itr) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, xf_id) || !valid_vec(C, xp_id)) return uft_set_error(C, UFT_ERR_ARG, "uft_heaviside: bad ids");
  if (!C->P.proj) {
    uftk::copy(C->s, C->ne, C->vec[xf_id], C->vec[xp_id]);
    return UFT_OK;
  }
  if (itr < 30)
    C->beta = C->P.beta0;
  else if (itr % 10 == 0)
    C->beta = std::min(C->P.betamax, C->btemp + 1);
  uftk::heaviside(C->s, C->ne, C->vec[xf_id], C->vec[xp_id], C->beta, C->P.x0, C->P.Xmin);
  C->btemp = C->beta;
  return UFT_OK;
}

int uft_projected_sens(uft_ctx *C, int xf_id, int dfdx_id, int m, const int *dgdx_ids) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, xf_id) || !valid_vec(C, dfdx_id)) return uft_set_error(C, UFT_ERR_ARG, "uft_projected_sens: bad ids");
  uftk::projected_sens(C->s, C->ne, C->vec[xf_id], C->beta, C->P.x0, C->vec[dfdx_id]);
  for (int j = 0; j < m; ++j) {
    if (!valid_vec(C, dgdx_ids[j])) return uft_set_error(C, UFT_ERR_ARG, "uft_projected_sens: bad id");
    uftk::projected_sens(C->s, C->ne, C->vec[xf_id], C->beta, C->P.x0, C->vec[dgdx_ids[j]]);
  }
  return UFT_OK;
}

int uft_mma_movelimit(uft_ctx *C, int x_id, int xmin_id, int xmax_id) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, x_id) || !valid_vec(C, xmin_id) || !valid_vec(C, xmax_id))
    return uft_set_error(C, UFT_ERR_ARG, "uft_mma_movelimit: bad ids");
  uftk::mma_movelimit(C->s, C->ne, C->vec[x_id], C->P.Xmin, C->P.Xmax, C->P.movlim, C->vec[xmin_id], C->vec[xmax_id]);
  return UFT_OK;
}

int uft_mma_update(uft_ctx *C, int x_id, int dfdx_id, const double *gx, const int *dgdx_ids, int xmin_id,
                   int xmax_id) {
  int rc = check_setup(C);
  if (rc) return rc;
  const int m = C->P.m;
  const int64_t n = C->ne;
  if (!gx || !dgdx_ids || !valid_vec(C, x_id) || !valid_vec(C, dfdx_id) || !valid_vec(C, xmin_id) ||
      !valid_vec(C, xmax_id))
    return uft_set_error(C, UFT_ERR_ARG, "uft_mma_update: bad arguments");
  const double *dg[UFT_MAXM];
  for (int j = 0; j < m; ++j) {
    if (!valid_vec(C, dgdx_ids[j])) return uft_set_error(C, UFT_ERR_ARG, "uft_mma_update: bad dgdx id");
    dg[j] = C->vec[dgdx_ids[j]];
  }
  hipStream_t s = C->s;
  double *x = C->vec[x_id];
  if (C->bt.nseg) return mma_update_batch(C, x, C->vec[dfdx_id], dg[0], C->vec[xmin_id], C->vec[xmax_id]);
  // GenSub (MMA.cpp:506-637)
  C->mk++;
  uftk::mma_gensub(s, C, n, m, C->mk, x, C->vec[xmin_id], C->vec[xmax_id], C->vec[dfdx_id], dg, C->d_part);
  uftk::final_sum(s, C->d_part, uftk::red_blocks(n), m, C->d_scal + S_RED);
  if ((rc = fetch(C, C->d_scal + S_RED, m, C->h_scal))) return rc;
  for (int j = 0; j < m; ++j) C->b[j] = C->h_scal[j] - gx[j];
  // xo2 := xo1; xo1 := x (MMA.cpp:496-497)
  std::swap(C->d_xo1, C->d_xo2);
  uftk::copy(s, n, x, C->d_xo1);

  // SolveDIP (MMA.cpp:639-676).  One fused device pass per lambda gives x(lambda),
  // the constraint sums (shared by DualGrad and DualResidual) and the Hessian sum.
  const int K = m + m * (m + 1) / 2;
  double S[UFT_MAXM], Hs[UFT_MAXM * UFT_MAXM];
  auto pass = [&]() -> int {
    uftk::mma_dual_pass(s, C, n, m, C->lam, x, C->d_part);
    uftk::final_sum(s, C->d_part, uftk::red_blocks(n), K, C->d_scal + S_RED);
    int r2 = fetch(C, C->d_scal + S_RED, K, C->h_scal);
    if (r2) return r2;
    for (int j = 0; j < m; ++j) S[j] = C->h_scal[j];
    int h = m;
    for (int j = 0; j < m; ++j)
      for (int k = j; k < m; ++k, ++h) Hs[j * m + k] = Hs[k * m + j] = C->h_scal[h];
    return UFT_OK;
  };
  auto yz = [&]() {  // XYZofLAMBDA's scalar part (MMA.cpp:693-701)
    double lamai = 0.0;
    for (int j = 0; j < m; ++j) {
      if (C->lam[j] < 0.0) C->lam[j] = 0.0;
      C->y[j] = std::max(0.0, C->lam[j] - C->c[j]);
      lamai += C->lam[j] * C->a[j];
    }
    C->z = std::max(0.0, 10.0 * (lamai - 1.0));
  };
  for (int j = 0; j < m; ++j) {
    C->lam[j] = C->c[j] / 2.0;
    C->mu[j] = 1.0;
  }
  const double tol = 1.0e-9 * std::sqrt((double)(m + n));
  double epsi = 1.0, err = 1.0;
  int steps = 0;
  yz();
  if ((rc = pass())) return rc;
  while (epsi > tol) {
    int loop = 0;
    while (err > 0.9 * epsi && loop < 100) {
      loop++;
      steps++;
      bool neg = false;
      for (int j = 0; j < m; ++j) neg |= C->lam[j] < 0.0;
      yz();
      if (neg && (rc = pass())) return rc;
      // DualGrad (MMA.cpp:730-765) and the Newton right-hand side
      for (int j = 0; j < m; ++j) C->grad[j] = S[j] - C->b[j] - C->a[j] * C->z - C->y[j];
      for (int j = 0; j < m; ++j) C->grad[j] = -1.0 * C->grad[j] - epsi / C->lam[j];
      // DualHess (MMA.cpp:767-868)
      for (int j = 0; j < m * m; ++j) C->Hess[j] = Hs[j];
      double lamai = 0.0;
      for (int j = 0; j < m; ++j) {
        lamai += C->lam[j] * C->a[j];
        if (C->lam[j] > C->c[j]) C->Hess[j * m + j] += -1.0;
        C->Hess[j * m + j] += -C->mu[j] / C->lam[j];
      }
      if (lamai > 0.0)
        for (int j = 0; j < m; ++j)
          for (int k = 0; k < m; ++k) C->Hess[j * m + k] += -10.0 * C->a[j] * C->a[k];
      double tr = 0.0;
      for (int j = 0; j < m; ++j) tr += C->Hess[j * m + j];
      double corr = 1e-4 * tr / m;
      if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
      for (int j = 0; j < m; ++j) C->Hess[j * m + j] += corr;
      // LU without pivoting + solve (MMA.cpp:936-969)
      double *Kh = C->Hess, *xv = C->grad;
      for (int ss = 0; ss < m - 1; ++ss)
        for (int i = ss + 1; i < m; ++i) {
          Kh[i * m + ss] = Kh[i * m + ss] / Kh[ss * m + ss];
          for (int j = ss + 1; j < m; ++j) Kh[i * m + j] -= Kh[i * m + ss] * Kh[ss * m + j];
        }
      for (int i = 1; i < m; ++i) {
        double acc = 0.0;
        for (int j = 0; j < i; ++j) acc -= Kh[i * m + j] * xv[j];
        xv[i] += acc;
      }
      xv[m - 1] = xv[m - 1] / Kh[(m - 1) * m + (m - 1)];
      for (int i = m - 2; i >= 0; --i) {
        double acc = xv[i];
        for (int j = i + 1; j < m; ++j) acc -= Kh[i * m + j] * xv[j];
        xv[i] = acc / Kh[i * m + i];
      }
      for (int j = 0; j < m; ++j) C->sv[j] = C->grad[j];
      for (int j = 0; j < m; ++j) C->sv[m + j] = -C->mu[j] + epsi / C->lam[j] - C->sv[j] * C->mu[j] / C->lam[j];
      // DualLineSearch (MMA.cpp:870-888)
      double theta = 1.005;
      for (int j = 0; j < m; ++j) {
        if (theta < -1.01 * C->sv[j] / C->lam[j]) theta = -1.01 * C->sv[j] / C->lam[j];
        if (theta < -1.01 * C->sv[j + m] / C->mu[j]) theta = -1.01 * C->sv[j + m] / C->mu[j];
      }
      theta = 1.0 / theta;
      for (int j = 0; j < m; ++j) {
        C->lam[j] += theta * C->sv[j];
        C->mu[j] += theta * C->sv[j + m];
      }
      // XYZofLAMBDA + DualResidual (MMA.cpp:890-934)
      yz();
      if ((rc = pass())) return rc;
      double nrI = 0.0;
      for (int j = 0; j < m; ++j) {
        const double r1 = S[j] - C->b[j] - C->a[j] * C->z - C->y[j] + C->mu[j];
        const double r2 = C->mu[j] * C->lam[j] - epsi;
        nrI = std::max(nrI, std::max(std::fabs(r1), std::fabs(r2)));
      }
      err = nrI;
    }
    epsi = epsi * 0.1;
  }
  C->dip_steps = steps;
  return UFT_OK;
}

int uft_mma_design_change(uft_ctx *C, int x_id, int xold_id, double *ch) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!ch || !valid_vec(C, x_id) || !valid_vec(C, xold_id))
    return uft_set_error(C, UFT_ERR_ARG, "uft_mma_design_change: bad arguments");
  if (C->bt.nseg) {
    BatchState &B = C->bt;
    uftk::design_change(C->s, C->ne, C->vec[x_id], C->vec[xold_id], C->d_part, B.E.g());
    uftk::seg_final_max(C->s, C->d_part, B.E.d_dfirst, B.nseg, B.sc, B_RED);
    if ((rc = fetch(C, B.sc, B.nseg * UFT_BSC, B.h))) return rc;
    *ch = 0.0;
    for (int d = 0; d < B.nseg; ++d) {
      B.ch[d] = B.h[(size_t)d * UFT_BSC + B_RED];
      *ch = std::max(*ch, B.ch[d]);
    }
    return UFT_OK;
  }
  uftk::design_change(C->s, C->ne, C->vec[x_id], C->vec[xold_id], C->d_part);
  uftk::final_max(C->s, C->d_part, uftk::red_blocks(C->ne), C->d_scal + S_RED);
  rc = fetch(C, C->d_scal + S_RED, 1, C->h_scal);
  if (rc) return rc;
  *ch = C->h_scal[0];
  return UFT_OK;
}

int uft_design_iteration(uft_ctx *C, int itr, double *fx, double *gx, double *ch) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!fx || !gx || !ch) return uft_set_error(C, UFT_ERR_ARG, "uft_design_iteration: null output");
  const int m = C->P.m;
  int dg[UFT_MAXM];
  for (int j = 0; j < m; ++j) dg[j] = UFT_DGDX0 + j;
  int status = UFT_OK;
  const double t0 = now_ms();
  rc = uft_fem_obj_sens(C, fx, gx, UFT_DFDX, m, dg, UFT_XPHYS);
  if (rc == UFT_ERR_NOCONV) status = rc; else if (rc) return rc;
  const double t1 = now_ms();
  if (C->P.proj && (rc = uft_projected_sens(C, UFT_XFILED, UFT_DFDX, m, dg))) return rc;
  rc = uft_filter_sens(C, UFT_DFDX, m, dg);
  if (rc == UFT_ERR_NOCONV) status = rc; else if (rc) return rc;
  if ((rc = sync(C))) return rc;
  const double t2 = now_ms();
  if ((rc = uft_mma_movelimit(C, UFT_X, UFT_XMIN, UFT_XMAX))) return rc;
  if ((rc = uft_mma_update(C, UFT_X, UFT_DFDX, gx, dg, UFT_XMIN, UFT_XMAX))) return rc;
  if ((rc = uft_mma_design_change(C, UFT_X, UFT_XOLD, ch))) return rc;
  const double t3 = now_ms();
  rc = uft_filter(C, UFT_X, UFT_XFILED);
  if (rc == UFT_ERR_NOCONV) status = rc; else if (rc) return rc;
  if ((rc = uft_heaviside(C, UFT_XFILED, UFT_XPHYS, itr))) return rc;
  if ((rc = sync(C))) return rc;
  const double t4 = now_ms();
  C->t_phase[3] = t2 - t1;
  C->t_phase[4] = t3 - t2;
  C->t_phase[5] = t4 - t3;
  C->t_phase[6] = t4 - t0;
  return status;
}

int uft_solve_stats(const uft_ctx *C, int *it_el, double *relres_el, int *it_f, int *outer_el) {
  if (!C) return UFT_ERR_ARG;
  if (it_el) *it_el = C->it_el;
  if (relres_el) *relres_el = C->relres_el;
  if (it_f) *it_f = C->it_f;
  if (outer_el) *outer_el = C->outer_el;
  return UFT_OK;
}

int uft_apply_K(uft_ctx *C, int xphys_id, const double *u, double *y) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (!valid_vec(C, xphys_id) || !u || !y) return uft_set_error(C, UFT_ERR_ARG, "uft_apply_K: bad arguments");
  UFT_HIP(C, hipMemcpyAsync(C->d_p, u, C->ndof * sizeof(double), hipMemcpyHostToDevice, C->s));
  uftk::apply_K(C->s, C, C->vec[xphys_id], C->d_p, C->d_q, nullptr, nullptr);
  UFT_HIP(C, hipMemcpyAsync(y, C->d_q, C->ndof * sizeof(double), hipMemcpyDeviceToHost, C->s));
  return sync(C);
}

static int pk_index_h(int r, int c, int nd) {
  if (r > c) std::swap(r, c);
  return r * nd - (r * (r - 1)) / 2 + (c - r);
}

int uft_element_ke(uft_ctx *C, int64_t e, double *ke) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (e < 0 || e >= C->ne || !ke) return uft_set_error(C, UFT_ERR_ARG, "uft_element_ke: bad element");
  const int nd = C->nd, np = nd * (nd + 1) / 2;
  std::vector<double> pk(np);
  for (int p = 0; p < np; ++p)
    UFT_HIP(C, hipMemcpyAsync(&pk[p], C->d_KE + ((size_t)(p >> 1) * C->ne + e) * 2 + (p & 1), sizeof(double),
                              hipMemcpyDeviceToHost, C->s));
  if ((rc = sync(C))) return rc;
  for (int r = 0; r < nd; ++r)
    for (int c = 0; c < nd; ++c) ke[r * nd + c] = pk[pk_index_h(r, c, nd)];
  return UFT_OK;
}

int uft_element_kf(uft_ctx *C, int64_t e, double *kf, double *t) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (e < 0 || e >= C->ne || !kf || !t) return uft_set_error(C, UFT_ERR_ARG, "uft_element_kf: bad element");
  const int npe = C->npe, np = npe * (npe + 1) / 2;
  std::vector<double> pk(np);
  for (int p = 0; p < np; ++p)
    UFT_HIP(C, hipMemcpyAsync(&pk[p], C->d_KF + ((size_t)(p >> 1) * C->ne + e) * 2 + (p & 1), sizeof(double),
                              hipMemcpyDeviceToHost, C->s));
  for (int a = 0; a < npe; ++a)
    UFT_HIP(C, hipMemcpyAsync(&t[a], C->d_T + (size_t)a * C->ne + e, sizeof(double), hipMemcpyDeviceToHost, C->s));
  if ((rc = sync(C))) return rc;
  for (int r = 0; r < npe; ++r)
    for (int c = 0; c < npe; ++c) kf[r * npe + c] = pk[pk_index_h(r, c, npe)];
  return UFT_OK;
}

int uft_sync(uft_ctx *C) {
  if (!C) return UFT_ERR_ARG;
  if (C->amg.lag_pending) UFT_HIP(C, hipStreamSynchronize(C->side[2]));  // a lagged coarsest inverse
  return sync(C);
}

}  // extern "C"

extern "C" int uft_profile_enable(uft_ctx *C, int on) {
  if (!C) return UFT_ERR_ARG;
  C->prof = on > 0;
  C->prof_every = on > 0 ? on : 1;
  C->prof_tick[0] = C->prof_tick[1] = C->prof_tick[2] = 0;
  C->evused = 0;
  C->solve_bytes = C->solve_ms = 0.0;
  C->solve_applies = C->solve_count = 0;
  for (int k = 0; k < 6; ++k) {
    C->prof_ms[k] = 0.0;
    C->prof_n[k] = 0;
  }
  return UFT_OK;
}

extern "C" int uft_profile_read(uft_ctx *C, int kind, int64_t *count, double *total_ms) {
  if (!C || kind < 0 || kind > 5 || !count || !total_ms) return UFT_ERR_ARG;
  UFT_HIP(C, hipStreamSynchronize(C->s));
  // fold the recorded triples into the running totals
  for (size_t i = 0; i + 3 <= C->evused; i += 3) {
    float a = 0.f, b = 0.f;
    UFT_HIP(C, hipEventElapsedTime(&a, C->evpool[i], C->evpool[i + 1]));
    UFT_HIP(C, hipEventElapsedTime(&b, C->evpool[i + 1], C->evpool[i + 2]));
    const int k = 2 * C->evkind[i / 3];
    C->prof_ms[k] += a;
    C->prof_ms[k + 1] += b;
    C->prof_n[k]++;
    C->prof_n[k + 1]++;
  }
  C->evused = 0;
  *count = C->prof_n[kind];
  *total_ms = C->prof_ms[kind];
  return UFT_OK;
}

extern "C" int uft_solver_info(const uft_ctx *C, int *nlev, int64_t *n_per_level, int64_t *blocks_per_level) {
  if (!C || !nlev) return UFT_ERR_ARG;
  if (!C->is_setup) return UFT_ERR_STATE;
  *nlev = C->amg.nlev;
  for (int l = 0; l < C->amg.nlev; ++l) {
    if (n_per_level) n_per_level[l] = C->amg.L[l].n;
    if (blocks_per_level) blocks_per_level[l] = l == 0 ? C->ne : C->amg.L[l].nblk;
  }
  return UFT_OK;
}

extern "C" int uft_solve_traffic(const uft_ctx *C, double *bytes, double *ms, int64_t *applies, int64_t *solves) {
  if (!C) return UFT_ERR_ARG;
  if (bytes) *bytes = C->solve_bytes;
  if (ms) *ms = C->solve_ms;
  if (applies) *applies = C->solve_applies;
  if (solves) *solves = C->solve_count;
  return UFT_OK;
}

extern "C" int uft_fine_operator_info(const uft_ctx *C, int64_t *slots, int64_t *blocks, int *col_bytes,
                                      int *value_bytes) {
  if (!C || !slots || !blocks || !col_bytes || !value_bytes) return UFT_ERR_ARG;
  if (!C->is_setup) return UFT_ERR_STATE;
  const bool on = fine_blocks(C) || C->fell.d_e8;
  *slots = on ? C->fell.nslot : 0;
  *blocks = on ? C->fell.nnzb : 0;
  *col_bytes = on ? (C->fell.d_cd ? 2 : 4) : 0;
  *value_bytes = on ? (C->fell.d_q8 ? 12 : (C->fell.d_v8 ? 18 : (C->fell.d_e8 ? 72 : 36))) : 0;
  return UFT_OK;
}

extern "C" int uft_fine_tiles_info(const uft_ctx *C, int64_t *tiles, int64_t *slots, int *max_halo) {
  if (!C || !tiles || !slots || !max_halo) return UFT_ERR_ARG;
  if (!C->is_setup) return UFT_ERR_STATE;
  *tiles = C->fell.d_tnode ? C->fell.ntile : 0;
  *slots = C->fell.d_tnode ? C->fell.tslot : 0;
  *max_halo = C->fell.d_tnode ? C->fell.thmax : 0;
  return UFT_OK;
}

// ---- MMA settings, restart state, KKT residual (MMA.h:8-52)
extern "C" int uft_mma_set_asymptotes(uft_ctx *C, double init, double decrease, double increase) {
  if (int rc = check_setup(C)) return rc;
  C->asyminit = init;
  C->asymdec = decrease;
  C->asyminc = increase;
  return UFT_OK;
}

extern "C" int uft_mma_set_robust_asymptotes(uft_ctx *C, int type) {
  if (int rc = check_setup(C)) return rc;
  if (type != 0 && type != 1) {
    C->mma_robust = 0;
    return uft_set_error(C, UFT_ERR_ARG, "RobustAsymptotesType cannot be set to: %d (using 0)", type);
  }
  C->mma_robust = type;
  return UFT_OK;
}

extern "C" int uft_mma_constraint_modification(uft_ctx *C, int on) {
  if (int rc = check_setup(C)) return rc;
  C->mma_conmod = on ? 1 : 0;
  return UFT_OK;
}

extern "C" int uft_mma_set_penalization(uft_ctx *C, const double *a, const double *c, const double *d) {
  if (int rc = check_setup(C)) return rc;
  if (!a || !c || !d) return uft_set_error(C, UFT_ERR_ARG, "uft_mma_set_penalization: null array");
  for (int j = 0; j < C->P.m; ++j) {
    C->a[j] = a[j];
    C->c[j] = c[j];
    C->dd[j] = d[j];
  }
  return UFT_OK;
}

extern "C" int uft_mma_state_get(uft_ctx *C, int *k, double *xo1, double *xo2, double *U, double *L) {
  if (int rc = check_setup(C)) return rc;
  if (!k || !xo1 || !xo2 || !U || !L) return uft_set_error(C, UFT_ERR_ARG, "uft_mma_state_get: null output");
  const size_t b = C->ne * sizeof(double);
  *k = C->mk;
  UFT_HIP(C, hipMemcpyAsync(xo1, C->d_xo1, b, hipMemcpyDeviceToHost, C->s));
  UFT_HIP(C, hipMemcpyAsync(xo2, C->d_xo2, b, hipMemcpyDeviceToHost, C->s));
  UFT_HIP(C, hipMemcpyAsync(U, C->d_U, b, hipMemcpyDeviceToHost, C->s));
  UFT_HIP(C, hipMemcpyAsync(L, C->d_L, b, hipMemcpyDeviceToHost, C->s));
  return sync(C);
}

extern "C" int uft_mma_state_set(uft_ctx *C, int k, const double *xo1, const double *xo2, const double *U,
                                 const double *L) {
  if (int rc = check_setup(C)) return rc;
  if (!xo1 || !xo2 || !U || !L || k < 0) return uft_set_error(C, UFT_ERR_ARG, "uft_mma_state_set: bad arguments");
  const size_t b = C->ne * sizeof(double);
  C->mk = k;
  UFT_HIP(C, hipMemcpyAsync(C->d_xo1, xo1, b, hipMemcpyHostToDevice, C->s));
  UFT_HIP(C, hipMemcpyAsync(C->d_xo2, xo2, b, hipMemcpyHostToDevice, C->s));
  UFT_HIP(C, hipMemcpyAsync(C->d_U, U, b, hipMemcpyHostToDevice, C->s));
  UFT_HIP(C, hipMemcpyAsync(C->d_L, L, b, hipMemcpyHostToDevice, C->s));
  if (int rc = sync(C)) return rc;
  if (k < 3) return uft_set_error(C, UFT_OK, "not a legal restart point (k < 3): expect breakdown");
  return UFT_OK;
}

extern "C" int uft_mma_kkt_residual(uft_ctx *C, int x_id, int dfdx_id, const double *gx, const int *dgdx_ids,
                                    int xmin_id, int xmax_id, double *norm2, double *norminf) {
  if (int rc = check_setup(C)) return rc;
  const int m = C->P.m;
  if (!gx || !dgdx_ids || !norm2 || !norminf || !valid_vec(C, x_id) || !valid_vec(C, dfdx_id) ||
      !valid_vec(C, xmin_id) || !valid_vec(C, xmax_id))
    return uft_set_error(C, UFT_ERR_ARG, "uft_mma_kkt_residual: bad arguments");
  const double *dg[UFT_MAXM];
  for (int j = 0; j < m; ++j) {
    if (!valid_vec(C, dgdx_ids[j])) return uft_set_error(C, UFT_ERR_ARG, "uft_mma_kkt_residual: bad dgdx id");
    dg[j] = C->vec[dgdx_ids[j]];
  }
  uftk::mma_kkt(C->s, C->ne, m, C->lam, C->vec[x_id], C->vec[dfdx_id], dg, C->vec[xmin_id], C->vec[xmax_id],
                C->d_part, C->d_scal + S_RED);
  if (int rc = fetch(C, C->d_scal + S_RED, 2, C->h_scal)) return rc;
  double rl = 0.0;
  for (int j = 0; j < m; ++j) rl += C->lam[j] * (C->a[j] * C->z + C->y[j] - gx[j]);
  *norm2 = std::sqrt(C->h_scal[0] + rl * rl);
  *norminf = std::max(C->h_scal[1], std::fabs(rl));
  return UFT_OK;
}

// ---- von Mises stress of the current displacement (stress/ComputeStress.cpp)
extern "C" int uft_stress(uft_ctx *C, int xphys_id, int vms_id) {
  if (int rc = check_setup(C)) return rc;
  if (!valid_vec(C, xphys_id) || !valid_vec(C, vms_id) || C->vlen[xphys_id] != C->ne || C->vlen[vms_id] != C->ne)
    return uft_set_error(C, UFT_ERR_ARG, "uft_stress: bad ids");
  uftk::stress(C->s, C, C->vec[xphys_id], C->vec[UFT_U], C->vec[vms_id]);
  return UFT_OK;
}

// ---- multi-load-case form: per-case results
extern "C" int uft_load_case_stats(const uft_ctx *C, int *iters, double *relres) {
  if (!C || !C->is_setup) return -1;
  const int n = C->mr.nr ? C->mr.ncase : 0;
  for (int k = 0; k < n; ++k) {
    if (iters) iters[k] = C->mr.it[k];
    if (relres) relres[k] = C->mr.relres[k];
  }
  return n;
}

// y_k = K(XPhys) u_k for every load case through the multi-load-case
// operator kernels: kind 0 the FP64 kernel of the outer loop, kind 1 the
// preconditioner's (FP32 vectors; matrix-core BF16 element products, or the
// assembled FP16 node blocks in UFT_MIXED_H).  u, y: host arrays of
// ncase x 3nn, case-major.
extern "C" int uft_apply_K_cases(uft_ctx *C, int kind, const double *u, double *y) {
  if (int rc = check_setup(C)) return rc;
  MrState &M = C->mr;
  if (!u || !y || !M.nr || (kind != 0 && kind != 1))
    return uft_set_error(C, UFT_ERR_ARG, "uft_apply_K_cases: bad arguments or not the multi-load-case form");
  const int nr = M.nr;
  const int64_t n = C->ndof;
  std::vector<double> ui((size_t)n * nr, 0.0), yi((size_t)n * nr);
  for (int k = 0; k < M.ncase; ++k)
    for (int64_t i = 0; i < n; ++i) ui[(size_t)i * nr + k] = u[(size_t)k * n + i];
  C->cur_xphys = C->vec[UFT_XPHYS];
  uftk::simp_moduli(C->s, C, C->cur_xphys, C->d_Ee);  // the matrix-core kernel reads E(x_e) from d_Ee
  if (fine_blocks(C)) {  // the assembled blocks of this density (as a hierarchy refresh makes them)
    AmgLevel &L0 = C->amg.L[0];
    uftk::diag_K(C->s, C, C->cur_xphys, L0.d_dinv, L0.d_dinvf);
    uftk::fine_ell_fill(C->s, C, C->d_Ee, L0.d_dinv);
  }
  if (kind == 0 || !is_mixed(C)) {
    UFT_HIP(C, hipMemcpyAsync(M.p, ui.data(), ui.size() * sizeof(double), hipMemcpyHostToDevice, C->s));
    RowOp<double> op;
    op.mode = ROW_STORE;
    op.y = M.q;
    mr_lev_op<double>(C, 0, M.p, op);
    UFT_HIP(C, hipMemcpyAsync(yi.data(), M.q, yi.size() * sizeof(double), hipMemcpyDeviceToHost, C->s));
  } else {
    std::vector<float> uf(ui.begin(), ui.end()), yf(yi.size());
    UFT_HIP(C, hipMemcpyAsync(M.rf, uf.data(), uf.size() * sizeof(float), hipMemcpyHostToDevice, C->s));
    RowOp<float> op;
    op.mode = ROW_STORE;
    op.y = M.zf;
    mr_lev_op<float>(C, 0, M.rf, op);
    UFT_HIP(C, hipMemcpyAsync(yf.data(), M.zf, yf.size() * sizeof(float), hipMemcpyDeviceToHost, C->s));
    if (int rc = sync(C)) return rc;
    for (size_t i = 0; i < yi.size(); ++i) yi[i] = yf[i];
  }
  if (int rc = sync(C)) return rc;
  for (int k = 0; k < M.ncase; ++k)
    for (int64_t i = 0; i < n; ++i) y[(size_t)k * n + i] = yi[(size_t)i * nr + k];
  return UFT_OK;
}

extern "C" int uft_level_apply(uft_ctx *C, int level, int kind, const double *x, double *y) {
  if (int rc = check_setup(C)) return rc;
  if (!x || !y || level < 0 || level >= C->amg.nlev || kind < 0 || kind > 2)
    return uft_set_error(C, UFT_ERR_ARG, "uft_level_apply: bad arguments");
  AmgLevel &Lv = C->amg.L[level];
  if (kind > 0 && !is_mixed(C)) return uft_set_error(C, UFT_ERR_ARG, "uft_level_apply: FP64 form");
  if (kind == 2 && !(level == 0 ? (void *)C->d_KEh : (void *)Lv.d_valh) && !(level == 0 && smoother_lo(C)))
    return uft_set_error(C, UFT_ERR_ARG, "uft_level_apply: no FP16 operator on level %d", level);
  if (C->refreshes == 0) return uft_set_error(C, UFT_ERR_STATE, "uft_level_apply: no hierarchy refresh yet");
  hipStream_t s = C->s;
  const int64_t n = Lv.n;
  UFT_HIP(C, hipMemcpyAsync(Lv.x, x, n * sizeof(double), hipMemcpyHostToDevice, s));
  if (kind == 0) {
    level_apply(C, level, Lv.x, Lv.t);
  } else {
    uftk::d2f(s, n, Lv.x, Lv.xf);
    RowOp<float> op;
    op.mode = ROW_STORE;
    op.y = Lv.tf;
    if (level == 0 && kind == 2 && smoother_lo(C)) {
      uftk::fine_ell_op(s, C, Lv.xf, op);
    } else if (level == 0 && kind == 1 && C->fell.d_w8) {
      uftk::fine_ell_op32(s, C, Lv.xf, op);
    } else if (level == 0) {
      if (kind == 2)
        uftk::elem_phase_h(s, C, C->elem_variant_h, C->cur_xphys, Lv.xf, C->d_Yef);
      else
        uftk::elem_phase_f(s, C, C->elem_variant_f, C->cur_xphys, Lv.xf, C->d_Yef);
      uftk::gather_op(s, C, Lv.xf, C->d_Yef, op);
    } else if (kind == 2) {
      uftk::bsr6_op_h(s, Lv.nunit, Lv.d_rp, Lv.d_ci, Lv.d_valh, Lv.d_dh, Lv.xf, op);
    } else {
      uftk::bsr6_op(s, Lv.nunit, Lv.d_rp, Lv.d_ci, Lv.d_valf, Lv.xf, op);
    }
    uftk::f2d(s, n, Lv.tf, Lv.t);
  }
  UFT_HIP(C, hipMemcpyAsync(y, Lv.t, n * sizeof(double), hipMemcpyDeviceToHost, s));
  return sync(C);
}

extern "C" int uft_time_fine_operator(uft_ctx *C, int kind, int reps, double *us) {
  if (int rc = check_setup(C)) return rc;
  if (!us || reps < 1 || kind < 0 || kind > 3) return uft_set_error(C, UFT_ERR_ARG, "uft_time_fine_operator: bad arguments");
  if (C->refreshes == 0) return uft_set_error(C, UFT_ERR_STATE, "uft_time_fine_operator: no hierarchy refresh yet");
  if ((kind == 1 || kind == 2) != is_mixed(C))
    return uft_set_error(C, UFT_ERR_ARG, "uft_time_fine_operator: kind %d does not match the form", kind);
  if (kind == 2 && C->P.precision != UFT_MIXED_H && !C->fell.d_w8)
    return uft_set_error(C, UFT_ERR_ARG, "uft_time_fine_operator: kind 2 needs UFT_MIXED_H or assembled FP32 blocks");
  if (C->bt.nseg) return uft_set_error(C, UFT_ERR_ARG, "uft_time_fine_operator: one design");
  if (C->mr.nr && kind == 3)
    return uft_set_error(C, UFT_ERR_ARG, "uft_time_fine_operator: kind 3 takes one load case");
  hipStream_t s = C->s;
  AmgLevel &L0 = C->amg.L[0];
  MrState &M = C->mr;
  const bool cnt = C->count_on;
  C->count_on = false;
  const bool pe = C->prof;  // no per-launch timers inside the measurement
  C->prof = false;
  auto launch = [&]() {
    if (M.nr) {  // several load cases: the same kinds on all right-hand sides at once
      if (kind == 0) {
        RowOp<double> op;
        op.mode = ROW_STORE;
        op.y = M.q;
        mr_lev_op<double>(C, 0, M.p, op);
      } else if (kind == 1) {
        mr_outer_apply_lo(C, M.pf, M.qf);
      } else {
        MLV<float> v = mlv<float>(C, 0);
        const ChebCoef cc = cheb_coef(L0.lmax, std::max(2, level_degree(C, 0)), C->P.cheb_lo / C->P.cheb_hi);
        RowOp<float> st;
        st.mode = ROW_STEP;
        st.r = v.r;
        st.dinv = v.dinv;
        st.x = v.x;
        st.c1 = cc.c1[1];
        st.c2 = cc.c2[1];
        st.dout = v.t;
        st.addprev = 0;
        mr_lev_op<float>(C, 0, (const float *)v.d, st);
      }
      return;
    }
    if (kind == 0) {
      RowOp<double> op;
      op.mode = ROW_STORE;
      op.y = L0.t;
      lev_op(C, 0, (const double *)L0.x, op);
    } else if (kind == 3) {  // the FP64 element phase alone
      uftk::elem_phase(s, C, C->elem_variant, C->cur_xphys, L0.x, C->d_Ye);
    } else if (kind == 1) {
      fine_apply_lo(C, C->cur_xphys, C->d_p, -1, true);
    } else {  // a Chebyshev step of the fine smoother: r -= A d; d = c1 d + c2 D^-1 r; x += d
      LV<float> v = lv(L0, (float *)nullptr);
      const ChebCoef cc = cheb_coef(L0.lmax, std::max(2, level_degree(C, 0)), C->P.cheb_lo / C->P.cheb_hi);
      RowOp<float> st;
      st.mode = ROW_STEP;
      st.r = v.r;
      st.dinv = v.dinv;
      st.x = v.x;
      st.c1 = cc.c1[1];
      st.c2 = cc.c2[1];
      st.dout = v.t;
      st.addprev = 0;
      lev_op(C, 0, (const float *)v.d, st);
    }
  };
  launch();  // warm-up
  hipEvent_t e0, e1;
  UFT_HIP(C, hipEventCreate(&e0));
  UFT_HIP(C, hipEventCreate(&e1));
  UFT_HIP(C, hipEventRecord(e0, s));
  for (int k = 0; k < reps; ++k) launch();
  UFT_HIP(C, hipEventRecord(e1, s));
  int rc = sync(C);
  float ms = 0.f;
  if (!rc) UFT_HIP(C, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  C->prof = pe;
  C->count_on = cnt;
  if (rc) return rc;
  *us = 1e3 * ms / reps;
  return UFT_OK;
}

extern "C" int uft_load_case_u(uft_ctx *C, int k, double *host) {
  if (int rc = check_setup(C)) return rc;
  if (!host || !C->mr.nr || k < 0 || k >= C->mr.ncase)
    return uft_set_error(C, UFT_ERR_ARG, "uft_load_case_u: no load case %d in the multi-load-case form", k);
  uftk::mr_extract(C->s, C->ndof, C->mr.nr, k, C->mr.U, C->d_q);
  UFT_HIP(C, hipMemcpyAsync(host, C->d_q, C->ndof * sizeof(double), hipMemcpyDeviceToHost, C->s));
  return sync(C);
}

extern "C" int uft_amg_operator_get(uft_ctx *C, int level, int32_t *rp, int32_t *ci, double *val) {
  if (int rc = check_setup(C)) return rc;
  if (level < 1 || level >= C->amg.nlev) return uft_set_error(C, UFT_ERR_ARG, "level %d out of range", level);
  const AmgLevel &Lv = C->amg.L[level];
  if (rp) UFT_HIP(C, hipMemcpyAsync(rp, Lv.d_rp, (Lv.nunit + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, C->s));
  if (ci) UFT_HIP(C, hipMemcpyAsync(ci, Lv.d_ci, Lv.nblk * sizeof(int32_t), hipMemcpyDeviceToHost, C->s));
  if (val) UFT_HIP(C, hipMemcpyAsync(val, Lv.d_val, Lv.nblk * 36 * sizeof(double), hipMemcpyDeviceToHost, C->s));
  return sync(C);
}

extern "C" int uft_amg_transfer_get(uft_ctx *C, int level, int32_t *parent, double *P) {
  if (int rc = check_setup(C)) return rc;
  if (level < 0 || level + 1 >= C->amg.nlev) return uft_set_error(C, UFT_ERR_ARG, "level %d out of range", level);
  const AmgLevel &Lv = C->amg.L[level];
  if (parent)
    UFT_HIP(C, hipMemcpyAsync(parent, Lv.d_parent, Lv.nunit * sizeof(int32_t), hipMemcpyDeviceToHost, C->s));
  if (P) UFT_HIP(C, hipMemcpyAsync(P, Lv.d_P, Lv.nunit * Lv.bs * 6 * sizeof(double), hipMemcpyDeviceToHost, C->s));
  return sync(C);
}

// IEEE binary16 bits -> double (host side of the FP16 prolongator copies)
static double half_to_double(uint16_t h) {
  const int sgn = (h >> 15) & 1, ex = (h >> 10) & 31, man = h & 1023;
  double v;
  if (ex == 0) v = std::ldexp((double)man, -24);
  else if (ex == 31) v = man ? NAN : INFINITY;
  else v = std::ldexp(1.0 + man / 1024.0, ex - 15);
  return sgn ? -v : v;
}

extern "C" int uft_amg_prolongator_get(uft_ctx *C, int64_t *nent, int32_t *pptr, int32_t *pagg, double *P,
                                       double *omega) {
  if (int rc = check_setup(C)) return rc;
  if (!nent) return UFT_ERR_ARG;
  const AmgLevel &L0 = C->amg.L[0];
  *nent = L0.sa ? L0.sa_np : 0;
  if (omega) *omega = L0.sa ? L0.sa_omega_used : 0.0;
  if (!L0.sa) return UFT_OK;
  if (pptr) UFT_HIP(C, hipMemcpyAsync(pptr, L0.d_sa_pptr, (L0.nunit + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, C->s));
  if (pagg) UFT_HIP(C, hipMemcpyAsync(pagg, L0.d_sa_pagg, L0.sa_np * sizeof(int32_t), hipMemcpyDeviceToHost, C->s));
  if (int rc = sync(C)) return rc;
  if (P) {  // support order, in the cycle's precision -> entry order, FP64
    std::vector<int32_t> spos(L0.sa_np);
    std::vector<double> ps(18 * L0.sa_np);
    UFT_HIP(C, hipMemcpyAsync(spos.data(), L0.d_sa_spos, L0.sa_np * sizeof(int32_t), hipMemcpyDeviceToHost, C->s));
    if (L0.d_sa_Psh) {
      std::vector<uint16_t> ph(18 * L0.sa_np);
      UFT_HIP(C, hipMemcpyAsync(ph.data(), L0.d_sa_Psh, ph.size() * sizeof(uint16_t), hipMemcpyDeviceToHost, C->s));
      if (int rc = sync(C)) return rc;
      for (size_t k = 0; k < ph.size(); ++k) ps[k] = half_to_double(ph[k]);
    } else if (L0.d_sa_Psf) {
      std::vector<float> pf(18 * L0.sa_np);
      UFT_HIP(C, hipMemcpyAsync(pf.data(), L0.d_sa_Psf, pf.size() * sizeof(float), hipMemcpyDeviceToHost, C->s));
      if (int rc = sync(C)) return rc;
      std::copy(pf.begin(), pf.end(), ps.begin());
    } else {
      UFT_HIP(C, hipMemcpyAsync(ps.data(), L0.d_sa_Ps, ps.size() * sizeof(double), hipMemcpyDeviceToHost, C->s));
      if (int rc = sync(C)) return rc;
    }
    for (int64_t t = 0; t < L0.sa_np; ++t) std::memcpy(P + 18 * t, &ps[18 * (size_t)spos[t]], 18 * sizeof(double));
  }
  return UFT_OK;
}

extern "C" int uft_amg_coarse_inverse_get(uft_ctx *C, double *inv) {
  if (int rc = check_setup(C)) return rc;
  if (!inv) return UFT_ERR_ARG;
  const int64_t n = C->amg.ncoarse;
  if (C->bt.nseg) {  // block-diagonal: scatter each design's block
    const BatchState &B = C->bt;
    std::vector<double> blk(B.coff[B.nseg]);
    UFT_HIP(C, hipMemcpyAsync(blk.data(), C->amg.d_cinv, blk.size() * sizeof(double), hipMemcpyDeviceToHost, C->s));
    if (int rc = sync(C)) return rc;
    std::memset(inv, 0, (size_t)(n * n) * sizeof(double));
    const std::vector<int64_t> &cu = B.uoff[C->amg.nlev - 1];
    for (int d = 0; d < B.nseg; ++d) {
      const int64_t r0 = 6 * cu[d], nd = 6 * (cu[d + 1] - cu[d]);
      for (int64_t i = 0; i < nd; ++i)
        for (int64_t j = 0; j < nd; ++j) inv[(r0 + i) * n + r0 + j] = blk[B.coff[d] + i * nd + j];
    }
    return UFT_OK;
  }
  UFT_HIP(C, hipMemcpyAsync(inv, C->amg.d_cinv, n * n * sizeof(double), hipMemcpyDeviceToHost, C->s));
  return sync(C);
}

// Smoother-schedule autotuner: one elasticity solve at the current XPhys per
// candidate (Chebyshev degree 1-3 x cycle type), the fastest kept.
extern "C" int uft_autotune_smoother(uft_ctx *C, int *degree, int *cycle_type, double *ms) {
  if (int rc = check_setup(C)) return rc;
  const int deg0 = C->P.cheb_degree, ct0 = C->P.cycle_type;
  const int nct = 3;
  const double *xp = C->vec[UFT_XPHYS];
  double *U = C->vec[UFT_U];
  double best = 1e300;
  int bd = deg0, bc = ct0;
  for (int d = 1; d <= 3; ++d)
    for (int ct = 0; ct < nct; ++ct) {
      C->P.cheb_degree = d;
      C->P.cycle_type = ct;
      if (int rc = sync(C)) return rc;
      const double t0 = now_ms();
      const int rc = C->bt.nseg  ? solve_el_batch(C, xp, U)
                     : C->mr.nr  ? solve_el_block(C, xp)
                     : is_mixed(C) ? solve_el_mixed(C, xp, U)
                                   : solve_el(C, xp, U);
      if (rc != UFT_OK && rc != UFT_ERR_NOCONV) {
        C->P.cheb_degree = deg0;
        C->P.cycle_type = ct0;
        return rc;
      }
      if (int r2 = sync(C)) return r2;
      const double t = rc == UFT_OK ? now_ms() - t0 : 1e300;
      if (ms) ms[(d - 1) * 4 + ct] = t;
      if (t < best) {
        best = t;
        bd = d;
        bc = ct;
      }
    }
  C->P.cheb_degree = bd;
  C->P.cycle_type = bc;
  if (degree) *degree = bd;
  if (cycle_type) *cycle_type = bc;
  if (best >= 1e300) return uft_set_error(C, UFT_ERR_NOCONV, "uft_autotune_smoother: no schedule converged");
  return UFT_OK;
}

extern "C" int uft_batch_set(uft_ctx *C, int nd, const int64_t *eoff, const int64_t *noff, const double *volfrac,
                             const double *rmin) {
  if (!C) return UFT_ERR_ARG;
  if (C->is_setup || !C->have_mesh) return uft_set_error(C, UFT_ERR_STATE, "uft_batch_set: after the mesh, before setup");
  if (C->bt.nseg) return uft_set_error(C, UFT_ERR_STATE, "uft_batch_set: already set");
  if (nd < 1 || nd > UFT_MAXSEG || !eoff || !noff || !volfrac || !rmin)
    return uft_set_error(C, UFT_ERR_ARG, "uft_batch_set: 1 to %d designs with offsets, volfrac and rmin", UFT_MAXSEG);
  if (eoff[0] != 0 || noff[0] != 0 || eoff[nd] != C->ne || noff[nd] != C->nn)
    return uft_set_error(C, UFT_ERR_ARG, "uft_batch_set: offsets must run from 0 to ne and nn");
  for (int d = 0; d < nd; ++d) {
    if (eoff[d + 1] <= eoff[d] || noff[d + 1] <= noff[d])
      return uft_set_error(C, UFT_ERR_ARG, "uft_batch_set: design %d is empty", d);
    if (!(volfrac[d] > 0.0 && volfrac[d] <= 1.0) || !(rmin[d] > 0.0))
      return uft_set_error(C, UFT_ERR_ARG, "uft_batch_set: design %d: volfrac in (0, 1] and rmin > 0", d);
    for (int64_t e = eoff[d]; e < eoff[d + 1]; ++e)
      for (int a = 0; a < C->npe; ++a) {
        const int64_t v = C->elem[e * C->npe + a];
        if (v < noff[d] || v >= noff[d + 1])
          return uft_set_error(C, UFT_ERR_ARG, "uft_batch_set: element %lld uses node %lld outside design %d",
                               (long long)e, (long long)v, d);
      }
  }
  BatchState &B = C->bt;
  if (hipHostMalloc((void **)&B.h, ((size_t)nd * UFT_BSC + 2 * UFT_MAXSEG) * sizeof(double), hipHostMallocDefault) !=
      hipSuccess)
    return uft_set_error(C, UFT_ERR_OOM, "uft_batch_set: pinned host buffer");
  B.eoff.assign(eoff, eoff + nd + 1);
  B.noff.assign(noff, noff + nd + 1);
  B.volfrac.assign(volfrac, volfrac + nd);
  B.rmin.assign(rmin, rmin + nd);
  B.fx.assign(nd, 0.0);
  B.gx.assign(nd, 0.0);
  B.ch.assign(nd, 0.0);
  B.relres.assign(nd, 0.0);
  B.it.assign(nd, 0);
  B.itf.assign(nd, 0);
  B.lam.assign(nd, 0.0);
  B.mu.assign(nd, 0.0);
  B.b.assign(nd, 0.0);
  B.nseg = nd;
  return UFT_OK;
}

extern "C" int uft_batch_results(const uft_ctx *C, double *fx, double *gx, double *ch, int *it_el, double *relres_el,
                                 int *it_f) {
  if (!C) return UFT_ERR_ARG;
  const BatchState &B = C->bt;
  for (int d = 0; d < B.nseg; ++d) {
    if (fx) fx[d] = B.fx[d];
    if (gx) gx[d] = B.gx[d];
    if (ch) ch[d] = B.ch[d];
    if (it_el) it_el[d] = B.it[d];
    if (relres_el) relres_el[d] = B.relres[d];
    if (it_f) it_f[d] = B.itf[d];
  }
  return B.nseg;
}

extern "C" int uft_phase_times(const uft_ctx *C, double *ms) {
  if (!C || !ms) return UFT_ERR_ARG;
  for (int k = 0; k < 7; ++k) ms[k] = C->t_phase[k];
  return UFT_OK;
}

// Element-phase autotuner: every variant computes bit-identical results, so
// the choice only affects speed.  Each variant is timed over `reps` launches
// with HIP events on the context stream (one untimed launch first).
extern "C" int uft_autotune(uft_ctx *C, int reps, int *variant, double *us_per_variant, int nmax) {
  int rc = check_setup(C);
  if (rc) return rc;
  if (reps < 1) reps = 1;
  const int nv = uftk::elem_variant_count();
  hipEvent_t e0, e1;
  UFT_HIP(C, hipEventCreate(&e0));
  UFT_HIP(C, hipEventCreate(&e1));
  const double *xp = C->vec[UFT_XPHYS];
  // kind 0: FP64 blocks, 1: FP32 blocks, 2: FP16 smoother blocks
  auto run = [&](int kind, int v) {
    if (kind == 1) uftk::elem_phase_f(C->s, C, v, xp, C->d_rf, C->d_Yef);
    else if (kind == 2) uftk::elem_phase_h(C->s, C, v, xp, C->d_rf, C->d_Yef);
    else uftk::elem_phase(C->s, C, v, xp, C->d_r, C->d_Ye);
  };
  auto tune = [&](int kind, int &best, double *us) -> int {
    double bt = 1e300;
    for (int v = 0; v < nv; ++v) {
      run(kind, v);
      UFT_HIP(C, hipEventRecord(e0, C->s));
      for (int k = 0; k < reps; ++k) run(kind, v);
      UFT_HIP(C, hipEventRecord(e1, C->s));
      UFT_HIP(C, hipEventSynchronize(e1));
      float ms = 0.f;
      UFT_HIP(C, hipEventElapsedTime(&ms, e0, e1));
      us[v] = 1e3 * ms / reps;
      if (us[v] < bt) {
        bt = us[v];
        best = v;
      }
    }
    return UFT_OK;
  };
  uftk::fill(C->s, C->ndof, 1.0, C->d_r);
  if ((rc = tune(0, C->elem_variant, C->tune_us))) return rc;
  if (C->d_KEf) {
    uftk::fillf(C->s, C->ndof, 1.0f, C->d_rf);
    if ((rc = tune(1, C->elem_variant_f, C->tune_us_f))) return rc;
  }
  if (C->d_KEh)
    if ((rc = tune(2, C->elem_variant_h, C->tune_us_h))) return rc;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (variant) {
    variant[0] = C->elem_variant;
    if (nmax > 1) variant[1] = C->elem_variant_f;
    if (nmax > 2) variant[2] = C->elem_variant_h;
  }
  if (us_per_variant)
    for (int v = 0; v < nv && v < nmax; ++v) us_per_variant[v] = C->tune_us[v];
  return sync(C);
}
