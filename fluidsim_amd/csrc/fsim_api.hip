// fsim_api.hip -- host side of libfsim_hip: handle, grid, initial state, the
// step orchestration and the C ABI declared in include/fsim.h.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "fsim.h"
#include "fsim_internal.h"
#include "fsim_learn.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, const char* a = "", int b = 0) {
  char buf[512];
  snprintf(buf, sizeof(buf), fmt, a, b);
  g_err = buf;
  return code;
}

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return fail(FS_E_HIP, "HIP error %s at line %d",        \
                                      hipGetErrorString(e_), __LINE__);          \
  } while (0)

// x0**6 correctly rounded (gfortran folds the parameter expression of
// ptcl_props.F90:18 exactly); x0 has a 24-bit significand so x0**4 is exact in
// binary128 and one rounding follows.
double pow6_exact(double x) {
  __float128 q = (__float128)x;
  __float128 q2 = q * q, q4 = q2 * q2;
  return (double)(q4 * q2);
}

double flog32(float v) { return (double)(float)log((double)v); }

}  // namespace

// ---------------------------------------------------------------------------
// constants: props.F90:24-44 and ptcl_props.F90:5-25 (p and gam at runtime)
void fs_phys_init(FsPhys* K, double p, double gam) {
  memset(K, 0, sizeof(*K));
  K->pi = 4.0 * atan(1.0);
  K->eps0 = F32(8.85418782e-12);
  K->mu0 = 4.0 * K->pi * F32(1e-7);
  K->c0 = 1.0 / sqrt(K->eps0 * K->mu0);
  K->e = F32(1.60217646e-19);
  K->kb = F32(1.3806503e-23);
  K->me = 9.10938188e-31;
  K->x0 = F32(1e-4);
  K->x0_2 = K->x0 * K->x0;
  K->x0_3 = K->x0_2 * K->x0;
  K->x0_6 = pow6_exact(K->x0);
  K->ph0 = K->e / (K->eps0 * K->x0);
  K->t0 = F32(1e-6);
  K->wp02 = (K->t0 * K->t0) * (K->e * K->e) / (K->eps0 * K->x0_3 * K->me);
  double tw = K->t0 * F32(14e9);
  K->w2 = tw * tw;
  K->Tg = 300.0;
  K->p = p;
  K->ninf = p * 101325.0 / 760.0 / K->kb / K->Tg * K->x0_3;
  K->n_zero = F32(1e8) * K->x0_3;
  K->n_init = F32(1e11) * K->x0_3;
  K->mi = 1.67262178e-27 * 39.948;
  K->mui = 1.45e3 / p * 1e-4 * K->ph0 * K->t0 / K->x0_2;
  K->vi = sqrt((8.0 * K->kb * K->Tg) / (K->pi * K->mi)) * K->t0 / K->x0;
  K->Di = K->mui * (K->kb * K->Tg / K->e) / K->ph0;
  K->Dm = 2.42e18 * K->x0 / K->ninf / 1e2 * K->t0;
  K->k_r = 2e-7 / 1e6 / K->x0_3 * K->t0;
  K->k_mp = 6.2e-10 / 1.0e6 / K->x0_3 * K->t0;
  K->k_2q = 3e-15 / 1e6 / K->x0_3 * K->t0;
  K->k_3q = 1.1e-31 / 1.0e12 / K->x0_6 * K->t0;
  K->H_ir = 15.80 / K->ph0;
  K->H_ex = 11.56 / K->ph0;
  K->H_si = 4.14 / K->ph0;
  K->H_sc = -11.56 / K->ph0;
  K->gam = gam;
  K->log5119 = flog32(5.119f);
  K->log5667 = flog32(5.667f);
  K->log75 = flog32(7.5f);
  K->ftol = F32(1e-12);
  K->Cap = 1e-11 * K->ph0 / K->e;
  K->kTe = K->Tg * K->kb / K->e / K->ph0;
  K->nmin_t = K->n_zero / K->ph0 / 1e2;
  K->c53 = 5.0 / 3.0;
  K->ve_num = 16.0 * K->e * K->ph0;
  K->ve_den = 3.0 * K->pi * K->me;
  K->qvi = 2.5e-1 * K->vi;
  K->hvi = 5e-1 * K->vi;
  K->k2q_ninf = K->k_2q * K->ninf;
  K->k3q_ninf2 = K->k_3q * (K->ninf * K->ninf);
  K->kmp2 = 2.0 * K->k_mp;
  K->log_kex_lo = log(9.611e-1);
  K->log_kir_lo = log(1.37);
  K->log_s_lo = log(5e-1);
  K->log_s_hi = log(16.0);
  K->exp_m100 = exp(-100.0);
  K->c_rate = K->t0 / K->x0_3;
  K->c_rate6 = K->t0 / (1.0e6 * K->x0_3);
  K->c_nu = K->ninf * K->t0 / K->x0_3;
  K->c_beta = 1e-13 * K->t0 / K->x0_3;
  K->me_mi = K->me / K->mi;
  K->c_mue = K->x0 / K->ninf * K->t0 * K->ph0;
  K->c_de = K->x0 / K->ninf * K->t0;
  K->r_Di = 1.0 / K->Di;
  {
    const long double ln10 = logl(10.0L);
    K->beta_a = (double)(ln10 * (0.944L + 0.665L * (long double)2.48f) -
                         0.665L * logl(11.604505e3L / 1.5L));
    const double x = log10(5e4);
    K->beta_hi = exp10(0.944 - 0.665 * (x - (double)2.48f)) * K->c_beta;
  }
}

void fs_phys_table(const FsPhys* K, double* o) {
  const double v[40] = {K->pi,   K->eps0, K->mu0,  K->e,    K->kb,    K->me,      K->x0,
                        K->ph0,  K->t0,   K->wp02, K->w2,   K->Tg,    K->p,       K->ninf,
                        K->n_zero, K->n_init, K->mi, K->mui, K->vi,   K->Di,      K->Dm,
                        K->k_r,  K->k_mp, K->k_2q, K->k_3q, K->H_ir,  K->H_ex,    K->H_si,
                        K->H_sc, K->gam,  K->log5119, K->log5667, K->log75, K->ftol, K->Cap,
                        K->x0_3, K->c0,   K->kTe,  K->nmin_t, K->x0_6};
  memcpy(o, v, sizeof(v));
}

// ---------------------------------------------------------------------------
constexpr int kChainLog = 5;  // MG-PCG iteration graphs of 1, 2, 4, 8, 16 iterations
constexpr int kSolveGraphs = 4;  // whole-first-chunk solve graphs: 1 .. 4 iterations after 0

struct fs_handle {
  fs_grid_desc g;
  fs_params p;
  FsPhys K;
  int bx, by, P;
  size_t PL;
  std::vector<double> x, y, dx, dlx, dy, dly, r;  // reference (1-based shifted to 0)
  std::vector<int32_t> type_x, type_y;
  // device
  hipStream_t st = nullptr;
  double* d_lev = nullptr;  // 4 buffers x 4 species x PL
  double* d_k = nullptr;    // k1,k3,k4 x 4 species x PL
  double* d_kdbg = nullptr; // 5 stages x 4 species x PL (debug)
  double* d_ph = nullptr;
  double* d_sig = nullptr;  // dielectric wall: sig_pl, sig_mi, sig_or, ks(1..5), bx+2 each
  double* d_geo = nullptr;  // dx,dlx,dy,dly,r and 1/dx,1/dlx,1/dy,1/dly
  int8_t* d_types = nullptr;  // ltype, rtype, gl, gr (by each)
  double* d_dr = nullptr;     // dra, drc
  double* d_cur = nullptr;  // electrode face currents, 2*by per stage parity
  double* d_part = nullptr;
  double* d_lspart = nullptr;  // projected warm start: nine sums per block
  DevScalars* d_sc = nullptr;
  DevScalars* h_sc = nullptr;  // pinned mirror
  DevScalars* h_pub = nullptr; // host-mapped copy written by k_publish (and its device address)
  DevScalars* d_pub = nullptr;
  uint32_t* d_pubcnt = nullptr;  // publications run (device)
  uint32_t pub_tok = 0;          // publications enqueued (host)
  bool spec_taken = false;       // the next pass's k_begin_step and stage 1 are enqueued and run
  double* d_pois = nullptr;    // rhs0 r0 cx cy dg x p r1 w0 w1 s0 s1 (kPois planes)
  double* d_dbg = nullptr;     // b (N) + A5 (5N)
  double* d_replay = nullptr;
  int nreplay = 0;
  int rot = 0;        // level rotation: A = buf[rot], L1 = buf[rot+1], L3 = buf[rot+2]
  bool alias = true;  // level 1 currently equals level 2 (after accept/reject/init)
  int nbx = 0, nby = 0, nblk = 0;   // assembly tiles (64x8)
  int sty = 8;                      // stage tile height (4, 8 or 16: fs_params.stage_tile)
  size_t npart = 0;                 // entries of one partial-sum buffer
  int pnby = 0, pnblk = 0;          // PCG tiles (64x8)
  std::vector<MgLevel> mg;          // multigrid hierarchy; level 0 aliases the PCG planes
  double* d_mg = nullptr;
  double* d_mginv = nullptr;        // dense inverse of the coarsest operator
  float* d_mgpk = nullptr;          // packed FP32 rows of the one-workgroup levels
  size_t mg_first = 0;              // first level of the one-workgroup coarse kernel
  int last_iters[3] = {0, 0, 0};    // iterations of the last solve, per solver
  fs_learn* net = nullptr;          // learned coarse solver (lapl_solver = 2)
  // device-resident kernel arguments: stage [(rot*2 + alias)*3 + variant][stage],
  // variant 0 step, 1 slopes only (debug copy), 2 slopes only (no copy);
  // assembly [rot*2 + alias]; the physics constants
  StageArgs* d_stab = nullptr;
  LaplArgs* d_ltab = nullptr;
  FsPhys* d_phys = nullptr;
  hipGraphExec_t mg_pro[2] = {};      // MG-PCG set-up and iteration 0 (from zero, warm)
  // set-up, iteration 0, k = 1 .. kSolveGraphs further iterations and the
  // finishing kernels in one graph (from zero, warm)
  hipGraphExec_t mg_solve[2][kSolveGraphs] = {};
  bool warm_ok = false;               // x holds the last converged MG-PCG correction
  bool moved = false;                 // an accepted transport step since the last solve
  bool in_ens = false;                // a member of an ensemble (its tables are the ensemble's)
  bool solve_warm = false;            // the MG-PCG solve in flight started from the last correction
  int mg_it = 0, mg_chunk = 0;        // next MG-PCG iteration / chunk of a solve in progress
  hipGraphExec_t mg_chain[2][kChainLog] = {};  // 2^m MG-PCG iterations from parity p
  bool mg_graphs = false;             // every mg_pro / mg_solve / mg_chain graph is captured
  hipGraphExec_t learn_graph[2] = {}; // one learned-FCG iteration, by parity
  // what the learned graphs were captured over: the network (or the handle:
  // none attached) and the allocation number of its inference activations,
  // which net_apply reallocates when a handle of another level size uses it
  const void* learn_graph_net = nullptr;
  uint64_t learn_graph_acts = 0;
  Geo G;
  std::vector<double> log;
  bool timing = false;
  hipEvent_t ev[8] = {};
  double tms[4] = {0, 0, 0, 0};
  std::vector<double> stage_buf;

  double* lev(int buf, int sp) const { return d_lev + ((size_t)buf * 4 + sp) * PL; }
  int bufA() const { return rot; }
  int bufL1() const { return (rot + 1) % 3; }
  int bufL3() const { return (rot + 2) % 3; }
  int bufB() const { return 3; }
  double* kk(int which, int sp) const { return d_k + ((size_t)which * 4 + sp) * PL; }
  double* pois(int which) const { return d_pois + (size_t)which * PL; }
  double* part(int which) const { return d_part + (size_t)which * 8 * npart; }
};

constexpr int kPois = 19;
enum { PO_RHS0 = 0, PO_R0 = 1, PO_CX = 2, PO_CY = 3, PO_DG = 4, PO_X = 5, PO_P = 6, PO_R1 = 7,
       PO_W0 = 8, PO_W1 = 9, PO_S0 = 10, PO_S1 = 11,
       // multigrid-preconditioned PCG: p (double-buffered), q, u = B r, pre-smoothed x1
       PO_MP0 = 12, PO_MP1 = 13, PO_MQ = 14, PO_MU = 15, PO_MX1 = 16,
       // the MG-PCG correction before the last (second-order warm start)
       PO_X2 = 17, PO_X3 = 18 };
constexpr double kMgOver = 1.8;    // scale of the piecewise-constant coarse correction

namespace {

// props.F90:96-124 (x) / :134-163 (y); kind 2 = geometric spacing ratio**min(k, n-k)
void build_axis(int kind, double ratio, int n, double len, int is_y, std::vector<double>& x) {
  x.assign(n + 2, 0.0);
  if (kind == 0) {
    double fl = (double)(float)(n + 1);
    for (int i = 0; i < n + 2; i++) x[i] = len / fl * (double)i;
  } else if (kind == 1) {
    float xt = is_y ? 0.886f / (float)(n + 1) : (2.0f * 1.175f) / (float)(n + 1);
    double c = is_y ? F32(-0.886) : F32(-1.175);
    for (int i = 0; i < n + 2; i++) x[i] = tanh(c + (double)xt * (double)i);
  } else {
    x[0] = 0.0;
    for (int k = 0; k <= n; k++) {
      int m = k < n - k ? k : n - k;
      x[k + 1] = x[k] + pow(ratio, (double)m);
    }
  }
  double t = x[0];
  for (auto& v : x) v = v - t;
  t = x[n + 1];
  for (auto& v : x) v = v / t;
  for (auto& v : x) v = v * len;
}

int upload(fs_handle* h, double* dst, const double* src, size_t n) {
  HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, h->st));
  HIP_OK(hipStreamSynchronize(h->st));
  return 0;
}

// Waits until the last publication enqueued (k_publish) has reached host
// memory, and copies it to h_sc.  The poll asks the stream every 1024 rounds,
// so that an error, or a stream that drained without publishing (which would
// be a bug), ends the wait.
int wait_published(fs_handle* h) {
  const uint32_t tok = h->pub_tok;
  const volatile uint32_t* seq = &h->h_pub->pub_seq;
  for (unsigned k = 1; *seq != tok; k++) {
    if ((k & 1023) == 0) {
      const hipError_t q = hipStreamQuery(h->st);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) HIP_OK(q);
      if (*seq != tok) return fail(FS_E_HIP, "scalars were not published%s");
      break;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  memcpy(h->h_sc, const_cast<const DevScalars*>(h->h_pub), sizeof(DevScalars));
  return 0;
}

void enqueue_publish(fs_handle* h) {
  launch_publish(h->d_sc, h->d_pub, h->d_pubcnt, h->st);
  h->pub_tok++;
}

// Reads the device scalars once the work enqueued so far has run.  With
// host-mapped memory (and no timing events to read) a kernel publishes them:
// no copy command and no stream synchronisation on the path from the last
// kernel to the next launch.
int sync_scalars(fs_handle* h) {
  if (!h->h_pub || h->timing) {
    HIP_OK(hipMemcpyAsync(h->h_sc, h->d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, h->st));
    HIP_OK(hipStreamSynchronize(h->st));
    return 0;
  }
  enqueue_publish(h);
  HIP_OK(hipGetLastError());
  return wait_published(h);
}

int push_scalars(fs_handle* h) {
  HIP_OK(hipMemcpyAsync(h->d_sc, h->h_sc, sizeof(DevScalars), hipMemcpyHostToDevice, h->st));
  HIP_OK(hipStreamSynchronize(h->st));
  return 0;
}

int copy_level(fs_handle* h, int dst, int src) {
  for (int s = 0; s < 4; s++)
    HIP_OK(hipMemcpyAsync(h->lev(dst, s), h->lev(src, s), h->PL * sizeof(double),
                          hipMemcpyDeviceToDevice, h->st));
  return 0;
}

void time_mark(fs_handle* h, int e) {
  if (h->timing) (void)hipEventRecord(h->ev[e], h->st);
}

// The electrode face currents of a stage (Ida terms, then Idc terms), by stage
// parity: a stage launch that also does the previous stage's circuit update
// reads that stage's currents while writing its own.
double* cur_buf(fs_handle* h, int stage) { return h->d_cur + (size_t)(stage & 1) * 2 * h->by; }

CircArgs circ_args(fs_handle* h, int stage);

// fuse_circ: the launch of stages 2-5 also does the previous stage's circuit
// update (StageArgs::cc)
StageArgs stage_args(fs_handle* h, int stage, int rhs_only, bool fuse_circ = false) {
  StageArgs a;
  memset(&a, 0, sizeof(a));
  // stage 1 reads A writes B; 2 reads B writes A; 3 A->B; 4 B->A; 5 A->L1
  int rd = (stage == 2 || stage == 4) ? h->bufB() : h->bufA();
  int wr = stage == 5 ? h->bufL1() : ((stage == 1 || stage == 3) ? h->bufB() : h->bufA());
  if (rhs_only) rd = h->bufA();
  for (int s = 0; s < 4; s++) {
    a.rd[s] = h->lev(rd, s);
    a.wr[s] = h->lev(wr, s);
    a.n3[s] = h->lev(h->bufL3(), s);
    a.k1[s] = h->kk(0, s);
    a.k3[s] = h->kk(1, s);
    a.k4[s] = h->kk(2, s);
    a.kdbg[s] = h->d_kdbg ? h->d_kdbg + ((size_t)(stage - 1) * 4 + s) * h->PL : nullptr;
  }
  a.ph = h->d_ph;
  a.cur = cur_buf(h, stage);
  a.part = h->part(0);
  a.sc = h->d_sc;
  a.G = h->G;
  a.stage = stage;
  a.rhs_only = rhs_only;
  a.nbx = h->nbx;
  a.nby = h->nby;
  if (fuse_circ && stage >= 2) a.cc = circ_args(h, stage - 1);
  return a;
}

CircArgs circ_args(fs_handle* h, int stage) {
  CircArgs c;
  memset(&c, 0, sizeof(c));
  c.ph = h->d_ph;
  c.cur = cur_buf(h, stage);
  c.dra = h->d_dr;
  c.drc = h->d_dr + h->by;
  c.gl = h->d_types + 2 * h->by;
  c.gr = h->d_types + 3 * h->by;
  c.sc = h->d_sc;
  c.vset = h->p.vset;
  c.res = h->p.res_ohm * h->K.e / (h->K.ph0 * h->K.t0);  // main.F90:286
  c.t_rise = h->p.t_rise;
  c.t_width = h->p.t_width;
  c.t_fall = h->p.t_fall;
  c.rf = h->p.rf;
  c.circuit = h->p.circuit;
  c.oned = h->G.oned;
  c.stage = stage;
  c.by = h->by;
  c.P = h->P;
  c.bx = h->bx;
  return c;
}

// PCG buffers for iteration `it` (r, w, s double-buffered by parity)
PcgArgs pcg_args(fs_handle* h, int it) {
  PcgArgs a;
  memset(&a, 0, sizeof(a));
  int par = it & 1;
  a.cx = h->pois(PO_CX);
  a.cy = h->pois(PO_CY);
  a.dg = h->pois(PO_DG);
  a.r_old = h->pois(par ? PO_R1 : PO_R0);
  a.r_new = h->pois(par ? PO_R0 : PO_R1);
  a.w_old = h->pois(par ? PO_W1 : PO_W0);
  a.w_new = h->pois(par ? PO_W0 : PO_W1);
  a.s_old = h->pois(par ? PO_S1 : PO_S0);
  a.s_new = h->pois(par ? PO_S0 : PO_S1);
  a.p = h->pois(PO_P);
  a.x = h->pois(PO_X);
  a.rhs0 = h->pois(PO_RHS0);
  a.ph = h->d_ph;
  a.part_in = h->part(par);
  a.part_out = h->part(par ^ 1);
  a.sc = h->d_sc;
  a.G = h->G;
  a.nbx = h->nbx;
  a.nby = h->pnby;
  a.it = it;
  return a;
}

int check_status(fs_handle* h) {
  int st = h->h_sc->status;
  if (st < 0) {
    const char* what = st == FS_E_NEGDENS ? "negative density in get_Te"
                       : st == FS_E_NAN   ? "NaN in a stage slope"
                       : st == FS_E_DTMIN ? "minimum timestep reached"
                       : st == FS_E_NANJ  ? "NaN in dFdPh"
                       : st == FS_E_REPLAY ? "replay table exhausted"
                                           : "device error";
    return fail(st, "fatal: %s (status %d)", what, st);
  }
  return 0;
}

// Aggregation hierarchy: halve both axes until the coarsest level has at most
// 64 cells (solved with its dense inverse).  Coarse planes: d (ghost 1), cx,
// cy, r, x1, e (ghost 0).
size_t mg_lds_first(const fs_handle* h);

int build_mg(fs_handle* h) {
  MgLevel L0;
  L0.d = h->pois(PO_DG);
  L0.cx = h->pois(PO_CX);
  L0.cy = h->pois(PO_CY);
  L0.r = h->pois(PO_R0);
  L0.x1 = h->pois(PO_MX1);
  L0.e = h->pois(PO_MU);
  L0.nx = h->bx;
  L0.ny = h->by;
  L0.P = h->P;
  L0.pad_ = 0;
  h->mg.assign(1, L0);
  std::vector<int> nxs, nys;
  int nx = h->bx, ny = h->by;
  size_t total = 0;
  HIP_OK(hipMalloc(&h->d_mginv, kMgBottomMax * kMgBottomMax * sizeof(double)));
  HIP_OK(hipMalloc(&h->d_mgpk, kMgPackFloats * sizeof(float)));
  // one dimension (ny = 1): the one-level hierarchy, a Jacobi-preconditioned
  // solve of the tridiagonal system
  if (h->by == 1) return 0;
  while ((size_t)nx * ny > (size_t)kMgBottomMax) {
    nx = (nx + 1) / 2;
    ny = (ny + 1) / 2;
    nxs.push_back(nx);
    nys.push_back(ny);
    total += 6 * (size_t)(nx + 2) * (ny + 2);
  }
  // very elongated grids would need more levels than the coarse-cycle kernel
  // takes: they keep the one-level hierarchy (Jacobi-preconditioned solve)
  if (total == 0 || nxs.size() + 1 > (size_t)kMgMaxLevels) return 0;
  HIP_OK(hipMalloc(&h->d_mg, total * sizeof(double)));
  HIP_OK(hipMemsetAsync(h->d_mg, 0, total * sizeof(double), h->st));
  double* p = h->d_mg;
  for (size_t l = 0; l < nxs.size(); l++) {
    MgLevel L;
    L.nx = nxs[l];
    L.ny = nys[l];
    L.P = L.nx + 2;
    L.pad_ = 0;
    size_t pl = (size_t)L.P * (L.ny + 2);
    L.d = p;
    L.cx = p + pl;
    L.cy = p + 2 * pl;
    L.r = p + 3 * pl;
    L.x1 = p + 4 * pl;
    L.e = p + 5 * pl;
    p += 6 * pl;
    launch_fill(L.d, 1.0, pl, h->st);
    h->mg.push_back(L);
  }
  h->mg_first = mg_lds_first(h);
  return 0;
}

MgArgs mg_args(fs_handle* h, size_t l) {
  MgArgs a;
  memset(&a, 0, sizeof(a));
  a.F = h->mg[l];
  if (l + 1 < h->mg.size()) a.C = h->mg[l + 1];
  a.sc = h->d_sc;
  a.G = h->G;
  a.oc = kMgOver;
  a.level0 = l == 0;
  return a;
}

MgPcgArgs mgpcg_args(fs_handle* h, int it) {
  MgPcgArgs a;
  memset(&a, 0, sizeof(a));
  a.cx = h->pois(PO_CX);
  a.cy = h->pois(PO_CY);
  a.dg = h->pois(PO_DG);
  a.u = h->pois(PO_MU);
  a.p_old = h->pois((it & 1) ? PO_MP1 : PO_MP0);
  a.p_new = h->pois((it & 1) ? PO_MP0 : PO_MP1);
  a.q = h->pois(PO_MQ);
  a.x = h->pois(PO_X);
  a.x2 = h->pois(PO_X2);
  a.x3 = h->pois(PO_X3);
  a.r = h->pois(PO_R0);
  a.lspart = h->d_lspart;
  a.sc = h->d_sc;
  a.G = h->G;
  a.nbx = h->nbx;
  a.nby = h->pnby;
  a.it = it;
  a.first = it == 0;
  return a;
}

// The first level run by the one-workgroup kernel: at most kMgLdsCells cells,
// the next at most 1024 (padded, at most kMgLdsR), at most kMgLdsLevels
// levels from there down, padded iterates within kMgLdsX doubles (those below
// it within kMgSetupX); never level 0 (its kernels carry the PCG reductions).
size_t mg_lds_first(const fs_handle* h) {
  const size_t L = h->mg.size();
  auto cells = [&](size_t q) { return (size_t)h->mg[q].nx * h->mg[q].ny; };
  for (size_t f = 1; f + 1 < L; f++) {
    size_t x = 0;
    for (size_t q = f; q < L; q++) x += (size_t)(h->mg[q].nx + 2) * (h->mg[q].ny + 2);
    const size_t padded = (size_t)(h->mg[f].nx + 2) * (h->mg[f].ny + 2);
    const size_t padded1 = (size_t)(h->mg[f + 1].nx + 2) * (h->mg[f + 1].ny + 2);
    if (cells(f) <= (size_t)kMgLdsCells && cells(f + 1) <= 1024 &&
        L - f <= (size_t)kMgLdsLevels && x <= (size_t)kMgLdsX && padded1 <= (size_t)kMgLdsR &&
        x - padded <= (size_t)kMgSetupX)
      return f;
  }
  return L - 1;
}

// Partial-sum buffers of the MG-PCG chain: 0 = |r/V|^2, 1 = r.u and u.w
// (flexible CG: r.z and z.q), 2 = p.q (flexible CG) or the learned input scale.
// The one-workgroup part of V-cycle `it` (-1: the initial cycle): levels
// mg_first .. L-1; when it starts at level 1 it also tests convergence.
MgCycleArgs cycle_args(fs_handle* h, int it) {
  const size_t L = h->mg.size(), f = h->mg_first;
  MgCycleArgs c;
  memset(&c, 0, sizeof(c));
  c.nq = (int)(L - f);
  int off = 0;
  for (int q = 0; q < c.nq; q++) {
    c.lv[q] = h->mg[f + q];
    c.off[q] = off;
    off += (c.lv[q].nx + 2) * (c.lv[q].ny + 2);
  }
  c.inv = h->d_mginv;
  c.pk = h->d_mgpk;
  c.sc = h->d_sc;
  c.oc = kMgOver;
  if (f == 1 && it >= 0) {
    c.check = 1;
    c.it = it;
    c.part_in = h->part(0);
    c.nb_in = mg_grid(h->mg[0]);
  }
  return c;
}

// Level-0 residual buffers alternate by iteration parity: iteration it reads
// r_buf(it) and writes r_buf(it + 1); the right-hand side starts in r_buf(0).
double* r_buf(fs_handle* h, int k) { return h->pois((k & 1) ? PO_R1 : PO_R0); }

MgCg cg_vectors(fs_handle* h, int it) {
  MgCg c;
  memset(&c, 0, sizeof(c));
  c.x = h->pois(PO_X);
  c.p = h->pois(PO_MP0);
  c.u = h->pois(PO_MU);
  c.w = h->pois(PO_W0);
  c.s_old = h->pois((it & 1) ? PO_S1 : PO_S0);
  c.s_new = h->pois((it & 1) ? PO_S0 : PO_S1);
  c.r_old = r_buf(h, it);
  c.on = 1;
  c.first = it == 0;
  return c;
}

// The down kernel of level l < mg_first in V-cycle `it` of the multigrid-PCG.
// it = -1, the initial cycle: level 0 tests the right-hand side (partials of
// the init kernel).  it >= 0: level 0 applies the CG update of iteration it
// (partials of the previous cycle's level-0 up kernel) and level 1 tests the
// updated residual.
MgArgs mg_down_args(fs_handle* h, size_t l, int it) {
  MgArgs a = mg_args(h, l);
  const int nb0 = mg_grid(h->mg[0]);
  if (l == 0 && it < 0) {
    a.check = 1;
    a.init = 1;
    a.part_in = h->part(0);
    a.nb_in = capped_grid(h->nbx * h->pnby);
    a.xs = h->pois(PO_X);
    a.xs2 = h->pois(PO_X2);
    a.xs3 = h->pois(PO_X3);
    a.rhs0 = h->pois(PO_RHS0);
  } else if (l == 0) {
    a.it = it;
    a.F.r = r_buf(h, it + 1);
    a.part_in = h->part(1);
    a.nb_in = nb0;
    a.part_out = h->part(0);
    a.cg = cg_vectors(h, it);
  } else if (l == 1 && it >= 0) {
    a.check = 1;
    a.it = it;
    a.part_in = h->part(0);
    a.nb_in = nb0;
  }
  return a;
}

// The up kernel of level l in V-cycle `it`: level 0 smooths the residual of
// that cycle, writes u, w = M u and the partials of r.u and u.w.
MgArgs mg_up_args(fs_handle* h, size_t l, int it) {
  MgArgs a = mg_args(h, l);
  if (l == 0) {
    a.F.r = r_buf(h, it + 1);
    a.part_out = h->part(1);
    a.cg = cg_vectors(h, it < 0 ? 0 : it);
  }
  return a;
}

// Multigrid levels the assembly coarsens itself (those of the tile levels
// whose aggregates lie inside one assembly tile); k_mg_coarsen makes the rest.
size_t lapl_coarse_levels(const fs_handle* h) {
  return h->mg.size() > 1 ? std::min<size_t>(kLaplCoarse, h->mg_first) : 0;
}

// The level whose equation the learned coarse solver takes: the first level
// below the finest with at most 64 x 64 cells.
size_t learned_level(const fs_handle* h) {
  size_t c = 1;
  while (c + 1 < h->mg.size() && (size_t)h->mg[c].nx * h->mg[c].ny > 4096) c++;
  return c;
}

// V-cycle `it` of the multigrid-PCG (-1: the initial cycle; it >= 0: one
// whole CG iteration, its update fused into the finest down kernel).  Only
// the parity of `it` matters after it = 0.  With `harvest`, the residual
// handed to the learned solver's level is appended to its training set (eager
// use only: the stream is synchronised).
int launch_vcycle(fs_handle* h, int it, fsl::Net* harvest = nullptr) {
  const size_t f = h->mg_first;
  const size_t hl = harvest ? learned_level(h) : 0;
  for (size_t l = 0; l < f; l++) {
    launch_mg_down(mg_down_args(h, l, it), h->st);
    if (harvest && l + 1 == hl) {
      if (sync_scalars(h)) return FS_E_HIP;
      if (!h->h_sc->done && fsl::net_store_sample(harvest, h->mg[hl], h->st) < 0) return FS_E_HIP;
    }
  }
  launch_mg_coarse_cycle(cycle_args(h, it), h->st);
  for (size_t l = f; l-- > 0;) launch_mg_up(mg_up_args(h, l, it), h->st);
  return 0;
}

// Learned preconditioner: the multigrid cycle down to the learned level c
// (red-black smoothing and aggregation on the finer levels; level c-1 also
// writes the partial sums of r_c^2 / d_c that scale the network input, into
// buffer 2, which is free at this point of the chain), the network as the
// solver of level c (or a zero correction without a network), and the cycle
// back up (no over-correction at level c-1; the finest level also writes the
// partials of r.z and z.q for the flexible beta).
int launch_learned_cycle(fs_handle* h, int it, bool init, int nb_rho) {
  const size_t c = learned_level(h);
  for (size_t l = 0; l < c; l++) {
    MgArgs a = mg_args(h, l);
    a.it = it;
    a.init = init;
    a.check = l == 0;
    a.part_in = h->part(0);
    a.nb_in = nb_rho;
    if (l + 1 == c) a.part_scale = h->part(2);
    launch_mg_down(a, h->st);
  }
  if (h->net) {
    if (fsl::net_apply(&h->net->net, h->mg[c], h->part(2), mg_grid(h->mg[c - 1]), h->d_sc, h->st))
      return FS_E_HIP;
  } else {
    const MgLevel& C = h->mg[c];
    launch_fill(C.e, 0.0, (size_t)C.P * (C.ny + 2), h->st);
  }
  for (size_t l = c; l-- > 0;) {
    MgArgs u = mg_args(h, l);
    u.part_out = h->part(1);
    if (l + 1 == c) u.oc = 1.0;
    if (l == 0) u.q = init ? nullptr : h->pois(PO_MQ);
    launch_mg_up(u, h->st);
  }
  return 0;
}

int launch_learned_iteration(fs_handle* h, int it) {
  const int nb_pcg = capped_grid(h->nbx * h->pnby);
  MgPcgArgs a = mgpcg_args(h, it);
  a.flex = 1;
  a.part_in = h->part(1);
  a.nb_in = mg_grid(h->mg[0]);
  a.part_out = h->part(2);
  launch_mgpcg_pq(a, h->st);
  a.part_in = h->part(2);
  a.nb_in = nb_pcg;
  a.part_out = h->part(0);
  launch_mgpcg_xr(a, h->st);
  return launch_learned_cycle(h, it, false, nb_pcg);
}

// Graphs are captured once per handle: every buffer is fixed for the
// handle's lifetime, and after the first iteration only the parity of the
// iteration number matters, so they serve all solves.
template <class F>
int capture(fs_handle* h, hipGraphExec_t* out, F&& launch) {
  hipGraph_t g = nullptr;
  HIP_OK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
  if (launch()) {
    (void)hipStreamEndCapture(h->st, &g);
    if (g) (void)hipGraphDestroy(g);
    return FS_E_HIP;
  }
  HIP_OK(hipStreamEndCapture(h->st, &g));
  HIP_OK(hipGraphInstantiate(out, g, nullptr, nullptr, 0));
  HIP_OK(hipGraphDestroy(g));
  return 0;
}

int capture_iteration(fs_handle* h, hipGraphExec_t* out, int (*launch)(fs_handle*, int), int it) {
  return capture(h, out, [&]() { return launch(h, it); });
}

// MG-PCG set-up: Galerkin hierarchy, coarsest inverse, the start (x = 0, or
// with `warm` the x in place), the initial V-cycle and iteration 0.
int mg_prologue(fs_handle* h, bool warm) {
  for (size_t l = lapl_coarse_levels(h); l < h->mg_first; l++)
    launch_mg_coarsen(mg_args(h, l), h->st);
  // (the one-workgroup set-up stays on the handle's stream: run beside the
  // start kernel on a second stream, the graph's parallel branch and its join
  // cost more than the 13 us it would hide, measured on the 512^2 pulse)
  launch_mg_coarse_setup(cycle_args(h, -1), h->st);
  MgPcgArgs ia = mgpcg_args(h, 1);
  ia.part_out = h->part(0);
  ia.warm = warm;
  if (warm && h->p.lapl_warm == 2) {
    ia.ls = 1;
    launch_mgpcg_warm_ls(ia, h->st);
  }
  launch_mgpcg_init(ia, h->st);
  return launch_vcycle(h, -1) || launch_vcycle(h, 0) ? FS_E_HIP : 0;
}

// The graph of 2^m MG-PCG iterations from an iteration of parity p.
int capture_chain(fs_handle* h, int p, int m) {
  const int base = p ? 1 : 2;
  return capture(h, &h->mg_chain[p][m], [&]() {
    for (int k = 0; k < (1 << m); k++)
      if (launch_vcycle(h, base + k)) return FS_E_HIP;
    return 0;
  });
}

// MG-PCG iterations it .. it+n-1 (it >= 1): with graphs, as chains of 2^m
// iterations (largest first), so a chunk costs a few graph launches.
int launch_mg_chunk(fs_handle* h, int it, int n) {
  while (n > 0) {
    if (!h->p.use_graphs) {
      if (launch_vcycle(h, it)) return FS_E_HIP;
      it++, n--;
      continue;
    }
    int m = kChainLog - 1;
    while ((1 << m) > n) m--;
    hipGraphExec_t& g = h->mg_chain[it & 1][m];
    if (!g && capture_chain(h, it & 1, m)) return FS_E_HIP;
    HIP_OK(hipGraphLaunch(g, h->st));
    it += 1 << m;
    n -= 1 << m;
  }
  return 0;
}

// The handle's graphs: the multigrid-PCG ones and, with `learned`, the
// learned-FCG iterations.
void drop_graphs(fs_handle* h, bool learned = true) {
  auto drop = [](hipGraphExec_t& g) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  };
  h->mg_graphs = false;
  drop(h->mg_pro[0]);
  drop(h->mg_pro[1]);
  for (int k = 0; k < 2; k++) {
    for (int c = 0; c < kSolveGraphs; c++) drop(h->mg_solve[k][c]);
    if (learned) drop(h->learn_graph[k]);
    for (int m = 0; m < kChainLog; m++) drop(h->mg_chain[k][m]);
  }
}

int make_learn_graph(fs_handle* h) {
  for (int k = 0; k < 2; k++) {
    if (h->learn_graph[k]) HIP_OK(hipGraphExecDestroy(h->learn_graph[k]));
    h->learn_graph[k] = nullptr;
  }
  if (capture_iteration(h, &h->learn_graph[1], launch_learned_iteration, 1) ||
      capture_iteration(h, &h->learn_graph[0], launch_learned_iteration, 2))
    return FS_E_HIP;
  h->learn_graph_net = h->net ? (const void*)h->net : (const void*)h;
  h->learn_graph_acts = h->net ? h->net->net.infer.id : 0;
  return 0;
}

LaplArgs lapl_args(fs_handle* h) {
  LaplArgs a;
  memset(&a, 0, sizeof(a));
  int l1 = h->alias ? h->bufA() : h->bufL1();
  a.n[0] = h->lev(l1, 0);
  a.n[1] = h->lev(l1, 1);
  a.n[2] = h->lev(l1, 2);
  a.ph = h->d_ph;
  a.rhs0 = h->pois(PO_RHS0);
  a.r = h->pois(PO_R0);
  a.cx = h->pois(PO_CX);
  a.cy = h->pois(PO_CY);
  a.dg = h->pois(PO_DG);
  a.part = h->part(2);
  a.sc = h->d_sc;
  a.G = h->G;
  a.nbx = h->nbx;
  a.nby = h->nby;
  if (h->d_dbg) {
    a.bdbg = h->d_dbg;
    a.adbg = h->d_dbg + (size_t)h->bx * h->by;
  }
  a.ncl = (int32_t)lapl_coarse_levels(h);
  for (int l = 0; l < a.ncl; l++) a.cl[l] = h->mg[l + 1];
  return a;
}

// The argument tables of the handle's stage and assembly kernels (rebuilt when
// the scalars move, see the ensemble).
int build_arg_tables(fs_handle* h) {
  std::vector<StageArgs> st(6 * 15);
  std::vector<LaplArgs> la(12);  // [gate * 6 + rot * 2 + alias]
  const int rot0 = h->rot;
  const bool alias0 = h->alias;
  for (int rot = 0; rot < 3; rot++)
    for (int al = 0; al < 2; al++) {
      h->rot = rot;
      h->alias = al != 0;
      for (int v = 0; v < 3; v++)
        for (int stg = 1; stg <= 5; stg++) {
          StageArgs a = stage_args(h, stg, v == 0 ? 0 : 1, v == 0);
          if (v == 2)
            for (int s = 0; s < 4; s++) a.kdbg[s] = nullptr;
          st[((rot * 2 + al) * 3 + v) * 5 + stg - 1] = a;
        }
      la[rot * 2 + al] = lapl_args(h);
      la[6 + rot * 2 + al] = la[rot * 2 + al];
      la[6 + rot * 2 + al].gate = 1;
    }
  h->rot = rot0;
  h->alias = alias0;
  if (!h->d_stab) HIP_OK(hipMalloc(&h->d_stab, st.size() * sizeof(StageArgs)));
  if (!h->d_ltab) HIP_OK(hipMalloc(&h->d_ltab, la.size() * sizeof(LaplArgs)));
  if (!h->d_phys) HIP_OK(hipMalloc(&h->d_phys, sizeof(FsPhys)));
  HIP_OK(hipMemcpy(h->d_stab, st.data(), st.size() * sizeof(StageArgs), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(h->d_ltab, la.data(), la.size() * sizeof(LaplArgs), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(h->d_phys, &h->K, sizeof(FsPhys), hipMemcpyHostToDevice));
  return 0;
}

const StageArgs* stab_at(const fs_handle* h, int rot, bool alias, int stage, int variant) {
  return h->d_stab + ((rot * 2 + (alias ? 1 : 0)) * 3 + variant) * 5 + stage - 1;
}
const StageArgs* stab(const fs_handle* h, int stage, int variant) {
  return stab_at(h, h->rot, h->alias, stage, variant);
}
const LaplArgs* ltab(const fs_handle* h, bool gated = false) {
  return h->d_ltab + (gated ? 6 : 0) + h->rot * 2 + (h->alias ? 1 : 0);
}

void launch_assembly(fs_handle* h, bool gated = false) {
  launch_lapl_assemble(ltab(h, gated), h->d_phys, h->nbx * h->nby, h->st);
}

// A pass's record of a ptcl_step attempt (the Poisson fields follow later).
void log_attempt(fs_handle* h) {
  DevScalars& s = *h->h_sc;
  double rec[FS_REC_LEN] = {s.ts,      s.t,     s.dt_stage, s.err_cur, s.dt,
                            (double)s.accept, -1.0, 0.0, 0.0, s.t_m};
  h->log.insert(h->log.end(), rec, rec + FS_REC_LEN);
}

// The Poisson fields of the pending record, on the last ptcl_step attempt.
void log_solve(fs_handle* h, int lerr) {
  size_t n = h->log.size() / FS_REC_LEN;
  if (n > 0 && h->log[(n - 1) * FS_REC_LEN + 6] < 0) {
    double* rec = &h->log[(n - 1) * FS_REC_LEN];
    rec[6] = lerr;
    rec[7] = h->h_sc->iters;
    rec[8] = h->h_sc->tres;
  }
}

// Iterations launched before the first convergence check: as many as the
// previous solve took, so that a solve usually needs one host round trip and
// launches no iteration past its last (an iteration after convergence costs
// its kernels' launch latency).  Without a previous solve: `dflt`.
int first_chunk(int last, int dflt) { return last > 0 ? last : dflt; }

// Every multigrid-PCG graph of the handle (set-up alone, set-up with the first
// chunk and the finishing kernels, iteration chains; from zero and warm),
// captured together at the handle's first graphed multigrid solve, so that no
// later solve waits for a capture when its iteration count first takes a new
// value.
int ensure_mg_graphs(fs_handle* h) {
  if (h->mg_graphs) return 0;
  drop_graphs(h, false);
  PcgArgs pa = pcg_args(h, 0);
  pa.part_out = h->part(2);
  pa.part_in = h->part(2);
  for (int w = 0; w < 2; w++) {
    const bool warm = w == 1;
    if (capture(h, &h->mg_pro[w], [&]() { return mg_prologue(h, warm); })) return FS_E_HIP;
    for (int n = 1; n <= kSolveGraphs; n++)
      if (capture(h, &h->mg_solve[w][n - 1], [&]() {
            if (mg_prologue(h, warm)) return FS_E_HIP;
            for (int k = 0; k < n; k++)
              if (launch_vcycle(h, 1 + k)) return FS_E_HIP;
            launch_pcg_finish(pa, h->st);
            return 0;
          }))
        return FS_E_HIP;
  }
  for (int p = 0; p < 2; p++)
    for (int m = 0; m < kChainLog; m++)
      if (capture_chain(h, p, m)) return FS_E_HIP;
  h->mg_graphs = true;
  return 0;
}

// Multigrid-PCG with graphs can be issued ahead of the host: after the
// controller's kernel (mode 1) the assembly, the set-up and the first chunk
// with its finishing kernels are enqueued, the assembly skipping the whole
// solve on the device unless the attempt is accepted; after the host has read
// the scalars, mode 2 continues that solve.  Mode 0: a whole solve.
bool can_issue_ahead(const fs_handle* h) {
  return h->p.use_graphs && h->p.lapl_solver == 1 && h->mg.size() > 1;
}

int lapl_solve_impl(fs_handle* h, int* lerr, fs_solve_stats* stats, int mode = 0) {
  auto t0 = std::chrono::steady_clock::now();
  if (mode != 2) {
    time_mark(h, 4);
    launch_assembly(h, mode == 1);
    HIP_OK(hipGetLastError());
    time_mark(h, 5);
  }
  const int chunk = h->p.lapl_chunk > 0 ? h->p.lapl_chunk : 32;
  int it = 0;
  // After each chunk of iterations the finishing kernels run, then the host
  // reads the scalars.  They change the potential only once the solve has
  // stopped converged, so launching them before knowing saves a host round
  // trip after the last chunk.
  PcgArgs pa = pcg_args(h, 0);
  pa.part_out = h->part(2);
  pa.part_in = h->part(2);
  auto end_chunk = [&](bool read = true) -> int {
    HIP_OK(hipGetLastError());
    launch_pcg_finish(pa, h->st);
    time_mark(h, 6);
    return read && sync_scalars(h) ? FS_E_HIP : 0;
  };
  if (h->p.lapl_solver == 2) {
    // flexible CG with the learned coarse solver (eager launches); without an
    // attached network the learned level gets a zero correction
    if (h->mg.size() < 2)
      return fail(FS_E_ARG,
                  "learned solver: the grid has no coarse level (<= 64 cells, or 1-D)%s");
    for (size_t l = lapl_coarse_levels(h); l < learned_level(h); l++)
      launch_mg_coarsen(mg_args(h, l), h->st);
    MgPcgArgs ia = mgpcg_args(h, 1);
    ia.part_out = h->part(0);
    launch_mgpcg_init(ia, h->st);
    const int nb_pcg = capped_grid(h->nbx * h->pnby);
    if (launch_learned_cycle(h, 0, true, nb_pcg) || launch_learned_iteration(h, 0))
      return FS_E_HIP;
    it = 1;
    const void* key = h->net ? (const void*)h->net : (const void*)h;
    const uint64_t acts = h->net ? h->net->net.infer.id : 0;
    if (h->p.use_graphs &&
        (!h->learn_graph[0] || h->learn_graph_net != key || h->learn_graph_acts != acts) &&
        make_learn_graph(h))
      return FS_E_HIP;
    // iteration 0 is launched already; then 2 more, then chunk / 4 at a time
    int lchunk = std::max(1, first_chunk(h->last_iters[2], chunk > 8 ? chunk / 4 : chunk) - 1);
    for (int round = 0;; round++) {
      for (int k = 0; k < lchunk; k++, it++) {
        if (h->p.use_graphs) HIP_OK(hipGraphLaunch(h->learn_graph[it & 1], h->st));
        else if (launch_learned_iteration(h, it)) return FS_E_HIP;
      }
      lchunk = round == 0 ? 2 : chunk > 8 ? chunk / 4 : chunk;
      if (int e = end_chunk()) return e;
      if (h->h_sc->status < 0) return check_status(h);
      if (h->h_sc->done) break;
    }
  } else if (h->p.lapl_solver == 1 && h->mg.size() > 1) {
    // multigrid-preconditioned PCG (grids of <= 64 cells use Jacobi-PCG)
    int& mg_it = h->mg_it;
    int& mg_chunk = h->mg_chunk;
    if (mode != 2) {
      // the last correction is a useful start only for the next pass's solve:
      // densities have moved since (not for a repeated or a Poisson-only solve)
      const bool warm = h->p.lapl_warm && h->warm_ok && h->moved;
      h->solve_warm = warm;
      // iteration 0 comes with the set-up; then as many as the last solve
      // took, then 2 more, then chunk / 4 at a time.  A short first chunk is
      // one graph with the set-up and the finishing kernels (each graph
      // boundary costs microseconds of idle device).
      mg_it = 1;
      mg_chunk = std::max(1, first_chunk(h->last_iters[1], chunk > 8 ? chunk / 4 : chunk) - 1);
      if (h->p.use_graphs && ensure_mg_graphs(h)) return FS_E_HIP;
      if (h->p.use_graphs && mg_chunk <= kSolveGraphs) {
        hipGraphExec_t& g = h->mg_solve[warm ? 1 : 0][mg_chunk - 1];
        HIP_OK(hipGraphLaunch(g, h->st));
        time_mark(h, 6);
        if (mode == 0 && sync_scalars(h)) return FS_E_HIP;
      } else {
        if (!h->p.use_graphs) {
          if (mg_prologue(h, warm)) return FS_E_HIP;
        } else {
          HIP_OK(hipGraphLaunch(h->mg_pro[warm ? 1 : 0], h->st));
        }
        if (launch_mg_chunk(h, mg_it, mg_chunk)) return FS_E_HIP;
        if (int e = end_chunk(mode == 0)) return e;
      }
      mg_it += mg_chunk;
      mg_chunk = 2;
      if (mode == 1) return 0;
    }
    for (;;) {
      if (h->h_sc->status < 0) return check_status(h);
      if (h->h_sc->done) break;
      if (launch_mg_chunk(h, mg_it, mg_chunk)) return FS_E_HIP;
      mg_it += mg_chunk;
      mg_chunk = chunk > 8 ? chunk / 4 : chunk;
      if (int e = end_chunk()) return e;
    }
  } else {
    PcgArgs ia = pcg_args(h, 0);
    ia.part_out = h->part(0);  // what iteration 0 reads
    ia.w_new = h->pois(PO_W0);  // w_0 is iteration 0's w_old
    launch_pcg_init(ia, h->st);
    for (;;) {
      const int n = it == 0 ? first_chunk(h->last_iters[0], chunk) : chunk;
      for (int k = 0; k < n; k++, it++) launch_pcg_it(pcg_args(h, it), h->st);
      if (int e = end_chunk()) return e;
      if (h->h_sc->status < 0) return check_status(h);
      if (h->h_sc->done) break;
    }
  }
  if (int e = check_status(h)) return e;
  int ok = h->h_sc->done == 1;
  if (!ok && h->p.lapl_solver == 1 && h->mg.size() > 1) {
    if (h->solve_warm) {
      // a warm start that did not converge (the potential is unchanged): the
      // same system again from zero
      h->solve_warm = false;
      h->warm_ok = false;
      return lapl_solve_impl(h, lerr, stats, 0);
    }
    // From zero too: the same system by Jacobi-PCG before the step is
    // rejected.  dFdPh's cancellation at a near-zero field can turn a face
    // coupling negative and the operator indefinite; the multigrid-PCG
    // iteration steps along negative curvature, but a smoother that amplifies
    // the negative mode strongly can still keep it from converging.
    h->p.lapl_solver = 0;
    const int e = lapl_solve_impl(h, lerr, stats, 0);
    h->p.lapl_solver = 1;
    return e;
  }
  *lerr = ok ? 0 : 1;
  const int sv = h->p.lapl_solver == 2 ? 2 : (h->p.lapl_solver == 1 && h->mg.size() > 1) ? 1 : 0;
  h->last_iters[sv] = h->h_sc->iters;
  h->warm_ok = sv == 1 && ok;
  h->moved = false;
  if (h->timing) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, h->ev[4], h->ev[5]);
    h->tms[2] += ms;
    (void)hipEventElapsedTime(&ms, h->ev[5], h->ev[6]);
    h->tms[3] += ms;
  }
  if (stats) {
    stats->iters = h->h_sc->iters;
    stats->lerr = *lerr;
    stats->resid = h->h_sc->tres;
    stats->bnorm = sqrt(h->h_sc->bnorm2);
    stats->seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  // log the solve on the last ptcl_step attempt (as the oracle does)
  log_solve(h, *lerr);
  return 0;
}

CtrlArgs ctrl_args(fs_handle* h) {
  CtrlArgs ca;
  memset(&ca, 0, sizeof(ca));
  ca.part = h->part(0);
  ca.nblocks = h->nbx * ((h->by + kStageRows - 1) / kStageRows);  // the stage kernel's row groups
  ca.sc = h->d_sc;
  ca.replay = h->d_replay;
  return ca;
}

// ahead: after each attempt's controller kernel, issue the Poisson solve that
// follows an accepted attempt (lapl_solve_impl mode 1) before reading the
// controller's decision.  spec_next (another pass follows in the same run):
// behind the solve's first chunk, also issue the next pass's accept and begin
// (k_accept_spec, which applies them only if the pass is accepted) and its
// first stage, gated on k_accept_spec's decision; the host reads the same
// scalars and knows whether they ran.  The device does not then wait for the
// host between passes that converge in their first chunk.
int ptcl_step_impl(fs_handle* h, fs_step_info* info, bool ahead = false,
                   bool spec_next = false) {
  const bool pre = h->spec_taken;  // this pass's begin and stage 1 are issued already
  h->spec_taken = false;
  spec_next = spec_next && ahead && h->h_pub && !h->timing;
  if (!pre) launch_begin_step(h->d_sc, h->st);
  int attempts = 0;
  for (;;) {
    time_mark(h, 0);
    for (int stage = 1; stage <= 5; stage++)  // (stages 2-5 also do circuit updates 1-4)
      if (!(pre && attempts == 0 && stage == 1))
        launch_stage(stab(h, stage, 0), h->d_phys, h->nbx, h->by, h->sty, h->st);
    time_mark(h, 1);
    launch_circ_ctrl(circ_args(h, 5), h->K, ctrl_args(h), h->st);
    HIP_OK(hipGetLastError());
    time_mark(h, 2);
    if (ahead) {
      int lerr_ahead = 0;
      h->alias = false;  // the solve reads level 1 (as below, after the read)
      h->moved = true;   // it runs only if this attempt is accepted
      if (int e = lapl_solve_impl(h, &lerr_ahead, nullptr, 1)) return e;
    }
    if (spec_next) {
      enqueue_publish(h);
      launch_accept_spec(h->d_sc, h->st);
      launch_stage(stab_at(h, (h->rot + 1) % 3, true, 1, 0), h->d_phys, h->nbx, h->by, h->sty, h->st,
                   1);
      HIP_OK(hipGetLastError());
      if (wait_published(h)) return FS_E_HIP;
      const DevScalars& s = *h->h_sc;  // what k_accept_spec reads
      h->spec_taken = s.status >= 0 && s.accept == 1 && s.done == 1;
    } else if (sync_scalars(h)) {
      return FS_E_HIP;
    }
    if (h->timing) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
      h->tms[0] += ms;
      (void)hipEventElapsedTime(&ms, h->ev[1], h->ev[2]);
      h->tms[1] += ms;
    }
    attempts++;
    DevScalars& s = *h->h_sc;
    log_attempt(h);
    if (int e = check_status(h)) return e;
    h->alias = false;
    if (s.accept) {
      h->moved = true;
      break;
    }
    // internal restart, ptcl_lib.F90:76-93: n1 = n2 = n3 (level 1 is rewritten at stage 5)
    if (copy_level(h, h->bufA(), h->bufL3())) return FS_E_HIP;
  }
  if (info) {
    DevScalars& s = *h->h_sc;
    info->err_cur = s.err_cur;
    info->dt_stage = s.dt_stage;
    info->dt_next = s.dt;
    info->t_m = s.t_m;
    info->Vd_pl = s.Vd_pl;
    info->Ida = s.Ida;
    info->Idc = s.Idc;
    info->Vsrc = s.Vsrc;
    info->attempts = attempts;
    info->status = s.status;
  }
  return 0;
}

int accept_impl(fs_handle* h) {
  if (!h->spec_taken) launch_accept(h->d_sc, h->st);  // else k_accept_spec did it
  HIP_OK(hipGetLastError());
  h->rot = (h->rot + 1) % 3;
  h->alias = true;
  return 0;
}

int reject_impl(fs_handle* h) {
  launch_reject(h->d_sc, h->st);
  HIP_OK(hipGetLastError());
  if (copy_level(h, h->bufA(), h->bufL3())) return FS_E_HIP;
  h->alias = true;
  return 0;
}


// ---------------------------------------------------------------------------
// Ensembles: independent discharges advanced in lockstep rounds by launches
// that cover every member (EnsMap).  A round gives each member one ptcl_step
// attempt (five stages, circuit, controller) or, once its attempt is accepted
// (or before the first transport step), its Poisson solve and the accept /
// reject of the pass -- exactly the sequence fs_run issues for the member
// alone, so each member's fields are bitwise those of its solo run.
struct EnsMapH {
  int32_t* d = nullptr;  // inst[nblocks], first[n], nb[n]
  EnsMap m{};
};

struct EnsRound {  // uploaded once per round
  int32_t* sel;
  int32_t* sel2;  // the MG-PCG start of each member: i * 2 + warm
  uint8_t *ptcl, *pois, *begin, *acc, *rej;
};

// ---------------------------------------------------------------------------
// The stage kernel's tile height (fs_params.stage_tile).  With 0 the handle
// times the three heights on its own grid and keeps the fastest; the choice is
// cached per grid shape for the process.  The stage kernel's results do not
// depend on the height (its error sums are reduced per row segment of kTX
// cells, the same tree for every height), only its speed does.
constexpr int kStageTiles[3] = {4, 8, 16};
std::mutex g_tile_mu;
std::map<std::tuple<int, int, int, int>, int> g_tile_cache;  // (nx, ny, cyl, rwall) -> height
// set while an ensemble creates its members: the ensemble times its own
// batched launch instead
thread_local bool t_tile_deferred = false;

bool valid_stage_tile(int ty) { return ty == 4 || ty == 8 || ty == 16; }

// Average time of `reps` launches made by `launch` on stream st (after two
// untimed ones).
template <class F>
int time_launches(hipStream_t st, int reps, F&& launch, float* ms) {
  hipEvent_t a = nullptr, b = nullptr;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  launch();
  launch();
  (void)hipEventRecord(a, st);
  for (int k = 0; k < reps; k++) launch();
  (void)hipEventRecord(b, st);
  const hipError_t e = hipEventSynchronize(b);
  float t = 0.0f;
  if (e == hipSuccess) (void)hipEventElapsedTime(&t, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  HIP_OK(e);
  HIP_OK(hipGetLastError());
  *ms = t / reps;
  return 0;
}

int choose_stage_tile(fs_handle* h) {
  const int req = h->p.stage_tile;
  if (req != 0 && !valid_stage_tile(req)) return fail(FS_E_ARG, "stage_tile must be 0, 4, 8 or 16%s");
  if (req != 0 || t_tile_deferred) {
    h->sty = req != 0 ? req : 8;
    return 0;
  }
  const auto key = std::make_tuple(h->bx, h->by, h->g.cyl, h->g.rwall);
  {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    auto it = g_tile_cache.find(key);
    if (it != g_tile_cache.end()) {
      h->sty = it->second;
      return 0;
    }
  }
  // right-hand-side launches of stage 1 on the initial state: they write
  // only the electrode-face currents, cleared afterwards
  float best = 0.0f;
  int bt = 8;
  for (int ty : kStageTiles) {
    float ms = 0.0f;
    if (time_launches(h->st, 10, [&]() {
          launch_stage(stab(h, 1, 2), h->d_phys, h->nbx, h->by, ty, h->st);
        }, &ms))
      return FS_E_HIP;
    if (ty == kStageTiles[0] || ms < best) {
      best = ms;
      bt = ty;
    }
  }
  HIP_OK(hipMemsetAsync(h->d_cur, 0, 4 * (size_t)h->by * sizeof(double), h->st));
  HIP_OK(hipStreamSynchronize(h->st));
  h->sty = bt;
  std::lock_guard<std::mutex> lk(g_tile_mu);
  g_tile_cache[key] = bt;
  return 0;
}

}  // namespace

struct fs_ensemble {
  std::vector<fs_handle*> m;
  std::vector<DevScalars*> own_d, own_h;  // members' own scalar buffers
  std::vector<hipStream_t> own_st;
  DevScalars* d_sc = nullptr;
  DevScalars* h_sc = nullptr;
  hipStream_t st = nullptr;
  int n = 0, maxL = 0, maxF = 0;
  bool projected = false;  // some member takes the projected warm start (lapl_warm = 2)
  // argument tables
  FsPhys* t_phys = nullptr;
  StageArgs* t_stage = nullptr;   // [(i*6 + rot*2 + alias)*5 + stage-1]
  CircArgs* t_circ = nullptr;     // [i*5 + stage-1]
  CtrlArgs* t_ctrl = nullptr;     // [i]
  LaplArgs* t_lapl = nullptr;     // [i*6 + rot*2 + alias]
  MgArgs* t_coarsen = nullptr;    // [i*kMgMaxLevels + l]
  MgPcgArgs* t_init = nullptr;    // [i * 2 + warm]
  // V-cycle variants v: 0 the initial cycle, 1 odd, 2 even, 3 first iteration
  MgArgs* t_down = nullptr;       // [(i*kMgMaxLevels + l)*4 + v]
  MgArgs* t_up = nullptr;         // [(i*kMgMaxLevels + l)*4 + v]
  MgCycleArgs* t_cc = nullptr;    // [i*4 + v]
  PcgArgs* t_fin = nullptr;       // [i]
  EnsMapH m_stage, m_one, m_pcg, m_coarsen[kMgMaxLevels], m_down[kMgMaxLevels], m_up[kMgMaxLevels];
  EnsMapH m_st;                   // the stage kernel's tiles (kTX x sty) of every member
  int sty = 8;                    // the stage kernel's tile height
  char* d_round = nullptr;
  char* h_round = nullptr;
  size_t round_bytes = 0;
  EnsRound r{};
  std::vector<char> phase;        // 0 start of a pass, 1 ptcl attempt, 2 Poisson
  // A round's fixed launch sequences as HIP graphs (when every member has
  // use_graphs), captured on first use in a run: each run uploads its
  // argument tables anew, so each run drops them.  Which members a launch
  // covers is read from the round buffer on the device.
  bool graphs = false;
  hipGraphExec_t g_ptcl = nullptr;   // begin, five stages with their circuit updates, controller
  hipGraphExec_t g_solve = nullptr;  // assembly, coarsening, set-up, start, cycles to iteration 8
  hipGraphExec_t g_iter = nullptr;   // eight iterations (from an odd one)
  hipGraphExec_t g_fin = nullptr;    // the finishing kernels
  hipGraphExec_t g_end = nullptr;    // accept / reject of the members' passes
};

namespace {

bool bad(hipError_t e) { return e != hipSuccess; }

template <class T>
int upload_table(T** d, const std::vector<T>& h) {
  if (*d) HIP_OK(hipFree(*d));
  *d = nullptr;
  if (h.empty()) return 0;
  HIP_OK(hipMalloc(d, h.size() * sizeof(T)));
  HIP_OK(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Block map from per-member block counts (0 = not in this launch).
int make_map(EnsMapH& mh, const std::vector<int>& nb) {
  const int n = (int)nb.size();
  std::vector<int32_t> first(n), inst;
  for (int i = 0; i < n; i++) {
    first[i] = (int32_t)inst.size();
    for (int b = 0; b < nb[i]; b++) inst.push_back(i);
  }
  std::vector<int32_t> all(inst);
  all.insert(all.end(), first.begin(), first.end());
  all.insert(all.end(), nb.begin(), nb.end());
  if (mh.d) HIP_OK(hipFree(mh.d));
  HIP_OK(hipMalloc(&mh.d, all.size() * sizeof(int32_t)));
  HIP_OK(hipMemcpy(mh.d, all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  mh.m.inst = mh.d;
  mh.m.first = mh.d + inst.size();
  mh.m.nb = mh.d + inst.size() + n;
  mh.m.nblocks = (int32_t)inst.size();
  return 0;
}

EnsMap with_mask(const EnsMapH& mh, const uint8_t* active) {
  EnsMap m = mh.m;
  m.active = active;
  return m;
}

// Argument tables (they hold device pointers only; rebuilt when a run starts,
// since a member's replay table may have changed).
int ens_tables(fs_ensemble* E) {
  const int n = E->n;
  std::vector<FsPhys> phys(n);
  std::vector<StageArgs> st(n * 30);
  std::vector<CircArgs> ci(n * 5);
  std::vector<CtrlArgs> ct(n);
  std::vector<LaplArgs> la(n * 6);
  std::vector<MgArgs> co(n * kMgMaxLevels), dn(n * kMgMaxLevels * 4), up(n * kMgMaxLevels * 4);
  std::vector<MgPcgArgs> in(2 * n);
  std::vector<MgCycleArgs> cc(n * 4);
  std::vector<PcgArgs> fi(n);
  for (int i = 0; i < n; i++) {
    fs_handle* h = E->m[i];
    phys[i] = h->K;
    const int rot0 = h->rot;
    const bool alias0 = h->alias;
    for (int rot = 0; rot < 3; rot++)
      for (int al = 0; al < 2; al++) {
        h->rot = rot;
        h->alias = al != 0;
        for (int stg = 1; stg <= 5; stg++) st[(i * 6 + rot * 2 + al) * 5 + stg - 1] = stage_args(h, stg, 0);
        la[i * 6 + rot * 2 + al] = lapl_args(h);
      }
    h->rot = rot0;
    h->alias = alias0;
    for (int stg = 1; stg <= 5; stg++) ci[i * 5 + stg - 1] = circ_args(h, stg);
    ct[i] = ctrl_args(h);
    const int L = (int)h->mg.size();
    for (int l = 0; l + 1 < L; l++) co[i * kMgMaxLevels + l] = mg_args(h, l);
    for (int w = 0; w < 2; w++) {
      in[i * 2 + w] = mgpcg_args(h, 1);
      in[i * 2 + w].part_out = h->part(0);
      in[i * 2 + w].warm = w;
      in[i * 2 + w].ls = w && h->p.lapl_warm == 2;
    }
    static const int kVariantIt[4] = {-1, 1, 2, 0};
    for (int v = 0; v < 4; v++) {
      for (int l = 0; l < (int)h->mg_first; l++) {
        dn[(i * kMgMaxLevels + l) * 4 + v] = mg_down_args(h, l, kVariantIt[v]);
        up[(i * kMgMaxLevels + l) * 4 + v] = mg_up_args(h, l, kVariantIt[v]);
      }
      cc[i * 4 + v] = cycle_args(h, kVariantIt[v]);
    }
    PcgArgs f = pcg_args(h, 0);
    f.part_out = h->part(2);
    f.part_in = h->part(2);
    fi[i] = f;
  }
  if (upload_table(&E->t_phys, phys) || upload_table(&E->t_stage, st) ||
      upload_table(&E->t_circ, ci) || upload_table(&E->t_ctrl, ct) ||
      upload_table(&E->t_lapl, la) || upload_table(&E->t_coarsen, co) ||
      upload_table(&E->t_init, in) || upload_table(&E->t_down, dn) ||
      upload_table(&E->t_up, up) || upload_table(&E->t_cc, cc) || upload_table(&E->t_fin, fi))
    return FS_E_HIP;
  return 0;
}

int ens_maps(fs_ensemble* E) {
  const int n = E->n;
  std::vector<int> stage(n), one(n, 1), pcg(n);
  E->maxL = E->maxF = 0;
  for (int i = 0; i < n; i++) {
    fs_handle* h = E->m[i];
    stage[i] = h->nbx * h->nby;
    pcg[i] = capped_grid(h->nbx * h->pnby);
    E->maxL = std::max(E->maxL, (int)h->mg.size());
    E->maxF = std::max(E->maxF, (int)h->mg_first);
  }
  std::vector<int> st(n);
  for (int i = 0; i < n; i++) st[i] = stage_blocks(E->m[i]->nbx, E->m[i]->by, E->sty);
  if (make_map(E->m_stage, stage) || make_map(E->m_one, one) || make_map(E->m_pcg, pcg) ||
      make_map(E->m_st, st))
    return FS_E_HIP;
  for (int l = 0; l < kMgMaxLevels; l++) {
    std::vector<int> co(n, 0), dn(n, 0);
    for (int i = 0; i < n; i++) {
      fs_handle* h = E->m[i];
      if (l < (int)h->mg_first) {  // the set-up kernel coarsens the levels below
        const MgLevel& C = h->mg[l + 1];
        co[i] = (C.nx * C.ny + 255) / 256;
      }
      if (l < (int)h->mg_first) dn[i] = mg_grid(h->mg[l]);
    }
    if (make_map(E->m_coarsen[l], co) || make_map(E->m_down[l], dn) || make_map(E->m_up[l], dn))
      return FS_E_HIP;
  }
  return 0;
}

// The ensemble's stage tile height: the one requested, or the fastest of the
// three for the batched launch over every member (right-hand-side launches of
// stage 1 on the members' initial states; they write only the electrode-face
// currents, cleared afterwards).
int ens_choose_stage_tile(fs_ensemble* E, int req) {
  const int n = E->n;
  if (req != 0) {
    E->sty = req;
    return ens_maps(E);
  }
  std::vector<StageArgs> st(n);
  std::vector<FsPhys> phys(n);
  for (int i = 0; i < n; i++) {
    st[i] = stage_args(E->m[i], 1, 1);
    for (int s = 0; s < 4; s++) st[i].kdbg[s] = nullptr;
    phys[i] = E->m[i]->K;
  }
  StageArgs* d_st = nullptr;
  FsPhys* d_phys = nullptr;
  uint8_t* d_on = nullptr;
  int err = upload_table(&d_st, st) || upload_table(&d_phys, phys) ? FS_E_HIP : 0;
  if (!err && (hipMalloc(&d_on, n) != hipSuccess || hipMemset(d_on, 1, n) != hipSuccess))
    err = fail(FS_E_HIP, "ensemble tile choice: allocation%s");
  float best = 0.0f;
  int bt = 8;
  for (int ty : kStageTiles) {
    if (err) break;
    E->sty = ty;
    float ms = 0.0f;
    if ((err = ens_maps(E))) break;
    const EnsMap m = with_mask(E->m_st, d_on);
    if ((err = time_launches(E->st, 5, [&]() {
           launch_ens_stage(d_st, d_phys, m, EnsSel{nullptr, 1, 0}, ty, E->st);
         }, &ms)))
      break;
    if (ty == kStageTiles[0] || ms < best) {
      best = ms;
      bt = ty;
    }
  }
  if (d_st) (void)hipFree(d_st);
  if (d_phys) (void)hipFree(d_phys);
  if (d_on) (void)hipFree(d_on);
  if (err) return err;
  for (fs_handle* h : E->m)
    HIP_OK(hipMemsetAsync(h->d_cur, 0, 4 * (size_t)h->by * sizeof(double), E->st));
  HIP_OK(hipStreamSynchronize(E->st));
  E->sty = bt;
  return ens_maps(E);
}

int ens_sync(fs_ensemble* E) {
  HIP_OK(hipMemcpyAsync(E->h_sc, E->d_sc, E->n * sizeof(DevScalars), hipMemcpyDeviceToHost, E->st));
  HIP_OK(hipStreamSynchronize(E->st));
  return 0;
}

// The host side of the round buffer is rewritten after each upload, so the
// copy completes before this returns (the stream is idle at each call).
int ens_upload_round(fs_ensemble* E) {
  HIP_OK(hipMemcpyAsync(E->d_round, E->h_round, E->round_bytes, hipMemcpyHostToDevice, E->st));
  HIP_OK(hipStreamSynchronize(E->st));
  return 0;
}

int ens_fail(fs_ensemble* E, int i) {
  fs_handle* h = E->m[i];
  int e = check_status(h);
  if (e) {
    std::string msg = "ensemble member " + std::to_string(i) + ": " + g_err;
    g_err = msg;
  }
  return e;
}

// One V-cycle of every Poisson member: down over the tile levels, the
// one-workgroup coarse cycle, up.  v: 0 the initial cycle, then one CG
// iteration each: 1 odd, 2 even, 3 the first.
void ens_vcycle(fs_ensemble* E, int v, const uint8_t* mask) {
  for (int l = 0; l < E->maxF; l++)
    launch_ens_mg_down(E->t_down, with_mask(E->m_down[l], mask),
                       EnsSel{nullptr, kMgMaxLevels * 4, l * 4 + v}, l == 0 && v != 0, E->st);
  launch_ens_mg_coarse_cycle(E->t_cc, with_mask(E->m_one, mask), EnsSel{nullptr, 4, v}, E->st);
  for (int l = E->maxF; l-- > 0;)
    launch_ens_mg_up(E->t_up, with_mask(E->m_up[l], mask), EnsSel{nullptr, kMgMaxLevels * 4, l * 4 + v},
                     l == 0, E->st);
}

void ens_iteration(fs_ensemble* E, int it, const uint8_t* mask) {
  ens_vcycle(E, it == 0 ? 3 : (it & 1) ? 1 : 2, mask);
}

void ens_drop_graphs(fs_ensemble* E) {
  for (hipGraphExec_t* g : {&E->g_ptcl, &E->g_solve, &E->g_iter, &E->g_fin, &E->g_end}) {
    if (*g) (void)hipGraphExecDestroy(*g);
    *g = nullptr;
  }
}

// Runs `launch` (kernel launches on E->st only) eagerly, or as the graph *g,
// captured on its first use.
template <class F>
int ens_launch(fs_ensemble* E, hipGraphExec_t* g, F&& launch) {
  if (!E->graphs) {
    launch();
    HIP_OK(hipGetLastError());
    return 0;
  }
  if (!*g) {
    hipGraph_t gr = nullptr;
    HIP_OK(hipStreamBeginCapture(E->st, hipStreamCaptureModeThreadLocal));
    launch();
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(E->st, &gr);
    if (le != hipSuccess || ce != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      return fail(FS_E_HIP, "ensemble graph capture: %s", hipGetErrorString(le != hipSuccess ? le : ce));
    }
    const hipError_t ie = hipGraphInstantiate(g, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    HIP_OK(ie);
  }
  HIP_OK(hipGraphLaunch(*g, E->st));
  return 0;
}

int ens_round(fs_ensemble* E, std::vector<int>& left) {
  const int n = E->n;
  EnsRound& r = E->r;
  char* hr = E->h_round;
  int32_t* sel = (int32_t*)hr;
  uint8_t* ptcl = (uint8_t*)(hr + 8 * n);
  uint8_t *pois = ptcl + n, *begin = pois + n, *acc = begin + n, *rej = acc + n;
  memset(ptcl, 0, 5 * n);
  // passes that start now: begin_step, then a ptcl attempt (or, before the
  // first transport step, straight to Poisson: main.F90:75)
  int nbeg = 0;
  for (int i = 0; i < n; i++)
    if (left[i] > 0 && E->phase[i] == 0) {
      begin[i] = 1;
      nbeg++;
      E->phase[i] = E->h_sc[i].ts > 1.0 ? 1 : 2;
    }
  int nptcl = 0;
  for (int i = 0; i < n; i++) {
    ptcl[i] = left[i] > 0 && E->phase[i] == 1;
    nptcl += ptcl[i];
    sel[i] = i * 6 + E->m[i]->rot * 2 + (E->m[i]->alias ? 1 : 0);
  }
  if (ens_upload_round(E)) return FS_E_HIP;
  if (nptcl) {
    // (the begin kernel changes only the members marked in r.begin)
    if (ens_launch(E, &E->g_ptcl, [&]() {
          launch_ens_scalars(E->d_sc, r.begin, n, 0, E->st);
          for (int stg = 1; stg <= 5; stg++) {
            launch_ens_stage(E->t_stage, E->t_phys, with_mask(E->m_st, r.ptcl),
                             EnsSel{r.sel, 5, stg - 1}, E->sty, E->st);
            launch_ens_circ(E->t_circ, E->t_phys, with_mask(E->m_one, r.ptcl),
                            EnsSel{nullptr, 5, stg - 1}, E->st);
          }
          launch_ens_ctrl(E->t_ctrl, with_mask(E->m_one, r.ptcl), EnsSel{nullptr, 1, 0}, E->st);
        }))
      return FS_E_HIP;
    if (ens_sync(E)) return FS_E_HIP;
    for (int i = 0; i < n; i++) {
      if (!ptcl[i]) continue;
      fs_handle* h = E->m[i];
      log_attempt(h);
      if (int e = ens_fail(E, i)) return e;
      h->alias = false;
      if (E->h_sc[i].accept) {
        E->phase[i] = 2;
        h->moved = true;
      } else if (copy_level(h, h->bufA(), h->bufL3())) {  // ptcl_lib.F90:76-93
        return FS_E_HIP;
      }
    }
  }
  else if (nbeg) {
    launch_ens_scalars(E->d_sc, r.begin, n, 0, E->st);
  }
  int npois = 0;
  int32_t* sel2 = sel + n;
  for (int i = 0; i < n; i++) {
    pois[i] = left[i] > 0 && E->phase[i] == 2;
    npois += pois[i];
    sel[i] = i * 6 + E->m[i]->rot * 2 + (E->m[i]->alias ? 1 : 0);
    sel2[i] = i * 2 + (E->m[i]->p.lapl_warm && E->m[i]->warm_ok && E->m[i]->moved ? 1 : 0);
  }
  if (!npois) return 0;
  // The members' solves; then, as a member's own solve does, the warm solves
  // that did not converge again from zero (the pois mask then holds just
  // those members, and is restored for the finishing kernels).
  std::vector<uint8_t> all_pois(pois, pois + n);
  for (int round = 0; round < 2; round++) {
    if (ens_upload_round(E)) return FS_E_HIP;
    // the set-up and iterations 0..8, then eight at a time (from odd ones)
    if (ens_launch(E, &E->g_solve, [&]() {
          launch_ens_lapl(E->t_lapl, E->t_phys, with_mask(E->m_stage, r.pois), EnsSel{r.sel, 1, 0},
                          E->st);
          for (int l = 0; l < E->maxF; l++)
            launch_ens_mg_coarsen(E->t_coarsen, with_mask(E->m_coarsen[l], r.pois),
                                  EnsSel{nullptr, kMgMaxLevels, l}, E->st);
          launch_ens_mg_coarse_setup(E->t_cc, with_mask(E->m_one, r.pois), EnsSel{nullptr, 4, 0},
                                     E->st);
          if (E->projected)
            launch_ens_mgpcg_warm_ls(E->t_init, with_mask(E->m_pcg, r.pois),
                                     with_mask(E->m_one, r.pois), EnsSel{r.sel2, 1, 0}, E->st);
          launch_ens_mgpcg_init(E->t_init, with_mask(E->m_pcg, r.pois), EnsSel{r.sel2, 1, 0},
                                E->st);
          ens_vcycle(E, 0, r.pois);
          for (int it = 0; it <= 8; it++) ens_iteration(E, it, r.pois);
        }))
      return FS_E_HIP;
    for (bool first = true;; first = false) {
      if (!first &&
          ens_launch(E, &E->g_iter, [&]() {
            for (int k = 0; k < 8; k++) ens_iteration(E, 9 + k, r.pois);
          }))
        return FS_E_HIP;
      if (ens_sync(E)) return FS_E_HIP;
      bool all = true;
      for (int i = 0; i < n; i++) {
        if (!pois[i]) continue;
        if (E->h_sc[i].status < 0) return ens_fail(E, i);
        all = all && E->h_sc[i].done;
      }
      if (all) break;
    }
    if (round == 1) break;
    int nretry = 0;
    for (int i = 0; i < n; i++) {
      const bool warm = (sel2[i] & 1) != 0;
      pois[i] = pois[i] && warm && E->h_sc[i].done != 1;
      nretry += pois[i];
      sel2[i] = i * 2;  // from zero
    }
    if (!nretry) break;
  }
  memcpy(pois, all_pois.data(), n);
  if (ens_upload_round(E)) return FS_E_HIP;
  if (ens_launch(E, &E->g_fin, [&]() {
        launch_ens_pcg_finish(E->t_fin, with_mask(E->m_pcg, r.pois), with_mask(E->m_one, r.pois),
                              EnsSel{nullptr, 1, 0}, E->st);
      }))
    return FS_E_HIP;
  if (ens_sync(E)) return FS_E_HIP;
  for (int i = 0; i < n; i++) {
    if (!pois[i]) continue;
    fs_handle* h = E->m[i];
    if (int e = ens_fail(E, i)) return e;
    const int lerr = E->h_sc[i].done == 1 ? 0 : 1;
    log_solve(h, lerr);
    h->warm_ok = lerr == 0;
    h->moved = false;
    acc[i] = lerr == 0;
    rej[i] = lerr != 0;
  }
  if (ens_upload_round(E)) return FS_E_HIP;
  if (ens_launch(E, &E->g_end, [&]() {
        launch_ens_scalars(E->d_sc, r.acc, n, 1, E->st);
        launch_ens_scalars(E->d_sc, r.rej, n, 2, E->st);
      }))
    return FS_E_HIP;
  for (int i = 0; i < n; i++) {
    if (!pois[i]) continue;
    fs_handle* h = E->m[i];
    if (acc[i]) {
      h->rot = (h->rot + 1) % 3;
    } else if (copy_level(h, h->bufA(), h->bufL3())) {
      return FS_E_HIP;
    }
    h->alias = true;
    E->phase[i] = 0;
    left[i]--;
  }
  HIP_OK(hipGetLastError());
  return ens_sync(E);
}

void ens_release(fs_ensemble* E) {
  ens_drop_graphs(E);
  for (size_t i = 0; i < E->m.size(); i++) {
    fs_handle* h = E->m[i];
    if (i < E->own_d.size()) {
      h->d_sc = E->own_d[i];
      h->h_sc = E->own_h[i];
    }
    fs_destroy(h);
  }
  auto fr = [](void* p) {
    if (p) (void)hipFree(p);
  };
  fr(E->t_phys); fr(E->t_stage); fr(E->t_circ); fr(E->t_ctrl); fr(E->t_lapl); fr(E->t_coarsen);
  fr(E->t_init); fr(E->t_down); fr(E->t_up); fr(E->t_cc);
  fr(E->t_fin); fr(E->d_sc); fr(E->d_round);
  fr(E->m_stage.d); fr(E->m_one.d); fr(E->m_pcg.d); fr(E->m_st.d);
  for (int l = 0; l < kMgMaxLevels; l++) {
    fr(E->m_coarsen[l].d);
    fr(E->m_down[l].d);
    fr(E->m_up[l].d);
  }
  if (E->h_sc) (void)hipHostFree(E->h_sc);
  if (E->h_round) (void)hipHostFree(E->h_round);
  if (E->st) (void)hipStreamDestroy(E->st);
  delete E;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* fs_last_error(void) { return g_err.c_str(); }

int fs_constants(double p, double gam, double* out) {
  FsPhys K;
  fs_phys_init(&K, p, gam);
  fs_phys_table(&K, out);
  return 0;
}

int fs_create(const fs_grid_desc* g, const fs_params* p, fs_handle** out) {
  if (!g || !p || !out) return fail(FS_E_ARG, "null argument%s", "");
  if (g->nx < 3 || g->ny < 1) return fail(FS_E_ARG, "need nx >= 3 and ny >= 1%s", "");
  if (p->rf < 0 || p->rf > 4) return fail(FS_E_ARG, "unknown waveform rf%s", "");
  if (p->lapl_warm < 0 || p->lapl_warm > 2) return fail(FS_E_ARG, "lapl_warm must be 0, 1 or 2%s", "");
  if (p->p_torr <= 0) return fail(FS_E_ARG, "pressure must be positive%s", "");
  if (p->rf == 4 && (p->t_rise <= 0 || p->t_fall <= 0))
    return fail(FS_E_ARG, "trapezoid needs t_rise, t_fall > 0%s", "");
  fs_handle* h = new fs_handle();
  h->g = *g;
  h->p = *p;
  const bool oned = g->ny == 1;
  if (oned) {  // props.F90:64-65
    h->g.cyl = 0;
    h->g.rwall = 0;
    g = &h->g;
  }
  fs_phys_init(&h->K, p->p_torr, p->gam);
  h->bx = g->nx;
  h->by = g->ny;
  h->P = h->bx + 2;
  h->PL = (size_t)h->P * (h->by + 2);
  const int bx = h->bx, by = h->by;
  // grid: props.F90:96-218
  build_axis(g->grid_x, g->ratio_x, bx, g->l, 0, h->x);
  h->dx.resize(bx + 1);
  h->dlx.resize(bx);
  h->dy.resize(by + 1);
  h->dly.resize(by);
  for (int i = 0; i < bx + 1; i++) h->dx[i] = h->x[i + 1] - h->x[i];
  for (int i = 0; i < bx; i++) h->dlx[i] = 0.5 * (h->dx[i + 1] + h->dx[i]);
  if (!oned) {
    build_axis(g->grid_y, g->ratio_y, by, g->w, 1, h->y);
    for (int j = 0; j < by + 1; j++) h->dy[j] = h->y[j + 1] - h->y[j];
    for (int j = 0; j < by; j++) h->dly[j] = 0.5 * (h->dy[j + 1] + h->dy[j]);
  } else {
    // props.F90:175-178: y = 1, dy = dx(1); the kernels skip every y term, the
    // other y metrics only need to be finite (dly = dx(1), r = 1)
    h->y.assign(3, 1.0);
    h->dy.assign(2, h->dx[0]);
    h->dly.assign(1, h->dx[0]);
  }
  h->r = h->y;
  h->type_x.assign((size_t)bx * by, 0);
  h->type_y.assign((size_t)bx * by, 0);
  std::vector<int8_t> types(4 * (size_t)by, 0);
  std::vector<double> dr(2 * (size_t)by, 0.0);
  const bool active = p->circuit == 1;
  for (int j = 0; j < by; j++) {
    // reference uses y(j) (the node below cell j+1) in the electrode test (props.F90:201)
    bool el = (h->y[j] <= g->ew) || (h->y[j] >= g->gap);
    types[j] = el ? -2 : -1;
    types[by + j] = el ? 2 : 1;
    for (int i = 0; i < bx; i++) {
      size_t n = (size_t)j * bx + i;
      if (j == 0) h->type_y[n] = -1;
      if (j == by - 1) h->type_y[n] = g->rwall ? 3 : 1;  // props.F90:192-198
      if (i == 0) h->type_x[n] = types[j];
      if (i == bx - 1) h->type_x[n] = types[by + j];
    }
    // circ_lib.F90:89-151: dr weights; Fortran j = j+2 -> r(j) = r[j+1], dy(j-1) = dy[j]
    double rr = h->r[j + 1];
    double w = g->cyl ? h->dy[j] * h->r[j] * 2.0 * h->K.pi : h->dy[j] * g->w;
    if (!active && rr < g->ew) w = 0.0;
    if (oned) w = g->w * g->w * h->K.pi;  // circ_lib.F90:98-99, 131-132
    dr[j] = el ? w : 0.0;
    dr[by + j] = el ? w : 0.0;
    // ghost potential rows, circ_lib.F90:180-198 (one dimension: the
    // electrode cells' rows, circ_lib.F90:199-202)
    types[2 * by + j] = ((oned || rr <= g->ew) && el) ? 1 : 0;
    types[3 * by + j] = ((oned || rr <= g->ew) && el) ? 1 : 0;
  }
  h->nblk = 0;
  h->nbx = (bx + kTX - 1) / kTX;
  h->nby = (by + kSTY - 1) / kSTY;
  h->nblk = h->nbx * h->nby;
  h->pnby = (by + kPcgTY - 1) / kPcgTY;
  h->pnblk = h->nbx * h->pnby;
  // a partial-sum buffer holds a value set per block of any of the handle's
  // launches, and the stage kernel's per-row-segment error sums
  h->npart = (size_t)std::max(h->nblk, h->nbx * by) + 64;

  auto bad = [&](hipError_t e) {
    if (e == hipSuccess) return false;
    fail(FS_E_HIP, "HIP allocation failed: %s", hipGetErrorString(e));
    return true;
  };
  size_t PL = h->PL;
  if (bad(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking)) ||
      bad(hipMalloc(&h->d_lev, 16 * PL * sizeof(double))) ||
      bad(hipMalloc(&h->d_k, 12 * PL * sizeof(double))) ||
      bad(hipMalloc(&h->d_ph, PL * sizeof(double))) ||
      bad(hipMalloc(&h->d_geo, (5 * (size_t)bx + 5 * (size_t)by + 10) * sizeof(double))) ||
      bad(hipMalloc(&h->d_types, 4 * (size_t)by)) ||
      bad(hipMalloc(&h->d_dr, 2 * (size_t)by * sizeof(double))) ||
      bad(hipMalloc(&h->d_cur, 4 * (size_t)by * sizeof(double))) ||
      bad(hipMalloc(&h->d_part, 3 * h->npart * 8 * sizeof(double))) ||
      bad(hipMalloc(&h->d_lspart, 9 * kMaxPartials * sizeof(double))) ||
      bad(hipMalloc(&h->d_sc, sizeof(DevScalars))) ||
      bad(hipHostMalloc(&h->h_sc, sizeof(DevScalars), hipHostMallocDefault)) ||
      bad(hipHostMalloc(&h->h_pub, sizeof(DevScalars),
                        hipHostMallocMapped | hipHostMallocCoherent)) ||
      bad(hipHostGetDevicePointer((void**)&h->d_pub, h->h_pub, 0)) ||
      bad(hipMalloc(&h->d_pubcnt, 256)) || bad(hipMemset(h->d_pubcnt, 0, 256)) ||
      bad(hipMalloc(&h->d_pois, kPois * PL * sizeof(double))) ||
      (g->rwall && bad(hipMalloc(&h->d_sig, 8 * ((size_t)bx + 2) * sizeof(double))))) {
    fs_destroy(h);
    return FS_E_HIP;
  }
  memset(h->h_pub, 0, sizeof(DevScalars));  // publication count 0, as *d_pubcnt
  // sfc_init (sfc_lib.F90:82-90): the wall starts uncharged
  if (h->d_sig) (void)hipMemsetAsync(h->d_sig, 0, 8 * ((size_t)bx + 2) * sizeof(double), h->st);
  if (p->debug_keep_k) {
    if (bad(hipMalloc(&h->d_kdbg, 20 * PL * sizeof(double))) ||
        bad(hipMalloc(&h->d_dbg, 6 * (size_t)bx * by * sizeof(double)))) {
      fs_destroy(h);
      return FS_E_HIP;
    }
    (void)hipMemsetAsync(h->d_kdbg, 0, 20 * PL * sizeof(double), h->st);
  }
  for (int e = 0; e < 8; e++) (void)hipEventCreate(&h->ev[e]);
  // metrics
  std::vector<double> geo;
  geo.insert(geo.end(), h->dx.begin(), h->dx.end());
  geo.insert(geo.end(), h->dlx.begin(), h->dlx.end());
  geo.insert(geo.end(), h->dy.begin(), h->dy.end());
  geo.insert(geo.end(), h->dly.begin(), h->dly.end());
  geo.insert(geo.end(), h->r.begin(), h->r.end());
  for (const std::vector<double>* v : {&h->dx, &h->dlx, &h->dy, &h->dly})
    for (double a : *v) geo.push_back(1.0 / a);
  (void)hipMemcpyAsync(h->d_geo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice,
                       h->st);
  (void)hipMemcpyAsync(h->d_types, types.data(), types.size(), hipMemcpyHostToDevice, h->st);
  (void)hipMemcpyAsync(h->d_dr, dr.data(), dr.size() * sizeof(double), hipMemcpyHostToDevice,
                       h->st);
  h->G.dx = h->d_geo;
  h->G.dlx = h->G.dx + (bx + 1);
  h->G.dy = h->G.dlx + bx;
  h->G.dly = h->G.dy + (by + 1);
  h->G.r = h->G.dly + by;
  h->G.rdx = h->G.r + (by + 2);
  h->G.rdlx = h->G.rdx + (bx + 1);
  h->G.rdy = h->G.rdlx + bx;
  h->G.rdly = h->G.rdy + (by + 1);
  h->G.ltype = h->d_types;
  h->G.rtype = h->d_types + by;
  h->G.bx = bx;
  h->G.by = by;
  h->G.P = h->P;
  h->G.cyl = g->cyl;
  h->G.rwall = g->rwall ? 1 : 0;
  h->G.oned = oned ? 1 : 0;
  h->G.sig = h->d_sig;
  // zero state; PCG ghosts: r = p = x = 0 and diagonal 1 so that neighbour
  // recomputation at the boundary yields p = 0
  (void)hipMemsetAsync(h->d_k, 0, 12 * PL * sizeof(double), h->st);
  (void)hipMemsetAsync(h->d_ph, 0, PL * sizeof(double), h->st);
  (void)hipMemsetAsync(h->d_cur, 0, 4 * (size_t)by * sizeof(double), h->st);
  (void)hipMemsetAsync(h->d_pois, 0, kPois * PL * sizeof(double), h->st);
  launch_fill(h->pois(PO_DG), 1.0, PL, h->st);
  if (build_mg(h)) {
    fs_destroy(h);
    return FS_E_HIP;
  }
  // circ_init (circ_lib.F90:213-223) and main.F90:65-66
  DevScalars& s = *h->h_sc;
  memset(&s, 0, sizeof(s));
  s.t = 0.0;
  s.dt = p->dt0;
  s.err_prev = 1.0;
  s.t_m = (double)1e9f;
  s.Vd_pl = s.Vd_mi = s.Vd_or = s.Vsrc = p->vset;
  s.tol = p->lapl_rtol > 0 ? p->lapl_rtol : 1e-10;
  s.maxit = p->lapl_maxit > 0 ? p->lapl_maxit : 20000;
  s.sig = h->d_sig;
  s.sig_n = h->d_sig ? bx + 2 : 0;
  if (push_scalars(h)) {
    fs_destroy(h);
    return FS_E_HIP;
  }
  h->stage_buf.resize(PL);
  if (build_arg_tables(h) || fs_init_uniform(h) != 0 || choose_stage_tile(h)) {
    fs_destroy(h);
    return FS_E_HIP;
  }
  *out = h;
  return 0;
}

int fs_destroy(fs_handle* h) {
  if (!h) return 0;
  if (h->st) (void)hipStreamSynchronize(h->st);
  (void)hipFree(h->d_lev);
  (void)hipFree(h->d_k);
  (void)hipFree(h->d_kdbg);
  (void)hipFree(h->d_ph);
  (void)hipFree(h->d_sig);
  (void)hipFree(h->d_geo);
  (void)hipFree(h->d_types);
  (void)hipFree(h->d_dr);
  (void)hipFree(h->d_cur);
  (void)hipFree(h->d_part);
  (void)hipFree(h->d_lspart);
  (void)hipFree(h->d_sc);
  (void)hipHostFree(h->h_sc);
  if (h->h_pub) (void)hipHostFree(h->h_pub);
  (void)hipFree(h->d_pubcnt);
  (void)hipFree(h->d_pois);
  (void)hipFree(h->d_dbg);
  (void)hipFree(h->d_replay);
  (void)hipFree(h->d_mg);
  (void)hipFree(h->d_mginv);
  (void)hipFree(h->d_mgpk);
  drop_graphs(h);
  (void)hipFree(h->d_stab);
  (void)hipFree(h->d_ltab);
  (void)hipFree(h->d_phys);
  for (int e = 0; e < 8; e++)
    if (h->ev[e]) (void)hipEventDestroy(h->ev[e]);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return 0;
}

int fs_get_grid(const fs_handle* h, double* x, double* y, double* dx, double* dlx, double* dy,
                double* dly, double* r, int32_t* tx, int32_t* ty) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  if (x) memcpy(x, h->x.data(), h->x.size() * sizeof(double));
  if (y) memcpy(y, h->y.data(), h->y.size() * sizeof(double));
  if (dx) memcpy(dx, h->dx.data(), h->dx.size() * sizeof(double));
  if (dlx) memcpy(dlx, h->dlx.data(), h->dlx.size() * sizeof(double));
  if (dy) memcpy(dy, h->dy.data(), h->dy.size() * sizeof(double));
  if (dly) memcpy(dly, h->dly.data(), h->dly.size() * sizeof(double));
  if (r) memcpy(r, h->r.data(), h->r.size() * sizeof(double));
  if (tx) memcpy(tx, h->type_x.data(), h->type_x.size() * sizeof(int32_t));
  if (ty) memcpy(ty, h->type_y.data(), h->type_y.size() * sizeof(int32_t));
  return 0;
}

static double* field_dev(const fs_handle* h, int field, int level, bool for_write) {
  if (field >= FS_NI && field <= FS_NM) {
    int buf;
    if (level == 1) buf = (h->alias && !for_write) ? h->bufA() : h->bufL1();
    else if (level == 2) buf = h->bufA();
    else if (level == 3) buf = h->bufL3();
    else return nullptr;
    return h->lev(buf, field);
  }
  if (field == FS_PH) return h->d_ph;
  if (field >= FS_KI && field <= FS_KM) {
    if (!h->d_kdbg || level < 1 || level > 5) return nullptr;
    return h->d_kdbg + ((size_t)(level - 1) * 4 + (field - FS_KI)) * h->PL;
  }
  return nullptr;
}

// the wall's surface charge (FS_SIG): level 1 sig_pl, 2 sig_mi, 3 sig_or, nx+2
// entries (x nodes; the ghost ends are carried but never used)
static double* sig_dev(const fs_handle* h, int level) {
  if (!h->d_sig || level < 1 || level > 3) return nullptr;
  const int which = level == 1 ? kSigPl : level == 2 ? kSigMi : kSigOr;
  return h->d_sig + (size_t)which * (h->bx + 2);
}

int fs_set_field(fs_handle* h, int32_t field, int32_t level, const double* host, int32_t ld) {
  if (!h || !host) return fail(FS_E_ARG, "null argument%s");
  h->warm_ok = false;
  if (field == FS_SIG) {
    double* d = sig_dev(h, level);
    if (!d) return fail(FS_E_ARG, "no dielectric wall, or bad level%s");
    return upload(h, d, host, (size_t)h->bx + 2);
  }
  if (ld < h->P) return fail(FS_E_ARG, "ld < nx+2%s");
  double* d = field_dev(h, field, level, true);
  if (!d) return fail(FS_E_ARG, "bad field/level%s");
  for (int j = 0; j < h->by + 2; j++)
    memcpy(&h->stage_buf[(size_t)j * h->P], host + (size_t)j * ld, h->P * sizeof(double));
  if (field <= FS_NM && (level == 1 || level == 2) && h->alias) {
    // level 1 stops aliasing level 2 (a write to either would change both):
    // give it its own copy first
    for (int s = 0; s < 4; s++)
      HIP_OK(hipMemcpyAsync(h->lev(h->bufL1(), s), h->lev(h->bufA(), s), h->PL * sizeof(double),
                            hipMemcpyDeviceToDevice, h->st));
    h->alias = false;
  }
  return upload(h, d, h->stage_buf.data(), h->PL);
}

int fs_get_field(const fs_handle* h, int32_t field, int32_t level, double* host, int32_t ld) {
  if (!h || !host) return fail(FS_E_ARG, "null argument%s");
  if (field == FS_SIG) {
    const double* d = sig_dev(h, level);
    if (!d) return fail(FS_E_ARG, "no dielectric wall, or bad level%s");
    HIP_OK(hipMemcpyAsync(host, d, ((size_t)h->bx + 2) * sizeof(double), hipMemcpyDeviceToHost,
                          h->st));
    HIP_OK(hipStreamSynchronize(h->st));
    return 0;
  }
  if (ld < h->P) return fail(FS_E_ARG, "ld < nx+2%s");
  double* d = field_dev(h, field, level, false);
  if (!d) return fail(FS_E_ARG, "bad field/level%s");
  fs_handle* hm = const_cast<fs_handle*>(h);
  HIP_OK(hipMemcpyAsync(hm->stage_buf.data(), d, h->PL * sizeof(double), hipMemcpyDeviceToHost,
                        h->st));
  HIP_OK(hipStreamSynchronize(h->st));
  for (int j = 0; j < h->by + 2; j++)
    memcpy(host + (size_t)j * ld, &hm->stage_buf[(size_t)j * h->P], h->P * sizeof(double));
  return 0;
}

static int set_all_levels(fs_handle* h, int sp, const std::vector<double>& v) {
  for (int b = 0; b < 4; b++)
    HIP_OK(hipMemcpyAsync(h->lev(b, sp), v.data(), h->PL * sizeof(double), hipMemcpyHostToDevice,
                          h->st));
  HIP_OK(hipStreamSynchronize(h->st));
  return 0;
}

int fs_init_uniform(fs_handle* h) {  // ptcl_lib.F90:22-29
  if (!h) return fail(FS_E_ARG, "null handle%s");
  const FsPhys& K = h->K;
  std::vector<double> v(h->PL, K.n_init);
  if (set_all_levels(h, 0, v) || set_all_levels(h, 1, v) || set_all_levels(h, 3, v))
    return FS_E_HIP;
  for (auto& a : v) a = K.n_init * K.Tg * K.kb / K.e / K.ph0;
  if (set_all_levels(h, 2, v)) return FS_E_HIP;
  h->alias = true;
  return 0;
}

namespace {
uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// uniform in [-1, 1) keyed by (seed, species, i, j), Fortran indices i, j
double unoise(uint64_t seed, int sp, int i, int j) {
  uint64_t h = splitmix64(splitmix64(seed) ^ ((uint64_t)sp << 56) ^ ((uint64_t)j << 28) ^
                          (uint64_t)i);
  return (double)(h >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}
}  // namespace

int fs_init_gaussian(fs_handle* h, uint64_t seed, double peak, double sigma, double noise) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  const FsPhys& K = h->K;
  double xc = h->g.l * 0.5, yc = h->g.w * 0.5;
  std::vector<double> f[4];
  for (auto& a : f) a.resize(h->PL);
  for (int j = 0; j < h->by + 2; j++)
    for (int i = 0; i < h->bx + 2; i++) {
      // in one dimension the blob is a profile in x alone
      double ddx = h->x[i] - xc, ddy = h->by > 1 ? h->y[j] - yc : 0.0;
      double base = peak * exp(-(ddx * ddx + ddy * ddy) / (2.0 * sigma * sigma));
      size_t m = (size_t)j * h->P + i;
      double ni = base * (1.0 + noise * unoise(seed, 0, i + 1, j + 1)) + K.n_zero;
      double ne = base * (1.0 + noise * unoise(seed, 1, i + 1, j + 1)) + K.n_zero;
      f[0][m] = ni;
      f[1][m] = ne;
      f[2][m] = ne * K.Tg * K.kb / K.e / K.ph0;
      f[3][m] = base + K.n_zero;
    }
  for (int s = 0; s < 4; s++)
    if (set_all_levels(h, s, f[s])) return FS_E_HIP;
  h->alias = true;
  return 0;
}

int fs_get_scalars(const fs_handle* h, double* o) {
  if (!h || !o) return fail(FS_E_ARG, "null argument%s");
  fs_handle* hm = const_cast<fs_handle*>(h);
  if (sync_scalars(hm)) return FS_E_HIP;
  const DevScalars& s = *h->h_sc;
  double v[11] = {s.t,     s.dt,  s.err_prev, s.t_m, s.Vd_pl, s.Vd_mi,
                  s.Vd_or, s.Ida, s.Idc,      s.Vsrc, s.ts};
  memcpy(o, v, sizeof(v));
  return 0;
}

int fs_stage_tile(const fs_handle* h) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  return h->sty;
}

int fs_circ_init(fs_handle* h, double vset) {  // circ_lib.F90:213-223
  if (!h) return fail(FS_E_ARG, "null handle%s");
  if (h->in_ens) return fail(FS_E_ARG, "circ_init on an ensemble member%s");
  if (sync_scalars(h)) return FS_E_HIP;
  h->p.vset = vset;
  // the stage launches carry the circuit's amplitude in their argument tables
  if (build_arg_tables(h)) return FS_E_HIP;
  DevScalars& s = *h->h_sc;
  s.Vsrc = s.Vd_pl = s.Vd_mi = s.Vd_or = vset;
  return push_scalars(h);
}

int fs_set_scalars(fs_handle* h, const double* o) {
  if (!h || !o) return fail(FS_E_ARG, "null argument%s");
  if (sync_scalars(h)) return FS_E_HIP;
  DevScalars& s = *h->h_sc;
  s.t = o[0];
  s.dt = o[1];
  s.err_prev = o[2];
  s.t_m = o[3];
  s.Vd_pl = o[4];
  s.Vd_mi = o[5];
  s.Vd_or = o[6];
  s.Ida = o[7];
  s.Idc = o[8];
  s.Vsrc = o[9];
  s.ts = o[10];
  return push_scalars(h);
}

int fs_ptcl_step(fs_handle* h, fs_step_info* info) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  return ptcl_step_impl(h, info);
}

int fs_lapl_solve(fs_handle* h, int32_t* lerr, fs_solve_stats* stats) {
  if (!h || !lerr) return fail(FS_E_ARG, "null argument%s");
  int l = 0;
  int e = lapl_solve_impl(h, &l, stats);
  *lerr = l;
  if (e) return e;
  return l ? FS_NOT_CONVERGED : 0;
}

int fs_accept(fs_handle* h) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  return accept_impl(h);
}

int fs_reject(fs_handle* h) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  return reject_impl(h);
}

// The host copy of the scalars is refreshed by the reads inside each pass (the
// controller's decision, the end of the solve); between passes only ts, which
// decides whether the next pass starts with a transport step, is advanced on
// the host as k_accept advances it on the device, and one read at the end
// leaves the copy current.
int fs_run(fs_handle* h, int32_t n_iters) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  for (int it = 0; it < n_iters; it++) {
    int lerr = 0;
    if (h->h_sc->ts > 1.0) {  // main.F90:75
      const bool ahead = can_issue_ahead(h);
      if (int e = ptcl_step_impl(h, nullptr, ahead, it + 1 < n_iters)) return e;
      if (int e = lapl_solve_impl(h, &lerr, nullptr, ahead ? 2 : 0)) return e;
    } else {
      launch_begin_step(h->d_sc, h->st);
      if (int e = lapl_solve_impl(h, &lerr, nullptr)) return e;
    }
    int e = lerr == 0 ? accept_impl(h) : reject_impl(h);
    if (e) return e;
    if (lerr == 0) h->h_sc->ts = h->h_sc->ts + 1.0;
  }
  return sync_scalars(h) ? FS_E_HIP : 0;
}

int fs_set_replay(fs_handle* h, const double* recs, int32_t n) {
  if (!h || (n > 0 && !recs)) return fail(FS_E_ARG, "null argument%s");
  (void)hipFree(h->d_replay);
  h->d_replay = nullptr;
  h->nreplay = 0;
  if (n > 0) {
    HIP_OK(hipMalloc(&h->d_replay, (size_t)n * FS_REC_LEN * sizeof(double)));
    HIP_OK(hipMemcpy(h->d_replay, recs, (size_t)n * FS_REC_LEN * sizeof(double),
                     hipMemcpyHostToDevice));
    h->nreplay = n;
  }
  if (sync_scalars(h)) return FS_E_HIP;
  h->h_sc->replay_n = n;
  h->h_sc->replay_idx = 0;
  return push_scalars(h);
}

int fs_log_count(const fs_handle* h) { return h ? (int)(h->log.size() / FS_REC_LEN) : 0; }

int fs_get_log(const fs_handle* h, double* out, int32_t max) {
  if (!h || !out) return fail(FS_E_ARG, "null argument%s");
  int n = fs_log_count(h);
  if (n > max) n = max;
  memcpy(out, h->log.data(), (size_t)n * FS_REC_LEN * sizeof(double));
  return n;
}

int fs_clear_log(fs_handle* h) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  h->log.clear();
  return 0;
}

int fs_rhs_sweep(fs_handle* h, int32_t stage) {
  if (!h || stage < 1 || stage > 5) return fail(FS_E_ARG, "bad argument%s");
  if (!h->d_kdbg) return fail(FS_E_ARG, "fs_rhs_sweep needs debug_keep_k%s");
  launch_stage(stab(h, stage, 1), h->d_phys, h->nbx, h->by, h->sty, h->st);
  HIP_OK(hipGetLastError());
  if (sync_scalars(h)) return FS_E_HIP;
  return check_status(h);
}

int fs_circ_step(fs_handle* h, int32_t stage) {
  if (!h || stage < 1 || stage > 5) return fail(FS_E_ARG, "bad argument%s");
  // currents need the stage-kernel face terms of the current level 2
  launch_stage(stab(h, stage, 2), h->d_phys, h->nbx, h->by, h->sty, h->st);
  CircArgs c = circ_args(h, stage);
  launch_circ(c, h->K, h->st);
  HIP_OK(hipGetLastError());
  if (sync_scalars(h)) return FS_E_HIP;
  return check_status(h);
}

int fs_lapl_assemble(fs_handle* h, double* b, double* A5) {
  if (!h || !b || !A5) return fail(FS_E_ARG, "null argument%s");
  if (!h->d_dbg) return fail(FS_E_ARG, "fs_lapl_assemble needs debug_keep_k%s");
  launch_assembly(h);
  HIP_OK(hipGetLastError());
  size_t N = (size_t)h->bx * h->by;
  HIP_OK(hipMemcpyAsync(b, h->d_dbg, N * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIP_OK(hipMemcpyAsync(A5, h->d_dbg + N, 5 * N * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIP_OK(hipStreamSynchronize(h->st));
  if (sync_scalars(h)) return FS_E_HIP;
  return check_status(h);
}

int fs_eval_fits(double p, int32_t n, const double* T, double* out) {
  if (n <= 0) return 0;
  FsPhys K;
  fs_phys_init(&K, p, 0.1);
  double *dT = nullptr, *dO = nullptr;
  HIP_OK(hipMalloc(&dT, n * sizeof(double)));
  HIP_OK(hipMalloc(&dO, 12 * (size_t)n * sizeof(double)));
  HIP_OK(hipMemcpy(dT, T, n * sizeof(double), hipMemcpyHostToDevice));
  launch_eval_fits(K, dT, dO, n, nullptr);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(out, dO, 12 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  (void)hipFree(dT);
  (void)hipFree(dO);
  return 0;
}

static int eval_flux_common(int32_t n, const double* a, double* out, int dfdph) {
  if (n <= 0) return 0;
  FsPhys K;
  fs_phys_init(&K, 10.0, 0.1);
  double *dA = nullptr, *dO = nullptr;
  HIP_OK(hipMalloc(&dA, 7 * (size_t)n * sizeof(double)));
  HIP_OK(hipMalloc(&dO, (size_t)n * sizeof(double)));
  HIP_OK(hipMemcpy(dA, a, 7 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  launch_eval_flux(K, dA, dO, n, dfdph, nullptr);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(out, dO, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(dO);
  return 0;
}

int fs_eval_flux(int32_t n, const double* a, double* out) { return eval_flux_common(n, a, out, 0); }
int fs_eval_dfdph(int32_t n, const double* a, double* out) {
  return eval_flux_common(n, a, out, 1);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Staging of fs_eval_flux_noalloc: a file-scope device array instead of the
// per-call hipMalloc of eval_flux_common.  Each device holds its own copy for
// the life of the process; its address comes from hipGetSymbolAddress once per
// device (the array's name is a host-side stand-in, never a device address).
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)2 << 20;  // 2 MiB per device
constexpr int32_t kNoAllocMaxN = (int32_t)(kNoAllocScratchBytes / (8 * sizeof(double)));
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kNoAllocMaxDevices];  // device address of `scratch`, per device
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

extern "C" {

int fs_eval_flux_noalloc(int32_t n, const double* a, double* out, int32_t dfdph) {
  if (n < 0 || n > kNoAllocMaxN)
    return fail(FS_E_ARG, "fs_eval_flux_noalloc: n must lie in 0..%s%d", "", kNoAllocMaxN);
  if (n == 0) return 0;
  if (!a || !out) return fail(FS_E_ARG, "null argument%s");
  int dev = -1;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(FS_E_ARG, "device id out of range%s");
  // calls in this process are serialised: they share the device's one array
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  double* ptr = g_noalloc_ptr[dev];
  if (!ptr) {
    HIP_OK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(FS_E_HIP, "no device address for the staging array%s");
    g_noalloc_ptr[dev] = ptr;
  }
  FsPhys K;
  fs_phys_init(&K, 10.0, 0.1);
  double* dA = ptr;                    // 7 n doubles of arguments
  double* dO = ptr + 7 * (size_t)n;    // n doubles of results
  HIP_OK(hipMemcpy(dA, a, 7 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  launch_eval_flux(K, dA, dO, n, dfdph ? 1 : 0, nullptr);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(out, dO, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int fs_time_kernel(fs_handle* h, int32_t which, int32_t n, double* avg_ms) {
  if (!h || !avg_ms || n <= 0) return fail(FS_E_ARG, "bad argument%s");
  h->warm_ok = false;  // the timed launches overwrite the solver's vectors
  if (sync_scalars(h)) return FS_E_HIP;
  const double tol0 = h->h_sc->tol;
  const int32_t maxit0 = h->h_sc->maxit;
  h->h_sc->done = 0;  // let the PCG kernels run
  h->h_sc->maxit = 1 << 30;
  h->h_sc->tol = 0.0;
  if (push_scalars(h)) return FS_E_HIP;
  auto launch = [&](int k) {
    if (which >= 1 && which <= 5) {
      launch_stage(stab(h, which, 0), h->d_phys, h->nbx, h->by, h->sty, h->st);
    } else if (which == 10) {
      launch_assembly(h);
    } else if (which == 11) {
      launch_pcg_it(pcg_args(h, 2 + (k & 1)), h->st);
    } else if (which == 15) {
      const size_t c = learned_level(h);
      (void)fsl::net_apply(&h->net->net, h->mg[c], h->part(2), mg_grid(h->mg[c - 1]), h->d_sc,
                           h->st);
    } else if (which == 16) {
      (void)launch_learned_iteration(h, 2 + (k & 1));
    } else if (which == 17) {
      launch_vcycle(h, 2 + k);
    } else if (which == 13) {
      // first-iteration form (beta = 0, p and s_old not read): repeated on the
      // same partials it stays a valid update, unlike the recurrence form
      launch_mg_down(mg_down_args(h, 0, 0), h->st);
    } else if (which == 14) {
      launch_mg_up(mg_up_args(h, 0, 2 + (k & 1)), h->st);
    } else if (which == 18) {
      launch_mg_coarse_cycle(cycle_args(h, 2 + (k & 1)), h->st);
    } else if (which >= 20 && which < 40) {
      launch_mg_down(mg_down_args(h, which - 20, 2 + (k & 1)), h->st);
    } else if (which >= 40 && which < 60) {
      launch_mg_up(mg_up_args(h, which - 40, 2 + (k & 1)), h->st);
    } else {
      PcgArgs ia = pcg_args(h, 0);
      ia.part_out = h->part(0);
      ia.w_new = h->pois(PO_W0);
      launch_pcg_init(ia, h->st);
    }
  };
  const bool level_id = (which >= 20 && which < 60);
  if (!(which >= 1 && which <= 5) && !(which >= 10 && which <= 18) && !level_id)
    return fail(FS_E_ARG, "unknown kernel id%s");
  if (level_id && h->mg.size() > 1 && (size_t)(which % 20) >= h->mg_first)
    return fail(FS_E_ARG, "no tile kernel on that multigrid level%s");
  if (which >= 13 && h->mg.size() < 2)
    return fail(FS_E_ARG, "grid too small for a multigrid hierarchy%s");
  if (which == 15 && !h->net) return fail(FS_E_ARG, "no network attached%s");
  if (which == 11) {  // consistent partials for the Jacobi-PCG iterations
    PcgArgs ia = pcg_args(h, 0);
    ia.part_out = h->part(0);
    ia.w_new = h->pois(PO_W0);
    launch_pcg_init(ia, h->st);
  }
  if (which == 13 || which == 14 || which == 17 || which == 18 || level_id) {
    // a consistent multigrid-PCG state: start, the initial cycle, iterations 0 and 1
    for (size_t l = 0; l < h->mg_first; l++) launch_mg_coarsen(mg_args(h, l), h->st);
    launch_mg_coarse_setup(cycle_args(h, -1), h->st);
    MgPcgArgs ia = mgpcg_args(h, 1);
    ia.part_out = h->part(0);
    launch_mgpcg_init(ia, h->st);
    if (launch_vcycle(h, -1) || launch_vcycle(h, 0) || launch_vcycle(h, 1)) return FS_E_HIP;
  }
  // iteration numbers keep alternating in parity through the warm-up launch
  int seq = 0;
  launch(seq++);  // warm
  HIP_OK(hipEventRecord(h->ev[6], h->st));
  for (int k = 0; k < n; k++) launch(seq++);
  HIP_OK(hipEventRecord(h->ev[7], h->st));
  HIP_OK(hipEventSynchronize(h->ev[7]));
  HIP_OK(hipGetLastError());
  float ms = 0;
  HIP_OK(hipEventElapsedTime(&ms, h->ev[6], h->ev[7]));
  *avg_ms = (double)ms / n;
  // the solver settings come back; the fields do not (see the declaration)
  if (sync_scalars(h)) return FS_E_HIP;
  h->h_sc->tol = tol0;
  h->h_sc->maxit = maxit0;
  h->h_sc->done = 0;
  return push_scalars(h);
}

int fs_copy_bench(int64_t bytes, int32_t reps, double* gbps) {
  if (bytes < 16 || reps <= 0 || !gbps) return fail(FS_E_ARG, "bad argument%s");
  size_t n2 = (size_t)bytes / 16;
  double *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  HIP_OK(hipMalloc(&a, n2 * 16));
  HIP_OK(hipMalloc(&b, n2 * 16));
  HIP_OK(hipMemset(a, 0, n2 * 16));
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  launch_copy(a, b, n2, nullptr);
  HIP_OK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; r++) launch_copy(a, b, n2, nullptr);
  HIP_OK(hipEventRecord(e1, nullptr));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  *gbps = 2.0 * (double)(n2 * 16) * reps / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int fs_get_timers(const fs_handle* h, double* out4) {
  if (!h || !out4) return fail(FS_E_ARG, "null argument%s");
  memcpy(out4, h->tms, sizeof(h->tms));
  return 0;
}

int fs_set_solver(fs_handle* h, int32_t solver) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  if (solver < 0 || solver > 2) return fail(FS_E_ARG, "unknown Poisson solver%s");
  h->p.lapl_solver = solver;
  h->warm_ok = false;
  return 0;
}

int fs_ensemble_create(int32_t n, const fs_grid_desc* g, const fs_params* p, fs_ensemble** out) {
  if (n <= 0 || !g || !p || !out) return fail(FS_E_ARG, "bad argument%s");
  if (p[0].stage_tile != 0 && !valid_stage_tile(p[0].stage_tile))
    return fail(FS_E_ARG, "stage_tile must be 0, 4, 8 or 16%s");
  for (int i = 0; i < n; i++) {
    if (g[i].rwall) return fail(FS_E_ARG, "ensemble members have no dielectric wall%s");
    if (g[i].ny < 2) return fail(FS_E_ARG, "ensemble members are two-dimensional (ny >= 2)%s");

  }
  fs_ensemble* E = new fs_ensemble();
  E->n = n;
  for (int i = 0; i < n; i++) E->projected = E->projected || p[i].lapl_warm == 2;
  for (int i = 0; i < n; i++) {
    if (p[i].lapl_solver != 1) {
      ens_release(E);
      return fail(FS_E_ARG, "ensemble members use the multigrid solver (lapl_solver 1)%s");
    }
    fs_handle* h = nullptr;
    t_tile_deferred = true;
    const int ce = fs_create(&g[i], &p[i], &h);
    t_tile_deferred = false;
    if (ce) {
      ens_release(E);
      return ce;
    }
    E->m.push_back(h);
    if (h->mg.size() < 2) {
      ens_release(E);
      return fail(FS_E_ARG, "ensemble member grid of <= 64 cells%s");
    }
  }
  // the members' scalars move into one array so that a round reads them with
  // one copy; the members' graphs (made with the old address) are dropped
  if (bad(hipStreamCreate(&E->st)) || bad(hipMalloc(&E->d_sc, n * sizeof(DevScalars))) ||
      bad(hipHostMalloc(&E->h_sc, n * sizeof(DevScalars)))) {
    ens_release(E);
    return fail(FS_E_HIP, "out of memory for the ensemble%s");
  }
  for (int i = 0; i < n; i++) {
    fs_handle* h = E->m[i];
    if (sync_scalars(h)) {
      ens_release(E);
      return FS_E_HIP;
    }
    E->h_sc[i] = *h->h_sc;
    E->own_d.push_back(h->d_sc);
    E->own_h.push_back(h->h_sc);
    h->d_sc = E->d_sc + i;
    h->h_sc = E->h_sc + i;
    h->in_ens = true;
    drop_graphs(h);
    if (build_arg_tables(h)) {
      ens_release(E);
      return FS_E_HIP;
    }
  }
  E->round_bytes = (size_t)n * 13;
  if (bad(hipMemcpy(E->d_sc, E->h_sc, n * sizeof(DevScalars), hipMemcpyHostToDevice)) ||
      bad(hipMalloc(&E->d_round, E->round_bytes)) ||
      bad(hipHostMalloc(&E->h_round, E->round_bytes)) || ens_maps(E) ||
      ens_choose_stage_tile(E, p[0].stage_tile)) {
    ens_release(E);
    return fail(FS_E_HIP, "ensemble set-up failed%s");
  }
  E->r.sel = (int32_t*)E->d_round;
  E->r.sel2 = E->r.sel + n;
  E->r.ptcl = (uint8_t*)(E->d_round + 8 * n);
  E->r.pois = E->r.ptcl + n;
  E->r.begin = E->r.pois + n;
  E->r.acc = E->r.begin + n;
  E->r.rej = E->r.acc + n;
  E->phase.assign(n, 0);
  *out = E;
  return 0;
}

int fs_ensemble_stage_tile(const fs_ensemble* E) {
  if (!E) return fail(FS_E_ARG, "null ensemble%s");
  return E->sty;
}

int fs_ensemble_destroy(fs_ensemble* E) {
  if (!E) return 0;
  if (E->st) (void)hipStreamSynchronize(E->st);
  ens_release(E);
  return 0;
}

fs_handle* fs_ensemble_member(fs_ensemble* E, int32_t i) {
  if (!E || i < 0 || i >= E->n) {
    fail(FS_E_ARG, "no such ensemble member%s");
    return nullptr;
  }
  return E->m[i];
}

int fs_ensemble_run(fs_ensemble* E, int32_t passes) {
  if (!E || passes < 0) return fail(FS_E_ARG, "bad argument%s");
  const int n = E->n;
  for (fs_handle* h : E->m) HIP_OK(hipStreamSynchronize(h->st));
  ens_drop_graphs(E);  // they hold the previous run's tables
  if (ens_tables(E)) return FS_E_HIP;
  E->graphs = true;
  for (fs_handle* h : E->m) E->graphs = E->graphs && h->p.use_graphs;
  // members work on the ensemble's stream for the run (restart copies)
  E->own_st.clear();
  for (fs_handle* h : E->m) {
    E->own_st.push_back(h->st);
    h->st = E->st;
  }
  std::vector<int> left(n, passes);
  int err = ens_sync(E);
  while (!err) {
    bool any = false;
    for (int i = 0; i < n; i++) any = any || left[i] > 0;
    if (!any) break;
    err = ens_round(E, left);
  }
  if (!err && hipStreamSynchronize(E->st) != hipSuccess) err = fail(FS_E_HIP, "stream%s");
  for (int i = 0; i < n; i++) E->m[i]->st = E->own_st[i];
  return err;
}

int fs_learn_create(int32_t capacity, int32_t H, int32_t W, uint64_t seed, fs_learn** out) {
  if (!out || capacity < 0 || H <= 0 || W <= 0 || H % fsl::kAlign || W % fsl::kAlign)
    return fail(FS_E_ARG, "bad argument (sides must be positive multiples of 8)%s");
  fs_learn* n = new fs_learn();
  HIP_OK(hipStreamCreate(&n->net.st));
  if (fsl::net_init(&n->net, capacity, H, W, seed, n->net.st)) {
    fsl::net_free(&n->net);
    delete n;
    return fail(FS_E_HIP, "out of device memory for the network%s");
  }
  *out = n;
  return 0;
}

int fs_learn_destroy(fs_learn* n) {
  if (!n) return 0;
  (void)hipStreamSynchronize(n->net.st);
  fsl::net_free(&n->net);
  (void)hipStreamDestroy(n->net.st);
  delete n;
  return 0;
}

int64_t fs_learn_nparams(const fs_learn* n) { return n ? (int64_t)n->net.nparam : -1; }

int fs_learn_get_weights(const fs_learn* n, float* out) {
  if (!n || !out) return fail(FS_E_ARG, "null argument%s");
  HIP_OK(hipMemcpy(out, n->net.w, n->net.nparam * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int fs_learn_set_weights(fs_learn* n, const float* in) {
  if (!n || !in) return fail(FS_E_ARG, "null argument%s");
  HIP_OK(hipMemcpy(n->net.w, in, n->net.nparam * sizeof(float), hipMemcpyHostToDevice));
  fsl::net_refresh(&n->net);
  HIP_OK(hipStreamSynchronize(n->net.st));
  return 0;
}

int fs_learn_samples(const fs_learn* n) { return n ? n->net.count : -1; }

int fs_learn_harvest(fs_learn* n, fs_handle* h, int32_t* added) {
  if (!n || !h) return fail(FS_E_ARG, "null argument%s");
  if (h->mg.size() < 2)
    return fail(FS_E_ARG, "the grid has no coarse level (<= 64 cells, or 1-D)%s");
  const MgLevel& lc = h->mg[learned_level(h)];
  if (lc.nx > n->net.W || lc.ny > n->net.H)
    return fail(FS_E_ARG, "learned level larger than the training grid%s");
  h->warm_ok = false;
  const int before = n->net.count;
  HIP_OK(hipStreamSynchronize(n->net.st));
  launch_assembly(h);
  for (size_t l = 0; l < h->mg_first; l++) launch_mg_coarsen(mg_args(h, l), h->st);
  launch_mg_coarse_setup(cycle_args(h, -1), h->st);
  MgPcgArgs ia = mgpcg_args(h, 1);
  ia.part_out = h->part(0);
  launch_mgpcg_init(ia, h->st);
  if (launch_vcycle(h, -1, &n->net)) return FS_E_HIP;
  for (int it = 0; it < h->h_sc->maxit; it++) {
    if (launch_vcycle(h, it, &n->net)) return FS_E_HIP;
    if (h->h_sc->done || n->net.count >= n->net.cap) break;
  }
  HIP_OK(hipStreamSynchronize(h->st));
  HIP_OK(hipGetLastError());
  if (added) *added = n->net.count - before;
  return 0;
}

int fs_learn_train(fs_learn* n, int32_t steps, int32_t batch, float lr, uint64_t seed,
                   double* loss) {
  if (!n || steps < 0 || batch <= 0) return fail(FS_E_ARG, "bad argument%s");
  if (n->net.count == 0) return fail(FS_E_ARG, "empty training set%s");
  std::vector<int> idx(batch);
  uint64_t s = seed;
  for (int k = 0; k < steps; k++) {
    for (int b = 0; b < batch; b++) {
      uint64_t z = (s += 0x9e3779b97f4a7c15ull);
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      idx[b] = (int)((z ^ (z >> 31)) % (uint64_t)n->net.count);
    }
    double l = 0.0;
    if (fsl::train_step(&n->net, idx.data(), batch, lr, &l))
      return fail(FS_E_HIP, "training step failed (batch too large?)%s");
    if (loss) loss[k] = l;
  }
  return 0;
}

int fs_learn_loss_grad(fs_learn* n, int32_t B, int32_t H, int32_t W, const float* in5,
                       const float* rhat, const float* op, double* loss, float* grad, float* out) {
  if (!n || !in5 || !rhat || !op || !loss || B <= 0 || H <= 0 || W <= 0 || H % fsl::kAlign ||
      W % fsl::kAlign)
    return fail(FS_E_ARG, "bad argument%s");
  if (fsl::loss_grad(&n->net, B, H, W, in5, rhat, op, loss, grad, out))
    return fail(FS_E_HIP, "network evaluation failed%s");
  return 0;
}

int fs_learn_attach(fs_handle* h, fs_learn* n) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  // learned-iteration graphs captured over the previous network's buffers are
  // never replayed again (a new network may sit at a freed one's address)
  for (int k = 0; k < 2; k++) {
    if (h->learn_graph[k]) (void)hipGraphExecDestroy(h->learn_graph[k]);
    h->learn_graph[k] = nullptr;
  }
  h->learn_graph_net = nullptr;
  h->learn_graph_acts = 0;
  h->net = n;
  return 0;
}

int fs_set_timing(fs_handle* h, int32_t on) {
  if (!h) return fail(FS_E_ARG, "null handle%s");
  h->timing = on != 0;
  for (double& v : h->tms) v = 0;
  return 0;
}

}  // extern "C"
