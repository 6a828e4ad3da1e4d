// fsim_internal.h -- device state layout and kernel launchers of libfsim_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include "fsim.h"
#include "fsim_phys.h"

// Scalars that live on the device so that a whole step can run without host
// round trips.  One instance per handle.
struct DevScalars {
  double t, dt, dt_stage, err_prev, err_cur, t_m;
  double Vd_pl, Vd_mi, Vd_or, Ida, Idc, Vsrc;
  double kv[6];
  double err4[4];
  double ts;
  // Poisson / PCG
  double cg_gam[2], cg_alpha[2], bnorm2, resid, tol, tres;
  double cg_pq[2];   // p.q of the last two iterations (flexible CG)
  double nn_scale;   // input scale of the learned coarse solver
  // warm start: x holds the correction of the last solve, after a transport
  // step of dt_warm, x2 (when x2_ok) the one before, after dt_warm2.  The next
  // solve starts from warm_a x + warm_b x2: the corrections per unit step
  // length extrapolated linearly to the new step (or, without x2, x scaled by
  // the step-length ratio).  warm_now: the solve in flight starts warm.
  // While the step length still changes, x3 (when x3_ok, after dt_warm3)
  // makes the extrapolation quadratic: + warm_c x3.
  double dt_warm, dt_warm2, warm_a, warm_b;
  double dt_warm3, warm_c;
  int32_t warm_now, x2_ok;
  int32_t x3_ok, pad_x;
  int32_t itmax, pad_w;  // iteration limit of the MG-PCG solve in flight (set by its init kernel)
  int32_t iters, done, maxit, pad_i;
  int32_t status, accept, attempts, replay_idx;
  int32_t replay_n, pad_j[3];
  // dielectric wall (rwall): the surface-charge arrays (sig_pl, sig_mi, sig_or
  // of sig_n = bx+2 entries each, Geo::sig) for the accept, reject and
  // restart copies; sig_n = 0 without the wall
  double* sig;
  int32_t sig_n, pad_s;
  int32_t itc[2], pad_k[2];   // MG-PCG iteration count by parity
  uint32_t ticket[8];
  uint32_t pub_seq;  // host-memory copy only: the publication count (k_publish)
  int32_t spec;      // k_accept_spec accepted the last pass: the gated stage kernel runs
};

// Grid metrics on the device, 0-based over interior cells ci in [0,bx), cj in
// [0,by).  Face k of the x axis lies between cells k-1 and k (k in [0,bx]).
struct Geo {
  const double* dx;   // bx+1  (Fortran dx(k+1))
  const double* dlx;  // bx
  const double* dy;   // by+1
  const double* dly;  // by
  const double* r;    // by+2  (Fortran r(k+1))
  const int8_t* ltype;  // by: type_x of the first cell of each row (-2 / -1)
  const int8_t* rtype;  // by: type_x of the last cell of each row (2 / 1)
  const double* rdx;    // 1/dx, 1/dlx, 1/dy, 1/dly
  const double* rdlx;
  const double* rdy;
  const double* rdly;
  int32_t bx, by, P, cyl;
  // dielectric top wall (props.F90:47, 192-198): the top row is type_y = 3.
  // sig: 8 arrays of bx+2 (x-node index i+1 for cell i): sig_pl, sig_mi,
  // sig_or (sfc_lib.F90:7), then the surface-charge slopes of stages 1-5
  double* sig;
  int32_t rwall;
  // one dimension (ny = 1, by = 1): every y term is left out (ptcl_lib.F90:285,
  // 455; lapl_lib.F90:255, 462, 577)
  int32_t oned;
};
constexpr int kSigPl = 0, kSigMi = 1, kSigOr = 2, kSigK = 2;  // ks of stage s: kSigK + s

// species order everywhere: 0 ni, 1 ne, 2 nt, 3 nm
struct CircArgs {
  double* ph;
  const double* cur;
  const double* dra;    // by: current weights at the left electrode face
  const double* drc;    // by: current weights at the right electrode face
  const int8_t* gl;     // by: ghost potential set on the left
  const int8_t* gr;     // by: ghost potential set on the right
  DevScalars* sc;
  double vset, res, t_rise, t_width, t_fall;
  int32_t rf, circuit, stage, by, P, bx;
  int32_t oned;  // one dimension: the ghost potential is the source voltage (circ_lib.F90:199-202)
  int32_t pad_c;
};

struct StageArgs {
  const double* rd[4];  // level-2 densities read by this stage
  double* wr[4];        // stage 1-4: next predictor, stage 5: level 1
  const double* n3[4];  // RK base (level 3)
  double* k1[4];
  double* k3[4];
  double* k4[4];
  double* kdbg[4];      // all-stage slope copy (debug), may be null
  const double* ph;
  double* cur;          // 2*by: Ida face terms, then Idc face terms
  double* part;         // per-block partials, 8 doubles per block
  DevScalars* sc;
  Geo G;
  int32_t stage, rhs_only, nbx, nby;
  // cc.stage > 0: the circuit update of that (the previous) stage is done by
  // this launch before its fluxes, as k_circ would (the circuit of stage 5
  // goes with the controller, k_circ_ctrl)
  CircArgs cc;
};


struct CtrlArgs {
  const double* part;
  int32_t nblocks;
  DevScalars* sc;
  const double* replay;  // FS_REC_LEN per record
};

struct MgLevel {
  double* d;    // diagonal of the (Galerkin) operator
  double* cx;   // coupling to the right neighbour
  double* cy;   // coupling to the upper neighbour
  double* r;    // residual handed down to this level (level 0: the PCG residual)
  double* x1;   // pre-smoothed iterate
  double* e;    // correction returned to the finer level (level 0: u = B r)
  int32_t nx, ny, P, pad_;
};

struct LaplArgs {
  const double* n[3];   // level-1 ni, ne, nt
  const double* ph;
  double* rhs0;         // V*b (kept for the true residual)
  double* r;            // CG residual (starts as V*b)
  double* cx;           // x-face coupling to the right neighbour (symmetrised)
  double* cy;           // y-face coupling to the upper neighbour
  double* dg;           // diagonal of -V*A
  double* bdbg;         // debug: reference-form b (nx*ny), may be null
  double* adbg;         // debug: reference-form A5 (nx*ny*5), may be null
  double* part;
  DevScalars* sc;
  Geo G;
  int32_t nbx, nby;
  int32_t gate, pad_;   // gate: skip the solve (done = 3) unless the last attempt was accepted
  // multigrid levels 1 .. ncl, whose 2^l x 2^l aggregates lie inside one
  // assembly tile (l <= kLaplCoarse): their Galerkin operators are coarsened by
  // the assembly kernel from the tile it has just assembled
  MgLevel cl[3];
  int32_t ncl, pad2_;
};

struct PcgArgs {
  const double* cx;
  const double* cy;
  const double* dg;
  const double* r_old;
  double* r_new;
  const double* w_old;
  double* w_new;
  double* s_old;        // s_{i-1} (read) -- zeroed by the init kernel
  double* s_new;
  double* p;
  double* x;
  const double* rhs0;
  double* ph;
  const double* part_in;   // previous kernel's partials (3 per block)
  double* part_out;
  DevScalars* sc;
  Geo G;
  int32_t nbx, nby, it, pad_;
};

// One level of the aggregation multigrid hierarchy: (nx+2)*(ny+2) planes
// with zero ghosts (d ghost = 1) so stencils need no bounds checks.

// Level-0 vectors of the multigrid-PCG in the Chronopoulos-Gear form: the
// finest down kernel applies the CG update before smoothing the new residual,
//   p = u + beta p,  s = w + beta s_old,  x += alpha p,  r = r_old - alpha s,
// and the finest up kernel writes u = B r, w = M u and the partials of r.u
// and u.w that give the next alpha and beta.
struct MgCg {
  double* x;
  double* p;
  const double* u;      // B r of the last cycle (level-0 e)
  double* w;            // M u
  const double* s_old;  // s = M p, double-buffered by iteration parity
  double* s_new;
  const double* r_old;  // residual before the update (F.r receives the new one)
  int32_t on, first;    // first iteration: p = u, s = w
};

struct MgArgs {
  MgLevel F;             // fine level of this kernel
  MgLevel C;             // coarse level
  const double* part_in; // convergence test / CG update: the previous kernel's partials
  int32_t nb_in;
  double* part_out;      // level-0 up: partials of r.u (and u.w or z.q); level-0 CG down: |r/V|^2
  DevScalars* sc;
  Geo G;                 // level-0 metrics (volumes for the residual norm)
  const double* q;       // level-0 up, flexible CG: q = M p of the last update (z.q partials)
  double* part_scale;    // level-0 down, learned solver: partials of sum r_c^2 / d_c
  double oc;
  int32_t it, level0, init, check;  // check: convergence test on part_in before the work
  MgCg cg;
  double* xs;            // level-0 initial test: the iterate x and the previous
  double* xs2;           // correction x2, advanced here (warm_a, warm_b)
  double* xs3;           // correction x3, advanced here (warm_c)
  const double* rhs0;    // level-0 initial test: the right-hand side (zero start)
};

// Kernels whose per-block partial sums every block of the next kernel reads
// run at most kMaxPartials blocks (each covering a contiguous range of tiles),
// so that the cross-kernel reduction stays O(kMaxPartials) per block.
constexpr int kMaxPartials = 1024;
inline int capped_grid(int ntiles) { return ntiles < kMaxPartials ? ntiles : kMaxPartials; }

// most levels a hierarchy may have
constexpr int kMgMaxLevels = 16;
constexpr int kMgBottomMax = 16;   // cells of the coarsest level (dense inverse)
constexpr int kMgLdsLevels = 5;    // levels run by the one-workgroup kernel
constexpr int kMgLdsCells = 4096;  // cells of the finest of those (the others <= 1024)
constexpr int kMgLdsX = 6144;      // LDS doubles for their padded iterates
constexpr int kMgLdsR = 4608;      // LDS doubles for the padded residuals of the others
constexpr int kMgSetupX = 2048;    // LDS doubles per coefficient for the levels below it
// packed FP32 rows: 5 values per cell, the 4 cells of a 2x2 aggregate of the
// finest level and one cell of each other non-coarsest level per thread, 1024
// threads
constexpr int kMgPackSlots = 4 + kMgLdsLevels - 2;
constexpr size_t kMgPackFloats = (size_t)kMgPackSlots * 5 * 1024;

struct MgCycleArgs {
  MgLevel lv[kMgLdsLevels];  // levels f .. L-1, the last one the coarsest
  int32_t nq;                // how many of them
  int32_t off[kMgLdsLevels]; // LDS offset of each level's (nx+2)(ny+2) iterate
  double* inv;               // dense inverse of the coarsest operator (set-up kernel)
  float* pk;                 // packed FP32 operator rows of the other levels (set-up kernel)
  DevScalars* sc;
  double oc;
  const double* part_in;     // check: |r/V|^2 partials of the level-0 CG update
  int32_t nb_in, it, check, pad_;
};

struct MgPcgArgs {
  const double* cx;
  const double* cy;
  const double* dg;
  const double* u;
  const double* p_old;
  double* p_new;
  double* q;
  double* x;
  double* r;
  const double* part_in;
  int32_t nb_in, it;      // `it` is used for its parity only
  double* part_out;
  DevScalars* sc;
  Geo G;
  int32_t nbx, nby, first, flex;  // flex: Polak-Ribiere beta (flexible CG)
  int32_t warm, ls;  // init: start from the x in place (r = rhs - M x); ls: with the
                     // projection's coefficients the warm-start kernels wrote
  const double* x2;  // init: the correction before x
  const double* x3;  // init: the one before x2
  double* lspart;    // projected warm start: nine sums per block
};

constexpr int kTX = 64;  // tile width (one wave64 per tile row)
constexpr int kTY = 4;   // tile height
constexpr int kNT = kTX * kTY;
// The stage and assembly kernels stage a tile plus a one-cell halo: taller
// tiles (64x8, 512 threads) recompute fewer halo cells per cell.
constexpr int kSTY = 8;
constexpr int kSNT = kTX * kSTY;
constexpr int kStageRows = 4;  // rows of the stage kernel's error-sum groups
constexpr int kLaplCoarse = 3;  // log2(kSTY): coarse levels an assembly tile holds whole
static_assert(kSTY == (1 << kLaplCoarse), "assembly tile height");
static_assert(sizeof(LaplArgs::cl) == kLaplCoarse * sizeof(MgLevel), "LaplArgs::cl");
constexpr int kPcgTY = 8;    // PCG tile height: each thread owns kPcgRows cells
constexpr int kPcgRows = kPcgTY / kTY;

// ---- ensemble launches --------------------------------------------------------
// One launch covers the blocks of every ensemble member that takes part.  Block b
// of a launch belongs to member inst[b] as that member's block b - first[member]
// of nb[member] -- the shape of the member's own launch -- so every member
// computes exactly what its own launch computes.
struct EnsMap {
  const int32_t* inst;
  const int32_t* first;
  const int32_t* nb;
  const uint8_t* active;  // per member: takes part in this launch
  int32_t nblocks;        // blocks of the launch
};
// A member's arguments: table[(sel ? sel[member] : member) * stride + off].
struct EnsSel {
  const int32_t* sel;
  int32_t stride, off;
};

void launch_ens_stage(const StageArgs* t, const FsPhys* K, const EnsMap& m, const EnsSel& s, int ty,
                      hipStream_t st);
void launch_ens_circ(const CircArgs* t, const FsPhys* K, const EnsMap& m, const EnsSel& s,
                     hipStream_t st);
void launch_ens_ctrl(const CtrlArgs* t, const EnsMap& m, const EnsSel& s, hipStream_t st);
// op: 0 begin step, 1 accept, 2 reject; one thread per member of mask
void launch_ens_scalars(DevScalars* sc, const uint8_t* mask, int n, int op, hipStream_t st);
void launch_ens_lapl(const LaplArgs* t, const FsPhys* K, const EnsMap& m, const EnsSel& s,
                     hipStream_t st);
void launch_ens_mg_coarsen(const MgArgs* t, const EnsMap& m, const EnsSel& s, hipStream_t st);
void launch_ens_mg_coarse_setup(const MgCycleArgs* t, const EnsMap& m, const EnsSel& s,
                                hipStream_t st);
// cg: the members' level-0 CG form (cg.on in every selected entry)
void launch_ens_mg_down(const MgArgs* t, const EnsMap& m, const EnsSel& s, bool cg,
                        hipStream_t st);
// cg: the level-0 up kernel of the multigrid-PCG (wider halo, w = M u)
void launch_ens_mg_up(const MgArgs* t, const EnsMap& m, const EnsSel& s, bool cg, hipStream_t st);
void launch_ens_mg_coarse_cycle(const MgCycleArgs* t, const EnsMap& m, const EnsSel& s,
                                hipStream_t st);
void launch_ens_mgpcg_init(const MgPcgArgs* t, const EnsMap& m, const EnsSel& s, hipStream_t st);
void launch_ens_mgpcg_warm_ls(const MgPcgArgs* t, const EnsMap& m, const EnsMap& m_one,
                              const EnsSel& s, hipStream_t st);
// finish over map m (the members' capped PCG grids), then the true residual
// with one block per member (map one, partial counts m.nb)
void launch_ens_pcg_finish(const PcgArgs* t, const EnsMap& m, const EnsMap& one, const EnsSel& s,
                           hipStream_t st);

// a, K: device pointers (argument tables kept per handle)
// the stage kernel on nbx x ceil(by / ty) tiles of kTX x ty cells (ty = 4, 8, 16)
int stage_blocks(int nbx, int by, int ty);
void launch_stage(const StageArgs* a, const FsPhys* K, int nbx, int by, int ty, hipStream_t s,
                  int gated = 0);
void launch_circ(const CircArgs& a, const FsPhys& K, hipStream_t s);
void launch_circ_ctrl(const CircArgs& c, const FsPhys& K, const CtrlArgs& a, hipStream_t s);
void launch_accept(DevScalars* sc, hipStream_t s);
void launch_reject(DevScalars* sc, hipStream_t s);
void launch_begin_step(DevScalars* sc, hipStream_t s);
void launch_accept_spec(DevScalars* sc, hipStream_t s);
void launch_publish(const DevScalars* sc, DevScalars* pub, uint32_t* cnt, hipStream_t s);
void launch_lapl_assemble(const LaplArgs* a, const FsPhys* K, int nblocks, hipStream_t s);
void launch_pcg_init(const PcgArgs& a, hipStream_t s);
void launch_pcg_it(const PcgArgs& a, hipStream_t s);
void launch_pcg_finish(const PcgArgs& a, hipStream_t s);
void launch_mg_coarsen(const MgArgs& a, hipStream_t s);
int mg_grid(const MgLevel& F);  // blocks (= partials) of a down/up launch
void launch_mg_down(const MgArgs& a, hipStream_t s);
void launch_mg_up(const MgArgs& a, hipStream_t s);
void launch_mg_coarse_setup(const MgCycleArgs& a, hipStream_t s);
void launch_mg_coarse_cycle(const MgCycleArgs& a, hipStream_t s);
void launch_mgpcg_init(const MgPcgArgs& a, hipStream_t s);
void launch_mgpcg_warm_ls(const MgPcgArgs& a, hipStream_t s);
void launch_mgpcg_pq(const MgPcgArgs& a, hipStream_t s);
void launch_mgpcg_xr(const MgPcgArgs& a, hipStream_t s);
void launch_fill(double* p, double v, size_t n, hipStream_t s);
void launch_copy(const double* a, double* b, size_t n2, hipStream_t s);
void launch_eval_fits(const FsPhys& K, const double* T, double* out, int n, hipStream_t s);
void launch_eval_flux(const FsPhys& K, const double* a, double* out, int n, int dfdph,
                      hipStream_t s);
