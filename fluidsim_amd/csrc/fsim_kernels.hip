// fsim_kernels.hip -- gfx950 kernels of the plasma-transport hot path.
//
//   k_stage      one RK-Merson stage: ptclEqn (ptcl_lib.F90:99-550) fused with
//                rk_step (ptcl_props.F90:71-138) and the embedded error sum;
//                stages 2-5 first make the previous stage's circuit update
//                (circ_step, circ_lib.F90:19-210).  64x8 tiles; densities,
//                potential and per-cell transport properties of the tile plus
//                a one-cell halo are staged in LDS, each face flux is computed
//                once and shared by the two cells it separates.
//   k_circ_ctrl  the circuit update of stage 5 and the dt controller of
//                ptcl_step (ptcl_lib.F90:59-94); k_circ alone serves the
//                ensemble and fs_circ_step.
//   k_lapl       lapl_solve assembly (laplEqn + laplJ, lapl_lib.F90:123-609)
//                into the volume-symmetrised operator -V*A (3 planes), and the
//                multigrid levels whose aggregates lie inside a tile.
//   k_mg_*       the multigrid-preconditioned CG that replaces KSPSolve
//                (lapl_lib.F90:103); k_pcg_* the Jacobi-preconditioned one.
//
// Reductions are deterministic: per-block partials in a fixed tree order,
// which every block of the next kernel sums in the same fixed order.
#include <hip/hip_runtime.h>
#include <string.h>

#include "fsim_internal.h"

namespace {

// stage / assembly tiles (kTX x kSTY cells, kSNT threads) and their halo
constexpr int kHX = kTX + 2, kHY = kSTY + 2, kNH = kHX * kHY;
constexpr int kNFX = (kTX + 1) * kSTY, kNFY = kTX * (kSTY + 1);
constexpr int kNF = kNFX > kNFY ? kNFX : kNFY;
// the assembly stages its tile's levels 0 .. kLaplCoarse (3 planes each) in
// its face buffer s_f[4][kNF]
static_assert(3 * (kSNT + kSNT / 4 + kSNT / 16 + kSNT / 64) <= 4 * kNF, "assembly LDS levels");
constexpr int kNW = kNT / 64;
constexpr int kSNW = kSNT / 64;

// The stage kernel's tile: kTX x TY cells (TY = 4, 8 or 16, chosen per grid by
// timing them), one thread per cell, a one-cell halo, its x and y faces.  Its
// error sums are reduced per group of kStageRows rows (a divisor of every TY).
template <int TY>
struct StageTile {
  static constexpr int kNT = kTX * TY, kNW = kNT / 64, kHY = TY + 2, kNH = kHX * kHY;
  static constexpr int kNFX = (kTX + 1) * TY, kNFY = kTX * (TY + 1);
  static constexpr int kNF = kNFX > kNFY ? kNFX : kNFY;
};

// Device pointers held in argument blocks in device memory are generic
// pointers, and accesses through them compile to flat instructions, whose
// completion every LDS wait also waits for.  These casts make them global
// accesses (all such buffers are device allocations).
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* to_global(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ T gld(const T* p, size_t i) {
  return to_global(p)[i];
}
template <class T>
__device__ __forceinline__ void gst(T* p, size_t i, T v) {
  to_global(p)[i] = v;
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous band of tiles so that halo rows are shared inside one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  int nb8 = (nb / 8) * 8;
  if (b >= nb8) return b;
  return (b % 8) * (nb8 / 8) + b / 8;
}

// Contiguous range [t0, t1) of the ntiles tiles for this block when a capped
// grid covers them (XCD-aware block order; trailing blocks may be empty).
struct TileRange {
  int t0, t1;
};
__device__ __forceinline__ TileRange tile_range(int ntiles, int bid, int nb) {
  const int b = xcd_tile(bid, nb);
  const int per = (ntiles + nb - 1) / nb;
  const int t0 = b * per;
  return {t0, t0 + per < ntiles ? t0 + per : ntiles};
}

// Block-wide sums (s_red: NV * NW doubles) and minimum of NW waves.
template <int NV, int NW = kNW>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* s_red) {
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double a = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    v[k] = a;
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) s_red[k * NW + w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double a = s_red[k * NW];
      for (int q = 1; q < NW; q++) a += s_red[k * NW + q];
      v[k] = a;
    }
  __syncthreads();
}

template <int NW = kNW>
__device__ __forceinline__ double block_min(double v, double* s_red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s_red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 1; q < NW; q++) v = fmin(v, s_red[q]);
  __syncthreads();
  return v;
}

// The source voltage at time t for the waveform C.rf (circ_lib.F90:26-50).
__device__ double vsource(const CircArgs& C, const FsPhys& K, double t) {
  const double Vmax = C.vset, pi = K.pi;
  switch (C.rf) {
    case 1: return Vmax * sin(pi * t);
    case 2: {
      if (t < 2.5e-1) {
        double s = sin(4.0 * pi * t);
        return Vmax * (s * (s * s));
      }
      return 0.0;
    }
    case 3: {
      if (t < 2.3e-1) {
        double s = sin(4.0 * pi * t);
        return Vmax * (s * s);
      }
      if (t < 3.35e-1) return Vmax * (sin(3.81 * pi * (t + 5.85e-2)) * 7.19e-1 + 2.81e-1);
      if (t < 8.2e-1) {
        double s = sin(3.81 * pi * (t + 1.2e-1) / 3.5);
        return Vmax * ((s * (s * s)) * (-4.38e-1));
      }
      return 0.0;
    }
    case 4: {
      double tr = C.t_rise, tw = C.t_width, tf = C.t_fall;
      if (t < tr) return Vmax * (t / tr);
      if (t < tr + tw) return Vmax;
      double te = tr + tw + tf;
      if (t < te) return Vmax * ((te - t) / tf);
      return 0.0;
    }
    default: return Vmax;
  }
}

// The electrode voltage after the circuit update of `stage` (Vd_mi for stages
// 1-4, Vd_pl for 5), circ_lib.F90:158-174, from kv[1..stage].
__device__ __forceinline__ double circ_vd(int stage, double Vd_or, double dt, const double* kv) {
  if (stage == 1) return Vd_or + dt * kv[1] / 3.0;
  if (stage == 2) return Vd_or + dt * (kv[1] + kv[2]) / 6.0;
  if (stage == 3) return Vd_or + dt * (kv[1] + kv[3] * 3.0) / 8.0;
  if (stage == 4) return Vd_or + dt * (kv[1] - kv[3] * 3.0 + kv[4] * 4.0) / 2.0;
  return Vd_or + dt * (kv[1] + kv[4] * 4.0 + kv[5]) / 6.0;
}

// ---------------------------------------------------------------------------
// Electrode boundary fluxes, ptcl_lib.F90:187-213 (left) and :247-272 (right)
struct Flux4 {
  double i, e, t, m;
};

__device__ __forceinline__ Flux4 electrode_left(const FsPhys& K, double E, double phL, double phR,
                                                double te, double mue, double ni, double ne,
                                                double nt, double nm) {
  Flux4 F;
  double a = phL > phR ? 1.0 : 0.0;
  double mut = K.c53 * mue;
  double ve = sqrt(K.ve_num * te / K.ve_den) * K.t0 / K.x0;
  F.i = (1.0 - a) * K.mui * E * ni - K.qvi * ni;
  F.e = -(a * mue * E * ne) - 2.5e-1 * ve * ne - K.gam * F.i;
  F.t = -(a * mut * E * nt) - 1.0 / 3.0 * ve * nt - K.gam * te * F.i;
  F.m = -(K.hvi * nm);
  return F;
}

__device__ __forceinline__ Flux4 electrode_right(const FsPhys& K, double E, double phL,
                                                 double phR, double te, double mue, double ni,
                                                 double ne, double nt, double nm) {
  Flux4 F;
  double a = phR > phL ? 1.0 : 0.0;
  double mut = K.c53 * mue;
  double ve = sqrt(K.ve_num * te / K.ve_den) * K.t0 / K.x0;
  F.i = (1.0 - a) * K.mui * E * ni + K.qvi * ni;
  F.e = -(a * mue * E * ne) + 2.5e-1 * ve * ne - K.gam * F.i;
  F.t = -(a * mut * E * nt) + 1.0 / 3.0 * ve * nt - K.gam * te * F.i;
  F.m = K.hvi * nm;
  return F;
}

// interior face between cells L and R (ptcl_lib.F90:131-161)
// rh = 1/h.
__device__ __forceinline__ Flux4 interior_face(const FsPhys& K, double E, double h, double rh,
                                               double mueL, double mueR, double deL, double deR,
                                               double niL, double niR, double neL, double neR,
                                               double ntL, double ntR, double nmL, double nmR) {
  Flux4 F;
  double mue = 5e-1 * (mueL + mueR);
  double de = 5e-1 * (deL + deR);
  double mut = K.c53 * mue;
  double dt = K.c53 * de;
  F.i = sg_flux_ion(K, E, h, rh, niL, niR);
  sg_flux_et(K, E, h, rh, mue, de, mut, dt, neL, neR, ntL, ntR, F.e, F.t);
  F.m = -(K.Dm * (nmR - nmL) * rh);
  return F;
}

// ---------------------------------------------------------------------------
// one instance per kernel: the LDS arrays declared here then belong to one kernel each
template <int kInstance, int TY>
__device__ __forceinline__ void stage_body(const StageArgs& A, const FsPhys& K, int bid, int nb) {
  using ST = StageTile<TY>;
  __shared__ double s_ni[ST::kNH], s_ne[ST::kNH], s_nt[ST::kNH], s_nm[ST::kNH], s_ph[ST::kNH];
  __shared__ double s_te[ST::kNH], s_mue[ST::kNH], s_de[ST::kNH], s_xa[ST::kNH];
  __shared__ double s_f[4][ST::kNF];
  __shared__ double s_red[8 * ST::kNW];
  __shared__ int s_bad;

  const Geo& G = A.G;
  const int bx = G.bx, by = G.by, P = G.P;
  const int t = xcd_tile(bid, nb);
  const int i0 = (t % A.nbx) * kTX, j0 = (t / A.nbx) * TY;
  const double dt = A.sc->dt;
  if (threadIdx.x == 0) s_bad = 0;
  // The previous stage's circuit update (cc.stage > 0), k_circ's arithmetic in
  // each block that needs it: the electrode currents are summed as k_circ's 256 threads sum
  // them (the other threads add zeros), the update is made after phase 1, and
  // the ghost potential of the electrode rows in this tile's halo is set in
  // LDS from it (k_circ_ctrl writes the final one to memory after stage 5).
  // (only the blocks that use it: those with an electrode ghost column in the
  // halo, and block 0, which records the scalars)
  const int cstage =
      kInstance == 0 && (bid == 0 || i0 == 0 || i0 + kTX >= bx) ? A.cc.stage : 0;
  double cv[2] = {0.0, 0.0};
  if (cstage > 0 && threadIdx.x < kNT)
    for (int j = threadIdx.x; j < A.cc.by; j += kNT) {
      cv[0] += gld(A.cc.dra, j) * gld(A.cc.cur, j);
      cv[1] += gld(A.cc.drc, j) * gld(A.cc.cur, A.cc.by + j);
    }
  __syncthreads();

  // ---- phase 1: stage densities + potential of tile and halo; cell properties
  for (int h = threadIdx.x; h < ST::kNH; h += ST::kNT) {
    int hx = h % kHX, hy = h / kHX;
    int ci = i0 - 1 + hx, cj = j0 - 1 + hy;
    double vi = 0, ve = 0, vt = 0, vm = 0, vp = 0, te = 0, mue = 0, de = 0, xa = 0;
    if (ci <= bx && cj <= by) {
      size_t m = (size_t)(cj + 1) * P + (ci + 1);
      vi = gld(A.rd[0], m);
      ve = gld(A.rd[1], m);
      vt = gld(A.rd[2], m);
      vm = gld(A.rd[3], m);
      vp = gld(A.ph, m);
      if (ci >= 0 && ci < bx && cj >= 0 && cj < by) {
        if (vt >= 0.0 && ve >= 0.0) {
          te = fs_div(vt, ve);  // get_Te, ptcl_props.F90:167-180
        } else {
          te = 1e-8;
          s_bad = 1;
        }
        xa = fit_x_stage(te * K.ph0);
        mue = mue_of_x_folded(K, xa);
        de = de_of_x_folded(K, xa);
      }
    }
    s_ni[h] = vi;
    s_ne[h] = ve;
    s_nt[h] = vt;
    s_nm[h] = vm;
    s_ph[h] = vp;
    s_te[h] = te;
    s_mue[h] = mue;
    s_de[h] = de;
    s_xa[h] = xa;
  }
  // this cell's RK base and earlier slopes, loaded here so that their latency
  // overlaps the flux phases
  const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
  const int ci = i0 + tx, cj = j0 + ty;
  const bool valid = (ci < bx) && (cj < by);
  const size_t mc = (size_t)(cj + 1) * P + (ci + 1);
  const int stage = A.stage;
  double pn3[4] = {0, 0, 0, 0}, pk1[4] = {0, 0, 0, 0}, pk3[4] = {0, 0, 0, 0},
         pk4[4] = {0, 0, 0, 0};
  if (valid && !A.rhs_only) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
      pn3[s] = gld(A.n3[s], mc);
      if (stage >= 2) pk1[s] = gld((const double*)A.k1[s], mc);
      if (stage >= 4) pk3[s] = gld((const double*)A.k3[s], mc);
      if (stage == 5) pk4[s] = gld((const double*)A.k4[s], mc);
    }
  }
  if (cstage > 0) {
    __shared__ double s_vd;
    block_sum<2, ST::kNW>(cv, s_red);
    if (threadIdx.x == 0) {
      const CircArgs& C = A.cc;
      DevScalars* sc = A.sc;
      const int k = cstage;
      const double Vsrc = vsource(C, K, sc->t);
      const bool pin = C.rf != 1 && C.circuit == 0;  // circ_lib.F90:176-177
      double kv[6];
      for (int q = 1; q < k; q++) kv[q] = sc->kv[q];
      // Vd_mi as the previous update left it (stored only from update 2 on:
      // the other blocks still read the stored value during update 1)
      const double vprev = k == 1 ? sc->Vd_mi : (pin ? Vsrc : circ_vd(k - 1, sc->Vd_or, dt, kv));
      kv[k] = -(cv[1] - (Vsrc - vprev) / C.res) / K.Cap;
      const double vd = pin ? Vsrc : circ_vd(k, sc->Vd_or, dt, kv);
      s_vd = C.oned ? Vsrc : vd;  // one dimension: circ_lib.F90:199-202
      if (bid == 0) {
        sc->kv[k] = kv[k];
        if (k > 1) sc->Vd_mi = vd;
        if (pin) sc->Vd_pl = Vsrc;
        sc->Vsrc = Vsrc;
        sc->Ida = cv[0];
        sc->Idc = cv[1];
      }
    }
    __syncthreads();
    const double vd = s_vd;
    for (int q = threadIdx.x; q < 2 * ST::kHY; q += ST::kNT) {  // circ_lib.F90:180-198
      const int hy = q % ST::kHY, cj = j0 - 1 + hy;
      if (cj < 0 || cj >= by) continue;
      if (q < ST::kHY) {
        if (i0 == 0 && gld(A.cc.gl, cj)) s_ph[hy * kHX] = vd / 2.0;
      } else if (i0 + kTX >= bx && gld(A.cc.gr, cj)) {
        s_ph[hy * kHX + bx - i0 + 1] = vd / (-2.0);
      }
    }
  }
  __syncthreads();

  // ---- phase 2: x faces (left face of cells i0..i0+TX, rows j0..j0+TY-1)
  for (int f = threadIdx.x; f < ST::kNFX; f += ST::kNT) {
    int fx = f % (kTX + 1), fy = f / (kTX + 1);
    int fi = i0 + fx, fj = j0 + fy;
    Flux4 F = {0, 0, 0, 0};
    if (fj < by && fi <= bx) {
      int hL = (fy + 1) * kHX + fx, hR = hL + 1;
      double h = gld(G.dx, fi), rh = gld(G.rdx, fi);
      double E = -(s_ph[hR] - s_ph[hL]) * rh;
      if (fi == 0) {
        if (gld(G.ltype, fj) == -2)
          F = electrode_left(K, E, s_ph[hL], s_ph[hR], s_te[hR], s_mue[hR], s_ni[hR], s_ne[hR],
                             s_nt[hR], s_nm[hR]);
      } else if (fi == bx) {
        if (gld(G.rtype, fj) == 2)
          F = electrode_right(K, E, s_ph[hL], s_ph[hR], s_te[hL], s_mue[hL], s_ni[hL],
                              s_ne[hL], s_nt[hL], s_nm[hL]);
      } else {
        F = interior_face(K, E, h, rh, s_mue[hL], s_mue[hR], s_de[hL], s_de[hR], s_ni[hL],
                          s_ni[hR], s_ne[hL], s_ne[hR], s_nt[hL], s_nt[hR], s_nm[hL], s_nm[hR]);
        if (fi == 1) gst(A.cur, fj, F.i - F.e);            // Ida face (circ_lib.F90:87-118)
        if (fi == bx - 1) gst(A.cur, by + fj, F.i - F.e);  // Idc face (circ_lib.F90:120-151)
      }
    }
    s_f[0][f] = F.i;
    s_f[1][f] = F.e;
    s_f[2][f] = F.t;
    s_f[3][f] = F.m;
  }
  __syncthreads();

  // ---- phase 3a: x divergence per cell (kept in registers)
  const int hc = (ty + 1) * kHX + tx + 1;
  double dfx[4] = {0, 0, 0, 0}, feX = 0, exA = 0, mue2 = 0;
  int typx = 0;
  if (valid) {
    typx = ci == 0 ? gld(G.ltype, cj) : (ci == bx - 1 ? gld(G.rtype, cj) : 0);
    int fL = ty * (kTX + 1) + tx, fR = fL + 1;
    double rdlx = gld(G.rdlx, ci);
#pragma unroll
    for (int s = 0; s < 4; s++) dfx[s] = (s_f[s][fR] - s_f[s][fL]) * rdlx;
    double exL, exR;
    if (typx == -1) {
      exL = 0.0;
      exR = -(s_ph[hc + 1] - s_ph[hc]) * gld(G.rdx, ci + 1);
    } else if (typx == 1) {
      exL = -(s_ph[hc] - s_ph[hc - 1]) * gld(G.rdx, ci);
      exR = 0.0;
    } else {
      exL = -(s_ph[hc] - s_ph[hc - 1]) * gld(G.rdx, ci);
      exR = -(s_ph[hc + 1] - s_ph[hc]) * gld(G.rdx, ci + 1);
    }
    feX = (s_f[1][fR] + s_f[1][fL]) / 2.0;
    exA = (exL + exR) / 2.0;
    // mue(2) left by the x branch (used by t_m on the top row)
    if (typx == 2) mue2 = s_mue[hc];
    else if (typx == 1) mue2 = 0.0;
    else mue2 = 5e-1 * (s_mue[hc] + s_mue[hc + 1]);
  }
  __syncthreads();

  // ---- phase 2b: y faces (bottom face of rows j0..j0+TY, columns i0..i0+TX-1)
  for (int f = threadIdx.x; f < ST::kNFY; f += ST::kNT) {
    int fx = f % kTX, fy = f / kTX;
    int fci = i0 + fx, fj = j0 + fy;
    Flux4 F = {0, 0, 0, 0};
    if (fci < bx && fj > 0 && fj < by) {  // faces on y = 0 and y = w carry no flux
      int hL = fy * kHX + fx + 1, hR = hL + kHX;
      double h = gld(G.dy, fj), rh = gld(G.rdy, fj);
      double E = -(s_ph[hR] - s_ph[hL]) * rh;
      F = interior_face(K, E, h, rh, s_mue[hL], s_mue[hR], s_de[hL], s_de[hR], s_ni[hL],
                        s_ni[hR], s_ne[hL], s_ne[hR], s_nt[hL], s_nt[hR], s_nm[hL], s_nm[hR]);
    } else if (G.rwall && fci < bx && fj == by) {
      // the dielectric wall above the top row: E = -sig_mi, fluxes into a zero
      // density, energy transport by get_mut / get_Dt (ptcl_lib.F90:292-294,
      // 393-409)
      const int hL = fy * kHX + fx + 1;
      const double h = gld(G.dy, by), rh = gld(G.rdy, by);
      const double E = -gld(G.sig + (size_t)kSigMi * (bx + 2), fci + 1);
      const double xa = s_xa[hL];
      F.i = sg_flux_ion(K, E, h, rh, s_ni[hL], 0.0);
      F.e = sg_flux(K, E, h, rh, -1.0, s_mue[hL], s_de[hL], s_ne[hL], 0.0);
      F.t = sg_flux(K, E, h, rh, -1.0, mut_of_x(K, xa), dt_of_x(K, xa), s_nt[hL], 0.0);
      F.m = K.Dm * s_nm[hL] / h;
    }
    s_f[0][f] = F.i;
    s_f[1][f] = F.e;
    s_f[2][f] = F.t;
    s_f[3][f] = F.m;
  }
  __syncthreads();

  // ---- phase 3b: y divergence, sources, stage slope, RK-Merson update
  double acc[4] = {0, 0, 0, 0};
  double tmin = 1e300;
  if (valid) {
    int fB = ty * kTX + tx, fT = fB + kTX;
    double rdly = gld(G.rdly, cj);
    double dfy[4];
    if (G.cyl) {
      double rj = gld(G.r, cj + 1), rp = gld(G.r, cj + 2), rm = gld(G.r, cj);
      double w = rdly / rj;
#pragma unroll
      for (int s = 0; s < 4; s++)
        dfy[s] = (s_f[s][fT] * (rp + rj) / 2.0 - s_f[s][fB] * (rj + rm) / 2.0) * w;
    } else {
#pragma unroll
      for (int s = 0; s < 4; s++) dfy[s] = (s_f[s][fT] - s_f[s][fB]) * rdly;
    }
    // (one dimension: no y branch at all, typy 2 is never assigned otherwise;
    // Ey and the y fluxes stay zero, mue(2) is the x branch's)
    const int typy = G.oned ? 2 : (cj == 0 ? -1 : (cj == by - 1 ? (G.rwall ? 3 : 1) : 0));
    double eyB = 0.0, eyT = 0.0;
    if (typy == -1) {
      eyT = -(s_ph[hc + kHX] - s_ph[hc]) * gld(G.rdy, cj + 1);
    } else if (typy == 1) {
      eyB = -(s_ph[hc] - s_ph[hc - kHX]) * gld(G.rdy, cj);
    } else if (typy == 3) {  // ptcl_lib.F90:292-294
      eyB = -(s_ph[hc] - s_ph[hc - kHX]) * gld(G.rdy, cj);
      eyT = -gld(G.sig + (size_t)kSigMi * (bx + 2), ci + 1);
    } else if (typy == 0) {
      eyB = -(s_ph[hc] - s_ph[hc - kHX]) * gld(G.rdy, cj);
      eyT = -(s_ph[hc + kHX] - s_ph[hc]) * gld(G.rdy, cj + 1);
    }
    double feY = (s_f[1][fT] + s_f[1][fB]) / 2.0;
    double eyA = (eyB + eyT) / 2.0;
    if (typy <= 0) mue2 = 5e-1 * (s_mue[hc] + s_mue[hc + kHX]);
    else if (typy == 3) mue2 = s_mue[hc];  // get_mue(Te(2)) of the wall branch (:398)

    // rates and sources, ptcl_lib.F90:483-514 (Em = 0: its +0 term is dropped)
    double te = s_te[hc];
    Rates R = cell_rates(K, te * K.ph0, s_xa[hc]);
    double ne = s_ne[hc], ni = s_ni[hc], nm = s_nm[hc], ninf = K.ninf;
    double nm2 = nm * nm;
    double term_sie = R.k_ir * ninf * ne - R.beta * ni * ne + R.k_si * nm * ne + K.k_mp * nm2;
    double term_sm = R.k_ex * ninf * ne - R.k_si * nm * ne - R.k_sc * nm * ne - K.k_r * nm * ne -
                     K.kmp2 * nm2 - K.k2q_ninf * nm - K.k3q_ninf2 * nm;
    double st1 = -(feX * exA + feY * eyA);
    double st2 = -(ne * R.nu * K.me_mi * (te * (1.0 / 1.5) - K.kTe));
    double st3 = -(K.H_ir * R.k_ir * ninf * ne) - K.H_si * R.k_si * nm * ne -
                 K.H_ex * R.k_ex * ninf * ne;
    double kv[4];
    kv[0] = -dfx[0] - dfy[0] + term_sie;
    kv[1] = -dfx[1] - dfy[1] + term_sie;
    kv[2] = -dfx[2] - dfy[2] + st1 + st2 + st3;
    kv[3] = -dfx[3] - dfy[3] + term_sm;
    if (isnan(kv[0]) || isnan(kv[1]) || isnan(kv[2]) || isnan(kv[3])) s_bad = 2;

    const size_t m = mc;
    if (A.kdbg[0]) {
#pragma unroll
      for (int s = 0; s < 4; s++) gst(A.kdbg[s], m, kv[s]);
    }
    if (!A.rhs_only) {
#pragma unroll
      for (int s = 0; s < 4; s++) {
        double n3 = pn3[s], kk = kv[s], nw;
        double nmin = s == 2 ? K.nmin_t : K.n_zero;
        if (stage == 1) {
          nw = n3 + kk * dt * (1.0 / 3.0);
          gst(A.k1[s], m, kk);
        } else if (stage == 2) {
          nw = n3 + dt * (pk1[s] + kk) * (1.0 / 6.0);
        } else if (stage == 3) {
          nw = n3 + dt * (pk1[s] + kk * 3.0) / 8.0;
          gst(A.k3[s], m, kk);
        } else if (stage == 4) {
          nw = n3 + dt * (pk1[s] - pk3[s] * 3.0 + kk * 4.0) / 2.0;
          gst(A.k4[s], m, kk);
        } else {
          double k1 = pk1[s], k3 = pk3[s], k4 = pk4[s];
          nw = n3 + dt * (k1 + k4 * 4.0 + kk) * (1.0 / 6.0);
          double e = fabs(dt * (k1 * (2.0 / 30.0) - k3 * (3.0 / 10.0) + k4 * (4.0 / 15.0) -
                                kk * (1.0 / 30.0)));
          double q = fs_div(e, 1e-3 + 1e-3 * fabs(n3));
          acc[s] = q * q;
        }
        gst(A.wr[s], m, nw > nmin ? nw : nmin);
      }
      if (stage == 5) tmin = fs_rcp(mue2 * ne + K.mui * ni);
      if (typy == 3) {
        // sfc_step (sfc_lib.F90:20-58) for this column, from the stage
        // densities and sig_mi before its update, with the reference's
        // argument binding: ptcl_step passes (ni, ne, nt) to the dummies
        // (ne, ni, nt) (ptcl_lib.F90:52), so Te = nt / ni, the "ion" flux
        // carries ne and the "electron" flux ni; Merson's tableau as the
        // circuit's
        double* sg = G.sig;
        const size_t W = (size_t)bx + 2, ic = (size_t)ci + 1;
        const double E = -gld(sg + kSigMi * W, ic);
        const double h = gld(G.dy, by), rh = gld(G.rdy, by);
        const double tes = fs_div(s_nt[hc], ni);
        const double xs = fit_x_stage(tes * K.ph0);
        const double ks = sg_flux_ion(K, E, h, rh, ne, 0.0) -
                          sg_flux(K, E, h, rh, -1.0, mue_of_x_folded(K, xs),
                                  de_of_x_folded(K, xs), ni, 0.0);
        double kq[6];
        for (int q = 1; q < stage; q++) kq[q] = gld(sg + (kSigK + q) * W, ic);
        kq[stage] = ks;
        gst(sg + (kSigK + stage) * W, ic, ks);
        const double v = circ_vd(stage, gld(sg + kSigOr * W, ic), dt, kq);
        gst(sg + (stage == 5 ? kSigPl : kSigMi) * W, ic, v);
      }
    }
  }
  // The error sums and t_m of each group of kStageRows rows of kTX cells, by
  // one reduction tree whatever the tile height (a shuffle tree within each
  // row's wave, then the group's rows in order), so that the controller's
  // inputs do not depend on the tile shape; at (row group, tile column).
  if (!A.rhs_only && stage == 5) {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc[s] += __shfl_down(acc[s], off, 64);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tmin = fmin(tmin, __shfl_down(tmin, off, 64));
    const int w = threadIdx.x >> 6;  // = ty: one wave per tile row
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int s = 0; s < 4; s++) s_red[w * 5 + s] = acc[s];
      s_red[w * 5 + 4] = tmin;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && w % kStageRows == 0 && cj < by) {
      double v[5];
#pragma unroll
      for (int k = 0; k < 5; k++) v[k] = s_red[w * 5 + k];
      for (int r = 1; r < kStageRows; r++) {
#pragma unroll
        for (int s = 0; s < 4; s++) v[s] += s_red[(w + r) * 5 + s];
        v[4] = fmin(v[4], s_red[(w + r) * 5 + 4]);
      }
      const size_t p = ((size_t)(cj / kStageRows) * A.nbx + (size_t)(i0 / kTX)) * 8;
#pragma unroll
      for (int k = 0; k < 5; k++) gst(A.part, p + k, v[k]);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_bad)
    atomicMin(&A.sc->status, s_bad == 1 ? FS_E_NEGDENS : FS_E_NAN);
}

// Arguments from device memory: kept there per handle, they are read with
// scalar loads where used (passed by value, the compiler keeps the whole
// argument block live in SGPRs and spills it to VGPR lanes).
// gated: issued before the host has read the last pass's outcome; runs only
// if k_accept_spec accepted that pass.
template <int TY>
__global__ void __launch_bounds__(kTX * TY, 4) k_stage(const StageArgs* __restrict__ A,
                                                   const FsPhys* __restrict__ K, int gated) {
  if (gated && A->sc->spec == 0) return;
  stage_body<0, TY>(*A, *K, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------


template <int kInstance>
__device__ __forceinline__ void circ_body(const CircArgs& C, const FsPhys& K) {
  __shared__ double s_red[8 * kNW];
  __shared__ double s_vd;
  DevScalars* sc = C.sc;
  double v[2] = {0.0, 0.0};
  for (int j = threadIdx.x; j < C.by; j += kNT) {
    v[0] += C.dra[j] * C.cur[j];
    v[1] += C.drc[j] * C.cur[C.by + j];
  }
  block_sum<2>(v, s_red);
  if (threadIdx.x == 0) {
    const int stage = C.stage;
    double Vsrc = vsource(C, K, sc->t);
    double Ida = v[0], Idc = v[1];
    double* kv = sc->kv;
    kv[stage] = -(Idc - (Vsrc - sc->Vd_mi) / C.res) / K.Cap;
    const double vd = circ_vd(stage, sc->Vd_or, sc->dt, kv);
    if (stage <= 4) sc->Vd_mi = vd;
    else sc->Vd_pl = vd;
    if (C.rf != 1 && C.circuit == 0) {  // circ_lib.F90:176-177
      sc->Vd_pl = Vsrc;
      sc->Vd_mi = Vsrc;
    }
    sc->Vsrc = Vsrc;
    sc->Ida = Ida;
    sc->Idc = Idc;
    // (one dimension: the source voltage, circ_lib.F90:199-202)
    s_vd = C.oned ? Vsrc : (stage == 5 ? sc->Vd_pl : sc->Vd_mi);
  }
  __syncthreads();
  const double vd = s_vd;
  for (int j = threadIdx.x; j < C.by; j += kNT) {  // circ_lib.F90:180-198
    size_t row = (size_t)(j + 1) * C.P;
    if (C.gl[j]) C.ph[row] = vd / 2.0;
    if (C.gr[j]) C.ph[row + C.bx + 1] = vd / (-2.0);
  }
}

__global__ void __launch_bounds__(kNT) k_circ(CircArgs C, FsPhys K) { circ_body<0>(C, K); }

// dt controller: ptcl_lib.F90:59-94 (plus replayed decisions)
template <int kInstance>
__device__ __forceinline__ void ctrl_body(const CtrlArgs& A) {
  __shared__ double s_red[8 * kNW];
  double v[4] = {0, 0, 0, 0};
  double tm = 1e300;
  for (int b = threadIdx.x; b < A.nblocks; b += kNT) {
    const double* p = A.part + (size_t)b * 8;
    v[0] += p[0];
    v[1] += p[1];
    v[2] += p[2];
    v[3] += p[3];
    tm = fmin(tm, p[4]);
  }
  block_sum<4>(v, s_red);
  tm = block_min(tm, s_red);
  if (threadIdx.x != 0) return;
  DevScalars* sc = A.sc;
  double err_cur = sqrt(v[0] + v[1] + v[3] + v[2]);  // (ni + ne + nm + nt)**0.5
  double scfac = 8e-1 * pow(err_cur, -7e-1 / 4.0) * pow(sc->err_prev, 4e-1 / 4.0);
  scfac = fmin(2.5, fmax(3e-1, scfac));
  sc->dt_stage = sc->dt;
  double dtn = scfac * sc->dt;
  int accept = err_cur <= 1.0;
  if (sc->replay_n > 0) {
    if (sc->replay_idx < sc->replay_n) {
      const double* rec = A.replay + (size_t)sc->replay_idx * 10;
      dtn = rec[4];
      accept = rec[5] != 0.0;
      sc->replay_idx++;
    } else {
      sc->status = min(sc->status, FS_E_REPLAY);
    }
  }
  sc->dt = dtn;
  sc->err_cur = err_cur;
  for (int k = 0; k < 4; k++) sc->err4[k] = v[k];
  sc->t_m = fmin(sc->t_m, tm);
  sc->accept = accept;
  sc->attempts++;
  if (dtn <= 1.1e-12) sc->status = min(sc->status, FS_E_DTMIN);
  if (accept) {
    sc->err_prev = err_cur;
  } else {
    sc->Vd_pl = sc->Vd_or;
    sc->Vd_mi = sc->Vd_or;
  }
}

// The wall charge's rotation (accept: sig_or <- sig_mi <- sig_pl, main.F90:
// 100-101) and restore (reject, restart: sig_pl = sig_mi = sig_or, main.F90:
// 117-118, ptcl_lib.F90:90-91); every thread of the launch takes a share.
__device__ __forceinline__ void sig_accept(const DevScalars* sc) {
  const int n = sc->sig_n;
  double* s = sc->sig;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    s[kSigOr * n + i] = s[kSigMi * n + i];
    s[kSigMi * n + i] = s[kSigPl * n + i];
  }
}
__device__ __forceinline__ void sig_restore(const DevScalars* sc) {
  const int n = sc->sig_n;
  double* s = sc->sig;
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    s[kSigPl * n + i] = s[kSigMi * n + i] = s[kSigOr * n + i];
}

// After the controller: a restart restores the wall charge too (every thread
// of the launch copies a share, on the decision thread 0 made)
__device__ __forceinline__ void ctrl_sig_restore(const CtrlArgs& A) {
  __shared__ int s_restart;
  if (threadIdx.x == 0) s_restart = A.sc->accept == 0;
  __syncthreads();
  if (s_restart) sig_restore(A.sc);
}

// The circuit update of stage 5 and the controller in one launch.
__global__ void __launch_bounds__(kNT) k_circ_ctrl(CircArgs C, FsPhys K, CtrlArgs A) {
  circ_body<0>(C, K);
  __syncthreads();
  ctrl_body<0>(A);
  ctrl_sig_restore(A);
}

__device__ __forceinline__ void begin_step_body(DevScalars* sc) {
  sc->t_m = (double)1e9f;  // main.F90:74
  sc->attempts = 0;
}

__global__ void k_begin_step(DevScalars* sc) { begin_step_body(sc); }

// Copies the scalars to host-mapped memory and then, after a system-scope
// fence, the count of publications so far (kept on the device in *cnt, so that
// a publication captured in a graph counts too): the host polls for the count
// it expects instead of synchronising the stream, then reads the copy (one
// wave, vector stores).
__global__ void __launch_bounds__(64) k_publish(const DevScalars* sc, DevScalars* pub,
                                                uint32_t* cnt) {
  constexpr int kWords = sizeof(DevScalars) / sizeof(uint64_t);
  static_assert(sizeof(DevScalars) % sizeof(uint64_t) == 0, "DevScalars is copied as 64-bit words");
  const uint32_t seq = *cnt + 1u;
  const uint64_t* src = reinterpret_cast<const uint64_t*>(sc);
  uint64_t* dst = reinterpret_cast<uint64_t*>(pub);
  for (int k = threadIdx.x; k < kWords; k += 64) dst[k] = src[k];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    *cnt = seq;
    *reinterpret_cast<volatile uint32_t*>(&pub->pub_seq) = seq;
    __threadfence_system();
  }
}

__device__ __forceinline__ void accept_body(DevScalars* sc) {  // main.F90:83-101
  sc->ts = sc->ts + 1.0;
  sc->t = sc->t + sc->dt;
  sc->Vd_or = sc->Vd_mi;
  sc->Vd_mi = sc->Vd_pl;
}

__global__ void __launch_bounds__(kNT) k_accept(DevScalars* sc) {
  sig_accept(sc);
  if (threadIdx.x == 0) accept_body(sc);
}

// The accept of a pass and the begin of the next, issued before the host has
// read the pass's outcome: applied when the host accepts the pass on the
// scalars this kernel reads (no error, the attempt accepted, its solve
// converged in the chunk just run), and recorded in spec for the gated stage
// kernel that follows.
__global__ void __launch_bounds__(kNT) k_accept_spec(DevScalars* sc) {
  // (the flags read here are not written by this kernel)
  const int go = sc->status >= 0 && sc->accept == 1 && sc->done == 1;
  if (go) sig_accept(sc);
  __syncthreads();
  if (threadIdx.x != 0) return;
  sc->spec = go;
  if (go) {
    accept_body(sc);
    begin_step_body(sc);
  }
}

__global__ void __launch_bounds__(kNT) k_reject(DevScalars* sc) {  // main.F90:103-121
  sig_restore(sc);
  if (threadIdx.x != 0) return;
  sc->dt = sc->dt / 2.0;
  sc->Vd_pl = sc->Vd_or;
  sc->Vd_mi = sc->Vd_or;
}

__device__ __forceinline__ size_t mg_at(const MgLevel& L, int i, int j) {
  return (size_t)(j + 1) * L.P + (i + 1);
}

// Galerkin coarsening (P = piecewise-constant 2x2 aggregation, R = P^T) of
// aggregate (I, J) of an nx x ny level into an NX x NY level; at(i, j) is the
// index of fine cell (i, j) in the coefficient arrays.
struct Coarse3 {
  double d, cx, cy;
};

template <class At>
__device__ __forceinline__ Coarse3 coarsen_agg(const double* fd, const double* fcx,
                                               const double* fcy, At at, int nx, int ny, int NX,
                                               int NY, int I, int J) {
  int i0 = 2 * I, j0 = 2 * J;
  bool hx = i0 + 1 < nx, hy = j0 + 1 < ny;
  double d = fd[at(i0, j0)];
  double internal = 0.0;
  if (hx) {
    d += fd[at(i0 + 1, j0)];
    internal += fcx[at(i0, j0)];
  }
  if (hy) {
    d += fd[at(i0, j0 + 1)];
    internal += fcy[at(i0, j0)];
  }
  if (hx && hy) {
    d += fd[at(i0 + 1, j0 + 1)];
    internal += fcx[at(i0, j0 + 1)] + fcy[at(i0 + 1, j0)];
  }
  Coarse3 c;
  c.d = d - 2.0 * internal;
  // faces leaving the aggregate to the right / upwards (zero at the domain edge)
  int ir = hx ? i0 + 1 : i0, ju = hy ? j0 + 1 : j0;
  c.cx = 0.0;
  c.cy = 0.0;
  if (I + 1 < NX) {
    c.cx = fcx[at(ir, j0)];
    if (hy) c.cx += fcx[at(ir, j0 + 1)];
  }
  if (J + 1 < NY) {
    c.cy = fcy[at(i0, ju)];
    if (hx) c.cy += fcy[at(i0 + 1, ju)];
  }
  return c;
}


// ---------------------------------------------------------------------------
// Poisson assembly.  Per x/y face: ion and electron SG fluxes (laplEqn) and
// P = (mu/h)*dFdPh (laplJ) from one shared exponential per species.
struct FaceJ {
  double fi, fe, pi, pe;
};

__device__ __forceinline__ FaceJ lapl_face(const FsPhys& K, double E, double h, double mueL,
                                           double mueR, double deL, double deR, double niL,
                                           double niR, double neL, double neR) {
  FaceJ F;
  double mue = 5e-1 * (mueL + mueR);
  double de = 5e-1 * (deL + deR);
  double dfi, dfe;
  sg_flux_dfdph(K, E, h, 1.0, K.mui, K.Di, niL, niR, F.fi, dfi);
  sg_flux_dfdph(K, E, h, -1.0, mue, de, neL, neR, F.fe, dfe);
  F.pi = K.mui / h * dfi;  // dfxi_mi (lapl_lib.F90:405); dfxi_pl = -pi (:440)
  F.pe = mue / h * dfe;    // dfxe_pl (lapl_lib.F90:438); dfxe_mi = -pe (:403)
  return F;
}

template <int kInstance>
__device__ __forceinline__ void lapl_body(const LaplArgs& A, const FsPhys& K, int bid, int nb) {
  __shared__ double s_ni[kNH], s_ne[kNH], s_ph[kNH], s_te[kNH], s_mue[kNH], s_de[kNH];
  __shared__ double s_f[4][kNF];
  __shared__ double s_red[8 * kSNW];
  __shared__ int s_bad;
  if (A.gate && A.sc->accept == 0) {  // the attempt was rejected: no solve
    if (bid == 0 && threadIdx.x == 0) A.sc->done = 3;
    return;
  }
  const Geo& G = A.G;
  const int bx = G.bx, by = G.by, P = G.P;
  const int t = xcd_tile(bid, nb);
  const int i0 = (t % A.nbx) * kTX, j0 = (t / A.nbx) * kSTY;
  const double dt = A.sc->dt;
  if (threadIdx.x == 0) s_bad = 0;
  __syncthreads();

  for (int h = threadIdx.x; h < kNH; h += kSNT) {
    int hx = h % kHX, hy = h / kHX;
    int ci = i0 - 1 + hx, cj = j0 - 1 + hy;
    double vi = 0, ve = 0, vp = 0, te = 0, mue = 0, de = 0;
    if (ci <= bx && cj <= by) {
      size_t m = (size_t)(cj + 1) * P + (ci + 1);
      vi = A.n[0][m];
      ve = A.n[1][m];
      double vt = A.n[2][m];
      vp = A.ph[m];
      if (ci >= 0 && ci < bx && cj >= 0 && cj < by) {
        if (vt >= 0.0 && ve >= 0.0) {
          te = vt / ve;
        } else {
          te = 1e-8;
          s_bad = 1;
        }
        double xa = fit_x(te * K.ph0);
        mue = mue_of_x(K, xa);
        de = de_of_x(K, xa);
      }
    }
    s_ni[h] = vi;
    s_ne[h] = ve;
    s_ph[h] = vp;
    s_te[h] = te;
    s_mue[h] = mue;
    s_de[h] = de;
  }
  __syncthreads();

  // x faces: fluxes (laplEqn) and Jacobian terms (laplJ)
  for (int f = threadIdx.x; f < kNFX; f += kSNT) {
    int fx = f % (kTX + 1), fy = f / (kTX + 1);
    int fi = i0 + fx, cj = j0 + fy;
    FaceJ F = {0, 0, 0, 0};
    if (cj < by && fi <= bx) {
      int hL = (fy + 1) * kHX + fx, hR = hL + 1;
      double h = G.dx[fi];
      double E = -(s_ph[hR] - s_ph[hL]) / h;
      if (fi == 0) {
        if (G.ltype[cj] == -2) {  // lapl_lib.F90:186-203 and :410-421
          double a = s_ph[hL] > s_ph[hR] ? 1.0 : 0.0;
          double te = s_te[hR], mue = s_mue[hR], ni = s_ni[hR], ne = s_ne[hR];
          double ve = sqrt(K.ve_num * te / K.ve_den) * K.t0 / K.x0;
          F.fi = (1.0 - a) * K.mui * E * ni - K.qvi * ni;
          F.fe = -(a * mue * E * ne) - 2.5e-1 * ve * ne - K.gam * F.fi;
          F.pi = (1.0 - a) * K.mui * ni / h;          // dfxi_mi
          F.pe = -(a * mue * ne / h) - K.gam * F.pi;  // dfxe_mi
        }
      } else if (fi == bx) {
        if (gld(G.rtype, cj) == 2) {  // lapl_lib.F90:224-240 and :445-456
          double a = s_ph[hR] > s_ph[hL] ? 1.0 : 0.0;
          double te = s_te[hL], mue = s_mue[hL], ni = s_ni[hL], ne = s_ne[hL];
          double ve = sqrt(K.ve_num * te / K.ve_den) * K.t0 / K.x0;
          F.fi = (1.0 - a) * K.mui * E * ni + K.qvi * ni;
          F.fe = -(a * mue * E * ne) + 2.5e-1 * ve * ne - K.gam * F.fi;
          F.pi = (a - 1.0) * K.mui * ni / h;       // dfxi_pl
          F.pe = a * mue * ne / h - K.gam * F.pi;  // dfxe_pl
        }
      } else {
        F = lapl_face(K, E, h, s_mue[hL], s_mue[hR], s_de[hL], s_de[hR], s_ni[hL], s_ni[hR],
                      s_ne[hL], s_ne[hR]);
      }
      if (isnan(F.pi) || isnan(F.pe)) s_bad = 3;
    }
    s_f[0][f] = F.fi;
    s_f[1][f] = F.fe;
    s_f[2][f] = F.pi;
    s_f[3][f] = F.pe;
  }
  __syncthreads();

  const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
  const int ci = i0 + tx, cj = j0 + ty;
  const bool valid = (ci < bx) && (cj < by);
  const int hc = (ty + 1) * kHX + tx + 1;
  double o_d = 0.0, o_cx = 0.0, o_cy = 0.0;  // this cell's operator row, as stored
  double dflxi_x = 0, dflxe_x = 0, dfdx = 0, Cx = 0, aW = 0, aE = 0, gR = 0;
  int typx = 0;
  if (valid) {
    typx = ci == 0 ? G.ltype[cj] : (ci == bx - 1 ? G.rtype[cj] : 0);
    int fL = ty * (kTX + 1) + tx, fR = fL + 1;
    double dlx = G.dlx[ci], hL = G.dx[ci], hR = G.dx[ci + 1];
    dflxi_x = (s_f[0][fR] - s_f[0][fL]) / dlx;
    dflxe_x = (s_f[1][fR] - s_f[1][fL]) / dlx;
    double exL, exR;
    if (typx == -1) {
      exL = 0.0;
      exR = -(s_ph[hc + 1] - s_ph[hc]) / hR;
    } else if (typx == 1) {
      exL = -(s_ph[hc] - s_ph[hc - 1]) / hL;
      exR = 0.0;
    } else {
      exL = -(s_ph[hc] - s_ph[hc - 1]) / hL;
      exR = -(s_ph[hc + 1] - s_ph[hc]) / hR;
    }
    dfdx = -(exR - exL) / dlx;
    double dfxi_mi, dfxe_mi, dfxi_pl, dfxe_pl;
    if (typx >= 0) {
      dfxi_mi = s_f[2][fL];
      dfxe_mi = -s_f[3][fL];
      double gL = 1.0 / hL + dt * (dfxi_mi - dfxe_mi);
      aW = gL / dlx;
    } else if (typx == -2) {
      dfxi_mi = s_f[2][fL];
      dfxe_mi = s_f[3][fL];
    } else {
      dfxi_mi = 0.0;
      dfxe_mi = 0.0;
    }
    if (typx <= 0) {
      dfxi_pl = -s_f[2][fR];
      dfxe_pl = s_f[3][fR];
      gR = 1.0 / hR - dt * (dfxi_pl - dfxe_pl);
      aE = gR / dlx;
    } else if (typx == 2) {
      dfxi_pl = s_f[2][fR];
      dfxe_pl = s_f[3][fR];
    } else {
      dfxi_pl = 0.0;
      dfxe_pl = 0.0;
    }
    Cx = dt * (dfxi_pl - dfxi_mi - dfxe_pl + dfxe_mi) / dlx;  // lapl_lib.F90:568-575
    if (typx == -1 || typx == 1) Cx = Cx - 1.0 / hL / dlx;
    else Cx = Cx - 1.0 / dlx * (1.0 / hR + 1.0 / hL);
  }
  __syncthreads();

  // y faces (those on y = 0 and y = w carry nothing: type_y = -1 / 1)
  for (int f = threadIdx.x; f < kNFY; f += kSNT) {
    int fx = f % kTX, fy = f / kTX;
    int fci = i0 + fx, fj = j0 + fy;
    FaceJ F = {0, 0, 0, 0};
    if (fci < bx && fj > 0 && fj < by) {
      int hL = fy * kHX + fx + 1, hR = hL + kHX;
      double h = G.dy[fj];
      double E = -(s_ph[hR] - s_ph[hL]) / h;
      F = lapl_face(K, E, h, s_mue[hL], s_mue[hR], s_de[hL], s_de[hR], s_ni[hL], s_ni[hR],
                    s_ne[hL], s_ne[hR]);
      if (isnan(F.pi) || isnan(F.pe)) s_bad = 3;
    } else if (G.rwall && fci < bx && fj == by) {
      // the dielectric wall: E = -sig_pl, fluxes into a zero density with the
      // top cell's own mobility (lapl_lib.F90:263-265, 324-335); its Jacobian
      // terms are zero (lapl_lib.F90:548-561)
      const int hL = fy * kHX + fx + 1;
      const double h = G.dy[by];
      const double E = -G.sig[(size_t)kSigPl * (bx + 2) + fci + 1];
      double dfi, dfe;
      sg_flux_dfdph(K, E, h, 1.0, K.mui, K.Di, s_ni[hL], 0.0, F.fi, dfi);
      sg_flux_dfdph(K, E, h, -1.0, s_mue[hL], s_de[hL], s_ne[hL], 0.0, F.fe, dfe);
    }
    s_f[0][f] = F.fi;
    s_f[1][f] = F.fe;
    s_f[2][f] = F.pi;
    s_f[3][f] = F.pe;
  }
  __syncthreads();

  if (valid) {
    int fB = ty * kTX + tx, fT = fB + kTX;
    double dly = G.dly[cj], hB = G.dy[cj], hT = G.dy[cj + 1];
    double rj = G.r[cj + 1], rp = G.r[cj + 2], rm = G.r[cj];
    int typy = cj == 0 ? -1 : (cj == by - 1 ? (G.rwall ? 3 : 1) : 0);
    // one dimension: the y part is left out (lapl_lib.F90:255, 462, 577); dfdy,
    // dflxi_y and dflxe_y are zero, there is no y column
    double dfdy = 0.0, dflxi_y = 0.0, dflxe_y = 0.0;
    double dfyi_mi = 0, dfye_mi = 0, dfyi_pl = 0, dfye_pl = 0, aS = 0, aN = 0, gT = 0;
    double C = Cx;
    if (!G.oned) {
      double eyB, eyT;
      if (typy == -1) {
        eyB = 0.0;
        eyT = -(s_ph[hc + kHX] - s_ph[hc]) / hT;
      } else if (typy == 1) {
        eyB = -(s_ph[hc] - s_ph[hc - kHX]) / hB;
        eyT = 0.0;
      } else if (typy == 3) {  // Ey(2) = -sig (lapl_lib.F90:263-265)
        eyB = -(s_ph[hc] - s_ph[hc - kHX]) / hB;
        eyT = -G.sig[(size_t)kSigPl * (bx + 2) + ci + 1];
      } else {
        eyB = -(s_ph[hc] - s_ph[hc - kHX]) / hB;
        eyT = -(s_ph[hc + kHX] - s_ph[hc]) / hT;
      }
      double fiB = s_f[0][fB], fiT = s_f[0][fT], feB = s_f[1][fB], feT = s_f[1][fT];
      if (G.cyl) {
        dfdy = -(eyT * (rp + rj) / 2.0 - eyB * (rj + rm) / 2.0) / dly / rj;
        dflxi_y = (fiT * (rp + rj) / 2.0 - fiB * (rj + rm) / 2.0) / dly / rj;
        dflxe_y = (feT * (rp + rj) / 2.0 - feB * (rj + rm) / 2.0) / dly / rj;
      } else {
        dfdy = -(eyT - eyB) / dly;
        dflxi_y = (fiT - fiB) / dly;
        dflxe_y = (feT - feB) / dly;
      }

      // laplJ y part and diagonal, lapl_lib.F90:463-607
      if (typy >= 0) {
        dfyi_mi = s_f[2][fB];
        dfye_mi = -s_f[3][fB];
        if (G.cyl)
          aS = 5e-1 * (rj + rm) / hB / dly / rj + dt * (dfyi_mi - dfye_mi) * (rj + rm) / 2.0 / dly / rj;
        else
          aS = 1.0 / hB / dly + dt * (dfyi_mi - dfye_mi) / dly;
      }
      if (typy <= 0) {
        dfyi_pl = -s_f[2][fT];
        dfye_pl = s_f[3][fT];
        gT = 1.0 / hT - dt * (dfyi_pl - dfye_pl);
        if (G.cyl)
          aN = 5e-1 * (rp + rj) / hT / dly / rj - dt * (dfyi_pl - dfye_pl) * (rp + rj) / 2.0 / dly / rj;
        else
          aN = 1.0 / hT / dly - dt * (dfyi_pl - dfye_pl) / dly;
      }
      if (G.cyl) {
        C = C + dt * ((dfyi_pl - dfye_pl) * (rp + rj) / 2.0 - (dfyi_mi - dfye_mi) * (rj + rm) / 2.0) /
                    dly / rj;
        if (typy == -1) C = C - 1.0 / hB / dly * (rp + rj) / 2.0 / rj;
        else if (typy == 1) C = C - 1.0 / hB / dly * (rj + rm) / 2.0 / rj;
        else C = C - 5e-1 * ((rp + rj) / hT + (rj + rm) / hB) / dly / rj;
      } else {
        C = C + dt * (dfyi_pl - dfyi_mi - dfye_pl + dfye_mi) / dly;
        if (typy == -1 || typy == 1) C = C - 1.0 / hB / dly;
        else C = C - 1.0 / dly * (1.0 / hT + 1.0 / hB);
      }
    }
    // laplEqn residual, lapl_lib.F90:374-375
    size_t m = (size_t)(cj + 1) * P + (ci + 1);
    double ni = s_ni[hc], ne = s_ne[hc];
    double b = dfdx + dfdy + ni - ne - dt * (dflxi_x + dflxi_y - dflxe_x - dflxe_y);
    // symmetrised storage: row (i,j) times V = dlx*dly[*r]
    double dlx = G.dlx[ci];
    double rw = G.cyl ? rj : 1.0;
    double V = dlx * dly * rw;
    double rhs = V * b;  // -V*A*delta = V*b
    double dgv = -V * C;
    o_d = dgv;
    o_cx = typx <= 0 ? gR * dly * rw : 0.0;
    o_cy = !G.oned && typy <= 0 ? gT * dlx * (G.cyl ? (rp + rj) / 2.0 : 1.0) : 0.0;
    A.rhs0[m] = rhs;
    A.r[m] = rhs;
    A.dg[m] = o_d;
    A.cx[m] = o_cx;
    A.cy[m] = o_cy;
    if (A.bdbg) {
      size_t n = (size_t)cj * bx + ci;
      A.bdbg[n] = b;
      double* a5 = A.adbg + 5 * n;
      a5[0] = aW;
      a5[1] = aE;
      a5[2] = aS;
      a5[3] = aN;
      a5[4] = C;
    }
    if (isnan(b)) s_bad = 4;
  }
  if (A.ncl > 0) {
    // Multigrid levels 1 .. ncl over this tile's 2^q x 2^q aggregates: the
    // Galerkin coarsening of k_mg_coarsen (coarsen_agg, same operands and
    // order), from the rows just assembled, staged level by level in LDS in the
    // face buffer (free once every cell has read its faces).
    double* s_lv = &s_f[0][0];
    __syncthreads();
    s_lv[threadIdx.x] = o_d;
    s_lv[kSNT + threadIdx.x] = o_cx;
    s_lv[2 * kSNT + threadIdx.x] = o_cy;
    __syncthreads();
    int off = 0, fnx = bx, fny = by;
    for (int q = 1; q <= A.ncl; q++) {
      const int wf = kTX >> (q - 1), nf = wf * (kSTY >> (q - 1));  // the tile on level q - 1
      const int wc = kTX >> q, nc = wc * (kSTY >> q);              // ... and on level q
      const int fi0 = i0 >> (q - 1), fj0 = j0 >> (q - 1);
      const MgLevel& C = A.cl[q - 1];
      const double* fd = s_lv + off;
      double* cd = s_lv + off + 3 * nf;
      const int t = threadIdx.x;
      if (t < nc) {
        const int I = (fi0 >> 1) + t % wc, J = (fj0 >> 1) + t / wc;
        Coarse3 c = {0.0, 0.0, 0.0};
        if (I < C.nx && J < C.ny) {
          c = coarsen_agg(fd, fd + nf, fd + 2 * nf,
                          [&](int i, int j) { return (j - fj0) * wf + (i - fi0); }, fnx, fny,
                          C.nx, C.ny, I, J);
          const size_t m = mg_at(C, I, J);
          C.d[m] = c.d;
          C.cx[m] = c.cx;
          C.cy[m] = c.cy;
        }
        cd[t] = c.d;
        cd[nc + t] = c.cx;
        cd[2 * nc + t] = c.cy;
      }
      __syncthreads();
      off += 3 * nf;
      fnx = C.nx;
      fny = C.ny;
    }
  }
  if (threadIdx.x == 0 && s_bad == 1) atomicMin(&A.sc->status, FS_E_NEGDENS);
  if (threadIdx.x == 0 && s_bad == 3) atomicMin(&A.sc->status, FS_E_NANJ);
  if (bid == 0 && threadIdx.x == 0) {
    A.sc->done = 0;
    A.sc->iters = 0;
  }
}

__global__ void __launch_bounds__(kSNT, 4) k_lapl(const LaplArgs* __restrict__ A,
                                              const FsPhys* __restrict__ K) {
  lapl_body<0>(*A, *K, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// PCG on M = -V*A (SPD), Chronopoulos-Gear form: one kernel per iteration and
// one fused reduction (gamma = r.u, delta = w.u, rho = |r/V|^2).  No kernel
// communicates across its own blocks: every block reduces the previous
// kernel's per-block partials itself (fixed order, so all blocks obtain the
// same scalars), which needs no device-scope fence inside a kernel.
//
//   p_i = u_i + b_i p_{i-1}      s_i = w_i + b_i s_{i-1}      (s = M p)
//   x_{i+1} = x_i + a_i p_i      r_{i+1} = r_i - a_i s_i
//   u_{i+1} = r_{i+1} / d        w_{i+1} = M u_{i+1}
// u at the four neighbours is recomputed from r_i, w_i, s_{i-1} of the
// neighbour, so w_{i+1} needs no second pass.  r, w, s are double-buffered.

// Sum the previous kernel's 3 partials per block, identically in every block.
// Partial sums of the previous kernel: this thread's share (loads only, so
// that a kernel can issue them together with its own loads) ...
// (nb <= kMaxPartials).  In two parts, so that a kernel can issue other
// loads between them: every load is issued before the first sum (threads past
// nb load the first partial and do not add it).
struct Part3 {
  double a[kMaxPartials / kNT][3];
};
__device__ __forceinline__ void part3_load(const double* part, int nb, Part3& P) {
#pragma unroll
  for (int q = 0; q < kMaxPartials / kNT; q++) {
    const int b = threadIdx.x + kNT * q;
    const size_t o = b < nb ? 3 * (size_t)b : 0;
    P.a[q][0] = gld(part, o);
    P.a[q][1] = gld(part, o + 1);
    P.a[q][2] = gld(part, o + 2);
  }
}
__device__ __forceinline__ void part3_sum(const Part3& P, int nb, double (&v)[3]) {
  __builtin_amdgcn_sched_barrier(0);  // keep the sums (and their waits) here
  v[0] = v[1] = v[2] = 0.0;
#pragma unroll
  for (int q = 0; q < kMaxPartials / kNT; q++)
    if (threadIdx.x + kNT * q < nb) {
      v[0] += P.a[q][0];
      v[1] += P.a[q][1];
      v[2] += P.a[q][2];
    }
}
__device__ __forceinline__ void load_partials3(const double* part, int nb, double (&v)[3]) {
  Part3 P;
  part3_load(part, nb, P);
  part3_sum(P, nb, v);
}

// ... and the block-wide total, identical in every block.
__device__ __forceinline__ void finish_partials3(double (&v)[3], double* s_red, double (&tot)[3]) {
  block_sum<3>(v, s_red);
  __shared__ double s_tot[3];
  if (threadIdx.x == 0) {
    s_tot[0] = v[0];
    s_tot[1] = v[1];
    s_tot[2] = v[2];
  }
  __syncthreads();
  tot[0] = s_tot[0];
  tot[1] = s_tot[1];
  tot[2] = s_tot[2];
}

__device__ __forceinline__ void reduce_partials3(const double* part, int nb, double* s_red,
                                                 double (&tot)[3]) {
  double v[3];
  load_partials3(part, nb, v);
  finish_partials3(v, s_red, tot);
}

__device__ __forceinline__ void write_partials3(double (&v)[3], double* part, double* s_red,
                                                int bid) {
  block_sum<3>(v, s_red);
  if (threadIdx.x == 0) {
    part[3 * bid] = v[0];
    part[3 * bid + 1] = v[1];
    part[3 * bid + 2] = v[2];
  }
}

// w0 = M u0 with u0 = r0/d, and the partials of gamma0, delta0, rho0.
__global__ void __launch_bounds__(kNT) k_pcg_init(PcgArgs A) {
  __shared__ double s_red[8 * kNW];
  const Geo& G = A.G;
  const int P = G.P;
  const TileRange tr = tile_range(A.nbx * A.nby, blockIdx.x, gridDim.x);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = tr.t0; t < tr.t1; t++) {
    for (int k = 0; k < kPcgRows; k++) {
      const int ci = (t % A.nbx) * kTX + (int)(threadIdx.x % kTX);
      const int cj = (t / A.nbx) * kPcgTY + (int)(threadIdx.x / kTX) + k * (kPcgTY / kPcgRows);
      if (ci < G.bx && cj < G.by) {
        size_t m = (size_t)(cj + 1) * P + (ci + 1);
        const double *r = A.r_old, *d = A.dg;
        double uc = r[m] / d[m];
        double w = d[m] * uc - A.cx[m - 1] * (r[m - 1] / d[m - 1]) - A.cx[m] * (r[m + 1] / d[m + 1]) -
                   A.cy[m - P] * (r[m - P] / d[m - P]) - A.cy[m] * (r[m + P] / d[m + P]);
        A.w_new[m] = w;
        A.p[m] = 0.0;
        A.s_old[m] = 0.0;
        A.x[m] = 0.0;
        double V = G.dlx[ci] * G.dly[cj] * (G.cyl ? G.r[cj + 1] : 1.0);
        double rv = r[m] / V;
        acc[0] += r[m] * uc;
        acc[1] += w * uc;
        acc[2] += rv * rv;
      }
    }
  }
  write_partials3(acc, A.part_out, s_red, blockIdx.x);
}

__global__ void __launch_bounds__(kNT) k_pcg_it(PcgArgs A) {
  __shared__ double s_red[8 * kNW];
  DevScalars* sc = A.sc;
  if (sc->done) return;
  const int it = A.it;
  double tot[3];
  reduce_partials3(A.part_in, gridDim.x, s_red, tot);
  const double gam = tot[0], del = tot[1], rho = tot[2];
  const double bnorm2 = it == 0 ? rho : sc->bnorm2;
  const double res = bnorm2 > 0 ? sqrt(rho / bnorm2) : 0.0;
  double alpha, beta;
  if (it == 0) {
    beta = 0.0;
    alpha = gam / del;
  } else {
    beta = gam / sc->cg_gam[(it + 1) & 1];
    alpha = gam / (del - beta * gam / sc->cg_alpha[(it + 1) & 1]);
  }
  const bool conv = (it > 0 && res <= sc->tol) || bnorm2 == 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (it == 0) sc->bnorm2 = rho;
    sc->cg_gam[it & 1] = gam;
    sc->cg_alpha[it & 1] = alpha;
    sc->iters = it;
    sc->resid = res;
    if (conv) sc->done = 1;
    else if (it >= sc->maxit) sc->done = 2;
  }
  if (conv || it >= sc->maxit) return;

  const Geo& G = A.G;
  const int P = G.P;
  const TileRange tr = tile_range(A.nbx * A.nby, blockIdx.x, gridDim.x);
  const double *r = A.r_old, *w = A.w_old, *so = A.s_old, *d = A.dg;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = tr.t0; t < tr.t1; t++) {
    for (int k = 0; k < kPcgRows; k++) {
      const int ci = (t % A.nbx) * kTX + (int)(threadIdx.x % kTX);
      const int cj = (t / A.nbx) * kPcgTY + (int)(threadIdx.x / kTX) + k * (kPcgTY / kPcgRows);
      if (ci < G.bx && cj < G.by) {
        size_t m = (size_t)(cj + 1) * P + (ci + 1);
        // u_{i+1} of this cell and its neighbours (zero outside: r = s = w = 0, d = 1)
        auto unext = [&](size_t q) { return (r[q] - alpha * (w[q] + beta * so[q])) / d[q]; };
        double sc_ = w[m] + beta * so[m];
        double pc = r[m] / d[m] + beta * A.p[m];
        double rn = r[m] - alpha * sc_;
        double un = rn / d[m];
        double wn = d[m] * un - A.cx[m - 1] * unext(m - 1) - A.cx[m] * unext(m + 1) -
                    A.cy[m - P] * unext(m - P) - A.cy[m] * unext(m + P);
        A.p[m] = pc;
        A.x[m] = A.x[m] + alpha * pc;
        A.s_new[m] = sc_;
        A.r_new[m] = rn;
        A.w_new[m] = wn;
        double V = G.dlx[ci] * G.dly[cj] * (G.cyl ? G.r[cj + 1] : 1.0);
        double rv = rn / V;
        acc[0] += rn * un;
        acc[1] += wn * un;
        acc[2] += rv * rv;
      }
    }
  }
  write_partials3(acc, A.part_out, s_red, blockIdx.x);
}

// After the solve: ph += delta (lapl_lib.F90:112-116) when converged, and the
// partials of the true residual |(rhs0 - M x)/V|^2.
template <int kInstance>
__device__ __forceinline__ void pcg_finish_body(const PcgArgs& A, int bid, int nb) {
  __shared__ double s_red[8 * kNW];
  DevScalars* sc = A.sc;
  const Geo& G = A.G;
  const int P = G.P;
  const TileRange tr = tile_range(A.nbx * A.nby, bid, nb);
  const bool ok = sc->done == 1;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = tr.t0; t < tr.t1; t++) {
    for (int k = 0; k < kPcgRows; k++) {
      const int ci = (t % A.nbx) * kTX + (int)(threadIdx.x % kTX);
      const int cj = (t / A.nbx) * kPcgTY + (int)(threadIdx.x / kTX) + k * (kPcgTY / kPcgRows);
      if (ci < G.bx && cj < G.by) {
        size_t m = (size_t)(cj + 1) * P + (ci + 1);
        const double* x = A.x;
        double mx = A.dg[m] * x[m] - A.cx[m - 1] * x[m - 1] - A.cx[m] * x[m + 1] -
                    A.cy[m - P] * x[m - P] - A.cy[m] * x[m + P];
        double V = G.dlx[ci] * G.dly[cj] * (G.cyl ? G.r[cj + 1] : 1.0);
        double rr = (A.rhs0[m] - mx) / V;
        acc[0] += rr * rr;
        if (ok) A.ph[m] = A.ph[m] + x[m];
      }
    }
  }
  write_partials3(acc, A.part_out, s_red, bid);
}

__global__ void __launch_bounds__(kNT) k_pcg_finish(PcgArgs A) { pcg_finish_body<0>(A, blockIdx.x, gridDim.x); }

// Single block: true residual from the finish partials.
template <int kInstance>
__device__ __forceinline__ void pcg_tres_body(const PcgArgs& A, int nb) {
  __shared__ double s_red[8 * kNW];
  double tot[3];
  reduce_partials3(A.part_in, nb, s_red, tot);
  if (threadIdx.x == 0) {
    DevScalars* sc = A.sc;
    sc->tres = sc->bnorm2 > 0 ? sqrt(tot[0] / sc->bnorm2) : sqrt(tot[0]);
  }
}

__global__ void __launch_bounds__(kNT) k_pcg_tres(PcgArgs A, int nb) { pcg_tres_body<0>(A, nb); }

// ---------------------------------------------------------------------------
// Aggregation multigrid V(2,2)-cycle used as the PCG preconditioner.
//
// Level l+1 aggregates 2x2 cells of level l.  With piecewise-constant
// prolongation P and restriction P^T the Galerkin operator P^T M P keeps the
// 5-point face storage: a coarse face coupling is the sum of the fine faces
// that cross it; the coarse diagonal is the sum of the four fine diagonals
// minus twice the four internal faces.  One cycle on a level, from a zero
// guess, with S = two red-black Gauss-Seidel sweeps (red, black, red, black)
// and S^T the same sweeps in reverse colour order:
//   x1 = S r,  r_c = P^T (r - M x1)                                  [down]
//   x2 = x1 + oc P e_c,  e = x2 + S^T (r - M x2)                     [up]
// The post-smoother is the adjoint of the pre-smoother, so the cycle is a
// symmetric positive definite operator and plain CG remains valid.  Levels
// above 1024 cells are swept in tiles staged in LDS with a halo of four cells
// (one per half-sweep), one down and one up kernel per level; the remaining
// levels run inside one workgroup with their iterates in LDS, and the
// coarsest (<= 64 cells) is solved with a dense inverse.


// A kernel argument is loaded where it is first used; a kernel that first
// tests a flag in device memory would then load its arguments one dependent
// round trip after another behind that test.  hold() takes values as scalar
// operands of an empty asm statement, so that their loads are all issued
// before it (together with the flag's).
__device__ __forceinline__ void hold(const void* p) { asm volatile("" ::"s"(p)); }
__device__ __forceinline__ void hold(int v) { asm volatile("" ::"s"(v)); }
__device__ __forceinline__ void hold(double v) { asm volatile("" ::"s"(v)); }
__device__ __forceinline__ void hold_level(const MgLevel& L) {
  hold(L.d);
  hold(L.cx);
  hold(L.cy);
  hold(L.r);
  hold(L.x1);
  hold(L.e);
  hold(L.nx);
  hold(L.ny);
  hold(L.P);
}
// the arguments of the tile kernels (kernel-argument copies only: the
// ensemble kernels read theirs from a table)
__device__ __forceinline__ void hold_mg_args(const MgArgs& A, bool cg) {
  hold_level(A.F);
  hold_level(A.C);
  hold(A.part_in);
  hold(A.nb_in);
  hold(A.part_out);
  hold(A.part_scale);
  hold(A.q);
  hold(A.oc);
  hold(A.check);
  hold(A.it);
  hold(A.level0);
  hold(A.cg.on);
  hold(A.cg.w);
  hold(A.xs);
  hold(A.xs2);
  hold(A.xs3);
  if (cg) {
    hold(A.cg.x);
    hold(A.cg.p);
    hold(A.cg.u);
    hold(A.cg.w);
    hold(A.cg.s_old);
    hold(A.cg.s_new);
    hold(A.cg.r_old);
    hold(A.cg.first);
    hold(A.G.dlx);
    hold(A.G.dly);
    hold(A.G.r);
    hold(A.G.cyl);
  }
}

// The Galerkin coefficients of coarse cell (I, J) from fine planes d, cx, cy
// (padded, row pitch P, nx x ny cells); coarse size NX x NY.
__device__ __forceinline__ Coarse3 coarsen_cell(const double* fd, const double* fcx,
                                                const double* fcy, int P, int nx, int ny, int NX,
                                                int NY, int I, int J) {
  return coarsen_agg(fd, fcx, fcy, [&](int i, int j) { return (size_t)(j + 1) * P + (i + 1); },
                     nx, ny, NX, NY, I, J);
}

template <int kInstance>
__device__ __forceinline__ void mg_coarsen_body(const MgArgs& A, int bid, int nb) {
  const MgLevel &F = A.F, &C = A.C;
  int n = bid * blockDim.x + threadIdx.x;
  if (n >= C.nx * C.ny) return;
  int I = n % C.nx, J = n / C.nx;
  const Coarse3 c = coarsen_cell(F.d, F.cx, F.cy, F.P, F.nx, F.ny, C.nx, C.ny, I, J);
  size_t m = mg_at(C, I, J);
  C.d[m] = c.d;
  C.cx[m] = c.cx;
  C.cy[m] = c.cy;
}

__global__ void __launch_bounds__(256) k_mg_coarsen(MgArgs A) { mg_coarsen_body<0>(A, blockIdx.x, gridDim.x); }

constexpr int kGsTX = 32, kGsTY = 16;  // tile
constexpr int kGsH = 4;                 // halo of the down kernels (four half-sweeps)
static_assert((kGsTX / 2) * kGsTY == kNT, "one thread per horizontal cell pair");

// A level's equation on one tile plus a halo of H cells, staged in LDS:
// right-hand side, diagonal and its inverse (the sweeps multiply by it),
// right/upper couplings and the iterate.  Cells outside the level carry r = 0,
// d = 1 and no couplings, so they stay zero.
template <int H>
struct GsTile {
  static constexpr int W = kGsTX + 2 * H, R = kGsTY + 2 * H;
  double x[R][W];
  double r[R][W];
  double d[R][W];
  double dinv[R][W];
  double cx[R][W];
  double cy[R][W];
};

__device__ __forceinline__ int gs_tiles_x(const MgLevel& F) { return (F.nx + kGsTX - 1) / kGsTX; }

// Stage r, d, cx, cy of the tile at (i0, j0) and its halo; with kUp also the
// iterate x1 + oc * e_c(parent) of the up kernel.  In two parts: gs_load issues
// every load of the tile into registers (cells outside the level load the
// padded plane's first element and discard it), gs_store writes LDS, so the
// staging costs one memory round trip and a kernel can issue other loads
// between the two.
template <int H, bool kUp>
struct GsRegs {
  static constexpr int NS = (GsTile<H>::R * GsTile<H>::W + kNT - 1) / kNT;
  double r[NS], d[NS], cx[NS], cy[NS], x[NS], e[NS];
  bool in[NS];
};
template <int H, bool kUp>
__device__ __forceinline__ void gs_load(const MgLevel& F, int i0, int j0, const MgLevel& C,
                                        GsRegs<H, kUp>& g) {
  constexpr int W = GsTile<H>::W, R = GsTile<H>::R;
#pragma unroll
  for (int q = 0; q < g.NS; q++) {
    const int k = threadIdx.x + kNT * q;
    const int b = k / W, a = k % W;
    const int i = i0 - H + a, j = j0 - H + b;
    g.in[q] = k < R * W && i >= 0 && i < F.nx && j >= 0 && j < F.ny;
    const size_t m = g.in[q] ? mg_at(F, i, j) : 0;
    g.r[q] = gld(F.r, m);
    g.d[q] = gld(F.d, m);
    g.cx[q] = gld(F.cx, m);
    g.cy[q] = gld(F.cy, m);
    if (kUp) {
      g.x[q] = gld(F.x1, m);
      g.e[q] = gld(C.e, g.in[q] ? mg_at(C, i >> 1, j >> 1) : 0);
    }
  }
}
template <int H, bool kUp>
__device__ __forceinline__ void gs_store(const GsRegs<H, kUp>& g, GsTile<H>& s, double oc) {
  constexpr int W = GsTile<H>::W, R = GsTile<H>::R;
#pragma unroll
  for (int q = 0; q < g.NS; q++) {
    const int k = threadIdx.x + kNT * q;
    if (k < R * W) {
      const int b = k / W, a = k % W;
      const bool in = g.in[q];
      const double dq = in ? g.d[q] : 1.0;
      s.r[b][a] = in ? g.r[q] : 0.0;
      s.d[b][a] = dq;
      s.dinv[b][a] = 1.0 / dq;
      s.cx[b][a] = in ? g.cx[q] : 0.0;
      s.cy[b][a] = in ? g.cy[q] : 0.0;
      if (kUp) s.x[b][a] = in ? g.x[q] + oc * g.e[q] : 0.0;
    }
  }
}
template <int H, bool kUp>
__device__ __forceinline__ void gs_stage(const MgLevel& F, GsTile<H>& s, int i0, int j0,
                                         const MgLevel& C, double oc) {
  GsRegs<H, kUp> g;
  gs_load<H, kUp>(F, i0, j0, C, g);
  gs_store<H, kUp>(g, s, oc);
}

// One Gauss-Seidel half-sweep over the cells of colour c ((i + j) & 1 == c)
// within h <= H - 1 cells of the tile at (i0, j0), entirely in LDS.  A thread
// updates up to kGsHalfIt cells.  Batched (kBatch): cells of one colour do not
// read each other, so every cell's operands are read before any new value is
// written, and the LDS reads of all of a thread's cells are in flight
// together; that holds their values in registers, which the ensemble's
// kernels, whose grids fill the chip, cannot spare (three waves per SIMD need
// at most 168 VGPRs), so they update one cell after the other.
template <int H>
constexpr int kGsHalfIt =
    ((kGsTY + 2 * (H - 1)) * ((kGsTX + 2 * (H - 1)) / 2) + kNT - 1) / kNT;
template <bool kBatch, int H>
__device__ __forceinline__ void gs_half(GsTile<H>& s, int i0, int j0, int c, int h) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
  const int w2 = (kGsTX + 2 * h) / 2, rows = kGsTY + 2 * h, a0 = H - h;
  if constexpr (!kBatch) {
    for (int k = threadIdx.x; k < rows * w2; k += kNT) {
      const int b = a0 + k / w2;
      const int a = a0 + 2 * (k % w2) + ((c + i0 - H + a0 + j0 - H + b) & 1);
      s.x[b][a] = (s.r[b][a] + s.cx[b][a - 1] * s.x[b][a - 1] + s.cx[b][a] * s.x[b][a + 1] +
                   s.cy[b - 1][a] * s.x[b - 1][a] + s.cy[b][a] * s.x[b + 1][a]) *
                  s.dinv[b][a];
    }
    return;
  }
  constexpr int NI = kGsHalfIt<H>;
  double v[NI];
  int at[NI];
#pragma unroll
  for (int q = 0; q < NI; q++) {
    const int k = threadIdx.x + kNT * q;
    at[q] = -1;
    if (k < rows * w2) {
      const int b = a0 + k / w2;
      const int a = a0 + 2 * (k % w2) + ((c + i0 - H + a0 + j0 - H + b) & 1);
      at[q] = b * GsTile<H>::W + a;
      v[q] = (s.r[b][a] + s.cx[b][a - 1] * s.x[b][a - 1] + s.cx[b][a] * s.x[b][a + 1] +
              s.cy[b - 1][a] * s.x[b - 1][a] + s.cy[b][a] * s.x[b + 1][a]) *
             s.dinv[b][a];
    }
  }
#pragma unroll
  for (int q = 0; q < NI; q++)
    if (at[q] >= 0) (&s.x[0][0])[at[q]] = v[q];
}

template <int H>
__device__ __forceinline__ double gs_mx(const GsTile<H>& s, int b, int a) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
  return s.d[b][a] * s.x[b][a] - s.cx[b][a - 1] * s.x[b][a - 1] - s.cx[b][a] * s.x[b][a + 1] -
         s.cy[b - 1][a] * s.x[b - 1][a] - s.cy[b][a] * s.x[b + 1][a];
}

__device__ __forceinline__ int gs_ntiles(const MgLevel& F) {
  return gs_tiles_x(F) * ((F.ny + kGsTY - 1) / kGsTY);
}

// The solver state a convergence test reads, loaded at the start of the
// kernel (with its arguments) rather than after the reduction that needs it.
struct CheckState {
  double tol, bnorm2;
  int32_t itc_prev, maxit;
};
__device__ __forceinline__ CheckState check_state(const DevScalars* sc, int it) {
  return {sc->tol, sc->bnorm2, sc->itc[(it + 1) & 1], sc->itmax};
}
__device__ __forceinline__ void hold(const CheckState& c) {
  asm volatile("" ::"s"(c.tol), "s"(c.bnorm2), "s"(c.itc_prev), "s"(c.maxit));
}

// Convergence test of the multigrid-PCG on rho = |r/V|^2: of the right-hand
// side (init) or after the update of iteration `it`.  Every block decides
// identically; `writer` (one thread) records the count, the residual and the
// outcome.  True when the solve stops here.
__device__ __forceinline__ bool mg_check(DevScalars* sc, const CheckState& cs, double rho,
                                         double b2, int init, int it, bool writer) {
  if (init) {  // b2 = |rhs/V|^2; rho < b2 after a warm start
    const double res = b2 > 0.0 ? sqrt(rho / b2) : 0.0;
    const bool conv = b2 == 0.0 || res <= cs.tol;
    if (writer) {
      sc->bnorm2 = b2;
      sc->resid = res;
      sc->iters = 0;
      sc->itc[0] = 0;
      sc->itc[1] = 0;
      if (conv) sc->done = 1;
    }
    return conv;
  }
  // iteration count kept in two slots by parity: the other slot was written
  // by the previous iteration, the writer fills this one
  const int cur = cs.itc_prev + 1;
  const double res = sqrt(rho / cs.bnorm2);
  const bool conv = res <= cs.tol;
  const bool stop = conv || cur >= cs.maxit;
  if (writer) {
    sc->itc[it & 1] = cur;
    sc->iters = cur;
    sc->resid = res;
    if (conv) sc->done = 1;
    else if (stop) sc->done = 2;
  }
  return stop;
}

// Level-0 CG vectors of the staged cells a thread holds in the down kernel,
// and the volume of its own cells.
constexpr int kGsSlots = (GsTile<kGsH>::R * GsTile<kGsH>::W + kNT - 1) / kNT;
struct CgCells {
  double r[kGsSlots], w[kGsSlots], s[kGsSlots], u[kGsSlots], p[kGsSlots], x[kGsSlots];
  double vol[kGsSlots];
};

// Loads of one tile for the CG update: the operator into LDS, the vectors
// into registers (u, p, x and the volume only for the tile's own cells).  As
// gs_load / gs_store: cg_load issues every load (cells outside the level load
// the padded planes' first element and discard it), cg_store writes LDS.
struct CgRegs {
  double d[kGsSlots], cx[kGsSlots], cy[kGsSlots], vx[kGsSlots], vy[kGsSlots], vr[kGsSlots];
  bool in[kGsSlots], own[kGsSlots];
};
__device__ __forceinline__ void cg_load(const MgArgs& A, int i0, int j0, CgCells& c, CgRegs& g) {
  constexpr int W = GsTile<kGsH>::W, R = GsTile<kGsH>::R;
  const MgLevel& F = A.F;
  const Geo& G = A.G;
#pragma unroll
  for (int e = 0; e < kGsSlots; e++) {
    const int k = threadIdx.x + kNT * e;
    const int b = k / W, a = k % W;
    const int i = i0 - kGsH + a, j = j0 - kGsH + b;
    const bool in = k < R * W && i >= 0 && i < F.nx && j >= 0 && j < F.ny;
    const bool own = in && a >= kGsH && a < kGsH + kGsTX && b >= kGsH && b < kGsH + kGsTY;
    g.in[e] = in;
    g.own[e] = own;
    const size_t m = in ? mg_at(F, i, j) : 0;
    const size_t mo = own ? m : 0;
    g.d[e] = gld(F.d, m);
    g.cx[e] = gld(F.cx, m);
    g.cy[e] = gld(F.cy, m);
    c.r[e] = gld((const double*)A.cg.r_old, m);
    c.w[e] = gld((const double*)A.cg.w, m);
    c.s[e] = A.cg.first ? 0.0 : gld((const double*)A.cg.s_old, m);
    c.u[e] = gld((const double*)A.cg.u, mo);
    c.p[e] = A.cg.first ? 0.0 : gld((const double*)A.cg.p, mo);
    c.x[e] = gld((const double*)A.cg.x, mo);
    g.vx[e] = gld(G.dlx, own ? i : 0);
    g.vy[e] = gld(G.dly, own ? j : 0);
    g.vr[e] = G.cyl ? gld(G.r, own ? j + 1 : 0) : 1.0;
  }
}
__device__ __forceinline__ void cg_store(const CgRegs& g, GsTile<kGsH>& s, CgCells& c) {
  constexpr int W = GsTile<kGsH>::W, R = GsTile<kGsH>::R;
#pragma unroll
  for (int e = 0; e < kGsSlots; e++) {
    const int k = threadIdx.x + kNT * e;
    if (!g.in[e]) c.r[e] = c.w[e] = c.s[e] = 0.0;
    if (!g.own[e]) c.u[e] = c.p[e] = c.x[e] = 0.0;
    c.vol[e] = g.vx[e] * g.vy[e] * g.vr[e];
    if (k < R * W) {
      const int b = k / W, a = k % W;
      const double dq = g.in[e] ? g.d[e] : 1.0;
      s.d[b][a] = dq;
      s.dinv[b][a] = 1.0 / dq;
      s.cx[b][a] = g.in[e] ? g.cx[e] : 0.0;
      s.cy[b][a] = g.in[e] ? g.cy[e] : 0.0;
    }
  }
}
__device__ __forceinline__ void cg_stage(const MgArgs& A, GsTile<kGsH>& s, int i0, int j0,
                                         CgCells& c) {
  CgRegs g;
  cg_load(A, i0, j0, c, g);
  cg_store(g, s, c);
}

// The CG update of the staged cells: the new residual goes to LDS (halo cells
// recompute their owners' values exactly); the tile's own cells also write
// p, s, x, r and add |r/V|^2 to acc.
__device__ __forceinline__ void cg_update(const MgArgs& A, GsTile<kGsH>& s, int i0, int j0,
                                          const CgCells& c, double alpha, double beta,
                                          double& acc) {
  constexpr int W = GsTile<kGsH>::W, R = GsTile<kGsH>::R;
  const MgLevel& F = A.F;
#pragma unroll
  for (int e = 0; e < kGsSlots; e++) {
    const int k = threadIdx.x + kNT * e;
    if (k < R * W) {
      const int b = k / W, a = k % W;
      const int i = i0 - kGsH + a, j = j0 - kGsH + b;
      double r = 0.0;
      if (i >= 0 && i < F.nx && j >= 0 && j < F.ny) {
        const double sn = A.cg.first ? c.w[e] : c.w[e] + beta * c.s[e];
        r = c.r[e] - alpha * sn;
        if (a >= kGsH && a < kGsH + kGsTX && b >= kGsH && b < kGsH + kGsTY) {
          const size_t m = mg_at(F, i, j);
          const double pn = A.cg.first ? c.u[e] : c.u[e] + beta * c.p[e];
          A.cg.p[m] = pn;
          A.cg.s_new[m] = sn;
          A.cg.x[m] = c.x[e] + alpha * pn;
          F.r[m] = r;
          const double rn = r / c.vol[e];
          acc += rn * rn;
        }
      }
      s.r[b][a] = r;
    }
  }
}

// Down: pre-smooth M x = r on each tile, keep x1, restrict the residual into
// the coarse level.  With `check` the kernel first tests convergence on the
// partials of the previous kernel (every block identically); in the level-0
// CG form it first applies the CG update (alpha and beta from the up kernel's
// partials) and writes the partials of the new |r/V|^2, which the next
// kernel tests.  The partial sums are loaded together with the first tile.
template <int kInstance, bool kCg>
__device__ __forceinline__ void mg_down_body(const MgArgs& A, int bid, int nb) {
  constexpr int W = GsTile<kGsH>::W, R = GsTile<kGsH>::R;
  __shared__ GsTile<kGsH> s;
  __shared__ double s_res[kGsTY][kGsTX / 2];
  __shared__ double s_red[8 * kNW];
  DevScalars* sc = A.sc;
  const MgLevel &F = A.F, &C = A.C;
  const int ntx = gs_tiles_x(F);
  const TileRange tr = tile_range(gs_ntiles(F), bid, nb);
  const bool red = A.check || kCg;
  // The first tile's loads and the partials'.  In the level-0 CG form of a
  // single discharge they are issued with the solver state, before the test
  // of `done`: every address is in bounds whatever the state (a block without
  // tiles loads cells outside the level, which read the planes' first
  // element), so that kernel waits for one memory round trip instead of two
  // (measured: 0.4 us faster at 512^2; the coarser levels' kernels were
  // measured 0.2-0.4 us slower so).  An ensemble's kernels issue them after
  // the test and after the blocks without tiles have left: their launches
  // fill the chip, where registers held across the warm-start block cost a
  // wave per SIMD (256 members: 47.5 ms per pass against 35.5 ms), and they
  // cover members whose solve has stopped.  The LDS stores follow the loads.
  constexpr bool kEarly = kCg && kInstance == 0;
  CgCells cc;
  CgRegs cgr;
  GsRegs<kGsH, false> gsr;
  Part3 pl;
  auto issue_first = [&]() {
    const int i0 = (tr.t0 % ntx) * kGsTX, j0 = (tr.t0 / ntx) * kGsTY;
    if (kCg) cg_load(A, i0, j0, cc, cgr);
    else gs_load<kGsH, false>(F, i0, j0, C, gsr);
    if (red) part3_load(A.part_in, A.nb_in, pl);
  };
  if (kEarly) issue_first();
  CheckState cs{};
  double gam_old = 1.0, alpha_old = 1.0;  // the previous CG iteration's gamma and alpha
  double wa = 1.0, wb = 0.0, wc = 0.0;  // warm start x = wa x + wb x2 + wc x3 (initial test)
  int wn = 0;
  {
    const int done = sc->done;
    if (A.check) cs = check_state(sc, A.it);
    if (A.xs) {
      wa = sc->warm_a;
      wb = sc->warm_b;
      wc = sc->warm_c;
      wn = sc->warm_now;
    }
    if (kCg && !A.cg.first) {
      gam_old = sc->cg_gam[(A.it + 1) & 1];
      alpha_old = sc->cg_alpha[(A.it + 1) & 1];
    }
    if (kInstance == 0) {
      hold_mg_args(A, kCg);
      hold(nb);
      hold(cs);
      asm volatile("" ::"s"(gam_old), "s"(alpha_old), "s"(wa), "s"(wb), "s"(wc), "s"(wn));
    }
    if (done) return;
  }
  if (A.xs) {
    // the initial test of MG-PCG.  A warm start is wa x + wb x2 (the init
    // kernel formed r = rhs - M (wa x + wb x2)); it is formed here, before the
    // test may stop the solve, since the init kernel's blocks read x and x2
    // across tiles; x2 takes the last correction.  This solve's step length
    // becomes the reference of the next warm start.
    if (wn) {
      for (int t = tr.t0; t < tr.t1; t++) {
        const int i0 = (t % ntx) * kGsTX, j0 = (t / ntx) * kGsTY;
        for (int k = threadIdx.x; k < kGsTX * kGsTY; k += kNT) {
          const int i = i0 + k % kGsTX, j = j0 + k / kGsTX;
          if (i < F.nx && j < F.ny) {
            const size_t m = mg_at(F, i, j);
            const double x = A.xs[m], y = A.xs2[m], z = A.xs3[m];
            A.xs3[m] = y;
            A.xs2[m] = x;
            double v = wb != 0.0 ? wa * x + wb * y : wa * x;
            if (wc != 0.0) v += wc * z;
            A.xs[m] = v;
          }
        }
      }
    }
    if (bid == 0 && threadIdx.x == 0) {
      sc->x3_ok = wn && sc->x2_ok;
      sc->dt_warm3 = sc->dt_warm2;
      sc->dt_warm2 = sc->dt_warm;
      sc->dt_warm = sc->dt_stage;
      sc->x2_ok = wn;
    }
  }
  double sacc[3] = {0.0, 0.0, 0.0};  // sum r_c^2 / d_c for the learned coarse solver
  double racc[3] = {0.0, 0.0, 0.0};  // sum |r/V|^2 after the CG update
  if (tr.t0 >= tr.t1) {  // block 0 always has tiles (it writes the scalars)
    if (A.part_scale) write_partials3(sacc, A.part_scale, s_red, bid);
    if (kCg) write_partials3(racc, A.part_out, s_red, bid);
    return;
  }
  if (!kEarly) issue_first();
  if (kCg) cg_store(cgr, s, cc);
  else gs_store<kGsH, false>(gsr, s, 0.0);
  double pv[3];
  if (red) part3_sum(pl, A.nb_in, pv);
  double alpha = 0.0, beta = 0.0;
  bool zero_start = false;  // a warm start no better than zero, discarded
  if (red) {
    double tot[3];
    finish_partials3(pv, s_red, tot);
    if (A.check) {
      // (every block decides identically from the same partials)
      if (A.xs && wn && !(tot[0] < tot[1])) {
        zero_start = true;
        tot[0] = tot[1];
        // the corrections before this solve did not predict it (e.g. the
        // first transport step's, after the Poisson-only passes, is many
        // times larger than the next ones): the next start extrapolates
        // from this solve's correction alone, as after a solve from zero
        if (bid == 0 && threadIdx.x == 0) {
          sc->x2_ok = 0;
          sc->x3_ok = 0;
        }
      }
      if (mg_check(sc, cs, tot[0], tot[1], A.init, A.it, bid == 0 && threadIdx.x == 0)) return;
    } else {
      // gamma = r.u, delta = u.M u; p.s = delta - beta gamma / alpha_old
      // (0 < p.s <= delta in exact arithmetic for a definite operator).  Near
      // convergence the difference can lose every digit; a p.s <= 0 of that
      // size (|p.s| <= kPsCancel |delta|; rounding leaves ~1e-16 |delta|)
      // restarts the direction from u (beta = 0, p.s = delta), as in the
      // first iteration.  The assembled operator is not always definite:
      // dFdPh's two terms cancel at a near-zero field (lapl_lib.F90:625-631)
      // and can leave a face coupling of the wrong sign, and with it an
      // eigenvalue below zero.  A larger negative p.s is then that negative
      // curvature, stepped along (alpha < 0) as CG does on a symmetric
      // indefinite system; restarting there instead throws the conjugacy
      // away at every such step (measured on such a 512^2 system: the solve
      // from zero diverged after 200+ restarts; kept, it converges in 16
      // iterations).  Only p.s = 0 stops the solve.
      constexpr double kPsCancel = 1e-8;
      const double gam = tot[0], del = tot[1];
      double ps = del;
      if (!A.cg.first) {
        beta = gam / gam_old;
        ps = del - beta * gam / alpha_old;
        if (!isfinite(beta) || !isfinite(ps) || (ps <= 0.0 && -ps <= kPsCancel * fabs(del))) {
          beta = 0.0;
          ps = del;
        }
      }
      alpha = gam / ps;
      if (ps == 0.0 || !isfinite(alpha)) {  // breakdown: stop, not converged
        if (bid == 0 && threadIdx.x == 0) sc->done = 2;
        return;
      }
      if (bid == 0 && threadIdx.x == 0) {
        sc->cg_gam[A.it & 1] = gam;
        sc->cg_alpha[A.it & 1] = alpha;
      }
    }
  }
  if (zero_start) {
    // x = 0 and r = rhs on this block's cells; every block stages r from rhs
    // (the neighbours' own r may not be rewritten yet)
    for (int t = tr.t0; t < tr.t1; t++) {
      const int i0 = (t % ntx) * kGsTX, j0 = (t / ntx) * kGsTY;
      for (int k = threadIdx.x; k < kGsTX * kGsTY; k += kNT) {
        const int i = i0 + k % kGsTX, j = j0 + k / kGsTX;
        if (i < F.nx && j < F.ny) {
          const size_t m = mg_at(F, i, j);
          A.xs[m] = 0.0;
          F.r[m] = A.rhs0[m];
        }
      }
    }
  }
  MgLevel Fs = F;  // the level as staged (r from rhs after a discarded start)
  if (zero_start) Fs.r = const_cast<double*>(A.rhs0);
  for (int t = tr.t0; t < tr.t1; t++) {
    const int i0 = (t % ntx) * kGsTX, j0 = (t / ntx) * kGsTY;
    if (t > tr.t0 || zero_start) {
      __syncthreads();
      if (kCg) cg_stage(A, s, i0, j0, cc);
      else gs_stage<kGsH, false>(Fs, s, i0, j0, C, 0.0);
    }
    if (kCg) cg_update(A, s, i0, j0, cc, alpha, beta, racc[0]);
    __syncthreads();
    // first (red) half-sweep from zero is pointwise: the whole staged halo
    // (batched as the half-sweeps are: every read of a thread's cells before
    // its writes)
    if constexpr (kInstance != 0) {
      for (int k = threadIdx.x; k < R * W; k += kNT) {
        const int b = k / W, a = k % W;
        s.x[b][a] = ((i0 + a + j0 + b) & 1) == 0 ? s.r[b][a] * s.dinv[b][a] : 0.0;
      }
    } else {
      constexpr int NP = (R * W + kNT - 1) / kNT;
      double v[NP];
#pragma unroll
      for (int q = 0; q < NP; q++) {
        const int k = threadIdx.x + kNT * q;
        if (k < R * W) {
          const int b = k / W, a = k % W;
          v[q] = ((i0 + a + j0 + b) & 1) == 0 ? s.r[b][a] * s.dinv[b][a] : 0.0;
        }
      }
#pragma unroll
      for (int q = 0; q < NP; q++) {
        const int k = threadIdx.x + kNT * q;
        if (k < R * W) (&s.x[0][0])[k] = v[q];
      }
    }
    __syncthreads();
    gs_half<kInstance == 0>(s, i0, j0, 1, 3);
    __syncthreads();
    gs_half<kInstance == 0>(s, i0, j0, 0, 2);
    __syncthreads();
    gs_half<kInstance == 0>(s, i0, j0, 1, 1);
    __syncthreads();
    // residual of one horizontal cell pair (half an aggregate) per thread
    const int cI = threadIdx.x % (kGsTX / 2), fj = threadIdx.x / (kGsTX / 2);
    double rs = 0.0;
    for (int q = 0; q < 2; q++) {
      const int a = kGsH + 2 * cI + q, b = kGsH + fj;
      const int i = i0 + 2 * cI + q, j = j0 + fj;
      if (i < F.nx && j < F.ny) {
        F.x1[mg_at(F, i, j)] = s.x[b][a];
        rs += s.r[b][a] - gs_mx(s, b, a);
      }
    }
    s_res[fj][cI] = rs;
    __syncthreads();
    if ((fj & 1) == 0) {
      const int I = i0 / 2 + cI, J = (j0 + fj) / 2;
      if (I < C.nx && J < C.ny) {
        const size_t mc = mg_at(C, I, J);
        const double rc = s_res[fj][cI] + s_res[fj + 1][cI];
        C.r[mc] = rc;
        if (A.part_scale) sacc[0] += rc * rc / C.d[mc];
      }
    }
  }
  if (A.part_scale) write_partials3(sacc, A.part_scale, s_red, bid);
  if (kCg) write_partials3(racc, A.part_out, s_red, bid);
}

// kCg: the level-0 CG form (its own instance: the CG vectors it keeps in
// registers would otherwise limit the occupancy of the other levels' launches)
template <bool kCg>
__global__ void __launch_bounds__(kNT) k_mg_down(MgArgs A) {
  mg_down_body<0, kCg>(A, blockIdx.x, gridDim.x);
}

// Up: x2 = x1 + oc * e_c(aggregate) on the staged halo, post-smooth, write e.
// Level 0 writes u = B r and the partials of r.u, and of z.q (flexible CG) or,
// in the CG form, w = M u and the partials of u.w; M u needs u one cell past
// the tile, so that kernel stages a halo of H = 5 and sweeps one cell wider.
template <int kInstance, int H>
__device__ __forceinline__ void mg_up_body(const MgArgs& A, int bid, int nb) {
  constexpr int W = GsTile<H>::W, R = GsTile<H>::R;
  __shared__ GsTile<H> s;
  __shared__ double s_red[8 * kNW];
  const MgLevel &F = A.F, &C = A.C;
  const int ntx = gs_tiles_x(F);
  const TileRange tr = tile_range(gs_ntiles(F), bid, nb);
  // A single discharge issues the first tile's loads before the test of
  // `done` (in bounds whatever the state, as in the down kernel).  An
  // ensemble stages every tile in the loop: its launches cover members whose
  // solve has stopped, and the first tile's registers, live across the loop,
  // would cost it a wave per SIMD.
  GsRegs<H, true> g0;
  if (kInstance == 0) gs_load<H, true>(F, (tr.t0 % ntx) * kGsTX, (tr.t0 / ntx) * kGsTY, C, g0);
  {
    const int done = A.sc->done;
    if (kInstance == 0) {
      hold_mg_args(A, false);
      hold(nb);
    }
    if (done) return;
  }
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = tr.t0; t < tr.t1; t++) {
    const int i0 = (t % ntx) * kGsTX, j0 = (t / ntx) * kGsTY;
    if (kInstance != 0 || t > tr.t0) {
      if (t > tr.t0) __syncthreads();
      gs_stage<H, true>(F, s, i0, j0, C, A.oc);
    } else {
      gs_store<H, true>(g0, s, A.oc);
    }
    __syncthreads();
    gs_half<kInstance == 0>(s, i0, j0, 1, H - 1);
    __syncthreads();
    gs_half<kInstance == 0>(s, i0, j0, 0, H - 2);
    __syncthreads();
    gs_half<kInstance == 0>(s, i0, j0, 1, H - 3);
    __syncthreads();
    gs_half<kInstance == 0>(s, i0, j0, 0, H - 4);
    __syncthreads();
    for (int k = threadIdx.x; k < kGsTX * kGsTY; k += kNT) {
      const int i = i0 + k % kGsTX, j = j0 + k / kGsTX;
      if (i < F.nx && j < F.ny) {
        const int b = H + k / kGsTX, a = H + k % kGsTX;
        const double e = s.x[b][a];
        const size_t m = mg_at(F, i, j);
        F.e[m] = e;
        acc[0] += s.r[b][a] * e;
        if (A.q) acc[1] += A.q[m] * e;
        if (H > kGsH && A.cg.on) {
          const double w = gs_mx(s, b, a);
          A.cg.w[m] = w;
          acc[1] += e * w;
        }
      }
    }
  }
  if (A.level0) write_partials3(acc, A.part_out, s_red, bid);
}

template <int H>
__global__ void __launch_bounds__(kNT) k_mg_up(MgArgs A) { mg_up_body<0, H>(A, blockIdx.x, gridDim.x); }

// Once per solve, for the one-workgroup levels f .. L-1 (1024 threads):
//  * levels f+1 .. L-1 are coarsened from level f (written by the tile
//    kernels) in LDS, and also stored;
//  * the coarsest level (<= kMgBottomMax cells) is inverted by Gauss-Jordan
//    elimination (SPD, so no pivoting): thread (row, col) < kB^2 keeps entry
//    (row, col) in a register; each pivot step publishes the pivot row and
//    column through LDS buffers alternated by step parity, one barrier per
//    step; the inverse is stored column-major;
//  * the operator rows of the other levels are packed as FP32 (inverse
//    diagonal, left, right, lower, upper coupling) in the order the cycle
//    kernel's threads hold them (slot-major, 1024 threads per slot), so that
//    the cycle loads half the bytes, coalesced.
template <int kInstance>
__device__ __forceinline__ void mg_coarse_setup_body(const MgCycleArgs& A) {
  constexpr int kB = kMgBottomMax;
  __shared__ double s_prow[2][kB], s_pcol[2][kB];
  // Galerkin coefficients of levels f+1 .. L-1, coarsened here from level f
  // (whose tile kernel wrote it) and kept in LDS (padded planes at s_off[q])
  __shared__ double s_c[3][kMgSetupX];
  __shared__ int s_off[kMgLdsLevels];
  const int t = threadIdx.x;
  if (t == 0) {
    int o = 0;
    for (int q = 1; q < A.nq; q++) {
      s_off[q] = o;
      o += (A.lv[q].nx + 2) * (A.lv[q].ny + 2);
    }
  }
  for (int k = t; k < kMgSetupX; k += 1024) s_c[0][k] = s_c[1][k] = s_c[2][k] = 0.0;
  __syncthreads();
  auto lvd = [&](int q) -> const double* { return q == 0 ? A.lv[0].d : s_c[0] + s_off[q]; };
  auto lvx = [&](int q) -> const double* { return q == 0 ? A.lv[0].cx : s_c[1] + s_off[q]; };
  auto lvy = [&](int q) -> const double* { return q == 0 ? A.lv[0].cy : s_c[2] + s_off[q]; };
  for (int q = 0; q + 1 < A.nq; q++) {
    const MgLevel &F = A.lv[q], &C = A.lv[q + 1];
    for (int n = t; n < C.nx * C.ny; n += 1024) {
      const int I = n % C.nx, J = n / C.nx;
      const Coarse3 c = coarsen_cell(lvd(q), lvx(q), lvy(q), F.nx + 2, F.nx, F.ny, C.nx, C.ny, I, J);
      const size_t m = mg_at(C, I, J);
      s_c[0][s_off[q + 1] + m] = c.d;
      s_c[1][s_off[q + 1] + m] = c.cx;
      s_c[2][s_off[q + 1] + m] = c.cy;
      C.d[m] = c.d;
      C.cx[m] = c.cx;
      C.cy[m] = c.cy;
    }
    __syncthreads();
  }
  {
    const int qb = A.nq - 1;
    const MgLevel& L = A.lv[qb];
    const double *Ld = lvd(qb), *Lx = lvx(qb), *Ly = lvy(qb);
    const int n = L.nx * L.ny, nx = L.nx;
    const bool own = t < kB * kB;
    const int row = t % kB, col = t / kB;
    double a = 0.0;
    if (own && row < n && col < n) {
      const size_t m = mg_at(L, row % nx, row / nx);
      if (col == row) a = Ld[m];
      else if (col == row - 1 && row % nx > 0) a = -Lx[m - 1];
      else if (col == row + 1 && row % nx + 1 < nx) a = -Lx[m];
      else if (col == row - nx) a = -Ly[m - L.P];
      else if (col == row + nx) a = -Ly[m];
    }
    for (int p = 0; p < n; p++) {
      const int buf = p & 1;
      if (own && row == p) s_prow[buf][col] = a;
      if (own && col == p) s_pcol[buf][row] = a;
      __syncthreads();
      if (own) {
        const double piv = s_prow[buf][p], f = s_pcol[buf][row];
        const double pk = (col == p ? 1.0 : s_prow[buf][col]) / piv;
        a = row == p ? pk : (col == p ? 0.0 : a) - f * pk;
      }
    }
    if (own && row < n && col < n) A.inv[(size_t)col * n + row] = a;
  }
  if (A.nq < 2) return;
  auto put = [&](int slot, int k, int lq, bool valid) {
    float* q = A.pk + (size_t)slot * 5 * 1024 + t;
    const int W = A.lv[lq].nx + 2;
    const double *Ld = lvd(lq), *Lx = lvx(lq), *Ly = lvy(lq);
    q[0] = valid ? (float)(1.0 / Ld[k]) : 1.0f;
    q[1024] = valid ? (float)Lx[k - 1] : 0.0f;
    q[2048] = valid ? (float)Lx[k] : 0.0f;
    q[3072] = valid ? (float)Ly[k - W] : 0.0f;
    q[4096] = valid ? (float)Ly[k] : 0.0f;
  };
  {  // level f by aggregates: child c = a + 2b of cell t of level f+1 in slot c
    const MgLevel &L = A.lv[0], &C = A.lv[1];
    const int W = L.nx + 2;
    const bool own = t < C.nx * C.ny;
    const int I = own ? t % C.nx : 0, J = own ? t / C.nx : 0;
    for (int c = 0; c < 4; c++) {
      const int i = 2 * I + (c & 1), j = 2 * J + (c >> 1);
      const bool valid = own && i < L.nx && j < L.ny;
      put(c, valid ? (j + 1) * W + i + 1 : 0, 0, valid);
    }
  }
  for (int q = 1; q + 1 < A.nq; q++) {
    const MgLevel& L = A.lv[q];
    const bool valid = t < L.nx * L.ny;
    const int k = valid ? (t / L.nx + 1) * (L.nx + 2) + t % L.nx + 1 : 0;
    put(4 + q - 1, k, q, valid);
  }
}

__global__ void __launch_bounds__(1024) k_mg_coarse_setup(MgCycleArgs A) {
  mg_coarse_setup_body<0>(A);
}

// Levels f .. L-1 in one workgroup of 1024 threads (at most kMgLdsLevels;
// level f has at most kMgLdsCells cells, the next at most 1024).  Iterates and
// residuals live in LDS in the levels' padded (nx+2)(ny+2) layout, which is
// also their layout in HBM; operator rows (the set-up kernel's FP32 table) and
// right-hand sides stay in registers.  Level f is held by aggregates: thread t
// owns the 2x2 children of cell t of level f+1, two of each colour, with their
// iterates in registers as well, so half of each update's neighbours are its
// own registers, and the restriction to and prolongation from level f+1 stay
// in the thread (it holds cell t there); the coarser levels hold one cell per
// thread.  The sweeps multiply by the inverse diagonal.  The operator rows are
// FP32 values but the arithmetic is FP64: the cycle must stay a symmetric
// operator to rounding level, which the CG recurrences (fused, Chronopoulos-
// Gear) rely on; with FP32 arithmetic here, warm-started solves broke down near
// convergence.  Same cycle as the tile kernels; level f's correction is
// written back for the levels above.
//
// Children c = a + 2b at (2I + a, 2J + b): colour (a + b) & 1, i.e. 0 and 3
// are red, 1 and 2 black.
struct AggRows {
  float dinv[4], l[4], r[4], b[4], t[4];  // held as the table's FP32 values
  double rhs[4], x[4];
  int k0;    // padded index of child 0, or -1 when thread t has no aggregate
  int mask;  // bit c: child c lies inside the level
};
struct CellRows {
  float dinv, l, r, b, t;
  double rhs;
  int k;   // padded index of cell t, or -1
  int ij;  // i | j << 16
};

// Rows of level f (aggregates) and of a coarser level q >= 1 (one cell) from
// the packed FP32 table; right-hand sides of level f from its residual plane.
__device__ __forceinline__ void lds_agg_rows(const MgLevel& L, const MgLevel& C, AggRows& R,
                                             const float* pk, int t) {
  const int W = L.nx + 2;
  R.k0 = -1;
  R.mask = 0;
  int I = 0, J = 0;
  if (t < C.nx * C.ny) {
    I = t % C.nx;
    J = t / C.nx;
    R.k0 = (2 * J + 1) * W + 2 * I + 1;
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int a = c & 1, b = c >> 1;
    const float* q = pk + (size_t)c * 5 * 1024 + t;
    R.dinv[c] = q[0];
    R.l[c] = q[1024];
    R.r[c] = q[2048];
    R.b[c] = q[3072];
    R.t[c] = q[4096];
    R.x[c] = 0.0;
    const bool in = R.k0 >= 0 && 2 * I + a < L.nx && 2 * J + b < L.ny;
    // loaded for every child (the padded plane's first element outside the
    // level): only children with their mask bit set are ever read
    R.rhs[c] = gld(L.r, in ? R.k0 + a + b * W : 0);
    if (in) R.mask |= 1 << c;
  }
}

__device__ __forceinline__ void lds_cell_rows(const MgLevel& L, CellRows& R, const float* pk,
                                              int slot, int t) {
  const float* q = pk + (size_t)slot * 5 * 1024 + t;
  R.dinv = q[0];
  R.l = q[1024];
  R.r = q[2048];
  R.b = q[3072];
  R.t = q[4096];
  R.rhs = 0.0;
  R.k = -1;
  R.ij = 0;
  if (t < L.nx * L.ny) {
    const int i = t % L.nx, j = t / L.nx;
    R.k = (j + 1) * (L.nx + 2) + i + 1;
    R.ij = i | (j << 16);
  }
}

// Position of cell t on the coarsest level (its rows are not needed).
__device__ __forceinline__ void lds_cell_pos(const MgLevel& L, CellRows& R, int t) {
  R.dinv = 1.0f;
  R.l = R.r = R.b = R.t = 0.0f;
  R.rhs = 0.0;
  R.k = -1;
  R.ij = 0;
  if (t < L.nx * L.ny) {
    const int i = t % L.nx, j = t / L.nx;
    R.k = (j + 1) * (L.nx + 2) + i + 1;
    R.ij = i | (j << 16);
  }
}

__device__ __forceinline__ double lds_gs(double rhs, double l, double r, double b, double t, double dinv,
                                        const double* x, int k, int W) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
  return (rhs + l * x[k - 1] + r * x[k + 1] + b * x[k - W] + t * x[k + W]) * dinv;
}
__device__ __forceinline__ double lds_res(double rhs, double l, double r, double b, double t, double dinv,
                                         const double* x, int k, int W) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
  return rhs - (x[k] / dinv - l * x[k - 1] - r * x[k + 1] - b * x[k - W] - t * x[k + W]);
}

// Neighbour values of child c (left, right, below, above): the other colour's
// children of the same aggregate from registers, the rest from LDS.
__device__ __forceinline__ void agg_nb(const AggRows& R, const double* x, int W, int c,
                                       double& L, double& Rt, double& B, double& T) {
  const int k = R.k0;
  if (c == 0) {
    L = x[k - 1];
    Rt = R.x[1];
    B = x[k - W];
    T = R.x[2];
  } else if (c == 1) {
    L = R.x[0];
    Rt = x[k + 2];
    B = x[k + 1 - W];
    T = R.x[3];
  } else if (c == 2) {
    L = x[k + W - 1];
    Rt = R.x[3];
    B = R.x[0];
    T = x[k + 2 * W];
  } else {
    L = R.x[2];
    Rt = x[k + W + 2];
    B = R.x[1];
    T = x[k + 2 * W + 1];
  }
}

__device__ __forceinline__ int agg_k(const AggRows& R, int W, int c) {
  return R.k0 + (c & 1) + (c >> 1) * W;
}

// One half-sweep of colour COL over the aggregates (children 0, 3 or 1, 2).
template <int COL>
__device__ __forceinline__ void agg_half(AggRows& R, double* x, int W) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
  constexpr int c0 = COL == 0 ? 0 : 1, c1 = COL == 0 ? 3 : 2;
  double v[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int c = q == 0 ? c0 : c1;
    v[q] = R.x[c];
    if (R.mask & (1 << c)) {
      double L, Rt, B, T;
      agg_nb(R, x, W, c, L, Rt, B, T);
      v[q] = (R.rhs[c] + R.l[c] * L + R.r[c] * Rt + R.b[c] * B + R.t[c] * T) * R.dinv[c];
    }
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int c = q == 0 ? c0 : c1;
    if (R.mask & (1 << c)) {
      R.x[c] = v[q];
      x[agg_k(R, W, c)] = v[q];
    }
  }
  __syncthreads();
}

// Pre-smoothing from zero (red pointwise, then black, red, black); returns the
// aggregate's residual sum, the right-hand side of cell t of level f+1.
__device__ __forceinline__ double agg_down(AggRows& R, double* x, int W) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
#pragma unroll
  for (int c = 0; c < 4; c += 3)
    if (R.mask & (1 << c)) {
      R.x[c] = R.rhs[c] * R.dinv[c];
      x[agg_k(R, W, c)] = R.x[c];
    }
  __syncthreads();
  agg_half<1>(R, x, W);
  agg_half<0>(R, x, W);
  agg_half<1>(R, x, W);
  double sum = 0.0;
#pragma unroll
  for (int c = 0; c < 4; c++)
    if (R.mask & (1 << c)) {
      double L, Rt, B, T;
      agg_nb(R, x, W, c, L, Rt, B, T);
      sum += R.rhs[c] - (R.x[c] / R.dinv[c] - R.l[c] * L - R.r[c] * Rt - R.b[c] * B - R.t[c] * T);
    }
  return sum;
}

// Prolongation of e (the correction of cell t of level f+1) with
// over-correction, post-smoothing (black, red, black, red), and the result.
__device__ __forceinline__ void agg_up(AggRows& R, double* x, int W, double e, double oc,
                                       double* out) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
#pragma unroll
  for (int c = 0; c < 4; c++)
    if (R.mask & (1 << c)) {
      R.x[c] += oc * e;
      x[agg_k(R, W, c)] = R.x[c];
    }
  __syncthreads();
  agg_half<1>(R, x, W);
  agg_half<0>(R, x, W);
  agg_half<1>(R, x, W);
  agg_half<0>(R, x, W);
#pragma unroll
  for (int c = 0; c < 4; c++)
    if (R.mask & (1 << c)) out[agg_k(R, W, c)] = R.x[c];
}

template <int COL>
__device__ __forceinline__ void cell_half(const CellRows& R, double* x, int W) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
  if (R.k >= 0 && (((R.ij & 0xffff) + (R.ij >> 16)) & 1) == COL)
    x[R.k] = lds_gs(R.rhs, R.l, R.r, R.b, R.t, R.dinv, x, R.k, W);
  __syncthreads();
}

__device__ __forceinline__ void cell_down(const CellRows& R, double* x, double* s_r, int W) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
  const bool red = R.k >= 0 && (((R.ij & 0xffff) + (R.ij >> 16)) & 1) == 0;
  if (red) x[R.k] = R.rhs * R.dinv;
  __syncthreads();
  cell_half<1>(R, x, W);
  cell_half<0>(R, x, W);
  cell_half<1>(R, x, W);
  if (R.k >= 0) s_r[R.k] = lds_res(R.rhs, R.l, R.r, R.b, R.t, R.dinv, x, R.k, W);
  __syncthreads();
}

// Restriction of level L's residual (s_r) to the coarse cell held in C.
__device__ __forceinline__ double lds_restrict(const MgLevel& L, const CellRows& C,
                                              const double* s_r) {
  double v = 0.0;
  if (C.k >= 0) {
    const int I = C.ij & 0xffff, J = C.ij >> 16, W = L.nx + 2;
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++)
        if (2 * I + a < L.nx && 2 * J + b < L.ny) v += s_r[(2 * J + b + 1) * W + 2 * I + a + 1];
  }
  return v;
}

__device__ __forceinline__ void cell_up(const CellRows& R, double* x, const double* xc, int W,
                                        int Wc, double oc) {
  #pragma clang fp contract(fast)  // smoother arithmetic: FMAs allowed
  if (R.k >= 0) {
    const int i = R.ij & 0xffff, j = R.ij >> 16;
    x[R.k] += oc * xc[((j >> 1) + 1) * Wc + (i >> 1) + 1];
  }
  __syncthreads();
  cell_half<1>(R, x, W);
  cell_half<0>(R, x, W);
  cell_half<1>(R, x, W);
  cell_half<0>(R, x, W);
}

template <int kInstance>
__device__ __forceinline__ void mg_coarse_cycle_body(const MgCycleArgs& A) {
  __shared__ double s_x[kMgLdsX];
  __shared__ double s_r[kMgLdsR];
  constexpr int kB = kMgBottomMax;
  __shared__ double s_bp[kB][kB];
  __shared__ double s_w[16];
  static_assert(kMgLdsLevels == 5, "the hold below lists five levels");
  if (kInstance == 0)  // one statement: one round of argument loads
    asm volatile("" ::"s"(A.lv[0].r), "s"(A.lv[0].e), "s"(A.lv[0].nx), "s"(A.lv[0].ny),
                 "s"(A.lv[1].nx), "s"(A.lv[1].ny), "s"(A.lv[2].nx), "s"(A.lv[2].ny),
                 "s"(A.lv[3].nx), "s"(A.lv[3].ny), "s"(A.lv[4].nx), "s"(A.lv[4].ny),
                 "s"(A.off[0]), "s"(A.off[1]), "s"(A.off[2]), "s"(A.off[3]), "s"(A.off[4]),
                 "s"(A.nq), "s"(A.inv), "s"(A.pk), "s"(A.oc), "s"(A.part_in), "s"(A.nb_in),
                 "s"(A.it), "s"(A.check), "s"(A.sc));
  // an ensemble member whose solve has stopped leaves before any load
  if (kInstance != 0 && A.sc->done) return;
  const int t = threadIdx.x;
  const int nq = A.nq;
  // The coarsest level's inverse: entry (brow, bcol) in thread brow + kB * bcol,
  // loaded with the partial (nb_in <= 1024: one per thread) and the operator
  // rows, whose loads overlap the convergence test.  Every load is issued
  // unconditionally, and before the test of `done` (rows of levels the
  // hierarchy does not have are loaded from the table and never used; without
  // the test the partial is read from the inverse; every address is in bounds
  // whatever the solver state), so that they all share one wait.
  const MgLevel& B = A.lv[nq - 1];  // coarsest
  const int nb = B.nx * B.ny;
  const int brow = t % kB, bcol = t / kB;
  const bool own_b = t < kB * kB && brow < nb && bcol < nb;
  const double inv_v = gld((const double*)A.inv, own_b ? (size_t)bcol * nb + brow : 0);
  const double* pin = A.check ? A.part_in : (const double*)A.inv;
  const double part_v = gld(pin, A.check && t < A.nb_in ? 3 * (size_t)t : 0);
  AggRows R0;                     // level f
  CellRows R1[kMgLdsLevels - 2];  // levels f+1 .. L-2
  CellRows Bc;                    // the coarsest level's cell t (its position only)
  lds_agg_rows(A.lv[0], A.lv[1], R0, A.pk, t);
#pragma unroll
  for (int q = 1; q < kMgLdsLevels - 1; q++) lds_cell_rows(A.lv[q], R1[q - 1], A.pk, 4 + q - 1, t);
  lds_cell_pos(A.lv[nq - 1], Bc, t);
  CheckState cs{};
  {
    const int done = A.sc->done;
    if (A.check) cs = check_state(A.sc, A.it);
    if (kInstance == 0) hold(cs);
    if (done) return;
  }
  if (A.check) {  // convergence test on the level-0 CG update's |r/V|^2 partials
    double v = t < A.nb_in ? part_v : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((t & 63) == 0) s_w[t >> 6] = v;
    __syncthreads();
    double rho = 0.0;
    for (int w = 0; w < 16; w++) rho += s_w[w];
    if (mg_check(A.sc, cs, rho, 0.0, 0, A.it, t == 0)) return;
  }
  const double binv = own_b ? inv_v : 0.0;
  for (int k = t; k < kMgLdsX; k += 1024) s_x[k] = 0.0;
  double rb = 0.0;  // right-hand side of the coarsest level at cell t
  if (nq == 1) {
    if (t < nb) rb = (double)B.r[(t / B.nx + 1) * (B.nx + 2) + t % B.nx + 1];
  } else {
    __syncthreads();
    // down: level f (its restriction is cell t's own), then f+1 .. L-2
    {
      const double v = agg_down(R0, s_x + A.off[0], A.lv[0].nx + 2);
      if (nq == 2) rb = t < nb ? v : 0.0;
      else R1[0].rhs = R1[0].k >= 0 ? v : 0.0;
    }
#pragma unroll
    for (int q = 1; q < kMgLdsLevels - 1; q++) {
      if (q + 1 >= nq) break;
      cell_down(R1[q - 1], s_x + A.off[q], s_r, A.lv[q].nx + 2);
      const double w = lds_restrict(A.lv[q], q + 2 == nq ? Bc : R1[q], s_r);
      if (q + 2 == nq) rb = w;
      else R1[q].rhs = w;
      __syncthreads();
    }
    // coarsest level: x = inv r (products in LDS, summed per row)
    if (t < nb) s_r[t] = rb;
    __syncthreads();
    if (t < kB * kB) s_bp[bcol][brow] = bcol < nb ? binv * s_r[bcol] : 0.0;
    __syncthreads();
    if (t < nb) {
      double v = 0.0;
      for (int c = 0; c < nb; c++) v += s_bp[c][t];
      s_x[A.off[nq - 1] + Bc.k] = v;
    }
    __syncthreads();
    // up: levels L-2 .. f+1, then level f
#pragma unroll
    for (int q = kMgLdsLevels - 2; q >= 1; q--) {
      if (q + 1 >= nq) continue;
      cell_up(R1[q - 1], s_x + A.off[q], s_x + A.off[q + 1], A.lv[q].nx + 2, A.lv[q + 1].nx + 2,
              A.oc);
    }
    // level f: the correction of cell t of level f+1 was last written by this
    // thread (cell t is its own there)
    const MgLevel& L = A.lv[0];
    const double e = Bc.k >= 0 && nq == 2 ? s_x[A.off[1] + Bc.k]
                     : (nq > 2 && R1[0].k >= 0 ? s_x[A.off[1] + R1[0].k] : 0.0);
    agg_up(R0, s_x + A.off[0], L.nx + 2, e, A.oc, L.e);
    return;
  }
  // a single level: the coarsest, solved directly
  if (t < nb) s_r[t] = rb;
  __syncthreads();
  if (t < kB * kB) s_bp[bcol][brow] = bcol < nb ? binv * s_r[bcol] : 0.0;
  __syncthreads();
  if (t < nb) {
    double v = 0.0;
    for (int c = 0; c < nb; c++) v += s_bp[c][t];
    B.e[(t / B.nx + 1) * (B.nx + 2) + t % B.nx + 1] = (double)v;
  }
}
__global__ void __launch_bounds__(1024) k_mg_coarse_cycle(MgCycleArgs A) { mg_coarse_cycle_body<0>(A); }

// MG-PCG start: r = rhs (from the assembly) with x = 0, or with `warm`
// r = rhs - M x for the x in place; p = 0; partials of |r/V|^2 and |rhs/V|^2.
template <int kInstance>
__device__ __forceinline__ void mgpcg_init_body(const MgPcgArgs& A, int bid, int nb) {
  __shared__ double s_red[8 * kNW];
  const Geo& G = A.G;
  const int P = G.P;
  const TileRange tr = tile_range(A.nbx * A.nby, bid, nb);
  // warm: start from a x + b x2 (the initial convergence test kernel forms it
  // in place, since other blocks read x and x2 here).  With g = correction /
  // step length sampled at the midpoints of the last two steps (dt1 = dt_warm,
  // dt2 = dt_warm2), extrapolated to the midpoint of this one (dt = dt_stage):
  // c = (dt + dt1) / (dt1 + dt2), a = (1 + c) dt / dt1, b = -c dt / dt2;
  // without x2, a = dt / dt1 (at most kWarmRatio), b = 0.  The extrapolation
  // only for step-length ratios dt / dt1 and dt1 / dt2 within [1/kWarmRatio,
  // kWarmRatio]: a correction of a much shorter step (after a rejected one)
  // carries the solver's tolerance in a large relative error, which the
  // extrapolation would amplify.  A warm solve stops after kWarmIters
  // iterations (its caller then solves from zero).
  constexpr double kWarmRatio = 3.0;
  constexpr int kWarmIters = 64;
  auto moderate = [](double q) { return q >= 1.0 / kWarmRatio && q <= kWarmRatio; };
  // While the step length still changes (|dt / dt1 - 1| > kQuadChange) and x3
  // is there, g is extrapolated quadratically through the last three
  // midpoints (Lagrange weights L_k at the new midpoint; x_k enters as
  // dt L_k / dt_k).  Measured on the 512^2 pulse, that start is 5-20x closer
  // while dt shrinks; at a steady dt both starts are at the solver-tolerance
  // floor and the quadratic weights amplify that noise more.
  constexpr double kQuadChange = 2e-3;
  double a = 1.0, b = 0.0, cq = 0.0;
  if (A.warm && A.ls) {  // the projection's coefficients (k_mgpcg_warm_solve)
    a = A.sc->warm_a;
    b = A.sc->warm_b;
    cq = A.sc->warm_c;
  } else if (A.warm) {
    const DevScalars* sc = A.sc;
    const double dt = sc->dt_stage, dt1 = sc->dt_warm, dt2 = sc->dt_warm2;
    const double a1 = dt / dt1;
    if (dt > 0.0 && dt1 > 0.0 && isfinite(a1)) {
      a = fmin(a1, kWarmRatio);
      const double q = dt1 / dt2;
      if (moderate(a1) && sc->x2_ok && dt2 > 0.0 && moderate(q)) {
        const double dt3 = sc->dt_warm3, q3 = dt2 / dt3;
        if (fabs(a1 - 1.0) > kQuadChange && sc->x3_ok && dt3 > 0.0 && moderate(q3)) {
          const double t0 = 0.5 * dt, t1 = -0.5 * dt1, t2 = -dt1 - 0.5 * dt2,
                       t3 = -dt1 - dt2 - 0.5 * dt3;
          a = dt * ((t0 - t2) * (t0 - t3) / ((t1 - t2) * (t1 - t3))) / dt1;
          b = dt * ((t0 - t1) * (t0 - t3) / ((t2 - t1) * (t2 - t3))) / dt2;
          cq = dt * ((t0 - t1) * (t0 - t2) / ((t3 - t1) * (t3 - t2))) / dt3;
        } else {
          const double c = (dt + dt1) / (dt1 + dt2);
          a = (1.0 + c) * a1;
          b = -c * (dt / dt2);
        }
      }
    }
  }
  if (bid == 0 && threadIdx.x == 0) {
    if (!A.ls) {
      A.sc->warm_a = a;
      A.sc->warm_b = b;
      A.sc->warm_c = cq;
    }
    A.sc->warm_now = A.warm;
    A.sc->itmax = A.warm ? min(A.sc->maxit, kWarmIters) : A.sc->maxit;
  }
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = tr.t0; t < tr.t1; t++) {
    for (int k = 0; k < kPcgRows; k++) {
      const int ci = (t % A.nbx) * kTX + (int)(threadIdx.x % kTX);
      const int cj = (t / A.nbx) * kPcgTY + (int)(threadIdx.x / kTX) + k * (kPcgTY / kPcgRows);
      if (ci < G.bx && cj < G.by) {
        size_t m = (size_t)(cj + 1) * P + (ci + 1);
        const double rhs = A.r[m];
        double r = rhs;
        if (A.warm) {
          const double* x = A.x;
          const double* y = A.x2;
          double mx = a * (A.dg[m] * x[m] - A.cx[m - 1] * x[m - 1] - A.cx[m] * x[m + 1] -
                           A.cy[m - P] * x[m - P] - A.cy[m] * x[m + P]);
          if (b != 0.0)
            mx += b * (A.dg[m] * y[m] - A.cx[m - 1] * y[m - 1] - A.cx[m] * y[m + 1] -
                       A.cy[m - P] * y[m - P] - A.cy[m] * y[m + P]);
          if (cq != 0.0) {
            const double* z = A.x3;
            mx += cq * (A.dg[m] * z[m] - A.cx[m - 1] * z[m - 1] - A.cx[m] * z[m + 1] -
                        A.cy[m - P] * z[m - P] - A.cy[m] * z[m + P]);
          }
          r = rhs - mx;
        }
        A.p_new[m] = 0.0;
        const double V = G.dlx[ci] * G.dly[cj] * (G.cyl ? G.r[cj + 1] : 1.0);
        const double rv = r / V, bv = rhs / V;
        acc[0] += rv * rv;
        acc[1] += bv * bv;
        if (A.warm) A.r[m] = r;  // x is only read in this kernel
        else A.x[m] = 0.0;       // x is not read in this kernel
      }
    }
  }
  write_partials3(acc, A.part_out, s_red, bid);
}

__global__ void __launch_bounds__(kNT) k_mgpcg_init(MgPcgArgs A) { mgpcg_init_body<0>(A, blockIdx.x, gridDim.x); }

// Projected warm start (fs_params.lapl_warm = 2): the start x0 = a x + b x2 +
// c x3 closest to the solution in the operator's energy norm |x0 - M^-1 rhs|_M
// over the span of the last converged corrections (x; x2 when x2_ok; x3 when
// x3_ok), a span that holds both the extrapolation and zero: the Galerkin
// projection of the new system onto the past corrections, the error norm CG
// itself minimises (a start that only minimises the residual norm leaves
// smooth error that costs the multigrid-PCG more iterations, measured on the
// 512^2 pulse).  Successive corrections are parallel in that norm to ~1e-11
// (on the 512^2 pulse), so the span is taken in the difference basis u0 = x,
// u1 = x - x2, u2 = x - 2 x2 + x3, each formed cell by cell before M is
// applied: its Gram matrix G_kl = u_k . M u_l is well conditioned where that
// of x, x2, x3 would hold the new directions below its rounding.  With
// g_k = u_k . rhs the coefficients solve G c = g; x0 = (c0 + c1 + c2) x -
// (c1 + 2 c2) x2 + c2 x3.  k_mgpcg_warm_ls writes each block's nine sums
// (every product is formed, the invalid ones are not used);
// k_mgpcg_warm_solve reduces them in a fixed order and solves by Cholesky,
// keeping the leading basis vectors whose pivots stay above kLsPivot of their
// diagonal (a direction already in the span of the earlier ones is dropped
// with the later ones; on an indefinite operator a non-positive pivot drops
// it too).  The init kernel then forms r = rhs - M x0 with these
// coefficients, and the initial test discards a start no better than zero as
// for the extrapolation.
__device__ __forceinline__ void mgpcg_warm_ls_body(const MgPcgArgs& A, int bid, int nb) {
  __shared__ double s_red[9 * kNW];
  const Geo& G = A.G;
  const int P = G.P;
  const TileRange tr = tile_range(A.nbx * A.nby, bid, nb);
  // the basis vectors at padded index q, and M applied to one of them at m
  auto u0 = [&](size_t q) { return A.x[q]; };
  auto u1 = [&](size_t q) { return A.x[q] - A.x2[q]; };
  auto u2 = [&](size_t q) { return (A.x[q] - A.x2[q]) - (A.x2[q] - A.x3[q]); };
  auto mu = [&](auto u, size_t m) {
    return A.dg[m] * u(m) - A.cx[m - 1] * u(m - 1) - A.cx[m] * u(m + 1) - A.cy[m - P] * u(m - P) -
           A.cy[m] * u(m + P);
  };
  double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = tr.t0; t < tr.t1; t++) {
    for (int k = 0; k < kPcgRows; k++) {
      const int ci = (t % A.nbx) * kTX + (int)(threadIdx.x % kTX);
      const int cj = (t / A.nbx) * kPcgTY + (int)(threadIdx.x / kTX) + k * (kPcgTY / kPcgRows);
      if (ci < G.bx && cj < G.by) {
        const size_t m = (size_t)(cj + 1) * P + (ci + 1);
        const double f = A.r[m], x0 = u0(m), x1 = u1(m), x2 = u2(m);
        const double y0 = mu(u0, m), y1 = mu(u1, m), y2 = mu(u2, m);
        acc[0] += x0 * y0;
        acc[1] += 0.5 * (x0 * y1 + x1 * y0);
        acc[2] += 0.5 * (x0 * y2 + x2 * y0);
        acc[3] += x1 * y1;
        acc[4] += 0.5 * (x1 * y2 + x2 * y1);
        acc[5] += x2 * y2;
        acc[6] += x0 * f;
        acc[7] += x1 * f;
        acc[8] += x2 * f;
      }
    }
  }
  block_sum<9>(acc, s_red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 9; k++) A.lspart[9 * (size_t)bid + k] = acc[k];
}
__global__ void __launch_bounds__(kNT) k_mgpcg_warm_ls(MgPcgArgs A) {
  mgpcg_warm_ls_body(A, blockIdx.x, gridDim.x);
}

// (one workgroup of kNT threads; nb = the blocks of the sums' launch, that of
// the start kernel)
__device__ __forceinline__ void mgpcg_warm_solve_body(const MgPcgArgs& A) {
  __shared__ double s_red[9 * kNW];
  constexpr double kLsPivot = 1e-10;
  const int nb = min(A.nbx * A.nby, kMaxPartials);
  double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += kNT)
    for (int k = 0; k < 9; k++) acc[k] += A.lspart[9 * (size_t)b + k];
  block_sum<9>(acc, s_red);
  if (threadIdx.x != 0) return;
  DevScalars* sc = A.sc;
  const double Gm[3][3] = {{acc[0], acc[1], acc[2]}, {acc[1], acc[3], acc[4]}, {acc[2], acc[4], acc[5]}};
  const double g[3] = {acc[6], acc[7], acc[8]};
  const int nv = sc->x2_ok ? (sc->x3_ok ? 3 : 2) : 1;
  double c[3] = {0.0, 0.0, 0.0};
  for (int n = nv; n >= 1; n--) {
    double L[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    bool ok = true;
    for (int i = 0; i < n && ok; i++)
      for (int j = 0; j <= i && ok; j++) {
        double v = Gm[i][j];
        for (int k = 0; k < j; k++) v -= L[i][k] * L[j][k];
        if (i == j) {
          ok = v > kLsPivot * Gm[i][i] && isfinite(v);  // (false for NaN)
          if (ok) L[i][i] = sqrt(v);
        } else {
          L[i][j] = v / L[j][j];
        }
      }
    if (!ok) continue;
    double z[3];
    for (int i = 0; i < n; i++) {
      double v = g[i];
      for (int k = 0; k < i; k++) v -= L[i][k] * z[k];
      z[i] = v / L[i][i];
    }
    for (int i = n - 1; i >= 0; i--) {
      double v = z[i];
      for (int k = i + 1; k < n; k++) v -= L[k][i] * c[k];
      c[i] = v / L[i][i];
    }
    break;
  }
  sc->warm_a = c[0] + c[1] + c[2];
  sc->warm_b = -c[1] - 2.0 * c[2];
  sc->warm_c = c[2];
}
__global__ void __launch_bounds__(kNT) k_mgpcg_warm_solve(MgPcgArgs A) { mgpcg_warm_solve_body(A); }

// Cells of one 64x8 PCG tile held by a thread (kPcgRows of them).
struct PqCells {
  double u[kPcgRows][5], po[kPcgRows][5], cf[kPcgRows][5];
  size_t m[kPcgRows];
  bool in[kPcgRows];
};

__device__ __forceinline__ void pq_load(const MgPcgArgs& A, int t, PqCells& c) {
  const Geo& G = A.G;
  const int P = G.P;
#pragma unroll
  for (int k = 0; k < kPcgRows; k++) {
    const int ci = (t % A.nbx) * kTX + (int)(threadIdx.x % kTX);
    const int cj = (t / A.nbx) * kPcgTY + (int)(threadIdx.x / kTX) + k * (kPcgTY / kPcgRows);
    c.in[k] = ci < G.bx && cj < G.by;
    const size_t m = c.in[k] ? (size_t)(cj + 1) * P + (ci + 1) : (size_t)P + 1;
    c.m[k] = m;
    const size_t nb[5] = {m, m - 1, m + 1, m - P, m + P};
#pragma unroll
    for (int q = 0; q < 5; q++) {
      c.u[k][q] = A.u[nb[q]];
      c.po[k][q] = A.p_old[nb[q]];
    }
    c.cf[k][0] = A.dg[m];
    c.cf[k][1] = A.cx[m - 1];
    c.cf[k][2] = A.cx[m];
    c.cf[k][3] = A.cy[m - P];
    c.cf[k][4] = A.cy[m];
  }
}

// p = u + beta p_old (neighbours recomputed), q = M p, partials of p.q.
// The first tile's loads are issued before the reduction that yields beta.
template <int kInstance>
__device__ __forceinline__ void mgpcg_pq_body(const MgPcgArgs& A, int bid, int nb) {
  __shared__ double s_red[8 * kNW];
  DevScalars* sc = A.sc;
  if (sc->done) return;
  const TileRange tr = tile_range(A.nbx * A.nby, bid, nb);
  double pv[3];
  load_partials3(A.part_in, A.nb_in, pv);
  PqCells c;
  if (tr.t0 < tr.t1) pq_load(A, tr.t0, c);
  double tot[3];
  finish_partials3(pv, s_red, tot);
  const double gam = tot[0];
  // Fletcher-Reeves form for a fixed preconditioner; Polak-Ribiere for a
  // varying one: beta = z.(r - r_old)/gamma_old = -(z.q_old)/(p.q)_old
  const double beta = A.first  ? 0.0
                      : A.flex ? -tot[1] / sc->cg_pq[(A.it + 1) & 1]
                               : gam / sc->cg_gam[(A.it + 1) & 1];
  if (bid == 0 && threadIdx.x == 0) sc->cg_gam[A.it & 1] = gam;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = tr.t0; t < tr.t1; t++) {
    if (t > tr.t0) pq_load(A, t, c);
#pragma unroll
    for (int k = 0; k < kPcgRows; k++) {
      if (!c.in[k]) continue;
      double pn[5];
#pragma unroll
      for (int q = 0; q < 5; q++) pn[q] = c.u[k][q] + beta * c.po[k][q];
      const double q = c.cf[k][0] * pn[0] - c.cf[k][1] * pn[1] - c.cf[k][2] * pn[2] -
                       c.cf[k][3] * pn[3] - c.cf[k][4] * pn[4];
      A.p_new[c.m[k]] = pn[0];
      A.q[c.m[k]] = q;
      acc[0] += pn[0] * q;
    }
  }
  write_partials3(acc, A.part_out, s_red, bid);
}

__global__ void __launch_bounds__(kNT) k_mgpcg_pq(MgPcgArgs A) { mgpcg_pq_body<0>(A, blockIdx.x, gridDim.x); }

struct XrCells {
  double x[kPcgRows], p[kPcgRows], r[kPcgRows], q[kPcgRows], vol[kPcgRows];
  size_t m[kPcgRows];
  bool in[kPcgRows];
};

__device__ __forceinline__ void xr_load(const MgPcgArgs& A, int t, XrCells& c) {
  const Geo& G = A.G;
#pragma unroll
  for (int k = 0; k < kPcgRows; k++) {
    const int ci = (t % A.nbx) * kTX + (int)(threadIdx.x % kTX);
    const int cj = (t / A.nbx) * kPcgTY + (int)(threadIdx.x / kTX) + k * (kPcgTY / kPcgRows);
    c.in[k] = ci < G.bx && cj < G.by;
    const int cic = c.in[k] ? ci : 0, cjc = c.in[k] ? cj : 0;
    const size_t m = (size_t)(cjc + 1) * G.P + (cic + 1);
    c.m[k] = m;
    c.x[k] = A.x[m];
    c.p[k] = A.p_new[m];
    c.r[k] = A.r[m];
    c.q[k] = A.q[m];
    c.vol[k] = G.dlx[cic] * G.dly[cjc] * (G.cyl ? G.r[cjc + 1] : 1.0);
  }
}

// alpha = gamma / p.q; x += alpha p; r -= alpha q; partials of |r/V|^2.
// The first tile's loads are issued before the reduction that yields alpha.
template <int kInstance>
__device__ __forceinline__ void mgpcg_xr_body(const MgPcgArgs& A, int bid, int nb) {
  __shared__ double s_red[8 * kNW];
  DevScalars* sc = A.sc;
  if (sc->done) return;
  const TileRange tr = tile_range(A.nbx * A.nby, bid, nb);
  double pv[3];
  load_partials3(A.part_in, A.nb_in, pv);
  XrCells c;
  if (tr.t0 < tr.t1) xr_load(A, tr.t0, c);
  double tot[3];
  finish_partials3(pv, s_red, tot);
  const double alpha = sc->cg_gam[A.it & 1] / tot[0];
  if (bid == 0 && threadIdx.x == 0) sc->cg_pq[A.it & 1] = tot[0];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = tr.t0; t < tr.t1; t++) {
    if (t > tr.t0) xr_load(A, t, c);
#pragma unroll
    for (int k = 0; k < kPcgRows; k++) {
      if (!c.in[k]) continue;
      A.x[c.m[k]] = c.x[k] + alpha * c.p[k];
      const double r = c.r[k] - alpha * c.q[k];
      A.r[c.m[k]] = r;
      const double rn = r / c.vol[k];
      acc[0] += rn * rn;
    }
  }
  write_partials3(acc, A.part_out, s_red, bid);
}

__global__ void __launch_bounds__(kNT) k_mgpcg_xr(MgPcgArgs A) { mgpcg_xr_body<0>(A, blockIdx.x, gridDim.x); }

__global__ void __launch_bounds__(256) k_copy(const double2* __restrict__ a,
                                             double2* __restrict__ b, size_t n) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

__global__ void k_fill(double* p, double v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_eval_fits(FsPhys K, const double* T, double* out, int n) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double t = T[m], v = t * K.ph0, xa = fit_x(v), xs = fit_x_stage(v);
  Rates R = cell_rates(K, v, xs);  // as the stage kernel evaluates them
  double* o = out + 12 * (size_t)m;
  o[0] = mue_of_x(K, xa);
  o[1] = mut_of_x(K, xa);
  o[2] = de_of_x(K, xa);
  o[3] = dt_of_x(K, xa);
  o[4] = R.k_ex;
  o[5] = R.k_ir;
  o[6] = R.nu;
  o[7] = R.k_sc;
  o[8] = R.k_si;
  o[9] = R.beta;
  o[10] = mue_of_x_folded(K, xs);  // the stage kernel's mobility and diffusion
  o[11] = de_of_x_folded(K, xs);
}

__global__ void k_eval_flux(FsPhys K, const double* a, double* out, int n, int dfdph) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double* v = a + 7 * (size_t)m;
  double f, df;
  sg_flux_dfdph(K, v[0], v[1], v[2], v[3], v[4], v[5], v[6], f, df);
  out[m] = dfdph ? df : sg_flux(K, v[0], v[1], 1.0 / v[1], v[2], v[3], v[4], v[5], v[6]);
}

// ---------------------------------------------------------------------------
// Ensemble kernels: each block finds its member and its block index within
// that member's own launch shape, then runs the same body as the solo kernel.
__device__ __forceinline__ bool ens_slot(const EnsMap& m, int& i, int& bid, int& nb) {
  i = m.inst[blockIdx.x];
  if (!m.active[i]) return false;
  bid = (int)blockIdx.x - m.first[i];
  nb = m.nb[i];
  return true;
}

template <class T>
__device__ __forceinline__ const T& ens_arg(const T* t, const EnsSel& s, int i) {
  return t[(size_t)(s.sel ? s.sel[i] : i) * s.stride + s.off];
}

template <int TY>
__global__ void __launch_bounds__(kTX * TY, 4) kb_stage(const StageArgs* t, const FsPhys* K,
                                                    EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) stage_body<1, TY>(ens_arg(t, s, i), K[i], bid, nb);
}
__global__ void __launch_bounds__(kNT) kb_circ(const CircArgs* t, const FsPhys* K, EnsMap m,
                                               EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) circ_body<1>(ens_arg(t, s, i), K[i]);
}
__global__ void __launch_bounds__(kNT) kb_ctrl(const CtrlArgs* t, EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) ctrl_body<1>(ens_arg(t, s, i));
}
__global__ void kb_scalars(DevScalars* sc, const uint8_t* mask, int n, int op) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !mask[i]) return;
  DevScalars* c = sc + i;
  if (op == 0) {
    c->t_m = (double)1e9f;  // main.F90:74
    c->attempts = 0;
  } else if (op == 1) {  // main.F90:83-101
    c->ts = c->ts + 1.0;
    c->t = c->t + c->dt;
    c->Vd_or = c->Vd_mi;
    c->Vd_mi = c->Vd_pl;
  } else {  // main.F90:103-121
    c->dt = c->dt / 2.0;
    c->Vd_pl = c->Vd_or;
    c->Vd_mi = c->Vd_or;
  }
}
__global__ void __launch_bounds__(kSNT, 4) kb_lapl(const LaplArgs* t, const FsPhys* K, EnsMap m,
                                               EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) lapl_body<1>(ens_arg(t, s, i), K[i], bid, nb);
}
__global__ void __launch_bounds__(256) kb_mg_coarsen(const MgArgs* t, EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) mg_coarsen_body<1>(ens_arg(t, s, i), bid, nb);
}
__global__ void __launch_bounds__(1024) kb_mg_coarse_setup(const MgCycleArgs* t, EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) mg_coarse_setup_body<1>(ens_arg(t, s, i));
}
template <bool kCg>
__global__ void __launch_bounds__(kNT) kb_mg_down(const MgArgs* t, EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) mg_down_body<1, kCg>(ens_arg(t, s, i), bid, nb);
}
template <int H>
__global__ void __launch_bounds__(kNT) kb_mg_up(const MgArgs* t, EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) mg_up_body<1, H>(ens_arg(t, s, i), bid, nb);
}
__global__ void __launch_bounds__(1024) kb_mg_coarse_cycle(const MgCycleArgs* t, EnsMap m,
                                                           EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) mg_coarse_cycle_body<1>(ens_arg(t, s, i));
}
__global__ void __launch_bounds__(kNT) kb_mgpcg_init(const MgPcgArgs* t, EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) mgpcg_init_body<1>(ens_arg(t, s, i), bid, nb);
}
// the projected warm start of the members whose start takes it (the others
// return at once): the start kernel's map, then one block per member
__global__ void __launch_bounds__(kNT) kb_mgpcg_warm_ls(const MgPcgArgs* t, EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (!ens_slot(m, i, bid, nb)) return;
  const MgPcgArgs& A = ens_arg(t, s, i);
  if (A.warm && A.ls) mgpcg_warm_ls_body(A, bid, nb);
}
__global__ void __launch_bounds__(kNT) kb_mgpcg_warm_solve(const MgPcgArgs* t, EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (!ens_slot(m, i, bid, nb)) return;
  const MgPcgArgs& A = ens_arg(t, s, i);
  if (A.warm && A.ls) mgpcg_warm_solve_body(A);
}
__global__ void __launch_bounds__(kNT) kb_pcg_finish(const PcgArgs* t, EnsMap m, EnsSel s) {
  int i, bid, nb;
  if (ens_slot(m, i, bid, nb)) pcg_finish_body<1>(ens_arg(t, s, i), bid, nb);
}
__global__ void __launch_bounds__(kNT) kb_pcg_tres(const PcgArgs* t, EnsMap one, EnsSel s,
                                                   const int32_t* nbfin) {
  int i, bid, nb;
  if (ens_slot(one, i, bid, nb)) pcg_tres_body<1>(ens_arg(t, s, i), nbfin[i]);
}

}  // namespace

static inline int nblocks_of(const Geo& G, int* nbx, int* nby) {
  *nbx = (G.bx + kTX - 1) / kTX;
  *nby = (G.by + kSTY - 1) / kSTY;
  return *nbx * *nby;
}

int stage_blocks(int nbx, int by, int ty) { return nbx * ((by + ty - 1) / ty); }
void launch_stage(const StageArgs* a, const FsPhys* K, int nbx, int by, int ty, hipStream_t s,
                  int gated) {
  const dim3 g(stage_blocks(nbx, by, ty));
  if (ty == 4) hipLaunchKernelGGL(k_stage<4>, g, dim3(kTX * 4), 0, s, a, K, gated);
  else if (ty == 16) hipLaunchKernelGGL(k_stage<16>, g, dim3(kTX * 16), 0, s, a, K, gated);
  else hipLaunchKernelGGL(k_stage<8>, g, dim3(kTX * 8), 0, s, a, K, gated);
}
void launch_circ(const CircArgs& a, const FsPhys& K, hipStream_t s) {
  hipLaunchKernelGGL(k_circ, dim3(1), dim3(kNT), 0, s, a, K);
}
void launch_circ_ctrl(const CircArgs& c, const FsPhys& K, const CtrlArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_circ_ctrl, dim3(1), dim3(kNT), 0, s, c, K, a);
}
void launch_accept(DevScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_accept, dim3(1), dim3(kNT), 0, s, sc);
}
void launch_reject(DevScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_reject, dim3(1), dim3(kNT), 0, s, sc);
}
void launch_begin_step(DevScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(1), 0, s, sc);
}
void launch_accept_spec(DevScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_accept_spec, dim3(1), dim3(kNT), 0, s, sc);
}
void launch_publish(const DevScalars* sc, DevScalars* pub, uint32_t* cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, s, sc, pub, cnt);
}
void launch_lapl_assemble(const LaplArgs* a, const FsPhys* K, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_lapl, dim3(nblocks), dim3(kSNT), 0, s, a, K);
}
void launch_pcg_init(const PcgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_init, dim3(capped_grid(a.nbx * a.nby)), dim3(kNT), 0, s, a);
}
void launch_pcg_it(const PcgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_it, dim3(capped_grid(a.nbx * a.nby)), dim3(kNT), 0, s, a);
}
void launch_pcg_finish(const PcgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_finish, dim3(capped_grid(a.nbx * a.nby)), dim3(kNT), 0, s, a);
  hipLaunchKernelGGL(k_pcg_tres, dim3(1), dim3(kNT), 0, s, a, capped_grid(a.nbx * a.nby));
}
static int cells_blocks(int n) { return (n + 255) / 256; }
void launch_mg_coarsen(const MgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_coarsen, dim3(cells_blocks(a.C.nx * a.C.ny)), dim3(256), 0, s, a);
}
int mg_grid(const MgLevel& F) {
  return capped_grid(((F.nx + kGsTX - 1) / kGsTX) * ((F.ny + kGsTY - 1) / kGsTY));
}
void launch_mg_down(const MgArgs& a, hipStream_t s) {
  if (a.cg.on) hipLaunchKernelGGL(k_mg_down<true>, dim3(mg_grid(a.F)), dim3(kNT), 0, s, a);
  else hipLaunchKernelGGL(k_mg_down<false>, dim3(mg_grid(a.F)), dim3(kNT), 0, s, a);
}
void launch_mg_up(const MgArgs& a, hipStream_t s) {
  if (a.cg.on) hipLaunchKernelGGL(k_mg_up<kGsH + 1>, dim3(mg_grid(a.F)), dim3(kNT), 0, s, a);
  else hipLaunchKernelGGL(k_mg_up<kGsH>, dim3(mg_grid(a.F)), dim3(kNT), 0, s, a);
}
void launch_mg_coarse_setup(const MgCycleArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_coarse_setup, dim3(1), dim3(1024), 0, s, a);
}
void launch_mg_coarse_cycle(const MgCycleArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_coarse_cycle, dim3(1), dim3(1024), 0, s, a);
}
void launch_mgpcg_warm_ls(const MgPcgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mgpcg_warm_ls, dim3(capped_grid(a.nbx * a.nby)), dim3(kNT), 0, s, a);
  hipLaunchKernelGGL(k_mgpcg_warm_solve, dim3(1), dim3(kNT), 0, s, a);
}
void launch_mgpcg_init(const MgPcgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mgpcg_init, dim3(capped_grid(a.nbx * a.nby)), dim3(kNT), 0, s, a);
}
void launch_mgpcg_pq(const MgPcgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mgpcg_pq, dim3(capped_grid(a.nbx * a.nby)), dim3(kNT), 0, s, a);
}
void launch_mgpcg_xr(const MgPcgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mgpcg_xr, dim3(capped_grid(a.nbx * a.nby)), dim3(kNT), 0, s, a);
}
void launch_copy(const double* a, double* b, size_t n2, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(256 * 16), dim3(256), 0, s, (const double2*)a, (double2*)b,
                     n2);
}
void launch_fill(double* p, double v, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
}
void launch_eval_fits(const FsPhys& K, const double* T, double* out, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_eval_fits, dim3((n + 255) / 256), dim3(256), 0, s, K, T, out, n);
}
void launch_eval_flux(const FsPhys& K, const double* a, double* out, int n, int dfdph,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_eval_flux, dim3((n + 255) / 256), dim3(256), 0, s, K, a, out, n, dfdph);
}

void launch_ens_stage(const StageArgs* t, const FsPhys* K, const EnsMap& m, const EnsSel& s, int ty,
                      hipStream_t st) {
  if (!m.nblocks) return;
  const dim3 g(m.nblocks);
  if (ty == 4) hipLaunchKernelGGL(kb_stage<4>, g, dim3(kTX * 4), 0, st, t, K, m, s);
  else if (ty == 16) hipLaunchKernelGGL(kb_stage<16>, g, dim3(kTX * 16), 0, st, t, K, m, s);
  else hipLaunchKernelGGL(kb_stage<8>, g, dim3(kTX * 8), 0, st, t, K, m, s);
}
void launch_ens_circ(const CircArgs* t, const FsPhys* K, const EnsMap& m, const EnsSel& s,
                     hipStream_t st) {
  if (m.nblocks) hipLaunchKernelGGL(kb_circ, dim3(m.nblocks), dim3(kNT), 0, st, t, K, m, s);
}
void launch_ens_ctrl(const CtrlArgs* t, const EnsMap& m, const EnsSel& s, hipStream_t st) {
  if (m.nblocks) hipLaunchKernelGGL(kb_ctrl, dim3(m.nblocks), dim3(kNT), 0, st, t, m, s);
}
void launch_ens_scalars(DevScalars* sc, const uint8_t* mask, int n, int op, hipStream_t st) {
  if (n) hipLaunchKernelGGL(kb_scalars, dim3((n + 255) / 256), dim3(256), 0, st, sc, mask, n, op);
}
void launch_ens_lapl(const LaplArgs* t, const FsPhys* K, const EnsMap& m, const EnsSel& s,
                     hipStream_t st) {
  if (m.nblocks) hipLaunchKernelGGL(kb_lapl, dim3(m.nblocks), dim3(kSNT), 0, st, t, K, m, s);
}
void launch_ens_mg_coarsen(const MgArgs* t, const EnsMap& m, const EnsSel& s, hipStream_t st) {
  if (m.nblocks) hipLaunchKernelGGL(kb_mg_coarsen, dim3(m.nblocks), dim3(256), 0, st, t, m, s);
}
void launch_ens_mg_coarse_setup(const MgCycleArgs* t, const EnsMap& m, const EnsSel& s,
                                hipStream_t st) {
  if (m.nblocks) hipLaunchKernelGGL(kb_mg_coarse_setup, dim3(m.nblocks), dim3(1024), 0, st, t, m, s);
}
void launch_ens_mg_down(const MgArgs* t, const EnsMap& m, const EnsSel& s, bool cg,
                        hipStream_t st) {
  if (!m.nblocks) return;
  if (cg) hipLaunchKernelGGL(kb_mg_down<true>, dim3(m.nblocks), dim3(kNT), 0, st, t, m, s);
  else hipLaunchKernelGGL(kb_mg_down<false>, dim3(m.nblocks), dim3(kNT), 0, st, t, m, s);
}
void launch_ens_mg_up(const MgArgs* t, const EnsMap& m, const EnsSel& s, bool cg, hipStream_t st) {
  if (!m.nblocks) return;
  if (cg) hipLaunchKernelGGL(kb_mg_up<kGsH + 1>, dim3(m.nblocks), dim3(kNT), 0, st, t, m, s);
  else hipLaunchKernelGGL(kb_mg_up<kGsH>, dim3(m.nblocks), dim3(kNT), 0, st, t, m, s);
}
void launch_ens_mg_coarse_cycle(const MgCycleArgs* t, const EnsMap& m, const EnsSel& s,
                                hipStream_t st) {
  if (m.nblocks)
    hipLaunchKernelGGL(kb_mg_coarse_cycle, dim3(m.nblocks), dim3(1024), 0, st, t, m, s);
}
void launch_ens_mgpcg_warm_ls(const MgPcgArgs* t, const EnsMap& m, const EnsMap& m_one,
                              const EnsSel& s, hipStream_t st) {
  if (m.nblocks) hipLaunchKernelGGL(kb_mgpcg_warm_ls, dim3(m.nblocks), dim3(kNT), 0, st, t, m, s);
  if (m_one.nblocks)
    hipLaunchKernelGGL(kb_mgpcg_warm_solve, dim3(m_one.nblocks), dim3(kNT), 0, st, t, m_one, s);
}
void launch_ens_mgpcg_init(const MgPcgArgs* t, const EnsMap& m, const EnsSel& s, hipStream_t st) {
  if (m.nblocks) hipLaunchKernelGGL(kb_mgpcg_init, dim3(m.nblocks), dim3(kNT), 0, st, t, m, s);
}
void launch_ens_pcg_finish(const PcgArgs* t, const EnsMap& m, const EnsMap& one, const EnsSel& s,
                           hipStream_t st) {
  if (m.nblocks) hipLaunchKernelGGL(kb_pcg_finish, dim3(m.nblocks), dim3(kNT), 0, st, t, m, s);
  if (one.nblocks)
    hipLaunchKernelGGL(kb_pcg_tres, dim3(one.nblocks), dim3(kNT), 0, st, t, one, s, m.nb);
}
