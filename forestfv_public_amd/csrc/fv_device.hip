// fv_device.hip -- MI355X (gfx950) kernels of the explicit finite-volume
// update and the C ABI declared in include/forestfv_dev.h.
//
// Data in HBM (n leaves in DFS order):
//   S, SN          [NEQ][n]  FP64 structure of arrays (coalesced per field)
//   SF             [6][NEQ][n] reconstructed face states of the last stage
//   nbr            [6][n]    neighbour leaf / fine-list row / -bc code
//   info           [n]       level | SI[0]!=0 | 3-bit face kind per face
//   fine4          [nfine][4] four finer neighbours of a 1:4 face
// Geometry of axis-aligned cubes comes from per-level tables passed by
// value (Geo), so no per-face geometry is streamed.
//
// Stage = recon (per leaf: 6 MUSCL face states + relaxation, m4_faceinterp_V1)
//       + flux_rk (per leaf: residual by gathering its own faces in face
//         order 0..5, RK stage, relaxation).  A face shared by two
//         same-level leaves outside interface layers is evaluated by the
//         later leaf exactly as the earlier one evaluates it (its frame, its
//         left state) and negated, which is the reference's mirror reuse
//         (m4_residual.c:62-85) without flags or atomics: the result is
//         deterministic and bit-identical to the sequential loop.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <type_traits>
#include "forestfv_dev.h"
#include "fv_internal.h"
#include "fv_fast.h"
#include <hipcub/hipcub.hpp>

using namespace fvk;

#define FV_BLOCK 256

// ---------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------

// gather of the MUSCL stencil value across face f (m4_facevalues_MUSCL.c:3-44)
template <int NM>
FV_DEV void stencil_value(const double *__restrict__ S, int n, int i, int kind, int nb,
                          const int *__restrict__ fine4, const double *Sc, const double *Araw, double a,
                          double *X) {
  using L = Lay<NM>;
  if (kind == FV_FK_BOUNDARY) {
    // ghost = own state with the clipped volume fraction (m4_facevalues.c:48-80),
    // clipped from the raw cell value as the reference does
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) X[v] = Sc[v];
    double at[NM], sum = 0.0;
#pragma unroll
    for (int k = 0; k < L::N3; ++k) {
      at[k] = Araw[k];
      if (at[k] <= a) at[k] = a;
      else if (at[k] >= (1.0 - (a))) at[k] = 1.0 - (a);
      sum += at[k];
    }
    if (sum > (1.0 - (a))) {
      double dif = sum - (1.0 - (a));
#pragma unroll
      for (int j = 0; j < L::N3; ++j) { double per = at[j] / sum; at[j] = at[j] - per * dif; }
    }
#pragma unroll
    for (int k = 0; k < L::N3; ++k) X[L::I3 + k] = at[k];
  } else if (kind == FV_FK_FINE) {
    const int4 q = *reinterpret_cast<const int4 *>(&fine4[4 * nb]);
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) {
      const double *Sv = S + (size_t)v * n;
      double acc = 0.0;
      acc += Sv[q.x] / ((double)4);
      acc += Sv[q.y] / ((double)4);
      acc += Sv[q.z] / ((double)4);
      acc += Sv[q.w] / ((double)4);
      X[v] = acc;
    }
  } else {
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) X[v] = S[(size_t)v * n + nb];
  }
  avf_limit<NM, double>(a, X);
}

FV_DEV void dist_of(const Geo &G, int kind, int lvl, double &dcf, double &dcc) {
  if (kind == FV_FK_SAME || kind == FV_FK_MIRROR) { dcf = G.dcf[0][lvl]; dcc = G.dcc[0][lvl]; }
  else if (kind == FV_FK_FINE) { dcf = G.dcf[1][lvl]; dcc = G.dcc[1][lvl]; }
  else if (kind == FV_FK_COARSE) { dcf = G.dcf[2][lvl]; dcc = G.dcc[2][lvl]; }
  else { dcf = 0.0; dcc = 0.0; }
}

// m4_faceinterp_V1 (m4_faceinterp.c:59-88) for one leaf per thread
template <int NM>
__global__ __launch_bounds__(FV_BLOCK) void k_recon(int n, Params P, Geo G, const double *__restrict__ S,
                                                    const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                    const int *__restrict__ fine4, const uint8_t *__restrict__ lim,
                                                    double *__restrict__ SF) {
  using L = Lay<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = P.amin / 100.0;
  const uint32_t inf = info[i];
  const int lvl = cell_level(inf);
  const int limiter = (P.limiter == 2) ? (int)lim[i] : P.limiter;
  double Sc[L::NEQ], Araw[NM];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sc[v] = S[(size_t)v * n + i];
#pragma unroll
  for (int k = 0; k < L::N3; k++) Araw[k] = Sc[L::I3 + k];
  avf_limit<NM, double>(a, Sc);
  // axis pairs (f, fo): (0,2), (1,3), (4,5); both faces of a pair share the two stencil values
#pragma unroll 1
  for (int ax = 0; ax < 3; ax++) {
    const int fa = (ax == 0) ? 0 : (ax == 1 ? 1 : 4);
    const int fb = FCOPP_D[fa];
    const int ka = face_kind(inf, fa), kb = face_kind(inf, fb);
    double Xa[L::NEQ], Xb[L::NEQ];
    stencil_value<NM>(S, n, i, ka, nbr[(size_t)fa * n + i], fine4, Sc, Araw, a, Xa);
    stencil_value<NM>(S, n, i, kb, nbr[(size_t)fb * n + i], fine4, Sc, Araw, a, Xb);
    double dcf_a, dcc_a, dcf_b, dcc_b;
    dist_of(G, ka, lvl, dcf_a, dcc_a);
    dist_of(G, kb, lvl, dcf_b, dcc_b);
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const int f = side ? fb : fa;
      const int kf = side ? kb : ka;
      const double *Xf = side ? Xb : Xa;
      const double *Xk = side ? Xa : Xb;
      const double dcf = side ? dcf_b : dcf_a;
      const double dp = side ? dcc_b : dcc_a;
      const double dm = side ? dcc_a : dcc_b;
      double F[L::NEQ];
      if (P.primtv) {
        recon_prim_face<NM>(P, a, limiter, kf == FV_FK_BOUNDARY, Sc, Xf, Xk, dcf, dp, dm, F);
      } else {
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) {
          double df = Xf[v] - Sc[v];
          double db = Xk[v] - Sc[v];
          double g = 0.0;
          if (kf != FV_FK_BOUNDARY) g = limited_slope<double>(limiter, df, db, dp, dm);
          F[v] = Sc[v] + dcf * g;
        }
        vf_limit_face<NM, double>(a, F);
      }
      relax<NM>(P, F);
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) SF[((size_t)f * L::NEQ + v) * n + i] = F[v];
    }
  }
}

template <int NM>
FV_DEV void load_face(const double *__restrict__ SF, int n, int cell, int f, double *X) {
  using L = Lay<NM>;
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) X[v] = SF[((size_t)f * L::NEQ + v) * n + cell];
}

// RK.c:17-45
FV_DEV double rk_combine(int cs, double dt, double Vol, double SN, double S, double RHS) {
  switch (cs) {
    case 1: return SN - dt * (1.0 / Vol) * RHS;
    case 2: return (3.0 / 4.0) * SN + (1.0 / 4.0) * S - (1.0 / 4.0) * dt * RHS / Vol;
    case 3: return (1.0 / 3.0) * SN + (2.0 / 3.0) * S - (2.0 / 3.0) * dt * RHS / Vol;
    default: return (1.0 / 2.0) * SN + (1.0 / 2.0) * S - (1.0 / 2.0) * dt * RHS / Vol;
  }
}

// m4_residual (m4_residual.c:7-197) + RK(irk) + m4_relaxation_fluid for one leaf per thread
template <int NM>
__global__ __launch_bounds__(FV_BLOCK) void k_flux_rk(int n, Params P, Geo G, int cs,
                                                      double *__restrict__ S, double *__restrict__ SN,
                                                      const double *__restrict__ SF, const int *__restrict__ nbr,
                                                      const uint32_t *__restrict__ info, const int *__restrict__ fine4,
                                                      int *__restrict__ nanflag) {
  using L = Lay<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t inf = info[i];
  const int lvl = cell_level(inf);
  double Sc[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sc[v] = S[(size_t)v * n + i];
  // cell state for the non-conservative sources (m4_residual.c:29-42)
  Sol<NM, double> sc;
  fill_sol<NM, double>(P, Sc, sc);
  double src_avf[NM], src_st[NM];
#pragma unroll
  for (int k = 0; k < L::N3; ++k) src_avf[k] = sc.avf[k];
#pragma unroll
  for (int k = 0; k < NM; ++k) src_st[k] = -sc.avf[k] * sc.p[k];

  double RHS[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) RHS[v] = 0.0;

#pragma unroll 1
  for (int f = 0; f < 6; f++) {
    const int kind = face_kind(inf, f);
    const int nb = nbr[(size_t)f * n + i];
    const int fo = FCOPP_D[f];
    double ff[L::NEQ], uf0 = 0.0, uf1 = 0.0, uf2 = 0.0;
    double Area_f;
    if (kind == FV_FK_MIRROR) {
      // the earlier leaf nb computes this face from its side; take the negation
      double X[L::NEQ];
      Sol<NM, double> sL, sR;
      load_face<NM>(SF, n, nb, fo, X);
      fill_sol<NM, double>(P, X, sL);
      load_face<NM>(SF, n, i, f, X);
      fill_sol<NM, double>(P, X, sR);
      FaceFlux<NM, double> r;
      hllc<NM, double>(P, G.frame[fo], 0, sL, sR, r);
      const double A = G.area[lvl];
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) { double t = 0.0; t += A * r.f[v]; ff[v] = -t; }
      double t0 = 0.0, t1 = 0.0, t2 = 0.0;
      t0 += r.un / ((double)1); t1 += r.vn / ((double)1); t2 += r.wn / ((double)1);
      if (P.fix & 1) { uf0 = t0; uf1 = t1; uf2 = t2; }   // FV_FIX_NONCONS: the star velocity is a global vector
      else { uf0 = -t0; uf1 = -t1; uf2 = -t2; }          // m4_residual.c:72-74
      Area_f = 0.0; Area_f += G.area[lvl];
    } else {
      const int Nneig = (kind == FV_FK_FINE) ? 4 : 1;
      const double A = (kind == FV_FK_FINE) ? G.area[lvl + 1] : G.area[lvl];
      Area_f = 0.0;
      for (int k = 0; k < Nneig; ++k) Area_f += A;
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) ff[v] = 0.0;
      double X[L::NEQ];
      Sol<NM, double> sL, sR;
      load_face<NM>(SF, n, i, f, X);
      fill_sol<NM, double>(P, X, sL);
#pragma unroll 1
      for (int k = 0; k < Nneig; ++k) {
        int bcode = 0;
        if (kind == FV_FK_BOUNDARY) {
          sR = sL;
          bcode = -nb;
          apply_bc<NM, double>(P, bcode, G.frame[f], sL, sR);
        } else {
          const int j = (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
          load_face<NM>(SF, n, j, fo, X);
          fill_sol<NM, double>(P, X, sR);
        }
        FaceFlux<NM, double> r;
        hllc<NM, double>(P, G.frame[f], bcode, sL, sR, r);
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) ff[v] += A * r.f[v];
        uf0 += r.un / ((double)Nneig);
        uf1 += r.vn / ((double)Nneig);
        uf2 += r.wn / ((double)Nneig);
      }
    }
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) RHS[v] += ff[v];
    const double nx = G.frame[f][0], ny = G.frame[f][1], nz = G.frame[f][2];
    // non-conservative terms (m4_residual.c:154-176); FV_FIX_NONCONS subtracts them
    const bool fixnc = (P.fix & 1) != 0;
#pragma unroll
    for (int k = 0; k < L::N3; ++k) {
      const double t = Area_f * (src_avf[k] * uf0 * nx + src_avf[k] * uf1 * ny + src_avf[k] * uf2 * nz);
      RHS[L::I3 + k] = fixnc ? RHS[L::I3 + k] - t : RHS[L::I3 + k] + t;
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      const double t = Area_f * (src_st[k] * uf0 * nx + src_st[k] * uf1 * ny + src_st[k] * uf2 * nz);
      RHS[L::I4 + k] = fixnc ? RHS[L::I4 + k] - t : RHS[L::I4 + k] + t;
    }
  }
  bool bad = false;
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) bad |= isnan(RHS[v]);
  if (bad) *nanflag = 1;

  const double Vol = G.vol[lvl];
  double Sn[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) {
    double sn;
    if (cs == 1) { sn = Sc[v]; SN[(size_t)v * n + i] = sn; }
    else sn = SN[(size_t)v * n + i];   // S^n is read here, not staged: less LDS, and no staging load waits on it
    Sn[v] = rk_combine(cs, P.dt, Vol, sn, Sc[v], RHS[v]);
  }
  relax<NM>(P, Sn);
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) S[(size_t)v * n + i] = Sn[v];
}

// ---------------------------------------------------------------------
// throughput kernels: plain FP64 (MIXED = false) and mixed precision
// (MIXED = true: FP32 slopes, face records and HLLC; FP64 relaxation,
// pressure recovery and RK accumulation)
// ---------------------------------------------------------------------
#pragma clang fp contract(fast)
// AX: one lane per (leaf, axis) instead of per leaf, for the short list of
// leaves outside blocks, whose launch holds too few waves to hide the latency
// of a lane's three axes in turn; the arithmetic of each face is the same
template <int NM, bool MIXED, bool XR, bool AX>
__global__ __launch_bounds__(FV_BLOCK, 2) void k_recon_fast(int n, Params P, Geo G, const double *__restrict__ S,
                                                         const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                         const int *__restrict__ fine4, const uint8_t *__restrict__ lim,
                                                         double *__restrict__ SF, float *__restrict__ SFm,
                                                         const int *__restrict__ list, int cnt) {
  using L = Lay<NM>;
  using R = typename std::conditional<MIXED, float, double>::type;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (AX ? 3 * cnt : cnt)) return;
  const int li = AX ? t / 3 : t;
  const int ax0 = AX ? t - 3 * li : 0, ax1 = AX ? ax0 + 1 : 3;
  const int i = list ? list[li] : li;   // every leaf, or the leaves of a list
  const fvf::Mat<NM, double> M = fvf::make_mat<NM, double>(P);
  const double a = P.amin / 100.0;
  const uint32_t inf = info[i];
  const int lvl = cell_level(inf);
  const int limiter = (P.limiter == 2) ? (int)lim[i] : P.limiter;
  double Sc[L::NEQ], Araw[NM];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sc[v] = S[(size_t)v * n + i];
#pragma unroll
  for (int k = 0; k < L::N3; k++) Araw[k] = Sc[L::I3 + k];
  avf_limit<NM, double>(a, Sc);
  // the neighbour ids are requested together with the cell state, so each
  // axis waits on one round trip (the stencil states) instead of two
  int nb0 = 0, nb1 = 0, nb2 = 0, nb3 = 0, nb4 = 0, nb5 = 0;
  if constexpr (AX) {
    const int fa = (ax0 == 0) ? 0 : (ax0 == 1 ? 1 : 4);
    nb0 = nbr[(size_t)fa * n + i];
    nb2 = nbr[(size_t)FCOPP_D[fa] * n + i];
  } else {
    nb0 = nbr[i]; nb1 = nbr[(size_t)n + i]; nb2 = nbr[(size_t)2 * n + i];
    nb3 = nbr[(size_t)3 * n + i]; nb4 = nbr[(size_t)4 * n + i]; nb5 = nbr[(size_t)5 * n + i];
  }
#pragma unroll 1
  for (int ax = ax0; ax < ax1; ax++) {
    const int fa = (ax == 0) ? 0 : (ax == 1 ? 1 : 4);
    const int fb = FCOPP_D[fa];
    const int ka = face_kind(inf, fa), kb = face_kind(inf, fb);
    const int na = AX ? nb0 : ((ax == 0) ? nb0 : (ax == 1 ? nb1 : nb4));
    const int nbb = AX ? nb2 : ((ax == 0) ? nb2 : (ax == 1 ? nb3 : nb5));
    double Xa[L::NEQ], Xb[L::NEQ];
    stencil_value<NM>(S, n, i, ka, na, fine4, Sc, Araw, a, Xa);
    stencil_value<NM>(S, n, i, kb, nbb, fine4, Sc, Araw, a, Xb);
    double dcf_a, dcc_a, dcf_b, dcc_b;
    dist_of(G, ka, lvl, dcf_a, dcc_a);
    dist_of(G, kb, lvl, dcf_b, dcc_b);
    const bool bnd = (ka == FV_FK_BOUNDARY) || (kb == FV_FK_BOUNDARY);
    const R ia = bnd ? (R)0 : (R)fvf::rcp(dcc_a), ib = bnd ? (R)0 : (R)fvf::rcp(dcc_b);
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const int f = side ? fb : fa;
      const double *Xf = side ? Xb : Xa;
      const double *Xk = side ? Xa : Xb;
      const double dcf = side ? dcf_b : dcf_a;
      const R idp = side ? ib : ia, idm = side ? ia : ib;
      double F[L::NEQ];
      if (P.primtv) {
        // primitive-variable reconstruction in the reference's arithmetic
        double dcc_f, dcc_k, dcf_k;
        dist_of(G, side ? kb : ka, lvl, dcf_k, dcc_f);
        dist_of(G, side ? ka : kb, lvl, dcf_k, dcc_k);
        recon_prim_face<NM>(P, a, limiter, (side ? kb : ka) == FV_FK_BOUNDARY, Sc, Xf, Xk, dcf, dcc_f, dcc_k, F);
      } else {
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) {
          const R g = bnd ? (R)0 : fvf::slope<R>(limiter, Xf[v] - Sc[v], Xk[v] - Sc[v], idp, idm);
          F[v] = Sc[v] + dcf * (double)g;
        }
        vf_limit_face<NM, double>(a, F);
      }
      if (MIXED) {
        float rec[fvf::Rec<NM>::N];
        fvf::relax_to_record<NM, XR>(P, M, a, F, rec);
#pragma unroll
        for (int q = 0; q < fvf::Rec<NM>::N; q++) SFm[((size_t)f * fvf::Rec<NM>::N + q) * n + i] = rec[q];
      } else {
        fvf::relax_fast<NM, XR>(P, M, F);
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) SF[((size_t)f * L::NEQ + v) * n + i] = F[v];
      }
    }
  }
}

template <int NM, bool MIXED, typename R>
FV_DEV void load_sol(const fvf::Mat<NM, double> &Md, const fvf::Mat<NM, R> &M, double a, const double *__restrict__ SF,
                     const float *__restrict__ SFm, int n, int cell, int f, fvf::SolF<NM, R> &s) {
  if constexpr (MIXED) {
    float rec[fvf::Rec<NM>::N];
#pragma unroll
    for (int q = 0; q < fvf::Rec<NM>::N; q++) rec[q] = SFm[((size_t)f * fvf::Rec<NM>::N + q) * n + cell];
    fvf::fill_rec<NM>(M, rec, s);
  } else {
    double X[Lay<NM>::NEQ];
    load_face<NM>(SF, n, cell, f, X);
    fvf::fill_cons<NM, R>(Md, a, X, s);
  }
}

template <int NM, bool MIXED, bool XR>
__global__ __launch_bounds__(FV_BLOCK) void k_flux_fast(int n, Params P, Geo G, int cs, const double *__restrict__ S,
                                                        double *__restrict__ Sout,
                                                        double *__restrict__ SN, const double *__restrict__ SF,
                                                        const float *__restrict__ SFm, const int *__restrict__ nbr,
                                                        const uint32_t *__restrict__ info, const int *__restrict__ fine4,
                                                        int *__restrict__ nanflag, const int *__restrict__ list, int cnt) {
  using L = Lay<NM>;
  using R = typename std::conditional<MIXED, float, double>::type;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int i = list ? list[t] : t;
  const fvf::Mat<NM, double> Md = fvf::make_mat<NM, double>(P);
  const fvf::Mat<NM, R> M = fvf::make_mat<NM, R>(P);
  const double a = P.amin / 100.0;
  const uint32_t inf = info[i];
  const int lvl = cell_level(inf);
  double Sc[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sc[v] = S[(size_t)v * n + i];
  fvf::SolF<NM, double> sc;
  fvf::fill_cons<NM, double>(Md, a, Sc, sc);
  double src_avf[NM], src_st[NM];
#pragma unroll
  for (int k = 0; k < L::N3; ++k) src_avf[k] = sc.avf[k];
#pragma unroll
  for (int k = 0; k < NM; ++k) src_st[k] = -sc.avf[k] * sc.p[k];
  double RHS[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) RHS[v] = 0.0;
  const bool fixnc = (P.fix & 1) != 0;
#pragma unroll 1
  for (int f = 0; f < 6; f++) {
    const int kind = face_kind(inf, f);
    const int nb = nbr[(size_t)f * n + i];
    const int fo = FCOPP_D[f];
    double ff[L::NEQ], uf0 = 0.0, uf1 = 0.0, uf2 = 0.0, Area_f;
    fvf::SolF<NM, R> sL, sR;
    R fr[9];
    if (kind == FV_FK_MIRROR) {
#pragma unroll
      for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[fo][q];
      load_sol<NM, MIXED, R>(Md, M, a, SF, SFm, n, nb, fo, sL);
      load_sol<NM, MIXED, R>(Md, M, a, SF, SFm, n, i, f, sR);
      R out[L::NEQ], un, vn, wn;
      fvf::hllc<NM, R>(M, P.fix, fr, 0, sL, sR, out, un, vn, wn);
      const double A = G.area[lvl];
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) ff[v] = -(A * (double)out[v]);
      const double sgn = fixnc ? 1.0 : -1.0;
      uf0 = sgn * (double)un; uf1 = sgn * (double)vn; uf2 = sgn * (double)wn;
      Area_f = A;
    } else {
#pragma unroll
      for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[f][q];
      const int Nneig = (kind == FV_FK_FINE) ? 4 : 1;
      const double A = (kind == FV_FK_FINE) ? G.area[lvl + 1] : G.area[lvl];
      Area_f = (double)Nneig * A;
      const double wn_ = (kind == FV_FK_FINE) ? 0.25 : 1.0;
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) ff[v] = 0.0;
      load_sol<NM, MIXED, R>(Md, M, a, SF, SFm, n, i, f, sL);
      for (int k = 0; k < Nneig; ++k) {
        int bcode = 0;
        if (kind == FV_FK_BOUNDARY) {
          sR = sL;
          bcode = -nb;
          fvf::ghost<NM, R, XR>(M, P, bcode, fr, sL, sR);
        } else {
          const int j = (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
          load_sol<NM, MIXED, R>(Md, M, a, SF, SFm, n, j, fo, sR);
        }
        R out[L::NEQ], un, vn, wn;
        fvf::hllc<NM, R>(M, P.fix, fr, bcode, sL, sR, out, un, vn, wn);
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) ff[v] += A * (double)out[v];
        uf0 += wn_ * (double)un;
        uf1 += wn_ * (double)vn;
        uf2 += wn_ * (double)wn;
      }
    }
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) RHS[v] += ff[v];
    const double nx = G.frame[f][0], ny = G.frame[f][1], nz = G.frame[f][2];
    const double und = uf0 * nx + uf1 * ny + uf2 * nz;
#pragma unroll
    for (int k = 0; k < L::N3; ++k) {
      const double tt = Area_f * src_avf[k] * und;
      RHS[L::I3 + k] = fixnc ? RHS[L::I3 + k] - tt : RHS[L::I3 + k] + tt;
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      const double tt = Area_f * src_st[k] * und;
      RHS[L::I4 + k] = fixnc ? RHS[L::I4 + k] - tt : RHS[L::I4 + k] + tt;
    }
  }
  bool bad = false;
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) bad |= isnan(RHS[v]);
  if (bad) *nanflag = 1;
  const double dtv = P.dt * fvf::rcp(G.vol[lvl]);
  const double wa = (cs == 1) ? 0.0 : (cs == 2 ? 0.75 : (cs == 3 ? 1.0 / 3.0 : 0.5));
  const double wb = 1.0 - wa;
  double Sn[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) {
    double sn;
    if (cs == 1) { sn = Sc[v]; SN[(size_t)v * n + i] = sn; }
    else sn = SN[(size_t)v * n + i];   // S^n is read here, not staged: less LDS, and no staging load waits on it
    Sn[v] = wa * sn + wb * (Sc[v] - dtv * RHS[v]);
  }
  fvf::relax_fast<NM, XR>(P, Md, Sn);
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sout[(size_t)v * n + i] = Sn[v];
}


// k_flux_fast with the face loads software-pipelined: the neighbour index of
// face f+2 and the stored face states of face f+1 are requested while face f
// is evaluated, so each face waits on no memory round trip of its own (the
// 1:4 faces still gather their four states in place).  Same arithmetic and
// order as k_flux_fast.
template <int NM, bool MIXED>
struct RawFace {
  static constexpr int N = MIXED ? fvf::Rec<NM>::N : Lay<NM>::NEQ;
  typename std::conditional<MIXED, float, double>::type x[N];
};

template <int NM, bool MIXED>
FV_DEV void load_raw(const double *__restrict__ SF, const float *__restrict__ SFm, int n, int cell, int f,
                     RawFace<NM, MIXED> &r) {
  constexpr int N = RawFace<NM, MIXED>::N;
#pragma unroll
  for (int q = 0; q < N; q++) {
    if constexpr (MIXED) r.x[q] = SFm[((size_t)f * N + q) * n + cell];
    else r.x[q] = SF[((size_t)f * N + q) * n + cell];
  }
}

template <int NM, bool MIXED, typename R>
FV_DEV void raw_sol(const fvf::Mat<NM, double> &Md, const fvf::Mat<NM, R> &M, double a, const RawFace<NM, MIXED> &r,
                    fvf::SolF<NM, R> &s) {
  if constexpr (MIXED) fvf::fill_rec<NM>(M, r.x, s);
  else fvf::fill_cons<NM, R>(Md, a, r.x, s);
}

FV_DEV bool one_neighbour(int kind) {
  return kind == FV_FK_SAME || kind == FV_FK_MIRROR || kind == FV_FK_COARSE;
}

template <int NM, bool MIXED, bool XR>
__global__ __launch_bounds__(FV_BLOCK) void k_flux_pipe(int n, Params P, Geo G, int cs, const double *__restrict__ S,
                                                        double *__restrict__ Sout,
                                                        double *__restrict__ SN, const double *__restrict__ SF,
                                                        const float *__restrict__ SFm, const int *__restrict__ nbr,
                                                        const uint32_t *__restrict__ info, const int *__restrict__ fine4,
                                                        int *__restrict__ nanflag, const int *__restrict__ list, int cnt) {
  using L = Lay<NM>;
  using R = typename std::conditional<MIXED, float, double>::type;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int i = list ? list[t] : t;
  const fvf::Mat<NM, double> Md = fvf::make_mat<NM, double>(P);
  const fvf::Mat<NM, R> M = fvf::make_mat<NM, R>(P);
  const double a = P.amin / 100.0;
  const uint32_t inf = info[i];
  const int lvl = cell_level(inf);
  int nbA = nbr[i], nbB = nbr[(size_t)n + i];
  RawFace<NM, MIXED> ownA, nbrA;
  load_raw<NM, MIXED>(SF, SFm, n, i, 0, ownA);
  if (one_neighbour(face_kind(inf, 0))) load_raw<NM, MIXED>(SF, SFm, n, nbA, FCOPP_D[0], nbrA);
  double Sc[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sc[v] = S[(size_t)v * n + i];
  fvf::SolF<NM, double> sc;
  fvf::fill_cons<NM, double>(Md, a, Sc, sc);
  double src_avf[NM], src_st[NM];
#pragma unroll
  for (int k = 0; k < L::N3; ++k) src_avf[k] = sc.avf[k];
#pragma unroll
  for (int k = 0; k < NM; ++k) src_st[k] = -sc.avf[k] * sc.p[k];
  double RHS[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) RHS[v] = 0.0;
  const bool fixnc = (P.fix & 1) != 0;
#pragma unroll 1
  for (int f = 0; f < 6; f++) {
    // requests for the faces ahead
    int nbC = 0;
    if (f + 2 < 6) nbC = nbr[(size_t)(f + 2) * n + i];
    RawFace<NM, MIXED> ownB, nbrB;
    if (f + 1 < 6) {
      load_raw<NM, MIXED>(SF, SFm, n, i, f + 1, ownB);
      if (one_neighbour(face_kind(inf, f + 1))) load_raw<NM, MIXED>(SF, SFm, n, nbB, FCOPP_D[f + 1], nbrB);
    }
    // face f
    const int kind = face_kind(inf, f);
    const int nb = nbA;
    const int fo = FCOPP_D[f];
    double ff[L::NEQ], uf0 = 0.0, uf1 = 0.0, uf2 = 0.0, Area_f;
    fvf::SolF<NM, R> sL, sR;
    R fr[9];
    if (kind == FV_FK_MIRROR) {
#pragma unroll
      for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[fo][q];
      raw_sol<NM, MIXED, R>(Md, M, a, nbrA, sL);
      raw_sol<NM, MIXED, R>(Md, M, a, ownA, sR);
      R out[L::NEQ], un, vn, wn;
      fvf::hllc<NM, R>(M, P.fix, fr, 0, sL, sR, out, un, vn, wn);
      const double A = G.area[lvl];
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) ff[v] = -(A * (double)out[v]);
      const double sgn = fixnc ? 1.0 : -1.0;
      uf0 = sgn * (double)un; uf1 = sgn * (double)vn; uf2 = sgn * (double)wn;
      Area_f = A;
    } else {
#pragma unroll
      for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[f][q];
      const int Nneig = (kind == FV_FK_FINE) ? 4 : 1;
      const double A = (kind == FV_FK_FINE) ? G.area[lvl + 1] : G.area[lvl];
      Area_f = (double)Nneig * A;
      const double wn_ = (kind == FV_FK_FINE) ? 0.25 : 1.0;
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) ff[v] = 0.0;
      raw_sol<NM, MIXED, R>(Md, M, a, ownA, sL);
      for (int k = 0; k < Nneig; ++k) {
        int bcode = 0;
        if (kind == FV_FK_BOUNDARY) {
          sR = sL;
          bcode = -nb;
          fvf::ghost<NM, R, XR>(M, P, bcode, fr, sL, sR);
        } else if (kind == FV_FK_FINE) {
          load_sol<NM, MIXED, R>(Md, M, a, SF, SFm, n, fine4[4 * nb + k], fo, sR);
        } else {
          raw_sol<NM, MIXED, R>(Md, M, a, nbrA, sR);
        }
        R out[L::NEQ], un, vn, wn;
        fvf::hllc<NM, R>(M, P.fix, fr, bcode, sL, sR, out, un, vn, wn);
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) ff[v] += A * (double)out[v];
        uf0 += wn_ * (double)un;
        uf1 += wn_ * (double)vn;
        uf2 += wn_ * (double)wn;
      }
    }
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) RHS[v] += ff[v];
    const double nx = G.frame[f][0], ny = G.frame[f][1], nz = G.frame[f][2];
    const double und = uf0 * nx + uf1 * ny + uf2 * nz;
#pragma unroll
    for (int k = 0; k < L::N3; ++k) {
      const double tt = Area_f * src_avf[k] * und;
      RHS[L::I3 + k] = fixnc ? RHS[L::I3 + k] - tt : RHS[L::I3 + k] + tt;
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      const double tt = Area_f * src_st[k] * und;
      RHS[L::I4 + k] = fixnc ? RHS[L::I4 + k] - tt : RHS[L::I4 + k] + tt;
    }
    // rotate the pipeline
    ownA = ownB;
    nbrA = nbrB;
    nbA = nbB;
    nbB = nbC;
  }
  bool bad = false;
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) bad |= isnan(RHS[v]);
  if (bad) *nanflag = 1;
  const double dtv = P.dt * fvf::rcp(G.vol[lvl]);
  const double wa = (cs == 1) ? 0.0 : (cs == 2 ? 0.75 : (cs == 3 ? 1.0 / 3.0 : 0.5));
  const double wb = 1.0 - wa;
  double Sn[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) {
    double sn;
    if (cs == 1) { sn = Sc[v]; SN[(size_t)v * n + i] = sn; }
    else sn = SN[(size_t)v * n + i];   // S^n is read here, not staged: less LDS, and no staging load waits on it
    Sn[v] = wa * sn + wb * (Sc[v] - dtv * RHS[v]);
  }
  fvf::relax_fast<NM, XR>(P, Md, Sn);
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sout[(size_t)v * n + i] = Sn[v];
}

// ---------------------------------------------------------------------
// fused stage (plain FP64 and mixed precision): one leaf per thread, no
// face-state array.  A leaf reconstructs its own state at each face and the
// neighbour's state at the shared face from the cell states (the
// neighbour's two stencil values along the face normal are the leaf's own
// state and the second neighbour), so HBM traffic per stage is the state
// read, SN and the new state; neighbour and second-neighbour states come
// from the caches.  Faces are taken in order 0..5 and every face state is
// computed by the same device function as in k_recon_fast, so the result
// equals the two-kernel path.  S_in and S_out are different buffers.
// Measured on the 2.2 M-leaf bubble mesh it is slower than the two-kernel
// path (2.4 against 1.0 ms per FP64 stage): the stage is bound by VALU work
// and latency, not HBM bytes, and the fused form doubles the reconstruction
// work at 2 waves per SIMD (a launch bound that spills about 40 VGPRs).  It
// is kept, off by default, as the no-face-state variant.
// ---------------------------------------------------------------------

// k_recon_fast's face-state body: Sc avf-limited cell state, Xg / Xo the
// stencil values across face g and across its opposite face
template <int NM, bool MIXED, bool XR>
FV_DEV void recon_one(const Params &P, const fvf::Mat<NM, double> &M, const Geo &G, double a, int limiter, int lvl,
                      int kg, int ko, const double *Sc, const double *Xg, const double *Xo, double *F) {
  using L = Lay<NM>;
  using R = typename std::conditional<MIXED, float, double>::type;
  double dcf_g, dcc_g, dcf_o, dcc_o;
  dist_of(G, kg, lvl, dcf_g, dcc_g);
  dist_of(G, ko, lvl, dcf_o, dcc_o);
  const bool bnd = (kg == FV_FK_BOUNDARY) || (ko == FV_FK_BOUNDARY);
  if (P.primtv) {
    recon_prim_face<NM>(P, a, limiter, kg == FV_FK_BOUNDARY, Sc, Xg, Xo, dcf_g, dcc_g, dcc_o, F);
  } else {
    const R idp = bnd ? (R)0 : (R)fvf::rcp(dcc_g), idm = bnd ? (R)0 : (R)fvf::rcp(dcc_o);
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) {
      const R g = bnd ? (R)0 : fvf::slope<R>(limiter, Xg[v] - Sc[v], Xo[v] - Sc[v], idp, idm);
      F[v] = Sc[v] + dcf_g * (double)g;
    }
    vf_limit_face<NM, double>(a, F);
  }
  // the mixed path relaxes while it forms the FP32 record (state_sol)
  if constexpr (!MIXED) fvf::relax_fast<NM, XR>(P, M, F);
}

// face state -> solution used by HLLC (through the FP32 record when MIXED,
// where the state arrives unrelaxed and is relaxed here, as k_recon_fast does)
template <int NM, bool MIXED, typename R, bool XR>
FV_DEV void state_sol(const Params &P, const fvf::Mat<NM, double> &Md, const fvf::Mat<NM, R> &M, double a, double *F,
                      fvf::SolF<NM, R> &s) {
  if constexpr (MIXED) {
    float rec[fvf::Rec<NM>::N];
    fvf::relax_to_record<NM, XR>(P, Md, a, F, rec);
    fvf::fill_rec<NM>(M, rec, s);
  } else {
    fvf::fill_cons<NM, R>(Md, a, F, s);
  }
}

FV_DEV int leaf_limiter(const Params &P, const uint8_t *__restrict__ lim, int j) {
  return (P.limiter == 2) ? (int)lim[j] : P.limiter;
}

// reconstructed state of any leaf j at its face g, from the cell states
template <int NM, bool MIXED, bool XR>
FV_DEV void face_state_of(const Params &P, const fvf::Mat<NM, double> &M, const Geo &G, double a,
                          const double *__restrict__ S, int n, const int *__restrict__ nbr,
                          const uint32_t *__restrict__ info, const int *__restrict__ fine4,
                          const uint8_t *__restrict__ lim, int j, int g, double *F) {
  using L = Lay<NM>;
  double Sc[L::NEQ], Araw[NM], Xg[L::NEQ], Xo[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sc[v] = S[(size_t)v * n + j];
#pragma unroll
  for (int k = 0; k < L::N3; k++) Araw[k] = Sc[L::I3 + k];
  avf_limit<NM, double>(a, Sc);
  const uint32_t inf = info[j];
  const int go = FCOPP_D[g];
  const int kg = face_kind(inf, g), ko = face_kind(inf, go);
  stencil_value<NM>(S, n, j, kg, nbr[(size_t)g * n + j], fine4, Sc, Araw, a, Xg);
  stencil_value<NM>(S, n, j, ko, nbr[(size_t)go * n + j], fine4, Sc, Araw, a, Xo);
  recon_one<NM, MIXED, XR>(P, M, G, a, leaf_limiter(P, lim, j), cell_level(inf), kg, ko, Sc, Xg, Xo, F);
}

// blocks of one XCD (workgroups are dealt to the 8 XCDs round robin) take a
// contiguous range of leaves, so face neighbours share an L2
FV_DEV int xcd_leaf(void) {
  const unsigned per = gridDim.x >> 3;
  const unsigned b = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  return (int)(b * blockDim.x + threadIdx.x);
}

template <int NM, bool MIXED, bool XR>
__global__ __launch_bounds__(FV_BLOCK, 2) void k_stage_fused(int n, Params P, const Geo *__restrict__ Gp, int cs,
                                                          const double *__restrict__ S, double *__restrict__ Sout,
                                                          double *__restrict__ SN, const int *__restrict__ nbr,
                                                          const uint32_t *__restrict__ info,
                                                          const int *__restrict__ fine4,
                                                          const uint8_t *__restrict__ lim, int *__restrict__ nanflag) {
  using L = Lay<NM>;
  using R = typename std::conditional<MIXED, float, double>::type;
  const int i = xcd_leaf();
  if (i >= n) return;
  // the per-level tables are read through a pointer: a by-value copy indexed
  // by the per-leaf level would be placed in scratch memory
  const Geo &G = *Gp;
  const fvf::Mat<NM, double> Md = fvf::make_mat<NM, double>(P);
  const fvf::Mat<NM, R> M = fvf::make_mat<NM, R>(P);
  const double a = P.amin / 100.0;
  const uint32_t inf = info[i];
  const int lvl = cell_level(inf);
  const int limiter = leaf_limiter(P, lim, i);
  double Sc[L::NEQ], Araw[NM];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sc[v] = S[(size_t)v * n + i];
#pragma unroll
  for (int k = 0; k < L::N3; k++) Araw[k] = Sc[L::I3 + k];
  double src_avf[NM], src_st[NM];
  {
    fvf::SolF<NM, double> sc;
    fvf::fill_cons<NM, double>(Md, a, Sc, sc);
#pragma unroll
    for (int k = 0; k < L::N3; ++k) src_avf[k] = sc.avf[k];
#pragma unroll
    for (int k = 0; k < NM; ++k) src_st[k] = -sc.avf[k] * sc.p[k];
  }
  avf_limit<NM, double>(a, Sc);
  double RHS[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) RHS[v] = 0.0;
  const bool fixnc = (P.fix & 1) != 0;
#pragma unroll 1
  for (int f = 0; f < 6; f++) {
    const int fo = FCOPP_D[f];
    const int kind = face_kind(inf, f), kopp = face_kind(inf, fo);
    const int nb = nbr[(size_t)f * n + i];
    double Xf[L::NEQ];
    stencil_value<NM>(S, n, i, kind, nb, fine4, Sc, Araw, a, Xf);
    double ff[L::NEQ], uf0 = 0.0, uf1 = 0.0, uf2 = 0.0, Area_f;
    R fr[9];
    if (kind == FV_FK_SAME || kind == FV_FK_MIRROR) {
      // same-level neighbour j: its cell state is Xf, its stencil value
      // towards this leaf is Sc, the other one is the second neighbour
      fvf::SolF<NM, R> sA, sB;
      {
        const uint32_t infj = info[nb];
        const int kj = face_kind(infj, f), kjo = face_kind(infj, fo);
        double X2[L::NEQ], Arj[NM] = {}, Fnb[L::NEQ];
        if (kj == FV_FK_BOUNDARY) {
#pragma unroll
          for (int k = 0; k < L::N3; k++) Arj[k] = S[(size_t)(L::I3 + k) * n + nb];
        }
        stencil_value<NM>(S, n, nb, kj, nbr[(size_t)f * n + nb], fine4, Xf, Arj, a, X2);
        recon_one<NM, MIXED, XR>(P, Md, G, a, leaf_limiter(P, lim, nb), lvl, kjo, kj, Xf, Sc, X2, Fnb);
        state_sol<NM, MIXED, R, XR>(P, Md, M, a, Fnb, sA);
      }
      {
        double Xo[L::NEQ], Fown[L::NEQ];
        stencil_value<NM>(S, n, i, kopp, nbr[(size_t)fo * n + i], fine4, Sc, Araw, a, Xo);
        recon_one<NM, MIXED, XR>(P, Md, G, a, limiter, lvl, kind, kopp, Sc, Xf, Xo, Fown);
        state_sol<NM, MIXED, R, XR>(P, Md, M, a, Fown, sB);
      }
      R out[L::NEQ], un, vn, wn;
      const double A = G.area[lvl];
      Area_f = A;
      if (kind == FV_FK_MIRROR) {
        // evaluated exactly as the neighbour evaluates it, then negated
#pragma unroll
        for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[fo][q];
        fvf::hllc<NM, R>(M, P.fix, fr, 0, sA, sB, out, un, vn, wn);
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) ff[v] = -(A * (double)out[v]);
        const double sgn = fixnc ? 1.0 : -1.0;
        uf0 = sgn * (double)un; uf1 = sgn * (double)vn; uf2 = sgn * (double)wn;
      } else {
        // same accumulation form as the one-neighbour case of k_flux_fast
#pragma unroll
        for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[f][q];
        fvf::hllc<NM, R>(M, P.fix, fr, 0, sB, sA, out, un, vn, wn);
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) { ff[v] = 0.0; ff[v] += A * (double)out[v]; }
        uf0 += 1.0 * (double)un;
        uf1 += 1.0 * (double)vn;
        uf2 += 1.0 * (double)wn;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[f][q];
      const int Nneig = (kind == FV_FK_FINE) ? 4 : 1;
      const double A = (kind == FV_FK_FINE) ? G.area[lvl + 1] : G.area[lvl];
      Area_f = (double)Nneig * A;
      const double wn_ = (kind == FV_FK_FINE) ? 0.25 : 1.0;
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) ff[v] = 0.0;
      fvf::SolF<NM, R> sL, sR;
      {
        double Xo[L::NEQ], Fown[L::NEQ];
        stencil_value<NM>(S, n, i, kopp, nbr[(size_t)fo * n + i], fine4, Sc, Araw, a, Xo);
        recon_one<NM, MIXED, XR>(P, Md, G, a, limiter, lvl, kind, kopp, Sc, Xf, Xo, Fown);
        state_sol<NM, MIXED, R, XR>(P, Md, M, a, Fown, sL);
      }
      for (int k = 0; k < Nneig; ++k) {
        int bcode = 0;
        if (kind == FV_FK_BOUNDARY) {
          sR = sL;
          bcode = -nb;
          fvf::ghost<NM, R, XR>(M, P, bcode, fr, sL, sR);
        } else {
          const int j = (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
          double Fnb[L::NEQ];
          face_state_of<NM, MIXED, XR>(P, Md, G, a, S, n, nbr, info, fine4, lim, j, fo, Fnb);
          state_sol<NM, MIXED, R, XR>(P, Md, M, a, Fnb, sR);
        }
        R out[L::NEQ], un, vn, wn;
        fvf::hllc<NM, R>(M, P.fix, fr, bcode, sL, sR, out, un, vn, wn);
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) ff[v] += A * (double)out[v];
        uf0 += wn_ * (double)un;
        uf1 += wn_ * (double)vn;
        uf2 += wn_ * (double)wn;
      }
    }
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) RHS[v] += ff[v];
    const double nx = G.frame[f][0], ny = G.frame[f][1], nz = G.frame[f][2];
    const double und = uf0 * nx + uf1 * ny + uf2 * nz;
#pragma unroll
    for (int k = 0; k < L::N3; ++k) {
      const double tt = Area_f * src_avf[k] * und;
      RHS[L::I3 + k] = fixnc ? RHS[L::I3 + k] - tt : RHS[L::I3 + k] + tt;
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      const double tt = Area_f * src_st[k] * und;
      RHS[L::I4 + k] = fixnc ? RHS[L::I4 + k] - tt : RHS[L::I4 + k] + tt;
    }
  }
  bool bad = false;
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) bad |= isnan(RHS[v]);
  if (bad) *nanflag = 1;
  const double dtv = P.dt * fvf::rcp(G.vol[lvl]);
  const double wa = (cs == 1) ? 0.0 : (cs == 2 ? 0.75 : (cs == 3 ? 1.0 / 3.0 : 0.5));
  const double wb = 1.0 - wa;
  double Sn[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) {
    const double s0 = S[(size_t)v * n + i];
    double sn;
    if (cs == 1) { sn = s0; SN[(size_t)v * n + i] = sn; }
    else sn = SN[(size_t)v * n + i];   // S^n is read here, not staged: less LDS, and no staging load waits on it
    Sn[v] = wa * sn + wb * (s0 - dtv * RHS[v]);
  }
  fvf::relax_fast<NM, XR>(P, Md, Sn);
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sout[(size_t)v * n + i] = Sn[v];
}

// ---------------------------------------------------------------------
// block stage (plain FP64 and mixed precision): one workgroup per brick,
// 64 consecutive leaves in DFS order that form a 4x4x4 block of one level
// and whose faces on the block surface have a same-level neighbour or a
// domain boundary.  The cell states, the states across the surface and the
// stencil values beyond those are staged in LDS once.  Then one lane per face
// (144 inside the brick, 96 on its surface) reconstructs both face states of
// its face from the staged states and runs HLLC, so every face is evaluated
// once and no face state goes through HBM except the ones a non-brick
// neighbour's flux reads.  Each cell then gathers its six faces in face
// order and applies the RK stage and the relaxation.  A face takes one
// evaluation, from the side with the lower position along the face normal
// (the brick cell's side on the surface), which equals the reference's
// evaluation by either side (computed once by the earlier leaf, or by both
// for interface leaves, m4_residual.c:62-103) to rounding: the HLLC flux is
// antisymmetric under exchange of the sides with the normal reversed.
// ---------------------------------------------------------------------
constexpr int BR_THREADS = 256;
constexpr int BR_SLOTS = 96;              // surface slots: 6 faces x 16 cells
constexpr int BR_TASKS = BR_SLOTS + 144;  // faces: one per surface slot + 3 x 48 inside the brick

// a position in the brick; components are selected with compares rather than
// array indexing, so that no private array is left in scratch memory
struct BrPos {
  int x, y, z;
};
FV_DEV int br_get(const BrPos &p, int ax) { return ax == 0 ? p.x : (ax == 1 ? p.y : p.z); }
FV_DEV void br_add(BrPos &p, int ax, int d) {
  p.x += (ax == 0) ? d : 0;
  p.y += (ax == 1) ? d : 0;
  p.z += (ax == 2) ? d : 0;
}
// local index of a position: Morton, x fastest (the DFS order)
FV_DEV int br_index(const BrPos &p) {
  return (p.x & 1) | ((p.y & 1) << 1) | ((p.z & 1) << 2) | ((p.x >> 1) << 3) | ((p.y >> 1) << 4) | ((p.z >> 1) << 5);
}
FV_DEV BrPos br_pos(int l) {
  return BrPos{(l & 1) | (((l >> 3) & 1) << 1), ((l >> 1) & 1) | (((l >> 4) & 1) << 1),
               ((l >> 2) & 1) | (((l >> 5) & 1) << 1)};
}
// axis and direction of face f (f0 y-, f1 x+, f2 y+, f3 x-, f4 z-, f5 z+)
FV_DEV int br_axis(int f) { return (f == 1 || f == 3) ? 0 : ((f == 0 || f == 2) ? 1 : 2); }
FV_DEV int br_plus(int f) { return f == 1 || f == 2 || f == 5; }
FV_DEV int br_plus_face(int ax) { return ax == 0 ? 1 : (ax == 1 ? 2 : 5); }
// position of a cell in the 4x4 sheet across the axis: the other two coordinates
FV_DEV int br_sheet(int ax, const BrPos &p) {
  return ax == 0 ? p.y + 4 * p.z : (ax == 1 ? p.x + 4 * p.z : p.x + 4 * p.y);
}
FV_DEV BrPos br_sheet_pos(int ax, int k, int along) {
  const int u = k & 3, w = k >> 2;
  return ax == 0 ? BrPos{along, u, w} : (ax == 1 ? BrPos{u, along, w} : BrPos{u, w, along});
}
// cell of surface slot s = 16 f + k
FV_DEV int br_slot_cell(int s) {
  const int f = s >> 4;
  return br_index(br_sheet_pos(br_axis(f), s & 15, br_plus(f) ? 3 : 0));
}

// XCD-aware brick order: workgroups are dealt to the 8 XCDs round robin, so
// workgroup w takes brick (w mod 8) * per + w / 8 and each XCD walks a
// contiguous range of bricks, whose surface neighbours then share its L2
FV_DEV int br_of_block(int nb) {
  const unsigned per = gridDim.x >> 3;
  const unsigned b = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  return (int)b < nb ? (int)b : -1;
}

template <int NM, bool MIXED = false>
struct BrickLds {
  using L = Lay<NM>;
  // the face fluxes of the tasks: A F in FP64, or in mixed precision the FP32
  // HLLC output F itself (A F is formed in FP64 when a cell gathers it, the
  // same product the FP64 store would hold), which halves this array
  using FT = typename std::conditional<MIXED, float, double>::type;
  // staged cell states [v][c], avf-limited as stencil values and the cell's
  // own reconstruction base use them: own cells c < 64, the neighbour across
  // surface slot s at 64 + s (that neighbour's stencil value beyond it stays in
  // the registers of the slot's lane)
  static constexpr int NC = 64 + BR_SLOTS;
  double st[L::NEQ * NC];
  // the own cells' raw volume fractions and source coefficients (alpha_k, then
  // -alpha_k p_k); the mixed kernel forms them in the cell phase from the raw
  // state read there, which with fu in FP32 and idmh formed where it is used
  // leaves 22.4 KB of LDS, room for a seventh workgroup
  double araw[MIXED ? 1 : L::N3 * 64];
  double src[MIXED ? 1 : (L::N3 + NM) * 64];
  FT ff[L::NEQ * BR_TASKS];             // face fluxes per face task
  FT fu[BR_TASKS];                      // u*.n per face task
  uint32_t info[64];
  uint32_t infoh[BR_SLOTS];             // info word of the neighbour across each slot
  int nb[BR_SLOTS];                     // that neighbour (leaf index, or -BC code)
  uint8_t need[BR_SLOTS];               // it is no brick leaf: its flux reads this face state from HBM
  uint8_t lim[64 + BR_SLOTS];           // per-leaf limiter of the own cells and the slot neighbours
  double idmh[MIXED ? 1 : BR_SLOTS];    // 1/dcc of the slot neighbour's face beyond it (any kind)
};

template <int NM, bool MIXED>
FV_DEV void br_state(const BrickLds<NM, MIXED> &B, int c, double *X) {
#pragma unroll
  for (int v = 0; v < Lay<NM>::NEQ; v++) X[v] = B.st[v * BrickLds<NM, MIXED>::NC + c];
}

// face state (FP64 vector, relaxed) or FP32 record of a same-level cell at a
// face, from its staged state (column c) and the stencil values across the
// face (column cg) and across the opposite face (column co, or Xo in registers
// when xo_reg): k_recon_fast's arithmetic
// with the brick's one level (dcf, 1/dcc), and zero slope when either face is
// a domain boundary (bnd)
template <int NM, bool MIXED, bool XR>
FV_DEV void br_face_state(const Params &P, const fvf::Mat<NM, double> &Md, double a, const BrickLds<NM, MIXED> &B, int c,
                          int cg, int co, const double *Xo, bool xo_reg, int limiter, bool bnd, double dcf,
                          double idp, double idm, double *F, float *rec) {
  using L = Lay<NM>;
  using R = typename std::conditional<MIXED, float, double>::type;
  br_state<NM, MIXED>(B, c, F);
  if (!bnd) {
    const R ip = (R)idp, im = (R)idm;
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) {
      const double sc = F[v];
      const double xs = B.st[v * BrickLds<NM, MIXED>::NC + co];
      const double xg = B.st[v * BrickLds<NM, MIXED>::NC + cg], xo = xo_reg ? Xo[v] : xs;
      const R g = fvf::slope<R>(limiter, xg - sc, xo - sc, ip, im);
      F[v] = sc + dcf * (double)g;
    }
  }
  vf_limit_face<NM, double>(a, F);
  if constexpr (MIXED) fvf::relax_to_record<NM, XR>(P, Md, a, F, rec);
  else fvf::relax_fast<NM, XR>(P, Md, F);
}

template <int NM, bool MIXED, typename R>
FV_DEV void br_sol(const fvf::Mat<NM, double> &Md, const fvf::Mat<NM, R> &M, double a, const double *F,
                   const float *rec, fvf::SolF<NM, R> &s) {
  if constexpr (MIXED) fvf::fill_rec<NM>(M, rec, s);
  else fvf::fill_cons<NM, R>(Md, a, F, s);
}

template <int NM, bool MIXED>
FV_DEV void br_write_sf(int n, int i, int f, const double *F, const float *rec, double *__restrict__ SF,
                        float *__restrict__ SFm) {
  if constexpr (MIXED) {
#pragma unroll
    for (int q = 0; q < fvf::Rec<NM>::N; q++) SFm[((size_t)f * fvf::Rec<NM>::N + q) * n + i] = rec[q];
  } else {
#pragma unroll
    for (int v = 0; v < Lay<NM>::NEQ; v++) SF[((size_t)f * Lay<NM>::NEQ + v) * n + i] = F[v];
  }
}

// 1/dcc and dcf of face f of a brick cell with info word inf on a graded
// brick: the face kind's distances (same-level, 1:4 or 2:1 tables); a domain
// boundary takes the same-level values (its slope is zero)
FV_DEV void br_face_dist(const Geo &G, uint32_t inf, int f, int lvl, double &dcf, double &idc) {
  const int k = face_kind(inf, f);
  double dcc;
  dist_of(G, (k == FV_FK_BOUNDARY) ? FV_FK_SAME : k, lvl, dcf, dcc);
  idc = fvf::rcp(dcc);
}

// GR: a graded brick, whose surface faces may also be 2:1 (the brick cell's
// neighbour one level coarser) or 1:4 (four neighbours one level finer).  The
// neighbours across those faces are leaves outside bricks, so their face
// states are in SF (written by k_recon_fast before this kernel); the brick
// cell's own face state is reconstructed here with the face kind's distances
// and its stencil value across the face (the coarser leaf, or the mean of the
// four finer ones), and the face is evaluated as k_flux_fast evaluates it from
// this cell: one HLLC against the coarser leaf's face state, or four with the
// finer area against the finer leaves' face states.
template <int NM, bool MIXED, bool XR, bool GR>
__global__ __launch_bounds__(BR_THREADS, GR ? 3 : ((MIXED && !XR) ? 7 : 4)) void k_stage_brick(int n, Params P, const Geo *__restrict__ Gp, int cs,
                                                               const int *__restrict__ bricks, int nbricks,
                                                               const double *__restrict__ S, double *__restrict__ Sout,
                                                               double *__restrict__ SN, const int *__restrict__ nbr,
                                                               const uint32_t *__restrict__ info,
                                                               const int *__restrict__ fine4,
                                                               const uint8_t *__restrict__ lim,
                                                               const uint8_t *__restrict__ inbrick,
                                                               double *__restrict__ SF, float *__restrict__ SFm,
                                                               int *__restrict__ nanflag) {
  using L = Lay<NM>;
  using R = typename std::conditional<MIXED, float, double>::type;
  using LD = BrickLds<NM, MIXED>;
  constexpr int NC = LD::NC;
  __shared__ LD B;
  const int br = br_of_block(nbricks);
  if (br < 0) return;   // whole workgroups only: no barrier is reached by part of one
  const int b0 = bricks[br];
  const int tid = threadIdx.x;
  const Geo &G = *Gp;
  const fvf::Mat<NM, double> Md = fvf::make_mat<NM, double>(P);
  const double a = P.amin / 100.0;

  // ---- one staging phase: the own states (lanes 96..159,
  // one cell each: raw fractions, source coefficients and the avf-limited
  // state), the info words and limiters (lanes 160..223), and per surface slot
  // (lanes 0..95) the neighbour across it, whose stencil value beyond it the
  // slot's lane keeps in registers for its face task.  No load waits for
  // another lane's.
  double X2[L::NEQ];   // slot lanes: the slot neighbour's stencil value beyond it
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) X2[v] = 0.0;
  if (tid >= 96 && tid < 160) {
    const int l = tid - 96;
    double X[L::NEQ];
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) X[v] = S[(size_t)v * n + b0 + l];
    if constexpr (!MIXED) {
#pragma unroll
      for (int k = 0; k < L::N3; k++) B.araw[k * 64 + l] = X[L::I3 + k];
      // the sources of the cell update take the raw state (k_flux_fast's fill_cons of S)
      fvf::SolF<NM, double> sc;
      fvf::fill_cons<NM, double>(Md, a, X, sc);
#pragma unroll
      for (int k = 0; k < L::N3; ++k) B.src[k * 64 + l] = sc.avf[k];
#pragma unroll
      for (int k = 0; k < NM; ++k) B.src[(L::N3 + k) * 64 + l] = -sc.avf[k] * sc.p[k];
    }
    avf_limit<NM, double>(a, X);
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) B.st[v * NC + l] = X[v];
  } else if (tid >= 160 && tid < 224) {
    const int l = tid - 160;
    B.info[l] = info[b0 + l];
    B.lim[l] = (uint8_t)leaf_limiter(P, lim, b0 + l);
  } else if (tid < BR_SLOTS) {
    const int s = tid, f = s >> 4, i = b0 + br_slot_cell(s);
    const int nb = nbr[(size_t)f * n + i];
    B.nb[s] = nb;
    const int kf = face_kind(info[i], f);
    if (GR && (kf == FV_FK_COARSE || kf == FV_FK_FINE)) {
      // the brick cell's stencil value across a 2:1 or 1:4 face; the face state
      // beyond comes from SF, and the brick cell's own one goes there for the
      // neighbours' fluxes
      double Xs[L::NEQ], Z[NM] = {};
      stencil_value<NM>(S, n, i, kf, nb, fine4, X2, Z, a, Xs);
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) B.st[v * NC + 64 + s] = Xs[v];
      B.need[s] = 1;
      B.infoh[s] = 0;
      B.lim[64 + s] = 0;
      if constexpr (!MIXED) B.idmh[s] = 0.0;
    } else if (kf == FV_FK_BOUNDARY) {
      // no state across a boundary face: its slope is zero and its ghost comes from the face state
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) B.st[v * NC + 64 + s] = 0.0;
      B.need[s] = 0;
      B.infoh[s] = 0;
      B.lim[64 + s] = 0;
      if constexpr (!MIXED) B.idmh[s] = 0.0;
    } else {
      const uint32_t infj = info[nb];
      B.infoh[s] = infj;
      B.need[s] = inbrick[nb] ? 0 : 1;
      B.lim[64 + s] = (uint8_t)leaf_limiter(P, lim, nb);
      double Xj[L::NEQ], Araw[NM] = {};
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) Xj[v] = S[(size_t)v * n + nb];
      const int kj = face_kind(infj, f);
      if (kj == FV_FK_BOUNDARY) {
#pragma unroll
        for (int k = 0; k < L::N3; k++) Araw[k] = Xj[L::I3 + k];
      }
      if constexpr (!MIXED) {
        double dcf_j, dcc_j;
        dist_of(G, kj, cell_level(infj), dcf_j, dcc_j);
        B.idmh[s] = (kj == FV_FK_BOUNDARY) ? 0.0 : fvf::rcp(dcc_j);
      }
      avf_limit<NM, double>(a, Xj);
      stencil_value<NM>(S, n, nb, kj, nbr[(size_t)f * n + nb], fine4, Xj, Araw, a, X2);
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) B.st[v * NC + 64 + s] = Xj[v];
    }
  }
  __syncthreads();
  // the brick's one level: distances and areas are the same for every face
  const int lvl = cell_level(B.info[0]);
  const double dcf = G.dcf[0][lvl], idcc = fvf::rcp(G.dcc[0][lvl]), A = G.area[lvl];

  // ---- faces: one lane per face, 96 across the surface slots and 144 inside
  // the brick; a lane reconstructs both face states of its face and runs HLLC.
  // Surface and inside lanes differ only in the LDS columns and constants they
  // pass, so that a wave holding both kinds runs one instruction stream.
  if (tid < BR_TASKS) {
    const fvf::Mat<NM, R> M = fvf::make_mat<NM, R>(P);
    const bool surf = tid < BR_SLOTS;
    int fL, bcode = 0, lown;
    // left side (outward normal frame[fL]) and right side: the cell column, its
    // stencil value columns across the face and across the opposite face, its
    // limiter, zero-slope flag and 1/dcc beyond it
    int cA, gA, oA, cB, gB, oB, limA, limB;
    bool bndA, bndB, need = false, oreg = false;
    double idmB = idcc;
    // graded bricks: the left face state's distances (dcfA, idpA) and 1/dcc
    // beyond each side (idmA, idmB) follow the kinds of the surface faces;
    // kf21 is a surface face's 2:1 or 1:4 kind (0 otherwise) and nb21 its
    // neighbour or 1:4 row
    double dcfA = dcf, idpA = idcc, idmA = idcc;
    int kf21 = 0, nb21 = 0;
    if (surf) {
      const int s = tid, f = s >> 4, fo = FCOPP_D[f], ax = br_axis(f);
      const int l = br_slot_cell(s);
      const uint32_t inf = B.info[l];
      fL = f;
      lown = l;
      BrPos pos = br_pos(l);
      br_add(pos, ax, br_plus(f) ? -1 : 1);   // the brick cell's neighbour on the other side, inside the brick
      const bool bndf = face_kind(inf, f) == FV_FK_BOUNDARY;
      cA = l; gA = 64 + s; oA = br_index(pos); limA = B.lim[l];
      bndA = bndf || face_kind(inf, fo) == FV_FK_BOUNDARY;
      need = B.need[s] != 0;
      if constexpr (GR) {
        const int kf = face_kind(inf, f);
        if (kf == FV_FK_COARSE || kf == FV_FK_FINE) { kf21 = kf; nb21 = B.nb[s]; }
        br_face_dist(G, inf, f, lvl, dcfA, idpA);
      }
      if (GR && kf21 != 0) {
        // no reconstruction on the far side: its face states are in SF
        cB = cA; gB = gA; oB = oA; limB = limA; bndB = bndA;
      } else if (bndf) {
        // no cell across a domain boundary: the right state is the ghost of the left
        // one, so the right reconstruction repeats the left one on valid data
        bcode = -B.nb[s];
        cB = cA; gB = gA; oB = oA; limB = limA; bndB = bndA;
      } else {
        // the neighbour's face state towards the brick: its stencil value across
        // that face is the brick cell, beyond it the staged second layer
        const uint32_t infj = B.infoh[s];
        cB = 64 + s; gB = l; oB = l; limB = B.lim[64 + s];   // the column beyond is X2
        oreg = true;
        bndB = face_kind(infj, fo) == FV_FK_BOUNDARY || face_kind(infj, f) == FV_FK_BOUNDARY;
        if constexpr (MIXED) {
          const int kj = face_kind(infj, f);
          double dcf_j, dcc_j;
          dist_of(G, kj, cell_level(infj), dcf_j, dcc_j);
          idmB = (kj == FV_FK_BOUNDARY) ? 0.0 : fvf::rcp(dcc_j);
        } else {
          idmB = B.idmh[s];
        }
      }
    } else {
      const int q = tid - BR_SLOTS, ax = q / 48, p = (q % 48) >> 4, k = q & 15;
      const int f = br_plus_face(ax), fo = FCOPP_D[f];
      const int lL = br_index(br_sheet_pos(ax, k, p)), lR = br_index(br_sheet_pos(ax, k, p + 1));
      fL = f;
      lown = lL;
      // the cells before the left one and after the right one along the axis:
      // inside the brick, or the staged neighbour of the - / + surface slot
      cA = lL; gA = lR; oA = (p > 0) ? br_index(br_sheet_pos(ax, k, p - 1)) : 64 + 16 * fo + k; limA = B.lim[lL];
      cB = lR; gB = lL; oB = (p < 2) ? br_index(br_sheet_pos(ax, k, p + 2)) : 64 + 16 * f + k; limB = B.lim[lR];
      // inside the brick only the far faces of the two cells can be domain boundaries
      bndA = face_kind(B.info[lL], fo) == FV_FK_BOUNDARY;
      bndB = face_kind(B.info[lR], f) == FV_FK_BOUNDARY;
      if constexpr (GR) {
        double d;
        if (p == 0) br_face_dist(G, B.info[lL], fo, lvl, d, idmA);
        if (p == 2) br_face_dist(G, B.info[lR], f, lvl, d, idmB);
      }
    }
    fvf::SolF<NM, R> sL, sR;
    double F[L::NEQ];
    float rec[MIXED ? fvf::Rec<NM>::N : 1];
    // the right side first: its stencil value beyond may be X2, which then dies
    if (!GR || kf21 == 0) {
      br_face_state<NM, MIXED, XR>(P, Md, a, B, cB, gB, oB, X2, oreg, limB, bndB, dcf, idcc, idmB, F, rec);
      br_sol<NM, MIXED, R>(Md, M, a, F, rec, sR);
    }
    br_face_state<NM, MIXED, XR>(P, Md, a, B, cA, gA, oA, X2, false, limA, bndA, dcfA, idpA, idmA, F, rec);
    if (need) br_write_sf<NM, MIXED>(n, b0 + lown, fL, F, rec, SF, SFm);
    br_sol<NM, MIXED, R>(Md, M, a, F, rec, sL);
    if (bcode != 0) sR = sL;
    R fr[9];
#pragma unroll
    for (int qq = 0; qq < 9; qq++) fr[qq] = (R)G.frame[fL][qq];
    if (bcode != 0) fvf::ghost<NM, R, XR>(M, P, bcode, fr, sL, sR);
    if (!GR || kf21 == 0) {
      R out[L::NEQ], un, vn, wn;
      fvf::hllc<NM, R>(M, P.fix, fr, bcode, sL, sR, out, un, vn, wn);
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) {
        if constexpr (MIXED) B.ff[v * BR_TASKS + tid] = out[v];
        else B.ff[v * BR_TASKS + tid] = A * (double)out[v];
      }
      B.fu[tid] =
          (double)un * G.frame[fL][0] + (double)vn * G.frame[fL][1] + (double)wn * G.frame[fL][2];
    } else {
      // a 2:1 face (one neighbour) or a 1:4 face (four sub-faces of the finer
      // area, star velocity their mean), in k_flux_fast's order
      const int fo = FCOPP_D[fL];
      const bool fine = kf21 == FV_FK_FINE;
      const int nsub = fine ? 4 : 1;
      const double As = fine ? G.area[lvl + 1] : A, wsub = fine ? 0.25 : 1.0;
      double acc[L::NEQ], u0 = 0.0, u1 = 0.0, u2 = 0.0;
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) acc[v] = 0.0;
      for (int k = 0; k < nsub; ++k) {
        const int j = fine ? fine4[4 * nb21 + k] : nb21;
        load_sol<NM, MIXED, R>(Md, M, a, SF, SFm, n, j, fo, sR);
        R out[L::NEQ], un, vn, wn;
        fvf::hllc<NM, R>(M, P.fix, fr, 0, sL, sR, out, un, vn, wn);
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) {
          if constexpr (MIXED) acc[v] += (double)out[v];
          else acc[v] += As * (double)out[v];
        }
        u0 += wsub * (double)un;
        u1 += wsub * (double)vn;
        u2 += wsub * (double)wn;
      }
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) B.ff[v * BR_TASKS + tid] = (typename LD::FT)acc[v];
      B.fu[tid] = u0 * G.frame[fL][0] + u1 * G.frame[fL][1] + u2 * G.frame[fL][2];
    }
  }
  __syncthreads();

  // ---- per cell: gather the six faces in face order, RK stage, relaxation
  if (tid >= 64) return;
  const int l = tid, i = b0 + l;
  const uint32_t inf = B.info[l];
  double Sc[L::NEQ];   // the raw state
  double src_avf[NM], src_st[NM];
  if constexpr (MIXED) {
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) Sc[v] = S[(size_t)v * n + i];
    fvf::SolF<NM, double> sc;
    fvf::fill_cons<NM, double>(Md, a, Sc, sc);
#pragma unroll
    for (int k = 0; k < L::N3; ++k) src_avf[k] = sc.avf[k];
#pragma unroll
    for (int k = 0; k < NM; ++k) src_st[k] = -sc.avf[k] * sc.p[k];
  } else {
    br_state<NM, MIXED>(B, l, Sc);
#pragma unroll
    for (int k = 0; k < L::N3; k++) Sc[L::I3 + k] = B.araw[k * 64 + l];
#pragma unroll
    for (int k = 0; k < L::N3; ++k) src_avf[k] = B.src[k * 64 + l];
#pragma unroll
    for (int k = 0; k < NM; ++k) src_st[k] = B.src[(L::N3 + k) * 64 + l];
  }
  const bool fixnc = (P.fix & 1) != 0;
  const BrPos pos = br_pos(l);
  double RHS[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) RHS[v] = 0.0;
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const int ax = br_axis(f), along = br_get(pos, ax), k = br_sheet(ax, pos);
    const int kind = face_kind(inf, f);
    int t;
    bool right = false;
    if (br_plus(f)) t = (along == 3) ? 16 * f + k : BR_SLOTS + 48 * ax + 16 * along + k;
    else if (along == 0) t = 16 * f + k;
    else { t = BR_SLOTS + 48 * ax + 16 * (along - 1) + k; right = true; }
    const double sgn = right ? -1.0 : 1.0;
    // a 1:4 face on a graded brick's surface: four sub-faces of the finer area
    double Asub = A, Area_f = A;
    if (GR && kind == FV_FK_FINE) { Asub = G.area[lvl + 1]; Area_f = 4.0 * Asub; }
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) {
      if constexpr (MIXED) RHS[v] += sgn * (Asub * (double)B.ff[v * BR_TASKS + t]);
      else RHS[v] += sgn * B.ff[v * BR_TASKS + t];
    }
    // u*.n of this leaf's face: the star velocity is shared by both sides; the
    // published code negates the reused global-frame vector on the mirrored side
    // (m4_residual.c:72-74), which FV_FIX_NONCONS leaves out
    const double un = B.fu[t];
    double und;
    if (right) und = (!fixnc && kind == FV_FK_MIRROR) ? un : -un;
    else und = (!fixnc && kind == FV_FK_MIRROR) ? -un : un;
#pragma unroll
    for (int q = 0; q < L::N3; ++q) {
      const double tt = Area_f * src_avf[q] * und;
      RHS[L::I3 + q] = fixnc ? RHS[L::I3 + q] - tt : RHS[L::I3 + q] + tt;
    }
#pragma unroll
    for (int q = 0; q < NM; ++q) {
      const double tt = Area_f * src_st[q] * und;
      RHS[L::I4 + q] = fixnc ? RHS[L::I4 + q] - tt : RHS[L::I4 + q] + tt;
    }
  }
  bool bad = false;
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) bad |= isnan(RHS[v]);
  if (bad) *nanflag = 1;
  const double dtv = P.dt * fvf::rcp(G.vol[lvl]);
  const double wa = (cs == 1) ? 0.0 : (cs == 2 ? 0.75 : (cs == 3 ? 1.0 / 3.0 : 0.5));
  const double wb = 1.0 - wa;
  double Sn[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) {
    double sn;
    if (cs == 1) { sn = Sc[v]; SN[(size_t)v * n + i] = sn; }
    else sn = SN[(size_t)v * n + i];   // S^n is read here, not staged: less LDS, and no staging load waits on it
    Sn[v] = wa * sn + wb * (Sc[v] - dtv * RHS[v]);
  }
  fvf::relax_fast<NM, XR>(P, Md, Sn);
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sout[(size_t)v * n + i] = Sn[v];
}

// A leaf i starts a brick when leaves i .. i+63 are of one level and connected
// as a 4x4x4 block in DFS (Morton) order, and each face of the block surface
// has a same-level neighbour or is a domain boundary.  First a cheap test per
// leaf: one level at both ends, and the links a block's first cells have to
// their +x, +y, +z neighbours and to the next octants (cells 1, 2, 4 to the
// octants at 8, 16, 32), which only a block's first leaf satisfies; then one
// wave per candidate checks all 64 cells' faces.
__global__ __launch_bounds__(FV_BLOCK) void k_brick_quick(int n, const int *__restrict__ nbr,
                                                          const uint32_t *__restrict__ info, int *__restrict__ mark) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int ok = 0;
  if (i + 63 < n) {
    const size_t N = (size_t)n;
    ok = cell_level(info[i + 63]) == cell_level(info[i]) && nbr[1 * N + i] == i + 1 && nbr[2 * N + i] == i + 2 &&
         nbr[5 * N + i] == i + 4 && nbr[1 * N + i + 1] == i + 8 && nbr[2 * N + i + 2] == i + 16 &&
         nbr[5 * N + i + 4] == i + 32;
  }
  mark[i] = ok;
}

// mark: 1 for a brick whose surface faces are all same-level or boundary
// faces, 2 (graded on) for one with 2:1 or 1:4 faces on its surface
__global__ __launch_bounds__(64) void k_brick_full(int n, const int *__restrict__ cand, const int *__restrict__ nbr,
                                                   const uint32_t *__restrict__ info, int graded,
                                                   int *__restrict__ mark) {
  const int i = cand[blockIdx.x];
  const int l = threadIdx.x;
  const uint32_t il = info[i + l];
  bool ok = cell_level(il) == cell_level(info[i]), plain = true;
  const BrPos pos = br_pos(l);
  for (int f = 0; f < 6; f++) {
    const int k = face_kind(il, f), ax = br_axis(f);
    BrPos q = pos;
    br_add(q, ax, br_plus(f) ? 1 : -1);
    const int along = br_get(q, ax);
    if (along >= 0 && along <= 3)
      ok = ok && (k == FV_FK_SAME || k == FV_FK_MIRROR) && nbr[(size_t)f * n + i + l] == i + br_index(q);
    else
      plain = plain && (k == FV_FK_SAME || k == FV_FK_MIRROR || k == FV_FK_BOUNDARY);
  }
  const bool all = __all(ok), allplain = __all(plain);
  if (l == 0 && all) mark[i] = allplain ? 1 : (graded ? 2 : 0);
}

// A graded candidate is dropped when a leaf across one of its 1:4 faces is in
// any candidate brick (inbrick marks every candidate's leaves).  The kept
// graded bricks then have only leaves outside bricks across their 2:1 and 1:4
// faces: a coarser candidate across a 2:1 face has that face as a 1:4 face to
// this candidate's leaf and is dropped itself, and a plain brick has no such
// face.  So every face state a graded brick reads from SF is written by the
// per-leaf reconstruction before the block kernels run.
__global__ __launch_bounds__(64) void k_brick_resolve(int n, const int *__restrict__ cand, const int *__restrict__ nbr,
                                                      const uint32_t *__restrict__ info,
                                                      const int *__restrict__ fine4,
                                                      const uint8_t *__restrict__ inbrick, int *__restrict__ mark) {
  const int i = cand[blockIdx.x];
  if (mark[i] != 2) return;   // the same for the whole workgroup
  const int l = threadIdx.x;
  const uint32_t il = info[i + l];
  bool clash = false;
  for (int f = 0; f < 6; f++) {
    if (face_kind(il, f) != FV_FK_FINE) continue;
    const int row = nbr[(size_t)f * n + i + l];
    for (int k = 0; k < 4; k++) clash = clash || inbrick[fine4[4 * row + k]] != 0;
  }
  const bool any = __any(clash);
  if (l == 0 && any) mark[i] = 0;
}

__global__ __launch_bounds__(FV_BLOCK) void k_mark_is(int n, const int *__restrict__ mark, int v,
                                                      uint8_t *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = mark[i] == v;
}

// leaves of the listed bricks: inbrick = 1 (0 elsewhere, set beforehand); rest = !inbrick
__global__ __launch_bounds__(FV_BLOCK) void k_brick_fill(int nbricks, const int *__restrict__ bricks,
                                                         uint8_t *__restrict__ inbrick) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbricks * 64) return;
  inbrick[bricks[t >> 6] + (t & 63)] = 1;
}

__global__ __launch_bounds__(FV_BLOCK) void k_not(int n, const uint8_t *__restrict__ a, uint8_t *__restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i] ? 0 : 1;
}

#pragma clang fp contract(off)

// CFL: per-block minimum of h / (max|u_i| + c) with a wavefront shuffle reduction
template <int NM>
__global__ __launch_bounds__(FV_BLOCK) void k_cfl(int n, Params P, Geo G, const double *__restrict__ S,
                                                  const uint32_t *__restrict__ info, double *__restrict__ part) {
  using L = Lay<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = INFINITY;
  if (i < n) {
    double X[L::NEQ];
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) X[v] = S[(size_t)v * n + i];
    Sol<NM, double> s;
    fill_sol<NM, double>(P, X, s);
    double c2 = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) c2 += s.c2Y[k];
    double lam = maxr(maxr(fabs(s.u), fabs(s.v)), fabs(s.w)) + sqrt(c2);
    d = G.h[cell_level(info[i])] / lam;
  }
  for (int off = 32; off > 0; off >>= 1) d = fmin(d, __shfl_xor(d, off, 64));
  __shared__ double wmin[FV_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmin[0];
    for (int w = 1; w < FV_BLOCK / 64; w++) m = fmin(m, wmin[w]);
    part[blockIdx.x] = m;
  }
}

// minimum of m partial minima, times scale, into *out (one block of 1024 lanes)
__global__ void k_min_final(int m, const double *__restrict__ part, double *__restrict__ out, double scale) {
  double d = INFINITY;
  for (int k = threadIdx.x; k < m; k += blockDim.x) d = fmin(d, part[k]);
  for (int off = 32; off > 0; off >>= 1) d = fmin(d, __shfl_xor(d, off, 64));
  __shared__ double wmin[16];
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = wmin[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = fmin(r, wmin[w]);
    *out = scale * r;
  }
}

// ---- indicator: interface_location.c:7-86 and criterionsplt.c:3-140 ----
struct CritDev {
  int uniform_on, uniform_lvl, prs_on, prs_lvl, rho_on, rho_lvl, vel_on, vel_lvl;
  int inter0_on, inter0_lvl, inter0_lrs, inter1_on, inter1_lvl, inter1_lrs;
  double prs_thr, rho_thr, vel_thr;
};

FV_DEV double rel_jump(double x_cur, double x_nb) {
  const double eps = 1.0e-8;
  if ((fabs(x_nb) > eps) || (fabs(x_cur) > eps)) return fabs(x_nb - x_cur) / maxr(fabs(x_nb), fabs(x_cur));
  return 0.0;
}

FV_DEV int sub_neighbour(int kind, int nb, int k, const int *__restrict__ fine4) {
  return (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
}

template <int NM>
__global__ __launch_bounds__(FV_BLOCK) void k_si_first(int n, Params P, const double *__restrict__ S,
                                                       const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                       const int *__restrict__ fine4, int *__restrict__ SI) {
  using L = Lay<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = P.amin / 100.0;
  const uint32_t inf = info[i];
  int si0 = 0, si1 = 0, flag = 0;
  auto phi = [&](int c, int k) {
    double x = S[(size_t)(L::I3 + k) * n + c];
    if (x <= a) x = a;
    else if (x >= (1.0 - (a))) x = 1.0 - (a);
    return x - P.avf_lim;
  };
  for (int f = 0; f < 6 && !flag; f++) {
    const int kind = face_kind(inf, f);
    if (kind == FV_FK_BOUNDARY) continue;
    const int nb = nbr[(size_t)f * n + i];
    const int ns = (kind == FV_FK_FINE) ? 4 : 1;
    for (int k = 0; k < ns; k++) {
      const int j = sub_neighbour(kind, nb, k, fine4);
#pragma unroll
      for (int v = 0; v < L::N3; ++v) {
        if (phi(i, v) * phi(j, v) < 0.0) {
          if (v == 0) si0 = 1; else si1 = 1;
          flag = 1;
        }
      }
    }
  }
  SI[2 * i] = si0;
  SI[2 * i + 1] = si1;
}

// New SI: the info word's interface bit and the mirror / same choice of
// every equal-level face, by the rule of the host packer (a face is computed
// once, by the later leaf, only when neither leaf is an interface leaf).
__global__ __launch_bounds__(FV_BLOCK) void k_refresh_kinds(int n, const int *__restrict__ nbr,
                                                            const int *__restrict__ SI, uint32_t *__restrict__ info) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int si = SI[2 * i];
  uint32_t inf = (info[i] & ~16u) | (si != 0 ? 16u : 0u);
  for (int f = 0; f < 6; f++) {
    const int kind = face_kind(inf, f);
    if (kind != FV_FK_SAME && kind != FV_FK_MIRROR) continue;
    const int j = nbr[(size_t)f * n + i];
    const uint32_t nk = (si == 0 && SI[2 * j] == 0 && j < i) ? FV_FK_MIRROR : FV_FK_SAME;
    inf = (inf & ~(7u << (8 + 3 * f))) | (nk << (8 + 3 * f));
  }
  info[i] = inf;
}

template <int NM>
__global__ __launch_bounds__(FV_BLOCK) void k_split(int n, Params P, CritDev C, const double *__restrict__ S,
                                                    const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                    const int *__restrict__ fine4, const int *__restrict__ SI,
                                                    int *__restrict__ split) {
  using L = Lay<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t inf = info[i];
  int lt = 1;
  if (C.uniform_on == 1) lt = max(lt, C.uniform_lvl);
  if (C.prs_on == 1 || C.rho_on == 1 || C.vel_on == 1) {
    double X[L::NEQ];
    Sol<NM, double> sL, sR;
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) X[v] = S[(size_t)v * n + i];
    fill_sol<NM, double>(P, X, sL);
    double pl = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) pl += sL.avf[k] * sL.p[k];
    double gp = 0.0, gr = 0.0, gv = 0.0, grho = 0.0, lg = 1;
    for (int f = 0; f < 6; f++) {
      const int kind = face_kind(inf, f);
      if (kind == FV_FK_BOUNDARY) continue;
      const int nb = nbr[(size_t)f * n + i];
      const int ns = (kind == FV_FK_FINE) ? 4 : 1;
      for (int k = 0; k < ns; k++) {
        const int j = sub_neighbour(kind, nb, k, fine4);
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) X[v] = S[(size_t)v * n + j];
        fill_sol<NM, double>(P, X, sR);
        if (C.prs_on == 1) {
          double pr = 0.0;
#pragma unroll
          for (int q = 0; q < NM; ++q) pr += sR.avf[q] * sR.p[q];
          gp = maxr(rel_jump(pl, pr), gp);
        }
        if (C.rho_on == 1) { grho = rel_jump(sL.r, sR.r); gr = maxr(grho, gr); }
        if (C.vel_on == 1) gv = maxr(grho, gv);   // criterionsplt.c:113 reads the density jump
      }
    }
    if (C.prs_on == 1 && fabs(gp) >= C.prs_thr) lg = C.prs_lvl;
    if (C.rho_on == 1 && fabs(gr) >= C.rho_thr) lg = C.rho_lvl;
    if (C.vel_on == 1 && fabs(gv) >= C.vel_thr) lg = C.vel_lvl;
    lt = (int)maxr((double)lt, lg);
  }
  if (C.inter0_on != 0) {
    for (int v = 0; v < 2; ++v) {
      const int s = SI[2 * i + v];
      if (s >= 1 && s < C.inter0_lrs) lt = max(lt, C.inter0_lvl);
      if (s >= C.inter0_lrs && s < C.inter1_lrs) lt = max(lt, C.inter1_lvl);
    }
  }
  split[i] = lt;
}

// ---------------------------------------------------------------------
// host side of the C ABI
// ---------------------------------------------------------------------
thread_local char fv_g_err[256];

int fv_set_err(int code, const char *msg) {
  snprintf(fv_g_err, sizeof fv_g_err, "%s", msg);
  return code;
}
static int set_err(int code, const char *msg) { return fv_set_err(code, msg); }

static void to_dev_params(const fv_params *p, Params &P) {
  memset(&P, 0, sizeof P);
  for (int k = 0; k < 4; k++) { P.gamma[k] = p->gamma[k]; P.pinf[k] = p->pinf[k]; }
  P.amin = p->amin; P.avf_lim = p->avf_lim; P.dt = p->dt;
  P.prelax = p->prelax; P.limiter = p->limiter; P.rk = p->rk; P.nmat = p->nmat; P.fix = p->fix;
  P.primtv = p->primtv;
  P.bc_fixed_p = p->bc_fixed_p; P.p_inlet = p->p_inlet; P.p_outlet = p->p_outlet;
  for (int k = 0; k < 4; k++) {
    const double g = p->gamma[k], pi = p->pinf[k];
    const double gm1 = (g != 1.0) ? g - 1.0 : 1.0, gg = (g != 0.0) ? g : 1.0;
    P.GP[k] = g * pi; P.Gm1[k] = g - 1.0; P.iGm1[k] = 1.0 / gm1; P.C[k] = g * pi / gm1; P.iG[k] = 1.0 / gg;
    P.Gf[k] = (float)g; P.Pf[k] = (float)pi; P.GPf[k] = (float)P.GP[k]; P.Gm1f[k] = (float)P.Gm1[k];
    P.iGm1f[k] = (float)P.iGm1[k]; P.Cf[k] = (float)P.C[k]; P.iGf[k] = (float)P.iG[k];
  }
}

static void to_dev_geo(const fv_geo *g, Geo &G) {
  memset(&G, 0, sizeof G);
  memcpy(G.frame, g->frame, sizeof G.frame);
  for (int l = 0; l < 12; l++) {
    G.area[l] = g->area[l]; G.vol[l] = g->vol[l]; G.h[l] = g->h[l];
    for (int k = 0; k < 3; k++) { G.dcf[k][l] = g->dcf[k][l]; G.dcc[k][l] = g->dcc[k][l]; }
  }
}

static void free_buffers(fv_dev *c) {
  hipFree(c->S); hipFree(c->S2); hipFree(c->SN); hipFree(c->SF); hipFree(c->part); hipFree(c->dtmp); hipFree(c->SFm);
  c->SFm = nullptr;
  c->cap_m = 0;
  hipFree(c->nbr); hipFree(c->fine4); hipFree(c->nanflag); hipFree(c->info); hipFree(c->iflag); hipFree(c->bfsq);
  c->iflag = c->bfsq = nullptr;
  hipFree(c->SIa); hipFree(c->SIb); hipFree(c->split); hipFree(c->lim);
  c->lim = nullptr;
  c->S = c->S2 = c->SN = c->SF = c->part = c->dtmp = nullptr;
  c->nbr = c->fine4 = c->nanflag = c->SIa = c->SIb = c->split = nullptr;
  c->info = nullptr;
  c->cap_n = c->cap_fine = 0;
}

// buffers for n leaves and nfine 1:4 faces; reallocates everything (contents
// lost) only when the capacity is too small.  The new set is allocated beside
// the old one and swapped in only when every allocation succeeded, so a failed
// reservation leaves the context with its old buffers, size and state.
int fv_dev_reserve(fv_dev *c, int n_leaves, int nfine) {
  if (n_leaves <= c->cap_n && nfine <= c->cap_fine) return FV_OK;
  const size_t n = (size_t)n_leaves;
  const size_t nf = (size_t)(nfine > 0 ? nfine : 1);
  const size_t nb = (n + FV_BLOCK - 1) / FV_BLOCK;
  fv_dev t;   // only its buffer pointers are used
  struct { void **p; size_t bytes; } want[] = {
      {(void **)&t.S, sizeof(double) * n * c->neq},      {(void **)&t.S2, sizeof(double) * n * c->neq},
      {(void **)&t.SN, sizeof(double) * n * c->neq},     {(void **)&t.SF, sizeof(double) * n * 6 * c->neq},
      {(void **)&t.part, sizeof(double) * (nb + 1)},     {(void **)&t.dtmp, sizeof(double) * 4},
      {(void **)&t.nbr, sizeof(int) * n * 6},            {(void **)&t.fine4, sizeof(int) * nf * 4},
      {(void **)&t.nanflag, sizeof(int)},                {(void **)&t.iflag, 8 * sizeof(int)},
      {(void **)&t.bfsq, 2 * n * sizeof(int)},           {(void **)&t.info, sizeof(uint32_t) * n},
      {(void **)&t.SIa, sizeof(int) * n * 2},            {(void **)&t.SIb, sizeof(int) * n * 2},
      {(void **)&t.split, sizeof(int) * n},              {(void **)&t.lim, n},
  };
  const int nw = (int)(sizeof want / sizeof want[0]);
  for (int k = 0; k < nw; k++) {
    // a topology-only context (neq = 0, fv_dev_attach_topology) holds no state arrays
    if (want[k].bytes == 0) { *want[k].p = nullptr; continue; }
    const hipError_t e = hipMalloc(want[k].p, want[k].bytes);
    if (e != hipSuccess) {
      for (int j = 0; j < k; j++) hipFree(*want[j].p);
      snprintf(fv_g_err, sizeof fv_g_err, "reserve for %d leaves: hipMalloc of %zu bytes: %s (the context keeps its "
               "%d leaves)", n_leaves, want[k].bytes, hipGetErrorString(e), c->n);
      return e == hipErrorOutOfMemory ? FV_ERR_NOMEM : FV_ERR_HIP;
    }
  }
  free_buffers(c);
  c->S = t.S; c->S2 = t.S2; c->SN = t.SN; c->SF = t.SF; c->part = t.part; c->dtmp = t.dtmp;
  c->nbr = t.nbr; c->fine4 = t.fine4; c->nanflag = t.nanflag; c->iflag = t.iflag; c->bfsq = t.bfsq;
  c->info = t.info; c->SIa = t.SIa; c->SIb = t.SIb; c->split = t.split; c->lim = t.lim;
  c->cap_n = n_leaves;
  c->cap_fine = (int)nf;
  return FV_OK;
}

static int upload_mesh(fv_dev *c, const fv_pack *pk) {
  if (pk->n <= 0) return set_err(FV_ERR_ARG, "empty mesh");
  const size_t n = (size_t)pk->n;
  int rc = fv_dev_reserve(c, pk->n, pk->nfine);
  if (rc != FV_OK) return rc;
  c->broken = 1;   // until every upload below has succeeded
  c->brick_ok = 0;
  c->forest = 0;   // the leaf addresses of an earlier fv_dev_set_forest no longer apply
  c->n = pk->n;
  c->nfine = pk->nfine;
  to_dev_geo(&pk->geo, c->G);
  if (!c->Gdev) HIPCHK(hipMalloc(&c->Gdev, sizeof(Geo)));
  HIPCHK(hipMemcpyAsync(c->Gdev, &c->G, sizeof(Geo), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->nbr, pk->nbr, sizeof(int) * n * 6, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->info, pk->info, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
  if (pk->nfine > 0)
    HIPCHK(hipMemcpyAsync(c->fine4, pk->fine4, sizeof(int) * 4 * (size_t)pk->nfine, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->nanflag, 0, sizeof(int), c->stream));
  HIPCHK(hipMemsetAsync(c->lim, 0, n, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->broken = 0;
  return FV_OK;
}

extern "C" int fv_dev_attach(fv_dev **out, const fv_params *p, const fv_pack *pk) {
  if (!out || !p || !pk) return set_err(FV_ERR_ARG, "null argument");
  if (p->nmat != 2 && p->nmat != 3) return set_err(FV_ERR_ARG, "nmat must be 2 or 3");
  if (p->prelax < 0 || p->prelax > 3) return set_err(FV_ERR_ARG, "PRELAX must be 0, 1, 2 or 3");
  if (p->primtv != 0 && p->primtv != 1) return set_err(FV_ERR_ARG, "PRIMTV must be 0 or 1");
  if (p->bc_fixed_p != 0 && p->bc_fixed_p != 1) return set_err(FV_ERR_ARG, "bc_fixed_p must be 0 or 1");
  if (p->bc_fixed_p && !(isfinite(p->p_inlet) && isfinite(p->p_outlet)))
    return set_err(FV_ERR_ARG, "the inlet and outlet pressures must be finite");
  if (p->prelax == 1 && p->nmat != 2) return set_err(FV_ERR_ARG, "PRELAX=1 is two-phase only");
  if (p->rk != FV_RK_REF2 && p->rk != FV_RK_SSP3 && p->rk != FV_RK_SSP2) return set_err(FV_ERR_ARG, "unknown RK table");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(FV_ERR_HIP, "no HIP device");
  fv_dev *c = new fv_dev();
  c->nmat = p->nmat;
  c->neq = fv_neq(p->nmat);
  to_dev_params(p, c->P);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_err(FV_ERR_HIP, "stream creation failed");
  }
  int rc = upload_mesh(c, pk);
  if (rc != FV_OK) { free_buffers(c); hipFree(c->Gdev); hipStreamDestroy(c->stream); delete c; return rc; }
  *out = c;
  return FV_OK;
}

// a context that holds a mesh's topology, forest and indicator buffers but no
// state (neq = 0): the mesh side of a model-5 context's regrid
int fv_dev_attach_topology(fv_dev **out, const fv_pack *pk, int nmat) {
  if (!out || !pk) return set_err(FV_ERR_ARG, "null argument");
  fv_dev *c = new fv_dev();
  c->nmat = nmat;
  c->neq = 0;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_err(FV_ERR_HIP, "stream creation failed");
  }
  int rc = upload_mesh(c, pk);
  if (rc != FV_OK) { free_buffers(c); hipFree(c->Gdev); hipStreamDestroy(c->stream); delete c; return rc; }
  *out = c;
  return FV_OK;
}

extern "C" int fv_dev_reattach(fv_dev *c, const fv_pack *pk) {
  FV_GUARD(c);
  if (!c || !pk) return set_err(FV_ERR_ARG, "null argument");
  return upload_mesh(c, pk);
}

extern "C" int fv_dev_detach(fv_dev *c) {
  if (!c) return FV_OK;
  hipStreamSynchronize(c->stream);
  free_buffers(c);
  hipFree(c->lroot); hipFree(c->llevel); hipFree(c->lci); hipFree(c->lSI); hipFree(c->tpart);
  hipFree(c->bricks); hipFree(c->rest); hipFree(c->inbrick); hipFree(c->notbrick); hipFree(c->bmark);
  hipFree(c->bcnt); hipFree(c->bcub);
  for (auto &b : c->pool) hipFree(b.first);
  c->pool.clear();
  hipFree(c->Gdev);
  if (c->pinned) hipHostFree(c->pinned);
  hipStreamDestroy(c->stream);
  delete c;
  return FV_OK;
}

extern "C" int fv_dev_size(const fv_dev *c, int *n, int *neq) {
  if (!c) return set_err(FV_ERR_ARG, "null context");
  if (n) *n = c->n;
  if (neq) *neq = c->neq;
  return FV_OK;
}

// [n][NEQ] <-> [NEQ][n] on the host side of the copy
extern "C" int fv_dev_set_state(fv_dev *c, const double *S) {
  FV_GUARD(c);
  if (!c || !S) return set_err(FV_ERR_ARG, "null argument");
  const size_t n = c->n, neq = c->neq;
  double *h = (double *)malloc(sizeof(double) * n * neq);
  if (!h) return set_err(FV_ERR_NOMEM, "host staging");
  for (size_t i = 0; i < n; i++)
    for (size_t v = 0; v < neq; v++) h[v * n + i] = S[i * neq + v];
  hipError_t e = hipMemcpy(c->S, h, sizeof(double) * n * neq, hipMemcpyHostToDevice);
  free(h);
  HIPCHK(e);
  return FV_OK;
}

extern "C" int fv_dev_get_state(fv_dev *c, double *S) {
  FV_GUARD(c);
  if (!c || !S) return set_err(FV_ERR_ARG, "null argument");
  const size_t n = c->n, neq = c->neq;
  double *h = (double *)malloc(sizeof(double) * n * neq);
  if (!h) return set_err(FV_ERR_NOMEM, "host staging");
  HIPCHK(hipStreamSynchronize(c->stream));
  hipError_t e = hipMemcpy(h, c->S, sizeof(double) * n * neq, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (size_t i = 0; i < n; i++)
      for (size_t v = 0; v < neq; v++) S[i * neq + v] = h[v * n + i];
  free(h);
  HIPCHK(e);
  return FV_OK;
}

extern "C" int fv_dev_get_faces(fv_dev *c, double *SF) {
  FV_GUARD(c);
  if (!c || !SF) return set_err(FV_ERR_ARG, "null argument");
  const size_t n = c->n, neq = c->neq;
  double *h = (double *)malloc(sizeof(double) * n * 6 * neq);
  if (!h) return set_err(FV_ERR_NOMEM, "host staging");
  HIPCHK(hipStreamSynchronize(c->stream));
  hipError_t e = hipMemcpy(h, c->SF, sizeof(double) * n * 6 * neq, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (size_t i = 0; i < n; i++)
      for (size_t f = 0; f < 6; f++)
        for (size_t v = 0; v < neq; v++) SF[(i * 6 + f) * neq + v] = h[(f * neq + v) * n + i];
  free(h);
  HIPCHK(e);
  return FV_OK;
}

extern "C" int fv_dev_set_dt(fv_dev *c, double dt) {
  FV_GUARD(c);
  if (!c) return set_err(FV_ERR_ARG, "null context");
  c->P.dt = dt;
  return FV_OK;
}

extern "C" int fv_dev_set_precision(fv_dev *c, int prec) {
  FV_GUARD(c);
  if (!c) return set_err(FV_ERR_ARG, "null context");
  if (prec != FV_PREC_FP64 && prec != FV_PREC_MIXED && prec != FV_PREC_EXACT) return set_err(FV_ERR_ARG, "unknown precision");
  c->prec = prec;
  return FV_OK;
}

static int classify_bricks(fv_dev *c);

extern "C" int fv_dev_set_bricks(fv_dev *c, int on) {
  FV_GUARD(c);
  if (on < 0 || on > 2) return set_err(FV_ERR_ARG, "fv_dev_set_bricks: on must be 0, 1 or 2");
  if (on != c->use_bricks) c->brick_ok = 0;   // the block list depends on the mode
  c->use_bricks = on;
  return FV_OK;
}

extern "C" int fv_dev_brick_kinds(fv_dev *c, int *nplain, int *ngraded) {
  FV_GUARD(c);
  if (!c->brick_ok) {
    int rc = classify_bricks(c);
    if (rc != FV_OK) return rc;
  }
  if (nplain) *nplain = c->nplain;
  if (ngraded) *ngraded = c->ngraded;
  return FV_OK;
}

extern "C" int fv_dev_brick_count(fv_dev *c, int *nbricks, int *nrest) {
  FV_GUARD(c);
  if (!c->brick_ok) {
    int rc = classify_bricks(c);
    if (rc != FV_OK) return rc;
  }
  if (nbricks) *nbricks = c->nbricks;
  if (nrest) *nrest = c->nrest;
  return FV_OK;
}

extern "C" int fv_dev_set_fused(fv_dev *c, int on) {
  FV_GUARD(c);
  if (!c) return set_err(FV_ERR_ARG, "null context");
  c->fused = on ? 1 : 0;
  return FV_OK;
}

extern "C" int fv_dev_begin_step(fv_dev *c) {
  FV_GUARD(c);
  if (!c) return set_err(FV_ERR_ARG, "null context");
  HIPCHK(hipMemcpyAsync(c->SN, c->S, sizeof(double) * (size_t)c->n * c->neq, hipMemcpyDeviceToDevice, c->stream));
  return FV_OK;
}

static int stage_case(int rk, int irk) {
  if (rk == FV_RK_SSP2) return irk == 1 ? 1 : 4;
  return irk;
}

// Bricks of the current mesh (k_brick_mark), their leaves (inbrick) and the
// list of the other leaves, which run the per-leaf kernels.  Computed at the
// first stage after a mesh change; face-kind refreshes between steps keep it
// valid, since a brick accepts same-level faces of either kind.
static int classify_bricks(fv_dev *c) {
  const int n = c->n;
  if (c->brick_cap < n) {
    hipFree(c->bricks); hipFree(c->rest); hipFree(c->inbrick); hipFree(c->notbrick); hipFree(c->bmark);
    hipFree(c->bcnt); hipFree(c->bcub);
    c->bricks = c->rest = c->bmark = c->bcnt = nullptr;
    c->inbrick = c->notbrick = nullptr;
    c->bcub = nullptr;
    c->brick_cap = 0;
    size_t b1 = 0, b2 = 0;
    hipcub::CountingInputIterator<int> idx(0);
    HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, b1, idx, (const int *)nullptr, (int *)nullptr, (int *)nullptr, n,
                                         c->stream));
    HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, b2, idx, (const uint8_t *)nullptr, (int *)nullptr, (int *)nullptr,
                                         n, c->stream));
    c->bcub_bytes = b1 > b2 ? b1 : b2;
    HIPCHK(hipMalloc(&c->bricks, sizeof(int) * ((size_t)n / 64 + 1)));
    HIPCHK(hipMalloc(&c->rest, sizeof(int) * (size_t)n));
    HIPCHK(hipMalloc(&c->bmark, sizeof(int) * (size_t)n));
    HIPCHK(hipMalloc(&c->inbrick, (size_t)n));
    HIPCHK(hipMalloc(&c->notbrick, (size_t)n));
    HIPCHK(hipMalloc(&c->bcnt, 2 * sizeof(int)));
    HIPCHK(hipMalloc(&c->bcub, c->bcub_bytes));
    c->brick_cap = n;
  }
  const dim3 grid((n + FV_BLOCK - 1) / FV_BLOCK), block(FV_BLOCK);
  // candidates (into the rest list's buffer, which holds n), then the full check
  hipLaunchKernelGGL(k_brick_quick, grid, block, 0, c->stream, n, c->nbr, c->info, c->bmark);
  hipcub::CountingInputIterator<int> idx(0);
  size_t bytes = c->bcub_bytes;
  HIPCHK(hipcub::DeviceSelect::Flagged(c->bcub, bytes, idx, c->bmark, c->rest, c->bcnt, n, c->stream));
  int ncand = 0;
  HIPCHK(hipMemcpyAsync(&ncand, c->bcnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemsetAsync(c->bmark, 0, sizeof(int) * (size_t)n, c->stream));
  const int graded = c->use_bricks == 2;
  if (ncand > 0)
    hipLaunchKernelGGL(k_brick_full, dim3(ncand), dim3(64), 0, c->stream, n, c->rest, c->nbr, c->info, graded,
                       c->bmark);
  // every candidate (plain or graded): its leaves into inbrick, then the
  // graded candidates that clash with another candidate are dropped
  bytes = c->bcub_bytes;
  HIPCHK(hipcub::DeviceSelect::Flagged(c->bcub, bytes, idx, c->bmark, c->bricks, c->bcnt, n, c->stream));
  int nb = 0;
  HIPCHK(hipMemcpyAsync(&nb, c->bcnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (nb > n / 64) return set_err(FV_ERR_MESH, "brick classification: more blocks than the leaves allow");
  HIPCHK(hipMemsetAsync(c->inbrick, 0, (size_t)n, c->stream));
  if (nb > 0)
    hipLaunchKernelGGL(k_brick_fill, dim3((64 * nb + FV_BLOCK - 1) / FV_BLOCK), block, 0, c->stream, nb, c->bricks,
                       c->inbrick);
  int np = nb, ng = 0;
  if (graded && nb > 0) {
    hipLaunchKernelGGL(k_brick_resolve, dim3(nb), dim3(64), 0, c->stream, n, c->bricks, c->nbr, c->info, c->fine4,
                       c->inbrick, c->bmark);
    // the plain bricks first, then the kept graded ones (both in leaf order)
    hipLaunchKernelGGL(k_mark_is, grid, block, 0, c->stream, n, c->bmark, 1, c->notbrick);
    bytes = c->bcub_bytes;
    HIPCHK(hipcub::DeviceSelect::Flagged(c->bcub, bytes, idx, c->notbrick, c->bricks, c->bcnt, n, c->stream));
    HIPCHK(hipMemcpyAsync(&np, c->bcnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (np > nb) return set_err(FV_ERR_MESH, "brick classification: more plain blocks than candidates");
    hipLaunchKernelGGL(k_mark_is, grid, block, 0, c->stream, n, c->bmark, 2, c->notbrick);
    bytes = c->bcub_bytes;
    HIPCHK(hipcub::DeviceSelect::Flagged(c->bcub, bytes, idx, c->notbrick, c->bricks + np, c->bcnt, n, c->stream));
    HIPCHK(hipMemcpyAsync(&ng, c->bcnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (np + ng > nb) return set_err(FV_ERR_MESH, "brick classification: more blocks than candidates");
    nb = np + ng;
    HIPCHK(hipMemsetAsync(c->inbrick, 0, (size_t)n, c->stream));
    if (nb > 0)
      hipLaunchKernelGGL(k_brick_fill, dim3((64 * nb + FV_BLOCK - 1) / FV_BLOCK), block, 0, c->stream, nb, c->bricks,
                         c->inbrick);
  }
  hipLaunchKernelGGL(k_not, grid, block, 0, c->stream, n, c->inbrick, c->notbrick);
  bytes = c->bcub_bytes;
  HIPCHK(hipcub::DeviceSelect::Flagged(c->bcub, bytes, idx, c->notbrick, c->rest, c->bcnt + 1, n, c->stream));
  int nr = 0;
  HIPCHK(hipMemcpyAsync(&nr, c->bcnt + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipGetLastError());
  if (nr + 64 * nb != n) return set_err(FV_ERR_MESH, "brick classification: leaves counted twice or not at all");
  c->nbricks = nb;
  c->nplain = np;
  c->ngraded = ng;
  c->nrest = nr;
  c->brick_ok = 1;
  return FV_OK;
}

// ev_flux0 / ev_flux1 bracket the dominant kernel: the block kernel where the mesh
// has bricks, else the flux (or fused stage) kernel
template <int NM, bool XR>
static int launch_stage_nm(fv_dev *c, int cs, hipEvent_t *ev_flux0, hipEvent_t *ev_flux1) {
  const int n = c->n;
  const dim3 grid((n + FV_BLOCK - 1) / FV_BLOCK), block(FV_BLOCK);
  if (c->prec == FV_PREC_EXACT) {
    hipLaunchKernelGGL(k_recon<NM>, grid, block, 0, c->stream, n, c->P, c->G, c->S, c->nbr, c->info, c->fine4, c->lim,
                       c->SF);
    if (ev_flux0) HIPCHK(hipEventRecord(*ev_flux0, c->stream));
    hipLaunchKernelGGL(k_flux_rk<NM>, grid, block, 0, c->stream, n, c->P, c->G, cs, c->S, c->SN, c->SF, c->nbr,
                       c->info, c->fine4, c->nanflag);
    if (ev_flux1) HIPCHK(hipEventRecord(*ev_flux1, c->stream));
  } else if (c->fused) {
    const unsigned nb = (unsigned)((n + FV_BLOCK - 1) / FV_BLOCK);
    const dim3 gx((nb + 7u) & ~7u);   // a multiple of 8 for the XCD-contiguous leaf mapping
    if (ev_flux0) HIPCHK(hipEventRecord(*ev_flux0, c->stream));
    if (c->prec == FV_PREC_FP64)
      hipLaunchKernelGGL((k_stage_fused<NM, false, XR>), gx, block, 0, c->stream, n, c->P, c->Gdev, cs, c->S, c->S2, c->SN,
                         c->nbr, c->info, c->fine4, c->lim, c->nanflag);
    else
      hipLaunchKernelGGL((k_stage_fused<NM, true, XR>), gx, block, 0, c->stream, n, c->P, c->Gdev, cs, c->S, c->S2, c->SN,
                         c->nbr, c->info, c->fine4, c->lim, c->nanflag);
    if (ev_flux1) HIPCHK(hipEventRecord(*ev_flux1, c->stream));
    double *t = c->S;
    c->S = c->S2;
    c->S2 = t;
  } else {
    const bool mixed = c->prec == FV_PREC_MIXED;
    if (mixed && c->cap_m < n) {
      HIPCHK(hipFree(c->SFm));
      c->SFm = nullptr;
      c->cap_m = 0;
      HIPCHK(hipMalloc(&c->SFm, sizeof(float) * (size_t)n * 6 * fvf::Rec<NM>::N));
      c->cap_m = n;
    }
    const bool bricks = c->use_bricks && !c->P.primtv && n >= 64;
    if (bricks && !c->brick_ok) {
      int rc = classify_bricks(c);
      if (rc != FV_OK) return rc;
    }
    // per-leaf kernels: every leaf, or the leaves outside bricks
    const int cnt = bricks ? c->nrest : n;
    const int *list = bricks ? c->rest : nullptr;
    const dim3 gl((cnt + FV_BLOCK - 1) / FV_BLOCK);
    if (cnt > 0) {
      // the leaves outside blocks: one lane per (leaf, axis); every leaf: one per leaf
      const dim3 ga((3 * cnt + FV_BLOCK - 1) / FV_BLOCK);
      if (mixed && bricks)
        hipLaunchKernelGGL((k_recon_fast<NM, true, XR, true>), ga, block, 0, c->stream, n, c->P, c->G, c->S, c->nbr,
                           c->info, c->fine4, c->lim, c->SF, c->SFm, list, cnt);
      else if (mixed)
        hipLaunchKernelGGL((k_recon_fast<NM, true, XR, false>), gl, block, 0, c->stream, n, c->P, c->G, c->S, c->nbr,
                           c->info, c->fine4, c->lim, c->SF, c->SFm, list, cnt);
      else if (bricks)
        hipLaunchKernelGGL((k_recon_fast<NM, false, XR, true>), ga, block, 0, c->stream, n, c->P, c->G, c->S, c->nbr,
                           c->info, c->fine4, c->lim, c->SF, c->SFm, list, cnt);
      else
        hipLaunchKernelGGL((k_recon_fast<NM, false, XR, false>), gl, block, 0, c->stream, n, c->P, c->G, c->S, c->nbr,
                           c->info, c->fine4, c->lim, c->SF, c->SFm, list, cnt);
    }
    if (ev_flux0) HIPCHK(hipEventRecord(*ev_flux0, c->stream));
    if (bricks && c->nbricks > 0) {
      // the plain bricks, then the graded ones (their SF reads need only
      // k_recon_fast).  The graded launch on a second stream beside the plain one
      // measured the same stage time (0.710 against 0.709 ms on C1): the plain
      // launch fills every CU's LDS, so the two only share the tails.
      bool timed = false;
      for (int gr = 0; gr < 2; gr++) {
        const int cntb = gr ? c->ngraded : c->nplain;
        if (cntb == 0) continue;
        const int *bl = c->bricks + (gr ? c->nplain : 0);
        const dim3 gb(((unsigned)cntb + 7u) & ~7u);   // a multiple of 8 for the XCD-aware brick order
        const hipStream_t st = c->stream;
        if (mixed && gr)
          hipLaunchKernelGGL((k_stage_brick<NM, true, XR, true>), gb, dim3(BR_THREADS), 0, st, n, c->P, c->Gdev,
                             cs, bl, cntb, c->S, c->S2, c->SN, c->nbr, c->info, c->fine4, c->lim, c->inbrick, c->SF,
                             c->SFm, c->nanflag);
        else if (mixed)
          hipLaunchKernelGGL((k_stage_brick<NM, true, XR, false>), gb, dim3(BR_THREADS), 0, st, n, c->P, c->Gdev,
                             cs, bl, cntb, c->S, c->S2, c->SN, c->nbr, c->info, c->fine4, c->lim, c->inbrick, c->SF,
                             c->SFm, c->nanflag);
        else if (gr)
          hipLaunchKernelGGL((k_stage_brick<NM, false, XR, true>), gb, dim3(BR_THREADS), 0, st, n, c->P, c->Gdev,
                             cs, bl, cntb, c->S, c->S2, c->SN, c->nbr, c->info, c->fine4, c->lim, c->inbrick, c->SF,
                             c->SFm, c->nanflag);
        else
          hipLaunchKernelGGL((k_stage_brick<NM, false, XR, false>), gb, dim3(BR_THREADS), 0, st, n, c->P,
                             c->Gdev, cs, bl, cntb, c->S, c->S2, c->SN, c->nbr, c->info, c->fine4, c->lim, c->inbrick,
                             c->SF, c->SFm, c->nanflag);
        // the events bracket the first launch: the plain bricks (the graded ones when there are none)
        if (ev_flux1 && !timed) HIPCHK(hipEventRecord(*ev_flux1, c->stream));
        timed = true;
      }
    }
    if (cnt > 0) {
      // the pipelined flux needs more registers: 2 waves per SIMD with FP32 face
      // records (8 % faster), 1 with FP64 ones (35 % slower), so FP64 keeps k_flux_fast
      if (mixed)
        hipLaunchKernelGGL((k_flux_pipe<NM, true, XR>), gl, block, 0, c->stream, n, c->P, c->G, cs, c->S, c->S2, c->SN,
                           c->SF, c->SFm, c->nbr, c->info, c->fine4, c->nanflag, list, cnt);
      else
        hipLaunchKernelGGL((k_flux_fast<NM, false, XR>), gl, block, 0, c->stream, n, c->P, c->G, cs, c->S, c->S2, c->SN,
                           c->SF, c->SFm, c->nbr, c->info, c->fine4, c->nanflag, list, cnt);
    }
    if (ev_flux1 && !(bricks && c->nbricks > 0)) HIPCHK(hipEventRecord(*ev_flux1, c->stream));
    // the new state is in S2; the old one stays intact until every kernel of the
    // stage has read its neighbours from it
    double *t = c->S;
    c->S = c->S2;
    c->S2 = t;
  }
  return FV_OK;
}

static int launch_stage(fv_dev *c, int irk, hipEvent_t *ev_flux0, hipEvent_t *ev_flux1) {
  const int cs = stage_case(c->P.rk, irk);
  // PRELAX 2 and 3 iterate; their kernels are separate instances so that the
  // closed-form relaxations do not carry the iterations' code and registers.
  // The fixed-pressure inlet and outlet ghosts live in those instances too.
  const bool xr = c->P.prelax >= 2 || c->P.bc_fixed_p;
  int rc = (c->nmat == 2)
               ? (xr ? launch_stage_nm<2, true>(c, cs, ev_flux0, ev_flux1) : launch_stage_nm<2, false>(c, cs, ev_flux0, ev_flux1))
               : (xr ? launch_stage_nm<3, true>(c, cs, ev_flux0, ev_flux1) : launch_stage_nm<3, false>(c, cs, ev_flux0, ev_flux1));
  if (rc != FV_OK) return rc;
  HIPCHK(hipGetLastError());
  return FV_OK;
}

static int check_nan(fv_dev *c) {
  int flag = 0;
  HIPCHK(hipMemcpyAsync(&flag, c->nanflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (flag) {
    HIPCHK(hipMemsetAsync(c->nanflag, 0, sizeof(int), c->stream));
    return set_err(FV_ERR_NAN, "NaN in the residual");
  }
  return FV_OK;
}

extern "C" int fv_dev_stage(fv_dev *c, int irk) {
  FV_GUARD(c);
  if (!c) return set_err(FV_ERR_ARG, "null context");
  int ns = (c->P.rk == FV_RK_SSP3) ? 3 : 2;
  if (irk < 1 || irk > ns) return set_err(FV_ERR_ARG, "stage index out of range");
  int rc = launch_stage(c, irk, nullptr, nullptr);
  if (rc != FV_OK) return rc;
  return check_nan(c);
}

extern "C" int fv_dev_advance(fv_dev *c, int nsteps) {
  FV_GUARD(c);
  if (!c) return set_err(FV_ERR_ARG, "null context");
  const int ns = (c->P.rk == FV_RK_SSP3) ? 3 : 2;
  for (int s = 0; s < nsteps; s++) {
    // RK(0) is folded into stage 1, which stores SN from the state it reads
    for (int irk = 1; irk <= ns; irk++) {
      int rc = launch_stage(c, irk, nullptr, nullptr);
      if (rc != FV_OK) return rc;
    }
  }
  return check_nan(c);
}

// volume integrals of every state field: per-block tree sums, then the blocks in
// order (the same bits on every run)
__global__ __launch_bounds__(FV_BLOCK) void k_totals(int n, int neq, const Geo *__restrict__ Gp,
                                                     const double *__restrict__ S, const uint32_t *__restrict__ info,
                                                     double *__restrict__ part) {
  __shared__ double red[FV_BLOCK];
  const int i = blockIdx.x * FV_BLOCK + threadIdx.x;
  const double vol = (i < n) ? Gp->vol[cell_level(info[i])] : 0.0;
  for (int v = 0; v < neq; v++) {
    red[threadIdx.x] = (i < n) ? S[(size_t)v * n + i] * vol : 0.0;
    __syncthreads();
    for (int w = FV_BLOCK / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)v * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ void k_totals_final(int nb, int neq, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double red[1024];
  for (int v = 0; v < neq; v++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(size_t)v * nb + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[v] = red[0];
    __syncthreads();
  }
}

extern "C" int fv_dev_totals(fv_dev *c, double *tot) {
  FV_GUARD(c);
  if (!c || !tot) return set_err(FV_ERR_ARG, "null argument");
  const int n = c->n, neq = c->neq;
  const int nb = (n + FV_BLOCK - 1) / FV_BLOCK;
  const size_t need = (size_t)nb * neq + neq;
  if (need > c->cap_tpart) {
    hipFree(c->tpart);
    c->tpart = nullptr;
    c->cap_tpart = 0;
    HIPCHK(hipMalloc(&c->tpart, sizeof(double) * need));
    c->cap_tpart = need;
  }
  hipLaunchKernelGGL(k_totals, dim3(nb), dim3(FV_BLOCK), 0, c->stream, n, neq, c->Gdev, c->S, c->info, c->tpart);
  hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(1024), 0, c->stream, nb, neq, c->tpart,
                     c->tpart + (size_t)nb * neq);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(tot, c->tpart + (size_t)nb * neq, sizeof(double) * neq, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FV_OK;
}

extern "C" int fv_dev_cfl_dt(fv_dev *c, double cfl, double *dt) {
  FV_GUARD(c);
  if (!c || !dt) return set_err(FV_ERR_ARG, "null argument");
  const int n = c->n;
  const int nb = (n + FV_BLOCK - 1) / FV_BLOCK;
  if (c->nmat == 2)
    hipLaunchKernelGGL(k_cfl<2>, dim3(nb), dim3(FV_BLOCK), 0, c->stream, n, c->P, c->G, c->S, c->info, c->part);
  else
    hipLaunchKernelGGL(k_cfl<3>, dim3(nb), dim3(FV_BLOCK), 0, c->stream, n, c->P, c->G, c->S, c->info, c->part);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(1024), 0, c->stream, nb, c->part, c->dtmp, cfl);
  HIPCHK(hipGetLastError());
  double h = 0.0;
  HIPCHK(hipMemcpyAsync(&h, c->dtmp, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *dt = h;
  return FV_OK;
}

// With two materials only SI[.][0] is ever set, and layer lvl takes every
// unset leaf with a face neighbour in layer lvl-1: a breadth-first search from
// the interface leaves, cut at imax-1 layers.  The search keeps the current
// layer as a queue and pushes from it: each queued leaf claims its unset
// neighbours with a compare-and-swap and queues them for the next layer.  A
// leaf gets layer lvl exactly when a neighbour has layer lvl-1, as in the
// layer-by-layer passes, but each layer touches only its own leaves.  The
// queue order varies from run to run; the layer values do not.
constexpr int BFS_INF = 1 << 28;
constexpr int BFS_GRID = 256;
constexpr int BFS3_GRID = 1024;   // the three-material candidates: one round of lanes

// wave-aggregated append of leaf i (when take) to queue Q with counter *cnt
__device__ __forceinline__ void bfs_append(bool take, int i, int *__restrict__ Q, int *__restrict__ cnt) {
  const unsigned long long m = __ballot(take);
  if (m == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(cnt, __popcll(m));
  base = __shfl(base, leader, 64);
  if (take) Q[base + __popcll(m & ((1ull << lane) - 1ull))] = i;
}

__global__ __launch_bounds__(FV_BLOCK) void k_si_bfs_seed(int n, const int *__restrict__ SI, int *__restrict__ X,
                                                          int *__restrict__ Q, int *__restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  const bool seed = live && SI[2 * i] != 0;
  if (live) X[i] = seed ? 1 : BFS_INF;
  bfs_append(seed, i, Q, cnt);
}

// one layer: leaves of Qin (layer lvl-1) claim their unset neighbours for layer lvl.
// A lane gathers its up to 24 neighbour indices, then loads their layers, then
// claims the unset ones with compare-and-swap, so the memory operations of its
// candidates overlap instead of forming one chain; the claimed leaves are
// appended with one queue reservation per wave (an exclusive scan of the
// lanes' counts).  A leaf's layer is the level of the pass that claims it
// whichever lane wins, so the result does not depend on the queue order.
__global__ __launch_bounds__(FV_BLOCK) void k_si_bfs_push(int n, int lvl, const int *__restrict__ nbr,
                                                          const uint32_t *__restrict__ info,
                                                          const int *__restrict__ fine4, int *__restrict__ X,
                                                          const int *__restrict__ Qin, const int *__restrict__ cin,
                                                          int *__restrict__ Qout, int *__restrict__ cout,
                                                          int *__restrict__ cnext) {
  const int total = *cin;
  const int stride = gridDim.x * blockDim.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) *cnext = 0;   // the counter two layers on is free again
  const int lane = threadIdx.x & 63;
  // every lane of a wave runs the same number of rounds, so the scans see full waves
  const int rounds = (total + stride - 1) / stride;
  for (int r = 0; r < rounds; r++) {
    const int t = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < total;
    const int j = live ? Qin[t] : 0;
    const uint32_t inf = live ? info[j] : 0u;
    int nb[6];
#pragma unroll
    for (int f = 0; f < 6; f++) nb[f] = (live && face_kind(inf, f) != FV_FK_BOUNDARY) ? nbr[(size_t)f * n + j] : -1;
    int ci[24];
#pragma unroll
    for (int f = 0; f < 6; f++) {
      const bool fine = face_kind(inf, f) == FV_FK_FINE;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (nb[f] < 0) ci[4 * f + q] = -1;
        else if (fine) ci[4 * f + q] = fine4[4 * nb[f] + q];
        else ci[4 * f + q] = (q == 0) ? nb[f] : -1;
      }
    }
    int xv[24];
#pragma unroll
    for (int s = 0; s < 24; s++) xv[s] = (ci[s] >= 0) ? X[ci[s]] : 0;
    unsigned took = 0u;
#pragma unroll
    for (int s = 0; s < 24; s++)
      if (ci[s] >= 0 && xv[s] == BFS_INF && atomicCAS(&X[ci[s]], BFS_INF, lvl) == BFS_INF) took |= 1u << s;
    const int cnt = __popc(took);
    int incl = cnt;   // inclusive scan of the counts over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    const int wsum = __shfl(incl, 63, 64);
    int base = 0;
    if (lane == 63 && wsum > 0) base = atomicAdd(cout, wsum);
    base = __shfl(base, 63, 64) + incl - cnt;
#pragma unroll
    for (int s = 0; s < 24; s++)
      if (took & (1u << s)) Qout[base++] = ci[s];
  }
}

__global__ __launch_bounds__(FV_BLOCK) void k_si_bfs_final(int n, int dmax, const int *__restrict__ X,
                                                           int *__restrict__ SI) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = X[i];
  SI[2 * i] = (x <= dmax) ? x : 0;
}

// Two volume fractions (three materials).  The reference's face loop stops at
// the first face where either fraction's layer is set (interface_location.c:
// 60-80), so the two layers are coupled and a leaf is decided by its faces in
// order.  Layer lvl still changes only leaves with a face neighbour in layer
// lvl-1 of some fraction, and such a neighbour is in the frontier of the
// previous layer (the leaves that took a layer there).  So per layer: the
// frontier pushes its neighbours with an unset fraction that the frontier leaf
// has at lvl-1 into a candidate list (each leaf once, by an atomic max on a
// stamp), then each candidate applies the reference's face-ordered rule in
// place, and the ones that took a layer form the next frontier.  In place is
// exact: the rule reads its neighbours' values equal to lvl-1, which no lane
// writes in this layer, and a value being written goes from 0 to lvl.
__global__ __launch_bounds__(FV_BLOCK) void k_si3_seed(int n, const int *__restrict__ SI, int *__restrict__ stamp,
                                                       int *__restrict__ Q, int *__restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  const bool seed = live && (SI[2 * i] == 1 || SI[2 * i + 1] == 1);
  if (live) stamp[i] = 0;
  bfs_append(seed, i, Q, cnt);
}

// the frontier (layer lvl-1) pushes candidates for layer lvl
__global__ __launch_bounds__(FV_BLOCK) void k_si3_push(int n, int lvl, const int *__restrict__ nbr,
                                                       const uint32_t *__restrict__ info,
                                                       const int *__restrict__ fine4, const int *__restrict__ SI,
                                                       int *__restrict__ stamp, const int *__restrict__ Qin,
                                                       const int *__restrict__ cin, int *__restrict__ C,
                                                       int *__restrict__ cc, int *__restrict__ cnext) {
  const int total = *cin;
  const int stride = gridDim.x * blockDim.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) *cnext = 0;   // the frontier counter two layers on is free again
  const int rounds = (total + stride - 1) / stride;       // the same for every lane: full waves append
  for (int r = 0; r < rounds; r++) {
    const int t = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < total;
    const int j = live ? Qin[t] : 0;
    const uint32_t inf = live ? info[j] : 0u;
    const bool w0 = live && SI[2 * j] == lvl - 1, w1 = live && SI[2 * j + 1] == lvl - 1;
    for (int f = 0; f < 6; f++) {
      const int kind = face_kind(inf, f);
      const bool on = live && kind != FV_FK_BOUNDARY;
      const int nb = on ? nbr[(size_t)f * n + j] : 0;
      const int ns = (on && kind == FV_FK_FINE) ? 4 : 1;
      for (int k = 0; k < 4; k++) {
        bool take = false;
        int c = 0;
        if (on && k < ns) {
          c = sub_neighbour(kind, nb, k, fine4);
          const bool want = (w0 && SI[2 * c] == 0) || (w1 && SI[2 * c + 1] == 0);
          take = want && atomicMax(&stamp[c], lvl) < lvl;
        }
        bfs_append(take, c, C, cc);
      }
    }
  }
}

// each candidate applies the rule of interface_location.c:60-80 in place; the
// ones that took a layer are the next frontier
__global__ __launch_bounds__(FV_BLOCK) void k_si3_pull(int n, int lvl, const int *__restrict__ nbr,
                                                       const uint32_t *__restrict__ info,
                                                       const int *__restrict__ fine4, int *__restrict__ SI,
                                                       const int *__restrict__ C, const int *__restrict__ cc,
                                                       int *__restrict__ Qout, int *__restrict__ cout,
                                                       int *__restrict__ ccnext) {
  const int total = *cc;
  const int stride = gridDim.x * blockDim.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) *ccnext = 0;   // the other candidate counter is free again
  const int rounds = (total + stride - 1) / stride;
  for (int r = 0; r < rounds; r++) {
    const int t = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < total;
    const int i = live ? C[t] : 0;
    bool took = false;
    if (live) {
      const uint32_t inf = info[i];
      int si[2] = {SI[2 * i], SI[2 * i + 1]};
      int flag = 0;
      for (int f = 0; f < 6 && !flag; f++) {
        const int kind = face_kind(inf, f);
        if (kind == FV_FK_BOUNDARY) continue;
        const int nb = nbr[(size_t)f * n + i];
        const int ns = (kind == FV_FK_FINE) ? 4 : 1;
        for (int k = 0; k < ns; k++) {
          const int j = sub_neighbour(kind, nb, k, fine4);
          for (int v = 0; v < 2; ++v)
            if (SI[2 * j + v] == (lvl - 1) && si[v] == 0) { si[v] = lvl; flag = 1; }
        }
      }
      if (flag) {
        SI[2 * i] = si[0];
        SI[2 * i + 1] = si[1];
        took = true;
      }
    }
    bfs_append(took, i, Qout, cout);
  }
}

// interface_location (interface_location.c:7-86): interface leaves, then
// imax-2 layers grown one face neighbour at a time; *out is the buffer that
// holds the final SI [n][2].
// layers 2 .. ADAPT_INTER_MAX - 1 of interface_location (interface_location.c:52-86)
// from the first layer in c->SIa; *out: the buffer that holds every layer
int fv_dev_si_layers(fv_dev *c, const fv_crit *cr, int **out) {
  const int n = c->n;
  const dim3 grid((n + FV_BLOCK - 1) / FV_BLOCK), block(FV_BLOCK);
  int imax = 0;   // ADAPT_INTER_MAX (m4_input.c:52-54)
  if (cr->inter0_on == 1 && cr->inter1_on == 1) imax = cr->inter0_lrs + cr->inter1_lrs;
  else if (cr->inter0_on == 1 && cr->inter1_on == 0) imax = cr->inter0_lrs;
  int *cur = c->SIa, *nxt = c->SIb;
  if (c->nmat == 2 && imax > 2) {
    const int dmax = imax - 1;
    int *X = nxt;                                   // [n] layer per leaf, in the spare SI buffer
    int *Q[2] = {c->bfsq, c->bfsq + n};             // current and next layer
    int *cnt = c->iflag;                            // three rotating queue counters
    HIPCHK(hipMemsetAsync(cnt, 0, 3 * sizeof(int), c->stream));
    hipLaunchKernelGGL(k_si_bfs_seed, grid, block, 0, c->stream, n, cur, X, Q[0], cnt);
    for (int lvl = 2; lvl <= dmax; lvl++) {
      const int a = (lvl - 2) % 3, b = (lvl - 1) % 3, z = lvl % 3;
      hipLaunchKernelGGL(k_si_bfs_push, dim3(BFS_GRID), block, 0, c->stream, n, lvl, c->nbr, c->info, c->fine4, X,
                         Q[lvl & 1], cnt + a, Q[(lvl + 1) & 1], cnt + b, cnt + z);
    }
    hipLaunchKernelGGL(k_si_bfs_final, grid, block, 0, c->stream, n, dmax, X, cur);
  } else if (imax > 2) {
    // three materials: frontier, candidates and in-place layers (k_si3_*)
    int *stamp = nxt, *C = nxt + n;                 // the spare SI buffer holds 2n ints
    int *Q[2] = {c->bfsq, c->bfsq + n};
    int *fc = c->iflag, *ccnt = c->iflag + 3;       // three rotating frontier and two candidate counters
    HIPCHK(hipMemsetAsync(c->iflag, 0, 5 * sizeof(int), c->stream));
    hipLaunchKernelGGL(k_si3_seed, grid, block, 0, c->stream, n, cur, stamp, Q[0], fc);
    for (int lvl = 2; lvl < imax; lvl++) {
      const int a = (lvl - 2) % 3, b = (lvl - 1) % 3, z = lvl % 3;
      hipLaunchKernelGGL(k_si3_push, dim3(BFS_GRID), block, 0, c->stream, n, lvl, c->nbr, c->info, c->fine4, cur,
                         stamp, Q[lvl & 1], fc + a, C, ccnt + (lvl & 1), fc + z);
      hipLaunchKernelGGL(k_si3_pull, dim3(BFS3_GRID), block, 0, c->stream, n, lvl, c->nbr, c->info, c->fine4, cur, C,
                         ccnt + (lvl & 1), Q[(lvl + 1) & 1], fc + b, ccnt + ((lvl + 1) & 1));
    }
  }
  HIPCHK(hipGetLastError());
  *out = cur;
  return FV_OK;
}

static int launch_interface_location(fv_dev *c, const fv_crit *cr, int **out) {
  const int n = c->n;
  const dim3 grid((n + FV_BLOCK - 1) / FV_BLOCK), block(FV_BLOCK);
  if (c->nmat == 2)
    hipLaunchKernelGGL(k_si_first<2>, grid, block, 0, c->stream, n, c->P, c->S, c->nbr, c->info, c->fine4, c->SIa);
  else
    hipLaunchKernelGGL(k_si_first<3>, grid, block, 0, c->stream, n, c->P, c->S, c->nbr, c->info, c->fine4, c->SIa);
  return fv_dev_si_layers(c, cr, out);
}

void fv_dev_refresh_kinds_launch(hipStream_t st, int n, const int *nbr, const int *SI, uint32_t *info) {
  hipLaunchKernelGGL(k_refresh_kinds, dim3((n + FV_BLOCK - 1) / FV_BLOCK), dim3(FV_BLOCK), 0, st, n, nbr, SI, info);
}

int fv_dev_indicator_raw(fv_dev *c, const fv_crit *cr, int **SIdev) {
  if (!c || !cr) return set_err(FV_ERR_ARG, "null argument");
  const int n = c->n;
  const dim3 grid((n + FV_BLOCK - 1) / FV_BLOCK), block(FV_BLOCK);
  CritDev C;
  C.uniform_on = cr->uniform_on; C.uniform_lvl = cr->uniform_lvl;
  C.prs_on = cr->prs_on; C.prs_lvl = cr->prs_lvl; C.prs_thr = cr->prs_thr;
  C.rho_on = cr->rho_on; C.rho_lvl = cr->rho_lvl; C.rho_thr = cr->rho_thr;
  C.vel_on = cr->vel_on; C.vel_lvl = cr->vel_lvl; C.vel_thr = cr->vel_thr;
  C.inter0_on = cr->inter0_on; C.inter0_lvl = cr->inter0_lvl; C.inter0_lrs = cr->inter0_lrs;
  C.inter1_on = cr->inter1_on; C.inter1_lvl = cr->inter1_lvl; C.inter1_lrs = cr->inter1_lrs;
  int *cur = nullptr;
  int rc = launch_interface_location(c, cr, &cur);
  if (rc != FV_OK) return rc;
  if (c->nmat == 2)
    hipLaunchKernelGGL(k_split<2>, grid, block, 0, c->stream, n, c->P, C, c->S, c->nbr, c->info, c->fine4, cur, c->split);
  else
    hipLaunchKernelGGL(k_split<3>, grid, block, 0, c->stream, n, c->P, C, c->S, c->nbr, c->info, c->fine4, cur, c->split);
  HIPCHK(hipGetLastError());
  *SIdev = cur;
  return FV_OK;
}

int fv_copy_to_host(fv_dev *c, void *dst, const void *src, size_t bytes) {
  const size_t chunk = (size_t)32 << 20;
  if (!c->pinned) {
    HIPCHK(hipHostMalloc(&c->pinned, chunk, hipHostMallocDefault));
    c->pinned_cap = chunk;
  }
  for (size_t off = 0; off < bytes; off += chunk) {
    const size_t b = (bytes - off < chunk) ? bytes - off : chunk;
    HIPCHK(hipMemcpyAsync(c->pinned, (const char *)src + off, b, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    memcpy((char *)dst + off, c->pinned, b);
  }
  return FV_OK;
}

extern "C" int fv_dev_indicator(fv_dev *c, const fv_crit *cr, int *SI, int *split) {
  FV_GUARD(c);
  if (!c || !cr) return set_err(FV_ERR_ARG, "null argument");
  int *cur = nullptr;
  int rc = fv_dev_indicator_raw(c, cr, &cur);
  if (rc != FV_OK) return rc;
  if (SI && (rc = fv_copy_to_host(c, SI, cur, sizeof(int) * 2 * (size_t)c->n)) != FV_OK) return rc;
  if (split && (rc = fv_copy_to_host(c, split, c->split, sizeof(int) * (size_t)c->n)) != FV_OK) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return FV_OK;
}

extern "C" int fv_dev_interface_location(fv_dev *c, const fv_crit *cr, int *SI) {
  FV_GUARD(c);
  if (!c || !cr) return set_err(FV_ERR_ARG, "null argument");
  const int n = c->n;
  int *cur = nullptr;
  int rc = launch_interface_location(c, cr, &cur);
  if (rc != FV_OK) return rc;
  hipLaunchKernelGGL(k_refresh_kinds, dim3((n + FV_BLOCK - 1) / FV_BLOCK), dim3(FV_BLOCK), 0, c->stream, n, c->nbr,
                     (const int *)cur, c->info);
  HIPCHK(hipGetLastError());
  if (SI && (rc = fv_copy_to_host(c, SI, cur, sizeof(int) * 2 * (size_t)n)) != FV_OK) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return FV_OK;
}

extern "C" int fv_dev_set_limiter_flags(fv_dev *c, const uint8_t *lim) {
  FV_GUARD(c);
  if (!c || !lim) return set_err(FV_ERR_ARG, "null argument");
  HIPCHK(hipMemcpy(c->lim, lim, (size_t)c->n, hipMemcpyHostToDevice));
  return FV_OK;
}

extern "C" int fv_dev_get_limiter_flags(fv_dev *c, uint8_t *lim) {
  FV_GUARD(c);
  if (!c || !lim) return set_err(FV_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(lim, c->lim, (size_t)c->n, hipMemcpyDeviceToHost));
  return FV_OK;
}



// nrep stages back to back (stage 1, 2, .. of the table, cycling); every stage's
// flux kernel is bracketed by its own event pair, and the host waits only
// once, after the last stage, so the overall window holds no host sync
extern "C" int fv_dev_time_stages(fv_dev *c, int nrep, double *ms_stage, double *ms_flux) {
  FV_GUARD(c);
  if (nrep < 1) return set_err(FV_ERR_ARG, "bad argument");
  const int ns = (c->P.rk == FV_RK_SSP3) ? 3 : 2;
  std::vector<hipEvent_t> ev(2 * (size_t)nrep + 2, nullptr);
  int rc = FV_OK;
  for (auto &e : ev) {
    if (hipEventCreate(&e) != hipSuccess) { rc = set_err(FV_ERR_HIP, "hipEventCreate failed"); break; }
  }
  double tot_flux = 0.0;
  float all = 0.f;
  if (rc == FV_OK && hipEventRecord(ev[0], c->stream) != hipSuccess) rc = set_err(FV_ERR_HIP, "hipEventRecord failed");
  for (int r = 0; r < nrep && rc == FV_OK; r++) rc = launch_stage(c, 1 + (r % ns), &ev[2 + 2 * r], &ev[3 + 2 * r]);
  if (rc == FV_OK && hipEventRecord(ev[1], c->stream) != hipSuccess) rc = set_err(FV_ERR_HIP, "hipEventRecord failed");
  if (rc == FV_OK && hipEventSynchronize(ev[1]) != hipSuccess) rc = set_err(FV_ERR_HIP, "hipEventSynchronize failed");
  for (int r = 0; r < nrep && rc == FV_OK; r++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[2 + 2 * r], ev[3 + 2 * r]) != hipSuccess) rc = set_err(FV_ERR_HIP, "event timing failed");
    tot_flux += ms;
  }
  if (rc == FV_OK && hipEventElapsedTime(&all, ev[0], ev[1]) != hipSuccess) rc = set_err(FV_ERR_HIP, "event timing failed");
  for (auto &e : ev)
    if (e) hipEventDestroy(e);
  if (rc != FV_OK) return rc;
  if (ms_stage) *ms_stage = all / nrep;
  if (ms_flux) *ms_flux = tot_flux / nrep;
  return check_nan(c);
}

// stream copy: measured HBM bandwidth on this device and a known-size dispatch
// for calibrating the FETCH_SIZE / WRITE_SIZE counters
__global__ __launch_bounds__(256) void k_copy(size_t n4, const double4 *__restrict__ a, double4 *__restrict__ b) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n4; i += stride) b[i] = a[i];
}

extern "C" int fv_dev_stream_copy(size_t bytes, int reps, double *gbps) {
  if (bytes < 4096 || reps < 1 || !gbps) return set_err(FV_ERR_ARG, "bad argument");
  const size_t n4 = bytes / sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, n4 * sizeof(double4)));
  if (hipMalloc(&b, n4 * sizeof(double4)) != hipSuccess) { (void)hipFree(a); return set_err(FV_ERR_HIP, "alloc"); }
  HIPCHK(hipMemset(a, 0, n4 * sizeof(double4)));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int blocks = 256 * 16;
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, n4, a, b);
  HIPCHK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, n4, a, b);
  HIPCHK(hipEventRecord(e1, 0));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *gbps = 2.0 * (double)(n4 * sizeof(double4)) * reps / (ms * 1e-3) / 1e9;
  hipEventDestroy(e0); hipEventDestroy(e1);
  HIPCHK(hipFree(a));
  HIPCHK(hipFree(b));
  return FV_OK;
}

extern "C" const char *fv_dev_last_error(void) { return fv_g_err; }

// Host-only self-check of the failure handling of a context (no counterpart in
// the reference).  1) A context marked broken is refused by the entry points
// that launch kernels or copy state, before any HIP call.  2) A reservation
// that cannot be satisfied leaves the context's buffers, capacity and size as
// they were.  Returns FV_OK when both hold, else the number of the failed check.
extern "C" int fv_dev_selftest_guard(void) {
  fv_dev *c = new fv_dev();
  c->n = 4; c->neq = 9; c->nmat = 2; c->cap_n = 4; c->cap_fine = 1;
  double buf[64] = {0};
  int SI[8], split[4], nout = 0;
  fv_crit cr;
  memset(&cr, 0, sizeof cr);
  double ms0 = 0.0, ms1 = 0.0, dt = 0.0;
  int bad = 0;
  c->broken = 1;
  if (fv_dev_stage(c, 1) != FV_ERR_HIP) bad = 1;
  else if (fv_dev_advance(c, 1) != FV_ERR_HIP) bad = 2;
  else if (fv_dev_begin_step(c) != FV_ERR_HIP) bad = 3;
  else if (fv_dev_set_state(c, buf) != FV_ERR_HIP) bad = 4;
  else if (fv_dev_get_state(c, buf) != FV_ERR_HIP) bad = 5;
  else if (fv_dev_cfl_dt(c, 0.4, &dt) != FV_ERR_HIP) bad = 6;
  else if (fv_dev_totals(c, buf) != FV_ERR_HIP) bad = 7;
  else if (fv_dev_indicator(c, &cr, SI, split) != FV_ERR_HIP) bad = 8;
  else if (fv_dev_interface_location(c, &cr, SI) != FV_ERR_HIP) bad = 9;
  else if (fv_dev_regrid(c, &cr, 5, &nout) != FV_ERR_HIP) bad = 10;
  else if (fv_dev_time_stages(c, 3, &ms0, &ms1) != FV_ERR_HIP) bad = 11;
  else if (fv_dev_criterionsmth(c, 5, split) != FV_ERR_HIP) bad = 12;
  if (!bad) {
    c->broken = 0;
    double *sentinel = reinterpret_cast<double *>(buf);
    c->S = sentinel;
    // INT_MAX leaves need far more than any device holds, so some hipMalloc fails
    const int rc = fv_dev_reserve(c, 0x7fffffff, 1);
    if (rc == FV_OK) bad = 13;
    else if (c->S != sentinel || c->cap_n != 4 || c->n != 4 || c->broken) bad = 14;
    c->S = nullptr;
  }
  delete c;
  return bad;
}

// ---------------------------------------------------------------------
// fv_cfl_dt_noalloc: fv_dev_cfl_dt's kernels on caller-owned device arrays,
// with the per-block partial minima in a fixed device array instead of a
// context buffer, so that the call allocates nothing.
// ---------------------------------------------------------------------
#include <mutex>

// per-block partial minima of k_cfl: FV_CFL_NOALLOC_MAX_LEAVES / FV_BLOCK doubles
// (2 MiB), held by every device the library runs on until the process ends
static __device__ __align__(256) double scratch[FV_CFL_NOALLOC_MAX_LEAVES / FV_BLOCK];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {nullptr};

// the address of this device's copy of `scratch`, looked up once per device
// with that device current (the array's name is only a host-side stand-in)
int scratch_of_current_device(double **out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_err(FV_ERR_ARG, "fv_cfl_dt_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return set_err(FV_ERR_HIP, "fv_cfl_dt_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return FV_OK;
}
}  // namespace

extern "C" int fv_cfl_dt_noalloc(const fv_params *p, const fv_geo *g, int n, const double *S, const uint32_t *info,
                                 double cfl, double *dt, void *stream) {
  if (!p || !g || !S || !info || !dt) return set_err(FV_ERR_ARG, "fv_cfl_dt_noalloc: null argument");
  if (p->nmat != 2 && p->nmat != 3) return set_err(FV_ERR_ARG, "fv_cfl_dt_noalloc: nmat must be 2 or 3");
  if (n < 1) return set_err(FV_ERR_ARG, "fv_cfl_dt_noalloc: no leaves");
  if (n > FV_CFL_NOALLOC_MAX_LEAVES) {
    snprintf(fv_g_err, sizeof fv_g_err, "fv_cfl_dt_noalloc: %d leaves exceed the scratch array's %d", n,
             FV_CFL_NOALLOC_MAX_LEAVES);
    return FV_ERR_ARG;
  }
  double *part = nullptr;
  int rc = scratch_of_current_device(&part);
  if (rc != FV_OK) return rc;
  Params P;
  Geo G;
  to_dev_params(p, P);
  to_dev_geo(g, G);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = (n + FV_BLOCK - 1) / FV_BLOCK;
  // every one of the nb partials is written by k_cfl before k_min_final reads it,
  // so the contents left by an earlier call need no clearing
  if (p->nmat == 2)
    hipLaunchKernelGGL(k_cfl<2>, dim3(nb), dim3(FV_BLOCK), 0, st, n, P, G, S, info, part);
  else
    hipLaunchKernelGGL(k_cfl<3>, dim3(nb), dim3(FV_BLOCK), 0, st, n, P, G, S, info, part);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(1024), 0, st, nb, part, dt, cfl);
  HIPCHK(hipGetLastError());
  return FV_OK;
}
