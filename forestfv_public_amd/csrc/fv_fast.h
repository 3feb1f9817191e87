// fv_fast.h -- throughput form of the model-4 face physics for gfx950.
//
// Same mathematics as fv_phys.h (which keeps the reference's operation
// order for bit-identical FP64 parity), rearranged for the VALU:
//  - divisions by a shared denominator become one reciprocal and products
//    (v_rcp_f64 + two Newton steps in FP64, v_rcp_f32 in FP32);
//  - per-material constants G, p_inf, G p_inf, 1/(G-1), G p_inf/(G-1) are
//    formed once per thread;
//  - the HLLC star state uses the identities ar*_k = ar_k f, r* = r f with
//    f = (S_K - u_K)/(S_K - u*), are*_k = alpha_k (p*_k + G p_inf)/(G - 1),
//    and the face frames are orthonormal (s12 = s13 = 0 for diagonal stress);
//  - total energy is carried as rE = rE' + Pi with Pi = sum_k alpha_k G_k
//    p_inf_k/(G_k - 1): the large stiffened-gas constant is fluxed
//    separately as Pi* u* (it is advected with the star density), so that an
//    FP32 face kernel never forms rE* - rE where both are ~4e8 J/m^3.
// The real type R of the face kernels is double (plain path) or float
// (mixed precision); pressure recovery from energy and the RK accumulation
// are always FP64.
#pragma once
#include "fv_phys.h"

// the throughput form lets the compiler fuse multiply-adds
#pragma clang fp contract(fast)

namespace fvf {
using fvk::Lay;
using fvk::Params;
using fvk::Geo;

FV_DEV double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
FV_DEV float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
FV_DEV double sqrtr(double x) { return sqrt(x); }
FV_DEV float sqrtr(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename R> FV_DEV R absr(R x) { return x < (R)0 ? -x : x; }

// per-material constants
template <int NM, typename R>
struct Mat {
  R G[NM], P[NM], GP[NM], Gm1[NM], iGm1[NM], C[NM], iG[NM];
};

template <int NM, typename R>
FV_DEV Mat<NM, R> make_mat(const Params &Pr) {
  Mat<NM, R> M;
#pragma unroll
  for (int k = 0; k < NM; k++) {
    if constexpr (sizeof(R) == 8) {
      M.G[k] = Pr.gamma[k]; M.P[k] = Pr.pinf[k]; M.GP[k] = Pr.GP[k]; M.Gm1[k] = Pr.Gm1[k];
      M.iGm1[k] = Pr.iGm1[k]; M.C[k] = Pr.C[k]; M.iG[k] = Pr.iG[k];
    } else {
      M.G[k] = Pr.Gf[k]; M.P[k] = Pr.Pf[k]; M.GP[k] = Pr.GPf[k]; M.Gm1[k] = Pr.Gm1f[k];
      M.iGm1[k] = Pr.iGm1f[k]; M.C[k] = Pr.Cf[k]; M.iG[k] = Pr.iGf[k];
    }
  }
  return M;
}

// face/cell state as the HLLC needs it
template <int NM, typename R>
struct SolF {
  R r, ra[NM], u, v, w;
  R avf[NM], p[NM], are[NM];
  R st, c2, rEp, Pi;
};

// from a conservative FP64 vector (m4_facevalues / m4_facercnstr)
template <int NM, typename R>
FV_DEV void fill_cons(const Mat<NM, double> &M, double amin100, const double *X, SolF<NM, R> &s) {
  using L = Lay<NM>;
  const double a = amin100;
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) r += X[k];
  const double ir = rcp(r);
  double at[NM], sum = 0.0;
#pragma unroll
  for (int k = 0; k < L::N3; ++k) {
    at[k] = X[L::I3 + k];
    if (at[k] <= a) at[k] = a;
    else if (at[k] >= 1.0 - a) at[k] = 1.0 - a;
    sum += at[k];
  }
  if (sum > 1.0 - a) {
    const double dif = sum - (1.0 - a), is = rcp(sum);
#pragma unroll
    for (int j = 0; j < L::N3; ++j) at[j] -= at[j] * is * dif;
    sum = 1.0 - a;
  }
  at[L::N3] = 1.0 - sum;
  double st = 0.0, c2 = 0.0, Pi = 0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    const double pk = M.Gm1[k] * X[L::I4 + k] * rcp(at[k]) - M.GP[k];
    c2 += at[k] * ir * M.G[k] * (pk + M.P[k]);
    st -= at[k] * pk;
    Pi += at[k] * M.C[k];
    s.ra[k] = (R)X[k];
    s.avf[k] = (R)at[k];
    s.p[k] = (R)pk;
    s.are[k] = (R)X[L::I4 + k];
  }
  s.r = (R)r;
  s.u = (R)(X[L::I1] * ir);
  s.v = (R)(X[L::I1 + 1] * ir);
  s.w = (R)(X[L::I1 + 2] * ir);
  s.st = (R)st;
  s.c2 = (R)c2;
  s.Pi = (R)Pi;
  s.rEp = (R)(X[L::I2] - Pi);
}

// Mixed-precision face record: ra[NM], u, v, w, alpha[N3] (clipped), p[NM], are[NM], rE - Pi
template <int NM>
struct Rec {
  static constexpr int N = 4 * NM + 3;
  static constexpr int RA = 0, U = NM, AL = NM + 3, P = NM + 3 + (NM - 1), ARE = P + NM, REP = ARE + NM;
};

template <int NM>
FV_DEV void to_record(const Mat<NM, double> &M, double amin100, const double *X, float *rec) {
  SolF<NM, double> s;
  fill_cons<NM, double>(M, amin100, X, s);
  using Q = Rec<NM>;
#pragma unroll
  for (int k = 0; k < NM; k++) {
    rec[Q::RA + k] = (float)s.ra[k];
    rec[Q::P + k] = (float)s.p[k];
    rec[Q::ARE + k] = (float)s.are[k];
  }
#pragma unroll
  for (int k = 0; k < NM - 1; k++) rec[Q::AL + k] = (float)s.avf[k];
  rec[Q::U] = (float)s.u;
  rec[Q::U + 1] = (float)s.v;
  rec[Q::U + 2] = (float)s.w;
  rec[Q::REP] = (float)s.rEp;
}

template <int NM>
FV_DEV void fill_rec(const Mat<NM, float> &M, const float *rec, SolF<NM, float> &s) {
  using Q = Rec<NM>;
  float r = 0.f, sum = 0.f;
#pragma unroll
  for (int k = 0; k < NM; k++) { s.ra[k] = rec[Q::RA + k]; r += s.ra[k]; }
#pragma unroll
  for (int k = 0; k < NM - 1; k++) { s.avf[k] = rec[Q::AL + k]; sum += s.avf[k]; }
  s.avf[NM - 1] = 1.f - sum;
  s.r = r;
  s.u = rec[Q::U];
  s.v = rec[Q::U + 1];
  s.w = rec[Q::U + 2];
  const float ir = rcp(r);
  float st = 0.f, c2 = 0.f, Pi = 0.f;
#pragma unroll
  for (int k = 0; k < NM; k++) {
    s.p[k] = rec[Q::P + k];
    s.are[k] = rec[Q::ARE + k];
    st -= s.avf[k] * s.p[k];
    c2 += s.avf[k] * ir * M.G[k] * (s.p[k] + M.P[k]);
    Pi += s.avf[k] * M.C[k];
  }
  s.st = st;
  s.c2 = c2;
  s.Pi = Pi;
  s.rEp = rec[Q::REP];
}

// The boundary ghost sR (a copy of sL on entry): the mirrored velocity of
// codes 1, 4, 5 (bc_symmetry.c, bc_reflective.c:12-28, bc_wall.c:12-26) and,
// with bc_fixed_p, every material at the outlet (2) / inlet (3) pressure of
// m4_bc_outlet / m4_bc_inlet for CASE 3400/3410 (m4_bc.c:44-113, 121-186).
// Only the XR kernel instances carry the fixed-pressure branch: compiled into
// the others it costs their flux kernels 14 VGPRs and a spill.
template <int NM, typename R, bool XR>
FV_DEV void ghost(const Mat<NM, R> &M, const Params &P, int bcode, const R *fr, const SolF<NM, R> &sL,
                  SolF<NM, R> &sR) {
  if (bcode == 1 || bcode == 4 || bcode == 5) {
    const R ul = sL.u * fr[0] + sL.v * fr[1] + sL.w * fr[2];
    const R vl = sL.u * fr[3] + sL.v * fr[4] + sL.w * fr[5];
    const R wl = sL.u * fr[6] + sL.v * fr[7] + sL.w * fr[8];
    const R vr = (bcode == 5) ? -vl : vl, wr = (bcode == 5) ? -wl : wl;
    sR.u = -ul * fr[0] + vr * fr[3] + wr * fr[6];
    sR.v = -ul * fr[1] + vr * fr[4] + wr * fr[7];
    sR.w = -ul * fr[2] + vr * fr[5] + wr * fr[8];
    return;
  }
  if constexpr (XR) {
    if (!(P.bc_fixed_p && (bcode == 2 || bcode == 3))) return;
    const R pb = (R)(bcode == 3 ? P.p_inlet : P.p_outlet);
    const R ir = rcp(sR.r);
    R re = (R)0, c2 = (R)0;
#pragma unroll
    for (int k = 0; k < NM; k++) {
      sR.p[k] = pb;
      sR.are[k] = sR.avf[k] * (pb + M.GP[k]) * M.iGm1[k];
      c2 += sR.avf[k] * ir * M.G[k] * (pb + M.P[k]);
      re += sR.are[k];
    }
    re += (R)0.5 * sR.r * (sR.u * sR.u + sR.v * sR.v + sR.w * sR.w);
    sR.st = -pb;
    sR.c2 = c2;
    sR.rEp = re - sR.Pi;
  }
}

// HLLC n-flux of one face (m4_hllc.c, m4_wave_speeds.c, m4_star_region.c, m4_flux.c)
template <int NM, typename R>
FV_DEV void hllc(const Mat<NM, R> &M, int fix, const R *F, int bc, const SolF<NM, R> &sL,
                 const SolF<NM, R> &sR, R *out, R &un, R &vn, R &wn) {
  using L = Lay<NM>;
  const R nx = F[0], ny = F[1], nz = F[2], mx = F[3], my = F[4], mz = F[5], lx = F[6], ly = F[7], lz = F[8];
  const R uL = sL.u * nx + sL.v * ny + sL.w * nz, vL = sL.u * mx + sL.v * my + sL.w * mz,
          wL = sL.u * lx + sL.v * ly + sL.w * lz;
  const R uR = sR.u * nx + sR.v * ny + sR.w * nz, vR = sR.u * mx + sR.v * my + sR.w * mz,
          wR = sR.u * lx + sR.v * ly + sR.w * lz;
  const R cL = sqrtr(sL.c2), cR = sqrtr(sR.c2);
  const R SRs = fvk::maxr(uL + cL, uR + cR);
  const R SLs = fvk::minr(uL - cL, uR - cR);
  const R s11L = sL.st, s11R = sR.st;
  const R ruL = sL.r * uL, ruR = sR.r * uR;
  R u_S = ((sL.r * uL * uL - s11L) - (sR.r * uR * uR - s11R) - SLs * ruL + SRs * ruR) *
          rcp(ruL - ruR - SLs * sL.r + SRs * sR.r);
  if (bc == 4) u_S = (R)0;
  int region;
  if ((R)0 <= SLs) region = 0;
  else if (SLs <= (R)0 && (R)0 <= u_S) region = 1;
  else if (u_S <= (R)0 && (R)0 <= SRs) region = 2;
  else if (SRs <= (R)0) region = 3;
  else region = -1;
  if (region == 0 || region == 3 || region < 0) {
    const SolF<NM, R> &s = (region == 3) ? sR : sL;
    const R uK = (region == 3) ? uR : uL;
#pragma unroll
    for (int k = 0; k < NM; k++) out[k] = s.ra[k] * uK;
    out[L::I1 + 0] = s.r * s.u * uK - s.st * nx;
    out[L::I1 + 1] = s.r * s.v * uK - s.st * ny;
    out[L::I1 + 2] = s.r * s.w * uK - s.st * nz;
    out[L::I2] = (s.rEp + s.Pi - s.st) * uK;
#pragma unroll
    for (int k = 0; k < L::N3; ++k) out[L::I3 + k] = uK * s.avf[k];
#pragma unroll
    for (int k = 0; k < NM; ++k) out[L::I4 + k] = uK * s.are[k];
    un = s.u; vn = s.v; wn = s.w;
    if (region < 0) {
#pragma unroll
      for (int k = 0; k < L::NEQ; k++) out[k] = (R)NAN;
    }
    return;
  }
  const bool left = (region == 1);
  const SolF<NM, R> &s = left ? sL : sR;
  const R SK = left ? SLs : SRs, uK = left ? uL : uR, vK = left ? vL : vR, wK = left ? wL : wR;
  const R s11K = left ? s11L : s11R;
  const R iD = rcp(SK - u_S);
  const R fac = (SK - uK) * iD;
  const R r_S = s.r * fac;
  const R s11_S = ((uR - SRs) * sR.r * s11L - (uL - SLs) * sL.r * s11R + (uL - SLs) * sL.r * (uR - SRs) * sR.r * (uR - uL)) *
                  rcp((uR - SRs) * sR.r - (uL - SLs) * sL.r);
  // star velocity back to the global frame (v*, w* = vK, wK for diagonal stress)
  const R vel0 = u_S * nx + vK * mx + wK * lx;
  const R vel1 = u_S * ny + vK * my + wK * ly;
  const R vel2 = u_S * nz + vK * mz + wK * lz;
  // r* E'* = (rE'(uK - SK) -/+ s11K uK + s11* u*) / (u* - SK)   (m4_star_region.c:46)
  const R sw = (fix & 2) ? -(s11K * uK) : (s11K * uK);
  const R rEp_S = -(s.rEp * (uK - SK) + sw + s11_S * u_S) * iD;
  // F_K + S_K (U*_K - U_K) with the physical flux of state K (m4_hllc.c:30-48)
#pragma unroll
  for (int k = 0; k < NM; k++) out[k] = s.ra[k] * uK + SK * (s.ra[k] * fac - s.ra[k]);
  out[L::I1 + 0] = s.r * s.u * uK - s.st * nx + SK * (r_S * vel0 - s.r * s.u);
  out[L::I1 + 1] = s.r * s.v * uK - s.st * ny + SK * (r_S * vel1 - s.r * s.v);
  out[L::I1 + 2] = s.r * s.w * uK - s.st * nz + SK * (r_S * vel2 - s.r * s.w);
  out[L::I2] = (s.rEp - s.st) * uK + SK * (rEp_S - s.rEp) + s.Pi * fac * u_S;
#pragma unroll
  for (int k = 0; k < L::N3; ++k) out[L::I3 + k] = u_S * s.avf[k];
  // p*_k = (p_k + pinf)(((G-1) + (G+1) f) / ((G-1) f + (G+1))) - pinf, are*_k = alpha_k (p*_k + G pinf)/(G-1)
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    const R Gp1 = M.Gm1[k] + (R)2;
    const R pS = (s.p[k] + M.P[k]) * ((M.Gm1[k] + Gp1 * fac) * rcp(M.Gm1[k] * fac + Gp1)) - M.P[k];
    out[L::I4 + k] = u_S * (s.avf[k] * (pS + M.GP[k]) * M.iGm1[k]);
  }
  un = vel0; vn = vel1; wn = vel2;
}

// m4_relaxationvec.c for PRELAX 0 and 1 (the Newton forms PRELAX 2 and 3 use fvk::relax)
// XR selects the iterative relaxations (PRELAX 2 and 3), which go to the exact
// Newton forms; a kernel instantiated without them carries none of their code
// and registers
template <int NM, bool XR>
FV_DEV void relax_fast(const Params &Pr, const Mat<NM, double> &M, double *vec) {
  using L = Lay<NM>;
  if constexpr (XR) {
    fvk::relax<NM>(Pr, vec);
    return;
  }
  const double a = Pr.amin / 100.0;
  double ar[NM], are[NM], avf[NM + 1], rho = 0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) { ar[k] = vec[k]; are[k] = vec[L::I4 + k]; rho += ar[k]; }
#pragma unroll
  for (int k = 0; k < L::N3; ++k) avf[k] = vec[L::I3 + k];
  fvk::vf_limit_relax<NM>(a, avf);
  if (NM == 2 && Pr.prelax == 1) {
    const double ia0 = rcp(avf[0]), ia1 = rcp(avf[1]);
    const double p00 = are[0] * M.Gm1[0] * ia0 - M.GP[0];
    const double p01 = are[1] * M.Gm1[1] * ia1 - M.GP[1];
    const double w0 = avf[0] * M.iG[0], w1 = avf[1] * M.iG[1];
    const double iw = rcp(w0 + w1);
    const double A1 = w0 * (p00 + M.P[0]) * iw, A2 = w1 * (p01 + M.P[1]) * iw;
    const double d = A2 - A1 - (M.P[1] - M.P[0]);
    const double pe = 0.5 * (A1 + A2 - (M.P[0] + M.P[1])) + sqrt(0.25 * d * d + A1 * A2);
    avf[0] = avf[0] * (p00 + M.GP[0] + M.Gm1[0] * pe) * rcp(M.G[0] * pe + M.GP[0]);
    avf[1] = avf[1] * (p01 + M.GP[1] + M.Gm1[1] * pe) * rcp(M.G[1] * pe + M.GP[1]);
    fvk::vf_limit_relax<NM>(a, avf);
    const double irho = rcp(rho);
    const double u = vec[L::I1] * irho, v = vec[L::I1 + 1] * irho, w = vec[L::I1 + 2] * irho;
    const double e_hyd = vec[L::I2] * irho - 0.5 * (u * u + v * v + w * w);
    double num = 0.0, den = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) { num += avf[k] * M.C[k]; den += avf[k] * M.iGm1[k]; }
    const double pr = (rho * e_hyd - num) * rcp(den);
#pragma unroll
    for (int k = 0; k < NM; ++k) are[k] = avf[k] * (pr + M.GP[k]) * M.iGm1[k];
  }
#pragma unroll
  for (int k = 0; k < L::N3; ++k) vec[L::I3 + k] = avf[k];
#pragma unroll
  for (int k = 0; k < NM; ++k) vec[L::I4 + k] = are[k];
}

// relax_fast for PRELAX=1 (two phases) followed by the FP32 face record,
// written from the relaxed quantities: both phasic pressures equal the
// recovered pressure, and rE - Pi uses the Pi of the recovery, so the record
// needs no second conversion of the face state.  vec is relaxed in place too.
template <int NM>
FV_DEV void relax_record(const Params &Pr, const Mat<NM, double> &M, double *vec, float *rec) {
  using L = Lay<NM>;
  using Q = Rec<NM>;
  const double a = Pr.amin / 100.0;
  double ar[NM], are[NM], avf[NM + 1], rho = 0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) { ar[k] = vec[k]; are[k] = vec[L::I4 + k]; rho += ar[k]; }
  avf[0] = vec[L::I3];
  fvk::vf_limit_relax<NM>(a, avf);
  const double ia0 = rcp(avf[0]), ia1 = rcp(avf[1]);
  const double p00 = are[0] * M.Gm1[0] * ia0 - M.GP[0];
  const double p01 = are[1] * M.Gm1[1] * ia1 - M.GP[1];
  const double w0 = avf[0] * M.iG[0], w1 = avf[1] * M.iG[1];
  const double iw = rcp(w0 + w1);
  const double A1 = w0 * (p00 + M.P[0]) * iw, A2 = w1 * (p01 + M.P[1]) * iw;
  const double d = A2 - A1 - (M.P[1] - M.P[0]);
  const double pe = 0.5 * (A1 + A2 - (M.P[0] + M.P[1])) + sqrt(0.25 * d * d + A1 * A2);
  avf[0] = avf[0] * (p00 + M.GP[0] + M.Gm1[0] * pe) * rcp(M.G[0] * pe + M.GP[0]);
  avf[1] = avf[1] * (p01 + M.GP[1] + M.Gm1[1] * pe) * rcp(M.G[1] * pe + M.GP[1]);
  fvk::vf_limit_relax<NM>(a, avf);
  const double irho = rcp(rho);
  const double u = vec[L::I1] * irho, v = vec[L::I1 + 1] * irho, w = vec[L::I1 + 2] * irho;
  const double e_hyd = vec[L::I2] * irho - 0.5 * (u * u + v * v + w * w);
  double num = 0.0, den = 0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) { num += avf[k] * M.C[k]; den += avf[k] * M.iGm1[k]; }
  const double pr = (rho * e_hyd - num) * rcp(den);
#pragma unroll
  for (int k = 0; k < NM; ++k) are[k] = avf[k] * (pr + M.GP[k]) * M.iGm1[k];
  vec[L::I3] = avf[0];
#pragma unroll
  for (int k = 0; k < NM; ++k) vec[L::I4 + k] = are[k];
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    rec[Q::RA + k] = (float)ar[k];
    rec[Q::P + k] = (float)pr;
    rec[Q::ARE + k] = (float)are[k];
  }
  rec[Q::AL] = (float)avf[0];
  rec[Q::U] = (float)u;
  rec[Q::U + 1] = (float)v;
  rec[Q::U + 2] = (float)w;
  rec[Q::REP] = (float)(vec[L::I2] - num);
}

// relaxation of a reconstructed face state and its FP32 record
template <int NM, bool XR>
FV_DEV void relax_to_record(const Params &Pr, const Mat<NM, double> &M, double amin100, double *vec, float *rec) {
  if constexpr (NM == 2 && !XR) {
    if (Pr.prelax == 1) { relax_record<NM>(Pr, M, vec, rec); return; }
  }
  relax_fast<NM, XR>(Pr, M, vec);
  to_record<NM>(M, amin100, vec, rec);
}

// limited slope (m4_reconctruct.c:105-128) with reciprocal distances; a
// boundary on either side gives zero slope, as the reference's bc test and
// its 0/0 on the zero boundary distance do
template <typename R>
FV_DEV R slope(int limiter, double df, double db, R idp, R idm) {
  if (!(fabs(df) > 1.0e-10)) return (R)0;
  const R a = -(R)db * idm, b = (R)df * idp;
  if (limiter == 0) {
    if (a * b > (R)0) return (absr(a) > absr(b)) ? b : a;
    return (R)0;
  }
  if (a * b > (R)0) {
    const R c = (R)0.5 * (a + b);
    const R m2 = fvk::minr((R)2 * absr(a), (R)2 * absr(b));
    const R mag = fvk::minr(m2, absr(c));
    return c > (R)0 ? mag : -mag;
  }
  return (R)0;
}

}  // namespace fvf

#pragma clang fp contract(off)
