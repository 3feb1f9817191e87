// fv_m5.hip -- model-5 ghost-fluid operators for gfx950 behind include/forestfv_m5.h.
//
// Three pieces of the reference's model-5 stage, each on its own data layout:
//  * k_m5_gfm_faces: one lane per interface sub-face, SoA face states
//    [NEQ][nf]; builds both face records (m5_facercnstr.c with the elastic EOS
//    of state_equation.c), the ghost state, wave speeds, star region and the
//    area-weighted face flux (m5_hllc_gfm.c, m5_wave_speeds_gfm.c,
//    m5_flux_gfm.c, m5_star_region.c, m5_residual.c:166-170).
//  * k_m5_rk: one lane per leaf, SoA state [NEQ][n]; the m5_RK stage
//    combinations, the positivity fallback and the ZEROU freezes.  Stage 1
//    marks the leaves whose solid fraction crosses AVF_LIM.
//  * k_m5_gfm_seq / k_m5_rel_seq: the marked leaves, compacted in leaf order
//    (hipcub DeviceSelect::Flagged), re-projected by one lane in leaf order.
//    The reference updates leaves in place in leaf order, so a marked leaf
//    sees its earlier neighbours' new states and its later neighbours' old
//    ones; the sequential pass over the (few) marked leaves reproduces that
//    exactly while every other leaf is updated in parallel.
//
// Every function keeps the reference's operation order and the library is
// built with -ffp-contract=off, so results are bit-identical to the CPU
// restatement except through pow(|G|, 1/3) and pow(|G|, 1/6) of the elastic
// EOS, where the device and glibc pow may differ in the last bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <vector>
#include "forestfv_m5.h"
#include "fv_internal.h"

namespace {

using fvk::Geo;

struct M5P {
  int nmat;
  double gamma[3], pinf[3], mush[3];
  double amin, avf_lim, dt;
  int zerou[3];
  int prelax, order;
  double yield;   // von Mises yield stress of perfect plasticity, 0 = elastic
  int fix;        // 1: the consistent update (FV_M5_FIX_SIGNS), 0: the published one
  int mix;        // 1: elastic energy and shear stress weighted by the volume fraction (FV_M5_FIX_MIXTURE)
  int limiter;    // MUSCL slope: 0 minmod, 1 monotonised central, 2 per leaf (fv_m5_dev.lim)
  double dtu;     // signed step of the update: S = SN + dtu RHS / V, +dt published, -dt consistent
  double sg;      // sign of sigma in the momentum flux: +1 published (rho u u + sigma), -1 consistent
  int fast;       // 1 (FP64 and mixed stages): the powers below in cheaper forms within a few ulp
};

template <int NM>
struct L5 {
  static constexpr int NEQ = 3 * NM + 12;
  static constexpr int NCONS = NM + 4;
  static constexpr int I0 = 0, I1 = NM, I2 = NM + 3, I3 = NM + 4, N3 = NM - 1, I4 = NM + 4 + NM - 1, I5 = I4 + NM;
};

template <typename R>
__device__ __forceinline__ R sq(R x) { return x * x; }
__device__ __forceinline__ bool has_solid(const M5P &P) { return P.mush[0] != 0.0 || P.mush[1] != 0.0 || P.mush[2] != 0.0; }
template <typename R>
__device__ __forceinline__ R maxr(R a, R b) { return (a > b) ? a : b; }   // strdata.h:27
template <typename R>
__device__ __forceinline__ R minr(R a, R b) { return (a < b) ? a : b; }   // strdata.h:31

// pow(|G|, 1/3) and pow(|G|, 1/6), which the elastic energy and the stress of
// every solid material of one state use: evaluated once per state
template <typename R = double>
struct GPow {
  R g13, g16;
};
template <typename R>
__device__ GPow<R> gpow(R Gmag) { return GPow<R>{pow(fabs(Gmag), R(1.0) / R(3.0)), pow(fabs(Gmag), R(1.0) / R(6.0))}; }
// P.fast: |G|^(1/3) = cbrt |G| and |G|^(1/6) = sqrt(cbrt |G|), each within two
// ulp of pow, at a fraction of its cost (the exact stage keeps pow)
template <typename R>
__device__ GPow<R> gpow(const M5P &P, R Gmag) {
  if (P.fast) {
    const R c = cbrt(fabs(Gmag));
    return GPow<R>{c, sqrt(c)};
  }
  return gpow(Gmag);
}

// FAST (the FP64 and mixed residual): a / b as a times a reciprocal from
// v_rcp and two Newton steps (FP64, within an ulp or two of the quotient) or
// v_rcp_f32, with one reciprocal shared by the quotients of one denominator;
// the denominators it is used on are nonzero in every admissible state
__device__ __forceinline__ double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
template <bool FAST, typename R>
__device__ __forceinline__ R dv(R a, R b) {
  if constexpr (FAST) return a * frcp(b);
  else return a / b;
}
// the powers of FAST: cbrt forms whatever P.fast says
template <bool FAST, typename R>
__device__ __forceinline__ GPow<R> gpow_t(const M5P &P, R Gmag) {
  if constexpr (FAST) {
    const R c = cbrt(fabs(Gmag));
    return GPow<R>{c, sqrt(c)};
  } else {
    return gpow(P, Gmag);
  }
}

// m5_elastic_energy_solid_V2, state_equation.c:269-316
template <typename R, bool FAST = false>
__device__ R elastic_energy_v2(const R *G, R Gmag, const GPow<R> &gp, R mu) {
  R eints;
  if (fabs(Gmag) != R(0.0)) {
    R M1[9], M2[9];
    R Gmag13 = gp.g13;
    if (isnan(Gmag13)) Gmag13 = R(0.0);
#pragma unroll
    for (int k = 0; k < 9; k++) { M1[k] = G[k]; M2[k] = R(0.0); }
    M1[0] = G[0] - Gmag13; M1[4] = G[4] - Gmag13; M1[8] = G[8] - Gmag13;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) M2[i * 3 + j] += M1[i * 3 + k] * M1[k * 3 + j];
    const R tr1 = M2[0] + M2[4] + M2[8];
    eints = dv<FAST>(mu, R(4.0) * gp.g16) * tr1;
  } else {
    eints = R(0.0);
  }
  return eints;
}

// G_det (state_equation.c:206-215) and m5_elastic_energy_solid (:217-267)
template <typename R, bool FAST = false>
__device__ R elastic_energy_v1(const M5P &P, const R *G, R mu) {
  const R Gmag = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) +
                      G[2] * (G[3] * G[7] - G[4] * G[6]);
  if ((fabs(Gmag) != R(0.0)) && (mu != R(0.0))) return elastic_energy_v2<R, FAST>(G, Gmag, gpow_t<FAST>(P, Gmag), mu);
  return R(0.0);
}

// m5_eos_stress_V2, state_equation.c:143-204: the stress s of one material
template <typename R, bool FAST = false>
__device__ void stress_v2(R *s, R avf, R pr, const R *G, R Gmag, const GPow<R> &gp, R mu) {
  if ((fabs(Gmag) != R(0.0)) && (mu != R(0.0))) {
    R G2[9], ss[9];
#pragma unroll
    for (int k = 0; k < 9; k++) G2[k] = R(0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) G2[i * 3 + j] += G[i * 3 + k] * G[k * 3 + j];
    const R J1 = G[0] + G[4] + G[8];
    const R J2 = G2[0] + G2[4] + G2[8];
    const R g13 = gp.g13;
    const R J2t = FAST ? J2 * R(1.0 / 3.0) : J2 / R(3.0), J1t = FAST ? J1 * R(1.0 / 3.0) : J1 / R(3.0);
    ss[0] = G2[0] - J2t - g13 * (G[0] - J1t);
    ss[1] = G2[1] - g13 * (G[1]);
    ss[2] = G2[2] - g13 * (G[2]);
    ss[3] = G2[3] - g13 * (G[3]);
    ss[4] = G2[4] - J2t - g13 * (G[4] - J1t);
    ss[5] = G2[5] - g13 * (G[5]);
    ss[6] = G2[6] - g13 * (G[6]);
    ss[7] = G2[7] - g13 * (G[7]);
    ss[8] = G2[8] - J2t - g13 * (G[8] - J1t);
    const R k = dv<FAST>(mu, gp.g16);
#pragma unroll
    for (int q = 0; q < 9; q++) s[q] = (q == 0 || q == 4 || q == 8) ? (-avf * pr - k * (ss[q])) : (-k * (ss[q]));
  } else {
    s[0] = -avf * pr; s[1] = R(0.0); s[2] = R(0.0);
    s[3] = R(0.0); s[4] = -avf * pr; s[5] = R(0.0);
    s[6] = R(0.0); s[7] = R(0.0); s[8] = -avf * pr;
  }
}

// the sum of the materials' stresses (m5_facercnstr.c:150-173)
template <typename R, bool FAST = false>
__device__ void add_stress_v2(R *st, R avf, R pr, const R *G, R Gmag, const GPow<R> &gp,
                              R mu) {
  R s[9];
  stress_v2<R, FAST>(s, avf, pr, G, Gmag, gp, mu);
#pragma unroll
  for (int q = 0; q < 9; q++) st[q] += s[q];
}

// G = A A^T and its determinant (m5_facevalues.c:123-136)
__device__ void gmat(const double *A, double *G, double &Gmag) {
  G[0] = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  G[1] = A[0] * A[3] + A[1] * A[4] + A[2] * A[5];
  G[2] = A[0] * A[6] + A[1] * A[7] + A[2] * A[8];
  G[3] = A[3] * A[0] + A[4] * A[1] + A[5] * A[2];
  G[4] = A[3] * A[3] + A[4] * A[4] + A[5] * A[5];
  G[5] = A[3] * A[6] + A[4] * A[7] + A[5] * A[8];
  G[6] = A[6] * A[0] + A[7] * A[1] + A[8] * A[2];
  G[7] = A[6] * A[3] + A[7] * A[4] + A[8] * A[5];
  G[8] = A[6] * A[6] + A[7] * A[7] + A[8] * A[8];
  Gmag = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) + G[2] * (G[3] * G[7] - G[4] * G[6]);
}

template <int NM, typename R = double>
struct Rec5 {
  R r, e, u, v, w, phi0;
  R ra[NM], avf[NM], are[NM], c2Y[NM], ph[NM];
  R A[9], st[9];
  R vm[3], vE;        // vec momentum and energy (rewritten for a ghost)
  R ut, vt, wt;       // u_temp, v_temp, w_temp
};

// s = left ? a : b, field by field (selects in registers)
template <int NM, typename R>
__device__ __forceinline__ void sel_rec(bool left, const Rec5<NM, R> &a, const Rec5<NM, R> &b, Rec5<NM, R> &s) {
  s.r = left ? a.r : b.r; s.e = left ? a.e : b.e; s.u = left ? a.u : b.u; s.v = left ? a.v : b.v;
  s.w = left ? a.w : b.w; s.phi0 = left ? a.phi0 : b.phi0;
#pragma unroll
  for (int k = 0; k < NM; k++) {
    s.ra[k] = left ? a.ra[k] : b.ra[k]; s.avf[k] = left ? a.avf[k] : b.avf[k]; s.are[k] = left ? a.are[k] : b.are[k];
    s.c2Y[k] = left ? a.c2Y[k] : b.c2Y[k]; s.ph[k] = left ? a.ph[k] : b.ph[k];
  }
#pragma unroll
  for (int k = 0; k < 9; k++) { s.A[k] = left ? a.A[k] : b.A[k]; s.st[k] = left ? a.st[k] : b.st[k]; }
#pragma unroll
  for (int k = 0; k < 3; k++) s.vm[k] = left ? a.vm[k] : b.vm[k];
  s.vE = left ? a.vE : b.vE; s.ut = left ? a.ut : b.ut; s.vt = left ? a.vt : b.vt; s.wt = left ? a.wt : b.wt;
}

// the record in the face arithmetic's precision (FP64 records are built first:
// the pressures carry p_inf of up to 2e10 Pa and are formed in FP64)
template <int NM, typename R>
__device__ void narrow(const Rec5<NM, double> &a, Rec5<NM, R> &b) {
  b.r = (R)a.r; b.e = (R)a.e; b.u = (R)a.u; b.v = (R)a.v; b.w = (R)a.w; b.phi0 = (R)a.phi0;
#pragma unroll
  for (int k = 0; k < NM; k++) {
    b.ra[k] = (R)a.ra[k]; b.avf[k] = (R)a.avf[k]; b.are[k] = (R)a.are[k]; b.c2Y[k] = (R)a.c2Y[k]; b.ph[k] = (R)a.ph[k];
  }
#pragma unroll
  for (int k = 0; k < 9; k++) { b.A[k] = (R)a.A[k]; b.st[k] = (R)a.st[k]; }
#pragma unroll
  for (int k = 0; k < 3; k++) b.vm[k] = (R)a.vm[k];
  b.vE = (R)a.vE; b.ut = (R)a.ut; b.vt = (R)a.vt; b.wt = (R)a.wt;
}

// Record of a state vector x, n_solids = 0.  FACE = false: m5_facevalues.c:3-188,
// x a cell state.  FACE = true: m5_facercnstr.c:3-190, x a face state whose phi
// comes from the cell's raw solid fraction cavf0, and whose pressure and c2Y
// divide by the re-summed fraction r.avf[i] (the two forms differ only for the
// last material after a renormalisation).
template <int NM, bool FACE = false, bool FAST = false>
__device__ void record(const M5P &P, const double *x, Rec5<NM> &r, double cavf0 = 0.0) {
  using L = L5<NM>;
  const double a = P.amin / 100.0;
  double avf_temp[NM], avf_c0 = 0.0;
  r.r = 0.0;
#pragma unroll
  for (int i = 0; i < NM; ++i) { r.r += x[i]; r.ra[i] = x[i]; }
  const double ir = FAST ? frcp(r.r) : 0.0;
  r.u = FAST ? x[L::I1 + 0] * ir : x[L::I1 + 0] / r.r;
  r.v = FAST ? x[L::I1 + 1] * ir : x[L::I1 + 1] / r.r;
  r.w = FAST ? x[L::I1 + 2] * ir : x[L::I1 + 2] / r.r;
  r.vm[0] = x[L::I1 + 0]; r.vm[1] = x[L::I1 + 1]; r.vm[2] = x[L::I1 + 2];
  r.e = FAST ? x[L::I2] * ir : x[L::I2] / r.r;
  r.vE = x[L::I2];
  double avfsum = 0.0;
#pragma unroll
  for (int i = 0; i < L::N3; ++i) {
    double t = x[L::I3 + i], c = (FACE && i == 0) ? cavf0 : x[L::I3 + i];
    if (t <= a) t = a;
    else if (t >= (1.0 - a)) t = 1.0 - a;
    if (c >= (1.0 - a)) c = 1.0 - a;
    else if (c <= a) c = a;
    if (i == 0) avf_c0 = c;
    avf_temp[i] = t;
    avfsum += t;
  }
  if (avfsum > (1.0 - a)) {
    const double avf_dif = avfsum - (1.0 - a);
#pragma unroll
    for (int j = 0; j < L::N3; ++j) {
      const double avf_per = avf_temp[j] / avfsum;
      avf_temp[j] = avf_temp[j] - avf_per * avf_dif;
    }
    avfsum = 1.0 - a;
  }
  avf_temp[L::N3] = 1.0 - avfsum;
  avfsum = 0.0;
#pragma unroll
  for (int i = 0; i < L::N3; ++i) { r.avf[i] = avf_temp[i]; avfsum += avf_temp[i]; }
  r.avf[L::N3] = 1.0 - avfsum;
  r.phi0 = avf_c0 - P.avf_lim;
#pragma unroll
  for (int i = 0; i < NM; ++i) r.are[i] = x[L::I4 + i];
  double A[9], G[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) { A[k] = x[L::I5 + k]; r.A[k] = A[k]; r.st[k] = 0.0; }
  G[0] = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  G[1] = A[0] * A[3] + A[1] * A[4] + A[2] * A[5];
  G[2] = A[0] * A[6] + A[1] * A[7] + A[2] * A[8];
  G[3] = A[3] * A[0] + A[4] * A[1] + A[5] * A[2];
  G[4] = A[3] * A[3] + A[4] * A[4] + A[5] * A[5];
  G[5] = A[3] * A[6] + A[4] * A[7] + A[5] * A[8];
  G[6] = A[6] * A[0] + A[7] * A[1] + A[8] * A[2];
  G[7] = A[6] * A[3] + A[7] * A[4] + A[8] * A[5];
  G[8] = A[6] * A[6] + A[7] * A[7] + A[8] * A[8];
  const double Gmag = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) +
                      G[2] * (G[3] * G[7] - G[4] * G[6]);
  const GPow<> gp = has_solid(P) ? gpow_t<FAST>(P, Gmag) : GPow<>{0.0, 0.0};
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    double are_elast = (P.mush[i] != 0.0) ? elastic_energy_v2<double, FAST>(G, Gmag, gp, P.mush[i]) : 0.0;
    const double g = P.gamma[i], pi = P.pinf[i];
    const double den = FACE ? r.avf[i] : avf_temp[i];
    // mixture mode: phase i holds alpha_i of the elastic energy and of the shear stress
    const double mu_i = P.mix ? r.avf[i] * P.mush[i] : P.mush[i];
    if (P.mix) are_elast *= den;
    const double press = dv<FAST>((g - 1.0) * (r.are[i] - are_elast), den) - g * pi;
    r.c2Y[i] = (FAST ? den * ir : den / r.r) * g * (press + pi);
    if (P.mix) r.c2Y[i] += FAST ? (4.0 / 3.0) * mu_i * ir : (4.0 / 3.0) * mu_i / r.r;   // longitudinal speed of the mixture
    r.ph[i] = press;
    add_stress_v2<double, FAST>(r.st, r.avf[i], r.ph[i], G, Gmag, gp, mu_i);
  }
}

// rotate_tensor (rotate_vectors.c:128-188): t = V st V^T, V rows a, b, c
template <typename R>
__device__ void rot_tensor(R *t, const R *st, const R *a, const R *b, const R *c) {
  R V[9] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
  R M1[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R s = R(0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) s += V[i * 3 + k] * st[k * 3 + j];
      M1[i * 3 + j] = s;
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R s = R(0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) s += M1[i * 3 + k] * V[j * 3 + k];   // vecT[k][j] = V[j][k]
      t[i * 3 + j] = s;
    }
}

template <int NM, typename R = double>
struct Star5 {
  R Us[NM + 4];
  R unc[3];
  R avf_S[NM], are_S[NM], AS[9];
};

// m5_star_region.c:3-157 for side s (flag 1: left, 2: right)
template <int NM, typename R, bool FAST = false>
__device__ void star_region(const M5P &P, const R *F, const Rec5<NM, R> &sL, const Rec5<NM, R> &sR,
                            const Rec5<NM, R> &s, int flag_ifs, R u_S, R SL, R SR, R uL,
                            R vL, R wL, R uR, R vR, R wR, R s11L, R s12L,
                            R s13L, R s11R, R s12R, R s13R, Star5<NM, R> &W) {
  const R *n = F, *m = F + 3, *l = F + 6;
  R SK, uK, vK, wK, s11K, s12K, s13K;
  if (flag_ifs == 1) { SK = SL; uK = uL; vK = vL; wK = wL; s11K = s11L; s12K = s12L; s13K = s13L; }
  else { SK = SR; uK = uR; vK = vR; wK = wR; s11K = s11R; s12K = s12R; s13K = s13R; }
  // FAST: the shared denominators u* - S_K, the stress jump's and (u_K - S_K) rho_K
  const R iUS = FAST ? frcp(u_S - SK) : R(0.0);
  const R Dst = (uR - SR) * sR.r - (uL - SL) * sL.r;
  const R iDst = FAST ? frcp(Dst) : R(0.0);
  const R iUK = FAST ? frcp((uK - SK) * s.r) : R(0.0);
  R ar_S[NM], r_S = R(0.0);
#pragma unroll
  for (int v = 0; v < NM; ++v) {
    ar_S[v] = FAST ? s.ra[v] * ((uK - SK) * iUS) : s.ra[v] * (SK - uK) / (SK - u_S);
    r_S += ar_S[v];
  }
  const R n11 = (uR - SR) * sR.r * (s11L) - (uL - SL) * sL.r * (s11R) + (uL - SL) * sL.r * (uR - SR) * sR.r * (uR - uL);
  const R s11_S = FAST ? n11 * iDst : n11 / Dst;
  R s12_S, s13_S;
  if ((sL.phi0 > R(0.0)) && (sR.phi0 > R(0.0))) {
    const R n12 = (uR - SR) * sR.r * (s12L) - (uL - SL) * sL.r * (s12R) + (uL - SL) * sL.r * (uR - SR) * sR.r * (vR - vL);
    const R n13 = (uR - SR) * sR.r * (s13L) - (uL - SL) * sL.r * (s13R) + (uL - SL) * sL.r * (uR - SR) * sR.r * (wR - wL);
    s12_S = FAST ? n12 * iDst : n12 / Dst;
    s13_S = FAST ? n13 * iDst : n13 / Dst;
  } else {
    s12_S = R(0.0);
    s13_S = R(0.0);
  }
  const R v_S = vK + (FAST ? (s12_S - s12K) * iUK : (s12_S - s12K) / ((uK - SK) * s.r));
  const R w_S = wK + (FAST ? (s13_S - s13K) * iUK : (s13_S - s13K) / ((uK - SK) * s.r));
  R vel[3];
  vel[0] = u_S * n[0] + v_S * m[0] + w_S * l[0];   // rotate_vector_ltg_v2
  vel[1] = u_S * n[1] + v_S * m[1] + w_S * l[1];
  vel[2] = u_S * n[2] + v_S * m[2] + w_S * l[2];
  const R re = s.r * s.e;
  // consistent: - s11K uK over r* (u* - S_K); published (m5_star_region.c:73): + s11K uK over r* (u* + S_K)
  const R Em_S = P.fix ? dv<FAST>(re * (uK - SK) - s11K * uK - s12K * vK - s13K * wK + s11_S * u_S + s12_S * v_S + s13_S * w_S,
                                  r_S * (u_S - SK))
                            : dv<FAST>(re * (uK - SK) + s11K * uK - s12K * vK - s13K * wK + s11_S * u_S + s12_S * v_S + s13_S * w_S,
                                  r_S * (u_S + SK));
  R A[9], t[9], AS[9];
  rot_tensor(t, s.A, n, m, l);
#pragma unroll
  for (int k = 0; k < 9; k++) A[k] = t[k];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const R na = P.fix ? (A[j] * (uK - SK) + A[3 + j] * (vK - v_S) + A[6 + j] * (wK - w_S))
                       : (A[j] * (uK - SK) - A[3 + j] * (vK - v_S) + A[6 + j] * (wK - w_S));
    AS[j] = FAST ? na * iUS : na / (u_S - SK);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) { AS[3 + j] = A[3 + j]; AS[6 + j] = A[6 + j]; }
  {
    const R c0[3] = {n[0], m[0], l[0]}, c1[3] = {n[1], m[1], l[1]}, c2[3] = {n[2], m[2], l[2]};
    rot_tensor(t, AS, c0, c1, c2);
  }
#pragma unroll
  for (int k = 0; k < 9; k++) AS[k] = t[k];
  R G[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G[i * 3 + j] = AS[i * 3 + 0] * AS[j * 3 + 0] + AS[i * 3 + 1] * AS[j * 3 + 1] + AS[i * 3 + 2] * AS[j * 3 + 2];
  R e_S[NM];
#pragma unroll
  for (int v = 0; v < NM; ++v) {
    const R g = (R)P.gamma[v], pi = (R)P.pinf[v];
    const R rm = dv<FAST>(s.ra[v], s.avf[v]);
    const R rm_S = FAST ? rm * ((uK - SK) * iUS) : (s.ra[v] / s.avf[v]) * (uK - SK) / (u_S - SK);
    const R p_k = P.fix ? s.ph[v] : -s.ph[v];   // published: -p_hydro
    const R p_S = (p_k + pi) * dv<FAST>((g - R(1.0)) * rm + (g + R(1.0)) * rm_S, (g - R(1.0)) * rm_S + (g + R(1.0)) * rm) - pi;
    R are_el = elastic_energy_v1<R, FAST>(P, G, (R)P.mush[v]);
    if (P.mix) are_el *= s.avf[v];   // mixture mode: alpha_v of the elastic energy
    const R e_el = dv<FAST>(are_el, s.avf[v] * rm_S);
    e_S[v] = dv<FAST>(p_S + g * pi, rm_S * (P.fix ? (g - R(1.0)) : (g + R(1.0)))) + e_el;   // published: gamma + 1
  }
#pragma unroll
  for (int v = 0; v < NM; ++v) W.Us[v] = ar_S[v];
#pragma unroll
  for (int v = 0; v < 3; ++v) W.Us[NM + v] = r_S * vel[v];
  W.Us[NM + 3] = r_S * Em_S;
  W.unc[0] = vel[0]; W.unc[1] = vel[1]; W.unc[2] = vel[2];
#pragma unroll
  for (int v = 0; v < NM - 1; ++v) W.avf_S[v] = s.avf[v];
#pragma unroll
  for (int v = 0; v < NM; ++v) W.are_S[v] = ar_S[v] * e_S[v];
#pragma unroll
  for (int k = 0; k < 9; k++) W.AS[k] = AS[k];
}

// One GFM sub-face (m5_hllc_gfm.c:3-100 with m5_wave_speeds_gfm.c and
// m5_flux_gfm.c): out[iv] = F.n + F_hllc as m5_residual.c:166-170 forms it
// before the area, and the non-conservative face velocity.  sL and sR are
// rewritten with the ghost.  Returns false where the reference prints NaN and exits.
template <int NM, typename R, bool FAST = false>
__device__ bool gfm_subface(const M5P &P, const R *F, Rec5<NM, R> &sL, Rec5<NM, R> &sR, R *out, R *unc) {
  using L = L5<NM>;
  const R *n = F, *m = F + 3, *l = F + 6;
  // face-frame velocities (rotate_vector_gtl_v2) and the ghost: v_R = v_L, w_R = w_L
  const R uL = sL.u * n[0] + sL.v * n[1] + sL.w * n[2];
  const R vL = sL.u * m[0] + sL.v * m[1] + sL.w * m[2];
  const R wL = sL.u * l[0] + sL.v * l[1] + sL.w * l[2];
  const R uR = sR.u * n[0] + sR.v * n[1] + sR.w * n[2];
  const R vR = vL, wR = wL;
  sL.ut = sL.u; sL.vt = sL.v; sL.wt = sL.w;
  sR.ut = uR * n[0] + vR * m[0] + wR * l[0];
  sR.vt = uR * n[1] + vR * m[1] + wR * l[1];
  sR.wt = uR * n[2] + vR * m[2] + wR * l[2];
  sR.vm[0] = sR.ut * sR.r; sR.vm[1] = sR.vt * sR.r; sR.vm[2] = sR.wt * sR.r;
  sR.vE = sR.r * sR.e;
  // m5_wave_speeds_gfm.c:16-58
  const R ruL = sL.r * (uL), ruR = sR.r * (uR);
  R c2YL = R(0.0), c2YR = R(0.0);
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) { c2YL += sL.c2Y[iv]; c2YR += sR.c2Y[iv]; }
  const R cL = sqrt(c2YL), cR = sqrt(c2YR);
  const R SR = maxr((uL + cL), (uR + cR));
  const R SL = minr((uL - cL), (uR - cR));
  R t[9];
  rot_tensor(t, sL.st, n, m, l);
  const R s11L = t[0], s12L = t[1], s13L = t[2];
  rot_tensor(t, sR.st, n, m, l);
  const R s11R = t[0], s12R = t[1], s13R = t[2];
  const R u_S = dv<FAST>((sL.r * sq((uL)) - (s11L)) - (sR.r * sq((uR)) - (s11R)) - ((SL)*ruL) + ((SR)*ruR),
                         ruL - ruR - ((SL)*sL.r) + ((SR)*sR.r));
  if (isnan(u_S) || isnan(SR) || isnan(SL)) return false;
  int flag_ifs;
  if (R(0.0) <= SL) flag_ifs = 0;
  else if ((SL < R(0.0)) && (R(0.0) <= u_S)) flag_ifs = 1;
  else if ((u_S < R(0.0)) && (R(0.0) < SR)) flag_ifs = 2;
  else flag_ifs = 3;   // SR <= 0, the only case left once no speed is NaN
  // the side's record by value, selected field by field: a reference chosen at run
  // time would place both records in scratch memory
  Rec5<NM, R> s;
  sel_rec<NM, R>(flag_ifs <= 1, sL, sR, s);
  Star5<NM, R> W;
  R SK = R(0.0);
  if (flag_ifs == 1 || flag_ifs == 2) {
    SK = (flag_ifs == 1) ? SL : SR;
    star_region<NM, R, FAST>(P, F, sL, sR, s, flag_ifs, u_S, SL, SR, uL, vL, wL, uR, vR, wR, s11L, s12L, s13L, s11R, s12R, s13R, W);
#pragma unroll
    for (int iv = 0; iv < L::NCONS; ++iv)
      if (isnan(W.Us[iv])) return false;
  } else {
    // m5_flux_gfm.c:43-56: the non-conservative face state of the side itself
    W.unc[0] = s.ut; W.unc[1] = s.vt; W.unc[2] = s.wt;
#pragma unroll
    for (int v = 0; v < NM - 1; ++v) W.avf_S[v] = s.avf[v];
#pragma unroll
    for (int v = 0; v < NM; ++v) W.are_S[v] = s.are[v];
#pragma unroll
    for (int k = 0; k < 9; k++) W.AS[k] = s.A[k];
  }
  const bool star = (flag_ifs == 1 || flag_ifs == 2);
  bool ok = true;
  auto emit = [&](int iv, R f0, R f1, R f2, R vec_iv) {
    const R h = (star && iv < L::NCONS) ? SK * (W.Us[iv] - vec_iv) : R(0.0);
    if (isnan(h)) ok = false;
    out[iv] = ((f0 - R(0.0)) * n[0] + (f1 - R(0.0)) * n[1] + (f2 - R(0.0)) * n[2]) + h;
  };
  // m5_flux_gfm.c:10-39: physical flux of side s with its u_temp velocity
#pragma unroll
  for (int v = 0; v < NM; ++v) emit(L::I0 + v, s.ut * s.ra[v], s.vt * s.ra[v], s.wt * s.ra[v], s.ra[v]);
  emit(L::I1 + 0, s.ut * s.ut * s.r - s.st[0], s.vt * s.ut * s.r - s.st[1], s.wt * s.ut * s.r - s.st[2], s.vm[0]);
  emit(L::I1 + 1, s.ut * s.vt * s.r - s.st[3], s.vt * s.vt * s.r - s.st[4], s.wt * s.vt * s.r - s.st[5], s.vm[1]);
  emit(L::I1 + 2, s.ut * s.wt * s.r - s.st[6], s.vt * s.wt * s.r - s.st[7], s.wt * s.wt * s.r - s.st[8], s.vm[2]);
  emit(L::I2, s.r * s.e * s.ut - s.st[0] * s.ut - s.st[1] * s.vt - s.st[2] * s.wt,
       s.r * s.e * s.vt - s.st[3] * s.ut - s.st[4] * s.vt - s.st[5] * s.wt,
       s.r * s.e * s.wt - s.st[6] * s.ut - s.st[7] * s.vt - s.st[8] * s.wt, s.vE);
#pragma unroll
  for (int v = 0; v < NM - 1; ++v) emit(L::I3 + v, W.unc[0] * W.avf_S[v], W.unc[1] * W.avf_S[v], W.unc[2] * W.avf_S[v], R(0.0));
#pragma unroll
  for (int v = 0; v < NM; ++v) emit(L::I4 + v, W.unc[0] * W.are_S[v], W.unc[1] * W.are_S[v], W.unc[2] * W.are_S[v], R(0.0));
#pragma unroll
  for (int k = 0; k < 9; k++) emit(L::I5 + k, W.unc[0] * W.AS[k], W.unc[1] * W.AS[k], W.unc[2] * W.AS[k], R(0.0));
  unc[0] = W.unc[0]; unc[1] = W.unc[1]; unc[2] = W.unc[2];
  return ok;
}

// one interface sub-face of a list: gfm_subface times the sub-face area
template <int NM>
__global__ __launch_bounds__(256) void k_m5_gfm_faces(int nf, M5P P, const double *__restrict__ UL,
                                                      const double *__restrict__ UR, const double *__restrict__ FR,
                                                      const double *__restrict__ area, double *__restrict__ flux,
                                                      double *__restrict__ unc, int *__restrict__ nanflag) {
  using L = L5<NM>;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nf) return;
  double x[L::NEQ], F[9];
  Rec5<NM> sL, sR;
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) x[v] = UL[(size_t)v * nf + q];
  record<NM>(P, x, sL);
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) x[v] = UR[(size_t)v * nf + q];
  record<NM>(P, x, sR);
#pragma unroll
  for (int k = 0; k < 9; k++) F[k] = FR[(size_t)k * nf + q];
  double out[L::NEQ], un[3];
  if (!gfm_subface<NM, double>(P, F, sL, sR, out, un)) { atomicOr(nanflag, 1); return; }
  const double A = area[q];
#pragma unroll
  for (int iv = 0; iv < L::NEQ; iv++) {
    double fl = 0.0;
    fl += (A) * out[iv];
    flux[(size_t)iv * nf + q] = fl;
  }
  unc[(size_t)0 * nf + q] = un[0];
  unc[(size_t)1 * nf + q] = un[1];
  unc[(size_t)2 * nf + q] = un[2];
}

// ---- m5_RK ----------------------------------------------------------------

template <int NM>
__device__ void fallback_and_freeze(const M5P &P, int n, int i, const double *__restrict__ SN, double *__restrict__ S) {
  using L = L5<NM>;
  // m5_RK.c:90-122
#pragma unroll
  for (int v = 0; v < NM; v++) {
    const size_t e = (size_t)(L::I0 + v) * n + i;
    if (S[e] < 0.0) S[e] = SN[e];
  }
  auto keep = [&](int e) { S[(size_t)e * n + i] = SN[(size_t)e * n + i]; };
  if (P.zerou[0] == 1) { keep(L::I1 + 0); keep(L::I5 + 1); keep(L::I5 + 2); keep(L::I5 + 3); keep(L::I5 + 6); }
  if (P.zerou[1] == 1) { keep(L::I1 + 1); keep(L::I5 + 1); keep(L::I5 + 3); keep(L::I5 + 5); keep(L::I5 + 7); }
  if (P.zerou[2] == 1) { keep(L::I1 + 2); keep(L::I5 + 2); keep(L::I5 + 5); keep(L::I5 + 6); keep(L::I5 + 7); }
}

// CFL step of model 5 (new, the model-4 k_cfl's form): per-block minimum of
// h / (max|u_i| + c) with c^2 = sum_k c2Y_k + (4/3) sum_k alpha_k mu_k / rho, the
// longitudinal speed of the mixture (c2Y_k carries the shear part already in
// the mixture mode), reduced with wavefront shuffles
template <int NM>
__global__ __launch_bounds__(256) void k_m5_cfl(int n, M5P P, const Geo *__restrict__ Gp,
                                                const uint32_t *__restrict__ info, const double *__restrict__ S,
                                                double *__restrict__ part) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = INFINITY;
  if (i < n) {
    double x[L::NEQ];
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) x[v] = S[(size_t)v * n + i];
    Rec5<NM> r;
    record<NM>(P, x, r);
    double c2 = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) c2 += r.c2Y[k];
    if (!P.mix) {
      double mu = 0.0;
#pragma unroll
      for (int k = 0; k < NM; ++k) mu += r.avf[k] * P.mush[k];
      c2 += (4.0 / 3.0) * mu / r.r;
    }
    const double lam = maxr(maxr(fabs(r.u), fabs(r.v)), fabs(r.w)) + sqrt(c2);
    d = Gp->h[fvk::cell_level(info[i])] / lam;
  }
  for (int off = 32; off > 0; off >>= 1) d = fmin(d, __shfl_xor(d, off, 64));
  __shared__ double wmin[4];
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmin(fmin(wmin[0], wmin[1]), fmin(wmin[2], wmin[3]));
}

// minimum of m partial minima, times scale, into *out (one block of 1024 lanes)
__global__ __launch_bounds__(1024) void k_m5_min_final(int m, const double *__restrict__ part,
                                                       double *__restrict__ out, double scale) {
  double d = INFINITY;
  for (int k = threadIdx.x; k < m; k += blockDim.x) d = fmin(d, part[k]);
  for (int off = 32; off > 0; off >>= 1) d = fmin(d, __shfl_xor(d, off, 64));
  __shared__ double wmin[16];
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = wmin[0];
    for (int w = 1; w < 16; w++) r = fmin(r, wmin[w]);
    *out = scale * r;
  }
}

template <int NM>
__global__ __launch_bounds__(256) void k_m5_rk(int n, M5P P, int irk, const Geo *__restrict__ Gp,
                                               const uint32_t *__restrict__ info, double *__restrict__ SN,
                                               const double *__restrict__ Sin, double *__restrict__ Sout,
                                               const double *__restrict__ RHS, uint8_t *__restrict__ mark) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double Vol = Gp->vol[fvk::cell_level(info[i])];
  auto at = [&](int e) { return (size_t)e * n + i; };
  uint8_t flag_rk = 0;
  switch (irk) {
    case 0:
#pragma unroll
      for (int e = 0; e < L::NEQ; e++) { SN[at(e)] = Sin[at(e)]; Sout[at(e)] = Sin[at(e)]; }
      break;
    case 1: {
#pragma unroll
      for (int v = 0; v < L::N3; ++v) {
        if (P.mush[v] != 0.0) {
          const int e = L::I3 + v;
          const double phi_old = Sin[at(e)] - P.avf_lim;
          const double buf = (1.0 / Vol) * RHS[at(e)];
          const double as_new = SN[at(e)] + P.dtu * buf;
          const double phi_new = as_new - P.avf_lim;
          if (phi_old * phi_new < 0.0) flag_rk = 1;
        }
      }
#pragma unroll
      for (int e = 0; e < L::NEQ; e++) {
        if (flag_rk && e < L::I2) { Sout[at(e)] = Sin[at(e)]; continue; }   // set by m5_gfm
        const double buf = (1.0 / Vol) * RHS[at(e)];
        Sout[at(e)] = SN[at(e)] + P.dtu * buf;
      }
      break;
    }
    case 2:
#pragma unroll
      for (int e = 0; e < L::NEQ; e++)
        Sout[at(e)] = (3.0 / 4.0) * SN[at(e)] + (1.0 / 4.0) * Sin[at(e)] + (1.0 / 4.0) * P.dtu * RHS[at(e)] / Vol;
      break;
    case 3:
#pragma unroll
      for (int e = 0; e < L::NEQ; e++)
        Sout[at(e)] = (1.0 / 3.0) * SN[at(e)] + (2.0 / 3.0) * Sin[at(e)] + (2.0 / 3.0) * P.dtu * RHS[at(e)] / Vol;
      break;
    case 4:
#pragma unroll
      for (int e = 0; e < L::NEQ; e++)
        Sout[at(e)] = (1.0 / 2.0) * SN[at(e)] + (1.0 / 2.0) * Sin[at(e)] + (1.0 / 2.0) * P.dtu * RHS[at(e)] / Vol;
      break;
    default:
#pragma unroll
      for (int e = 0; e < L::NEQ; e++) Sout[at(e)] = Sin[at(e)];
      break;
  }
  mark[i] = flag_rk;
  if (!flag_rk) fallback_and_freeze<NM>(P, n, i, SN, Sout);
}

// relaxation trigger of m5_relaxation_solid.c:18-32 (phi_new[0] only exists for a solid)
template <int NM>
__global__ __launch_bounds__(256) void k_m5_rel_mark(int n, M5P P, const double *__restrict__ Sb,
                                                     const double *__restrict__ S, uint8_t *__restrict__ mark) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double phi_old = Sb[(size_t)L::I3 * n + i] - P.avf_lim;
  const double phi_new = S[(size_t)L::I3 * n + i] - P.avf_lim;
  mark[i] = (P.mush[0] != 0.0 && phi_old * phi_new < 0.0) ? 1 : 0;
}

// interface normal, tangent frame and new global velocity shared by m5_gfm.c:76-196
// and m5_gfm_rel.c:60-219 (accumulate selects the latter's summed normal)
template <int NM>
__device__ int gfm_velocity(const M5P &P, const Geo *__restrict__ Gp, int n, int i, const int *__restrict__ nbr,
                            const uint32_t *__restrict__ info, const int *__restrict__ fine4,
                            const uint8_t *__restrict__ gf, const double *__restrict__ Snew,
                            const double *__restrict__ Sold, int accumulate, double ut, double vt, double wt,
                            double ng[3], int *err) {
  using L = L5<NM>;
  const uint32_t inf = info[i];
  const int lvl = fvk::cell_level(inf);
  double nxi = 0.0, nyi = 0.0, nzi = 0.0, total_area = 0.0;
  double us = 0.0, vs = 0.0, ws = 0.0, ar = 0.0;   // the last interface face's neighbour velocity and area
  double uss[6], vss[6], wss[6], ars[6];
  int fi = 0;
  double xo[L::NEQ], xn[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) xo[v] = Snew[(size_t)v * n + i];
  Rec5<NM> rL;
  record<NM>(P, xo, rL);
  for (int f = 0; f < 6; f++) {
    if (!gf[6 * i + f]) continue;
    const int kind = fvk::face_kind(inf, f);
    if (kind == FV_FK_BOUNDARY) { *err = 1; return 0; }
    const int nb = nbr[(size_t)f * n + i];
    const int j = (kind == FV_FK_FINE) ? fine4[4 * nb] : nb;   // neigtr[f][0]
    const double A = (kind == FV_FK_FINE) ? Gp->area[lvl + 1] : Gp->area[lvl];
    const double *Sj = (j < i) ? Snew : Sold;
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) xn[v] = Sj[(size_t)v * n + j];
    Rec5<NM> rR;
    record<NM>(P, xn, rR);
    const double phi = (rL.phi0 + rR.phi0) / 2.0;
    const double *nf = Gp->frame[f];
    if (accumulate) {
      nxi += phi * A * nf[0];
      nyi += phi * A * nf[1];
      nzi += phi * A * nf[2];
    } else {
      nxi = phi * A * nf[0];
      nyi = phi * A * nf[1];
      nzi = phi * A * nf[2];
    }
    uss[fi] = rR.u; vss[fi] = rR.v; wss[fi] = rR.w; ars[fi] = A;
    total_area += A;
    fi++;
  }
  if (fi == 0) return 0;
  us = uss[fi - 1]; vs = vss[fi - 1]; ws = wss[fi - 1]; ar = ars[fi - 1];
  const double norm = sqrt(sq(nxi) + sq(nyi) + sq(nzi));
  nxi = -nxi / norm; nyi = -nyi / norm; nzi = -nzi / norm;
  double nx1, ny1, nz1;
  if (fabs(nzi) < 0.99) { nx1 = nyi; ny1 = -nxi; nz1 = 0.0; }
  else { nx1 = nzi; ny1 = 0.0; nz1 = -nxi; }
  const double nx2 = nyi * nz1 - nzi * ny1;
  const double ny2 = nzi * nx1 - nxi * nz1;
  const double nz2 = nxi * ny1 - nyi * nx1;
  const double u_lc = ut * nxi + vt * nyi + wt * nzi;     // rotate_vector (rotate_vectors.c:3-13)
  // v_new_lc and w_new_lc are overwritten per interface face (m5_gfm.c:160-171): the last one stays
  const double v_lc = us * nx1 + vs * ny1 + ws * nz1;
  const double w_lc = us * nx2 + vs * ny2 + ws * nz2;
  const double v_new = v_lc * (ar / total_area);
  const double w_new = w_lc * (ar / total_area);
  ng[0] = u_lc * nxi + v_new * nx1 + w_new * nx2;
  ng[1] = u_lc * nyi + v_new * ny1 + w_new * ny2;
  ng[2] = u_lc * nzi + v_new * nz1 + w_new * nz2;
  return fi;
}

// m5_gfm.c:3-254 over the marked leaves in leaf order (one lane: the list is short
// and each leaf may read the new state of an earlier marked neighbour)
template <int NM>
__global__ void k_m5_gfm_seq(int n, const int *__restrict__ nlist_p, const int *__restrict__ list, M5P P,
                             const Geo *__restrict__ Gp,
                             const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                             const int *__restrict__ fine4, const uint8_t *__restrict__ gf,
                             const double *__restrict__ SN, const double *__restrict__ Sin, double *__restrict__ S,
                             const double *__restrict__ RHS, int *__restrict__ err) {
  using L = L5<NM>;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int nlist = *nlist_p;
  for (int t = 0; t < nlist; t++) {
    const int i = list[t];
    auto at = [&](int e) { return (size_t)e * n + i; };
    const double dV = P.dtu / Gp->vol[fvk::cell_level(info[i])];
    double rho_new = 0.0;
#pragma unroll
    for (int v = 0; v < NM; ++v) rho_new += S[at(L::I0 + v)] + (dV)*RHS[at(L::I0 + v)];
    double ut = S[at(L::I1 + 0)] + (dV)*RHS[at(L::I1 + 0)];
    double vt = S[at(L::I1 + 1)] + (dV)*RHS[at(L::I1 + 1)];
    double wt = S[at(L::I1 + 2)] + (dV)*RHS[at(L::I1 + 2)];
    ut = ut / rho_new; vt = vt / rho_new; wt = wt / rho_new;
    double ng[3];
    int e1 = 0;
    const int fi = gfm_velocity<NM>(P, Gp, n, i, nbr, info, fine4, gf, S, Sin, 0, ut, vt, wt, ng, &e1);
    if (e1) { *err = 1; return; }
    if (fi != 0) {
#pragma unroll
      for (int v = 0; v < NM; ++v) S[at(L::I0 + v)] = S[at(L::I0 + v)] + (dV)*RHS[at(L::I0 + v)];
      S[at(L::I1 + 0)] = rho_new * ng[0];
      S[at(L::I1 + 1)] = rho_new * ng[1];
      S[at(L::I1 + 2)] = rho_new * ng[2];
      double E = S[at(L::I2)];
      E -= rho_new * 0.5 * (sq(ut) + sq(vt) + sq(wt));
      E += rho_new * 0.5 * (sq(ng[0]) + sq(ng[1]) + sq(ng[2]));
      S[at(L::I2)] = E;
    } else {
#pragma unroll
      for (int v = 0; v < NM; ++v) S[at(L::I0 + v)] = S[at(L::I0 + v)] + (dV)*RHS[at(L::I0 + v)];
#pragma unroll
      for (int k = 0; k < 3; ++k) S[at(L::I1 + k)] = S[at(L::I1 + k)] + (dV)*RHS[at(L::I1 + k)];
    }
    fallback_and_freeze<NM>(P, n, i, SN, S);
  }
}

// m5_gfm_rel.c:3-246 over the marked leaves in leaf order
template <int NM>
__global__ void k_m5_rel_seq(int n, const int *__restrict__ nlist_p, const int *__restrict__ list, M5P P,
                             const Geo *__restrict__ Gp,
                             const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                             const int *__restrict__ fine4, const uint8_t *__restrict__ gf,
                             const double *__restrict__ Sb, double *__restrict__ S, int *__restrict__ err) {
  using L = L5<NM>;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int nlist = *nlist_p;
  for (int t = 0; t < nlist; t++) {
    const int i = list[t];
    auto at = [&](int e) { return (size_t)e * n + i; };
    double rho = 0.0;
#pragma unroll
    for (int v = 0; v < NM; ++v) rho += S[at(v)];
    const double ut = S[at(L::I1 + 0)] / rho, vt = S[at(L::I1 + 1)] / rho, wt = S[at(L::I1 + 2)] / rho;
    double ng[3];
    int e1 = 0;
    const int fi = gfm_velocity<NM>(P, Gp, n, i, nbr, info, fine4, gf, S, Sb, 1, ut, vt, wt, ng, &e1);
    if (e1) { *err = 1; return; }
    if (fi != 0) {
      S[at(L::I1 + 0)] = rho * ng[0];
      S[at(L::I1 + 1)] = rho * ng[1];
      S[at(L::I1 + 2)] = rho * ng[2];
      double E = S[at(L::I2)];
      E -= rho * 0.5 * (sq(ut) + sq(vt) + sq(wt));
      E += rho * 0.5 * (sq(ng[0]) + sq(ng[1]) + sq(ng[2]));
      S[at(L::I2)] = E;
    }
  }
}


// ---- the whole stage: MUSCL face states, residual, relaxation ---------------

__constant__ static const int FCOPP5[6] = {2, 3, 0, 1, 5, 4};   // operatetp.c:268-279

// m5_bc.c:3-83 on the right record: 1 symmetry and 4 reflective flip the
// normal velocity of the left record, 5 flips all three components
template <int NM, typename R>
__device__ void apply_bc5(int code, const R *F, const Rec5<NM, R> &sL, Rec5<NM, R> &sR) {
  if (code != 1 && code != 4 && code != 5) return;
  const R *n = F, *m = F + 3, *l = F + 6;
  const R ul = sL.u * n[0] + sL.v * n[1] + sL.w * n[2];
  const R vl = sL.u * m[0] + sL.v * m[1] + sL.w * m[2];
  const R wl = sL.u * l[0] + sL.v * l[1] + sL.w * l[2];
  const R u1 = -ul, v1 = (code == 5) ? -vl : vl, w1 = (code == 5) ? -wl : wl;
  sR.u = u1 * n[0] + v1 * m[0] + w1 * l[0];
  sR.v = u1 * n[1] + v1 * m[1] + w1 * l[1];
  sR.w = u1 * n[2] + v1 * m[2] + w1 * l[2];
}

// m5_flux.c:10-39: the physical flux rows of a record, F.n per equation for
// the mass, momentum (sg sigma: + published, - consistent) and energy rows
template <int NM, typename R>
__device__ void phys_rows(const Rec5<NM, R> &s, const R *n, R sg, R *adv) {
  using L = L5<NM>;
  auto dot = [&](R f0, R f1, R f2) { return (f0 - R(0.0)) * n[0] + (f1 - R(0.0)) * n[1] + (f2 - R(0.0)) * n[2]; };
#pragma unroll
  for (int v = 0; v < NM; ++v) adv[L::I0 + v] = dot(s.u * s.ra[v], s.v * s.ra[v], s.w * s.ra[v]);
  adv[L::I1 + 0] = dot(s.u * s.u * s.r + sg * s.st[0], s.v * s.u * s.r + sg * s.st[1], s.w * s.u * s.r + sg * s.st[2]);
  adv[L::I1 + 1] = dot(s.u * s.v * s.r + sg * s.st[3], s.v * s.v * s.r + sg * s.st[4], s.w * s.v * s.r + sg * s.st[5]);
  adv[L::I1 + 2] = dot(s.u * s.w * s.r + sg * s.st[6], s.v * s.w * s.r + sg * s.st[7], s.w * s.w * s.r + sg * s.st[8]);
  adv[L::I2] = dot(s.r * s.e * s.u - s.st[0] * s.u - s.st[1] * s.v - s.st[2] * s.w,
                   s.r * s.e * s.v - s.st[3] * s.u - s.st[4] * s.v - s.st[5] * s.w,
                   s.r * s.e * s.w - s.st[6] * s.u - s.st[7] * s.v - s.st[8] * s.w);
}

// One HLLC sub-face (m5_hllc.c:3-94 with m5_wave_speeds.c:3-131 and
// m5_flux.c): out[iv] = F.n + F_hllc before the area, and the
// non-conservative face velocity.  bc 4 forces u* = 0; bc 1 (symmetry)
// replaces the flux by the mean of the physical fluxes of the cell record of
// xc and its mirror.  Returns false where the reference prints NaN and exits.
template <int NM, typename R, bool FAST = false>
__device__ bool hllc_subface(const M5P &P, const R *F, int bc, const Rec5<NM, R> &sL, const Rec5<NM, R> &sR,
                             const double *__restrict__ S, int n, int cell, R *out, R *unc) {
  using L = L5<NM>;
  const R *nv = F, *m = F + 3, *l = F + 6;
  const R uL = sL.u * nv[0] + sL.v * nv[1] + sL.w * nv[2];
  const R vL = sL.u * m[0] + sL.v * m[1] + sL.w * m[2];
  const R wL = sL.u * l[0] + sL.v * l[1] + sL.w * l[2];
  const R uR = sR.u * nv[0] + sR.v * nv[1] + sR.w * nv[2];
  const R vR = sR.u * m[0] + sR.v * m[1] + sR.w * m[2];
  const R wR = sR.u * l[0] + sR.v * l[1] + sR.w * l[2];
  const R ruL = sL.r * (uL), ruR = sR.r * (uR);
  R c2YL = R(0.0), c2YR = R(0.0);
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) { c2YL += sL.c2Y[iv]; c2YR += sR.c2Y[iv]; }
  const R cL = sqrt(c2YL), cR = sqrt(c2YR);
  const R SR = maxr((uL + cL), (uR + cR));
  const R SL = minr((uL - cL), (uR - cR));
  R t[9];
  rot_tensor(t, sL.st, nv, m, l);
  const R s11L = t[0], s12L = t[1], s13L = t[2];
  rot_tensor(t, sR.st, nv, m, l);
  const R s11R = t[0], s12R = t[1], s13R = t[2];
  R u_S = dv<FAST>((sL.r * sq((uL)) - (s11L)) - (sR.r * sq((uR)) - (s11R)) - ((SL)*ruL) + ((SR)*ruR),
                   ruL - ruR - ((SL)*sL.r) + ((SR)*sR.r));
  if (isnan(u_S) || isnan(SR) || isnan(SL)) return false;
  if (bc == 4) u_S = R(0.0);
  int flag_ifs;
  if (R(0.0) <= SL) flag_ifs = 0;
  else if ((SL <= R(0.0)) && (R(0.0) <= u_S)) flag_ifs = 1;
  else if ((u_S <= R(0.0)) && (R(0.0) <= SR)) flag_ifs = 2;
  else flag_ifs = 3;   // SR <= 0
  // the side's record by value, selected field by field: a reference chosen at run
  // time would place both records in scratch memory
  Rec5<NM, R> s;
  sel_rec<NM, R>(flag_ifs <= 1, sL, sR, s);
  Star5<NM, R> W;
  R SK = R(0.0);
  const bool star = (flag_ifs == 1 || flag_ifs == 2);
  if (star) {
    SK = (flag_ifs == 1) ? SL : SR;
    star_region<NM, R, FAST>(P, F, sL, sR, s, flag_ifs, u_S, SL, SR, uL, vL, wL, uR, vR, wR, s11L, s12L, s13L, s11R, s12R, s13R, W);
#pragma unroll
    for (int iv = 0; iv < L::NCONS; ++iv)
      if (isnan(W.Us[iv])) return false;
  } else {
    // m5_flux.c:44-54: the side's own velocity, fractions and -A (+A consistent)
    W.unc[0] = s.u; W.unc[1] = s.v; W.unc[2] = s.w;
#pragma unroll
    for (int v = 0; v < NM - 1; ++v) W.avf_S[v] = s.avf[v];
#pragma unroll
    for (int v = 0; v < NM; ++v) W.are_S[v] = s.are[v];
#pragma unroll
    for (int k = 0; k < 9; k++) W.AS[k] = P.fix ? s.A[k] : -s.A[k];
  }
  if (bc == 1) {
    // m5_hllc.c:66-88: F_hllc = 0, the face velocity and flux are the means of
    // the cell record and its mirror (flag 4 of m5_flux: +A, the side's fractions).
    // The region branch before it only matters through its NaN exits.
    double x[L::NEQ];
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) x[v] = S[(size_t)v * n + cell];
    Rec5<NM, R> cl, cr;
    {
      Rec5<NM> c64;
      record<NM, false, FAST>(P, x, c64);
      narrow<NM, R>(c64, cl);
    }
    cr = cl;
    apply_bc5<NM, R>(1, F, cl, cr);
    const R un = R(0.5) * (cl.u + cr.u), vn = R(0.5) * (cl.v + cr.v), wn = R(0.5) * (cl.w + cr.w);
    R a0[L::NEQ], a1[L::NEQ];
    auto rows = [&](const Rec5<NM, R> &s, R *adv, int k) {
      // the three components of row k of flag 4, k = 0, 1, 2
      const R uu = (k == 0) ? un : (k == 1 ? vn : wn);
      const R su = (k == 0) ? s.u : (k == 1 ? s.v : s.w);
#pragma unroll
      for (int v = 0; v < NM; ++v) adv[L::I0 + v] = su * s.ra[v];
      adv[L::I1 + 0] = su * s.u * s.r + (R)P.sg * s.st[0 + k];
      adv[L::I1 + 1] = su * s.v * s.r + (R)P.sg * s.st[3 + k];
      adv[L::I1 + 2] = su * s.w * s.r + (R)P.sg * s.st[6 + k];
      adv[L::I2] = s.r * s.e * su - s.st[3 * k + 0] * s.u - s.st[3 * k + 1] * s.v - s.st[3 * k + 2] * s.w;
#pragma unroll
      for (int v = 0; v < NM - 1; ++v) adv[L::I3 + v] = uu * s.avf[v];
#pragma unroll
      for (int v = 0; v < NM; ++v) adv[L::I4 + v] = uu * s.are[v];
#pragma unroll
      for (int q = 0; q < 9; ++q) adv[L::I5 + q] = uu * s.A[q];
    };
    R acc[3][L::NEQ];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      rows(cl, a0, k);
      rows(cr, a1, k);
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) {
        R sym = R(0.0);
        sym += R(0.5) * a0[v];
        sym += R(0.5) * a1[v];
        acc[k][v] = sym;
      }
    }
#pragma unroll
    for (int v = 0; v < L::NEQ; v++)
      out[v] = ((acc[0][v] - R(0.0)) * nv[0] + (acc[1][v] - R(0.0)) * nv[1] + (acc[2][v] - R(0.0)) * nv[2]) + R(0.0);
    unc[0] = un; unc[1] = vn; unc[2] = wn;
    return true;
  }
  R adv[L::NEQ];
  phys_rows<NM, R>(s, nv, (R)P.sg, adv);
  bool ok = true;
#pragma unroll
  for (int iv = 0; iv < L::NCONS; ++iv) {
    const R vec_iv = (iv < NM) ? s.ra[iv] : (iv < NM + 3 ? s.vm[iv - NM] : s.vE);
    const R h = star ? SK * (W.Us[iv] - vec_iv) : R(0.0);
    if (isnan(h)) ok = false;
    out[iv] = adv[iv] + h;
  }
  auto dot = [&](R f0, R f1, R f2) { return (f0 - R(0.0)) * nv[0] + (f1 - R(0.0)) * nv[1] + (f2 - R(0.0)) * nv[2]; };
#pragma unroll
  for (int v = 0; v < NM - 1; ++v)
    out[L::I3 + v] = dot(W.unc[0] * W.avf_S[v], W.unc[1] * W.avf_S[v], W.unc[2] * W.avf_S[v]) + R(0.0);
#pragma unroll
  for (int v = 0; v < NM; ++v)
    out[L::I4 + v] = dot(W.unc[0] * W.are_S[v], W.unc[1] * W.are_S[v], W.unc[2] * W.are_S[v]) + R(0.0);
#pragma unroll
  for (int k = 0; k < 9; k++) out[L::I5 + k] = dot(W.unc[0] * W.AS[k], W.unc[1] * W.AS[k], W.unc[2] * W.AS[k]) + R(0.0);
  unc[0] = W.unc[0]; unc[1] = W.unc[1]; unc[2] = W.unc[2];
  return ok;
}

// ---- relaxation (m5_relaxationvec.c) ----

template <int NM>
__device__ void vf_limit2(double a, double *avf) {   // m5_relaxationvec.c:131-150: two independent clips
  double avfsum = 0.0;
#pragma unroll
  for (int iv = 0; iv < NM - 1; ++iv) {
    if (avf[iv] <= a) avf[iv] = a;
    if (avf[iv] >= (1.0 - a)) avf[iv] = 1.0 - a;
    avfsum += avf[iv];
  }
  if (avfsum >= (1.0 - a)) {
    const double dif = avfsum - (1.0 - a);
#pragma unroll
    for (int j = 0; j < NM - 1; ++j) { const double per = avf[j] / avfsum; avf[j] = avf[j] - per * dif; }
    avfsum = 1.0 - a;
  }
  avf[NM - 1] = 1.0 - avfsum;
}

template <int NM>
__device__ void vf_limit1(double a, double *avf) {   // m5_relaxationvec.c:44-60: else-if clips
  double avfsum = 0.0;
#pragma unroll
  for (int iv = 0; iv < NM - 1; ++iv) {
    if (avf[iv] <= a) avf[iv] = a;
    else if (avf[iv] >= (1.0 - a)) avf[iv] = 1.0 - a;
    avfsum += avf[iv];
  }
  if (avfsum >= (1.0 - a)) {
    const double dif = avfsum - (1.0 - a);
#pragma unroll
    for (int j = 0; j < NM - 1; ++j) { const double per = avf[j] / avfsum; avf[j] = avf[j] - per * dif; }
    avfsum = 1.0 - a;
  }
  avf[NM - 1] = 1.0 - avfsum;
}

// F_pe.c (i_paper = 1) on the PRELAX 2 branch's own arrays: the published F_pe
// reads RUN fields that are never written
template <int NM>
__device__ double fpe5(const M5P &P, const double *ar, const double *t0, const double *p0, double pe) {
  double F = 0.0;
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) {
    const double G = P.gamma[iv], Pi = P.pinf[iv];
    F += ar[iv] * t0[iv] * ((p0[iv] + G * Pi + (G - 1.0) * pe) / (pe + G * Pi + (G - 1.0) * pe));
  }
  return F + -1.0;
}
template <int NM>
__device__ double dfpe5(const M5P &P, const double *ar, const double *t0, const double *p0, double pe) {
  double D = 0.0;
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) {
    const double G = P.gamma[iv], Pi = P.pinf[iv];
    const double c1 = ar[iv] * t0[iv];
    const double f = (p0[iv] + G * Pi + (G - 1.0) * pe);
    const double g = (pe + G * Pi + (G - 1.0) * pe);
    D += (c1 * (G - 1.0) * g - c1 * f * G) / sq(g);
  }
  return D;
}

// m5_relaxationvec.c in pieces.  RelaxIn holds what every branch reads
// (m5_relaxationvec.c:29-87); PRELAX 3's Newton iteration (:365-560) can be
// stopped after a number of iterations and resumed from (pr, F, dF, iter) with
// the same arithmetic, so the few slow states can be finished apart from the rest.
template <int NM>
struct RelaxIn {
  double ar[NM], are[NM], avf[NM], avf_init[NM], A[9], G[9], Gmag, rho, u, v, w, rE;
  GPow<> gp;
};

template <int NM>
__device__ void relax_in(const M5P &P, const double *x, RelaxIn<NM> &R) {
  using L = L5<NM>;
  R.rho = 0.0;
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) { R.ar[iv] = x[L::I0 + iv]; R.are[iv] = x[L::I4 + iv]; R.rho += R.ar[iv]; }
  if (P.fast) {
    const double ir = frcp(R.rho);
    R.u = x[L::I1 + 0] * ir; R.v = x[L::I1 + 1] * ir; R.w = x[L::I1 + 2] * ir;
  } else {
    R.u = x[L::I1 + 0] / R.rho; R.v = x[L::I1 + 1] / R.rho; R.w = x[L::I1 + 2] / R.rho;
  }
  R.rE = x[L::I2];
#pragma unroll
  for (int iv = 0; iv < NM - 1; ++iv) R.avf[iv] = x[L::I3 + iv];
  vf_limit1<NM>(P.amin / 100.0, R.avf);
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) R.avf_init[iv] = R.avf[iv];
#pragma unroll
  for (int k = 0; k < 9; ++k) R.A[k] = x[L::I5 + k];
  gmat(R.A, R.G, R.Gmag);
  R.gp = has_solid(P) ? gpow(P, R.Gmag) : GPow<>{0.0, 0.0};
}

// the re-initialisation after PRELAX 1, 2, 3 (m5_relaxationvec.c:157-189 and
// the same in the other branches) and the write-back (:580-620)
template <int NM>
__device__ int relax_out(const M5P &P, const RelaxIn<NM> &R, double *are, double *x) {
  using L = L5<NM>;
  if (P.prelax >= 1 && P.prelax <= 3) {
    double e_elast = 0.0, are_el[NM];
    // mixture mode: alpha_iv of the elastic energy of the relaxed A
    double Gx[9], Gmx = R.Gmag;
    GPow<> gpx = R.gp;
#pragma unroll
    for (int k = 0; k < 9; ++k) Gx[k] = R.G[k];
    if (P.mix) {
      gmat(R.A, Gx, Gmx);
      gpx = gpow(P, Gmx);
    }
    const double irho = P.fast ? frcp(R.rho) : 0.0;
#pragma unroll
    for (int iv = 0; iv < NM; ++iv) {
      are_el[iv] = (P.mush[iv] != 0.0) ? elastic_energy_v2(Gx, Gmx, gpx, P.mush[iv]) : 0.0;
      if (P.mix) are_el[iv] *= R.avf[iv];
      // Ys e_K = (ar / rho) (are_el / ar) = are_el / rho (P.fast)
      e_elast += P.fast ? are_el[iv] * irho : (R.ar[iv] / R.rho) * (are_el[iv] / (R.ar[iv]));
    }
    const double e_hyd = (P.fast ? R.rE * irho : R.rE / R.rho) - 0.5 * (sq((R.u)) + sq((R.v)) + sq((R.w))) - e_elast;
    double num = 0.0, den = 0.0, igm[NM];
#pragma unroll
    for (int iv = 0; iv < NM; ++iv) {
      igm[iv] = P.fast ? frcp(P.gamma[iv] - 1.0) : 0.0;
      num += P.fast ? R.avf[iv] * P.gamma[iv] * P.pinf[iv] * igm[iv] : R.avf[iv] * P.gamma[iv] * P.pinf[iv] / (P.gamma[iv] - 1.0);
      den += P.fast ? R.avf[iv] * igm[iv] : R.avf[iv] / (P.gamma[iv] - 1.0);
    }
    const double p_relaxed = P.fast ? (R.rho * e_hyd - num) * frcp(den) : (R.rho * e_hyd - num) / den;
    if (isnan(p_relaxed)) return -1;
#pragma unroll
    for (int iv = 0; iv < NM; ++iv)
      are[iv] = (P.fast ? R.avf[iv] * (p_relaxed + P.gamma[iv] * P.pinf[iv]) * igm[iv]
                        : R.avf[iv] * (p_relaxed + P.gamma[iv] * P.pinf[iv]) / (P.gamma[iv] - 1.0)) +
                are_el[iv];
  }
  bool bad = false;
#pragma unroll
  for (int iv = 0; iv < NM - 1; ++iv) { bad |= isnan(R.avf[iv]); x[L::I3 + iv] = R.avf[iv]; }
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) { bad |= isnan(are[iv]); x[L::I4 + iv] = are[iv]; }
#pragma unroll
  for (int k = 0; k < 9; ++k) x[L::I5 + k] = R.A[k];
  return bad ? -1 : 0;
}

// PRELAX 3 set-up (m5_relaxationvec.c:365-400): clamps, the phasic pressures
// pk and the starting pressure; returns irelax
template <int NM>
__device__ int r3_setup(const M5P &P, RelaxIn<NM> &R, double *pk, double &pr) {
  const double eps_sml = 1.0e-12, eps_big = 1.0e-8;
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) {
    if (R.avf[iv] < 0.0 || R.ar[iv] < 0.0) { R.avf[iv] = eps_sml; R.ar[iv] = eps_sml; R.are[iv] = eps_sml; }
    if (R.avf[iv] >= 1.0) R.avf[iv] = 1.0 - eps_sml;
  }
  int irelax = 1;
#pragma unroll
  for (int iv = 0; iv < NM; ++iv)
    if (R.avf[iv] > 1.0 - eps_sml) irelax = 0;
  pr = 0.0;
  if (irelax == 1) {
#pragma unroll
    for (int iv = 0; iv < NM; ++iv) {
      if (R.avf[iv] > eps_sml) {
        const double ael = ((P.mush[iv] != 0.0) ? elastic_energy_v2(R.G, R.Gmag, R.gp, P.mush[iv]) : 0.0) *
                           (P.mix ? R.avf[iv] : 1.0);
        pk[iv] = (P.fast ? (P.gamma[iv] - 1.0) * (R.are[iv] - ael) * frcp(R.avf[iv])
                         : (P.gamma[iv] - 1.0) * (R.are[iv] - ael) / R.avf[iv]) -
                 P.gamma[iv] * P.pinf[iv];
        if (pk[iv] <= -(1.0 - eps_big) * P.pinf[iv] + eps_big) pk[iv] = -(1.0 - eps_big) * P.pinf[iv] + eps_big;
      } else {
        pk[iv] = 0.0;
      }
      pr += R.avf[iv] * pk[iv];
    }
  }
  return irelax;
}

template <int NM>
__device__ double r3_rhoK(const M5P &P, const RelaxIn<NM> &R, const double *pk, double pr, int iv) {
  return R.ar[iv] / maxr(R.avf[iv], 1.0e-12) * pow((pr + P.pinf[iv]) / (pk[iv] + P.pinf[iv]), 1.0 / P.gamma[iv]);
}

// P.fast: ar_k / rho_k = max(alpha_k, 1e-12) x^(-1/gamma_k) with x the ratio
// (pr + pinf) / (pk + pinf) formed with the reciprocal ipk of its denominator,
// and x^y as exp(y ln x): a few ulp from the quotient of r3_rhoK
template <int NM>
__device__ double r3_q_fast(const M5P &P, const RelaxIn<NM> &R, double ipk, double pr, int iv) {
  return maxr(R.avf[iv], 1.0e-12) * exp(log((pr + P.pinf[iv]) * ipk) * (-1.0 / P.gamma[iv]));
}

// at most cap iterations of the Newton loop (m5_relaxationvec.c:409-470)
// from (pr, F, dF, iter): 1 when the loop ended, 0 when cap was reached first,
// -1 on NaN
template <int NM>
__device__ int r3_newton(const M5P &P, const RelaxIn<NM> &R, const double *pk, double &pr, double &Fv, double &dF,
                         int &iter, int cap) {
  const double eps_sml = 1.0e-12, eps_big = 1.0e-8;
  double ipk[NM];
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) ipk[iv] = P.fast ? frcp(pk[iv] + P.pinf[iv]) : 0.0;
  for (int k = 0; k < cap; k++) {
    if (!((fabs(Fv) > 1.0e-12) && (iter < 500))) return 1;
    const double pr0 = pr, F0 = Fv, dF0 = dF;
    pr = pr - Fv / dF;
    iter = iter + 1;
#pragma unroll
    for (int iv = 0; iv < NM; ++iv)
      if (pr < -(1.0 - eps_big) * P.pinf[iv] + eps_big) pr = -(1.0 - eps_big) * P.pinf[iv] + 1.0;
    Fv = -1.0;
    dF = 0.0;
#pragma unroll
    for (int iv = 0; iv < NM; ++iv) {
      if (R.avf[iv] > eps_sml) {
        if (P.fast) {
          const double q = r3_q_fast<NM>(P, R, ipk[iv], pr, iv);
          Fv = Fv + q;
          dF = dF - q * frcp(P.gamma[iv] * (pr + P.pinf[iv]));
        } else {
          const double rk = r3_rhoK<NM>(P, R, pk, pr, iv);
          Fv = Fv + R.ar[iv] / rk;
          dF = dF - R.ar[iv] / (P.gamma[iv] * rk * (pr + P.pinf[iv]));
        }
      }
    }
    if (isnan(pr) || isnan(Fv) || isnan(dF)) return -1;
    // an iteration that reproduces its own input (the clamp holding pr at the same
    // value) repeats itself up to the limit of 500: end there with the same values
    if (pr == pr0 && Fv == F0 && dF == dF0) { iter = 500; return 1; }
  }
  return ((fabs(Fv) > 1.0e-12) && (iter < 500)) ? 0 : 1;
}

// after the Newton loop (m5_relaxationvec.c:472-578)
template <int NM>
__device__ int r3_finish(const M5P &P, RelaxIn<NM> &R, const double *pk, int irelax, double pr, double *x) {
  if (irelax == 1) {
#pragma unroll
    for (int iv = 0; iv < NM; ++iv)
      if (R.avf[iv] > 1.0e-12)
        R.avf[iv] = P.fast ? r3_q_fast<NM>(P, R, frcp(pk[iv] + P.pinf[iv]), pr, iv)
                           : R.ar[iv] / r3_rhoK<NM>(P, R, pk, pr, iv);
    vf_limit1<NM>(P.amin / 100.0, R.avf);
  }
  const int i_as = ((R.avf[0] < R.avf[1]) && (P.mush[1] != 0.0)) ? 1 : 0;
  const double ratio = (i_as == 1) ? (R.avf[1] / R.avf_init[1]) : (R.avf[0] / R.avf_init[0]);
  const double sA = P.fix ? cbrt(1.0 / ratio) : ratio;   // consistent: det A follows 1 / alpha
#pragma unroll
  for (int i = 0; i < 9; i++) R.A[i] = R.A[i] * sA;
  double are[NM];
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) are[iv] = R.are[iv];
  return relax_out<NM>(P, R, are, x);
}

// m5_relaxationvec.c:3-620 on x in place: 0, -1 NaN, -3 PRELAX 2 Newton failure
template <int NM>
__device__ int relax5(const M5P &P, double *x) {
  const double a = P.amin / 100.0, eps_big = 1.0e-8;
  RelaxIn<NM> R;
  relax_in<NM>(P, x, R);
  if (P.prelax == 3) {
    double pk[NM], pr, Fv = 1.0e-9, dF = 1.0e9;
    int iter = 0;
    const int irelax = r3_setup<NM>(P, R, pk, pr);
    if (irelax == 1 && r3_newton<NM>(P, R, pk, pr, Fv, dF, iter, 1 << 30) < 0) return -1;
    return r3_finish<NM>(P, R, pk, irelax, pr, x);
  }
  double are[NM];
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) are[iv] = R.are[iv];
  if (P.prelax == 1 || P.prelax == 2) {
    double t0[NM], p0[NM], pe;
    if (P.prelax == 1) {   // m5_relaxationvec.c:89-190
#pragma unroll
      for (int iv = 0; iv < NM; ++iv) {
        const double ael = ((P.mush[iv] != 0.0) ? elastic_energy_v2(R.G, R.Gmag, R.gp, P.mush[iv]) : 0.0) *
                           (P.mix ? R.avf[iv] : 1.0);
        t0[iv] = 1.0 / (R.ar[iv] / R.avf[iv]);
        p0[iv] = ((P.gamma[iv] - 1.0) * (R.are[iv] - ael) / R.avf[iv]) - P.gamma[iv] * P.pinf[iv];
      }
      const double G0 = P.gamma[0], G1 = P.gamma[1], P0 = P.pinf[0], P1 = P.pinf[1];
      const double A_1 = (R.avf[0] / G0) * (p0[0] + P0) / ((R.avf[0] / G0) + (R.avf[1] / G1));
      const double A_2 = (R.avf[1] / G1) * (p0[1] + P1) / ((R.avf[0] / G0) + (R.avf[1] / G1));
      pe = 0.5 * (A_1 + A_2 - (P0 + P1)) + sqrt(0.25 * (sq((A_2 - A_1 - (P1 - P0)))) + A_1 * A_2);
      if (isnan(pe)) return -1;
    } else {               // m5_relaxationvec.c:192-363
      double max_avf = -1.0;
      int principal = 0;
#pragma unroll
      for (int iv = 0; iv < NM; ++iv) {
        const double ael = ((P.mush[iv] != 0.0) ? elastic_energy_v2(R.G, R.Gmag, R.gp, P.mush[iv]) : 0.0) *
                           (P.mix ? R.avf[iv] : 1.0);
        t0[iv] = 1.0 / (R.ar[iv] / R.avf[iv]);
        p0[iv] = ((P.gamma[iv] - 1.0) * (R.are[iv] - ael) / R.avf[iv]) - P.gamma[iv] * P.pinf[iv];
        if (p0[iv] < -(1.0 - eps_big) * P.pinf[iv] + eps_big) p0[iv] = -(1.0 - eps_big) * P.pinf[iv] + 1.0;
        if (R.avf[iv] > max_avf) { max_avf = R.avf[iv]; principal = iv; }
      }
      double pe_old = 0.001, pe_new = pe_old;
      int iter = 0;
      double cond_1 = fabs(pe_new - pe_old);
      double cond_2 = fabs(fpe5<NM>(P, R.ar, t0, p0, pe_new));
      while (((cond_2 >= 1.0e-12) || (cond_1 >= 1.0e-4)) && (iter <= 101)) {
        pe_new = pe_old - 0.8 * (fpe5<NM>(P, R.ar, t0, p0, pe_old) / dfpe5<NM>(P, R.ar, t0, p0, pe_old));
        cond_1 = fabs(pe_new - pe_old);
        cond_2 = fabs(fpe5<NM>(P, R.ar, t0, p0, pe_new));
        if (iter >= 100 && ((fabs(pe_old - pe_new) < 1.0e3) || isnan(pe_new))) return -3;
        pe_old = pe_new;
        iter++;
      }
      if (isnan(pe_new)) return -1;
      double lim = -(1.0 - eps_big) * P.pinf[0];
#pragma unroll
      for (int iv = 1; iv < NM; ++iv)
        if (iv == principal) lim = -(1.0 - eps_big) * P.pinf[iv];
      if (pe_new < (lim + eps_big)) pe_new = lim + 1.0;
      pe = pe_new;
    }
#pragma unroll
    for (int iv = 0; iv < NM; ++iv) {
      const double Gk = P.gamma[iv], Pk = P.pinf[iv];
      const double t1 = t0[iv] * (p0[iv] + Gk * Pk + (Gk - 1.0) * pe) / (pe + Gk * Pk + (Gk - 1.0) * pe);
      const double r_m = 1.0 / t1;
      R.avf[iv] = R.ar[iv] / r_m;
    }
    vf_limit2<NM>(a, R.avf);
    const double ratio = R.avf[0] / R.avf_init[0];
    const double sA = P.fix ? cbrt(1.0 / ratio) : ratio;   // consistent: det A follows 1 / alpha
#pragma unroll
    for (int i = 0; i < 9; i++) R.A[i] = R.A[i] * sA;
  }
  return relax_out<NM>(P, R, are, x);
}

// ---- perfect plasticity ----------------------------------------------------
// New physics: the reference's plastic_rlx (m5_relaxation_solid.c:14) is
// commented out.  With G = A A^T = Q diag(l) Q^T, d = det G and log strains
// k_i = ln l_i - ln(d)/3 (sum 0), the state moves along l_i(t) = d^(1/3)
// exp(t k_i), which keeps det G, from t = 1 towards the stress-free t = 0, to
// the largest t whose von Mises stress of the deviatoric stress of
// state_equation.c:143-204 (coefficient sum_k mu_k) is at most the yield
// stress: A becomes Q diag(exp((t - 1) k_i / 2)) Q^T A.  The energies stay, so
// the released elastic energy reappears as pressure.

// cyclic Jacobi on the symmetric 3x3 a (destroyed): a = Q diag(l) Q^T
__device__ void sym3_eigen(double a[9], double l[3], double Q[9]) {
#pragma unroll
  for (int k = 0; k < 9; k++) Q[k] = (k % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 12; sweep++) {
    const double off = fabs(a[1]) + fabs(a[2]) + fabs(a[5]);
    const double dia = fabs(a[0]) + fabs(a[4]) + fabs(a[8]);
    if (off <= 1e-300 || off <= 1e-18 * dia) break;
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const int P = (r == 2) ? 1 : 0, q = (r == 0) ? 1 : 2;
      const double apq = a[P * 3 + q];
      if (apq == 0.0) continue;
      const double th = (a[q * 3 + q] - a[P * 3 + P]) / (2.0 * apq);
      const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double akp = a[k * 3 + P], akq = a[k * 3 + q];
        a[k * 3 + P] = c * akp - s * akq;
        a[k * 3 + q] = s * akp + c * akq;
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double apk = a[P * 3 + k], aqk = a[q * 3 + k];
        a[P * 3 + k] = c * apk - s * aqk;
        a[q * 3 + k] = s * apk + c * aqk;
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double qkp = Q[k * 3 + P], qkq = Q[k * 3 + q];
        Q[k * 3 + P] = c * qkp - s * qkq;
        Q[k * 3 + q] = s * qkp + c * qkq;
      }
    }
  }
  l[0] = a[0]; l[1] = a[4]; l[2] = a[8];
}

// von Mises stress along the path, divided by mu d^(1/2)
__device__ double vm_path(const double k[3], double t) {
  double e1[3], e2[3], s1 = 0.0, s2 = 0.0, q = 0.0;
#pragma unroll
  for (int i = 0; i < 3; i++) { e1[i] = exp(t * k[i]); e2[i] = e1[i] * e1[i]; s1 += e1[i]; s2 += e2[i]; }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double si = (e2[i] - s2 / 3.0) - (e1[i] - s1 / 3.0);
    q += si * si;
  }
  return sqrt(1.5 * q);
}

// the return map on A (9 values, row-major) in place: 1 changed, 0 not, -1 NaN
__device__ int plastic_A(const M5P &P, double mu, double *A) {
  double G[9], d;
  gmat(A, G, d);
  if (!(d > 0.0)) return 0;
  double G2[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G2[i * 3 + j] = G[i * 3 + 0] * G[0 * 3 + j] + G[i * 3 + 1] * G[1 * 3 + j] + G[i * 3 + 2] * G[2 * 3 + j];
  const double J1 = G[0] + G[4] + G[8], J2 = G2[0] + G2[4] + G2[8], d13 = cbrt(d);
  double q = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    double ssk = G2[k] - d13 * G[k];
    if (k % 4 == 0) ssk -= J2 / 3.0 - d13 * J1 / 3.0;
    q += ssk * ssk;
  }
  const double vm = mu / pow(d, 1.0 / 6.0) * sqrt(1.5 * q);
  if (!(vm > P.yield)) return 0;
  double l[3], Q[9], kap[3];
  sym3_eigen(G, l, Q);
  const double lnd3 = (log(l[0]) + log(l[1]) + log(l[2])) / 3.0;
#pragma unroll
  for (int i = 0; i < 3; i++) kap[i] = log(l[i]) - lnd3;
  const double target = P.yield / (mu * sqrt(d));
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 64; it++) {
    const double mid = 0.5 * (lo + hi);
    if (vm_path(kap, mid) > target) hi = mid; else lo = mid;
  }
  double f[3], M[9], An[9];
#pragma unroll
  for (int i = 0; i < 3; i++) f[i] = exp(0.5 * (lo - 1.0) * kap[i]);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      M[i * 3 + j] = Q[i * 3 + 0] * f[0] * Q[j * 3 + 0] + Q[i * 3 + 1] * f[1] * Q[j * 3 + 1] + Q[i * 3 + 2] * f[2] * Q[j * 3 + 2];
  int bad = 0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      An[i * 3 + j] = M[i * 3 + 0] * A[0 * 3 + j] + M[i * 3 + 1] * A[1 * 3 + j] + M[i * 3 + 2] * A[2 * 3 + j];
      bad |= isnan(An[i * 3 + j]);
    }
  if (bad) return -1;
#pragma unroll
  for (int k = 0; k < 9; k++) A[k] = An[k];
  return 1;
}

// the return map on the A rows of every leaf of S [NEQ][n], run between m5_RK
// and the relaxation (where m5_relaxation_solid.c:14 would call plastic_rlx);
// the launch is skipped when the yield stress is 0
template <int NM>
__global__ __launch_bounds__(256) void k_m5_plastic(int n, M5P P, double *__restrict__ S, int *__restrict__ flag) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double mu = 0.0;
#pragma unroll
  for (int k = 0; k < NM; k++) mu += P.mush[k];
  if (mu == 0.0) return;
  double A[9];
#pragma unroll
  for (int k = 0; k < 9; k++) A[k] = S[(size_t)(L::I5 + k) * n + i];
  const int rc = plastic_A(P, mu, A);
  if (rc < 0) { atomicOr(flag, 1); return; }
  if (rc == 0) return;
#pragma unroll
  for (int k = 0; k < 9; k++) S[(size_t)(L::I5 + k) * n + i] = A[k];
}

// PRELAX 3 over many states (face states or cell states) with the slow ones
// finished apart: state q of `count` is element v at X[(q / n) * NEQ * n + v * n + q % n]
// (SF [6][NEQ][n] for face states, S [NEQ][n] for cell states).  The start
// pass gives every state R3_CAP iterations; a state still iterating keeps its
// unrelaxed vector in X, saves (pr, F, dF, iter) and is marked; resume
// passes continue the marked ones from there.  Sb, when given, receives the
// state before relaxation (m5_relaxation_solid.c:18-24).
constexpr int R3_CAP = 8;

template <int NM>
__device__ void load_state(const double *X, int n, long q, double *x) {
  using L = L5<NM>;
  const long f = q / n, i = q - f * n;
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) x[v] = X[(size_t)f * L::NEQ * n + (size_t)v * n + i];
}
template <int NM>
__device__ void store_state(double *X, int n, long q, const double *x) {
  using L = L5<NM>;
  const long f = q / n, i = q - f * n;
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) X[(size_t)f * L::NEQ * n + (size_t)v * n + i] = x[v];
}

// appends val to list at *cnt for the lanes with want, one atomic per wave
// (the list's order is free: every listed state is resumed on its own)
__device__ __forceinline__ void wave_append(int *cnt, int *list, int val, bool want) {
  const unsigned long long m = __ballot(want);
  if (m == 0ull) return;
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(cnt, __popcll(m));
  base = __shfl(base, leader);
  if (want) list[base + __popcll(m & ((1ull << lane) - 1ull))] = val;
}

// the first R3_CAP iterations of state q held in x: 1 when the loop ended and
// x is relaxed; 0 when x keeps its unrelaxed values and the Newton state is
// saved for the resume passes; -1 on NaN (x unrelaxed, the flag raised)
template <int NM>
__device__ int r3_first(const M5P &P, double *x, long q, double *__restrict__ nst, int *__restrict__ iters,
                        int *__restrict__ flag) {
  RelaxIn<NM> R;
  relax_in<NM>(P, x, R);
  double pk[NM], pr, Fv = 1.0e-9, dF = 1.0e9;
  int iter = 0;
  const int irelax = r3_setup<NM>(P, R, pk, pr);
  int rc = 1;
  if (irelax == 1) rc = r3_newton<NM>(P, R, pk, pr, Fv, dF, iter, R3_CAP);
  if (rc < 0) { atomicOr(flag, 1); return -1; }
  if (rc == 0) {
    nst[3 * q + 0] = pr; nst[3 * q + 1] = Fv; nst[3 * q + 2] = dF;
    iters[q] = iter;
    return 0;
  }
  if (r3_finish<NM>(P, R, pk, irelax, pr, x) < 0) atomicOr(flag, 1);
  return 1;
}

template <int NM>
__global__ __launch_bounds__(256) void k_m5_relax3_start(long count, int n, M5P P, double *__restrict__ X,
                                                         double *__restrict__ Sb, double *__restrict__ nst,
                                                         int *__restrict__ iters, int *__restrict__ cnt,
                                                         int *__restrict__ list, int *__restrict__ flag) {
  using L = L5<NM>;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  double x[L::NEQ];
  load_state<NM>(X, n, q, x);
  if (Sb) store_state<NM>(Sb, n, q, x);
  const int st = r3_first<NM>(P, x, q, nst, iters, flag);
  if (st == 1) store_state<NM>(X, n, q, x);
  wave_append(cnt, list, (int)q, st == 0);
}

// resumes the listed states for at most cap more iterations; those still
// iterating save their Newton state again and go to olist (when given)
template <int NM>
__global__ __launch_bounds__(256) void k_m5_relax3_resume(const int *__restrict__ nlist_p,
                                                          const int *__restrict__ list, int n, M5P P,
                                                          double *__restrict__ X, double *__restrict__ nst,
                                                          int *__restrict__ iters, int *__restrict__ ocnt,
                                                          int *__restrict__ olist, int cap, int *__restrict__ flag) {
  using L = L5<NM>;
  const int nlist = *nlist_p;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nlist; t += gridDim.x * blockDim.x) {
    const long q = list[t];
    double x[L::NEQ];
    load_state<NM>(X, n, q, x);
    RelaxIn<NM> R;
    relax_in<NM>(P, x, R);
    double pk[NM], pr;
    const int irelax = r3_setup<NM>(P, R, pk, pr);
    pr = nst[3 * q + 0];
    double Fv = nst[3 * q + 1], dF = nst[3 * q + 2];
    int iter = iters[q];
    const int rc = r3_newton<NM>(P, R, pk, pr, Fv, dF, iter, cap);
    if (rc < 0) {
      atomicOr(flag, 1);
    } else if (rc == 0) {
      nst[3 * q + 0] = pr; nst[3 * q + 1] = Fv; nst[3 * q + 2] = dF;
      iters[q] = iter;
    } else {
      if (r3_finish<NM>(P, R, pk, irelax, pr, x) < 0) atomicOr(flag, 1);
      store_state<NM>(X, n, q, x);
    }
    if (olist) wave_append(ocnt, olist, (int)q, rc == 0);
  }
}

__global__ void k_zero_ints(int *p, int k) {
  if ((int)threadIdx.x < k) p[threadIdx.x] = 0;
}

// avf_limit.c:3-21 with max_limit 100 on the fractions of X
template <int NM>
__device__ void avf_limit5(double a, double *X) {
  using L = L5<NM>;
  double avf[NM], avfsum = 0.0;
#pragma unroll
  for (int v = 0; v < NM - 1; ++v) {
    avf[v] = X[L::I3 + v];
    if (avf[v] <= a) avf[v] = a;
    else if (avf[v] >= (1.0 - a)) avf[v] = 1.0 - a;
    avfsum += avf[v];
  }
  if (avfsum > (1.0 - a)) {
    const double dif = avfsum - (1.0 - a);
#pragma unroll
    for (int j = 0; j < NM - 1; ++j) avf[j] -= (avf[j] / avfsum) * dif;
  }
#pragma unroll
  for (int v = 0; v < NM - 1; ++v) X[L::I3 + v] = avf[v];
}

// m5_facevalues_MUSCL.c:3-44, iside = 1: the value across face f of leaf i
template <int NM>
__device__ void stencil5(const M5P &P, const double *__restrict__ S, int n, int i, int kind, int nb,
                         const int *__restrict__ fine4, double *X) {
  using L = L5<NM>;
  if (kind == FV_FK_BOUNDARY) {
    // the own cell's record vector: raw state with the clipped fractions (m5_bc moves only u, v, w)
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) X[v] = S[(size_t)v * n + i];
    const double a = P.amin / 100.0;
    double at[NM], sum = 0.0;
#pragma unroll
    for (int k = 0; k < NM - 1; ++k) {
      at[k] = X[L::I3 + k];
      if (at[k] <= a) at[k] = a;
      else if (at[k] >= (1.0 - a)) at[k] = 1.0 - a;
      sum += at[k];
    }
    if (sum > (1.0 - a)) {
      const double dif = sum - (1.0 - a);
#pragma unroll
      for (int j = 0; j < NM - 1; ++j) { const double per = at[j] / sum; at[j] = at[j] - per * dif; }
    }
#pragma unroll
    for (int k = 0; k < NM - 1; ++k) X[L::I3 + k] = at[k];
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
  avf_limit5<NM>(P.amin / 100.0, X);
}

__device__ void dist5(const Geo &G, int kind, int lvl, double &dcf, double &dcc) {
  if (kind == FV_FK_SAME || kind == FV_FK_MIRROR) { dcf = G.dcf[0][lvl]; dcc = G.dcc[0][lvl]; }
  else if (kind == FV_FK_FINE) { dcf = G.dcf[1][lvl]; dcc = G.dcc[1][lvl]; }
  else if (kind == FV_FK_COARSE) { dcf = G.dcf[2][lvl]; dcc = G.dcc[2][lvl]; }
  else { dcf = 0.0; dcc = 0.0; }
}

// m5_faceinterp_V1 (m5_faceinterp.c:60-93) for one leaf per thread: MUSCL face
// states with minmod (m5_reconctruct.c:93-190; g_limitedvars is 0 on the
// alpha-rho-e block, m5_input.c:196-204), the VF limit and m5_relaxationvec
// (RX3: PRELAX 3's first R3_CAP Newton iterations of each face state, the
// states still iterating left marked for k_m5_relax3_resume; the face states
// then go to HBM once).  R = float (mixed precision)
// forms the limited slopes in FP32 from the FP64 differences and adds them to
// the FP64 cell state; the relaxation stays FP64.
// the limited slope of one variable: minmod as m5_reconctruct.c:110-140 has it,
// or (new) the monotonised central slope minmod(2a, 2b, (a + b) / 2) with a the
// backward and b the forward difference quotient
// FAST (the FP64 and mixed stages): the two difference quotients with the
// reciprocals of the face's distances, formed once per face
template <typename R>
__device__ __forceinline__ R m5_slope_q(bool mc, R a, R b) {
  if (!mc) {
    if (a * b > R(0.0)) return (fabs(a) > fabs(b)) ? b : a;
    return R(0.0);
  }
  if (a * b > R(0.0)) {
    const R c = R(0.5) * (a + b);
    const R mag = minr(minr(R(2.0) * fabs(a), R(2.0) * fabs(b)), fabs(c));
    return (c > R(0.0)) ? mag : -mag;
  }
  return R(0.0);
}

template <typename R>
__device__ __forceinline__ R m5_slope(bool mc, R df, R db, R dp, R dm) {
  if (!mc) {
    const R product = (-db / dm) * (df / dp);
    if (product > R(0.0)) {
      if (fabs(-db / dm) > fabs(df / dp)) return df / dp;
      return -db / dm;
    }
    return R(0.0);
  }
  const R a = -db / dm, b = df / dp;
  if (a * b > R(0.0)) {
    const R c = R(0.5) * (a + b);
    const R mag = minr(minr(R(2.0) * fabs(a), R(2.0) * fabs(b)), fabs(c));
    return (c > R(0.0)) ? mag : -mag;
  }
  return R(0.0);
}

template <int NM, typename R, bool RX3, bool FAST>
__global__ __launch_bounds__(256) void k_m5_recon(int n, M5P P, const Geo *__restrict__ Gp,
                                                  const double *__restrict__ S, const int *__restrict__ nbr,
                                                  const uint32_t *__restrict__ info, const int *__restrict__ fine4,
                                                  const uint8_t *__restrict__ limf, double *__restrict__ SF,
                                                  double *__restrict__ nst, int *__restrict__ iters,
                                                  int *__restrict__ cnt, int *__restrict__ list,
                                                  int *__restrict__ flag) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = P.amin / 100.0;
  const uint32_t inf = info[i];
  const int lvl = fvk::cell_level(inf);
  const bool mc = (P.limiter == 1) || (P.limiter == 2 && limf[i] != 0);
  double Sc[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) Sc[v] = S[(size_t)v * n + i];
  avf_limit5<NM>(a, Sc);
  // per axis, the two stencil values across its faces are gathered once and
  // serve both faces (each face's forward value is the other's backward one)
#pragma unroll 1
  for (int ax = 0; ax < 3; ax++) {
  const int fpl = (ax == 0) ? 1 : (ax == 1 ? 2 : 5), fmi = FCOPP5[fpl];   // the + and - faces
  const int kpl = fvk::face_kind(inf, fpl), kmi = fvk::face_kind(inf, fmi);
  double Xp[L::NEQ], Xm[L::NEQ];
  stencil5<NM>(P, S, n, i, kpl, nbr[(size_t)fpl * n + i], fine4, Xp);
  stencil5<NM>(P, S, n, i, kmi, nbr[(size_t)fmi * n + i], fine4, Xm);
  double dcf_p, dcc_p, dcf_m, dcc_m;
  dist5(*Gp, kpl, lvl, dcf_p, dcc_p);
  dist5(*Gp, kmi, lvl, dcf_m, dcc_m);
#pragma unroll 1
  for (int side = 0; side < 2; side++) {
    const int f = side ? fmi : fpl;
    const int kf = side ? kmi : kpl;
    const double dcf = side ? dcf_m : dcf_p, dp = side ? dcc_m : dcc_p, dm = side ? dcc_p : dcc_m;
    double Fv[L::NEQ];
    const R idp = FAST ? frcp((R)dp) : R(0.0), idm = FAST ? frcp((R)dm) : R(0.0);
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) {
      const double xf = side ? Xm[v] : Xp[v], xb = side ? Xp[v] : Xm[v];
      const R df = (R)(xf - Sc[v]), db = (R)(xb - Sc[v]);
      const R lim = (v >= L::I4 && v < L::I5) ? R(0.0) : R(1.0);
      const R dpr = (R)dp, dmr = (R)dm;
      R g = R(0.0);
      if (kf != FV_FK_BOUNDARY && fabs(df) > R(1.0e-10))
        g = lim * (FAST ? m5_slope_q<R>(mc, -db * idm, df * idp) : m5_slope<R>(mc, df, db, dpr, dmr));
      Fv[v] = Sc[v] + dcf * (double)g;
    }
    // VF limit (m5_reconctruct.c:142-170)
    double avf[NM], avfsum = 0.0;
#pragma unroll
    for (int k = 0; k < NM - 1; ++k) {
      avf[k] = Fv[L::I3 + k];
      if (avf[k] <= a) avf[k] = a;
      else if (avf[k] >= (1.0 - (2.0 * a))) avf[k] = 1.0 - (2.0 * a);
      avfsum += avf[k];
    }
    if (avfsum >= (1.0 - a)) {
      const double dif = avfsum - (1.0 - a);
#pragma unroll
      for (int j = 0; j < NM - 1; ++j) { const double per = avf[j] / avfsum; avf[j] = avf[j] - per * dif; }
    }
#pragma unroll
    for (int k = 0; k < NM - 1; ++k) Fv[L::I3 + k] = avf[k];
    if (P.fix) {
      // consistent: each phase's alpha-rho-e follows its reconstructed fraction, which
      // keeps the phase's cell pressure at the face
      double sf = 0.0, sc = 0.0;
#pragma unroll
      for (int k = 0; k < NM - 1; ++k) { sf += Fv[L::I3 + k]; sc += Sc[L::I3 + k]; }
#pragma unroll
      for (int k = 0; k < NM; ++k) {
        const double af = (k < NM - 1) ? Fv[L::I3 + k] : 1.0 - sf;
        const double ac = (k < NM - 1) ? Sc[L::I3 + k] : 1.0 - sc;
        Fv[L::I4 + k] = Sc[L::I4 + k] * (af / ac);
      }
    }
    if constexpr (RX3) {
      const int st = r3_first<NM>(P, Fv, (long)f * n + i, nst, iters, flag);
      wave_append(cnt, list, f * n + i, st == 0);
    } else {
      const int rc = relax5<NM>(P, Fv);
      if (rc) atomicOr(flag, rc == -3 ? 4 : 1);
    }
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) SF[((size_t)f * L::NEQ + v) * n + i] = Fv[v];
  }
  }
}

// the record a face flux reads: the face state (ORDER 2, m5_facercnstr) or
// the cell state (ORDER 1, m5_facevalues) of leaf j across its face fj
template <int NM, typename R, bool FAST = false>
__device__ void side_record(const M5P &P, const double *__restrict__ S, const double *__restrict__ SF, int n, int j,
                            int fj, Rec5<NM, R> &r) {
  using L = L5<NM>;
  double x[L::NEQ];
  Rec5<NM> r64;
  if (P.order == 2) {
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) x[v] = SF[((size_t)fj * L::NEQ + v) * n + j];
    record<NM, true, FAST>(P, x, r64, S[(size_t)L::I3 * n + j]);
  } else {
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) x[v] = S[(size_t)v * n + j];
    record<NM, false, FAST>(P, x, r64);
  }
  narrow<NM, R>(r64, r);
}

// m5_residual.c:7-238 for one leaf per thread.  A mirrored face (the earlier
// leaf's flux_flag protocol, a face kind of the pack) recomputes the earlier
// leaf's evaluation of that face and negates it, as the reference reuses it.
// gfm bit f: the face used the GFM flux.
template <int NM, typename R>
__global__ __launch_bounds__(256) void k_m5_residual(int n, M5P P, const Geo *__restrict__ Gp,
                                                     const double *__restrict__ S, const double *__restrict__ SF,
                                                     const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                     const int *__restrict__ fine4, double *__restrict__ RHS,
                                                     uint8_t *__restrict__ gfm, int *__restrict__ flag) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Geo &G = *Gp;
  const uint32_t inf = info[i];
  const int lvl = fvk::cell_level(inf);
  // sources of the cell record (m5_residual.c:31-59)
  double src_avf[NM - 1], src_st[NM][9], Am[9];
  {
    double x[L::NEQ];
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) x[v] = S[(size_t)v * n + i];
    Rec5<NM> sc;
    record<NM>(P, x, sc);
#pragma unroll
    for (int k = 0; k < NM - 1; ++k) src_avf[k] = sc.avf[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) Am[k] = sc.A[k];
    double Gm[9], Gmag;
    gmat(sc.A, Gm, Gmag);
    const GPow<> gps = has_solid(P) ? gpow(P, Gmag) : GPow<>{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NM; ++k)
      stress_v2(src_st[k], sc.avf[k], sc.ph[k], Gm, Gmag, gps, P.mix ? sc.avf[k] * P.mush[k] : P.mush[k]);
  }
  const double A1 = Am[0], B1 = Am[3], C1 = Am[6];
  const double A2 = Am[1], B2 = Am[4], C2 = Am[7];
  const double A3 = Am[2], B3 = Am[5], C3 = Am[8];
  double rhs[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) rhs[v] = 0.0;
  uint8_t gbits = 0;
  bool ok = true;
#pragma unroll 1
  for (int f = 0; f < 6; f++) {
    const int kind = fvk::face_kind(inf, f);
    const int nb = nbr[(size_t)f * n + i];
    const int fo = FCOPP5[f];
    double ff[L::NEQ], uf0 = 0.0, uf1 = 0.0, uf2 = 0.0, Area_f = 0.0;
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) ff[v] = 0.0;
    if (kind == FV_FK_MIRROR) {
      Rec5<NM, R> sL, sR;
      side_record<NM, R>(P, S, SF, n, nb, fo, sL);
      side_record<NM, R>(P, S, SF, n, i, f, sR);
      R out[L::NEQ], un[3], fr[9];
#pragma unroll
      for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[fo][q];
      const bool g = sL.phi0 * sR.phi0 < R(0.0);
      ok &= g ? gfm_subface<NM, R>(P, fr, sL, sR, out, un)
              : hllc_subface<NM, R>(P, fr, 0, sL, sR, S, n, nb, out, un);
      const double A = G.area[lvl];
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) {
        double t = 0.0;
        t += (A) * (double)out[v];
        ff[v] = -t;
      }
      // the star velocity is a global vector: the published code negates it with the flux
      const double us = P.fix ? 1.0 : -1.0;
      uf0 = us * ((double)un[0] / ((double)1)); uf1 = us * ((double)un[1] / ((double)1));
      uf2 = us * ((double)un[2] / ((double)1));
      Area_f += A;
    } else {
      const int Nneig = (kind == FV_FK_FINE) ? 4 : 1;
      const double A = (kind == FV_FK_FINE) ? G.area[lvl + 1] : G.area[lvl];
      for (int k = 0; k < Nneig; ++k) Area_f += A;
      R fr[9];
#pragma unroll
      for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[f][q];
      Rec5<NM, R> sL;
      side_record<NM, R>(P, S, SF, n, i, f, sL);
      for (int k = 0; k < Nneig; ++k) {
        Rec5<NM, R> sR;
        int bcode = 0;
        if (kind == FV_FK_BOUNDARY) {
          bcode = -nb;
          side_record<NM, R>(P, S, SF, n, i, f, sR);
          apply_bc5<NM, R>(bcode, fr, sL, sR);
        } else {
          const int j = (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
          side_record<NM, R>(P, S, SF, n, j, fo, sR);
        }
        R out[L::NEQ], un[3];
        if (sL.phi0 * sR.phi0 < R(0.0)) {
          gbits |= (uint8_t)(1u << f);
          Rec5<NM, R> gL = sL;   // the ghost rewrites only this sub-face's copies
          ok &= gfm_subface<NM, R>(P, fr, gL, sR, out, un);
        } else {
          ok &= hllc_subface<NM, R>(P, fr, bcode, sL, sR, S, n, i, out, un);
        }
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) ff[v] += (A) * (double)out[v];
        uf0 = (double)un[0] / ((double)Nneig);   // overwritten per sub-face (m5_residual.c:163-165)
        uf1 = (double)un[1] / ((double)Nneig);
        uf2 = (double)un[2] / ((double)Nneig);
      }
    }
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) rhs[v] += ff[v];
    // non-conservative terms (m5_residual.c:176-214)
    const double u_nc = uf0, nx = G.frame[f][0];
    const double v_nc = uf1, ny = G.frame[f][1];
    const double w_nc = uf2, nz = G.frame[f][2];
    if (P.fix) {
      // consistent sources (the residual is subtracted in this mode): the published alpha and
      // alpha-rho-e terms negated, and for A_ij n_i (A_kj u_k) - A_ij (u.n), all from the cell
#pragma unroll
      for (int k = 0; k < NM - 1; ++k)
        rhs[L::I3 + k] -= Area_f * (src_avf[k] * u_nc * nx + src_avf[k] * v_nc * ny + src_avf[k] * w_nc * nz);
#pragma unroll
      for (int k = 0; k < NM; ++k) {
        const double *s = src_st[k];
        rhs[L::I4 + k] -= Area_f * (s[0] * u_nc * nx + s[1] * v_nc * nx + s[2] * w_nc * nx + s[3] * u_nc * ny +
                                    s[4] * v_nc * ny + s[5] * w_nc * ny + s[6] * u_nc * nz + s[7] * v_nc * nz +
                                    s[8] * w_nc * nz);
      }
      const double nn[3] = {nx, ny, nz}, uu[3] = {u_nc, v_nc, w_nc};
      const double un = u_nc * nx + v_nc * ny + w_nc * nz;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double cu = Am[0 * 3 + j] * uu[0] + Am[1 * 3 + j] * uu[1] + Am[2 * 3 + j] * uu[2];
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) rhs[L::I5 + 3 * ii + j] += Area_f * (nn[ii] * cu - Am[ii * 3 + j] * un);
      }
      continue;
    }
#pragma unroll
    for (int k = 0; k < NM - 1; ++k)
      rhs[L::I3 + k] += Area_f * (src_avf[k] * u_nc * nx + src_avf[k] * v_nc * ny + src_avf[k] * w_nc * nz);
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      const double *s = src_st[k];
      rhs[L::I4 + k] += Area_f * (s[0] * u_nc * nx + s[1] * v_nc * nx + s[2] * w_nc * nx + s[3] * u_nc * ny +
                                  s[4] * v_nc * ny + s[5] * w_nc * ny + s[6] * u_nc * nz + s[7] * v_nc * nz +
                                  s[8] * w_nc * nz);
    }
    const int v5 = L::I5;
    rhs[v5 + 0] += Area_f * ((A1 * v_nc * ny + A1 * w_nc * nz) + B1 * v_nc * nx - C1 * w_nc * nx);
    rhs[v5 + 1] += Area_f * ((A2 * v_nc * ny + A2 * w_nc * nz) + B2 * v_nc * nx - C2 * w_nc * nx);
    rhs[v5 + 2] += Area_f * ((A3 * v_nc * ny + A3 * w_nc * nz) + B3 * v_nc * nx - C3 * w_nc * nx);
    rhs[v5 + 3] += Area_f * ((B1 * u_nc * nx + B1 * w_nc * nz) + A1 * u_nc * ny - C1 * w_nc * ny);
    rhs[v5 + 4] += Area_f * ((B2 * u_nc * nx + B2 * w_nc * nz) + A2 * u_nc * ny - C2 * w_nc * ny);
    rhs[v5 + 5] += Area_f * ((B3 * u_nc * nx + B3 * w_nc * nz) + A3 * u_nc * ny - C3 * w_nc * ny);
    rhs[v5 + 6] += Area_f * ((C1 * u_nc * nx + C1 * v_nc * ny) + A1 * u_nc * nz - B1 * v_nc * nz);
    rhs[v5 + 7] += Area_f * ((C2 * u_nc * nx + C2 * v_nc * ny) + A2 * u_nc * nz - B2 * v_nc * nz);
    rhs[v5 + 8] += Area_f * ((C3 * u_nc * nx + C3 * v_nc * ny) + A3 * u_nc * nz - B3 * v_nc * nz);
  }
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) {
    ok &= !isnan(rhs[v]);
    RHS[(size_t)v * n + i] = rhs[v];
  }
#pragma unroll
  for (int f = 0; f < 6; f++) gfm[6 * i + f] = (gbits >> f) & 1u;
  if (!ok) atomicOr(flag, 1);
}

// The residual with the non-conservative terms formed once after the face loop
// (R = float: the mixed-precision residual with FP32 face records, HLLC and GFM
// flux; R = double: the default FP64 residual).  k_m5_residual's face loop, the
// area-weighted face fluxes added in FP64 straight into the residual.  Every
// source term of m5_residual.c:176-214 (and of the consistent mode) is linear
// in the face's u* n^T, so they follow from T_ij = sum_f Area_f n_i u*_j and the
// cell record, which is built after the loop and holds no registers during it.
// The sums differ from k_m5_residual's face-by-face source terms by rounding.
template <int NM, typename R>
__global__ __launch_bounds__(256) void k_m5_residual_mx(int n, M5P P, const Geo *__restrict__ Gp,
                                                        const double *__restrict__ S, const double *__restrict__ SF,
                                                        const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                        const int *__restrict__ fine4, double *__restrict__ RHS,
                                                        uint8_t *__restrict__ gfm, int *__restrict__ flag) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Geo &G = *Gp;
  const uint32_t inf = info[i];
  const int lvl = fvk::cell_level(inf);
  double rhs[L::NEQ], T[9];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) rhs[v] = 0.0;
#pragma unroll
  for (int q = 0; q < 9; q++) T[q] = 0.0;
  uint8_t gbits = 0;
  bool ok = true;
#pragma unroll 1
  for (int f = 0; f < 6; f++) {
    const int kind = fvk::face_kind(inf, f);
    const int nb = nbr[(size_t)f * n + i];
    const int fo = FCOPP5[f];
    double uf0 = 0.0, uf1 = 0.0, uf2 = 0.0, Area_f = 0.0;
    if (kind == FV_FK_MIRROR) {
      Rec5<NM, R> sL, sR;
      side_record<NM, R, true>(P, S, SF, n, nb, fo, sL);
      side_record<NM, R, true>(P, S, SF, n, i, f, sR);
      R out[L::NEQ], un[3], fr[9];
#pragma unroll
      for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[fo][q];
      const bool g = sL.phi0 * sR.phi0 < R(0.0);
      ok &= g ? gfm_subface<NM, R, true>(P, fr, sL, sR, out, un)
              : hllc_subface<NM, R, true>(P, fr, 0, sL, sR, S, n, nb, out, un);
      const double A = G.area[lvl];
#pragma unroll
      for (int v = 0; v < L::NEQ; v++) rhs[v] -= A * (double)out[v];
      const double us = P.fix ? 1.0 : -1.0;   // the published code negates the reused vector
      uf0 = us * (double)un[0]; uf1 = us * (double)un[1]; uf2 = us * (double)un[2];
      Area_f = A;
    } else {
      const int Nneig = (kind == FV_FK_FINE) ? 4 : 1;
      const double A = (kind == FV_FK_FINE) ? G.area[lvl + 1] : G.area[lvl];
      Area_f = (double)Nneig * A;
      R fr[9];
#pragma unroll
      for (int q = 0; q < 9; q++) fr[q] = (R)G.frame[f][q];
      Rec5<NM, R> sL;
      side_record<NM, R, true>(P, S, SF, n, i, f, sL);
      for (int k = 0; k < Nneig; ++k) {
        Rec5<NM, R> sR;
        int bcode = 0;
        if (kind == FV_FK_BOUNDARY) {
          bcode = -nb;
          side_record<NM, R, true>(P, S, SF, n, i, f, sR);
          apply_bc5<NM, R>(bcode, fr, sL, sR);
        } else {
          const int j = (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
          side_record<NM, R, true>(P, S, SF, n, j, fo, sR);
        }
        R out[L::NEQ], un[3];
        if (sL.phi0 * sR.phi0 < R(0.0)) {
          gbits |= (uint8_t)(1u << f);
          Rec5<NM, R> gL = sL;
          ok &= gfm_subface<NM, R, true>(P, fr, gL, sR, out, un);
        } else {
          ok &= hllc_subface<NM, R, true>(P, fr, bcode, sL, sR, S, n, i, out, un);
        }
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) rhs[v] += A * (double)out[v];
        uf0 = (double)un[0] / (double)Nneig;   // the last sub-face's velocity (m5_residual.c:163-165)
        uf1 = (double)un[1] / (double)Nneig;
        uf2 = (double)un[2] / (double)Nneig;
      }
    }
    const double nn[3] = {G.frame[f][0], G.frame[f][1], G.frame[f][2]}, uu[3] = {uf0, uf1, uf2};
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) T[3 * a + b] += Area_f * nn[a] * uu[b];
  }
  // the cell record's source coefficients (m5_residual.c:31-59)
  double src_avf[NM - 1], src_st[NM][9], Am[9];
  {
    double x[L::NEQ];
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) x[v] = S[(size_t)v * n + i];
    Rec5<NM> sc;
    record<NM, false, true>(P, x, sc);
#pragma unroll
    for (int k = 0; k < NM - 1; ++k) src_avf[k] = sc.avf[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) Am[k] = sc.A[k];
    double Gm[9], Gmag;
    gmat(sc.A, Gm, Gmag);
    const GPow<> gps = has_solid(P) ? gpow_t<true>(P, Gmag) : GPow<>{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NM; ++k)
      stress_v2<double, true>(src_st[k], sc.avf[k], sc.ph[k], Gm, Gmag, gps, P.mix ? sc.avf[k] * P.mush[k] : P.mush[k]);
  }
  const double tr = T[0] + T[4] + T[8];   // sum_f Area_f (u*.n)
  const double sgn = P.fix ? -1.0 : 1.0;
#pragma unroll
  for (int k = 0; k < NM - 1; ++k) rhs[L::I3 + k] += sgn * src_avf[k] * tr;
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) acc += src_st[k][q] * T[q];   // s[3a+b] n_a u_b
    rhs[L::I4 + k] += sgn * acc;
  }
  if (P.fix) {
    // A_ij: n_i (A_kj u_k) - A_ij (u.n)
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int j = 0; j < 3; j++)
        rhs[L::I5 + 3 * a + j] += Am[j] * T[3 * a + 0] + Am[3 + j] * T[3 * a + 1] + Am[6 + j] * T[3 * a + 2] -
                                  Am[3 * a + j] * tr;
  } else {
    // m5_residual.c:197-209 with T_ab = sum_f Area_f n_a u_b
    const double A1 = Am[0], B1 = Am[3], C1 = Am[6];
    const double A2 = Am[1], B2 = Am[4], C2 = Am[7];
    const double A3 = Am[2], B3 = Am[5], C3 = Am[8];
    const int v5 = L::I5;
    rhs[v5 + 0] += A1 * (T[4] + T[8]) + B1 * T[1] - C1 * T[2];
    rhs[v5 + 1] += A2 * (T[4] + T[8]) + B2 * T[1] - C2 * T[2];
    rhs[v5 + 2] += A3 * (T[4] + T[8]) + B3 * T[1] - C3 * T[2];
    rhs[v5 + 3] += B1 * (T[0] + T[8]) + A1 * T[3] - C1 * T[5];
    rhs[v5 + 4] += B2 * (T[0] + T[8]) + A2 * T[3] - C2 * T[5];
    rhs[v5 + 5] += B3 * (T[0] + T[8]) + A3 * T[3] - C3 * T[5];
    rhs[v5 + 6] += C1 * (T[0] + T[4]) + A1 * T[6] - B1 * T[7];
    rhs[v5 + 7] += C2 * (T[0] + T[4]) + A2 * T[6] - B2 * T[7];
    rhs[v5 + 8] += C3 * (T[0] + T[4]) + A3 * T[6] - B3 * T[7];
  }
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) {
    ok &= !isnan(rhs[v]);
    RHS[(size_t)v * n + i] = rhs[v];
  }
#pragma unroll
  for (int f = 0; f < 6; f++) gfm[6 * i + f] = (gbits >> f) & 1u;
  if (!ok) atomicOr(flag, 1);
}

// m5_relaxation_solid.c:18-24: Sb keeps the state before relaxation for the re-projection trigger
template <int NM>
__global__ __launch_bounds__(256) void k_m5_relax(int n, M5P P, double *__restrict__ S, double *__restrict__ Sb,
                                                  int *__restrict__ flag) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) { x[v] = S[(size_t)v * n + i]; Sb[(size_t)v * n + i] = x[v]; }
  const int rc = relax5<NM>(P, x);
  if (rc) atomicOr(flag, rc == -3 ? 4 : 1);
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) S[(size_t)v * n + i] = x[v];
}

// ---- host side --------------------------------------------------------------

int m5_params(const fv_m5_params *p, M5P &P) {
  if (!p) return fv_set_err(FV_ERR_ARG, "null argument");
  if (p->nmat != 2 && p->nmat != 3) return fv_set_err(FV_ERR_ARG, "model 5: nmat must be 2 or 3");
  if (!(p->dt > 0.0)) return fv_set_err(FV_ERR_ARG, "model 5: dt must be positive");
  P.nmat = p->nmat;
  for (int k = 0; k < 3; k++) {
    P.gamma[k] = k < p->nmat ? p->gamma[k] : 0.0;
    P.pinf[k] = k < p->nmat ? p->pinf[k] : 0.0;
    P.mush[k] = k < p->nmat ? p->mush[k] : 0.0;
    P.zerou[k] = p->zerou[k];
  }
  P.amin = p->amin; P.avf_lim = p->avf_lim; P.dt = p->dt;
  P.prelax = p->prelax; P.order = p->order;
  if (!(p->yield >= 0.0)) return fv_set_err(FV_ERR_ARG, "model 5: the yield stress must be >= 0");
  P.yield = p->yield;
  if (p->fix & ~(FV_M5_FIX_SIGNS | FV_M5_FIX_MIXTURE))
    return fv_set_err(FV_ERR_ARG, "model 5: fix must be a combination of FV_M5_FIX_SIGNS and FV_M5_FIX_MIXTURE");
  P.fix = (p->fix & FV_M5_FIX_SIGNS) ? 1 : 0;
  P.mix = (p->fix & FV_M5_FIX_MIXTURE) ? 1 : 0;
  if (p->limiter < 0 || p->limiter > 2) return fv_set_err(FV_ERR_ARG, "model 5: limiter must be 0, 1 or 2");
  P.limiter = p->limiter;
  P.dtu = P.fix ? -p->dt : p->dt;
  P.sg = P.fix ? -1.0 : 1.0;
  P.fast = 0;
  return FV_OK;
}

// device buffers of one call, freed on every return path
struct Bufs {
  std::vector<void *> ptrs;
  ~Bufs() { for (void *q : ptrs) hipFree(q); }
  template <class T>
  hipError_t alloc(T **p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 1);
    if (e == hipSuccess) ptrs.push_back(q);
    *p = (T *)q;
    return e;
  }
};

// row-major [count][w] host rows -> SoA [w][count] host buffer
void to_soa(const double *rows, int count, int w, std::vector<double> &out) {
  out.resize((size_t)count * w);
  for (int q = 0; q < count; q++)
    for (int v = 0; v < w; v++) out[(size_t)v * count + q] = rows[(size_t)q * w + v];
}
void from_soa(const std::vector<double> &soa, int count, int w, double *rows) {
  for (int q = 0; q < count; q++)
    for (int v = 0; v < w; v++) rows[(size_t)q * w + v] = soa[(size_t)v * count + q];
}

struct MeshDev {
  int *nbr = nullptr, *fine4 = nullptr;
  uint32_t *info = nullptr;
  Geo *G = nullptr;
};

int upload_pack(Bufs &B, const fv_pack *pk, MeshDev &M) {
  if (!pk || pk->n <= 0 || !pk->nbr || !pk->info) return fv_set_err(FV_ERR_ARG, "model 5: empty mesh");
  const size_t n = (size_t)pk->n, nf = (size_t)(pk->nfine > 0 ? pk->nfine : 1);
  Geo G;
  for (int f = 0; f < 6; f++)
    for (int k = 0; k < 9; k++) G.frame[f][k] = pk->geo.frame[f][k];
  for (int l = 0; l < FV_MAXLEVEL + 2; l++) {
    G.h[l] = pk->geo.h[l]; G.area[l] = pk->geo.area[l]; G.vol[l] = pk->geo.vol[l];
    for (int k = 0; k < 3; k++) { G.dcf[k][l] = pk->geo.dcf[k][l]; G.dcc[k][l] = pk->geo.dcc[k][l]; }
  }
  HIPCHK(B.alloc(&M.nbr, 6 * n));
  HIPCHK(B.alloc(&M.info, n));
  HIPCHK(B.alloc(&M.fine4, 4 * nf));
  HIPCHK(B.alloc(&M.G, 1));
  HIPCHK(hipMemcpy(M.nbr, pk->nbr, sizeof(int) * 6 * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(M.info, pk->info, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
  if (pk->nfine > 0) HIPCHK(hipMemcpy(M.fine4, pk->fine4, sizeof(int) * 4 * (size_t)pk->nfine, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(M.G, &G, sizeof(Geo), hipMemcpyHostToDevice));
  return FV_OK;
}

// leaves with mark[i] != 0 in leaf order; returns their count in *count and on the device in *dcount
int compact_marked(Bufs &B, int n, const uint8_t *mark, int **list, int *count, int **dcount) {
  int *d_list = nullptr, *d_num = nullptr;
  HIPCHK(B.alloc(&d_list, (size_t)n));
  HIPCHK(B.alloc(&d_num, 1));
  hipcub::CountingInputIterator<int> idx(0);
  size_t tmp_bytes = 0;
  HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, idx, mark, d_list, d_num, n));
  void *tmp = nullptr;
  HIPCHK(B.alloc((char **)&tmp, tmp_bytes));
  HIPCHK(hipcub::DeviceSelect::Flagged(tmp, tmp_bytes, idx, mark, d_list, d_num, n));
  HIPCHK(hipMemcpy(count, d_num, sizeof(int), hipMemcpyDeviceToHost));
  *list = d_list;
  *dcount = d_num;
  return FV_OK;
}

template <int NM>
int gfm_faces_t(const M5P &P, int nf, const double *UL, const double *UR, const double *frame, const double *area,
                double *flux, double *unc) {
  constexpr int NEQ = L5<NM>::NEQ;
  Bufs B;
  std::vector<double> hL, hR, hF, hflux, hunc;
  to_soa(UL, nf, NEQ, hL);
  to_soa(UR, nf, NEQ, hR);
  to_soa(frame, nf, 9, hF);
  double *dL, *dR, *dF, *dA, *dflux, *dunc;
  int *dnan;
  HIPCHK(B.alloc(&dL, hL.size())); HIPCHK(B.alloc(&dR, hR.size())); HIPCHK(B.alloc(&dF, hF.size()));
  HIPCHK(B.alloc(&dA, (size_t)nf)); HIPCHK(B.alloc(&dflux, (size_t)nf * NEQ)); HIPCHK(B.alloc(&dunc, (size_t)nf * 3));
  HIPCHK(B.alloc(&dnan, 1));
  HIPCHK(hipMemcpy(dL, hL.data(), sizeof(double) * hL.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dR, hR.data(), sizeof(double) * hR.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dF, hF.data(), sizeof(double) * hF.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dA, area, sizeof(double) * nf, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(dnan, 0, sizeof(int)));
  hipLaunchKernelGGL(k_m5_gfm_faces<NM>, dim3((nf + 255) / 256), dim3(256), 0, 0, nf, P, dL, dR, dF, dA, dflux, dunc,
                     dnan);
  HIPCHK(hipGetLastError());
  int nan = 0;
  HIPCHK(hipMemcpy(&nan, dnan, sizeof(int), hipMemcpyDeviceToHost));
  if (nan) return fv_set_err(FV_ERR_NAN, "model 5: NaN in a GFM face flux (m5_hllc_gfm.c:92-97)");
  hflux.resize((size_t)nf * NEQ);
  hunc.resize((size_t)nf * 3);
  HIPCHK(hipMemcpy(hflux.data(), dflux, sizeof(double) * hflux.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hunc.data(), dunc, sizeof(double) * hunc.size(), hipMemcpyDeviceToHost));
  from_soa(hflux, nf, NEQ, flux);
  from_soa(hunc, nf, 3, unc);
  return FV_OK;
}

template <int NM>
int rk_t(const M5P &P, const fv_pack *pk, int irk, double *SN, double *S, const double *RHS, const uint8_t *gfm) {
  constexpr int NEQ = L5<NM>::NEQ;
  const int n = pk->n;
  Bufs B;
  MeshDev M;
  int rc = upload_pack(B, pk, M);
  if (rc != FV_OK) return rc;
  std::vector<double> hSN, hS, hR;
  to_soa(SN, n, NEQ, hSN);
  to_soa(S, n, NEQ, hS);
  to_soa(RHS, n, NEQ, hR);
  double *dSN, *dSin, *dS, *dR;
  uint8_t *dgf, *dmark;
  int *derr;
  const size_t sz = (size_t)n * NEQ;
  HIPCHK(B.alloc(&dSN, sz)); HIPCHK(B.alloc(&dSin, sz)); HIPCHK(B.alloc(&dS, sz)); HIPCHK(B.alloc(&dR, sz));
  HIPCHK(B.alloc(&dgf, (size_t)n * 6)); HIPCHK(B.alloc(&dmark, (size_t)n)); HIPCHK(B.alloc(&derr, 1));
  HIPCHK(hipMemcpy(dSN, hSN.data(), sizeof(double) * sz, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dSin, hS.data(), sizeof(double) * sz, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dR, hR.data(), sizeof(double) * sz, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dgf, gfm, (size_t)n * 6, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(derr, 0, sizeof(int)));
  hipLaunchKernelGGL(k_m5_rk<NM>, dim3((n + 255) / 256), dim3(256), 0, 0, n, P, irk, M.G, M.info, dSN, dSin, dS, dR,
                     dmark);
  HIPCHK(hipGetLastError());
  if (irk == 1) {
    int *list = nullptr, count = 0, *dcount = nullptr;
    if ((rc = compact_marked(B, n, dmark, &list, &count, &dcount)) != FV_OK) return rc;
    if (count > 0) {
      hipLaunchKernelGGL(k_m5_gfm_seq<NM>, dim3(1), dim3(64), 0, 0, n, dcount, list, P, M.G, M.nbr, M.info, M.fine4,
                         dgf, dSN, dSin, dS, dR, derr);
      HIPCHK(hipGetLastError());
      int err = 0;
      HIPCHK(hipMemcpy(&err, derr, sizeof(int), hipMemcpyDeviceToHost));
      if (err) return fv_set_err(FV_ERR_ARG, "model 5: gfm_face set on a boundary face");
    }
  }
  HIPCHK(hipMemcpy(hS.data(), dS, sizeof(double) * sz, hipMemcpyDeviceToHost));
  from_soa(hS, n, NEQ, S);
  if (irk == 0) {
    HIPCHK(hipMemcpy(hSN.data(), dSN, sizeof(double) * sz, hipMemcpyDeviceToHost));
    from_soa(hSN, n, NEQ, SN);
  }
  return FV_OK;
}

template <int NM>
int rel_t(const M5P &P, const fv_pack *pk, const double *Sb, double *S, const uint8_t *gfm) {
  constexpr int NEQ = L5<NM>::NEQ;
  const int n = pk->n;
  Bufs B;
  MeshDev M;
  int rc = upload_pack(B, pk, M);
  if (rc != FV_OK) return rc;
  std::vector<double> hSb, hS;
  to_soa(Sb, n, NEQ, hSb);
  to_soa(S, n, NEQ, hS);
  double *dSb, *dS;
  uint8_t *dgf, *dmark;
  int *derr;
  const size_t sz = (size_t)n * NEQ;
  HIPCHK(B.alloc(&dSb, sz)); HIPCHK(B.alloc(&dS, sz));
  HIPCHK(B.alloc(&dgf, (size_t)n * 6)); HIPCHK(B.alloc(&dmark, (size_t)n)); HIPCHK(B.alloc(&derr, 1));
  HIPCHK(hipMemcpy(dSb, hSb.data(), sizeof(double) * sz, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dS, hS.data(), sizeof(double) * sz, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dgf, gfm, (size_t)n * 6, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(derr, 0, sizeof(int)));
  hipLaunchKernelGGL(k_m5_rel_mark<NM>, dim3((n + 255) / 256), dim3(256), 0, 0, n, P, dSb, dS, dmark);
  HIPCHK(hipGetLastError());
  int *list = nullptr, count = 0, *dcount = nullptr;
  if ((rc = compact_marked(B, n, dmark, &list, &count, &dcount)) != FV_OK) return rc;
  if (count > 0) {
    hipLaunchKernelGGL(k_m5_rel_seq<NM>, dim3(1), dim3(64), 0, 0, n, dcount, list, P, M.G, M.nbr, M.info, M.fine4,
                       dgf, dSb, dS, derr);
    HIPCHK(hipGetLastError());
    int err = 0;
    HIPCHK(hipMemcpy(&err, derr, sizeof(int), hipMemcpyDeviceToHost));
    if (err) return fv_set_err(FV_ERR_ARG, "model 5: gfm_face set on a boundary face");
  }
  HIPCHK(hipMemcpy(hS.data(), dS, sizeof(double) * sz, hipMemcpyDeviceToHost));
  from_soa(hS, n, NEQ, S);
  return FV_OK;
}


// ---- device-resident model-5 stage (fv_m5_dev_*) -----------------------------

template <class T>
static hipError_t dalloc(T **p, size_t count) {
  return hipMalloc((void **)p, (count > 0 ? count : 1) * sizeof(T));
}

}  // namespace

struct fv_m5_dev {
  M5P P{};
  int n = 0, nfine = 0, neq = 0, rk_steps = 3;
  // FV_PREC_EXACT: the reference's face-by-face source terms (bit for bit where
  // no pow runs); FV_PREC_FP64: the sources formed once per leaf; FV_PREC_MIXED:
  // that with FP32 face records, HLLC and GFM flux (fv_m5_dev_set_precision)
  int prec = FV_PREC_FP64;
  hipStream_t stream = nullptr;
  Geo *G = nullptr;
  int *nbr = nullptr, *fine4 = nullptr;
  uint32_t *info = nullptr;
  double *S = nullptr, *S2 = nullptr, *SN = nullptr, *SF = nullptr, *RHS = nullptr, *Sb = nullptr;
  uint8_t *gfm = nullptr, *mark = nullptr;
  int *flag = nullptr;        // [0] NaN (1) / Newton failure (4) bits, [1] GFM flag on a boundary face
  int *list = nullptr, *count = nullptr;   // count [3]: compacted leaves, the two PRELAX 3 lists
  void *cub_tmp = nullptr;
  size_t cub_bytes = 0;
  // PRELAX 3 over the 6n face states (and the n cell states): saved Newton state of the slow ones
  double *nst = nullptr;      // [6n][3] pr, F, dF
  int *iters = nullptr;       // [6n]
  int *list6 = nullptr;       // [6n] states still iterating after the first pass (count[1] of them)
  int *list6b = nullptr;      // [6n] those still iterating after the second (count[2])
  double *part = nullptr;     // [n/256 + 1] per-block minima of the CFL step, then the step
  uint8_t *lim = nullptr;     // [n] per-leaf slope flags of limiter 2 (0 minmod, 1 MC)
  int cap = 0, cap_fine = 0;  // leaves and 1:4 rows the per-leaf buffers hold
  // mesh side of regrids and interface_location (fv_m5_dev_set_forest): a
  // topology-only model-4 context over the same leaves
  fv_dev *topo = nullptr;
  double *Snew = nullptr;     // [NEQ][nn] remapped state of a regrid, until installed
  size_t cap_new = 0;
};

namespace {

void m5_free(fv_m5_dev *c) {
  void *ptrs[] = {c->G, c->nbr, c->fine4, c->info, c->S, c->S2, c->SN, c->SF, c->RHS, c->Sb,
                  c->gfm, c->mark, c->flag, c->list, c->count, c->cub_tmp, c->nst, c->iters, c->list6b, c->list6,
                  c->part, c->Snew, c->lim};
  for (void *q : ptrs)
    if (q) hipFree(q);
  if (c->topo) fv_dev_detach(c->topo);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

// items with mark != 0, in order, into list, their number into c->count (device side only)
int m5_compact(fv_m5_dev *c, const uint8_t *mark, int *list, int items) {
  hipcub::CountingInputIterator<int> idx(0);
  size_t bytes = c->cub_bytes;
  HIPCHK(hipcub::DeviceSelect::Flagged(c->cub_tmp, bytes, idx, mark, list, c->count, items, c->stream));
  return FV_OK;
}

// PRELAX 3 on `count` states of X (see k_m5_relax3_start)
// (started: the first pass ran inside the producing kernel)
template <int NM>
int m5_relax3(fv_m5_dev *c, double *X, long count, double *Sb, bool started = false) {
  const dim3 g((unsigned)((count + 255) / 256)), b(256);
  if (!started) {
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, c->stream, c->count + 1, 2);
    hipLaunchKernelGGL(k_m5_relax3_start<NM>, g, b, 0, c->stream, count, c->n, c->P, X, Sb, c->nst, c->iters,
                       c->count + 1, c->list6, c->flag);
    HIPCHK(hipGetLastError());
  }
  // the states still iterating: 24 more iterations for all of them, then the
  // few that stagnate run to the reference's limit of 500 in full waves
  const dim3 gr((unsigned)std::min<long>(4096, (count + 255) / 256));
  hipLaunchKernelGGL(k_m5_relax3_resume<NM>, gr, b, 0, c->stream, c->count + 1, c->list6, c->n, c->P, X, c->nst,
                     c->iters, c->count + 2, c->list6b, 24, c->flag);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_m5_relax3_resume<NM>, gr, b, 0, c->stream, c->count + 2, c->list6b, c->n, c->P, X, c->nst,
                     c->iters, (int *)nullptr, (int *)nullptr, 1 << 30, c->flag);
  HIPCHK(hipGetLastError());
  return FV_OK;
}

dim3 grid_of(int n) { return dim3((unsigned)((n + 255) / 256)); }

template <int NM, typename R, bool RX3>
void m5_recon_launch(fv_m5_dev *c) {
  if (RX3) hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, c->stream, c->count + 1, 2);
  if (c->P.fast)
    hipLaunchKernelGGL((k_m5_recon<NM, R, RX3, true>), grid_of(c->n), dim3(256), 0, c->stream, c->n, c->P, c->G, c->S,
                       c->nbr, c->info, c->fine4, c->lim, c->SF, c->nst, c->iters, c->count + 1, c->list6, c->flag);
  else
    hipLaunchKernelGGL((k_m5_recon<NM, R, RX3, false>), grid_of(c->n), dim3(256), 0, c->stream, c->n, c->P, c->G, c->S,
                       c->nbr, c->info, c->fine4, c->lim, c->SF, c->nst, c->iters, c->count + 1, c->list6, c->flag);
}

template <int NM>
int m5_faceinterp_t(fv_m5_dev *c) {
  const bool rx3 = c->P.prelax == 3;
  if (c->prec == FV_PREC_MIXED)
    rx3 ? m5_recon_launch<NM, float, true>(c) : m5_recon_launch<NM, float, false>(c);
  else
    rx3 ? m5_recon_launch<NM, double, true>(c) : m5_recon_launch<NM, double, false>(c);
  HIPCHK(hipGetLastError());
  if (rx3) return m5_relax3<NM>(c, c->SF, 6L * c->n, nullptr, true);
  return FV_OK;
}

template <int NM>
int m5_residual_t(fv_m5_dev *c) {
  if (c->prec == FV_PREC_MIXED)
    hipLaunchKernelGGL((k_m5_residual_mx<NM, float>), grid_of(c->n), dim3(256), 0, c->stream, c->n, c->P, c->G, c->S,
                       c->SF, c->nbr, c->info, c->fine4, c->RHS, c->gfm, c->flag);
  else if (c->prec == FV_PREC_FP64)
    hipLaunchKernelGGL((k_m5_residual_mx<NM, double>), grid_of(c->n), dim3(256), 0, c->stream, c->n, c->P, c->G, c->S,
                       c->SF, c->nbr, c->info, c->fine4, c->RHS, c->gfm, c->flag);
  else
    hipLaunchKernelGGL((k_m5_residual<NM, double>), grid_of(c->n), dim3(256), 0, c->stream, c->n, c->P, c->G, c->S,
                       c->SF, c->nbr, c->info, c->fine4, c->RHS, c->gfm, c->flag);
  HIPCHK(hipGetLastError());
  return FV_OK;
}

// one stage irk (advancetimeexplicit.c:11-24 with MODEL 5), no host synchronisation;
// ev brackets the residual kernel when given
template <int NM>
int m5_stage_t(fv_m5_dev *c, int irk, hipEvent_t *ev) {
  int rc;
  const dim3 g = grid_of(c->n), b(256);
  if (c->P.order == 2 && (rc = m5_faceinterp_t<NM>(c)) != FV_OK) return rc;
  if (ev) HIPCHK(hipEventRecord(ev[0], c->stream));
  if ((rc = m5_residual_t<NM>(c)) != FV_OK) return rc;
  if (ev) HIPCHK(hipEventRecord(ev[1], c->stream));
  // m5_RK(irk): S2 from S, SN and RHS; stage 1 re-projects the crossing leaves in leaf order
  hipLaunchKernelGGL(k_m5_rk<NM>, g, b, 0, c->stream, c->n, c->P, irk, c->G, c->info, c->SN, c->S, c->S2, c->RHS,
                     c->mark);
  HIPCHK(hipGetLastError());
  if (irk == 1) {
    if ((rc = m5_compact(c, c->mark, c->list, c->n)) != FV_OK) return rc;
    hipLaunchKernelGGL(k_m5_gfm_seq<NM>, dim3(1), dim3(64), 0, c->stream, c->n, c->count, c->list, c->P, c->G,
                       c->nbr, c->info, c->fine4, c->gfm, c->SN, c->S, c->S2, c->RHS, c->flag + 1);
    HIPCHK(hipGetLastError());
  }
  std::swap(c->S, c->S2);
  // m5_relaxation_solid: the plastic return map (new), relax every leaf, then
  // re-project the leaves whose solid fraction crossed.  The relaxation saves
  // the state it starts from in Sb; the re-projection reads only the fractions
  // and velocities of Sb, which the map does not change.
  if (c->P.yield > 0.0) {
    hipLaunchKernelGGL(k_m5_plastic<NM>, g, b, 0, c->stream, c->n, c->P, c->S, c->flag);
    HIPCHK(hipGetLastError());
  }
  if (c->P.prelax == 3) {
    if ((rc = m5_relax3<NM>(c, c->S, c->n, c->Sb)) != FV_OK) return rc;
  } else {
    hipLaunchKernelGGL(k_m5_relax<NM>, g, b, 0, c->stream, c->n, c->P, c->S, c->Sb, c->flag);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_m5_rel_mark<NM>, g, b, 0, c->stream, c->n, c->P, c->Sb, c->S, c->mark);
  HIPCHK(hipGetLastError());
  if ((rc = m5_compact(c, c->mark, c->list, c->n)) != FV_OK) return rc;
  hipLaunchKernelGGL(k_m5_rel_seq<NM>, dim3(1), dim3(64), 0, c->stream, c->n, c->count, c->list, c->P, c->G, c->nbr,
                     c->info, c->fine4, c->gfm, c->Sb, c->S, c->flag + 1);
  HIPCHK(hipGetLastError());
  return FV_OK;
}

int m5_stage(fv_m5_dev *c, int irk, hipEvent_t *ev) {
  return c->P.nmat == 2 ? m5_stage_t<2>(c, irk, ev) : m5_stage_t<3>(c, irk, ev);
}

// the flags the kernels raised since the last check, as a status (cleared)
int m5_check(fv_m5_dev *c) {
  int h[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(h, c->flag, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (h[0] || h[1]) HIPCHK(hipMemsetAsync(c->flag, 0, 2 * sizeof(int), c->stream));
  if (h[1]) return fv_set_err(FV_ERR_ARG, "model 5: a GFM face flag on a boundary face");
  if (h[0] & 4) return fv_set_err(FV_ERR_NAN, "model 5: PRELAX 2 Newton failed (m5_relaxationvec.c:244-272)");
  if (h[0]) return fv_set_err(FV_ERR_NAN, "model 5: NaN in a face state, flux or relaxation");
  return FV_OK;
}


// ---- mesh adaptation of the model-5 state ----------------------------------
// The forest side (criterionsmth, merge/split plan, connectivity) runs in the
// context's topology-only model-4 context; this side gives it the model-5
// indicator and remaps the model-5 state.

// interface_location.c:7-50 with m5_facevalues (m5_facevalues.c:82-98,
// n_solids = 0): SI[0] for the leaves with a face across which the solid's
// phi[0] = clipped alpha_0 - AVF_LIM changes sign, SI[1] for the other fractions
template <int NM>
__global__ __launch_bounds__(256) void k_m5_si_first(int n, M5P P, const double *__restrict__ S,
                                                     const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                     const int *__restrict__ fine4, int *__restrict__ SI) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = P.amin / 100.0;
  const uint32_t inf = info[i];
  auto phi = [&](int c, int v) {
    double x = S[(size_t)(L::I3 + v) * n + c];
    if (x >= (1.0 - a)) x = 1.0 - a;
    else if (x <= a) x = a;
    return x - P.avf_lim;
  };
  int si0 = 0, si1 = 0, flag = 0;
  for (int f = 0; f < 6 && !flag; f++) {
    const int kind = fvk::face_kind(inf, f);
    if (kind == FV_FK_BOUNDARY) continue;
    const int nb = nbr[(size_t)f * n + i];
    const int ns = (kind == FV_FK_FINE) ? 4 : 1;
    for (int k = 0; k < ns; k++) {
      const int j = (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
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

struct M5Crit {
  int uniform_on, uniform_lvl, prs_on, prs_lvl, rho_on, rho_lvl, vel_on, vel_lvl;
  int inter0_on, inter0_lvl, inter0_lrs, inter1_on, inter1_lvl, inter1_lrs;
  double prs_thr, rho_thr, vel_thr;
};

__device__ __forceinline__ double m5_rel_jump(double x_cur, double x_nb) {   // criterionsplt.c:81-88
  const double eps = 1.0e-8;
  if ((fabs(x_nb) > eps) || (fabs(x_cur) > eps)) return fabs(x_nb - x_cur) / maxr(fabs(x_nb), fabs(x_cur));
  return 0.0;
}

// criterionsplt.c:3-37 with criterion_grad (:40-140) on m5_facevalues cell
// records: the mixture pressure sum_k alpha_k p_hydro_k, the density, and the
// velocity criterion that reads the density jump (:113); then the interface bands
template <int NM>
__global__ __launch_bounds__(256) void k_m5_split(int n, M5P P, M5Crit C, const double *__restrict__ S,
                                                  const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                  const int *__restrict__ fine4, const int *__restrict__ SI,
                                                  int *__restrict__ split) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t inf = info[i];
  int lt = 1;
  if (C.uniform_on == 1) lt = max(lt, C.uniform_lvl);
  if (C.prs_on == 1 || C.rho_on == 1 || C.vel_on == 1) {
    double x[L::NEQ];
#pragma unroll
    for (int v = 0; v < L::NEQ; v++) x[v] = S[(size_t)v * n + i];
    Rec5<NM> sL;
    record<NM>(P, x, sL);
    double pl = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) pl += sL.avf[k] * sL.ph[k];
    double gp = 0.0, gr = 0.0, gv = 0.0, grho = 0.0, lg = 1;
    for (int f = 0; f < 6; f++) {
      const int kind = fvk::face_kind(inf, f);
      if (kind == FV_FK_BOUNDARY) continue;
      const int nb = nbr[(size_t)f * n + i];
      const int ns = (kind == FV_FK_FINE) ? 4 : 1;
      for (int k = 0; k < ns; k++) {
        const int j = (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
#pragma unroll
        for (int v = 0; v < L::NEQ; v++) x[v] = S[(size_t)v * n + j];
        Rec5<NM> sR;
        record<NM>(P, x, sR);
        if (C.prs_on == 1) {
          double pr = 0.0;
#pragma unroll
          for (int q = 0; q < NM; ++q) pr += sR.avf[q] * sR.ph[q];
          gp = maxr(m5_rel_jump(pl, pr), gp);
        }
        if (C.rho_on == 1) { grho = m5_rel_jump(sL.r, sR.r); gr = maxr(grho, gr); }
        if (C.vel_on == 1) gv = maxr(grho, gv);
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

// the state of the new leaves (merge.c:20-59 with m5_relaxationvec; split.c:425-434):
// a merged parent is the mean of its eight kids in kid order, relaxed; a split
// leaf's kids copy it; a kept leaf moves to its new position
template <int NM>
__global__ __launch_bounds__(256) void k_m5_adapt_state(int n, int nn, M5P P, const uint8_t *__restrict__ kind,
                                                        const int *__restrict__ off, const int *__restrict__ ci,
                                                        const double *__restrict__ S, double *__restrict__ S2,
                                                        int *__restrict__ flag) {
  using L = L5<NM>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = off[i];
  if (kind[i] == 1) {
    const int kid = (ci[3 * i] & 1) + 2 * (ci[3 * i + 1] & 1) + 4 * (ci[3 * i + 2] & 1);
    if (kid != 0) return;   // the first kid writes the parent
    double d[L::NEQ];
#pragma unroll
    for (int iv = 0; iv < L::NEQ; iv++) {
      double s = 0.0;
      for (int k = 0; k < 8; k++) s += S[(size_t)iv * n + i + k] / ((double)8);
      d[iv] = s;
    }
    const int rc = relax5<NM>(P, d);
    if (rc == -3) atomicOr(flag, 4);
    else if (rc != 0) atomicOr(flag, 1);
#pragma unroll
    for (int iv = 0; iv < L::NEQ; iv++) S2[(size_t)iv * nn + o] = d[iv];
  } else {
    const int cnt = (kind[i] == 2) ? 8 : 1;
    for (int k = 0; k < cnt; k++)
#pragma unroll
      for (int iv = 0; iv < L::NEQ; iv++) S2[(size_t)iv * nn + o + k] = S[(size_t)iv * n + i];
  }
}

// fields of the learned indicator (the model-4 ones, fv_mlp.hip, with the
// liquid as the fraction): density, mixture pressure sum_k alpha_k p_hydro_k,
// alpha of material 1, |u|, and c^2 = sum_k c2Y_k + (4/3) sum_k alpha_k mu_k / rho
template <int NM>
__device__ void m5_cell_fields(const M5P &P, const double *__restrict__ S, int n, int j, double &rho, double &p,
                               double &al, double &um, double &c) {
  using L = L5<NM>;
  double x[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) x[v] = S[(size_t)v * n + j];
  Rec5<NM> r;
  record<NM>(P, x, r);
  double pm = 0.0, c2 = 0.0, mu = 0.0;
#pragma unroll
  for (int k = 0; k < NM; k++) { pm += r.avf[k] * r.ph[k]; c2 += r.c2Y[k]; mu += r.avf[k] * P.mush[k]; }
  if (!P.mix) c2 += (4.0 / 3.0) * mu / r.r;
  rho = r.r;
  p = pm;
  al = r.avf[1];
  um = sqrt(r.u * r.u + r.v * r.v + r.w * r.w);
  c = sqrt(c2);
}

template <int NM>
__global__ __launch_bounds__(256) void k_m5_mlp_fields(int n, M5P P, const double *__restrict__ S,
                                                       float4 *__restrict__ fld, float *__restrict__ snd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double rho, p, al, um, c;
  m5_cell_fields<NM>(P, S, n, i, rho, p, al, um, c);
  fld[i] = make_float4((float)rho, (float)p, (float)al, (float)um);
  snd[i] = (float)c;
}

template <int NM>
__global__ __launch_bounds__(256) void k_m5_mlp_labels(int n, M5P P, const double *__restrict__ S,
                                                       const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                       const int *__restrict__ fine4, const int *__restrict__ split,
                                                       float thr, int *__restrict__ cls, float *__restrict__ smooth,
                                                       long row0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t inf = info[i];
  const int lvl = fvk::cell_level(inf);
  cls[row0 + i] = split[i] > lvl ? 0 : (split[i] < lvl ? 2 : 1);
  double rho, p0, al, um, c;
  m5_cell_fields<NM>(P, S, n, i, rho, p0, al, um, c);
  double jmax = 0.0;
  for (int f = 0; f < 6; f++) {
    const int kind = fvk::face_kind(inf, f);
    if (kind == FV_FK_BOUNDARY) continue;
    const int nb = nbr[(size_t)f * n + i];
    const int ns = (kind == FV_FK_FINE) ? 4 : 1;
    for (int k = 0; k < ns; k++) {
      const int j = (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
      double pj;
      m5_cell_fields<NM>(P, S, n, j, rho, pj, al, um, c);
      jmax = fmax(jmax, fabs(pj - p0) / fmax(fmax(fabs(pj), fabs(p0)), 1e-30));
    }
  }
  smooth[row0 + i] = (jmax < (double)thr) ? 1.f : 0.f;
}

M5Crit m5_crit(const fv_crit *cr) {
  M5Crit C;
  C.uniform_on = cr->uniform_on; C.uniform_lvl = cr->uniform_lvl;
  C.prs_on = cr->prs_on; C.prs_lvl = cr->prs_lvl; C.prs_thr = cr->prs_thr;
  C.rho_on = cr->rho_on; C.rho_lvl = cr->rho_lvl; C.rho_thr = cr->rho_thr;
  C.vel_on = cr->vel_on; C.vel_lvl = cr->vel_lvl; C.vel_thr = cr->vel_thr;
  C.inter0_on = cr->inter0_on; C.inter0_lvl = cr->inter0_lvl; C.inter0_lrs = cr->inter0_lrs;
  C.inter1_on = cr->inter1_on; C.inter1_lvl = cr->inter1_lvl; C.inter1_lrs = cr->inter1_lrs;
  return C;
}

// per-leaf buffers for n leaves and nfine 1:4 rows: a new set, swapped in only
// when every allocation succeeded (contents are not kept)
int m5_reserve(fv_m5_dev *c, int n_leaves, int nfine, int prelax3) {
  if (n_leaves <= c->cap && nfine <= c->cap_fine) return FV_OK;
  const size_t n = (size_t)n_leaves, nf = (size_t)(nfine > 0 ? nfine : 1), sz = n * c->neq;
  fv_m5_dev t;
  std::vector<void *> got;
  hipError_t e = hipSuccess;
  auto take = [&](hipError_t x, void *p) { if (e == hipSuccess) e = x; if (x == hipSuccess) got.push_back(p); };
  take(dalloc(&t.nbr, 6 * n), t.nbr); take(dalloc(&t.fine4, 4 * nf), t.fine4); take(dalloc(&t.info, n), t.info);
  take(dalloc(&t.S, sz), t.S); take(dalloc(&t.S2, sz), t.S2); take(dalloc(&t.SN, sz), t.SN);
  take(dalloc(&t.Sb, sz), t.Sb); take(dalloc(&t.RHS, sz), t.RHS); take(dalloc(&t.SF, 6 * sz), t.SF);
  take(dalloc(&t.gfm, 6 * n), t.gfm); take(dalloc(&t.mark, n), t.mark); take(dalloc(&t.list, n), t.list);
  take(dalloc(&t.part, (n + 255) / 256 + 1), t.part);
  take(dalloc(&t.lim, n), t.lim);
  if (prelax3) {
    take(dalloc(&t.nst, 18 * n), t.nst); take(dalloc(&t.iters, 6 * n), t.iters);
    take(dalloc(&t.list6b, 6 * n), t.list6b); take(dalloc(&t.list6, 6 * n), t.list6);
  }
  size_t cub = 0;
  if (e == hipSuccess) {
    hipcub::CountingInputIterator<int> idx(0);
    e = hipcub::DeviceSelect::Flagged(nullptr, cub, idx, t.mark, t.list, c->count, n_leaves, c->stream);
    if (e == hipSuccess) take(dalloc((char **)&t.cub_tmp, cub), t.cub_tmp);
  }
  if (e != hipSuccess) {
    for (void *q : got) hipFree(q);
    return fv_set_err(e == hipErrorOutOfMemory ? FV_ERR_NOMEM : FV_ERR_HIP, hipGetErrorString(e));
  }
  void *old[] = {c->nbr, c->fine4, c->info, c->S, c->S2, c->SN, c->Sb, c->RHS, c->SF, c->gfm, c->mark, c->list,
                 c->part, c->lim, c->nst, c->iters, c->list6b, c->list6, c->cub_tmp};
  for (void *q : old)
    if (q) hipFree(q);
  c->nbr = t.nbr; c->fine4 = t.fine4; c->info = t.info; c->S = t.S; c->S2 = t.S2; c->SN = t.SN; c->Sb = t.Sb;
  c->RHS = t.RHS; c->SF = t.SF; c->gfm = t.gfm; c->mark = t.mark; c->list = t.list; c->part = t.part; c->lim = t.lim;
  c->nst = t.nst; c->iters = t.iters; c->list6b = t.list6b; c->list6 = t.list6; c->cub_tmp = t.cub_tmp;
  c->cub_bytes = cub;
  t = fv_m5_dev();   // its pointers now belong to c
  c->cap = n_leaves;
  c->cap_fine = (int)nf;
  return FV_OK;
}

int m5_remap_hook(void *ctx, hipStream_t st, int n, int nn, const uint8_t *kind, const int *off, const int *ci) {
  fv_m5_dev *c = (fv_m5_dev *)ctx;
  const size_t need = (size_t)nn * c->neq;
  if (need > c->cap_new) {
    if (c->Snew) hipFree(c->Snew);
    c->Snew = nullptr;
    c->cap_new = 0;
    HIPCHK(dalloc(&c->Snew, need + need / 4));
    c->cap_new = need + need / 4;
  }
  if (c->P.nmat == 2)
    hipLaunchKernelGGL(k_m5_adapt_state<2>, dim3((n + 255) / 256), dim3(256), 0, st, n, nn, c->P, kind, off, ci, c->S,
                       c->Snew, c->flag);
  else
    hipLaunchKernelGGL(k_m5_adapt_state<3>, dim3((n + 255) / 256), dim3(256), 0, st, n, nn, c->P, kind, off, ci, c->S,
                       c->Snew, c->flag);
  HIPCHK(hipGetLastError());
  return FV_OK;
}

int m5_install_hook(void *ctx, fv_dev *t) {
  fv_m5_dev *c = (fv_m5_dev *)ctx;
  const int nn = t->n, nfine = t->nfine;
  // headroom, so that a slowly growing mesh does not reallocate at every regrid
  if (nn > c->cap || nfine > c->cap_fine) {
    const int rc = m5_reserve(c, nn + nn / 4, nfine + nfine / 4, c->P.prelax == 3);
    if (rc != FV_OK) return rc;
  }
  c->n = nn;
  c->nfine = nfine;
  const size_t n = (size_t)nn;
  HIPCHK(hipMemcpyAsync(c->S, c->Snew, sizeof(double) * n * c->neq, hipMemcpyDeviceToDevice, t->stream));
  HIPCHK(hipMemcpyAsync(c->nbr, t->nbr, sizeof(int) * 6 * n, hipMemcpyDeviceToDevice, t->stream));
  HIPCHK(hipMemcpyAsync(c->info, t->info, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, t->stream));
  if (nfine > 0)
    HIPCHK(hipMemcpyAsync(c->fine4, t->fine4, sizeof(int) * 4 * (size_t)nfine, hipMemcpyDeviceToDevice, t->stream));
  HIPCHK(hipMemsetAsync(c->lim, 0, n, t->stream));   // minmod until the flags are set for the new leaves
  HIPCHK(hipStreamSynchronize(t->stream));
  return FV_OK;
}

// the model-5 interface flags of the current state into the topology context
// (first layer here, the layers there); *SIdev: the device buffer holding them
int m5_interface_layers(fv_m5_dev *c, const fv_crit *cr, int **SIdev) {
  fv_dev *t = c->topo;
  HIPCHK(hipStreamSynchronize(c->stream));
  const dim3 g((c->n + 255) / 256), b(256);
  if (c->P.nmat == 2)
    hipLaunchKernelGGL(k_m5_si_first<2>, g, b, 0, t->stream, c->n, c->P, c->S, c->nbr, c->info, c->fine4, t->SIa);
  else
    hipLaunchKernelGGL(k_m5_si_first<3>, g, b, 0, t->stream, c->n, c->P, c->S, c->nbr, c->info, c->fine4, t->SIa);
  HIPCHK(hipGetLastError());
  return fv_dev_si_layers(t, cr, SIdev);
}

}  // namespace

extern "C" int fv_m5_dev_attach(const fv_m5_params *p, const fv_pack *pk, fv_m5_dev **out) {
  if (!out) return fv_set_err(FV_ERR_ARG, "null argument");
  *out = nullptr;
  M5P P;
  int rc = m5_params(p, P);
  if (rc != FV_OK) return rc;
  if (p->prelax < 0 || p->prelax > 3) return fv_set_err(FV_ERR_ARG, "model 5: prelax must be 0..3");
  if (p->order != 1 && p->order != 2) return fv_set_err(FV_ERR_ARG, "model 5: order must be 1 or 2");
  if (p->rk_steps < 1 || p->rk_steps > 3) return fv_set_err(FV_ERR_ARG, "model 5: rk_steps must be 1..3");
  if (!pk || pk->n <= 0 || !pk->nbr || !pk->info) return fv_set_err(FV_ERR_ARG, "model 5: empty mesh");
  fv_m5_dev *c = new fv_m5_dev();
  c->P = P;
  c->P.fast = c->prec != FV_PREC_EXACT;
  c->n = pk->n;
  c->nfine = pk->nfine;
  c->neq = 3 * P.nmat + 12;
  c->rk_steps = p->rk_steps;
  const size_t n = (size_t)c->n, nf = (size_t)(pk->nfine > 0 ? pk->nfine : 1), sz = n * c->neq;
  c->cap = c->n;
  c->cap_fine = (int)nf;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  chk(hipStreamCreate(&c->stream));
  chk(dalloc(&c->G, 1));
  chk(dalloc(&c->nbr, 6 * n));
  chk(dalloc(&c->fine4, 4 * nf));
  chk(dalloc(&c->info, n));
  chk(dalloc(&c->S, sz)); chk(dalloc(&c->S2, sz)); chk(dalloc(&c->SN, sz)); chk(dalloc(&c->Sb, sz));
  chk(dalloc(&c->RHS, sz));
  chk(dalloc(&c->SF, 6 * sz));
  chk(dalloc(&c->gfm, 6 * n)); chk(dalloc(&c->mark, n));
  chk(dalloc(&c->flag, 2)); chk(dalloc(&c->list, n)); chk(dalloc(&c->count, 3));
  chk(dalloc(&c->part, (n + 255) / 256 + 1));
  chk(dalloc(&c->lim, n));
  if (p->prelax == 3) {
    chk(dalloc(&c->nst, 18 * n)); chk(dalloc(&c->iters, 6 * n));
    chk(dalloc(&c->list6b, 6 * n)); chk(dalloc(&c->list6, 6 * n));
  }
  if (e == hipSuccess) {
    // temporary storage of the compaction of n leaves
    hipcub::CountingInputIterator<int> idx(0);
    chk(hipcub::DeviceSelect::Flagged(nullptr, c->cub_bytes, idx, c->mark, c->list, c->count, c->n, c->stream));
    chk(dalloc((char **)&c->cub_tmp, c->cub_bytes));
  }
  if (e == hipSuccess) {
    Geo G;
    for (int f = 0; f < 6; f++)
      for (int k = 0; k < 9; k++) G.frame[f][k] = pk->geo.frame[f][k];
    for (int l = 0; l < FV_MAXLEVEL + 2; l++) {
      G.h[l] = pk->geo.h[l]; G.area[l] = pk->geo.area[l]; G.vol[l] = pk->geo.vol[l];
      for (int k = 0; k < 3; k++) { G.dcf[k][l] = pk->geo.dcf[k][l]; G.dcc[k][l] = pk->geo.dcc[k][l]; }
    }
    chk(hipMemcpy(c->G, &G, sizeof(Geo), hipMemcpyHostToDevice));
    chk(hipMemcpy(c->nbr, pk->nbr, sizeof(int) * 6 * n, hipMemcpyHostToDevice));
    chk(hipMemcpy(c->info, pk->info, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    if (pk->nfine > 0) chk(hipMemcpy(c->fine4, pk->fine4, sizeof(int) * 4 * (size_t)pk->nfine, hipMemcpyHostToDevice));
    chk(hipMemset(c->flag, 0, 2 * sizeof(int)));
    chk(hipMemset(c->S, 0, sizeof(double) * sz));
    chk(hipMemset(c->SN, 0, sizeof(double) * sz));
    chk(hipMemset(c->SF, 0, sizeof(double) * 6 * sz));
    chk(hipMemset(c->lim, 0, n));
  }
  if (e != hipSuccess) {
    m5_free(c);
    return fv_set_err(e == hipErrorOutOfMemory ? FV_ERR_NOMEM : FV_ERR_HIP, hipGetErrorString(e));
  }
  *out = c;
  return FV_OK;
}

extern "C" int fv_m5_dev_detach(fv_m5_dev *c) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  m5_free(c);
  return FV_OK;
}

extern "C" int fv_m5_dev_set_state(fv_m5_dev *c, const double *S) {
  if (!c || !S) return fv_set_err(FV_ERR_ARG, "null argument");
  std::vector<double> h;
  to_soa(S, c->n, c->neq, h);
  HIPCHK(hipMemcpyAsync(c->S, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FV_OK;
}

extern "C" int fv_m5_dev_get_state(fv_m5_dev *c, double *S) {
  if (!c || !S) return fv_set_err(FV_ERR_ARG, "null argument");
  std::vector<double> h((size_t)c->n * c->neq);
  HIPCHK(hipMemcpyAsync(h.data(), c->S, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  from_soa(h, c->n, c->neq, S);
  return FV_OK;
}

extern "C" int fv_m5_dev_begin_step(fv_m5_dev *c) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(c->SN, c->S, sizeof(double) * (size_t)c->n * c->neq, hipMemcpyDeviceToDevice, c->stream));
  return FV_OK;
}

extern "C" int fv_m5_dev_stage(fv_m5_dev *c, int irk) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  if (irk < 1 || irk > 4) return fv_set_err(FV_ERR_ARG, "model 5: irk must be 1..4 (m5_RK.c cases)");
  int rc = m5_stage(c, irk, nullptr);
  return rc != FV_OK ? rc : m5_check(c);
}

extern "C" int fv_m5_dev_advance(fv_m5_dev *c, int nsteps) {
  if (!c || nsteps < 0) return fv_set_err(FV_ERR_ARG, "bad argument");
  for (int s = 0; s < nsteps; s++) {
    int rc = fv_m5_dev_begin_step(c);
    for (int irk = 1; irk <= c->rk_steps && rc == FV_OK; irk++) rc = m5_stage(c, irk, nullptr);
    if (rc != FV_OK) return rc;
    if ((rc = m5_check(c)) != FV_OK) return rc;
  }
  return FV_OK;
}

extern "C" int fv_m5_dev_set_dt(fv_m5_dev *c, double dt) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  if (!(dt > 0.0) || !isfinite(dt)) return fv_set_err(FV_ERR_ARG, "model 5: dt must be positive and finite");
  c->P.dt = dt;
  c->P.dtu = c->P.fix ? -dt : dt;
  return FV_OK;
}

extern "C" int fv_m5_dev_cfl_dt(fv_m5_dev *c, double cfl, double *dt) {
  if (!c || !dt) return fv_set_err(FV_ERR_ARG, "null argument");
  const int nb = (c->n + 255) / 256;
  if (c->P.nmat == 2)
    hipLaunchKernelGGL(k_m5_cfl<2>, dim3(nb), dim3(256), 0, c->stream, c->n, c->P, c->G, c->info, c->S, c->part);
  else
    hipLaunchKernelGGL(k_m5_cfl<3>, dim3(nb), dim3(256), 0, c->stream, c->n, c->P, c->G, c->info, c->S, c->part);
  hipLaunchKernelGGL(k_m5_min_final, dim3(1), dim3(1024), 0, c->stream, nb, c->part, c->part + nb, cfl);
  HIPCHK(hipGetLastError());
  double h = 0.0;
  HIPCHK(hipMemcpyAsync(&h, c->part + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *dt = h;
  return FV_OK;
}

extern "C" int fv_m5_dev_set_limiter_flags(fv_m5_dev *c, const uint8_t *lim) {
  if (!c || !lim) return fv_set_err(FV_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(c->lim, lim, (size_t)c->n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FV_OK;
}

extern "C" int fv_m5_dev_get_limiter_flags(fv_m5_dev *c, uint8_t *lim) {
  if (!c || !lim) return fv_set_err(FV_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(lim, c->lim, (size_t)c->n, hipMemcpyDeviceToHost));
  return FV_OK;
}

extern "C" int fv_m5_dev_set_precision(fv_m5_dev *c, int prec) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  if (prec != FV_PREC_FP64 && prec != FV_PREC_EXACT && prec != FV_PREC_MIXED)
    return fv_set_err(FV_ERR_ARG, "model 5: precision must be FV_PREC_FP64, FV_PREC_EXACT or FV_PREC_MIXED");
  c->prec = prec;
  c->P.fast = prec != FV_PREC_EXACT;
  return FV_OK;
}

extern "C" int fv_m5_dev_faceinterp(fv_m5_dev *c) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  int rc = c->P.nmat == 2 ? m5_faceinterp_t<2>(c) : m5_faceinterp_t<3>(c);
  return rc != FV_OK ? rc : m5_check(c);
}

extern "C" int fv_m5_dev_residual(fv_m5_dev *c) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  int rc = c->P.nmat == 2 ? m5_residual_t<2>(c) : m5_residual_t<3>(c);
  return rc != FV_OK ? rc : m5_check(c);
}

extern "C" int fv_m5_dev_get_stage_arrays(fv_m5_dev *c, double *SF, double *RHS, uint8_t *gfm_face) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  const size_t n = (size_t)c->n, neq = (size_t)c->neq;
  if (SF) {
    std::vector<double> h(6 * n * neq);
    HIPCHK(hipMemcpyAsync(h.data(), c->SF, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++)
      for (int f = 0; f < 6; f++)
        for (size_t v = 0; v < neq; v++) SF[(i * 6 + f) * neq + v] = h[((size_t)f * neq + v) * n + i];
  }
  if (RHS) {
    std::vector<double> h(n * neq);
    HIPCHK(hipMemcpyAsync(h.data(), c->RHS, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    from_soa(h, c->n, c->neq, RHS);
  }
  if (gfm_face) {
    HIPCHK(hipMemcpyAsync(gfm_face, c->gfm, 6 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return FV_OK;
}

extern "C" int fv_m5_dev_time_steps(fv_m5_dev *c, int nsteps, double *ms_stage, double *ms_residual) {
  if (!c || nsteps < 1) return fv_set_err(FV_ERR_ARG, "bad argument");
  hipEvent_t a0, a1, ev[2];
  HIPCHK(hipEventCreate(&a0)); HIPCHK(hipEventCreate(&a1));
  HIPCHK(hipEventCreate(&ev[0])); HIPCHK(hipEventCreate(&ev[1]));
  double tot_res = 0.0;
  int rc = FV_OK;
  HIPCHK(hipEventRecord(a0, c->stream));
  for (int s = 0; s < nsteps && rc == FV_OK; s++) {
    rc = fv_m5_dev_begin_step(c);
    for (int irk = 1; irk <= c->rk_steps && rc == FV_OK; irk++) {
      rc = m5_stage(c, irk, ev);
      if (rc != FV_OK) break;
      HIPCHK(hipEventSynchronize(ev[1]));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
      tot_res += ms;
    }
  }
  HIPCHK(hipEventRecord(a1, c->stream));
  HIPCHK(hipEventSynchronize(a1));
  float all = 0.f;
  HIPCHK(hipEventElapsedTime(&all, a0, a1));
  hipEventDestroy(a0); hipEventDestroy(a1); hipEventDestroy(ev[0]); hipEventDestroy(ev[1]);
  if (rc != FV_OK) return rc;
  const int nst = nsteps * c->rk_steps;
  if (ms_stage) *ms_stage = all / nst;
  if (ms_residual) *ms_residual = tot_res / nst;
  return m5_check(c);
}


int fv_m5_view_of(fv_m5_dev *c, fv_m5_view *v) {
  if (!c || !v) return fv_set_err(FV_ERR_ARG, "null argument");
  if (!c->topo) return fv_set_err(FV_ERR_ARG, "no forest: call fv_m5_dev_set_forest after attach");
  v->n = c->n; v->nbr = c->nbr; v->info = c->info; v->fine4 = c->fine4; v->lim = c->lim;
  v->stream = c->topo->stream;
  return FV_OK;
}

int fv_m5_indicator_raw(fv_m5_dev *c, const fv_crit *cr, int **SI, int **split) {
  if (!c || !cr) return fv_set_err(FV_ERR_ARG, "null argument");
  if (!c->topo) return fv_set_err(FV_ERR_ARG, "no forest: call fv_m5_dev_set_forest after attach");
  fv_dev *t = c->topo;
  int *cur = nullptr;
  int rc = m5_interface_layers(c, cr, &cur);
  if (rc != FV_OK) return rc;
  const dim3 g((c->n + 255) / 256), b(256);
  const M5Crit C = m5_crit(cr);
  if (c->P.nmat == 2)
    hipLaunchKernelGGL(k_m5_split<2>, g, b, 0, t->stream, c->n, c->P, C, c->S, c->nbr, c->info, c->fine4, cur, t->split);
  else
    hipLaunchKernelGGL(k_m5_split<3>, g, b, 0, t->stream, c->n, c->P, C, c->S, c->nbr, c->info, c->fine4, cur, t->split);
  HIPCHK(hipGetLastError());
  *SI = cur;
  *split = t->split;
  return FV_OK;
}

int fv_m5_mlp_fields(fv_m5_dev *c, float4 *fld, float *snd) {
  if (!c || !c->topo) return fv_set_err(FV_ERR_ARG, "no forest: call fv_m5_dev_set_forest after attach");
  const dim3 g((c->n + 255) / 256), b(256);
  if (c->P.nmat == 2)
    hipLaunchKernelGGL(k_m5_mlp_fields<2>, g, b, 0, c->topo->stream, c->n, c->P, c->S, fld, snd);
  else
    hipLaunchKernelGGL(k_m5_mlp_fields<3>, g, b, 0, c->topo->stream, c->n, c->P, c->S, fld, snd);
  HIPCHK(hipGetLastError());
  return FV_OK;
}

int fv_m5_mlp_labels(fv_m5_dev *c, const int *split, float thr, int *cls, float *smooth, long row0) {
  if (!c || !c->topo) return fv_set_err(FV_ERR_ARG, "no forest: call fv_m5_dev_set_forest after attach");
  const dim3 g((c->n + 255) / 256), b(256);
  if (c->P.nmat == 2)
    hipLaunchKernelGGL(k_m5_mlp_labels<2>, g, b, 0, c->topo->stream, c->n, c->P, c->S, c->nbr, c->info, c->fine4,
                       split, thr, cls, smooth, row0);
  else
    hipLaunchKernelGGL(k_m5_mlp_labels<3>, g, b, 0, c->topo->stream, c->n, c->P, c->S, c->nbr, c->info, c->fine4,
                       split, thr, cls, smooth, row0);
  HIPCHK(hipGetLastError());
  return FV_OK;
}

int fv_m5_regrid_targets(fv_m5_dev *c, int *split, const int *SI, int maxlevel, int *n_out) {
  if (!c || !c->topo) return fv_set_err(FV_ERR_ARG, "no forest: call fv_m5_dev_set_forest after attach");
  fv_state_hook h{c, m5_remap_hook, m5_install_hook};
  int rc = fv_dev_regrid_hook(c->topo, split, SI, maxlevel, &h, n_out);
  if (rc != FV_OK) return rc;
  return m5_check(c);
}

extern "C" int fv_m5_dev_set_forest(fv_m5_dev *c, const fv_mesh *m) {
  if (!c || !m) return fv_set_err(FV_ERR_ARG, "null argument");
  if (m->n != c->n) return fv_set_err(FV_ERR_ARG, "fv_m5_dev_set_forest: mesh and context differ in leaf count");
  fv_pack pk;
  memset(&pk, 0, sizeof pk);
  if (fv_mesh_pack(m, &pk) != FV_OK) return fv_set_err(FV_ERR_MESH, "fv_m5_dev_set_forest: packing the mesh failed");
  if (c->topo) { fv_dev_detach(c->topo); c->topo = nullptr; }
  int rc = fv_dev_attach_topology(&c->topo, &pk, c->P.nmat);
  fv_pack_free(&pk);
  if (rc != FV_OK) return rc;
  if ((rc = fv_dev_set_forest(c->topo, m)) != FV_OK) { fv_dev_detach(c->topo); c->topo = nullptr; return rc; }
  return FV_OK;
}

extern "C" int fv_m5_dev_get_forest(fv_m5_dev *c, int *root, int *level, int *ci, int *SI) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  if (!c->topo) return fv_set_err(FV_ERR_ARG, "no forest: call fv_m5_dev_set_forest after attach");
  return fv_dev_get_forest(c->topo, root, level, ci, SI);
}

extern "C" int fv_m5_dev_interface_location(fv_m5_dev *c, const fv_crit *cr, int *SI) {
  if (!c || !cr) return fv_set_err(FV_ERR_ARG, "null argument");
  if (!c->topo) return fv_set_err(FV_ERR_ARG, "no forest: call fv_m5_dev_set_forest after attach");
  fv_dev *t = c->topo;
  int *cur = nullptr;
  int rc = m5_interface_layers(c, cr, &cur);
  if (rc != FV_OK) return rc;
  // the new flags become the mesh's, and the face kinds of both contexts follow them
  fv_dev_refresh_kinds_launch(t->stream, c->n, c->nbr, cur, c->info);
  fv_dev_refresh_kinds_launch(t->stream, t->n, t->nbr, cur, t->info);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(t->lSI, cur, sizeof(int) * 2 * (size_t)c->n, hipMemcpyDeviceToDevice, t->stream));
  if (SI) HIPCHK(hipMemcpyAsync(SI, cur, sizeof(int) * 2 * (size_t)c->n, hipMemcpyDeviceToHost, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  return FV_OK;
}

extern "C" int fv_m5_dev_indicator(fv_m5_dev *c, const fv_crit *cr, int *SI, int *split) {
  int *cur = nullptr, *sp = nullptr;
  int rc = fv_m5_indicator_raw(c, cr, &cur, &sp);
  if (rc != FV_OK) return rc;
  hipStream_t st = c->topo->stream;
  if (SI) HIPCHK(hipMemcpyAsync(SI, cur, sizeof(int) * 2 * (size_t)c->n, hipMemcpyDeviceToHost, st));
  if (split) HIPCHK(hipMemcpyAsync(split, sp, sizeof(int) * (size_t)c->n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return FV_OK;
}

extern "C" int fv_m5_dev_regrid(fv_m5_dev *c, const fv_crit *cr, int maxlevel, int *n_out) {
  int *cur = nullptr, *sp = nullptr;
  int rc = fv_m5_indicator_raw(c, cr, &cur, &sp);
  if (rc != FV_OK) return rc;
  return fv_m5_regrid_targets(c, sp, cur, maxlevel, n_out);
}

extern "C" int fv_m5_neq(int nmat) { return 3 * nmat + 12; }

extern "C" int fv_m5_gfm_faces(const fv_m5_params *p, int nf, const double *UL, const double *UR,
                               const double *frame, const double *area, double *flux, double *unc) {
  M5P P;
  int rc = m5_params(p, P);
  if (rc != FV_OK) return rc;
  if (nf < 0 || (nf > 0 && (!UL || !UR || !frame || !area || !flux || !unc)))
    return fv_set_err(FV_ERR_ARG, "model 5: bad face arrays");
  if (nf == 0) return FV_OK;
  return P.nmat == 2 ? gfm_faces_t<2>(P, nf, UL, UR, frame, area, flux, unc)
                     : gfm_faces_t<3>(P, nf, UL, UR, frame, area, flux, unc);
}

extern "C" int fv_m5_rk(const fv_m5_params *p, const fv_pack *pk, int irk, double *SN, double *S, const double *RHS,
                        const uint8_t *gfm_face) {
  M5P P;
  int rc = m5_params(p, P);
  if (rc != FV_OK) return rc;
  if (!pk || !SN || !S || !RHS || !gfm_face) return fv_set_err(FV_ERR_ARG, "null argument");
  if (irk < 0 || irk > 4) return fv_set_err(FV_ERR_ARG, "model 5: irk must be 0..4 (m5_RK.c cases)");
  return P.nmat == 2 ? rk_t<2>(P, pk, irk, SN, S, RHS, gfm_face) : rk_t<3>(P, pk, irk, SN, S, RHS, gfm_face);
}

extern "C" int fv_m5_gfm_rel(const fv_m5_params *p, const fv_pack *pk, const double *S_before, double *S,
                             const uint8_t *gfm_face) {
  M5P P;
  int rc = m5_params(p, P);
  if (rc != FV_OK) return rc;
  if (!pk || !S_before || !S || !gfm_face) return fv_set_err(FV_ERR_ARG, "null argument");
  return P.nmat == 2 ? rel_t<2>(P, pk, S_before, S, gfm_face) : rel_t<3>(P, pk, S_before, S, gfm_face);
}
