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
};

template <int NM>
struct L5 {
  static constexpr int NEQ = 3 * NM + 12;
  static constexpr int NCONS = NM + 4;
  static constexpr int I0 = 0, I1 = NM, I2 = NM + 3, I3 = NM + 4, N3 = NM - 1, I4 = NM + 4 + NM - 1, I5 = I4 + NM;
};

__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ double maxr(double a, double b) { return (a > b) ? a : b; }   // strdata.h:27
__device__ __forceinline__ double minr(double a, double b) { return (a < b) ? a : b; }   // strdata.h:31

// m5_elastic_energy_solid_V2, state_equation.c:269-316
__device__ double elastic_energy_v2(const double *G, double Gmag, double mu) {
  double eints;
  if (fabs(Gmag) != 0.0) {
    double M1[9], M2[9];
    double Gmag13 = pow(fabs(Gmag), 1.0 / 3.0);
    if (isnan(Gmag13)) Gmag13 = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) { M1[k] = G[k]; M2[k] = 0.0; }
    M1[0] = G[0] - Gmag13; M1[4] = G[4] - Gmag13; M1[8] = G[8] - Gmag13;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) M2[i * 3 + j] += M1[i * 3 + k] * M1[k * 3 + j];
    const double tr1 = M2[0] + M2[4] + M2[8];
    eints = mu / (4.0 * pow(fabs(Gmag), 1.0 / 6.0)) * tr1;
  } else {
    eints = 0.0;
  }
  return eints;
}

// G_det (state_equation.c:206-215) and m5_elastic_energy_solid (:217-267)
__device__ double elastic_energy_v1(const double *G, double mu) {
  const double Gmag = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) +
                      G[2] * (G[3] * G[7] - G[4] * G[6]);
  if ((fabs(Gmag) != 0.0) && (mu != 0.0)) return elastic_energy_v2(G, Gmag, mu);
  return 0.0;
}

// m5_eos_stress_V2, state_equation.c:143-204, accumulated into st
__device__ void add_stress_v2(double *st, double avf, double pr, const double *G, double Gmag, double mu) {
  if ((fabs(Gmag) != 0.0) && (mu != 0.0)) {
    double G2[9], ss[9];
#pragma unroll
    for (int k = 0; k < 9; k++) G2[k] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) G2[i * 3 + j] += G[i * 3 + k] * G[k * 3 + j];
    const double J1 = G[0] + G[4] + G[8];
    const double J2 = G2[0] + G2[4] + G2[8];
    const double g13 = pow(fabs(Gmag), 1.0 / 3.0);
    ss[0] = G2[0] - J2 / 3.0 - g13 * (G[0] - J1 / 3.0);
    ss[1] = G2[1] - g13 * (G[1]);
    ss[2] = G2[2] - g13 * (G[2]);
    ss[3] = G2[3] - g13 * (G[3]);
    ss[4] = G2[4] - J2 / 3.0 - g13 * (G[4] - J1 / 3.0);
    ss[5] = G2[5] - g13 * (G[5]);
    ss[6] = G2[6] - g13 * (G[6]);
    ss[7] = G2[7] - g13 * (G[7]);
    ss[8] = G2[8] - J2 / 3.0 - g13 * (G[8] - J1 / 3.0);
    const double k = mu / pow(fabs(Gmag), 1.0 / 6.0);
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const double s = (q == 0 || q == 4 || q == 8) ? (-avf * pr - k * (ss[q])) : (-k * (ss[q]));
      st[q] += s;
    }
  } else {
    st[0] += -avf * pr; st[1] += 0.0; st[2] += 0.0;
    st[3] += 0.0; st[4] += -avf * pr; st[5] += 0.0;
    st[6] += 0.0; st[7] += 0.0; st[8] += -avf * pr;
  }
}

template <int NM>
struct Rec5 {
  double r, e, u, v, w, phi0;
  double ra[NM], avf[NM], are[NM], c2Y[NM], ph[NM];
  double A[9], st[9];
  double vm[3], vE;        // vec momentum and energy (rewritten for a ghost)
  double ut, vt, wt;       // u_temp, v_temp, w_temp
};

// face record of one state vector x (m5_facercnstr.c:3-190, n_solids = 0)
template <int NM>
__device__ void record(const M5P &P, const double *x, Rec5<NM> &r) {
  using L = L5<NM>;
  const double a = P.amin / 100.0;
  double avf_temp[NM], avf_c0 = 0.0;
  r.r = 0.0;
#pragma unroll
  for (int i = 0; i < NM; ++i) { r.r += x[i]; r.ra[i] = x[i]; }
  r.u = x[L::I1 + 0] / r.r;
  r.v = x[L::I1 + 1] / r.r;
  r.w = x[L::I1 + 2] / r.r;
  r.vm[0] = x[L::I1 + 0]; r.vm[1] = x[L::I1 + 1]; r.vm[2] = x[L::I1 + 2];
  r.e = x[L::I2] / r.r;
  r.vE = x[L::I2];
  double avfsum = 0.0;
#pragma unroll
  for (int i = 0; i < L::N3; ++i) {
    double t = x[L::I3 + i], c = x[L::I3 + i];
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
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    const double are_elast = (P.mush[i] != 0.0) ? elastic_energy_v2(G, Gmag, P.mush[i]) : 0.0;
    const double g = P.gamma[i], pi = P.pinf[i];
    const double press = ((g - 1.0) * (r.are[i] - are_elast) / avf_temp[i]) - g * pi;
    r.c2Y[i] = (avf_temp[i] / r.r) * g * (press + pi);
    r.ph[i] = press;
    add_stress_v2(r.st, r.avf[i], r.ph[i], G, Gmag, P.mush[i]);
  }
}

// rotate_tensor (rotate_vectors.c:128-188): t = V st V^T, V rows a, b, c
__device__ void rot_tensor(double *t, const double *st, const double *a, const double *b, const double *c) {
  double V[9] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
  double M1[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += V[i * 3 + k] * st[k * 3 + j];
      M1[i * 3 + j] = s;
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += M1[i * 3 + k] * V[j * 3 + k];   // vecT[k][j] = V[j][k]
      t[i * 3 + j] = s;
    }
}

template <int NM>
struct Star5 {
  double Us[NM + 4];
  double unc[3];
  double avf_S[NM], are_S[NM], AS[9];
};

// m5_star_region.c:3-157 for side s (flag 1: left, 2: right)
template <int NM>
__device__ void star_region(const M5P &P, const double *F, const Rec5<NM> &sL, const Rec5<NM> &sR,
                            const Rec5<NM> &s, int flag_ifs, double u_S, double SL, double SR, double uL,
                            double vL, double wL, double uR, double vR, double wR, double s11L, double s12L,
                            double s13L, double s11R, double s12R, double s13R, Star5<NM> &W) {
  const double *n = F, *m = F + 3, *l = F + 6;
  double SK, uK, vK, wK, s11K, s12K, s13K;
  if (flag_ifs == 1) { SK = SL; uK = uL; vK = vL; wK = wL; s11K = s11L; s12K = s12L; s13K = s13L; }
  else { SK = SR; uK = uR; vK = vR; wK = wR; s11K = s11R; s12K = s12R; s13K = s13R; }
  double ar_S[NM], r_S = 0.0;
#pragma unroll
  for (int v = 0; v < NM; ++v) { ar_S[v] = s.ra[v] * (SK - uK) / (SK - u_S); r_S += ar_S[v]; }
  const double s11_S = ((uR - SR) * sR.r * (s11L) - (uL - SL) * sL.r * (s11R) + (uL - SL) * sL.r * (uR - SR) * sR.r * (uR - uL)) /
                       ((uR - SR) * sR.r - (uL - SL) * sL.r);
  double s12_S, s13_S;
  if ((sL.phi0 > 0.0) && (sR.phi0 > 0.0)) {
    s12_S = ((uR - SR) * sR.r * (s12L) - (uL - SL) * sL.r * (s12R) + (uL - SL) * sL.r * (uR - SR) * sR.r * (vR - vL)) /
            ((uR - SR) * sR.r - (uL - SL) * sL.r);
    s13_S = ((uR - SR) * sR.r * (s13L) - (uL - SL) * sL.r * (s13R) + (uL - SL) * sL.r * (uR - SR) * sR.r * (wR - wL)) /
            ((uR - SR) * sR.r - (uL - SL) * sL.r);
  } else {
    s12_S = 0.0;
    s13_S = 0.0;
  }
  const double v_S = vK + (s12_S - s12K) / ((uK - SK) * s.r);
  const double w_S = wK + (s13_S - s13K) / ((uK - SK) * s.r);
  double vel[3];
  vel[0] = u_S * n[0] + v_S * m[0] + w_S * l[0];   // rotate_vector_ltg_v2
  vel[1] = u_S * n[1] + v_S * m[1] + w_S * l[1];
  vel[2] = u_S * n[2] + v_S * m[2] + w_S * l[2];
  const double re = s.r * s.e;
  const double Em_S = (re * (uK - SK) + s11K * uK - s12K * vK - s13K * wK + s11_S * u_S + s12_S * v_S + s13_S * w_S) /
                      (r_S * (u_S + SK));
  double A[9], t[9], AS[9];
  rot_tensor(t, s.A, n, m, l);
#pragma unroll
  for (int k = 0; k < 9; k++) A[k] = t[k];
#pragma unroll
  for (int j = 0; j < 3; ++j) AS[j] = (A[j] * (uK - SK) - A[3 + j] * (vK - v_S) + A[6 + j] * (wK - w_S)) / (u_S - SK);
#pragma unroll
  for (int j = 0; j < 3; ++j) { AS[3 + j] = A[3 + j]; AS[6 + j] = A[6 + j]; }
  {
    const double c0[3] = {n[0], m[0], l[0]}, c1[3] = {n[1], m[1], l[1]}, c2[3] = {n[2], m[2], l[2]};
    rot_tensor(t, AS, c0, c1, c2);
  }
#pragma unroll
  for (int k = 0; k < 9; k++) AS[k] = t[k];
  double G[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G[i * 3 + j] = AS[i * 3 + 0] * AS[j * 3 + 0] + AS[i * 3 + 1] * AS[j * 3 + 1] + AS[i * 3 + 2] * AS[j * 3 + 2];
  double e_S[NM];
#pragma unroll
  for (int v = 0; v < NM; ++v) {
    const double g = P.gamma[v], pi = P.pinf[v];
    const double rm = (s.ra[v] / s.avf[v]);
    const double rm_S = (s.ra[v] / s.avf[v]) * (uK - SK) / (u_S - SK);
    const double p_k = -s.ph[v];
    const double p_S = (p_k + pi) * (((g - 1.0) * rm + (g + 1.0) * rm_S) / ((g - 1.0) * rm_S + (g + 1.0) * rm)) - pi;
    const double are_el = elastic_energy_v1(G, P.mush[v]);
    const double e_el = are_el / (s.avf[v] * rm_S);
    e_S[v] = (p_S + g * pi) / (rm_S * (g + 1.0)) + e_el;
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

// one interface sub-face: m5_hllc_gfm + m5_residual.c:166-170
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
  const double *n = F, *m = F + 3, *l = F + 6;
  // face-frame velocities (rotate_vector_gtl_v2) and the ghost: v_R = v_L, w_R = w_L
  const double uL = sL.u * n[0] + sL.v * n[1] + sL.w * n[2];
  const double vL = sL.u * m[0] + sL.v * m[1] + sL.w * m[2];
  const double wL = sL.u * l[0] + sL.v * l[1] + sL.w * l[2];
  const double uR = sR.u * n[0] + sR.v * n[1] + sR.w * n[2];
  const double vR = vL, wR = wL;
  sL.ut = sL.u; sL.vt = sL.v; sL.wt = sL.w;
  sR.ut = uR * n[0] + vR * m[0] + wR * l[0];
  sR.vt = uR * n[1] + vR * m[1] + wR * l[1];
  sR.wt = uR * n[2] + vR * m[2] + wR * l[2];
  sR.vm[0] = sR.ut * sR.r; sR.vm[1] = sR.vt * sR.r; sR.vm[2] = sR.wt * sR.r;
  sR.vE = sR.r * sR.e;
  // m5_wave_speeds_gfm.c:16-58
  const double ruL = sL.r * (uL), ruR = sR.r * (uR);
  double c2YL = 0.0, c2YR = 0.0;
#pragma unroll
  for (int iv = 0; iv < NM; ++iv) { c2YL += sL.c2Y[iv]; c2YR += sR.c2Y[iv]; }
  const double cL = sqrt(c2YL), cR = sqrt(c2YR);
  const double SR = maxr((uL + cL), (uR + cR));
  const double SL = minr((uL - cL), (uR - cR));
  double t[9];
  rot_tensor(t, sL.st, n, m, l);
  const double s11L = t[0], s12L = t[1], s13L = t[2];
  rot_tensor(t, sR.st, n, m, l);
  const double s11R = t[0], s12R = t[1], s13R = t[2];
  const double u_S = ((sL.r * sq((uL)) - (s11L)) - (sR.r * sq((uR)) - (s11R)) - ((SL)*ruL) + ((SR)*ruR)) /
                     (ruL - ruR - ((SL)*sL.r) + ((SR)*sR.r));
  if (isnan(u_S) || isnan(SR) || isnan(SL)) { atomicOr(nanflag, 1); return; }
  int flag_ifs;
  if (0.0 <= SL) flag_ifs = 0;
  else if ((SL < 0.0) && (0.0 <= u_S)) flag_ifs = 1;
  else if ((u_S < 0.0) && (0.0 < SR)) flag_ifs = 2;
  else flag_ifs = 3;   // SR <= 0.0, the only case left once no speed is NaN
  const Rec5<NM> &s = (flag_ifs <= 1) ? sL : sR;
  Star5<NM> W;
  double SK = 0.0;
  if (flag_ifs == 1 || flag_ifs == 2) {
    SK = (flag_ifs == 1) ? SL : SR;
    star_region<NM>(P, F, sL, sR, s, flag_ifs, u_S, SL, SR, uL, vL, wL, uR, vR, wR, s11L, s12L, s13L, s11R, s12R, s13R, W);
#pragma unroll
    for (int iv = 0; iv < L::NCONS; ++iv)
      if (isnan(W.Us[iv])) { atomicOr(nanflag, 1); return; }
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
  auto emit = [&](int iv, double f0, double f1, double f2, double vec_iv) {
    const double h = (star && iv < L::NCONS) ? SK * (W.Us[iv] - vec_iv) : 0.0;
    if (isnan(h)) atomicOr(nanflag, 1);
    double fl = 0.0;
    fl += (area[q]) * (((f0 - 0.0) * n[0] + (f1 - 0.0) * n[1] + (f2 - 0.0) * n[2]) + h);
    flux[(size_t)iv * nf + q] = fl;
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
  for (int v = 0; v < NM - 1; ++v) emit(L::I3 + v, W.unc[0] * W.avf_S[v], W.unc[1] * W.avf_S[v], W.unc[2] * W.avf_S[v], 0.0);
#pragma unroll
  for (int v = 0; v < NM; ++v) emit(L::I4 + v, W.unc[0] * W.are_S[v], W.unc[1] * W.are_S[v], W.unc[2] * W.are_S[v], 0.0);
#pragma unroll
  for (int k = 0; k < 9; k++) emit(L::I5 + k, W.unc[0] * W.AS[k], W.unc[1] * W.AS[k], W.unc[2] * W.AS[k], 0.0);
  unc[(size_t)0 * nf + q] = W.unc[0];
  unc[(size_t)1 * nf + q] = W.unc[1];
  unc[(size_t)2 * nf + q] = W.unc[2];
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
          const double as_new = SN[at(e)] + P.dt * buf;
          const double phi_new = as_new - P.avf_lim;
          if (phi_old * phi_new < 0.0) flag_rk = 1;
        }
      }
#pragma unroll
      for (int e = 0; e < L::NEQ; e++) {
        if (flag_rk && e < L::I2) { Sout[at(e)] = Sin[at(e)]; continue; }   // set by m5_gfm
        const double buf = (1.0 / Vol) * RHS[at(e)];
        Sout[at(e)] = SN[at(e)] + P.dt * buf;
      }
      break;
    }
    case 2:
#pragma unroll
      for (int e = 0; e < L::NEQ; e++)
        Sout[at(e)] = (3.0 / 4.0) * SN[at(e)] + (1.0 / 4.0) * Sin[at(e)] + (1.0 / 4.0) * P.dt * RHS[at(e)] / Vol;
      break;
    case 3:
#pragma unroll
      for (int e = 0; e < L::NEQ; e++)
        Sout[at(e)] = (1.0 / 3.0) * SN[at(e)] + (2.0 / 3.0) * Sin[at(e)] + (2.0 / 3.0) * P.dt * RHS[at(e)] / Vol;
      break;
    case 4:
#pragma unroll
      for (int e = 0; e < L::NEQ; e++)
        Sout[at(e)] = (1.0 / 2.0) * SN[at(e)] + (1.0 / 2.0) * Sin[at(e)] + (1.0 / 2.0) * P.dt * RHS[at(e)] / Vol;
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
__global__ void k_m5_gfm_seq(int n, int nlist, const int *__restrict__ list, M5P P, const Geo *__restrict__ Gp,
                             const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                             const int *__restrict__ fine4, const uint8_t *__restrict__ gf,
                             const double *__restrict__ SN, const double *__restrict__ Sin, double *__restrict__ S,
                             const double *__restrict__ RHS, int *__restrict__ err) {
  using L = L5<NM>;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int t = 0; t < nlist; t++) {
    const int i = list[t];
    auto at = [&](int e) { return (size_t)e * n + i; };
    const double dV = P.dt / Gp->vol[fvk::cell_level(info[i])];
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
__global__ void k_m5_rel_seq(int n, int nlist, const int *__restrict__ list, M5P P, const Geo *__restrict__ Gp,
                             const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                             const int *__restrict__ fine4, const uint8_t *__restrict__ gf,
                             const double *__restrict__ Sb, double *__restrict__ S, int *__restrict__ err) {
  using L = L5<NM>;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
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

// leaves with mark[i] != 0 in leaf order; returns their count in *count
int compact_marked(Bufs &B, int n, const uint8_t *mark, int **list, int *count) {
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
    int *list = nullptr, count = 0;
    if ((rc = compact_marked(B, n, dmark, &list, &count)) != FV_OK) return rc;
    if (count > 0) {
      hipLaunchKernelGGL(k_m5_gfm_seq<NM>, dim3(1), dim3(64), 0, 0, n, count, list, P, M.G, M.nbr, M.info, M.fine4,
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
  int *list = nullptr, count = 0;
  if ((rc = compact_marked(B, n, dmark, &list, &count)) != FV_OK) return rc;
  if (count > 0) {
    hipLaunchKernelGGL(k_m5_rel_seq<NM>, dim3(1), dim3(64), 0, 0, n, count, list, P, M.G, M.nbr, M.info, M.fine4, dgf,
                       dSb, dS, derr);
    HIPCHK(hipGetLastError());
    int err = 0;
    HIPCHK(hipMemcpy(&err, derr, sizeof(int), hipMemcpyDeviceToHost));
    if (err) return fv_set_err(FV_ERR_ARG, "model 5: gfm_face set on a boundary face");
  }
  HIPCHK(hipMemcpy(hS.data(), dS, sizeof(double) * sz, hipMemcpyDeviceToHost));
  from_soa(hS, n, NEQ, S);
  return FV_OK;
}

}  // namespace

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
