// fv_phys.h -- device restatement of the model-4 face physics for gfx950.
//
// Templated on the number of materials NM; every function keeps the
// operation order of the reference routine it cites so that an FP64 build
// with -ffp-contract=off reproduces the CPU path bit for bit.  The real type
// R is double for the plain path and float for the mixed-precision face
// kernels (reconstruction, HLLC); pressure recovery from energy always runs
// in double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FV_DEV __device__ __forceinline__

namespace fvk {

template <int NM>
struct Lay {
  static constexpr int NEQ = 3 * NM + 3;
  static constexpr int I1 = NM, I2 = NM + 3, I3 = NM + 4, I4 = 2 * NM + 3, N3 = NM - 1;
  static constexpr int NCONS = NM + 4;
};

struct Params {
  double gamma[4], pinf[4];
  double amin, avf_lim, dt;
  int prelax, limiter, rk, nmat, fix, primtv;
  int bc_fixed_p;               // fixed-pressure ghost at outlet (2) / inlet (3) faces
  double p_inlet, p_outlet;
  // per-material constants formed once on the host: G p_inf, G - 1, 1/(G-1),
  // G p_inf/(G-1), 1/G (used by the throughput kernels)
  double GP[4], Gm1[4], iGm1[4], C[4], iG[4];
  float Gf[4], Pf[4], GPf[4], Gm1f[4], iGm1f[4], Cf[4], iGf[4];
};

struct Geo {
  double frame[6][9];
  double area[12], vol[12], h[12];
  double dcf[3][12], dcc[3][12];
};

__constant__ static const int FCOPP_D[6] = {2, 3, 0, 1, 5, 4};

FV_DEV int face_kind(uint32_t info, int f) { return (int)((info >> (8 + 3 * f)) & 7u); }
FV_DEV int cell_level(uint32_t info) { return (int)(info & 15u); }

template <typename R> FV_DEV R sq(R x) { return x * x; }
template <typename R> FV_DEV R maxr(R a, R b) { return ((a) > (b)) ? (a) : (b); }
template <typename R> FV_DEV R minr(R a, R b) { return ((a) < (b)) ? (a) : (b); }

// ---------------------------------------------------------------------
// SOL: what m4_facevalues / m4_facercnstr leave in run->sol_L / sol_R
// ---------------------------------------------------------------------
template <int NM, typename R>
struct Sol {
  R r, ra[NM], u, v, w, e;
  R mom[3], rE;          // vec[] of the conservative momentum and energy
  R avf[NM], are[NM];
  R st;                  // st[0][0] = st[1][1] = st[2][2]; off-diagonal entries are 0
  R c2Y[NM], p[NM];
};

// m4_facevalues.c:20-106 / m4_facercnstr.c:20-106 (phi omitted: model 4 never reads it)
template <int NM, typename R>
FV_DEV void fill_sol(const Params &P, const R *X, Sol<NM, R> &s) {
  using L = Lay<NM>;
  const R a = (R)(P.amin / 100.0);
  R at[NM];
  s.r = (R)0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) { s.r += X[k]; s.ra[k] = X[k]; }
  s.u = X[L::I1 + 0] / s.r;
  s.v = X[L::I1 + 1] / s.r;
  s.w = X[L::I1 + 2] / s.r;
  s.mom[0] = X[L::I1 + 0]; s.mom[1] = X[L::I1 + 1]; s.mom[2] = X[L::I1 + 2];
  s.e = X[L::I2] / s.r;
  s.rE = X[L::I2];
  R avfsum = (R)0.0;
#pragma unroll
  for (int k = 0; k < L::N3; ++k) {
    at[k] = X[L::I3 + k];
    if (at[k] <= a) at[k] = a;
    else if (at[k] >= ((R)1.0 - (a))) at[k] = (R)1.0 - (a);
    avfsum += at[k];
  }
  if (avfsum > ((R)1.0 - (a))) {
    R dif = avfsum - ((R)1.0 - (a));
#pragma unroll
    for (int j = 0; j < L::N3; ++j) { R per = at[j] / avfsum; at[j] = at[j] - per * dif; }
    avfsum = (R)1.0 - (a);
  }
  at[L::N3] = (R)1.0 - avfsum;
  R s2 = (R)0.0;
#pragma unroll
  for (int k = 0; k < L::N3; ++k) { s.avf[k] = at[k]; s2 += at[k]; }
  s.avf[L::N3] = (R)1.0 - s2;
#pragma unroll
  for (int k = 0; k < NM; ++k) s.are[k] = X[L::I4 + k];
  s.st = (R)0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    const R g = (R)P.gamma[k], pi = (R)P.pinf[k];
    R press = ((g - (R)1.0) * (s.are[k]) / at[k]) - g * pi;
    s.c2Y[k] = (at[k] / s.r) * g * (press + pi);
    s.st += -at[k] * press;
    s.p[k] = press;
  }
}

// m4_bc_inlet / m4_bc_outlet with CASE 3400 or 3410 (m4_bc.c:44-113, 121-186):
// every material of the ghost at pressure pb, with the ghost's own fractions,
// densities and velocity; are, c2Y, e and rho E follow (the reference writes
// materials 0 and 1, its injector cases having two; this covers all)
template <int NM, typename R>
FV_DEV void fixed_pressure_ghost(const Params &P, R pb, Sol<NM, R> &s) {
  s.st = -pb;
  R re = (R)0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    const R g = (R)P.gamma[k], pi = (R)P.pinf[k];
    s.p[k] = pb;
    s.are[k] = s.avf[k] * (pb + g * pi) / (g - (R)1.0);
    s.c2Y[k] = (s.avf[k] / s.r) * g * (pb + pi);
    re += s.are[k];
  }
  re = re + (R)0.5 * s.r * (s.u * s.u + s.v * s.v + s.w * s.w);
  s.e = re / s.r;
  s.rE = re;
}

// bc_symmetry.c / bc_reflective.c:12-28 / bc_wall.c:12-26: only sol_R->u,v,w move
template <int NM, typename R>
FV_DEV void apply_bc(const Params &P, int code, const R *F, const Sol<NM, R> &sL, Sol<NM, R> &sR) {
  if (code == 1 || code == 4 || code == 5) {
    R u = sL.u, v = sL.v, w = sL.w;
    R ul = u * F[0] + v * F[1] + w * F[2];
    R vl = u * F[3] + v * F[4] + w * F[5];
    R wl = u * F[6] + v * F[7] + w * F[8];
    R ur = -ul, vr = vl, wr = wl;
    if (code == 5) { vr = -vl; wr = -wl; }
    sR.u = ur * F[0] + vr * F[3] + wr * F[6];
    sR.v = ur * F[1] + vr * F[4] + wr * F[7];
    sR.w = ur * F[2] + vr * F[5] + wr * F[8];
  }
  if (P.bc_fixed_p && (code == 2 || code == 3))
    fixed_pressure_ghost<NM, R>(P, (R)(code == 3 ? P.p_inlet : P.p_outlet), sR);
}

// Output of one Riemann problem: n-flux (F.n + F_HLLC) and the star velocity.
template <int NM, typename R>
struct FaceFlux {
  R f[Lay<NM>::NEQ];
  R un, vn, wn;
};

// m4_flux.c:3-73 dotted with n (the caller adds the HLLC correction)
template <int NM, typename R>
FV_DEV void phys_flux_n(const Sol<NM, R> &s, const R *F, R u_nc, R v_nc, R w_nc,
                        const R *avf_S, const R *are_S, R *out) {
  using L = Lay<NM>;
  const R nx = F[0], ny = F[1], nz = F[2];
  const R z = (R)0.0;   // off-diagonal stress entries
  R f0, f1, f2;
#pragma unroll
  for (int k = 0; k < NM; k++) {
    f0 = s.u * s.ra[k]; f1 = s.v * s.ra[k]; f2 = s.w * s.ra[k];
    out[k] = ((f0 - z) * nx + (f1 - z) * ny + (f2 - z) * nz);
  }
  f0 = s.u * s.u * s.r - s.st; f1 = s.v * s.u * s.r - z; f2 = s.w * s.u * s.r - z;
  out[L::I1 + 0] = ((f0 - z) * nx + (f1 - z) * ny + (f2 - z) * nz);
  f0 = s.u * s.v * s.r - z; f1 = s.v * s.v * s.r - s.st; f2 = s.w * s.v * s.r - z;
  out[L::I1 + 1] = ((f0 - z) * nx + (f1 - z) * ny + (f2 - z) * nz);
  f0 = s.u * s.w * s.r - z; f1 = s.v * s.w * s.r - z; f2 = s.w * s.w * s.r - s.st;
  out[L::I1 + 2] = ((f0 - z) * nx + (f1 - z) * ny + (f2 - z) * nz);
  f0 = s.r * s.e * s.u - s.st * s.u - z * s.v - z * s.w;
  f1 = s.r * s.e * s.v - z * s.u - s.st * s.v - z * s.w;
  f2 = s.r * s.e * s.w - z * s.u - z * s.v - s.st * s.w;
  out[L::I2] = ((f0 - z) * nx + (f1 - z) * ny + (f2 - z) * nz);
#pragma unroll
  for (int k = 0; k < L::N3; ++k) {
    f0 = u_nc * avf_S[k]; f1 = v_nc * avf_S[k]; f2 = w_nc * avf_S[k];
    out[L::I3 + k] = ((f0 - z) * nx + (f1 - z) * ny + (f2 - z) * nz);
  }
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    f0 = u_nc * are_S[k]; f1 = v_nc * are_S[k]; f2 = w_nc * are_S[k];
    out[L::I4 + k] = ((f0 - z) * nx + (f1 - z) * ny + (f2 - z) * nz);
  }
}

// m4_hllc.c:3-61 with m4_wave_speeds.c:3-76, m4_star_region.c:3-87 and
// rotate_tensor (rotate_vectors.c:128-188) written out for a diagonal stress.
template <int NM, typename R>
FV_DEV void hllc(const Params &P, const R *F, int bc, const Sol<NM, R> &sL, const Sol<NM, R> &sR,
                 FaceFlux<NM, R> &out) {
  using L = Lay<NM>;
  const R nx = F[0], ny = F[1], nz = F[2], mx = F[3], my = F[4], mz = F[5], lx = F[6], ly = F[7], lz = F[8];
  R uL = sL.u * nx + sL.v * ny + sL.w * nz;
  R vL = sL.u * mx + sL.v * my + sL.w * mz;
  R wL = sL.u * lx + sL.v * ly + sL.w * lz;
  R uR = sR.u * nx + sR.v * ny + sR.w * nz;
  R vR = sR.u * mx + sR.v * my + sR.w * mz;
  R wR = sR.u * lx + sR.v * ly + sR.w * lz;
  R ruL = sL.r * (uL), ruR = sR.r * (uR);
  R c2YL = (R)0.0, c2YR = (R)0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) { c2YL += sL.c2Y[k]; c2YR += sR.c2Y[k]; }
  R cL = sqrt(c2YL), cR = sqrt(c2YR);
  R SR = maxr((uL + cL), (uR + cR));
  R SL = minr((uL - cL), (uR - cR));
  // row 0 of vec . st . vec^T for st = diag(s)
  R s11L = (nx * sL.st) * nx + (ny * sL.st) * ny + (nz * sL.st) * nz;
  R s12L = (nx * sL.st) * mx + (ny * sL.st) * my + (nz * sL.st) * mz;
  R s13L = (nx * sL.st) * lx + (ny * sL.st) * ly + (nz * sL.st) * lz;
  R s11R = (nx * sR.st) * nx + (ny * sR.st) * ny + (nz * sR.st) * nz;
  R s12R = (nx * sR.st) * mx + (ny * sR.st) * my + (nz * sR.st) * mz;
  R s13R = (nx * sR.st) * lx + (ny * sR.st) * ly + (nz * sR.st) * lz;
  R u_S = ((sL.r * sq((uL)) - (s11L)) - (sR.r * sq((uR)) - (s11R)) - ((SL)*ruL) + ((SR)*ruR)) /
          (ruL - ruR - ((SL)*sL.r) + ((SR)*sR.r));
  if (bc == 4) u_S = (R)0.0;

  int region;                     // 0 left, 1 left star, 2 right star, 3 right
  if ((R)0.0 <= SL) region = 0;
  else if ((SL <= (R)0.0) && ((R)0.0 <= u_S)) region = 1;
  else if ((u_S <= (R)0.0) && ((R)0.0 <= SR)) region = 2;
  else if (SR <= (R)0.0) region = 3;
  else region = -1;               // unordered (NaN) speeds: the reference leaves the flux unset

  if (region == 0 || region == 3 || region < 0) {
    const Sol<NM, R> &s = (region == 3) ? sR : sL;
    phys_flux_n<NM, R>(s, F, s.u, s.v, s.w, s.avf, s.are, out.f);
    out.un = s.u; out.vn = s.v; out.wn = s.w;
    if (region < 0) {
#pragma unroll
      for (int k = 0; k < L::NEQ; k++) out.f[k] = (R)NAN;
    }
    return;
  }
  const Sol<NM, R> &s = (region == 1) ? sL : sR;
  R SK, uK, vK, wK, s11K, s12K, s13K;
  if (region == 1) { SK = SL; uK = uL; vK = vL; wK = wL; s11K = s11L; s12K = s12L; s13K = s13L; }
  else { SK = SR; uK = uR; vK = vR; wK = wR; s11K = s11R; s12K = s12R; s13K = s13R; }
  R ar_S[NM], r_S = (R)0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) { ar_S[k] = s.ra[k] * (SK - uK) / (SK - u_S); r_S += ar_S[k]; }
  R s11_S = ((uR - SR) * sR.r * (s11L) - (uL - SL) * sL.r * (s11R) + (uL - SL) * sL.r * (uR - SR) * sR.r * (uR - uL)) /
            ((uR - SR) * sR.r - (uL - SL) * sL.r);
  const R s12_S = (R)0.0, s13_S = (R)0.0;
  R v_S = vK + (s12_S - (s12K)) / ((uK - SK) * s.r);
  R w_S = wK + (s13_S - (s13K)) / ((uK - SK) * s.r);
  R vel0 = u_S * nx + v_S * mx + w_S * lx;
  R vel1 = u_S * ny + v_S * my + w_S * ly;
  R vel2 = u_S * nz + v_S * mz + w_S * lz;
  R re = s.r * s.e;
  // FV_FIX_ENERGY (bit 2): -s11K uK as the Rankine-Hugoniot balance requires; the
  // published m4_star_region.c:46 has +s11K uK
  const R s11w = (P.fix & 2) ? -(s11K * uK) : (s11K * uK);
  R Em_S = (re * (uK - SK) + s11w - s12K * vK - s13K * wK + s11_S * u_S + s12_S * v_S + s13_S * w_S) / (r_S * (u_S - SK));
  R are_S[NM];
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    const R g = (R)P.gamma[k], pi = (R)P.pinf[k];
    R rm = (s.ra[k] / s.avf[k]);
    R rm_S = (s.ra[k] / s.avf[k]) * (uK - SK) / (u_S - SK);
    R p_S = (s.p[k] + pi) * (((g - (R)1.0) * rm + (g + (R)1.0) * rm_S) / ((g - (R)1.0) * rm_S + (g + (R)1.0) * rm)) - pi;
    R e_S = (p_S + g * pi) / (rm_S * (g - (R)1.0));
    are_S[k] = ar_S[k] * e_S;
  }
  phys_flux_n<NM, R>(s, F, vel0, vel1, vel2, s.avf, are_S, out.f);
  // F_HLLC = S_K (U* - U_K) on the NCONSEQ conservative equations
#pragma unroll
  for (int k = 0; k < NM; ++k) out.f[k] = out.f[k] + SK * (ar_S[k] - s.ra[k]);
  out.f[L::I1 + 0] = out.f[L::I1 + 0] + SK * (r_S * vel0 - s.mom[0]);
  out.f[L::I1 + 1] = out.f[L::I1 + 1] + SK * (r_S * vel1 - s.mom[1]);
  out.f[L::I1 + 2] = out.f[L::I1 + 2] + SK * (r_S * vel2 - s.mom[2]);
  out.f[L::I2] = out.f[L::I2] + SK * (r_S * Em_S - s.rE);
  out.un = vel0; out.vn = vel1; out.wn = vel2;
}

// ---------------------------------------------------------------------
// relaxation (m4_relaxationvec.c) -- always FP64
// ---------------------------------------------------------------------
template <int NM>
FV_DEV void vf_limit_relax(double a, double *avf) {
  using L = Lay<NM>;
  double sum = 0.0;
#pragma unroll
  for (int k = 0; k < L::N3; ++k) {
    if (avf[k] <= a) avf[k] = a;
    if (avf[k] >= (1.0 - a)) avf[k] = 1.0 - a;
    sum += avf[k];
  }
  if (sum >= (1.0 - a)) {
    double dif = sum - (1.0 - a);
#pragma unroll
    for (int j = 0; j < L::N3; ++j) { double per = avf[j] / sum; avf[j] = avf[j] - per * dif; }
    sum = 1.0 - a;
  }
  avf[L::N3] = 1.0 - sum;
}

template <int NM>
FV_DEV void relax(const Params &P, double *vec) {
  using L = Lay<NM>;
  const double a = P.amin / 100.0;
  double ar[NM], are[NM], avf[NM + 1], rho = 0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) { ar[k] = vec[k]; are[k] = vec[L::I4 + k]; rho += ar[k]; }
  const double u = vec[L::I1] / rho, v = vec[L::I1 + 1] / rho, w = vec[L::I1 + 2] / rho;
  const double rE = vec[L::I2];
#pragma unroll
  for (int k = 0; k < L::N3; ++k) avf[k] = vec[L::I3 + k];
  vf_limit_relax<NM>(a, avf);
  if (NM == 2 && P.prelax == 1) {   // m4_relaxationvec.c:63-127
    const double G0 = P.gamma[0], G1 = P.gamma[1], P0 = P.pinf[0], P1 = P.pinf[1];
    double t0_0 = 1.0 / (ar[0] / avf[0]);
    double t0_1 = 1.0 / (ar[1] / avf[1]);
    double p0_0 = (are[0] * (G0 - 1.0) - avf[0] * G0 * P0) / avf[0];
    double p0_1 = (are[1] * (G1 - 1.0) - avf[1] * G1 * P1) / avf[1];
    double A1 = (avf[0] / G0) * (p0_0 + P0) / ((avf[0] / G0) + (avf[1] / G1));
    double A2 = (avf[1] / G1) * (p0_1 + P1) / ((avf[0] / G0) + (avf[1] / G1));
    double pe = 0.5 * (A1 + A2 - (P0 + P1)) + sqrt(0.25 * (sq((A2 - A1 - (P1 - P0)))) + A1 * A2);
    double t1_0 = t0_0 * (p0_0 + G0 * P0 + (G0 - 1.0) * pe) / (pe + G0 * P0 + (G0 - 1.0) * pe);
    double t1_1 = t0_1 * (p0_1 + G1 * P1 + (G1 - 1.0) * pe) / (pe + G1 * P1 + (G1 - 1.0) * pe);
    avf[0] = ar[0] / (1.0 / t1_0);
    avf[1] = ar[1] / (1.0 / t1_1);
    vf_limit_relax<NM>(a, avf);
  } else if (P.prelax == 2) {       // m4_relaxationvec.c:128-297 with F_pe / DF_pe (F_pe.c:18-27, 64-73)
    // the published branch fills local arrays while F_pe reads never-written
    // RUN fields of the same names; the iteration uses the local values
    const double eb = 1.0e-8;
    double t0[NM], p0[NM], max_avf = -1.0;
    int principal = 0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      t0[k] = 1.0 / (ar[k] / avf[k]);
      p0[k] = (are[k] * (P.gamma[k] - 1.0) - avf[k] * P.gamma[k] * P.pinf[k]) / avf[k];
      if (p0[k] < -(1.0 - eb) * P.pinf[k] + eb) p0[k] = -(1.0 - eb) * P.pinf[k] + 1.0;
      if (avf[k] > max_avf) { max_avf = avf[k]; principal = k; }
    }
    auto Fpe = [&](double pe) {
      double F = 0.0;
#pragma unroll
      for (int k = 0; k < NM; ++k) {
        const double G = P.gamma[k], Pi = P.pinf[k];
        F += ar[k] * t0[k] * ((p0[k] + G * Pi + (G - 1.0) * pe) / (pe + G * Pi + (G - 1.0) * pe));
      }
      return F + -1.0;
    };
    auto DFpe = [&](double pe) {
      double D = 0.0;
#pragma unroll
      for (int k = 0; k < NM; ++k) {
        const double G = P.gamma[k], Pi = P.pinf[k];
        const double c1 = ar[k] * t0[k];
        const double f = (p0[k] + G * Pi + (G - 1.0) * pe);
        const double g = (pe + G * Pi + (G - 1.0) * pe);
        D += (c1 * (G - 1.0) * g - c1 * f * G) / sq(g);
      }
      return D;
    };
    double pe_old = 0.001, pe_new = pe_old;
    int it = 0;
    double c1 = fabs(pe_new - pe_old), c2 = fabs(Fpe(pe_new));
    while (((c2 >= 1.0e-12) || (c1 >= 1.0e-4)) && (it <= 101)) {
      pe_new = pe_old - 0.8 * (Fpe(pe_old) / DFpe(pe_old));
      c1 = fabs(pe_new - pe_old);
      c2 = fabs(Fpe(pe_new));
      if (it >= 100 && ((fabs(pe_old - pe_new) < 1.0e3) || isnan(pe_new))) { pe_new = NAN; break; }
      pe_old = pe_new;
      it++;
    }
    const double Pp = P.pinf[principal];
    if (pe_new < (-(1.0 - eb) * Pp + eb)) pe_new = -(1.0 - eb) * Pp + 1.0;
    const double pe = pe_new;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      const double G = P.gamma[k], Pi = P.pinf[k];
      const double t1 = t0[k] * (p0[k] + G * Pi + (G - 1.0) * pe) / (pe + G * Pi + (G - 1.0) * pe);
      avf[k] = ar[k] / (1.0 / t1);
    }
    vf_limit_relax<NM>(a, avf);
  } else if (P.prelax == 3) {       // m4_relaxationvec.c:298-479
    const double es = 1.0e-12, eb = 1.0e-8;
    double rk[NM], pk[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      if (avf[k] < 0.0 || ar[k] < 0.0) { avf[k] = es; ar[k] = es; are[k] = es; }
      if (avf[k] >= 1.0) avf[k] = 1.0 - es;
    }
    int irelax = 1;
#pragma unroll
    for (int k = 0; k < NM; ++k) if (avf[k] > 1.0 - es) irelax = 0;
    if (irelax) {
      double pr = 0.0;
#pragma unroll
      for (int k = 0; k < NM; ++k) {
        if (avf[k] > es) {
          pk[k] = (are[k]) / avf[k] * (P.gamma[k] - 1.0) - P.gamma[k] * P.pinf[k];
          if (pk[k] <= -(1.0 - eb) * P.pinf[k] + eb) pk[k] = -(1.0 - eb) * P.pinf[k] + eb;
        } else pk[k] = 0.0;
        pr += avf[k] * pk[k];
      }
      int it = 0;
      double Fv = 1.0e-9, dF = 1.0e9;
#pragma unroll
      for (int k = 0; k < NM; ++k) rk[k] = 0.0;
      while ((fabs(Fv) > 1.0e-12) && (it < 500)) {
        pr = pr - Fv / dF;
        it = it + 1;
#pragma unroll
        for (int k = 0; k < NM; ++k)
          if (pr < -(1.0 - eb) * P.pinf[k] + eb) pr = -(1.0 - eb) * P.pinf[k] + 1.0;
        Fv = -1.0; dF = 0.0;
#pragma unroll
        for (int k = 0; k < NM; ++k) {
          if (avf[k] > es) {
            rk[k] = ar[k] / maxr(avf[k], es) * pow((pr + P.pinf[k]) / (pk[k] + P.pinf[k]), 1.0 / P.gamma[k]);
            Fv = Fv + ar[k] / rk[k];
            dF = dF - ar[k] / (P.gamma[k] * rk[k] * (pr + P.pinf[k]));
          }
        }
      }
#pragma unroll
      for (int k = 0; k < NM; ++k) if (avf[k] > es) avf[k] = ar[k] / rk[k];
      vf_limit_relax<NM>(a, avf);
    }
  }
  if ((NM == 2 && P.prelax == 1) || P.prelax == 2 || P.prelax == 3) {
    double e_hyd = rE / rho - 0.5 * (sq((u)) + sq((v)) + sq((w)));
    double num = 0.0, den = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      num += avf[k] * P.gamma[k] * P.pinf[k] / (P.gamma[k] - 1.0);
      den += avf[k] / (P.gamma[k] - 1.0);
    }
    double pr = (rho * e_hyd - num) / den;
#pragma unroll
    for (int k = 0; k < NM; ++k) are[k] = (avf[k] * (pr + P.gamma[k] * P.pinf[k]) / (P.gamma[k] - 1.0));
  }
#pragma unroll
  for (int k = 0; k < L::N3; ++k) vec[L::I3 + k] = avf[k];
#pragma unroll
  for (int k = 0; k < NM; ++k) vec[L::I4 + k] = are[k];
}

// avf_limit.c:3-21
template <int NM, typename R>
FV_DEV void avf_limit(R a, R *X) {
  using L = Lay<NM>;
  R avf[NM], sum = (R)0.0;
#pragma unroll
  for (int k = 0; k < L::N3; ++k) {
    avf[k] = X[L::I3 + k];
    if (avf[k] <= a) avf[k] = a;
    else if (avf[k] >= ((R)1.0 - a)) avf[k] = (R)1.0 - a;
    sum += avf[k];
  }
  if (sum > ((R)1.0 - a)) {
    R dif = sum - ((R)1.0 - a);
#pragma unroll
    for (int j = 0; j < L::N3; ++j) avf[j] -= (avf[j] / sum) * dif;
  }
#pragma unroll
  for (int k = 0; k < L::N3; ++k) X[L::I3 + k] = avf[k];
}

// MUSCL slope of one variable (m4_reconctruct.c:105-128); limiter 1 is MC
template <typename R>
FV_DEV R limited_slope(int limiter, R df, R db, R dp, R dm) {
  R g = (R)0.0;
  if (fabs(df) > (R)1.0e-10) {
    if (limiter == 0) {
      R product = (-db / dm) * (df / dp);
      if (product > (R)0.0) {
        if (fabs(db / dm) > fabs(df / dp)) g = (R)1.0 * df / dp;
        else g = (R)-1.0 * db / dm;
      } else g = (R)0.0;
    } else {
      R a = -db / dm, b = df / dp;
      if (a * b > (R)0.0) {
        R c = (R)0.5 * (a + b);
        R m2 = minr((R)2.0 * fabs(a), (R)2.0 * fabs(b));
        R mag = minr(m2, (R)fabs(c));
        g = (c > (R)0.0) ? mag : -mag;
      } else g = (R)0.0;
    }
  }
  return g;
}

// m4_cons2primtv.c:3-40: [alpha (N3) | alpha rho_k | u v w | p_k]
template <int NM>
FV_DEV void cons2prim(const Params &P, const double *c, double *q) {
  using L = Lay<NM>;
  double avf[NM], sum = 0.0, rho = 0.0;
#pragma unroll
  for (int k = 0; k < L::N3; ++k) { q[k] = c[L::I3 + k]; avf[k] = c[L::I3 + k]; sum += avf[k]; }
  avf[L::N3] = 1.0 - sum;
#pragma unroll
  for (int k = 0; k < NM; ++k) { q[L::N3 + k] = c[k]; rho += c[k]; }
  constexpr int IU = L::N3 + NM, IP = L::N3 + NM + 3;
  q[IU + 0] = c[L::I1 + 0] / rho;
  q[IU + 1] = c[L::I1 + 1] / rho;
  q[IU + 2] = c[L::I1 + 2] / rho;
#pragma unroll
  for (int k = 0; k < NM; ++k) q[IP + k] = ((P.gamma[k] - 1.0) * (c[L::I4 + k]) / avf[k]) - P.gamma[k] * P.pinf[k];
}

// m4_prim2conc.c:3-74
template <int NM>
FV_DEV void prim2cons(const Params &P, const double *q, double *c) {
  using L = Lay<NM>;
  const double a10 = P.amin / 10.0, a100 = P.amin / 100.0;
  double avf[NM], sum = 0.0, rho = 0.0, sumare = 0.0;
#pragma unroll
  for (int k = 0; k < L::N3; ++k) {
    avf[k] = q[k];
    if (avf[k] <= a10) avf[k] = a10;
    if (avf[k] >= (1.0 - (2.0 * a10))) avf[k] = 1.0 - (2.0 * a10);
    sum += avf[k];
  }
  if (sum >= (1.0 - a100)) {
    const double dif = sum - (1.0 - a100);
#pragma unroll
    for (int j = 0; j < L::N3; ++j) { double per = avf[j] / sum; avf[j] = avf[j] - per * dif; }
    sum = 1.0 - a100;
  }
  avf[L::N3] = 1.0 - sum;
#pragma unroll
  for (int k = 0; k < NM; ++k) { c[k] = q[L::N3 + k]; rho += q[L::N3 + k]; }
  constexpr int IU = L::N3 + NM, IP = L::N3 + NM + 3;
#pragma unroll
  for (int k = 0; k < 3; ++k) c[L::I1 + k] = rho * q[IU + k];
#pragma unroll
  for (int k = 0; k < L::N3; ++k) c[L::I3 + k] = avf[k];
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    c[L::I4 + k] = (avf[k] * (q[IP + k] + P.gamma[k] * P.pinf[k]) / (P.gamma[k] - 1.0));
    sumare += c[L::I4 + k];
  }
  c[L::I2] = sumare + 0.5 * rho * (sq(q[IU + 0]) + sq(q[IU + 1]) + sq(q[IU + 2]));
}

// PRIMTV=1 face state before relaxation (m4_reconctruct.c:93-181): deltas and
// slopes of the primitive variables, the face value, the volume-fraction clip,
// m4_prim2conc.  Sc, Xf, Xo are the avf-limited cell and stencil values.
template <int NM>
FV_DEV void recon_prim_face(const Params &P, double a, int limiter, bool bnd_f, const double *Sc, const double *Xf,
                            const double *Xo, double dcf, double dp, double dm, double *F) {
  constexpr int NP = 3 * NM + 2;
  double qc[NP], qr[NP];
  cons2prim<NM>(P, Sc, qc);
  {
    double qf[NP], qo[NP];
    cons2prim<NM>(P, Xf, qf);
    cons2prim<NM>(P, Xo, qo);
#pragma unroll
    for (int v = 0; v < NP; ++v) {
      double g = 0.0;
      if (!bnd_f) g = limited_slope<double>(limiter, qf[v] - qc[v], qo[v] - qc[v], dp, dm);
      qr[v] = qc[v] + dcf * g;
    }
  }
  double avf[NM], sum = 0.0;
#pragma unroll
  for (int k = 0; k < NM - 1; ++k) {
    avf[k] = qr[k];
    if (avf[k] <= a) avf[k] = a;
    else if (avf[k] >= (1.0 - (2.0 * a))) avf[k] = 1.0 - (2.0 * a);
    sum += avf[k];
  }
  if (sum >= (1.0 - a)) {
    const double dif = sum - (1.0 - a);
#pragma unroll
    for (int j = 0; j < NM - 1; ++j) { double per = avf[j] / sum; avf[j] = avf[j] - per * dif; }
  }
#pragma unroll
  for (int k = 0; k < NM - 1; ++k) qr[k] = avf[k];
  prim2cons<NM>(P, qr, F);
}

// VF limit of a reconstructed face state (m4_reconctruct.c:147-171)
template <int NM, typename R>
FV_DEV void vf_limit_face(R a, R *SF) {
  using L = Lay<NM>;
  R avf[NM], sum = (R)0.0;
#pragma unroll
  for (int k = 0; k < L::N3; ++k) {
    avf[k] = SF[L::I3 + k];
    if (avf[k] <= a) avf[k] = a;
    else if (avf[k] >= ((R)1.0 - ((R)2.0 * a))) avf[k] = (R)1.0 - ((R)2.0 * a);
    sum += avf[k];
  }
  if (sum >= ((R)1.0 - a)) {
    R dif = sum - ((R)1.0 - a);
#pragma unroll
    for (int j = 0; j < L::N3; ++j) { R per = avf[j] / sum; avf[j] = avf[j] - per * dif; }
  }
#pragma unroll
  for (int k = 0; k < L::N3; ++k) SF[L::I3 + k] = avf[k];
}

}  // namespace fvk
