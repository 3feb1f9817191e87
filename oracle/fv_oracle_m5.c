/*
 * fv_oracle_m5.c -- CPU restatement of ForestFV's model-5 stage.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP model-5 kernels are tested
 * against (see fv_oracle_m5.h).  Built into libfvoracle.so with gcc -O2 -ffp-contract=off.
 *
 * Every function names the reference lines it restates and keeps their
 * operation order.  pow(x,2.0) is written x*x as in fv_oracle.c; the
 * pow(|G|,1/3) and pow(|G|,1/6) of the elastic EOS stay glibc pow calls.
 */
#include "fv_oracle_m5.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SQ(x) ((x) * (x))
#define MAXF(a, b) (((a) > (b)) ? (a) : (b))   /* strdata.h:27 */
#define MINF(a, b) (((a) < (b)) ? (a) : (b))   /* strdata.h:31 */

int orc_m5_neq(int nmat) { return nmat + 3 + 1 + (nmat - 1) + nmat + 9; }   /* m5_input.c */

typedef struct {
  int nmat, neq, ncons;
  int i0, i1, i2, i3, i4, i5;   /* eqtypi */
  int n0, n3, n4;               /* eqtypn[0], [3], [4] */
} lay5_t;

static lay5_t mk_lay5(int nmat) {
  lay5_t L;
  L.nmat = nmat;
  L.neq = orc_m5_neq(nmat);
  L.ncons = nmat + 3 + 1;                 /* NCONSEQ */
  L.n0 = nmat; L.n3 = nmat - 1; L.n4 = nmat;
  L.i0 = 0; L.i1 = nmat; L.i2 = L.i1 + 3; L.i3 = L.i2 + 1; L.i4 = L.i3 + L.n3; L.i5 = L.i4 + L.n4;
  return L;
}

typedef struct {
  double nx, ny, nz, mx, my, mz, lx, ly, lz;   /* n, t1, t2 */
} fr5_t;

typedef struct {
  double r, e, u, v, w, phi0;
  double ra[ORC_MAXMAT], avf[ORC_MAXMAT], are[ORC_MAXMAT], c2Y[ORC_MAXMAT], p_hydro[ORC_MAXMAT];
  double Amat[9], st[3][3];
  double ast[ORC_MAXMAT][3][3];   /* avf_stress of each material */
  double vec[ORC_M5_MAXNEQ];
  double u_temp, v_temp, w_temp;
} rec5_t;

/* ------------------------------------------------------------------ */
/* elastic EOS (state_equation.c)                                       */
/* ------------------------------------------------------------------ */
/* m5_elastic_energy_solid_V2, state_equation.c:269-316 */
static double elastic_energy_v2(const double *G, double Gmag, double mu) {
  double eints;
  if (fabs(Gmag) != 0.0) {
    double M1[9], M2[9] = {0};
    double Gmag13 = pow(fabs(Gmag), 1.0 / 3.0);
    if (isnan(Gmag13)) Gmag13 = 0.0;
    M1[0] = G[0] - Gmag13; M1[1] = G[1]; M1[2] = G[2];
    M1[3] = G[3]; M1[4] = G[4] - Gmag13; M1[5] = G[5];
    M1[6] = G[6]; M1[7] = G[7]; M1[8] = G[8] - Gmag13;
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) M2[i * 3 + j] += M1[i * 3 + k] * M1[k * 3 + j];
    double tr1 = M2[0] + M2[4] + M2[8];
    eints = mu / (4.0 * pow(fabs(Gmag), 1.0 / 6.0)) * tr1;
  } else {
    eints = 0.0;
  }
  return eints;
}

/* G_det, state_equation.c:206-215, and m5_elastic_energy_solid, :217-267 */
static double elastic_energy_v1(double G[3][3], double mu) {
  double Gmag = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
  double eints;
  if ((fabs(Gmag) != 0.0) && (mu != 0.0)) {
    double M1[3][3], M2[3][3] = {{0}};
    double Gmag13 = pow(fabs(Gmag), 1.0 / 3.0);
    if (isnan(Gmag13)) Gmag13 = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M1[i][j] = G[i][j];
    M1[0][0] = G[0][0] - Gmag13; M1[1][1] = G[1][1] - Gmag13; M1[2][2] = G[2][2] - Gmag13;
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) M2[i][j] += M1[i][k] * M1[k][j];
    double tr1 = M2[0][0] + M2[1][1] + M2[2][2];
    eints = mu / (4.0 * pow(fabs(Gmag), 1.0 / 6.0)) * tr1;
  } else {
    eints = 0.0;
  }
  return eints;
}

/* m5_eos_stress_V2, state_equation.c:143-204 */
static void eos_stress_v2(double s[3][3], double avf, double pr, const double *G, double Gmag, double mu) {
  if ((fabs(Gmag) != 0.0) && (mu != 0.0)) {
    double G2[9] = {0}, ss[9];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) G2[i * 3 + j] += G[i * 3 + k] * G[k * 3 + j];
    double J1 = G[0] + G[4] + G[8];
    double J2 = G2[0] + G2[4] + G2[8];
    ss[0] = G2[0] - J2 / 3.0 - pow(fabs(Gmag), 1.0 / 3.0) * (G[0] - J1 / 3.0);
    ss[1] = G2[1] - pow(fabs(Gmag), 1.0 / 3.0) * (G[1]);
    ss[2] = G2[2] - pow(fabs(Gmag), 1.0 / 3.0) * (G[2]);
    ss[3] = G2[3] - pow(fabs(Gmag), 1.0 / 3.0) * (G[3]);
    ss[4] = G2[4] - J2 / 3.0 - pow(fabs(Gmag), 1.0 / 3.0) * (G[4] - J1 / 3.0);
    ss[5] = G2[5] - pow(fabs(Gmag), 1.0 / 3.0) * (G[5]);
    ss[6] = G2[6] - pow(fabs(Gmag), 1.0 / 3.0) * (G[6]);
    ss[7] = G2[7] - pow(fabs(Gmag), 1.0 / 3.0) * (G[7]);
    ss[8] = G2[8] - J2 / 3.0 - pow(fabs(Gmag), 1.0 / 3.0) * (G[8] - J1 / 3.0);
    s[0][0] = -avf * pr - (mu / pow(fabs(Gmag), 1.0 / 6.0)) * (ss[0]);
    s[0][1] = -(mu / pow(fabs(Gmag), 1.0 / 6.0)) * (ss[1]);
    s[0][2] = -(mu / pow(fabs(Gmag), 1.0 / 6.0)) * (ss[2]);
    s[1][0] = -(mu / pow(fabs(Gmag), 1.0 / 6.0)) * (ss[3]);
    s[1][1] = -avf * pr - (mu / pow(fabs(Gmag), 1.0 / 6.0)) * (ss[4]);
    s[1][2] = -(mu / pow(fabs(Gmag), 1.0 / 6.0)) * (ss[5]);
    s[2][0] = -(mu / pow(fabs(Gmag), 1.0 / 6.0)) * (ss[6]);
    s[2][1] = -(mu / pow(fabs(Gmag), 1.0 / 6.0)) * (ss[7]);
    s[2][2] = -avf * pr - (mu / pow(fabs(Gmag), 1.0 / 6.0)) * (ss[8]);
  } else {
    s[0][0] = -avf * pr; s[0][1] = 0.0; s[0][2] = 0.0;
    s[1][0] = 0.0; s[1][1] = -avf * pr; s[1][2] = 0.0;
    s[2][0] = 0.0; s[2][1] = 0.0; s[2][2] = -avf * pr;
  }
}

/* ------------------------------------------------------------------ */
/* face records                                                         */
/* ------------------------------------------------------------------ */
/* face = 0: m5_facevalues.c:3-188, the record of a cell state S (C = S).
 * face = 1: m5_facercnstr.c:3-190, the record of a face state S whose phi
 * comes from the cell state C; its pressure and c2Y divide by the re-summed
 * fraction r->avf[i] where m5_facevalues divides by avf_temp[i] (the two
 * differ only for the last material after the renormalisation).  The two
 * clips of avf_c_temp come in opposite orders in the two files, which gives
 * the same value. */
static void m5_record_x(const lay5_t *L, const orc_m5_params *p, const double *S, const double *C, int face,
                        rec5_t *r) {
  double avf_temp[ORC_MAXMAT + 1], avf_c_temp[ORC_MAXMAT + 1];
  const double a = p->amin / 100.0;
  r->r = 0.0;
  for (int i = 0; i < L->n0; ++i) { r->r += S[i]; r->ra[i] = S[i]; r->vec[i] = S[i]; }
  r->u = S[L->i1 + 0] / r->r;
  r->v = S[L->i1 + 1] / r->r;
  r->w = S[L->i1 + 2] / r->r;
  r->vec[L->i1 + 0] = S[L->i1 + 0]; r->vec[L->i1 + 1] = S[L->i1 + 1]; r->vec[L->i1 + 2] = S[L->i1 + 2];
  r->e = S[L->i2] / r->r;
  r->vec[L->i2] = S[L->i2];
  for (int i = 0; i < L->n3; ++i) { avf_temp[i] = S[L->i3 + i]; avf_c_temp[i] = C[L->i3 + i]; }
  double avfsum = 0.0;
  for (int i = 0; i < L->n3; ++i) {
    if (avf_temp[i] <= a) avf_temp[i] = a;
    else if (avf_temp[i] >= (1.0 - a)) avf_temp[i] = 1.0 - a;
    if (avf_c_temp[i] >= (1.0 - a)) avf_c_temp[i] = 1.0 - a;
    else if (avf_c_temp[i] <= a) avf_c_temp[i] = a;
    avfsum += avf_temp[i];
  }
  if (avfsum > (1.0 - a)) {
    double avf_dif = avfsum - (1.0 - a);
    for (int j = 0; j < L->n3; ++j) {
      double avf_per = avf_temp[j] / avfsum;
      avf_temp[j] = avf_temp[j] - avf_per * avf_dif;
    }
    avfsum = 1.0 - a;
  }
  avf_temp[L->n3] = 1.0 - avfsum;
  avfsum = 0.0;
  for (int i = 0; i < L->n3; ++i) { r->avf[i] = avf_temp[i]; r->vec[L->i3 + i] = avf_temp[i]; avfsum += avf_temp[i]; }
  r->avf[L->n3] = 1.0 - avfsum;
  r->phi0 = avf_c_temp[0] - p->avf_lim;   /* n_solids is 0: the override branches never run */
  for (int i = 0; i < L->n4; ++i) { r->are[i] = S[L->i4 + i]; r->vec[L->i4 + i] = S[L->i4 + i]; }
  double A[9], G[9];
  for (int k = 0; k < 9; ++k) { A[k] = S[L->i5 + k]; r->Amat[k] = S[L->i5 + k]; r->vec[L->i5 + k] = S[L->i5 + k]; }
  G[0] = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  G[1] = A[0] * A[3] + A[1] * A[4] + A[2] * A[5];
  G[2] = A[0] * A[6] + A[1] * A[7] + A[2] * A[8];
  G[3] = A[3] * A[0] + A[4] * A[1] + A[5] * A[2];
  G[4] = A[3] * A[3] + A[4] * A[4] + A[5] * A[5];
  G[5] = A[3] * A[6] + A[4] * A[7] + A[5] * A[8];
  G[6] = A[6] * A[0] + A[7] * A[1] + A[8] * A[2];
  G[7] = A[6] * A[3] + A[7] * A[4] + A[8] * A[5];
  G[8] = A[6] * A[6] + A[7] * A[7] + A[8] * A[8];
  double Gmag = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) + G[2] * (G[3] * G[7] - G[4] * G[6]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r->st[i][j] = 0.0;
  const int mix = p->fix & ORC_M5_FIX_MIXTURE;
  for (int i = 0; i < L->n4; ++i) {
    double are_elast = (p->mush[i] != 0.0) ? elastic_energy_v2(G, Gmag, p->mush[i]) : 0.0;
    double g = p->gamma[i], pi = p->pinf[i];
    const double den = face ? r->avf[i] : avf_temp[i];
    /* mixture mode: phase k holds alpha_k of the elastic energy and of the shear stress */
    const double mu_k = mix ? r->avf[i] * p->mush[i] : p->mush[i];
    if (mix) are_elast *= den;
    double press = ((g - 1.0) * (r->are[i] - are_elast) / den) - g * pi;
    r->c2Y[i] = (den / r->r) * g * (press + pi);
    if (mix) r->c2Y[i] += (4.0 / 3.0) * mu_k / r->r;   /* longitudinal speed of the mixture */
    r->p_hydro[i] = press;
    eos_stress_v2(r->ast[i], r->avf[i], r->p_hydro[i], G, Gmag, mu_k);
    for (int a1 = 0; a1 < 3; ++a1)
      for (int b = 0; b < 3; ++b) r->st[a1][b] += r->ast[i][a1][b];
  }
}

static void m5_record(const lay5_t *L, const orc_m5_params *p, const double *S, rec5_t *r) {
  m5_record_x(L, p, S, S, 0, r);
}

void orc_m5_record(const orc_m5_params *p, const double *vec, double *out) {
  lay5_t L = mk_lay5(p->nmat);
  rec5_t r;
  memset(&r, 0, sizeof r);
  m5_record(&L, p, vec, &r);
  memset(out, 0, sizeof(double) * ORC_M5_REC);
  out[0] = r.r; out[1] = r.e; out[2] = r.u; out[3] = r.v; out[4] = r.w; out[5] = r.phi0;
  for (int i = 0; i < p->nmat; ++i) {
    out[6 + i] = r.ra[i]; out[9 + i] = r.avf[i]; out[12 + i] = r.are[i];
    out[15 + i] = r.c2Y[i]; out[18 + i] = r.p_hydro[i];
  }
  for (int k = 0; k < 9; ++k) { out[21 + k] = r.Amat[k]; out[30 + k] = r.st[k / 3][k % 3]; }
}

/* ------------------------------------------------------------------ */
/* GFM face flux                                                        */
/* ------------------------------------------------------------------ */
/* rotate_vectors.c:128-188: t = vec . st . vec^T with vec rows (n), (m), (l) */
static void rot_tensor(double t[9], double st[3][3], double nx, double ny, double nz, double mx, double my, double mz,
                       double lx, double ly, double lz) {
  double vec[3][3] = {{nx, ny, nz}, {mx, my, mz}, {lx, ly, lz}};
  double vecT[3][3], M1[3][3], M2[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { vecT[j][i] = vec[i][j]; M1[i][j] = 0.0; M2[i][j] = 0.0; }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) M1[i][j] += vec[i][k] * st[k][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) M2[i][j] += M1[i][k] * vecT[k][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[3 * i + j] = M2[i][j];
}

typedef struct {
  double flux_adv[ORC_M5_MAXNEQ][3];
  double flux_hllc[ORC_M5_MAXNEQ];
  double Us[ORC_M5_MAXNEQ];
  double u_nc, v_nc, w_nc;
  double avf_S[ORC_MAXMAT], are_S[ORC_MAXMAT], Amat_S[3][3];
} ws5_t;

/* m5_flux_gfm.c:3-87 */
static void m5_flux_gfm(const lay5_t *L, const rec5_t *s, int flag_ifs, ws5_t *W) {
  for (int v = 0; v < L->n0; ++v) {
    int iv = v + L->i0;
    W->flux_adv[iv][0] = s->u_temp * s->ra[v];
    W->flux_adv[iv][1] = s->v_temp * s->ra[v];
    W->flux_adv[iv][2] = s->w_temp * s->ra[v];
  }
  int iv = L->i1;
  W->flux_adv[iv][0] = s->u_temp * s->u_temp * s->r - s->st[0][0];
  W->flux_adv[iv][1] = s->v_temp * s->u_temp * s->r - s->st[0][1];
  W->flux_adv[iv][2] = s->w_temp * s->u_temp * s->r - s->st[0][2];
  iv = L->i1 + 1;
  W->flux_adv[iv][0] = s->u_temp * s->v_temp * s->r - s->st[1][0];
  W->flux_adv[iv][1] = s->v_temp * s->v_temp * s->r - s->st[1][1];
  W->flux_adv[iv][2] = s->w_temp * s->v_temp * s->r - s->st[1][2];
  iv = L->i1 + 2;
  W->flux_adv[iv][0] = s->u_temp * s->w_temp * s->r - s->st[2][0];
  W->flux_adv[iv][1] = s->v_temp * s->w_temp * s->r - s->st[2][1];
  W->flux_adv[iv][2] = s->w_temp * s->w_temp * s->r - s->st[2][2];
  iv = L->i2;
  W->flux_adv[iv][0] = s->r * s->e * s->u_temp - s->st[0][0] * s->u_temp - s->st[0][1] * s->v_temp - s->st[0][2] * s->w_temp;
  W->flux_adv[iv][1] = s->r * s->e * s->v_temp - s->st[1][0] * s->u_temp - s->st[1][1] * s->v_temp - s->st[1][2] * s->w_temp;
  W->flux_adv[iv][2] = s->r * s->e * s->w_temp - s->st[2][0] * s->u_temp - s->st[2][1] * s->v_temp - s->st[2][2] * s->w_temp;
  if ((flag_ifs == 0) || (flag_ifs == 3)) {
    W->u_nc = s->u_temp; W->v_nc = s->v_temp; W->w_nc = s->w_temp;
    for (int v = 0; v < L->n3; ++v) W->avf_S[v] = s->avf[v];
    for (int v = 0; v < L->n4; ++v) W->are_S[v] = s->are[v];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) W->Amat_S[i][j] = s->Amat[i * 3 + j];
  }
  for (int v = 0; v < L->n3; ++v) {
    iv = v + L->i3;
    W->flux_adv[iv][0] = W->u_nc * W->avf_S[v];
    W->flux_adv[iv][1] = W->v_nc * W->avf_S[v];
    W->flux_adv[iv][2] = W->w_nc * W->avf_S[v];
  }
  for (int v = 0; v < L->n4; ++v) {
    iv = v + L->i4;
    W->flux_adv[iv][0] = W->u_nc * W->are_S[v];
    W->flux_adv[iv][1] = W->v_nc * W->are_S[v];
    W->flux_adv[iv][2] = W->w_nc * W->are_S[v];
  }
  iv = L->i5;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      W->flux_adv[iv][0] = W->u_nc * W->Amat_S[i][j];
      W->flux_adv[iv][1] = W->v_nc * W->Amat_S[i][j];
      W->flux_adv[iv][2] = W->w_nc * W->Amat_S[i][j];
      iv++;
    }
}

/* m5_star_region.c:3-157 */
static int m5_star_region(const lay5_t *L, const orc_m5_params *p, const fr5_t *F, const rec5_t *sL,
                          const rec5_t *sR, const rec5_t *s, int flag_ifs, double u_S, double SL, double SR,
                          double uL, double vL, double wL, double uR, double vR, double wR, double s11L,
                          double s12L, double s13L, double s11R, double s12R, double s13R, ws5_t *W) {
  double SK, uK, vK, wK, s11K, s12K, s13K;
  if (flag_ifs == 1) { SK = SL; uK = uL; vK = vL; wK = wL; s11K = s11L; s12K = s12L; s13K = s13L; }
  else { SK = SR; uK = uR; vK = vR; wK = wR; s11K = s11R; s12K = s12R; s13K = s13R; }
  double ar_S[ORC_MAXMAT], r_S = 0.0;
  for (int v = 0; v < L->n0; ++v) { ar_S[v] = s->ra[v] * (SK - uK) / (SK - u_S); r_S += ar_S[v]; }
  double s11_S = ((uR - SR) * sR->r * (s11L) - (uL - SL) * sL->r * (s11R) + (uL - SL) * sL->r * (uR - SR) * sR->r * (uR - uL)) /
                 ((uR - SR) * sR->r - (uL - SL) * sL->r);
  double s12_S, s13_S;
  if ((sL->phi0 > 0.0) && (sR->phi0 > 0.0)) {
    s12_S = ((uR - SR) * sR->r * (s12L) - (uL - SL) * sL->r * (s12R) + (uL - SL) * sL->r * (uR - SR) * sR->r * (vR - vL)) /
            ((uR - SR) * sR->r - (uL - SL) * sL->r);
    s13_S = ((uR - SR) * sR->r * (s13L) - (uL - SL) * sL->r * (s13R) + (uL - SL) * sL->r * (uR - SR) * sR->r * (wR - wL)) /
            ((uR - SR) * sR->r - (uL - SL) * sL->r);
  } else {
    s12_S = 0.0;
    s13_S = 0.0;
  }
  double v_S = vK + (s12_S - s12K) / ((uK - SK) * s->r);
  double w_S = wK + (s13_S - s13K) / ((uK - SK) * s->r);
  double vel[3];
  vel[0] = u_S * F->nx + v_S * F->mx + w_S * F->lx;     /* rotate_vector_ltg_v2 */
  vel[1] = u_S * F->ny + v_S * F->my + w_S * F->ly;
  vel[2] = u_S * F->nz + v_S * F->mz + w_S * F->lz;
  double re = s->r * s->e;
  const int fix = p->fix & ORC_M5_FIX_SIGNS;
  double Em_S;
  if (fix)
    Em_S = (re * (uK - SK) - s11K * uK - s12K * vK - s13K * wK + s11_S * u_S + s12_S * v_S + s13_S * w_S) /
           (r_S * (u_S - SK));
  else
    Em_S = (re * (uK - SK) + s11K * uK - s12K * vK - s13K * wK + s11_S * u_S + s12_S * v_S + s13_S * w_S) /
           (r_S * (u_S + SK));
  double A[3][3], t[9], AS[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = s->Amat[i * 3 + j];
  rot_tensor(t, A, F->nx, F->ny, F->nz, F->mx, F->my, F->mz, F->lx, F->ly, F->lz);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = t[i * 3 + j];
  for (int j = 0; j < 3; ++j)
    AS[0][j] = fix ? (A[0][j] * (uK - SK) + A[1][j] * (vK - v_S) + A[2][j] * (wK - w_S)) / (u_S - SK)
                   : (A[0][j] * (uK - SK) - A[1][j] * (vK - v_S) + A[2][j] * (wK - w_S)) / (u_S - SK);
  for (int j = 0; j < 3; ++j) { AS[1][j] = A[1][j]; AS[2][j] = A[2][j]; }
  rot_tensor(t, AS, F->nx, F->mx, F->lx, F->ny, F->my, F->ly, F->nz, F->mz, F->lz);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) AS[i][j] = t[i * 3 + j];
  double G[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) G[i][j] = AS[i][0] * AS[j][0] + AS[i][1] * AS[j][1] + AS[i][2] * AS[j][2];
  double e_S[ORC_MAXMAT];
  for (int v = 0; v < L->n4; ++v) {
    double g = p->gamma[v], pi = p->pinf[v];
    double rm = (s->ra[v] / s->avf[v]);
    double rm_S = (s->ra[v] / s->avf[v]) * (uK - SK) / (u_S - SK);
    double p_k = fix ? s->p_hydro[v] : -s->p_hydro[v];
    double p_S = (p_k + pi) * (((g - 1.0) * rm + (g + 1.0) * rm_S) / ((g - 1.0) * rm_S + (g + 1.0) * rm)) - pi;
    double are_el = elastic_energy_v1(G, p->mush[v]);
    if (p->fix & ORC_M5_FIX_MIXTURE) are_el *= s->avf[v];
    double e_el = are_el / (s->avf[v] * rm_S);
    e_S[v] = (p_S + g * pi) / (rm_S * (fix ? (g - 1.0) : (g + 1.0))) + e_el;
  }
  for (int v = 0; v < L->n0; ++v) W->Us[L->i0 + v] = ar_S[v];
  for (int v = 0; v < 3; ++v) W->Us[L->i1 + v] = r_S * vel[v];
  W->Us[L->i2] = r_S * Em_S;
  W->u_nc = vel[0]; W->v_nc = vel[1]; W->w_nc = vel[2];
  for (int v = 0; v < L->n3; ++v) W->avf_S[v] = s->avf[v];
  for (int v = 0; v < L->n4; ++v) W->are_S[v] = ar_S[v] * e_S[v];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) W->Amat_S[i][j] = AS[i][j];
  for (int iv = 0; iv < L->ncons; ++iv)
    if (isnan(W->Us[iv])) return -1;
  return 0;
}

/* m5_hllc_gfm.c:3-100 with m5_wave_speeds_gfm.c:3-58 */
static int m5_hllc_gfm(const lay5_t *L, const orc_m5_params *p, const fr5_t *F, rec5_t *sL, rec5_t *sR, ws5_t *W) {
  double u, v, w;
  u = sL->u; v = sL->v; w = sL->w;
  double uL = u * F->nx + v * F->ny + w * F->nz;          /* rotate_vector_gtl_v2 */
  double vL = u * F->mx + v * F->my + w * F->mz;
  double wL = u * F->lx + v * F->ly + w * F->lz;
  u = sR->u; v = sR->v; w = sR->w;
  double uR = u * F->nx + v * F->ny + w * F->nz;
  double vR = u * F->mx + v * F->my + w * F->mz;
  double wR = u * F->lx + v * F->ly + w * F->lz;
  vR = vL;                                                 /* ghost: tangential velocity of the left side */
  wR = wL;
  sL->u_temp = sL->u; sL->v_temp = sL->v; sL->w_temp = sL->w;
  sR->u_temp = uR * F->nx + vR * F->mx + wR * F->lx;       /* rotate_vector_ltg_v2 */
  sR->v_temp = uR * F->ny + vR * F->my + wR * F->ly;
  sR->w_temp = uR * F->nz + vR * F->mz + wR * F->lz;
  sR->vec[L->i1 + 0] = sR->u_temp * sR->r;
  sR->vec[L->i1 + 1] = sR->v_temp * sR->r;
  sR->vec[L->i1 + 2] = sR->w_temp * sR->r;
  sR->vec[L->i2] = sR->r * sR->e;
  /* m5_wave_speeds_gfm */
  double ruL = sL->r * (uL), ruR = sR->r * (uR);
  double c2YL = 0.0, c2YR = 0.0;
  for (int iv = 0; iv < L->nmat; ++iv) { c2YL += sL->c2Y[iv]; c2YR += sR->c2Y[iv]; }
  double cL = sqrt(c2YL), cR = sqrt(c2YR);
  double SR = MAXF((uL + cL), (uR + cR));
  double SL = MINF((uL - cL), (uR - cR));
  double t[9];
  rot_tensor(t, sL->st, F->nx, F->ny, F->nz, F->mx, F->my, F->mz, F->lx, F->ly, F->lz);
  double s11L = t[0], s12L = t[1], s13L = t[2];
  rot_tensor(t, sR->st, F->nx, F->ny, F->nz, F->mx, F->my, F->mz, F->lx, F->ly, F->lz);
  double s11R = t[0], s12R = t[1], s13R = t[2];
  double u_S = ((sL->r * SQ((uL)) - (s11L)) - (sR->r * SQ((uR)) - (s11R)) - ((SL)*ruL) + ((SR)*ruR)) /
               (ruL - ruR - ((SL)*sL->r) + ((SR)*sR->r));
  if (isnan(u_S) || isnan(SR) || isnan(SL)) return -1;
  for (int iv = 0; iv < L->neq; ++iv) W->flux_hllc[iv] = 0.0;
  int end = 0;
  if ((0.0 <= SL) && !end) { end = 1; m5_flux_gfm(L, sL, 0, W); }
  if (((SL < 0.0) && (0.0 <= u_S)) && !end) {
    end = 1;
    if (m5_star_region(L, p, F, sL, sR, sL, 1, u_S, SL, SR, uL, vL, wL, uR, vR, wR, s11L, s12L, s13L, s11R, s12R,
                       s13R, W))
      return -1;
    m5_flux_gfm(L, sL, 1, W);
    for (int iv = 0; iv < L->ncons; ++iv) W->flux_hllc[iv] = SL * (W->Us[iv] - sL->vec[iv]);
  }
  if (((u_S < 0.0) && (0.0 < SR)) && !end) {
    end = 1;
    if (m5_star_region(L, p, F, sL, sR, sR, 2, u_S, SL, SR, uL, vL, wL, uR, vR, wR, s11L, s12L, s13L, s11R, s12R,
                       s13R, W))
      return -1;
    m5_flux_gfm(L, sR, 2, W);
    for (int iv = 0; iv < L->ncons; ++iv) W->flux_hllc[iv] = SR * (W->Us[iv] - sR->vec[iv]);
  }
  if ((SR <= 0.0) && !end) { end = 1; m5_flux_gfm(L, sR, 3, W); }
  for (int iv = 0; iv < L->neq; ++iv)
    if (isnan(W->flux_hllc[iv])) return -1;
  return 0;
}

int orc_m5_gfm_face(const orc_m5_params *p, const double *UL, const double *UR, const double *frame, double area,
                    double *flux, double *unc) {
  lay5_t L = mk_lay5(p->nmat);
  fr5_t F = {frame[0], frame[1], frame[2], frame[3], frame[4], frame[5], frame[6], frame[7], frame[8]};
  rec5_t sL, sR;
  ws5_t W;
  memset(&sL, 0, sizeof sL); memset(&sR, 0, sizeof sR); memset(&W, 0, sizeof W);
  m5_record(&L, p, UL, &sL);
  m5_record(&L, p, UR, &sR);
  if (m5_hllc_gfm(&L, p, &F, &sL, &sR, &W)) return -1;
  /* m5_residual.c:166-170 with NS = 0 (no viscous flux) on a zeroed flux_face */
  for (int iv = 0; iv < L.neq; ++iv) {
    double fl = 0.0;
    fl += (area) * (((W.flux_adv[iv][0] - 0.0) * F.nx + (W.flux_adv[iv][1] - 0.0) * F.ny +
                     (W.flux_adv[iv][2] - 0.0) * F.nz) + W.flux_hllc[iv]);
    flux[iv] = fl;
  }
  unc[0] = W.u_nc; unc[1] = W.v_nc; unc[2] = W.w_nc;
  return 0;
}

/* ------------------------------------------------------------------ */
/* velocity re-projection (m5_gfm.c, m5_gfm_rel.c) and m5_RK            */
/* ------------------------------------------------------------------ */
/* the part shared by m5_gfm.c:76-196 and m5_gfm_rel.c:60-219: interface
 * normal, tangent frame, and the new global velocity; returns f_inter */
static int gfm_velocity(const lay5_t *L, const orc_m5_params *p, const orc_mesh *m, const orc_cell *cl, int c,
                        const double *Sown, const double *const *Snb, const unsigned char *gf, int accumulate,
                        double ut, double vt, double wt, double ng[3], int *err) {
  double nxi = 0.0, nyi = 0.0, nzi = 0.0, total_area = 0.0;
  double us[6], vs[6], ws[6], ar[6];
  int fi = 0;
  *err = 0;
  for (int f = 0; f < 6; f++) {
    if (!gf[f]) continue;
    if (m->bc[6 * c + f] != 0) { *err = -2; return 0; }
    rec5_t rL, rR;
    m5_record(L, p, Sown, &rL);
    m5_record(L, p, Snb[f], &rR);
    double phi = (rL.phi0 + rR.phi0) / 2.0;
    const double A = cl->Area[f][0];
    if (accumulate) {
      nxi += phi * A * cl->nx[f];
      nyi += phi * A * cl->ny[f];
      nzi += phi * A * cl->nz[f];
    } else {
      nxi = phi * A * cl->nx[f];
      nyi = phi * A * cl->ny[f];
      nzi = phi * A * cl->nz[f];
    }
    us[fi] = rR.u; vs[fi] = rR.v; ws[fi] = rR.w; ar[fi] = A;
    total_area += A;
    fi++;
  }
  if (fi == 0) return 0;
  double norm = sqrt(SQ(nxi) + SQ(nyi) + SQ(nzi));
  nxi = -nxi / norm; nyi = -nyi / norm; nzi = -nzi / norm;
  double nx1, ny1, nz1;
  if (fabs(nzi) < 0.99) { nx1 = nyi; ny1 = -nxi; nz1 = 0.0; }
  else { nx1 = nzi; ny1 = 0.0; nz1 = -nxi; }
  double nx2 = nyi * nz1 - nzi * ny1;
  double ny2 = nzi * nx1 - nxi * nz1;
  double nz2 = nxi * ny1 - nyi * nx1;
  /* rotate_vector (rotate_vectors.c:3-13) */
  double u_lc = ut * nxi + vt * nyi + wt * nzi;
  double v_new = 0.0, w_new = 0.0;
  for (int k = 0; k < fi; k++) {
    double v_lc = us[k] * nx1 + vs[k] * ny1 + ws[k] * nz1;
    double w_lc = us[k] * nx2 + vs[k] * ny2 + ws[k] * nz2;
    v_new = v_lc * (ar[k] / total_area);       /* overwritten per interface face, as written */
    w_new = w_lc * (ar[k] / total_area);
  }
  ng[0] = u_lc * nxi + v_new * nx1 + w_new * nx2;
  ng[1] = u_lc * nyi + v_new * ny1 + w_new * ny2;
  ng[2] = u_lc * nzi + v_new * nz1 + w_new * nz2;
  return fi;
}

/* m5_gfm.c:3-254 on leaf c; S is the state array updated in leaf order */
static int m5_gfm(const lay5_t *L, const orc_m5_params *p, const orc_mesh *m, const orc_cell *cl, int c, double *S,
                  const double *RHS, const unsigned char *gf) {
  const int neq = L->neq;
  double *Sc = S + (size_t)c * neq;
  const double *R = RHS + (size_t)c * neq;
  const double dV = ((p->fix & ORC_M5_FIX_SIGNS) ? -p->dt : p->dt) / cl->Vol;
  double rho_new = 0.0;
  for (int v = 0; v < L->n0; ++v) rho_new += Sc[L->i0 + v] + (dV) * R[L->i0 + v];
  double ut = Sc[L->i1 + 0] + (dV) * R[L->i1 + 0];
  double vt = Sc[L->i1 + 1] + (dV) * R[L->i1 + 1];
  double wt = Sc[L->i1 + 2] + (dV) * R[L->i1 + 2];
  ut = ut / rho_new; vt = vt / rho_new; wt = wt / rho_new;
  const double *Snb[6];
  for (int f = 0; f < 6; f++) {
    int j = m->bc[6 * c + f] == 0 ? m->neigtr[24 * c + 4 * f] : c;
    Snb[f] = S + (size_t)j * neq;
  }
  double ng[3];
  int err;
  int fi = gfm_velocity(L, p, m, cl, c, Sc, Snb, gf + 6 * c, 0, ut, vt, wt, ng, &err);
  if (err) return err;
  if (fi != 0) {
    for (int v = 0; v < L->n0; ++v) Sc[L->i0 + v] = Sc[L->i0 + v] + (dV) * R[L->i0 + v];
    Sc[L->i1 + 0] = rho_new * ng[0];
    Sc[L->i1 + 1] = rho_new * ng[1];
    Sc[L->i1 + 2] = rho_new * ng[2];
    Sc[L->i2] -= rho_new * 0.5 * (SQ(ut) + SQ(vt) + SQ(wt));
    Sc[L->i2] += rho_new * 0.5 * (SQ(ng[0]) + SQ(ng[1]) + SQ(ng[2]));
  } else {
    for (int v = 0; v < L->n0; ++v) Sc[L->i0 + v] = Sc[L->i0 + v] + (dV) * R[L->i0 + v];
    Sc[L->i1 + 0] = Sc[L->i1 + 0] + (dV) * R[L->i1 + 0];
    Sc[L->i1 + 1] = Sc[L->i1 + 1] + (dV) * R[L->i1 + 1];
    Sc[L->i1 + 2] = Sc[L->i1 + 2] + (dV) * R[L->i1 + 2];
  }
  return 0;
}

int orc_m5_rk(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, int irk, const double *SN_in,
              double *S, const double *RHS, const unsigned char *gfm_face) {
  lay5_t L = mk_lay5(p->nmat);
  const int neq = L.neq;
  /* published S = SN + dt RHS / V; consistent S = SN - dt RHS / V */
  const double DT = (p->fix & ORC_M5_FIX_SIGNS) ? -p->dt : p->dt;
  double *SN = (double *)SN_in;   /* written only by case 0 (SN = S) */
  for (int c = 0; c < m->n; c++) {
    double *Sc = S + (size_t)c * neq, *SNc = SN + (size_t)c * neq;
    const double *R = RHS + (size_t)c * neq;
    const double Vol = cells[c].Vol;
    double phi_old[ORC_MAXMAT], phi_new[ORC_MAXMAT];
    for (int v = 0; v < L.n3; ++v) phi_old[v] = Sc[L.i3 + v] - p->avf_lim;
    switch (irk) {
      case 0:
        for (int e = 0; e < neq; e++) SNc[e] = Sc[e];
        break;
      case 1: {
        int flag_rk = 0;
        for (int v = 0; v < L.n3; ++v) {
          if (p->mush[v] != 0.0) {
            int e = L.i3 + v;
            double buf = (1.0 / Vol) * R[e];
            double as_new = SNc[e] + DT * buf;
            phi_new[v] = as_new - p->avf_lim;
            if (phi_old[v] * phi_new[v] < 0.0) flag_rk = 1;
          }
        }
        if (flag_rk == 1) {
          for (int e = 0; e < neq; e++) {
            if (e >= L.i2) {
              double buf = (1.0 / Vol) * R[e];
              Sc[e] = SNc[e] + DT * buf;
            }
          }
          int rc = m5_gfm(&L, p, m, &cells[c], c, S, RHS, gfm_face);
          if (rc) return rc;
        } else {
          for (int e = 0; e < neq; e++) {
            double buf = (1.0 / Vol) * R[e];
            Sc[e] = SNc[e] + DT * buf;
          }
        }
        break;
      }
      case 2:
        for (int e = 0; e < neq; e++) Sc[e] = (3.0 / 4.0) * SNc[e] + (1.0 / 4.0) * Sc[e] + (1.0 / 4.0) * DT * R[e] / Vol;
        break;
      case 3:
        for (int e = 0; e < neq; e++) Sc[e] = (1.0 / 3.0) * SNc[e] + (2.0 / 3.0) * Sc[e] + (2.0 / 3.0) * DT * R[e] / Vol;
        break;
      case 4:
        for (int e = 0; e < neq; e++) Sc[e] = (1.0 / 2.0) * SNc[e] + (1.0 / 2.0) * Sc[e] + (1.0 / 2.0) * DT * R[e] / Vol;
        break;
      default:
        break;
    }
    for (int v = 0; v < L.n0; v++) {
      int e = L.i0 + v;
      if (Sc[e] < 0.0) Sc[e] = SNc[e];
    }
    if (p->zerou[0] == 1) {
      Sc[L.i1 + 0] = SNc[L.i1 + 0];
      Sc[L.i5 + 1] = SNc[L.i5 + 1]; Sc[L.i5 + 2] = SNc[L.i5 + 2];
      Sc[L.i5 + 3] = SNc[L.i5 + 3]; Sc[L.i5 + 6] = SNc[L.i5 + 6];
    }
    if (p->zerou[1] == 1) {
      Sc[L.i1 + 1] = SNc[L.i1 + 1];
      Sc[L.i5 + 1] = SNc[L.i5 + 1]; Sc[L.i5 + 3] = SNc[L.i5 + 3];
      Sc[L.i5 + 5] = SNc[L.i5 + 5]; Sc[L.i5 + 7] = SNc[L.i5 + 7];
    }
    if (p->zerou[2] == 1) {
      Sc[L.i1 + 2] = SNc[L.i1 + 2];
      Sc[L.i5 + 2] = SNc[L.i5 + 2]; Sc[L.i5 + 5] = SNc[L.i5 + 5];
      Sc[L.i5 + 6] = SNc[L.i5 + 6]; Sc[L.i5 + 7] = SNc[L.i5 + 7];
    }
  }
  return 0;
}

int orc_m5_gfm_rel(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, const double *S_before,
                   double *S, const unsigned char *gfm_face) {
  lay5_t L = mk_lay5(p->nmat);
  const int neq = L.neq;
  for (int c = 0; c < m->n; c++) {
    double *Sc = S + (size_t)c * neq;
    const double phi_old = S_before[(size_t)c * neq + L.i3] - p->avf_lim;
    /* phi_new[0] is only assigned for a solid (m5_relaxation_solid.c:26-31) */
    if (p->mush[0] == 0.0) continue;
    const double phi_new = Sc[L.i3] - p->avf_lim;
    if (!(phi_old * phi_new < 0.0)) continue;
    /* m5_gfm_rel.c:44-60: velocity of the relaxed state */
    double rho = 0.0;
    for (int v = 0; v < L.n0; ++v) rho += Sc[v];
    double ut = Sc[L.i1 + 0] / rho, vt = Sc[L.i1 + 1] / rho, wt = Sc[L.i1 + 2] / rho;
    /* leaves before c are relaxed (and re-projected); leaves after c are not yet */
    const double *Snb[6];
    for (int f = 0; f < 6; f++) {
      int j = m->bc[6 * c + f] == 0 ? m->neigtr[24 * c + 4 * f] : c;
      Snb[f] = (j <= c ? S : S_before) + (size_t)j * neq;
    }
    double ng[3];
    int err;
    int fi = gfm_velocity(&L, p, m, &cells[c], c, Sc, Snb, gfm_face + 6 * c, 1, ut, vt, wt, ng, &err);
    if (err) return err;
    if (fi != 0) {
      Sc[L.i1 + 0] = rho * ng[0];
      Sc[L.i1 + 1] = rho * ng[1];
      Sc[L.i1 + 2] = rho * ng[2];
      Sc[L.i2] -= rho * 0.5 * (SQ(ut) + SQ(vt) + SQ(wt));
      Sc[L.i2] += rho * 0.5 * (SQ(ng[0]) + SQ(ng[1]) + SQ(ng[2]));
    }
  }
  return 0;
}

/* ================================================================== */
/* the whole model-5 stage                                              */
/* ================================================================== */
static const int FCOPP5[6] = {2, 3, 0, 1, 5, 4};   /* operatetp.c:268-279 */

static fr5_t frame_of_cell(const orc_cell *cl, int f) {
  fr5_t F = {cl->nx[f], cl->ny[f], cl->nz[f], cl->nxt1[f], cl->nyt1[f], cl->nzt1[f],
             cl->nxt2[f], cl->nyt2[f], cl->nzt2[f]};
  return F;
}

static void gmat_of(const double *A, double *G, double *Gmag) {
  G[0] = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  G[1] = A[0] * A[3] + A[1] * A[4] + A[2] * A[5];
  G[2] = A[0] * A[6] + A[1] * A[7] + A[2] * A[8];
  G[3] = A[3] * A[0] + A[4] * A[1] + A[5] * A[2];
  G[4] = A[3] * A[3] + A[4] * A[4] + A[5] * A[5];
  G[5] = A[3] * A[6] + A[4] * A[7] + A[5] * A[8];
  G[6] = A[6] * A[0] + A[7] * A[1] + A[8] * A[2];
  G[7] = A[6] * A[3] + A[7] * A[4] + A[8] * A[5];
  G[8] = A[6] * A[6] + A[7] * A[7] + A[8] * A[8];
  *Gmag = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) + G[2] * (G[3] * G[7] - G[4] * G[6]);
}

/* ------------------------------------------------------------------ */
/* relaxation (m5_relaxationvec.c)                                      */
/* ------------------------------------------------------------------ */
/* the volume-fraction limit after a relaxation (m5_relaxationvec.c:131-150):
 * two independent clips, renormalised when the sum reaches 1 - AMIN/100 */
static void vf_limit2(const lay5_t *L, double a, double *avf) {
  double avfsum = 0.0;
  for (int iv = 0; iv < L->n3; ++iv) {
    if (avf[iv] <= a) avf[iv] = a;
    if (avf[iv] >= (1.0 - a)) avf[iv] = 1.0 - a;
    avfsum += avf[iv];
  }
  if (avfsum >= (1.0 - a)) {
    double dif = avfsum - (1.0 - a);
    for (int j = 0; j < L->n3; ++j) { double per = avf[j] / avfsum; avf[j] = avf[j] - per * dif; }
    avfsum = 1.0 - a;
  }
  avf[L->n3] = 1.0 - avfsum;
}

/* the same limit with else-if clips (m5_relaxationvec.c:44-60 and :512-529) */
static void vf_limit1(const lay5_t *L, double a, double *avf) {
  double avfsum = 0.0;
  for (int iv = 0; iv < L->n3; ++iv) {
    if (avf[iv] <= a) avf[iv] = a;
    else if (avf[iv] >= (1.0 - a)) avf[iv] = 1.0 - a;
    avfsum += avf[iv];
  }
  if (avfsum >= (1.0 - a)) {
    double dif = avfsum - (1.0 - a);
    for (int j = 0; j < L->n3; ++j) { double per = avf[j] / avfsum; avf[j] = avf[j] - per * dif; }
    avfsum = 1.0 - a;
  }
  avf[L->n3] = 1.0 - avfsum;
}

/* F_pe.c (i_paper = 1, Favrie) on the branch's own arrays */
static double m5_fpe(int nm, const double *G, const double *P, const double *ar, const double *t0, const double *p0,
                     double pe) {
  double F = 0.0;
  for (int iv = 0; iv < nm; ++iv)
    F += ar[iv] * t0[iv] * ((p0[iv] + G[iv] * P[iv] + (G[iv] - 1.0) * pe) / (pe + G[iv] * P[iv] + (G[iv] - 1.0) * pe));
  return F + -1.0;
}
static double m5_dfpe(int nm, const double *G, const double *P, const double *ar, const double *t0, const double *p0,
                      double pe) {
  double D = 0.0;
  for (int iv = 0; iv < nm; ++iv) {
    double c1 = ar[iv] * t0[iv];
    double f = (p0[iv] + G[iv] * P[iv] + (G[iv] - 1.0) * pe);
    double g = (pe + G[iv] * P[iv] + (G[iv] - 1.0) * pe);
    D += (c1 * (G[iv] - 1.0) * g - c1 * f * G[iv]) / SQ(g);
  }
  return D;
}

static int m5_relaxvec(const lay5_t *L, const orc_m5_params *p, double *vec) {
  const double a = p->amin / 100.0, eps_sml = 1.0e-12, eps_big = 1.0e-8;
  const double *Gm = p->gamma, *Pi = p->pinf, *Mu = p->mush;
  const int nm = L->nmat;
  const int mix = p->fix & ORC_M5_FIX_MIXTURE;
  double ar[ORC_MAXMAT], are[ORC_MAXMAT], avf[ORC_MAXMAT + 1], avf_init[ORC_MAXMAT + 1];
  double are_el[ORC_MAXMAT] = {0}, A[9], G[9], Gmag;
  double rho = 0.0;
  for (int iv = 0; iv < L->n0; ++iv) { ar[iv] = vec[iv + L->i0]; are[iv] = vec[iv + L->i4]; rho += ar[iv]; }
  const double u = vec[L->i1 + 0] / rho, v = vec[L->i1 + 1] / rho, w = vec[L->i1 + 2] / rho;
  const double rE = vec[L->i2];
  for (int iv = 0; iv < L->n3; ++iv) avf[iv] = vec[iv + L->i3];
  vf_limit1(L, a, avf);
  for (int iv = 0; iv <= L->n3; ++iv) avf_init[iv] = avf[iv];
  for (int k = 0; k < 9; ++k) A[k] = vec[L->i5 + k];
  gmat_of(A, G, &Gmag);
  double p_relaxed;
  if (p->prelax == 1 || p->prelax == 2) {
    double t0[ORC_MAXMAT], p0[ORC_MAXMAT], pe;
    if (p->prelax == 1) {   /* m5_relaxationvec.c:89-190 */
      for (int iv = 0; iv < L->n0; ++iv) {
        double ael = (Mu[iv] != 0.0) ? elastic_energy_v2(G, Gmag, Mu[iv]) : 0.0;
        if (mix) ael *= avf[iv];
        t0[iv] = 1.0 / (ar[iv] / avf[iv]);
        p0[iv] = ((Gm[iv] - 1.0) * (are[iv] - ael) / avf[iv]) - Gm[iv] * Pi[iv];
      }
      double A_1 = (avf[0] / Gm[0]) * (p0[0] + Pi[0]) / ((avf[0] / Gm[0]) + (avf[1] / Gm[1]));
      double A_2 = (avf[1] / Gm[1]) * (p0[1] + Pi[1]) / ((avf[0] / Gm[0]) + (avf[1] / Gm[1]));
      pe = 0.5 * (A_1 + A_2 - (Pi[0] + Pi[1])) + sqrt(0.25 * (SQ((A_2 - A_1 - (Pi[1] - Pi[0])))) + A_1 * A_2);
      if (isnan(pe)) return -1;
    } else {                /* m5_relaxationvec.c:192-363 */
      double max_avf = -1.0;
      int principal = 0;
      for (int iv = 0; iv < L->n0; ++iv) {
        double ael = (Mu[iv] != 0.0) ? elastic_energy_v2(G, Gmag, Mu[iv]) : 0.0;
        if (mix) ael *= avf[iv];
        t0[iv] = 1.0 / (ar[iv] / avf[iv]);
        p0[iv] = ((Gm[iv] - 1.0) * (are[iv] - ael) / avf[iv]) - Gm[iv] * Pi[iv];
        if (p0[iv] < -(1.0 - eps_big) * Pi[iv] + eps_big) p0[iv] = -(1.0 - eps_big) * Pi[iv] + 1.0;
        if (avf[iv] > max_avf) { max_avf = avf[iv]; principal = iv; }
      }
      double pe_old = 0.001, pe_new = pe_old;
      int iter = 0;
      double cond_1 = fabs(pe_new - pe_old);
      double cond_2 = fabs(m5_fpe(nm, Gm, Pi, ar, t0, p0, pe_new));
      while (((cond_2 >= 1.0e-12) || (cond_1 >= 1.0e-4)) && (iter <= 101)) {
        pe_new = pe_old - 0.8 * (m5_fpe(nm, Gm, Pi, ar, t0, p0, pe_old) / m5_dfpe(nm, Gm, Pi, ar, t0, p0, pe_old));
        cond_1 = fabs(pe_new - pe_old);
        cond_2 = fabs(m5_fpe(nm, Gm, Pi, ar, t0, p0, pe_new));
        if (iter >= 100 && ((fabs(pe_old - pe_new) < 1.0e3) || isnan(pe_new))) return -3;
        pe_old = pe_new;
        iter++;
      }
      if (isnan(pe_new)) return -1;
      if (pe_new < (-(1.0 - eps_big) * Pi[principal] + eps_big)) pe_new = -(1.0 - eps_big) * Pi[principal] + 1.0;
      pe = pe_new;
    }
    for (int iv = 0; iv < L->n0; ++iv) {
      double t1 = t0[iv] * (p0[iv] + Gm[iv] * Pi[iv] + (Gm[iv] - 1.0) * pe) / (pe + Gm[iv] * Pi[iv] + (Gm[iv] - 1.0) * pe);
      double r_m = 1.0 / t1;
      avf[iv] = ar[iv] / r_m;
    }
    vf_limit2(L, a, avf);
    const double ratio = avf[0] / avf_init[0];
    for (int i = 0; i < 9; i++) A[i] = A[i] * ((p->fix & ORC_M5_FIX_SIGNS) ? cbrt(1.0 / ratio) : ratio);
  } else if (p->prelax == 3) {   /* m5_relaxationvec.c:365-560 */
    double rhoKs[ORC_MAXMAT] = {0}, pk[ORC_MAXMAT];
    for (int iv = 0; iv < nm; ++iv) {
      if (avf[iv] < 0.0 || ar[iv] < 0.0) { avf[iv] = eps_sml; ar[iv] = eps_sml; are[iv] = eps_sml; }
      if (avf[iv] >= 1.0) avf[iv] = 1.0 - eps_sml;
    }
    int irelax = 1;
    for (int iv = 0; iv < nm; ++iv)
      if (avf[iv] > 1.0 - eps_sml) irelax = 0;
    if (irelax == 1) {
      p_relaxed = 0.0;
      for (int iv = 0; iv < nm; ++iv) {
        if (avf[iv] > eps_sml) {
          double ael = (Mu[iv] != 0.0) ? elastic_energy_v2(G, Gmag, Mu[iv]) : 0.0;
        if (mix) ael *= avf[iv];
          pk[iv] = ((Gm[iv] - 1.0) * (are[iv] - ael) / avf[iv]) - Gm[iv] * Pi[iv];
          if (pk[iv] <= -(1.0 - eps_big) * Pi[iv] + eps_big) pk[iv] = -(1.0 - eps_big) * Pi[iv] + eps_big;
        } else {
          pk[iv] = 0.0;
        }
        p_relaxed += avf[iv] * pk[iv];
      }
      int iter = 0;
      double F = 1.0e-9, dF = 1.0e9;
      while ((fabs(F) > 1.0e-12) && (iter < 500)) {
        p_relaxed = p_relaxed - F / dF;
        iter = iter + 1;
        for (int iv = 0; iv < nm; ++iv)
          if (p_relaxed < -(1.0 - eps_big) * Pi[iv] + eps_big) p_relaxed = -(1.0 - eps_big) * Pi[iv] + 1.0;
        F = -1.0;
        dF = 0.0;
        for (int iv = 0; iv < nm; ++iv) {
          if (avf[iv] > eps_sml) {
            rhoKs[iv] = ar[iv] / MAXF(avf[iv], eps_sml) * pow((p_relaxed + Pi[iv]) / (pk[iv] + Pi[iv]), 1.0 / Gm[iv]);
            F = F + ar[iv] / rhoKs[iv];
            dF = dF - ar[iv] / (Gm[iv] * rhoKs[iv] * (p_relaxed + Pi[iv]));
          }
        }
        if (isnan(p_relaxed) || isnan(F) || isnan(dF)) return -1;
      }
      for (int iv = 0; iv < nm; ++iv)
        if (avf[iv] > eps_sml) avf[iv] = ar[iv] / rhoKs[iv];
      vf_limit1(L, a, avf);
    }
    int i_as = 0;
    if ((avf[0] < avf[1]) && (Mu[1] != 0.0)) i_as = 1;
    const double ratio = avf[i_as] / avf_init[i_as];
    for (int i = 0; i < 9; i++) A[i] = A[i] * ((p->fix & ORC_M5_FIX_SIGNS) ? cbrt(1.0 / ratio) : ratio);
  }
  if (p->prelax >= 1 && p->prelax <= 3) {
    /* re-initialisation: pressure from the total energy (m5_relaxationvec.c:157-189 and the same in each branch) */
    double e_elast = 0.0;
    if (mix) gmat_of(A, G, &Gmag);   /* the elastic energy of the relaxed A */
    for (int iv = 0; iv < nm; ++iv) {
      are_el[iv] = (Mu[iv] != 0.0) ? elastic_energy_v2(G, Gmag, Mu[iv]) : 0.0;
      if (mix) are_el[iv] *= avf[iv];
      double e_elastic_K = are_el[iv] / (ar[iv]);
      double Ys = ar[iv] / rho;
      e_elast += Ys * e_elastic_K;
    }
    double e_hyd = rE / rho - 0.5 * (SQ((u)) + SQ((v)) + SQ((w))) - e_elast;
    double num = 0.0, den = 0.0;
    for (int iv = 0; iv < L->n0; ++iv) {
      num += avf[iv] * Gm[iv] * Pi[iv] / (Gm[iv] - 1.0);
      den += avf[iv] / (Gm[iv] - 1.0);
    }
    p_relaxed = (rho * e_hyd - num) / den;
    if (isnan(p_relaxed)) return -1;
    for (int iv = 0; iv < L->n0; ++iv)
      are[iv] = (avf[iv] * (p_relaxed + Gm[iv] * Pi[iv]) / (Gm[iv] - 1.0)) + are_el[iv];
  }
  for (int iv = 0; iv < L->n3; ++iv) {
    if (isnan(avf[iv])) return -1;
    vec[iv + L->i3] = avf[iv];
  }
  for (int iv = 0; iv < L->n4; iv++) {
    if (isnan(are[iv])) return -1;
    vec[iv + L->i4] = are[iv];
  }
  for (int iv = 0; iv < 9; ++iv) vec[iv + L->i5] = A[iv];
  return 0;
}

int orc_m5_relax(const orc_m5_params *p, double *vec) {
  lay5_t L = mk_lay5(p->nmat);
  return m5_relaxvec(&L, p, vec);
}

/* ------------------------------------------------------------------ */
/* perfect plasticity (new: plastic_rlx, m5_relaxation_solid.c:14, is    */
/* commented out in the reference)                                      */
/* ------------------------------------------------------------------ */
/* cyclic Jacobi on the symmetric 3x3 matrix a (row-major, destroyed):
 * l = eigenvalues, Q columns = eigenvectors, a = Q diag(l) Q^T */
static void sym3_eigen(double a[9], double l[3], double Q[9]) {
  static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int k = 0; k < 9; k++) Q[k] = (k % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 12; sweep++) {
    double off = fabs(a[1]) + fabs(a[2]) + fabs(a[5]);
    double dia = fabs(a[0]) + fabs(a[4]) + fabs(a[8]);
    if (off <= 1e-300 || off <= 1e-18 * dia) break;
    for (int r = 0; r < 3; r++) {
      const int P = PQ[r][0], q = PQ[r][1];
      const double apq = a[P * 3 + q];
      if (apq == 0.0) continue;
      const double th = (a[q * 3 + q] - a[P * 3 + P]) / (2.0 * apq);
      const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; k++) {   /* a = a J (columns P, q) */
        const double akp = a[k * 3 + P], akq = a[k * 3 + q];
        a[k * 3 + P] = c * akp - s * akq;
        a[k * 3 + q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; k++) {   /* a = J^T a (rows P, q) */
        const double apk = a[P * 3 + k], aqk = a[q * 3 + k];
        a[P * 3 + k] = c * apk - s * aqk;
        a[q * 3 + k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; k++) {   /* Q = Q J */
        const double qkp = Q[k * 3 + P], qkq = Q[k * 3 + q];
        Q[k * 3 + P] = c * qkp - s * qkq;
        Q[k * 3 + q] = s * qkp + c * qkq;
      }
    }
  }
  l[0] = a[0]; l[1] = a[4]; l[2] = a[8];
}

/* von Mises stress of the deviatoric stress -mu d^(-1/6) (G^2 - tr(G^2)/3 I
 * - d^(1/3) (G - tr(G)/3 I)) of state_equation.c:143-204, for G with
 * eigenvalues d^(1/3) exp(t k_i) (sum k_i = 0), divided by mu d^(1/2) */
static double vm_path(const double k[3], double t) {
  double e1[3], e2[3], s1 = 0.0, s2 = 0.0, q = 0.0;
  for (int i = 0; i < 3; i++) { e1[i] = exp(t * k[i]); e2[i] = e1[i] * e1[i]; s1 += e1[i]; s2 += e2[i]; }
  for (int i = 0; i < 3; i++) {
    const double si = (e2[i] - s2 / 3.0) - (e1[i] - s1 / 3.0);
    q += si * si;
  }
  return sqrt(1.5 * q);
}

static int m5_plastic(const orc_m5_params *p, const lay5_t *L, double *vec) {
  double mu = 0.0;
  for (int i = 0; i < p->nmat; i++) mu += p->mush[i];
  if (!(p->yield > 0.0) || mu == 0.0) return 0;
  double A[9], G[9], d;
  for (int k = 0; k < 9; k++) A[k] = vec[L->i5 + k];
  gmat_of(A, G, &d);
  if (!(d > 0.0)) return 0;
  /* the von Mises stress of the current state, matrix form */
  double G2[9] = {0}, ss[9];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j) G2[i * 3 + j] += G[i * 3 + k] * G[k * 3 + j];
  const double J1 = G[0] + G[4] + G[8], J2 = G2[0] + G2[4] + G2[8], d13 = cbrt(d);
  double q = 0.0;
  for (int k = 0; k < 9; k++) {
    ss[k] = G2[k] - d13 * G[k];
    if (k % 4 == 0) ss[k] -= J2 / 3.0 - d13 * J1 / 3.0;
    q += ss[k] * ss[k];
  }
  const double vm = mu / pow(d, 1.0 / 6.0) * sqrt(1.5 * q);
  if (!(vm > p->yield)) return 0;
  /* log strains k_i = ln l_i - ln(d)/3 of G = Q diag(l) Q^T */
  double a[9], l[3], Q[9], kap[3];
  memcpy(a, G, sizeof a);
  sym3_eigen(a, l, Q);
  const double lnd3 = (log(l[0]) + log(l[1]) + log(l[2])) / 3.0;
  for (int i = 0; i < 3; i++) kap[i] = log(l[i]) - lnd3;
  /* the largest t in [0, 1] with von Mises <= yield along l_i(t) = d^(1/3) exp(t k_i) */
  const double target = p->yield / (mu * sqrt(d));
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 64; it++) {
    const double mid = 0.5 * (lo + hi);
    if (vm_path(kap, mid) > target) hi = mid; else lo = mid;
  }
  /* A = Q diag(exp((t - 1) k_i / 2)) Q^T A */
  double f[3], M[9], An[9];
  for (int i = 0; i < 3; i++) f[i] = exp(0.5 * (lo - 1.0) * kap[i]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      M[i * 3 + j] = Q[i * 3 + 0] * f[0] * Q[j * 3 + 0] + Q[i * 3 + 1] * f[1] * Q[j * 3 + 1] + Q[i * 3 + 2] * f[2] * Q[j * 3 + 2];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) An[i * 3 + j] = M[i * 3 + 0] * A[0 * 3 + j] + M[i * 3 + 1] * A[1 * 3 + j] + M[i * 3 + 2] * A[2 * 3 + j];
  for (int k = 0; k < 9; k++) {
    if (isnan(An[k])) return -1;
    vec[L->i5 + k] = An[k];
  }
  return 1;
}

int orc_m5_plastic(const orc_m5_params *p, double *vec) {
  lay5_t L = mk_lay5(p->nmat);
  return m5_plastic(p, &L, vec);
}

/* ------------------------------------------------------------------ */
/* MUSCL face states (m5_faceinterp_V1)                                 */
/* ------------------------------------------------------------------ */
/* avf_limit.c:3-21 with max_limit 100 on the fractions of X */
static void m5_avf_limit(const lay5_t *L, const orc_m5_params *p, double *X) {
  double avf[ORC_MAXMAT + 1], a = p->amin / 100.0, avfsum = 0.0;
  for (int iv = 0; iv < L->n3; ++iv) avf[iv] = X[iv + L->i3];
  for (int v = 0; v < L->n3; ++v) {
    if (avf[v] <= a) avf[v] = a;
    else if (avf[v] >= (1.0 - a)) avf[v] = 1.0 - a;
    avfsum += avf[v];
  }
  if (avfsum > (1.0 - a)) {
    double dif = avfsum - (1.0 - a);
    for (int j = 0; j < L->n3; ++j) avf[j] -= (avf[j] / avfsum) * dif;
  }
  for (int iv = 0; iv < L->n3; ++iv) X[iv + L->i3] = avf[iv];
}

/* m5_facevalues_MUSCL.c:3-44 with iside = 1: the neighbour's state, the mean
 * of four finer neighbours, or at a boundary the own cell's record vector
 * (m5_bc only moves the record's velocity, not its vector) */
static void m5_stencil(const lay5_t *L, const orc_m5_params *p, const orc_mesh *m, const double *S, int i, int f,
                       double *X) {
  const int neq = L->neq;
  const int bc = m->bc[6 * i + f];
  if (bc == 0) {
    const int lfc = m->lfc_neigs[6 * i + f];
    if (lfc == 1) {
      const int j = m->neigtr[24 * i + 4 * f];
      for (int iv = 0; iv < neq; iv++) X[iv] = S[(size_t)j * neq + iv];
    } else {
      for (int iv = 0; iv < neq; iv++) X[iv] = 0.0;
      for (int k = 0; k < lfc; k++) {
        const int j = m->neigtr[24 * i + 4 * f + k];
        for (int iv = 0; iv < neq; iv++) X[iv] += S[(size_t)j * neq + iv] / ((double)lfc);
      }
    }
  } else {
    rec5_t r;
    m5_record(L, p, &S[(size_t)i * neq], &r);
    for (int iv = 0; iv < neq; iv++) X[iv] = r.vec[iv];
  }
  m5_avf_limit(L, p, X);
}

static const unsigned char *g_m5_cell_limiter = NULL;
void orc_m5_set_cell_limiter(const unsigned char *lim) { g_m5_cell_limiter = lim; }

/* the limited slope of one variable: minmod as m5_reconctruct.c:110-140 has it,
 * or (new) the monotonised central slope minmod(2a, 2b, (a + b) / 2) of the
 * model-4 restatement, a the backward and b the forward difference quotient */
static double m5_slope(int mc, double df, double db, double dp, double dm) {
  if (!mc) {
    const double product = (-db / dm) * (df / dp);
    if (product > 0.0) {
      if (fabs(-db / dm) > fabs(df / dp)) return df / dp;
      return -db / dm;
    }
    return 0.0;
  }
  const double a = -db / dm, b = df / dp;
  if (a * b > 0.0) {
    const double c = 0.5 * (a + b);
    const double mag = MINF(MINF(2.0 * fabs(a), 2.0 * fabs(b)), fabs(c));
    return (c > 0.0) ? mag : -mag;
  }
  return 0.0;
}

/* m5_faceinterp.c:60-93 for leaf i, with m5_deltavalues_V1 (:20-56) and
 * m5_reconctruct_V1 (:93-190); g_limitedvars is 0 on the alpha-rho-e block
 * (m5_input.c:196-204) */
static int m5_faceinterp_leaf(const lay5_t *L, const orc_m5_params *p, const orc_mesh *m, const orc_cell *cl,
                              const double *S, int i, double *SFi) {
  const int neq = L->neq;
  const double a = p->amin / 100.0;
  double SLc[ORC_M5_MAXNEQ], Xf[ORC_M5_MAXNEQ], Xb[ORC_M5_MAXNEQ];
  for (int iv = 0; iv < neq; iv++) SLc[iv] = S[(size_t)i * neq + iv];
  m5_avf_limit(L, p, SLc);
  const int mc = (p->limiter == 1) || (p->limiter == 2 && g_m5_cell_limiter && g_m5_cell_limiter[i]);
  for (int f = 0; f < 6; ++f) {
    const int fo = FCOPP5[f];
    m5_stencil(L, p, m, S, i, f, Xf);
    m5_stencil(L, p, m, S, i, fo, Xb);
    double *SF = &SFi[f * neq];
    const int bcf = m->bc[6 * i + f];
    for (int iv = 0; iv < neq; ++iv) {
      const double df = Xf[iv] - SLc[iv], db = Xb[iv] - SLc[iv];
      const double lim = (iv >= L->i4 && iv < L->i5) ? 0.0 : 1.0;
      double g = 0.0;
      if (bcf == 0 && fabs(df) > 1.0e-10) g = lim * m5_slope(mc, df, db, cl->dist_cc[f], cl->dist_cc[fo]);
      SF[iv] = SLc[iv] + cl->dist_cf[f] * g;
    }
    /* VF limit (m5_reconctruct.c:142-170) */
    double avf[ORC_MAXMAT + 1], avfsum = 0.0;
    for (int iv = 0; iv < L->n3; ++iv) {
      avf[iv] = SF[iv + L->i3];
      if (avf[iv] <= a) avf[iv] = a;
      else if (avf[iv] >= (1.0 - (2.0 * a))) avf[iv] = 1.0 - (2.0 * a);
      avfsum += avf[iv];
    }
    if (avfsum >= (1.0 - a)) {
      double dif = avfsum - (1.0 - a);
      for (int j = 0; j < L->n3; ++j) { double per = avf[j] / avfsum; avf[j] = avf[j] - per * dif; }
    }
    for (int iv = 0; iv < L->n3; ++iv) SF[iv + L->i3] = avf[iv];
    if (p->fix & ORC_M5_FIX_SIGNS) {
      /* the phasic energies follow their fractions, so that each phase keeps its cell
       * pressure at the face (the published block keeps the cell's alpha-rho-e with
       * reconstructed fractions, a pressure error of order gamma p_inf d(alpha)/alpha) */
      double sf = 0.0, sc = 0.0;
      for (int k = 0; k < L->n3; ++k) { sf += SF[L->i3 + k]; sc += SLc[L->i3 + k]; }
      for (int k = 0; k < L->n4; ++k) {
        const double af = (k < L->n3) ? SF[L->i3 + k] : 1.0 - sf;
        const double ac = (k < L->n3) ? SLc[L->i3 + k] : 1.0 - sc;
        SF[L->i4 + k] = SLc[L->i4 + k] * (af / ac);
      }
    }
    int rc = m5_relaxvec(L, p, SF);
    if (rc) return rc;
  }
  return 0;
}

int orc_m5_faceinterp(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, const double *S, double *SF) {
  lay5_t L = mk_lay5(p->nmat);
  for (int i = 0; i < m->n; i++) {
    int rc = m5_faceinterp_leaf(&L, p, m, &cells[i], S, i, &SF[(size_t)i * 6 * L.neq]);
    if (rc) return rc;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* HLLC face flux (m5_hllc.c, m5_wave_speeds.c, m5_flux.c)              */
/* ------------------------------------------------------------------ */
/* m5_bc.c:3-83: 1 symmetry and 4 reflective flip the normal velocity of the
 * left record into the right one, 5 flips all three components; 2 and 3
 * leave the right record as it is */
static void m5_apply_bc(int code, const fr5_t *F, const rec5_t *sL, rec5_t *sR) {
  if (code != 1 && code != 4 && code != 5) return;
  const double ul = sL->u * F->nx + sL->v * F->ny + sL->w * F->nz;
  const double vl = sL->u * F->mx + sL->v * F->my + sL->w * F->mz;
  const double wl = sL->u * F->lx + sL->v * F->ly + sL->w * F->lz;
  const double u1 = -ul, v1 = (code == 5) ? -vl : vl, w1 = (code == 5) ? -wl : wl;
  sR->u = u1 * F->nx + v1 * F->mx + w1 * F->lx;
  sR->v = u1 * F->ny + v1 * F->my + w1 * F->ly;
  sR->w = u1 * F->nz + v1 * F->mz + w1 * F->lz;
}

/* m5_flux.c:3-95: flag 0/3 take the side's own velocity, fractions and -A;
 * 1/2 keep what m5_star_region set; 4 (symmetry) keeps the face velocity and
 * takes the side's fractions and +A */
static void m5_flux(const lay5_t *L, const rec5_t *s, int flag_ifs, int fix, ws5_t *W) {
  for (int v = 0; v < L->n0; v++) {
    int iv = v + L->i0;
    W->flux_adv[iv][0] = s->u * s->ra[v];
    W->flux_adv[iv][1] = s->v * s->ra[v];
    W->flux_adv[iv][2] = s->w * s->ra[v];
  }
  int iv = L->i1;
  /* published: rho u u + sigma; consistent: rho u u - sigma */
  const double sg = (fix & ORC_M5_FIX_SIGNS) ? -1.0 : 1.0;
  W->flux_adv[iv][0] = s->u * s->u * s->r + sg * s->st[0][0];
  W->flux_adv[iv][1] = s->v * s->u * s->r + sg * s->st[0][1];
  W->flux_adv[iv][2] = s->w * s->u * s->r + sg * s->st[0][2];
  iv = L->i1 + 1;
  W->flux_adv[iv][0] = s->u * s->v * s->r + sg * s->st[1][0];
  W->flux_adv[iv][1] = s->v * s->v * s->r + sg * s->st[1][1];
  W->flux_adv[iv][2] = s->w * s->v * s->r + sg * s->st[1][2];
  iv = L->i1 + 2;
  W->flux_adv[iv][0] = s->u * s->w * s->r + sg * s->st[2][0];
  W->flux_adv[iv][1] = s->v * s->w * s->r + sg * s->st[2][1];
  W->flux_adv[iv][2] = s->w * s->w * s->r + sg * s->st[2][2];
  iv = L->i2;
  W->flux_adv[iv][0] = s->r * s->e * s->u - s->st[0][0] * s->u - s->st[0][1] * s->v - s->st[0][2] * s->w;
  W->flux_adv[iv][1] = s->r * s->e * s->v - s->st[1][0] * s->u - s->st[1][1] * s->v - s->st[1][2] * s->w;
  W->flux_adv[iv][2] = s->r * s->e * s->w - s->st[2][0] * s->u - s->st[2][1] * s->v - s->st[2][2] * s->w;
  if ((flag_ifs == 0) || (flag_ifs == 3)) {
    W->u_nc = s->u; W->v_nc = s->v; W->w_nc = s->w;
    for (int v = 0; v < L->n3; ++v) W->avf_S[v] = s->avf[v];
    for (int v = 0; v < L->n4; ++v) W->are_S[v] = s->are[v];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) W->Amat_S[i][j] = (fix & ORC_M5_FIX_SIGNS) ? s->Amat[i * 3 + j] : -s->Amat[i * 3 + j];
  } else if (flag_ifs == 4) {
    for (int v = 0; v < L->n3; ++v) W->avf_S[v] = s->avf[v];
    for (int v = 0; v < L->n4; ++v) W->are_S[v] = s->are[v];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) W->Amat_S[i][j] = s->Amat[i * 3 + j];
  }
  for (int v = 0; v < L->n3; ++v) {
    iv = v + L->i3;
    W->flux_adv[iv][0] = W->u_nc * W->avf_S[v];
    W->flux_adv[iv][1] = W->v_nc * W->avf_S[v];
    W->flux_adv[iv][2] = W->w_nc * W->avf_S[v];
  }
  for (int v = 0; v < L->n4; ++v) {
    iv = v + L->i4;
    W->flux_adv[iv][0] = W->u_nc * W->are_S[v];
    W->flux_adv[iv][1] = W->v_nc * W->are_S[v];
    W->flux_adv[iv][2] = W->w_nc * W->are_S[v];
  }
  iv = L->i5;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      W->flux_adv[iv][0] = W->u_nc * W->Amat_S[i][j];
      W->flux_adv[iv][1] = W->v_nc * W->Amat_S[i][j];
      W->flux_adv[iv][2] = W->w_nc * W->Amat_S[i][j];
      iv++;
    }
}

/* m5_hllc.c:3-94 with m5_wave_speeds.c:3-131.  Xc: the own cell state, for
 * the symmetry flux that averages the cell record and its mirror. */
static int m5_hllc(const lay5_t *L, const orc_m5_params *p, const fr5_t *F, int bc, const rec5_t *sL,
                   const rec5_t *sR, const double *Xc, ws5_t *W) {
  double u, v, w;
  u = sL->u; v = sL->v; w = sL->w;
  double uL = u * F->nx + v * F->ny + w * F->nz;
  double vL = u * F->mx + v * F->my + w * F->mz;
  double wL = u * F->lx + v * F->ly + w * F->lz;
  u = sR->u; v = sR->v; w = sR->w;
  double uR = u * F->nx + v * F->ny + w * F->nz;
  double vR = u * F->mx + v * F->my + w * F->mz;
  double wR = u * F->lx + v * F->ly + w * F->lz;
  double ruL = sL->r * (uL), ruR = sR->r * (uR);
  double c2YL = 0.0, c2YR = 0.0;
  for (int iv = 0; iv < L->nmat; ++iv) { c2YL += sL->c2Y[iv]; c2YR += sR->c2Y[iv]; }
  double cL = sqrt(c2YL), cR = sqrt(c2YR);
  double SR = MAXF((uL + cL), (uR + cR));
  double SL = MINF((uL - cL), (uR - cR));
  double t[9];
  rot_tensor(t, (double(*)[3])sL->st, F->nx, F->ny, F->nz, F->mx, F->my, F->mz, F->lx, F->ly, F->lz);
  double s11L = t[0], s12L = t[1], s13L = t[2];
  rot_tensor(t, (double(*)[3])sR->st, F->nx, F->ny, F->nz, F->mx, F->my, F->mz, F->lx, F->ly, F->lz);
  double s11R = t[0], s12R = t[1], s13R = t[2];
  double u_S = ((sL->r * SQ((uL)) - (s11L)) - (sR->r * SQ((uR)) - (s11R)) - ((SL)*ruL) + ((SR)*ruR)) /
               (ruL - ruR - ((SL)*sL->r) + ((SR)*sR->r));
  if (isnan(u_S) || isnan(SR) || isnan(SL)) return -1;
  if (bc == 4) u_S = 0.0;
  for (int iv = 0; iv < L->neq; ++iv) W->flux_hllc[iv] = 0.0;
  int end = 0;
  if ((0.0 <= SL) && !end) { end = 1; m5_flux(L, sL, 0, p->fix, W); }
  if (((SL <= 0.0) && (0.0 <= u_S)) && !end) {
    end = 1;
    if (m5_star_region(L, p, F, sL, sR, sL, 1, u_S, SL, SR, uL, vL, wL, uR, vR, wR, s11L, s12L, s13L, s11R, s12R,
                       s13R, W))
      return -1;
    m5_flux(L, sL, 1, p->fix, W);
    for (int iv = 0; iv < L->ncons; ++iv) W->flux_hllc[iv] = SL * (W->Us[iv] - sL->vec[iv]);
  }
  if (((u_S <= 0.0) && (0.0 <= SR)) && !end) {
    end = 1;
    if (m5_star_region(L, p, F, sL, sR, sR, 2, u_S, SL, SR, uL, vL, wL, uR, vR, wR, s11L, s12L, s13L, s11R, s12R,
                       s13R, W))
      return -1;
    m5_flux(L, sR, 2, p->fix, W);
    for (int iv = 0; iv < L->ncons; ++iv) W->flux_hllc[iv] = SR * (W->Us[iv] - sR->vec[iv]);
  }
  if ((SR <= 0.0) && !end) { end = 1; m5_flux(L, sR, 3, p->fix, W); }
  if (bc == 1) {   /* symmetry: mean of the physical fluxes of the cell record and its mirror */
    double sym[ORC_M5_MAXNEQ][3];
    for (int iv = 0; iv < L->neq; ++iv) {
      W->flux_hllc[iv] = 0.0;
      for (int k = 0; k < 3; k++) sym[iv][k] = 0.0;
    }
    rec5_t cL, cR;
    m5_record(L, p, Xc, &cL);
    m5_record(L, p, Xc, &cR);
    m5_apply_bc(1, F, &cL, &cR);
    W->u_nc = 0.5 * (cL.u + cR.u);
    W->v_nc = 0.5 * (cL.v + cR.v);
    W->w_nc = 0.5 * (cL.w + cR.w);
    m5_flux(L, &cL, 4, p->fix, W);
    for (int iv = 0; iv < L->neq; iv++)
      for (int k = 0; k < 3; k++) sym[iv][k] += 0.5 * W->flux_adv[iv][k];
    m5_flux(L, &cR, 4, p->fix, W);
    for (int iv = 0; iv < L->neq; iv++)
      for (int k = 0; k < 3; k++) sym[iv][k] += 0.5 * W->flux_adv[iv][k];
    for (int iv = 0; iv < L->neq; iv++)
      for (int k = 0; k < 3; k++) W->flux_adv[iv][k] = sym[iv][k];
  }
  for (int iv = 0; iv < L->neq; ++iv)
    if (isnan(W->flux_hllc[iv])) return -1;
  return 0;
}

/* one sub-face of m5_residual.c:139-170: GFM where phi changes sign, else
 * HLLC; returns 1 for GFM, 0 for HLLC, -1 on NaN */
static int m5_subface(const lay5_t *L, const orc_m5_params *p, const fr5_t *F, int bc, rec5_t *sL, rec5_t *sR,
                      const double *Xc, ws5_t *W) {
  if (sL->phi0 * sR->phi0 < 0.0) return m5_hllc_gfm(L, p, F, sL, sR, W) ? -1 : 1;
  return m5_hllc(L, p, F, bc, sL, sR, Xc, W) ? -1 : 0;
}

int orc_m5_face(const orc_m5_params *p, const double *XL, const double *CL, const double *XR, const double *CR,
                const double *frame, int bc, double area, double *flux, double *unc) {
  lay5_t L = mk_lay5(p->nmat);
  fr5_t F = {frame[0], frame[1], frame[2], frame[3], frame[4], frame[5], frame[6], frame[7], frame[8]};
  rec5_t sL, sR;
  ws5_t W;
  memset(&sL, 0, sizeof sL); memset(&sR, 0, sizeof sR); memset(&W, 0, sizeof W);
  m5_record_x(&L, p, XL, CL, 1, &sL);
  if (bc == 0) {
    m5_record_x(&L, p, XR, CR, 1, &sR);
  } else {
    m5_record_x(&L, p, XL, CL, 1, &sR);
    m5_apply_bc(bc, &F, &sL, &sR);
  }
  int g = m5_subface(&L, p, &F, bc, &sL, &sR, CL, &W);
  if (g < 0) return -1;
  for (int iv = 0; iv < L.neq; ++iv)
    flux[iv] = 0.0 + (area) * (((W.flux_adv[iv][0] - 0.0) * F.nx + (W.flux_adv[iv][1] - 0.0) * F.ny +
                                (W.flux_adv[iv][2] - 0.0) * F.nz) + W.flux_hllc[iv]);
  unc[0] = W.u_nc; unc[1] = W.v_nc; unc[2] = W.w_nc;
  return g;
}

/* ------------------------------------------------------------------ */
/* residual (m5_residual.c:7-238)                                       */
/* ------------------------------------------------------------------ */
typedef struct {
  double *flux_face;   /* [n][6][NEQ] */
  double *u_face;      /* [n][6][3] */
  int *flux_flag;      /* [n][6] */
} m5_ws;

static int m5_residual_leaf(const lay5_t *L, const orc_m5_params *p, const orc_mesh *m, const orc_cell *cells,
                            const double *S, const double *SF, int i, m5_ws *R, double *RHS, unsigned char *gfm) {
  const orc_cell *cl = &cells[i];
  const int neq = L->neq;
  const double *Si = &S[(size_t)i * neq];
  double *rhs = &RHS[(size_t)i * neq];
  for (int iv = 0; iv < neq; iv++) rhs[iv] = 0.0;
  rec5_t sc;
  m5_record(L, p, Si, &sc);
  double source_avf[ORC_MAXMAT];
  for (int iv = 0; iv < L->n3; ++iv) source_avf[iv] = sc.avf[iv];
  const double A1 = sc.Amat[0], B1 = sc.Amat[3], C1 = sc.Amat[6];
  const double A2 = sc.Amat[1], B2 = sc.Amat[4], C2 = sc.Amat[7];
  const double A3 = sc.Amat[2], B3 = sc.Amat[5], C3 = sc.Amat[8];
  const int SI0 = m->SI[2 * i];
  for (int f = 0; f < 6; f++) {
    double *ff = &R->flux_face[((size_t)i * 6 + f) * neq];
    double *uf = &R->u_face[((size_t)i * 6 + f) * 3];
    for (int iv = 0; iv < neq; ++iv) ff[iv] = 0.0;
    gfm[6 * i + f] = 0;
    uf[0] = 0.0; uf[1] = 0.0; uf[2] = 0.0;
    const int bc = m->bc[6 * i + f];
    int Nneig = 1, compute = 1;
    if (bc == 0) {
      const int lfc = m->lfc_neigs[6 * i + f];
      if (lfc == 1) {
        const int j = m->neigtr[24 * i + 4 * f], fo = FCOPP5[f];
        if (m->lfc_neigs[6 * j + fo] == 1 && SI0 == 0) {   /* 1:1, one rank */
          if (R->flux_flag[6 * i + f] == 1) {
            const double *uo = &R->u_face[((size_t)j * 6 + fo) * 3];
            const double *fo_ = &R->flux_face[((size_t)j * 6 + fo) * neq];
            /* the star velocity is a global vector: the published code negates it on
             * the mirrored side (m5_residual.c:62-103), the consistent mode reuses it */
            const double us = (p->fix & ORC_M5_FIX_SIGNS) ? 1.0 : -1.0;
            uf[0] = us * uo[0]; uf[1] = us * uo[1]; uf[2] = us * uo[2];
            for (int iv = 0; iv < neq; ++iv) ff[iv] = -fo_[iv];
            compute = 0;
          } else {
            R->flux_flag[6 * j + fo] = 1;
          }
        }
      } else {
        Nneig = lfc;
      }
    }
    double Area_f = 0.0;
    for (int k = 0; k < Nneig; ++k) Area_f += cl->Area[f][k];
    if (compute) {
      const fr5_t F = frame_of_cell(cl, f);
      rec5_t sL, sR;
      if (p->order == 2) m5_record_x(L, p, &SF[((size_t)i * 6 + f) * neq], Si, 1, &sL);
      else m5_record(L, p, Si, &sL);
      for (int k = 0; k < Nneig; ++k) {
        if (bc == 0) {
          const int j = m->neigtr[24 * i + 4 * f + k];
          const double *Sj = &S[(size_t)j * neq];
          if (p->order == 2) m5_record_x(L, p, &SF[((size_t)j * 6 + FCOPP5[f]) * neq], Sj, 1, &sR);
          else m5_record(L, p, Sj, &sR);
        } else {
          if (p->order == 2) m5_record_x(L, p, &SF[((size_t)i * 6 + f) * neq], Si, 1, &sR);
          else m5_record(L, p, Si, &sR);
          m5_apply_bc(bc, &F, &sL, &sR);
        }
        ws5_t W;
        memset(&W, 0, sizeof W);
        const int g = m5_subface(L, p, &F, bc, &sL, &sR, Si, &W);
        if (g < 0) return -1;
        if (g == 1) gfm[6 * i + f] = 1;
        for (int iv = 0; iv < neq; ++iv)
          ff[iv] += (cl->Area[f][k]) * (((W.flux_adv[iv][0] - 0.0) * cl->nx[f] + (W.flux_adv[iv][1] - 0.0) * cl->ny[f] +
                                         (W.flux_adv[iv][2] - 0.0) * cl->nz[f]) + W.flux_hllc[iv]);
        uf[0] = W.u_nc / ((double)Nneig);   /* overwritten per sub-face (m5_residual.c:163-165) */
        uf[1] = W.v_nc / ((double)Nneig);
        uf[2] = W.w_nc / ((double)Nneig);
      }
    }
    for (int iv = 0; iv < neq; ++iv) rhs[iv] += ff[iv];
    /* non-conservative terms (m5_residual.c:176-214) */
    const double u_nc = uf[0], nx = cl->nx[f];
    const double v_nc = uf[1], ny = cl->ny[f];
    const double w_nc = uf[2], nz = cl->nz[f];
    int v = L->i3;
    if (p->fix & ORC_M5_FIX_SIGNS) {
      /* the consistent sources: -alpha (u.n), +n.(alpha_k sigma_k).u, and for A_ij
       * n_i (A_kj u_k) - A_ij (u.n), all with the cell's own values */
      for (int iv = 0; iv < L->n3; ++iv)
        rhs[v + iv] -= Area_f * (source_avf[iv] * u_nc * nx + source_avf[iv] * v_nc * ny + source_avf[iv] * w_nc * nz);
      v = L->i4;
      for (int iv = 0; iv < L->n4; ++iv) {
        double (*s)[3] = sc.ast[iv];
        rhs[v + iv] -= Area_f * (s[0][0] * u_nc * nx + s[0][1] * v_nc * nx + s[0][2] * w_nc * nx +
                                 s[1][0] * u_nc * ny + s[1][1] * v_nc * ny + s[1][2] * w_nc * ny +
                                 s[2][0] * u_nc * nz + s[2][1] * v_nc * nz + s[2][2] * w_nc * nz);
      }
      v = L->i5;
      const double nn[3] = {nx, ny, nz}, uu[3] = {u_nc, v_nc, w_nc};
      const double un = u_nc * nx + v_nc * ny + w_nc * nz;
      for (int j = 0; j < 3; ++j) {
        const double cu = sc.Amat[0 * 3 + j] * uu[0] + sc.Amat[1 * 3 + j] * uu[1] + sc.Amat[2 * 3 + j] * uu[2];
        for (int ii = 0; ii < 3; ++ii) rhs[v + 3 * ii + j] += Area_f * (nn[ii] * cu - sc.Amat[ii * 3 + j] * un);
      }
      continue;
    }
    for (int iv = 0; iv < L->n3; ++iv)
      rhs[v + iv] += Area_f * (source_avf[iv] * u_nc * nx + source_avf[iv] * v_nc * ny + source_avf[iv] * w_nc * nz);
    v = L->i4;
    for (int iv = 0; iv < L->n4; ++iv) {
      double (*s)[3] = sc.ast[iv];
      rhs[v + iv] += Area_f * (s[0][0] * u_nc * nx + s[0][1] * v_nc * nx + s[0][2] * w_nc * nx +
                               s[1][0] * u_nc * ny + s[1][1] * v_nc * ny + s[1][2] * w_nc * ny +
                               s[2][0] * u_nc * nz + s[2][1] * v_nc * nz + s[2][2] * w_nc * nz);
    }
    v = L->i5;
    rhs[v + 0] += Area_f * ((A1 * v_nc * ny + A1 * w_nc * nz) + B1 * v_nc * nx - C1 * w_nc * nx);
    rhs[v + 1] += Area_f * ((A2 * v_nc * ny + A2 * w_nc * nz) + B2 * v_nc * nx - C2 * w_nc * nx);
    rhs[v + 2] += Area_f * ((A3 * v_nc * ny + A3 * w_nc * nz) + B3 * v_nc * nx - C3 * w_nc * nx);
    rhs[v + 3] += Area_f * ((B1 * u_nc * nx + B1 * w_nc * nz) + A1 * u_nc * ny - C1 * w_nc * ny);
    rhs[v + 4] += Area_f * ((B2 * u_nc * nx + B2 * w_nc * nz) + A2 * u_nc * ny - C2 * w_nc * ny);
    rhs[v + 5] += Area_f * ((B3 * u_nc * nx + B3 * w_nc * nz) + A3 * u_nc * ny - C3 * w_nc * ny);
    rhs[v + 6] += Area_f * ((C1 * u_nc * nx + C1 * v_nc * ny) + A1 * u_nc * nz - B1 * v_nc * nz);
    rhs[v + 7] += Area_f * ((C2 * u_nc * nx + C2 * v_nc * ny) + A2 * u_nc * nz - B2 * v_nc * nz);
    rhs[v + 8] += Area_f * ((C3 * u_nc * nx + C3 * v_nc * ny) + A3 * u_nc * nz - B3 * v_nc * nz);
  }
  for (int iv = 0; iv < neq; iv++)
    if (isnan(rhs[iv])) return -1;
  return 0;
}

static int m5_ws_alloc(const orc_mesh *m, int neq, m5_ws *R) {
  R->flux_face = malloc(sizeof(double) * (size_t)m->n * 6 * neq);
  R->u_face = malloc(sizeof(double) * (size_t)m->n * 18);
  R->flux_flag = calloc((size_t)m->n * 6, sizeof(int));
  return (R->flux_face && R->u_face && R->flux_flag) ? 0 : -2;
}
static void m5_ws_free(m5_ws *R) { free(R->flux_face); free(R->u_face); free(R->flux_flag); }

static int m5_residual_all(const lay5_t *L, const orc_m5_params *p, const orc_mesh *m, const orc_cell *cells,
                           const double *S, const double *SF, double *RHS, unsigned char *gfm, m5_ws *R) {
  memset(R->flux_flag, 0, sizeof(int) * 6 * (size_t)m->n);   /* m5_RK.c:14-16 resets the flags of every leaf */
  for (int i = 0; i < m->n; i++) {
    int rc = m5_residual_leaf(L, p, m, cells, S, SF, i, R, RHS, gfm);
    if (rc) return rc;
  }
  return 0;
}

int orc_m5_residual(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, const double *S,
                    const double *SF, double *RHS, unsigned char *gfm_face) {
  lay5_t L = mk_lay5(p->nmat);
  m5_ws R;
  int rc = m5_ws_alloc(m, L.neq, &R);
  if (!rc) rc = m5_residual_all(&L, p, m, cells, S, SF, RHS, gfm_face, &R);
  m5_ws_free(&R);
  return rc;
}

/* ------------------------------------------------------------------ */
/* stage and steps (advancetimeexplicit.c:3-29, m5_relaxation_solid.c)  */
/* ------------------------------------------------------------------ */
static int m5_stage_impl(const lay5_t *L, const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, int irk,
                         double *SN, double *S, double *SF, double *RHS, unsigned char *gfm, double *Sb, m5_ws *R) {
  const int neq = L->neq;
  int rc;
  if (p->order == 2) {
    for (int i = 0; i < m->n; i++)
      if ((rc = m5_faceinterp_leaf(L, p, m, &cells[i], S, i, &SF[(size_t)i * 6 * neq])) != 0) return rc;
  }
  if ((rc = m5_residual_all(L, p, m, cells, S, SF, RHS, gfm, R)) != 0) return rc;
  if ((rc = orc_m5_rk(m, cells, p, irk, SN, S, RHS, gfm)) != 0) return rc;
  /* m5_relaxation_solid.c:12-37: relax, then re-project the leaves that crossed */
  memcpy(Sb, S, sizeof(double) * (size_t)m->n * neq);
  for (int i = 0; i < m->n; i++) {
    if (p->yield > 0.0 && (rc = m5_plastic(p, L, &S[(size_t)i * neq])) < 0) return rc;   /* where :14 calls plastic_rlx */
    if ((rc = m5_relaxvec(L, p, &S[(size_t)i * neq])) != 0) return rc;
  }
  return orc_m5_gfm_rel(m, cells, p, Sb, S, gfm);
}

typedef struct {
  double *SF, *RHS, *Sb;
  unsigned char *gfm;
  m5_ws R;
} m5_stage_ws;

static int m5_stage_ws_alloc(const orc_mesh *m, int neq, m5_stage_ws *W) {
  W->SF = malloc(sizeof(double) * (size_t)m->n * 6 * neq);
  W->RHS = malloc(sizeof(double) * (size_t)m->n * neq);
  W->Sb = malloc(sizeof(double) * (size_t)m->n * neq);
  W->gfm = malloc((size_t)m->n * 6);
  int rc = m5_ws_alloc(m, neq, &W->R);
  return (W->SF && W->RHS && W->Sb && W->gfm && !rc) ? 0 : -2;
}
static void m5_stage_ws_free(m5_stage_ws *W) {
  free(W->SF); free(W->RHS); free(W->Sb); free(W->gfm);
  m5_ws_free(&W->R);
}

static int m5_check(const orc_m5_params *p) {
  if (p->nmat != 2 && p->nmat != 3) return -2;
  if (p->prelax < 0 || p->prelax > 3 || (p->order != 1 && p->order != 2)) return -2;
  if (!(p->yield >= 0.0)) return -2;
  return 0;
}

int orc_m5_stage(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, int irk, const double *SN,
                 double *S, double *SF, double *RHS, unsigned char *gfm_face) {
  if (m5_check(p)) return -2;
  lay5_t L = mk_lay5(p->nmat);
  m5_stage_ws W;
  int rc = m5_stage_ws_alloc(m, L.neq, &W);
  if (!rc) rc = m5_stage_impl(&L, m, cells, p, irk, (double *)SN, S, W.SF, W.RHS, W.gfm, W.Sb, &W.R);
  if (!rc && SF) memcpy(SF, W.SF, sizeof(double) * (size_t)m->n * 6 * L.neq);
  if (!rc && RHS) memcpy(RHS, W.RHS, sizeof(double) * (size_t)m->n * L.neq);
  if (!rc && gfm_face) memcpy(gfm_face, W.gfm, (size_t)m->n * 6);
  m5_stage_ws_free(&W);
  return rc;
}

int orc_m5_advance(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, double *S, int nsteps) {
  if (m5_check(p) || p->rk_steps < 1 || p->rk_steps > 3) return -2;
  lay5_t L = mk_lay5(p->nmat);
  m5_stage_ws W;
  double *SN = malloc(sizeof(double) * (size_t)m->n * L.neq);
  int rc = m5_stage_ws_alloc(m, L.neq, &W);
  if (!SN) rc = -2;
  for (int s = 0; s < nsteps && !rc; s++) {
    memcpy(SN, S, sizeof(double) * (size_t)m->n * L.neq);   /* RK.c case 0 */
    for (int irk = 1; irk <= p->rk_steps && !rc; irk++)
      rc = m5_stage_impl(&L, m, cells, p, irk, SN, S, W.SF, W.RHS, W.gfm, W.Sb, &W.R);
  }
  free(SN);
  m5_stage_ws_free(&W);
  return rc;
}

/* CFL step of model 5 (new; the reference takes a fixed DT): cfl * min over
 * leaves of h / (max|u_i| + c), c^2 = sum_k c2Y_k + (4/3) sum_k alpha_k mu_k /
 * rho (c2Y_k of the cell record, m5_facevalues.c:140-150; in the mixture mode
 * it carries the shear part already) */
double orc_m5_cfl_dt(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, const double *S, double cfl) {
  (void)cells;
  lay5_t L = mk_lay5(p->nmat);
  double dt = INFINITY;
  for (int i = 0; i < m->n; i++) {
    rec5_t r;
    m5_record(&L, p, &S[(size_t)i * L.neq], &r);
    double c2 = 0.0;
    for (int k = 0; k < L.nmat; ++k) c2 += r.c2Y[k];
    if (!(p->fix & ORC_M5_FIX_MIXTURE)) {
      double mu = 0.0;
      for (int k = 0; k < L.nmat; ++k) mu += r.avf[k] * p->mush[k];
      c2 += (4.0 / 3.0) * mu / r.r;
    }
    const double lam = MAXF(MAXF(fabs(r.u), fabs(r.v)), fabs(r.w)) + sqrt(c2);
    const double h = m->rootsize / (double)(1 << (m->level[i] - 1));
    const double d = h / lam;
    if (d < dt) dt = d;
  }
  return cfl * dt;
}

/* ------------------------------------------------------------------ */
/* indicator with MODEL 5 (interface_location.c, criterionsplt.c)      */
/* ------------------------------------------------------------------ */
/* phi[v] of a cell record (m5_facevalues.c:82-98 with n_solids = 0): the
 * clipped cell fraction minus AVF_LIM */
static double m5_phi(const lay5_t *L, const orc_m5_params *p, const double *S, int v) {
  const double a = p->amin / 100.0;
  double c = S[L->i3 + v];
  if (c >= (1.0 - a)) c = 1.0 - a;
  else if (c <= a) c = a;
  return c - p->avf_lim;
}

static int m5_inter_max(const orc_crit *c) {   /* m5_input.c: ADAPT_INTER_MAX as in model 4 */
  if (c->inter0_on == 1 && c->inter1_on == 1) return c->inter0_lrs + c->inter1_lrs;
  if (c->inter0_on == 1 && c->inter1_on == 0) return c->inter0_lrs;
  return 0;
}

/* interface_location.c:7-90 with m5_facevalues: SI[0] marks the leaves with
 * a face across which phi[0] (the solid) changes sign, SI[1] those of the
 * other fractions, then the layers 2 .. ADAPT_INTER_MAX-1 in leaf order */
int orc_m5_interface_location(const orc_mesh *m, const orc_m5_params *p, const orc_crit *c, const double *S,
                              int *SI) {
  lay5_t L = mk_lay5(p->nmat);
  const size_t neq = (size_t)L.neq;
  for (int i = 0; i < m->n; i++) {
    SI[2 * i] = 0; SI[2 * i + 1] = 0;
    int flag = 0;
    for (int f = 0; f < 6 && !flag; f++) {
      if (m->bc[6 * i + f] != 0) continue;
      const int lfc = m->lfc_neigs[6 * i + f];
      const int nsfc = lfc > 1 ? lfc : 1;
      for (int k = 0; k < nsfc; k++) {
        const int j = m->neigtr[24 * i + 4 * f + k];
        for (int v = 0; v < L.n3; ++v) {
          if (m5_phi(&L, p, &S[i * neq], v) * m5_phi(&L, p, &S[j * neq], v) < 0.0) {
            if (v == 0) SI[2 * i] = 1; else SI[2 * i + 1] = 1;
            flag = 1;
          }
        }
      }
    }
  }
  for (int lvl = 2; lvl < m5_inter_max(c); lvl++) {
    for (int i = 0; i < m->n; i++) {
      int flag = 0;
      for (int f = 0; f < 6 && !flag; f++) {
        if (m->bc[6 * i + f] != 0) continue;
        const int lfc = m->lfc_neigs[6 * i + f];
        const int nsfc = lfc > 1 ? lfc : 1;
        for (int k = 0; k < nsfc; k++) {
          const int j = m->neigtr[24 * i + 4 * f + k];
          for (int v = 0; v < 2; ++v)
            if (SI[2 * j + v] == (lvl - 1) && SI[2 * i + v] == 0) { SI[2 * i + v] = lvl; flag = 1; }
        }
      }
    }
  }
  return 0;
}

static double m5_rel_jump(double x_cur, double x_nb) {   /* criterionsplt.c:81-88 */
  const double eps = 1.0e-8;
  if ((fabs(x_nb) > eps) || (fabs(x_cur) > eps)) return fabs(x_nb - x_cur) / MAXF(fabs(x_nb), fabs(x_cur));
  return 0.0;
}

/* criterionsplt.c:3-37 with criterion_grad (:40-140) on m5_facevalues cell
 * records: the mixture pressure sum_k alpha_k p_hydro_k, the density, and the
 * velocity criterion that reads the density jump (:113) */
int orc_m5_criterionsplt(const orc_mesh *m, const orc_m5_params *p, const orc_crit *c, const double *S,
                         const int *SI, int *split) {
  lay5_t L = mk_lay5(p->nmat);
  const size_t neq = (size_t)L.neq;
  for (int i = 0; i < m->n; i++) {
    int lt = 1;
    if (c->uniform_on == 1) lt = MAXF(lt, c->uniform_lvl);
    if (c->prs_on == 1 || c->rho_on == 1 || c->vel_on == 1) {
      double gp = 0.0, gr = 0.0, gv = 0.0, grho = 0.0, lg = 1;
      for (int f = 0; f < 6; f++) {
        if (m->bc[6 * i + f] != 0) continue;
        const int lfc = m->lfc_neigs[6 * i + f];
        const int nsfc = lfc > 1 ? lfc : 1;
        for (int k = 0; k < nsfc; k++) {
          const int j = m->neigtr[24 * i + 4 * f + k];
          rec5_t sL, sR;
          m5_record(&L, p, &S[i * neq], &sL);
          m5_record(&L, p, &S[j * neq], &sR);
          if (c->prs_on == 1) {
            double pl = 0.0, pr = 0.0;
            for (int iv = 0; iv < L.nmat; ++iv) { pl += sL.avf[iv] * sL.p_hydro[iv]; pr += sR.avf[iv] * sR.p_hydro[iv]; }
            gp = MAXF(m5_rel_jump(pl, pr), gp);
          }
          if (c->rho_on == 1) {
            grho = m5_rel_jump(sL.r, sR.r);
            gr = MAXF(grho, gr);
          }
          if (c->vel_on == 1) gv = MAXF(grho, gv);
        }
      }
      if (c->prs_on == 1 && fabs(gp) >= c->prs_thr) lg = c->prs_lvl;
      if (c->rho_on == 1 && fabs(gr) >= c->rho_thr) lg = c->rho_lvl;
      if (c->vel_on == 1 && fabs(gv) >= c->vel_thr) lg = c->vel_lvl;
      lt = (int)MAXF((double)lt, lg);
    }
    if (c->inter0_on != 0) {
      for (int v = 0; v < 2; ++v) {
        const int s = SI[2 * i + v];
        if (s >= 1 && s < c->inter0_lrs) lt = MAXF(lt, c->inter0_lvl);
        if (s >= c->inter0_lrs && s < c->inter1_lrs) lt = MAXF(lt, c->inter1_lvl);
      }
    }
    split[i] = lt;
  }
  return 0;
}
