/*
 * fv_oracle_m5.c -- CPU restatement of ForestFV's model-5 ghost-fluid pieces.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP model-5 operators (see
 * fv_oracle_m5.h).  Built into libfvoracle.so with gcc -O2 -ffp-contract=off.
 *
 * Every function names the reference lines it restates and keeps their
 * operation order.  pow(x,2.0) is written x*x as in fv_oracle.c; the
 * pow(|G|,1/3) and pow(|G|,1/6) of the elastic EOS stay glibc pow calls.
 */
#include "fv_oracle_m5.h"
#include <math.h>
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
/* face record (m5_facercnstr.c:3-190; m5_facevalues.c differs only in  */
/* the order of the two clips of avf_c_temp, which gives the same value) */
/* ------------------------------------------------------------------ */
static void m5_record(const lay5_t *L, const orc_m5_params *p, const double *S, rec5_t *r) {
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
  for (int i = 0; i < L->n3; ++i) { avf_temp[i] = S[L->i3 + i]; avf_c_temp[i] = S[L->i3 + i]; }
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
  for (int i = 0; i < L->n4; ++i) {
    double are_elast = (p->mush[i] != 0.0) ? elastic_energy_v2(G, Gmag, p->mush[i]) : 0.0;
    double g = p->gamma[i], pi = p->pinf[i];
    double press = ((g - 1.0) * (r->are[i] - are_elast) / avf_temp[i]) - g * pi;
    r->c2Y[i] = (avf_temp[i] / r->r) * g * (press + pi);
    r->p_hydro[i] = press;
    double s[3][3];
    eos_stress_v2(s, r->avf[i], r->p_hydro[i], G, Gmag, p->mush[i]);
    for (int a1 = 0; a1 < 3; ++a1)
      for (int b = 0; b < 3; ++b) r->st[a1][b] += s[a1][b];
  }
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
  double Em_S = (re * (uK - SK) + s11K * uK - s12K * vK - s13K * wK + s11_S * u_S + s12_S * v_S + s13_S * w_S) /
                (r_S * (u_S + SK));
  double A[3][3], t[9], AS[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = s->Amat[i * 3 + j];
  rot_tensor(t, A, F->nx, F->ny, F->nz, F->mx, F->my, F->mz, F->lx, F->ly, F->lz);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = t[i * 3 + j];
  for (int j = 0; j < 3; ++j)
    AS[0][j] = (A[0][j] * (uK - SK) - A[1][j] * (vK - v_S) + A[2][j] * (wK - w_S)) / (u_S - SK);
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
    double p_k = -s->p_hydro[v];
    double p_S = (p_k + pi) * (((g - 1.0) * rm + (g + 1.0) * rm_S) / ((g - 1.0) * rm_S + (g + 1.0) * rm)) - pi;
    double are_el = elastic_energy_v1(G, p->mush[v]);
    double e_el = are_el / (s->avf[v] * rm_S);
    e_S[v] = (p_S + g * pi) / (rm_S * (g + 1.0)) + e_el;
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
  const double dV = p->dt / cl->Vol;
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
            double as_new = SNc[e] + p->dt * buf;
            phi_new[v] = as_new - p->avf_lim;
            if (phi_old[v] * phi_new[v] < 0.0) flag_rk = 1;
          }
        }
        if (flag_rk == 1) {
          for (int e = 0; e < neq; e++) {
            if (e >= L.i2) {
              double buf = (1.0 / Vol) * R[e];
              Sc[e] = SNc[e] + p->dt * buf;
            }
          }
          int rc = m5_gfm(&L, p, m, &cells[c], c, S, RHS, gfm_face);
          if (rc) return rc;
        } else {
          for (int e = 0; e < neq; e++) {
            double buf = (1.0 / Vol) * R[e];
            Sc[e] = SNc[e] + p->dt * buf;
          }
        }
        break;
      }
      case 2:
        for (int e = 0; e < neq; e++) Sc[e] = (3.0 / 4.0) * SNc[e] + (1.0 / 4.0) * Sc[e] + (1.0 / 4.0) * p->dt * R[e] / Vol;
        break;
      case 3:
        for (int e = 0; e < neq; e++) Sc[e] = (1.0 / 3.0) * SNc[e] + (2.0 / 3.0) * Sc[e] + (2.0 / 3.0) * p->dt * R[e] / Vol;
        break;
      case 4:
        for (int e = 0; e < neq; e++) Sc[e] = (1.0 / 2.0) * SNc[e] + (1.0 / 2.0) * Sc[e] + (1.0 / 2.0) * p->dt * R[e] / Vol;
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
