/*
 * fv_init.c -- host-side physics used outside the time step: the state
 * layout, pressure relaxation for cells merged at regrid time, and initial
 * conditions.
 */
#include "forestfv.h"
#include "forestfv_m5.h"
#include <math.h>
#include <stdint.h>
#include <string.h>

#define SQ(x) ((x) * (x))

/* NEQ = NMATERIALS ar + 3 ru + 1 rE + (NMATERIALS-1) alpha + NMATERIALS are
 * (m4_input.c:121-131); layout eqtypi = {0, nm, nm+3, nm+4, 2nm+3} */
int fv_neq(int nmat) { return 3 * nmat + 3; }

/* m4_relaxationvec.c:3-505 for PRELAX 0, 1 and 3, applied in place to one
 * conservative vector.  Used on cells created by a merge (merge.c:52-55). */
void fv_relax_host(const fv_params *p, double *vec) {
  const int nm = p->nmat, i1 = nm, i2 = nm + 3, i3 = nm + 4, i4 = 2 * nm + 3, n3 = nm - 1;
  const double *G = p->gamma, *P = p->pinf;
  const double a = p->amin / 100.0;
  double ar[FV_MAXMAT], are[FV_MAXMAT], avf[FV_MAXMAT + 1], rho = 0.0, sum, dif;
  for (int k = 0; k < nm; k++) { ar[k] = vec[k]; are[k] = vec[i4 + k]; rho += ar[k]; }
  double u = vec[i1] / rho, v = vec[i1 + 1] / rho, w = vec[i1 + 2] / rho, rE = vec[i2];
  for (int k = 0; k < n3; k++) avf[k] = vec[i3 + k];
#define VFLIM()                                                                  \
  do {                                                                           \
    sum = 0.0;                                                                   \
    for (int k = 0; k < n3; k++) {                                               \
      if (avf[k] <= a) avf[k] = a;                                               \
      if (avf[k] >= (1.0 - a)) avf[k] = 1.0 - a;                                 \
      sum += avf[k];                                                             \
    }                                                                            \
    if (sum >= (1.0 - a)) {                                                      \
      dif = sum - (1.0 - a);                                                     \
      for (int k = 0; k < n3; k++) { double per = avf[k] / sum; avf[k] = avf[k] - per * dif; } \
      sum = 1.0 - a;                                                             \
    }                                                                            \
    avf[n3] = 1.0 - sum;                                                         \
  } while (0)
  VFLIM();
  if (p->prelax == 1 && nm == 2) {
    double t0[2], p0[2];
    for (int k = 0; k < 2; k++) {
      t0[k] = 1.0 / (ar[k] / avf[k]);
      p0[k] = (are[k] * (G[k] - 1.0) - avf[k] * G[k] * P[k]) / avf[k];
    }
    double A1 = (avf[0] / G[0]) * (p0[0] + P[0]) / ((avf[0] / G[0]) + (avf[1] / G[1]));
    double A2 = (avf[1] / G[1]) * (p0[1] + P[1]) / ((avf[0] / G[0]) + (avf[1] / G[1]));
    double pe = 0.5 * (A1 + A2 - (P[0] + P[1])) + sqrt(0.25 * (SQ((A2 - A1 - (P[1] - P[0])))) + A1 * A2);
    for (int k = 0; k < 2; k++) {
      double t1 = t0[k] * (p0[k] + G[k] * P[k] + (G[k] - 1.0) * pe) / (pe + G[k] * P[k] + (G[k] - 1.0) * pe);
      avf[k] = ar[k] / (1.0 / t1);
    }
    VFLIM();
  } else if (p->prelax == 2) {
    /* m4_relaxationvec.c:128-297 with F_pe / DF_pe (F_pe.c:18-27, 64-73) on the
     * branch's local arrays (the published F_pe reads never-written RUN fields) */
    const double eb = 1.0e-8;
    double t0[FV_MAXMAT], p0[FV_MAXMAT], max_avf = -1.0;
    int principal = 0;
    for (int k = 0; k < nm; k++) {
      t0[k] = 1.0 / (ar[k] / avf[k]);
      p0[k] = (are[k] * (G[k] - 1.0) - avf[k] * G[k] * P[k]) / avf[k];
      if (p0[k] < -(1.0 - eb) * P[k] + eb) p0[k] = -(1.0 - eb) * P[k] + 1.0;
      if (avf[k] > max_avf) { max_avf = avf[k]; principal = k; }
    }
#define FPE(pe_, out_)                                                                      \
    do {                                                                                    \
      double F_ = 0.0;                                                                      \
      for (int k = 0; k < nm; k++)                                                          \
        F_ += ar[k] * t0[k] * ((p0[k] + G[k] * P[k] + (G[k] - 1.0) * (pe_)) /                \
                               ((pe_) + G[k] * P[k] + (G[k] - 1.0) * (pe_)));               \
      (out_) = F_ + -1.0;                                                                   \
    } while (0)
#define DFPE(pe_, out_)                                                                     \
    do {                                                                                    \
      double D_ = 0.0;                                                                      \
      for (int k = 0; k < nm; k++) {                                                        \
        const double c1_ = ar[k] * t0[k];                                                   \
        const double f_ = (p0[k] + G[k] * P[k] + (G[k] - 1.0) * (pe_));                      \
        const double g_ = ((pe_) + G[k] * P[k] + (G[k] - 1.0) * (pe_));                      \
        D_ += (c1_ * (G[k] - 1.0) * g_ - c1_ * f_ * G[k]) / SQ(g_);                         \
      }                                                                                     \
      (out_) = D_;                                                                          \
    } while (0)
    double pe_old = 0.001, pe_new = pe_old, Fv, Dv;
    int it = 0;
    double c1 = fabs(pe_new - pe_old);
    FPE(pe_new, Fv);
    double c2 = fabs(Fv);
    while (((c2 >= 1.0e-12) || (c1 >= 1.0e-4)) && (it <= 101)) {
      FPE(pe_old, Fv);
      DFPE(pe_old, Dv);
      pe_new = pe_old - 0.8 * (Fv / Dv);
      c1 = fabs(pe_new - pe_old);
      FPE(pe_new, Fv);
      c2 = fabs(Fv);
      if (it >= 100 && ((fabs(pe_old - pe_new) < 1.0e3) || isnan(pe_new))) { pe_new = NAN; break; }
      pe_old = pe_new;
      it++;
    }
#undef FPE
#undef DFPE
    const double Pp = P[principal];
    if (pe_new < (-(1.0 - eb) * Pp + eb)) pe_new = -(1.0 - eb) * Pp + 1.0;
    for (int k = 0; k < nm; k++) {
      double t1 = t0[k] * (p0[k] + G[k] * P[k] + (G[k] - 1.0) * pe_new) / (pe_new + G[k] * P[k] + (G[k] - 1.0) * pe_new);
      avf[k] = ar[k] / (1.0 / t1);
    }
    VFLIM();
  } else if (p->prelax == 3) {
    const double es = 1.0e-12, eb = 1.0e-8;
    double rk[FV_MAXMAT], pk[FV_MAXMAT];
    for (int k = 0; k < nm; k++) {
      if (avf[k] < 0.0 || ar[k] < 0.0) { avf[k] = es; ar[k] = es; are[k] = es; }
      if (avf[k] >= 1.0) avf[k] = 1.0 - es;
    }
    int irelax = 1;
    for (int k = 0; k < nm; k++) if (avf[k] > 1.0 - es) irelax = 0;
    if (irelax) {
      double pr = 0.0;
      for (int k = 0; k < nm; k++) {
        if (avf[k] > es) {
          pk[k] = (are[k]) / avf[k] * (G[k] - 1.0) - G[k] * P[k];
          if (pk[k] <= -(1.0 - eb) * P[k] + eb) pk[k] = -(1.0 - eb) * P[k] + eb;
        } else pk[k] = 0.0;
        pr += avf[k] * pk[k];
      }
      int it = 0;
      double F = 1.0e-9, dF = 1.0e9;
      for (int k = 0; k < nm; k++) rk[k] = 0.0;
      while (fabs(F) > 1.0e-12 && it < 500) {
        pr = pr - F / dF;
        it++;
        for (int k = 0; k < nm; k++) if (pr < -(1.0 - eb) * P[k] + eb) pr = -(1.0 - eb) * P[k] + 1.0;
        F = -1.0; dF = 0.0;
        for (int k = 0; k < nm; k++) {
          if (avf[k] > es) {
            rk[k] = ar[k] / (avf[k] > es ? avf[k] : es) * pow((pr + P[k]) / (pk[k] + P[k]), 1.0 / G[k]);
            F = F + ar[k] / rk[k];
            dF = dF - ar[k] / (G[k] * rk[k] * (pr + P[k]));
          }
        }
      }
      for (int k = 0; k < nm; k++) if (avf[k] > es) avf[k] = ar[k] / rk[k];
      VFLIM();
    }
  }
  if ((p->prelax == 1 && nm == 2) || p->prelax == 2 || p->prelax == 3) {
    double e_hyd = rE / rho - 0.5 * (SQ((u)) + SQ((v)) + SQ((w)));
    double num = 0.0, den = 0.0;
    for (int k = 0; k < nm; k++) {
      num += avf[k] * G[k] * P[k] / (G[k] - 1.0);
      den += avf[k] / (G[k] - 1.0);
    }
    double pr = (rho * e_hyd - num) / den;
    for (int k = 0; k < nm; k++) are[k] = (avf[k] * (pr + G[k] * P[k]) / (G[k] - 1.0));
  }
#undef VFLIM
  for (int k = 0; k < n3; k++) vec[i3 + k] = avf[k];
  for (int k = 0; k < nm; k++) vec[i4 + k] = are[k];
}

/* init_S.c:70-107: pure material `mat` (others at AMIN) at pressure pres */
void fv_init_pure(const fv_params *p, const double *rho_init, int mat, double pres,
                  const double u[3], double *S) {
  const int nm = p->nmat, i1 = nm, i2 = nm + 3, i3 = nm + 4, i4 = 2 * nm + 3;
  double rho = 0.0, avf;
  S[i2] = 0.0;
  for (int k = 0; k < nm; k++) {
    avf = (mat == k) ? 1.0 - (double)(nm - 1) * p->amin : p->amin;
    S[k] = avf * rho_init[k];
    rho += avf * rho_init[k];
    double are = avf * (pres + p->gamma[k] * p->pinf[k]) / (p->gamma[k] - 1.0);
    S[i2] += are;
    S[i4 + k] = are;
  }
  S[i1 + 0] = u[0] * rho;
  S[i1 + 1] = u[1] * rho;
  S[i1 + 2] = u[2] * rho;
  S[i2] += 0.5 * rho * (SQ(u[0]) + SQ(u[1]) + SQ(u[2]));
  for (int k = 0; k < nm - 1; k++) S[i3 + k] = (mat == k) ? 1.0 - (double)(nm - 1) * p->amin : p->amin;
}

int fv_init_split_x(const fv_mesh *m, const fv_params *p, const double *rho_init,
                    const double *p_init, double xs, double *S) {
  int neq = fv_neq(p->nmat);
  const double u0[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < m->n; i++) {
    double c[3];
    fv_mesh_centre(m, i, c);
    int mat = (c[0] > xs) ? 1 : 0;        /* m4_init_domain.c:22-28 */
    fv_init_pure(p, rho_init, mat, p_init[mat], u0, &S[(size_t)i * neq]);
  }
  return FV_OK;
}

/* splitmix64: the documented generator behind the seeded perturbation */
static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Spherical bubble of material 1 (gas) in material 0 (liquid).  The gas
 * volume fraction is 1/2 (1 - tanh((r - R) / (smear h_min))) clipped to
 * [AMIN, 1 - AMIN], the pressure blends p_init[1] inside and p_init[0]
 * outside with the same profile, velocity is zero, and a seeded smooth
 * perturbation of relative size `perturb` multiplies the pressure. */
int fv_init_bubble(const fv_mesh *m, const fv_params *p, const double *rho_init,
                   const double *p_init, const double centre[3], double radius,
                   double smear, unsigned seed, double perturb, double *S) {
  if (p->nmat != 2) return FV_ERR_ARG;
  int neq = fv_neq(2);
  double hmin = m->rootsize / (double)(1 << (m->maxlevel - 1));
  double L = m->rootsize * (double)m->nroot[0];
  uint64_t st = (uint64_t)seed;
  double ph[3];
  for (int a = 0; a < 3; a++) ph[a] = (double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
  const double twopi = 6.283185307179586;
  for (int i = 0; i < m->n; i++) {
    double c[3];
    fv_mesh_centre(m, i, c);
    double r = sqrt(SQ(c[0] - centre[0]) + SQ(c[1] - centre[1]) + SQ(c[2] - centre[2]));
    double ag = 0.5 * (1.0 - tanh((r - radius) / (smear * hmin)));
    if (ag < p->amin) ag = p->amin;
    if (ag > 1.0 - p->amin) ag = 1.0 - p->amin;
    double al = 1.0 - ag;
    double pr = p_init[1] * ag + p_init[0] * al;
    if (perturb != 0.0)
      pr *= 1.0 + perturb * sin(twopi * ((c[0] - m->origin[0]) / L + ph[0])) *
                            sin(twopi * ((c[1] - m->origin[1]) / L + ph[1])) *
                            sin(twopi * ((c[2] - m->origin[2]) / L + ph[2]));
    double *s = &S[(size_t)i * neq];
    double avf[2] = {al, ag};
    double rE = 0.0;
    for (int k = 0; k < 2; k++) {
      s[k] = avf[k] * rho_init[k];
      double are = avf[k] * (pr + p->gamma[k] * p->pinf[k]) / (p->gamma[k] - 1.0);
      s[7 + k] = are;
      rE += are;
    }
    s[2] = 0.0; s[3] = 0.0; s[4] = 0.0;
    s[5] = rE;
    s[6] = al;
  }
  return FV_OK;
}

/* Model-5 initial data of a solid slab under a liquid with a gas bubble (new;
 * materials 0 solid, 1 liquid, 2 gas).  A cell of one material follows
 * m5_init_S (init_S.c:205-266): that material holds 1 - (nmat-1) AMIN of the
 * volume and the others AMIN, every material is at the cell's pressure, and
 * A = I where the solid fraction exceeds 10 AMIN, else A = 0
 * (compute_Amat.c:3-40), so that fluid cells carry no elastic energy or
 * shear stress.  The slab z < z_slab is solid with a sharp face (the GFM
 * level set phi = alpha_solid - AVF_LIM changes sign there); above it the gas
 * fraction is the model-4 bubble's tanh profile over smear * hmin
 * (fv_init_bubble) and the pressure mixes linearly with it.  At rest. */
int fv_m5_init_slab_bubble(const fv_mesh *m, const fv_m5_params *p, const double *rho_init, double z_slab,
                           const double centre[3], double radius, double p_liquid, double p_gas, double smear,
                           double *S) {
  if (p->nmat != 3 || !(radius > 0.0) || !(smear > 0.0)) return FV_ERR_ARG;
  const int nm = 3, neq = 3 * nm + 12;
  const int i1 = nm, i2 = nm + 3, i3 = nm + 4, i4 = i3 + nm - 1, i5 = i4 + nm;
  const double amin = p->amin;
  const double hmin = m->rootsize / (double)(1 << (m->maxlevel - 1));
  for (int i = 0; i < m->n; i++) {
    double c[3];
    fv_mesh_centre(m, i, c);
    double avf[3], pr;
    if (c[2] < z_slab) {
      avf[0] = 1.0 - (double)(nm - 1) * amin; avf[1] = amin; avf[2] = amin;
      pr = p_liquid;
    } else {
      double r = sqrt(SQ(c[0] - centre[0]) + SQ(c[1] - centre[1]) + SQ(c[2] - centre[2]));
      double ag = 0.5 * (1.0 - tanh((r - radius) / (smear * hmin)));
      if (ag < amin) ag = amin;
      if (ag > 1.0 - 2.0 * amin) ag = 1.0 - 2.0 * amin;
      avf[0] = amin; avf[2] = ag; avf[1] = 1.0 - amin - ag;
      pr = p_gas * ag + p_liquid * (1.0 - ag);
    }
    double *s = &S[(size_t)i * neq];
    for (int v = 0; v < neq; v++) s[v] = 0.0;
    double rE = 0.0;
    for (int k = 0; k < nm; k++) {
      s[k] = avf[k] * rho_init[k];
      /* A is I or 0, so the elastic energy of every material is 0 */
      double are = avf[k] * (pr + p->gamma[k] * p->pinf[k]) / (p->gamma[k] - 1.0);
      s[i4 + k] = are;
      rE += are;
    }
    s[i1 + 0] = 0.0; s[i1 + 1] = 0.0; s[i1 + 2] = 0.0;
    s[i2] = rE;
    for (int k = 0; k < nm - 1; k++) s[i3 + k] = avf[k];
    if (avf[0] > 10.0 * amin) { s[i5 + 0] = 1.0; s[i5 + 4] = 1.0; s[i5 + 8] = 1.0; }
  }
  return FV_OK;
}
