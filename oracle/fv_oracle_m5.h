/*
 * fv_oracle_m5.h -- CPU restatement of ForestFV's model-5 ghost-fluid pieces:
 * the GFM face flux (m5_hllc_gfm.c, m5_wave_speeds_gfm.c, m5_flux_gfm.c,
 * m5_star_region.c), the interface-cell velocity re-projection (m5_gfm.c,
 * m5_gfm_rel.c) and the model-5 RK stage (m5_RK.c), with the first-order
 * face record they read (m5_facevalues.c / m5_facercnstr.c and the elastic
 * EOS of state_equation.c:143-316).
 *
 * TEST INFRASTRUCTURE ONLY, like fv_oracle.h: only tests/ may load it.
 *
 * Model-5 state, NEQ = 3 nmat + 12 (m5_input.c): [alpha rho_k (nmat) |
 * rho u, rho v, rho w | rho E | alpha_k (nmat-1) | alpha rho e_k (nmat) |
 * A (9, row-major)].
 */
#ifndef FV_ORACLE_M5_H
#define FV_ORACLE_M5_H

#include "fv_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_M5_MAXNEQ 21

typedef struct {
  int    nmat;                 /* NMATERIALS, 2 or 3 */
  double gamma[ORC_MAXMAT];    /* MATERGAMA */
  double pinf[ORC_MAXMAT];     /* MATERPINF */
  double mush[ORC_MAXMAT];     /* MATERMUSH: shear modulus, 0 for a fluid */
  double amin;                 /* AMIN */
  double avf_lim;              /* AVF_LIM */
  double dt;                   /* DT */
  int    zerou[3];             /* ZEROU0, ZEROU1, ZEROU2 */
} orc_m5_params;

int orc_m5_neq(int nmat);

/* Face record of one state vector (m5_facercnstr.c, m5_facevalues.c with
 * n_solids = 0).  rec layout, ORC_M5_REC doubles:
 *   0 r, 1 e, 2 u, 3 v, 4 w, 5 phi0,
 *   6.. ra[3], 9.. avf[3], 12.. are[3], 15.. c2Y[3], 18.. p_hydro[3],
 *   21.. Amat[9], 30.. st[9] (row-major) */
#define ORC_M5_REC 39
void orc_m5_record(const orc_m5_params *p, const double *vec, double *rec);

/* One sub-face of m5_residual.c:139-170 at a solid interface: the GFM HLLC
 * flux (m5_hllc_gfm.c:3-100) times the sub-face area, dotted with the face
 * normal, plus the star jump term; unc is the non-conservative face velocity.
 * frame: n, t1, t2 (9 doubles).  Returns 0, or -1 when a NaN reaches the
 * flux (the reference prints and exits). */
int orc_m5_gfm_face(const orc_m5_params *p, const double *UL, const double *UR, const double *frame,
                    double area, double *flux, double *unc);

/* m5_RK(irk) over all leaves in leaf order (m5_RK.c:3-126), with m5_gfm for
 * the leaves whose solid volume fraction crosses AVF_LIM in stage 1.
 * gfm_face: [n][6], 1 where the last residual used the GFM flux.
 * Returns 0, or -2 when gfm_face marks a boundary face. */
int orc_m5_rk(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, int irk, const double *SN,
              double *S, const double *RHS, const unsigned char *gfm_face);

/* m5_relaxation_solid.c:12-37 after m5_relaxationvec: S_before is the state
 * before relaxation, S the relaxed state (updated in place, leaf order);
 * leaves whose solid fraction crosses AVF_LIM are re-projected by
 * m5_gfm_rel (m5_gfm_rel.c). */
int orc_m5_gfm_rel(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, const double *S_before,
                   double *S, const unsigned char *gfm_face);

#ifdef __cplusplus
}
#endif
#endif
