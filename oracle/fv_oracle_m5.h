/*
 * fv_oracle_m5.h -- CPU restatement of ForestFV's model-5 stage: MUSCL face
 * states (m5_faceinterp.c, m5_reconctruct.c), the residual (m5_residual.c)
 * with its HLLC (m5_hllc.c, m5_wave_speeds.c, m5_flux.c) and GFM face fluxes
 * (m5_hllc_gfm.c, m5_wave_speeds_gfm.c, m5_flux_gfm.c, m5_star_region.c),
 * the RK stage (m5_RK.c) with the interface re-projection (m5_gfm.c), and
 * the relaxation (m5_relaxationvec.c, m5_relaxation_solid.c, m5_gfm_rel.c),
 * with the face records they read (m5_facevalues.c / m5_facercnstr.c and the
 * elastic EOS of state_equation.c:143-316).
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
  int    prelax;               /* PRELAX of m5_relaxationvec: 0, 1, 2 or 3 */
  int    order;                /* ORDER: 1 cell states, 2 MUSCL face states */
  int    rk_steps;             /* RK_STEPS: stages irk = 1..rk_steps of m5_RK */
  double yield;                /* von Mises yield stress of the solid, 0 = elastic (new) */
  int    fix;                  /* ORC_M5_FIX_SIGNS: the consistent model-5 update below; 0 = literal */
  int    limiter;              /* MUSCL slope (new): 0 minmod (m5_reconctruct.c:93-140), 1 MC,
                                  2 per leaf from orc_m5_set_cell_limiter (0 minmod, 1 MC) */
} orc_m5_params;

/* per-leaf slope choice of limiter 2 (a leaf-order array the caller keeps alive) */
void orc_m5_set_cell_limiter(const unsigned char *lim);

/* The published model-5 update is inconsistent with itself in several places;
 * ORC_M5_FIX_SIGNS takes the consistent form of each, all as one mode (0
 * reproduces the published code):
 *  - m5_flux.c:21-33 forms the momentum flux rho u u + sigma: rho u u - sigma,
 *    as m4_flux.c:21, m5_flux_gfm.c:22-34 and the contact speed use;
 *  - m5_RK.c:49,58-79 update S = SN + dt RHS / V while the residual sums the
 *    outward face fluxes (m5_residual.c:156-161): S = SN - dt RHS / V, also in
 *    the stage-1 re-projection of m5_gfm.c;
 *  - the mirrored side of a face reuses the star velocity vector as it is
 *    (m5_residual.c:62-103 negates the global-frame vector with the flux);
 *  - the non-conservative sources follow that sign: -alpha div u for the
 *    fractions and +alpha_k sigma_k : grad u in the phasic energies (the
 *    published terms negated), and for the deformation gradient
 *    A_ij: n_i (A_kj u_k) - A_ij (u.n) per face from the cell's A, the
 *    remainder of the flux A_kj u_k n_i of dA_ij/dt + u.grad A_ij +
 *    A_kj d_i u_k = 0 after the advective part u* A_S (m5_residual.c:197-209
 *    has a sign that differs from this in one term of each row);
 *  - the MUSCL face states scale each phase's alpha-rho-e with its
 *    reconstructed fraction, which keeps the phase's pressure at the face
 *    (m5_input.c:196-204 leave the block at the cell value while the
 *    fractions are reconstructed);
 *  - the advective A flux takes +A on the supersonic sides, as the star
 *    states do (m5_flux.c:54-56 negates it there only);
 *  - m5_star_region.c: the star energy with -s11K uK over r* (u* - S_K)
 *    (:73 has + s11K uK over r* (u* + S_K)), the first row of the star A with
 *    + A_1j (vK - v*) (:84 has -), and the phasic star energy from +p_k with
 *    (gamma - 1) (:118-125 take -p_hydro and gamma + 1);
 *  - the relaxation scales A by cbrt(alpha0 / alpha) of the solid, so that
 *    det A follows its density (m5_relaxationvec.c scales A by alpha / alpha0). */
#define ORC_M5_FIX_SIGNS 1
/* ORC_M5_FIX_MIXTURE (new, a mixture model rather than a correction): the
 * reference gives every phase with a shear modulus the whole elastic energy
 * and shear stress of the cell (state_equation.c:143-316 are not weighted by
 * the volume fraction), so a fluid cell holding a trace of solid shears with
 * the solid's modulus while the wave speeds carry only the hydrodynamic sound
 * speed.  This mode gives phase k the fraction alpha_k of both (the mixture
 * stress sum_k alpha_k sigma_k), adds (4/3) alpha_k mu_k / rho to the phase's
 * squared sound speed in the HLLC and GFM wave speeds, and evaluates the
 * relaxation's elastic energy on its relaxed A. */
#define ORC_M5_FIX_MIXTURE 2

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

/* ---- the whole model-5 stage (advancetimeexplicit.c:3-29 with MODEL 5) ----
 * Arrays are row-major in leaf order: S, SN, RHS [n][NEQ]; SF [n][6][NEQ];
 * gfm_face [n][6].  Status: 0, -1 NaN (where the reference prints and exits),
 * -2 bad input, -3 Newton failure of PRELAX 2. */

/* m5_relaxationvec.c:3-620 on one state vector (PRELAX 2 iterates on the
 * branch's own t0_m / p0_m / ar_R: the published F_pe reads RUN fields that
 * are never written) */
int orc_m5_relax(const orc_m5_params *p, double *vec);

/* m5_faceinterp_V1 (m5_faceinterp.c:60-93): MUSCL face states of every leaf
 * (m5_facevalues_MUSCL.c, m5_deltavalues.c:20-56, m5_reconctruct.c:93-190),
 * each relaxed by m5_relaxationvec.  PRIMTV = 0. */
int orc_m5_faceinterp(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, const double *S, double *SF);

/* m5_residual.c:7-238 over all leaves in leaf order, with the flux_flag
 * mirror protocol; SF is read when p->order == 2. */
int orc_m5_residual(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, const double *S,
                    const double *SF, double *RHS, unsigned char *gfm_face);

/* One stage irk: face states, residual, m5_RK(irk), m5_relaxation_solid.
 * SF, RHS and gfm_face receive the stage's intermediate arrays (may be NULL). */
int orc_m5_stage(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, int irk, const double *SN,
                 double *S, double *SF, double *RHS, unsigned char *gfm_face);

/* Perfect plasticity of the solid (material 0), new: the reference's
 * plastic_rlx (m5_relaxation_solid.c:14) is commented out and undefined.
 * With G = A A^T = Q diag(l) Q^T and d = det G, the elastic state is moved
 * along l_i(t) = d^(1/3) (l_i / d^(1/3))^t, t in [0, 1] (volume preserving:
 * det G stays d; t = 0 is stress free), to the t where the von Mises stress of
 * the deviatoric stress of state_equation.c:143-204 equals the yield stress;
 * A becomes Q diag(sqrt(l(t) / l)) Q^T A.  Below yield nothing changes.  The
 * energies are kept, so the released elastic energy becomes heat through the
 * pressure recovery.  Returns 1 when A changed, 0 when not. */
int orc_m5_plastic(const orc_m5_params *p, double *vec);

/* nsteps explicit steps: SN = S (RK.c case 0), then stages 1..rk_steps. */
int orc_m5_advance(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, double *S, int nsteps);

/* Indicator with MODEL 5: interface_location.c:7-90 and criterionsplt.c:3-140
 * on m5_facevalues cell records (phi[v] = clipped alpha_v - AVF_LIM; the
 * mixture pressure of the elastic EOS).  SI [n][2] and split [n] out. */
int orc_m5_interface_location(const orc_mesh *m, const orc_m5_params *p, const orc_crit *c, const double *S,
                              int *SI);
int orc_m5_criterionsplt(const orc_mesh *m, const orc_m5_params *p, const orc_crit *c, const double *S,
                         const int *SI, int *split);
/* refineh with MODEL 5: orc_refineh's merge and split passes on NEQ = 3 nmat
 * + 12 states, a merged parent relaxed with m5_relaxationvec (merge.c:54-59). */
int orc_m5_refineh(const orc_mesh *m, const orc_m5_params *p, const int *split, const double *S, int *nout,
                   int *root_o, int *level_o, int *ci_o, double *S_o, int *SI_o, int *split_o);

/* CFL step (new): cfl * min h / (max|u_i| + c), c^2 = sum_k c2Y_k + (4/3)
 * sum_k alpha_k mu_k / rho, the longitudinal sound speed of the mixture. */
double orc_m5_cfl_dt(const orc_mesh *m, const orc_cell *cells, const orc_m5_params *p, const double *S, double cfl);

/* One face of m5_residual.c:139-170 between two face states (m5_facercnstr
 * records: face state X with phi from the cell state C), by m5_hllc or, when
 * phi changes sign, m5_hllc_gfm; bc is the face's BC code (0 interior, the
 * right state then mirrors the left one).  flux = area (F.n + F_hllc), unc
 * the non-conservative face velocity.  Returns 1 when the GFM flux was used,
 * 0 otherwise, -1 on NaN. */
int orc_m5_face(const orc_m5_params *p, const double *XL, const double *CL, const double *XR, const double *CR,
                const double *frame, int bc, double area, double *flux, double *unc);

#ifdef __cplusplus
}
#endif
#endif
