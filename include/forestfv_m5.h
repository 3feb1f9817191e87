/*
 * forestfv_m5.h -- C ABI of the model-5 ghost-fluid operators on the GPU.
 *
 * Model 5 couples an elastic solid with fluids.  At a face where the solid
 * indicator phi = alpha_solid - AVF_LIM changes sign, the reference replaces
 * the right state's tangential velocity by the left one (a ghost state) before
 * its HLLC flux, and after the RK stage or the relaxation that moves a leaf
 * across the interface it re-projects that leaf's velocity onto the interface
 * normal.  The operator entry points replace those pieces one to one; the
 * fv_m5_dev_* entry points run the whole model-5 stage on the device.
 *
 * State layout, NEQ = 3 nmat + 12 (m5_input.c): [alpha rho_k (nmat) |
 * rho u, rho v, rho w | rho E | alpha_k (nmat - 1) | alpha rho e_k (nmat) |
 * A (9, row-major)].  Host arrays are row-major [count][NEQ]; every call copies
 * its inputs to the device, runs the kernels and copies the outputs back.
 * Every call returns FV_OK or an FV_ERR_* status (forestfv.h).
 */
#ifndef FORESTFV_M5_H
#define FORESTFV_M5_H

#include <stdint.h>
#include "forestfv.h"
#include "forestfv_dev.h"   /* FV_PREC_* */

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int    nmat;                 /* NMATERIALS: 2 or 3 */
  double gamma[FV_MAXMAT];     /* MATERGAMA */
  double pinf[FV_MAXMAT];      /* MATERPINF */
  double mush[FV_MAXMAT];      /* MATERMUSH: shear modulus, 0 for a fluid */
  double amin;                 /* AMIN */
  double avf_lim;              /* AVF_LIM */
  double dt;                   /* DT */
  int    zerou[3];             /* ZEROU0, ZEROU1, ZEROU2: freeze a velocity component */
  int    prelax;               /* PRELAX of m5_relaxationvec: 0, 1, 2 or 3 (stage entry points) */
  int    order;                /* ORDER: 1 cell states, 2 MUSCL face states (stage entry points) */
  int    rk_steps;             /* RK_STEPS: stages 1..rk_steps of m5_RK per step (fv_m5_dev_advance) */
  double yield;                /* von Mises yield stress of the solid (material 0) for perfect
                                  plasticity, 0 = elastic; new: the reference's plastic_rlx is
                                  commented out (m5_relaxation_solid.c:14) */
  int    fix;                  /* FV_M5_FIX_SIGNS: the consistent model-5 update (new, below), with
                                  FV_M5_FIX_MIXTURE: the volume-fraction-weighted elastic mixture
                                  (new, below); 0 = the published code */
  int    limiter;              /* MUSCL slope (new): 0 minmod (m5_reconctruct.c:93-140), 1 monotonised
                                  central, 2 per leaf from fv_m5_dev_set_limiter_flags (0 minmod,
                                  1 MC; every leaf minmod after a regrid until set again) */
} fv_m5_params;

/* Corrections of the published model-5 update, one mode (0 reproduces it):
 * the momentum flux rho u u - sigma (m5_flux.c:21-33 adds sigma); S = SN -
 * dt RHS / V (m5_RK.c:49,58-79 add), also in the stage-1 re-projection of
 * m5_gfm.c; the mirrored side of a face reuses the star velocity as it is
 * (m5_residual.c:62-103 negates it); the non-conservative sources with the
 * sign of that update, and for the deformation tensor n_i (A_kj u_k) - A_ij
 * (u.n) per face (m5_residual.c:176-209); +A in the supersonic advective A
 * flux (m5_flux.c:54-56); MUSCL face states whose alpha-rho-e follow their
 * reconstructed fractions (m5_input.c:196-204 keep the cell value); in
 * m5_star_region.c the star energy with -s11K uK over r* (u* - S_K) (:73),
 * + A_1j (vK - v*) in the star A (:84) and the phasic star energy from +p_k
 * with gamma - 1 (:118-125); and the relaxation scales A by cbrt(alpha0 /
 * alpha), so that det A follows the solid's density (m5_relaxationvec.c
 * scales by alpha / alpha0). */
#define FV_M5_FIX_SIGNS 1
/* A mixture model (new, set as a bit of fix beside FV_M5_FIX_SIGNS): the
 * reference gives every phase with a shear modulus the cell's whole elastic
 * energy and shear stress (state_equation.c:143-316 carry no volume fraction),
 * so a fluid cell holding a trace of solid shears with the solid's modulus
 * while the wave speeds carry only the hydrodynamic sound speed.  This bit
 * gives phase k the fraction alpha_k of both (mixture stress sum_k alpha_k
 * sigma_k), adds (4/3) alpha_k mu_k / rho to each phase's squared sound speed
 * in the HLLC and GFM wave speeds, and evaluates the relaxation's elastic
 * energy on its relaxed A. */
#define FV_M5_FIX_MIXTURE 2

int fv_m5_neq(int nmat);

/* Initial data (new; materials 0 solid, 1 liquid, 2 gas) in the form of
 * m5_init_domain's solid cases (m5_init_domain.c:175-216, init_S.c:205-266,
 * compute_Amat.c): a solid slab z < z_slab with a sharp face, a gas bubble of
 * the given radius and centre smeared with tanh over smear * hmin in the
 * liquid, every material at the cell's pressure (p_liquid, p_gas mixed with
 * the bubble profile), A = I in the solid and 0 in the fluids, at rest.
 * rho_init[3]: material densities; S: [n][NEQ]. */
int fv_m5_init_slab_bubble(const fv_mesh *m, const fv_m5_params *p, const double *rho_init, double z_slab,
                           const double centre[3], double radius, double p_liquid, double p_gas, double smear,
                           double *S);

/* GFM flux of nf sub-faces at a solid interface: replaces m5_hllc_gfm
 * (m5_hllc_gfm.c:3-100, with m5_wave_speeds_gfm.c, m5_flux_gfm.c and
 * m5_star_region.c) and its accumulation into flux_face
 * (m5_residual.c:166-170).  UL, UR: [nf][NEQ] face states of the two sides
 * (the states m5_facercnstr reads); frame: [nf][9] n, t1, t2 of the face;
 * area: [nf] sub-face area.  Out: flux [nf][NEQ] = area (F.n + F_hllc), and
 * unc [nf][3] the face velocity of the non-conservative terms (before the
 * division by the number of sub-faces).  FV_ERR_NAN if a flux is NaN. */
int fv_m5_gfm_faces(const fv_m5_params *p, int nf, const double *UL, const double *UR, const double *frame,
                    const double *area, double *flux, double *unc);

/* RK stage irk of model 5 over every leaf of pk: replaces m5_RK
 * (m5_RK.c:3-126), including the stage-1 interface re-projection m5_gfm
 * (m5_gfm.c:3-254) of the leaves whose solid fraction crosses AVF_LIM, in leaf
 * order as the reference runs it.  SN, S, RHS: [n][NEQ] (SN is written by
 * irk 0, S in place); gfm_face: [n][6], 1 where the stage's residual used the
 * GFM flux.  FV_ERR_ARG if gfm_face marks a boundary face. */
int fv_m5_rk(const fv_m5_params *p, const fv_pack *pk, int irk, double *SN, double *S, const double *RHS,
             const uint8_t *gfm_face);

/* Interface re-projection after relaxation: replaces the m5_gfm_rel step of
 * m5_relaxation_solid (m5_relaxation_solid.c:16-34, m5_gfm_rel.c:3-246).
 * S_before: [n][NEQ] state before relaxation; S: the relaxed state, updated
 * in place. */
int fv_m5_gfm_rel(const fv_m5_params *p, const fv_pack *pk, const double *S_before, double *S,
                  const uint8_t *gfm_face);

/* ---- the whole model-5 stage, device resident ----
 * advancetimeexplicit (advancetimeexplicit.c:3-29) with MODEL 5: per stage
 * m5_faceinterp_V1 (m5_faceinterp.c:60-93, ORDER 2), m5_residual
 * (m5_residual.c:7-238, via RHS.c:57-66), m5_RK (m5_RK.c:3-126) and
 * m5_relaxation_solid (m5_relaxation_solid.c:3-40 with m5_relaxationvec.c).
 * The leaves and face kinds are those of pk (its SI decides which faces reuse
 * the earlier leaf's flux, m5_residual.c:62-103); p->prelax, p->order and
 * p->rk_steps select the branches.  PRELAX 2 iterates on its own arrays (the
 * published F_pe reads RUN fields that are never written).  Every call
 * returns FV_OK or FV_ERR_*: FV_ERR_NAN where the reference prints NaN and
 * exits, FV_ERR_ARG for a GFM flag on a boundary face. */
typedef struct fv_m5_dev fv_m5_dev;
int fv_m5_dev_attach(const fv_m5_params *p, const fv_pack *pk, fv_m5_dev **out);
int fv_m5_dev_detach(fv_m5_dev *c);
/* LEAF.S of every leaf, row-major [n][NEQ] in leaf order */
int fv_m5_dev_set_state(fv_m5_dev *c, const double *S);
int fv_m5_dev_get_state(fv_m5_dev *c, double *S);
/* RK(0) (RK.c:6-15): SN = S */
int fv_m5_dev_begin_step(fv_m5_dev *c);
/* one stage irk (1..4, the m5_RK cases) after fv_m5_dev_begin_step */
int fv_m5_dev_stage(fv_m5_dev *c, int irk);
/* nsteps steps of stages 1..rk_steps (loop.c:21) */
int fv_m5_dev_advance(fv_m5_dev *c, int nsteps);
/* Precision of the stage (no counterpart in the reference, which is FP64):
 * FV_PREC_EXACT keeps the reference's operation order (the non-conservative
 * sources added face by face, m5_residual.c:176-214); FV_PREC_FP64 (default)
 * forms them once per leaf from the face velocities' sum over faces, equal to
 * rounding; FV_PREC_MIXED adds FP32 limited slopes in the MUSCL face states,
 * FP32 face records, HLLC, star states and GFM ghost flux; the records are
 * formed in FP64 (pressures carry p_inf up to 2e10 Pa), and the residual, RK
 * accumulation and relaxation stay FP64. */
int fv_m5_dev_set_precision(fv_m5_dev *c, int prec);
/* Per-leaf slope flags of limiter 2 ([n], 0 minmod, 1 MC; new: the
 * reference reconstructs with minmod only, m5_reconctruct.c:93-140). */
int fv_m5_dev_set_limiter_flags(fv_m5_dev *c, const uint8_t *lim);
int fv_m5_dev_get_limiter_flags(fv_m5_dev *c, uint8_t *lim);
/* The time step of later stages (DT of the params at attach until then). */
int fv_m5_dev_set_dt(fv_m5_dev *c, double dt);
/* CFL step of the device state (new: the reference takes a fixed DT):
 * cfl * min over leaves of h / (max|u_i| + c), with c^2 = sum_k c2Y_k +
 * (4/3) sum_k alpha_k mu_k / rho, the longitudinal sound speed of the mixture
 * (c2Y of m5_wave_speeds.c:20-40 is the hydrodynamic part). */
int fv_m5_dev_cfl_dt(fv_m5_dev *c, double cfl, double *dt);

/* ---- mesh adaptation of the model-5 context ----
 * fv_m5_dev_set_forest gives the context the leaf addresses of the mesh it
 * was attached with (m: the same leaves), after which:
 *  - fv_m5_dev_interface_location is interface_location as loop() calls it
 *    after every step (loop.c:23; interface_location.c:7-90 with
 *    m5_facevalues): SI of the device state ([n][2] to SI when not NULL) and
 *    the face kinds that follow it (the mirror rule of m5_residual.c:62-103);
 *  - fv_m5_dev_indicator gives SI and the target levels (criterionsplt.c:3-140
 *    on m5_facevalues records) without changing the mesh;
 *  - fv_m5_dev_regrid is meshadaptation (meshadaptation.c:3-44) on the
 *    device: indicator, criterionsmth (criterionsmth.c:13-56), merge with the
 *    mean of the eight kids and m5_relaxationvec (merge.c:20-59), split with
 *    copies (split.c:425-434), the new DFS order, connectivity and face kinds;
 *    *n_out receives the new leaf count;
 *  - fv_m5_dev_get_forest returns root [n], level [n], ci [n][3], SI [n][2]
 *    of the device mesh (any may be NULL). */
int fv_m5_dev_set_forest(fv_m5_dev *c, const fv_mesh *m);
int fv_m5_dev_get_forest(fv_m5_dev *c, int *root, int *level, int *ci, int *SI);
int fv_m5_dev_interface_location(fv_m5_dev *c, const fv_crit *cr, int *SI);
int fv_m5_dev_indicator(fv_m5_dev *c, const fv_crit *cr, int *SI, int *split);
int fv_m5_dev_regrid(fv_m5_dev *c, const fv_crit *cr, int maxlevel, int *n_out);

/* ---- the learned indicator on a model-5 context (new; config 4) ----
 * The network of fv_mlp_* (forestfv_dev.h) with the model-5 fields: density,
 * mixture pressure, liquid fraction (material 1) and |u| over the sound
 * speed, in the same 3x3x3 stencil features.  Need fv_m5_dev_set_forest.
 *  - fv_mlp_sample_m5 appends the features and the plain indicator's labels
 *    of every leaf to the training set;
 *  - fv_m5_dev_indicator_mlp gives SI, the targets (the network's, with the
 *    plain refinement forced) and the smooth flags, which become the per-leaf
 *    slope flags of limiter 2;
 *  - fv_m5_dev_regrid_mlp adapts the mesh to those targets and sets the new
 *    leaves' slope flags from the network. */
int fv_mlp_sample_m5(fv_m5_dev *c, fv_mlp *m, const fv_crit *cr);
int fv_m5_dev_indicator_mlp(fv_m5_dev *c, fv_mlp *m, const fv_crit *cr, int *SI, int *split, uint8_t *smooth);
int fv_m5_dev_regrid_mlp(fv_m5_dev *c, fv_mlp *m, const fv_crit *cr, int maxlevel, int *n_out);
/* the stage's pieces alone: face states of the current S (m5_faceinterp_V1),
 * and the residual of S and those face states (m5_residual) */
int fv_m5_dev_faceinterp(fv_m5_dev *c);
int fv_m5_dev_residual(fv_m5_dev *c);
/* the last face states SF [n][6][NEQ], residual RHS [n][NEQ] and GFM flags
 * gfm_face [n][6] (LEAF.SF, LEAF.RHS, LEAF.gfm_face); any may be NULL */
int fv_m5_dev_get_stage_arrays(fv_m5_dev *c, double *SF, double *RHS, uint8_t *gfm_face);
/* Measurement utility with no counterpart in the reference: runs nsteps
 * steps and returns the mean milliseconds per stage and per residual kernel
 * launch (hipEvents on the context's stream). */
int fv_m5_dev_time_steps(fv_m5_dev *c, int nsteps, double *ms_stage, double *ms_residual);

#ifdef __cplusplus
}
#endif
#endif
