/*
 * fv_oracle.h -- CPU restatement of ForestFV's model-4 explicit update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the HIP path is tested against:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing that is measured or shipped.
 *
 * It restates, function by function, the reference's per-leaf loops
 * (src/c/advancetimeexplicit.c, RK.c, m4_faceinterp.c, m4_reconctruct.c,
 * m4_residual.c, m4_hllc.c, m4_wave_speeds.c, m4_star_region.c, m4_flux.c,
 * m4_relaxationvec.c, interface_location.c, criterionsplt.c) on a plain
 * array image of the reference's LEAF/CELL/BRANCH data, keeping the
 * reference's operation order so that a gcc -O2 -ffp-contract=off build is
 * reproducible bit for bit.  Geometry is recomputed here from the octree
 * node coordinates exactly as the reference builds them (split.c node rule,
 * normalvector.c, tagvector.c, meshproperties.c, distance.c), independently
 * of the product's closed-form geometry tables.
 */
#ifndef FV_ORACLE_H
#define FV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXMAT 4
#define ORC_MAXNEQ 16

typedef struct {
  int    nmat;                 /* NMATERIALS */
  double gamma[ORC_MAXMAT];    /* MATERGAMA  */
  double pinf[ORC_MAXMAT];     /* MATERPINF  */
  double amin;                 /* AMIN       */
  double avf_lim;              /* AVF_LIM    */
  int    prelax;               /* PRELAX: 0, 1, 2 or 3 */
  int    rk;                   /* 2: RK_STEPS=2 as the reference runs it (cases 1,2)
                                  3: SSP-RK3 (cases 1,2,3)
                                  22: SSP-RK2 / Heun (cases 1,4) -- not a reference table */
  int    limiter;              /* 0: minmod (reference); 1: monotonised central; 2: per leaf (new) */
  double dt;                   /* DT */
  int    bc[6];                /* BC code of the -x,+x,-y,+y,-z,+z domain sides */
  int    fix;                  /* corrections of reference defects, ORC_FIX_* bits; 0 = literal */
  int    primtv;               /* PRIMTV: 0 reconstruct conservative, 1 primitive variables */
  int    bc_fixed_p;           /* 1: fixed-pressure ghost at codes 2 / 3 (m4_bc.c:44-113, 121-186) */
  double p_inlet, p_outlet;
} orc_params;

/* m4_cons2primtv.c:3-40 and m4_prim2conc.c:3-74: primitive vector
 * [alpha (nmat-1) | alpha rho_k | u v w | p_k], 3 nmat + 2 values */
void orc_cons2prim(const orc_params *p, const double *cons, double *prim);
void orc_prim2cons(const orc_params *p, const double *prim, double *cons);

/* The published model-4 residual is inconsistent with its own RK sign
 * convention in two places; each correction is a bit so that the literal
 * reference stays reproducible:
 *  ORC_FIX_NONCONS: the non-conservative alpha and alpha-rho-e terms are
 *    subtracted (m4_residual.c:158-176 adds them, while RK.c:25 applies
 *    S = SN - dt RHS / V), and the mirrored side reuses the star velocity
 *    vector as is (m4_residual.c:72-74 negates a global-frame vector);
 *  ORC_FIX_ENERGY: the star-state energy uses -s11K uK as the
 *    Rankine-Hugoniot balance requires (m4_star_region.c:46 has +s11K uK). */
#define ORC_FIX_NONCONS 1
#define ORC_FIX_ENERGY  2

typedef struct {
  int        n;                /* number of leaves, in DFS (locl) order */
  int        nroot[3];         /* brick of cubic root trees, x fastest */
  double     origin[3];
  double     rootsize;
  const int *root;             /* [n] */
  const int *level;            /* [n] reference level, root = 1 */
  const int *ci;               /* [n][3] integer position at own level inside the root */
  const int *lfc_neigs;        /* [n][6] 0 boundary, 1 same/coarser, 4 finer */
  const int *neigtr;           /* [n][6][4] leaf index of each (sub)face neighbour */
  const int *bc;               /* [n][6] face BC code, 0 = interior */
  const int *SI;               /* [n][2] interface layer flags */
} orc_mesh;

/* Per-leaf geometry as the reference stores it in struct CELL. */
typedef struct {
  double nx[6], ny[6], nz[6];
  double nxt1[6], nyt1[6], nzt1[6];
  double nxt2[6], nyt2[6], nzt2[6];
  double Area[6][4];
  double Vol;
  double dist_cf[6], dist_cc[6];
  double xc, yc, zc;
} orc_cell;

int  orc_neq(int nmat);

/* Geometry of every leaf (meshproperties.c + distance.c). cells: [n]. */
int  orc_geometry(const orc_mesh *m, orc_cell *cells);

/* One explicit step (advancetimeexplicit.c:3-29) repeated nsteps times.
 * S is [n][NEQ] in leaf order, updated in place.  Returns 0 or a negative
 * status (-1 NaN in RHS, -2 bad input). */
int  orc_advance(const orc_mesh *m, const orc_cell *cells, const orc_params *p,
                 double *S, int nsteps);
/* orc_advance with the leaf loops run by OpenMP over nthreads threads (0: the
 * OpenMP default); mirrored faces are recomputed instead of reused, which gives
 * the same bits as orc_advance.  orc_num_threads: threads a parallel region uses. */
int orc_advance_par(const orc_mesh *m, const orc_cell *cells, const orc_params *p, double *S, int nsteps,
                    int nthreads);
int orc_num_threads(void);

/* A single RK stage irk (1..) of the chosen table, given SN.  Exposes the
 * intermediate arrays for stage-level tests: SF [n][6][NEQ] and RHS [n][NEQ]
 * may be NULL. */
int  orc_stage(const orc_mesh *m, const orc_cell *cells, const orc_params *p,
               int irk, const double *SN, double *S, double *SF, double *RHS);

/* Face states only (m4_faceinterp_V1), SF [n][6][NEQ]. */
int  orc_faceinterp(const orc_mesh *m, const orc_cell *cells, const orc_params *p,
                    const double *S, double *SF);

/* Per-leaf limiter used when p->limiter == 2 (0 minmod, 1 MC); [n], kept by pointer. */
void orc_set_cell_limiter(const unsigned char *lim);

/* Pointwise pieces, for closed-form tests. */
void orc_relax(const orc_params *p, double *vec);                       /* m4_relaxationvec */
void orc_hllc_state(const orc_params *p, const double *sfL, const double *sfR,
                    const double *n, const double *t1, const double *t2,
                    int bc, double *flux /*[NEQ]*/, double *ustar /*[3]*/); /* m4_hllc + flux */
void orc_physical_flux(const orc_params *p, const double *vec, const double *n,
                       double *flux /*[NEQ]*/);                           /* m4_flux . n */

/* Interface layers (interface_location.c) and target levels
 * (criterionsplt.c + criterion_grad).  SI: [n][2] out, split: [n] out. */
typedef struct {
  int    uniform_on, uniform_lvl;
  int    prs_on, prs_lvl;   double prs_thr;
  int    rho_on, rho_lvl;   double rho_thr;
  int    vel_on, vel_lvl;   double vel_thr;
  int    inter0_on, inter0_lvl, inter0_lrs;
  int    inter1_on, inter1_lvl, inter1_lrs;
} orc_crit;
int  orc_interface_location(const orc_mesh *m, const orc_params *p, const orc_crit *c,
                            const double *S, int *SI);
int  orc_criterionsplt(const orc_mesh *m, const orc_params *p, const orc_crit *c,
                       const double *S, const int *SI, int *split);

/* criterionsmth (criterionsmth.c:13-56): LEVEL + 1 passes of sibling consensus,
 * the one-level refinement cap and 2:1 smoothing, in place on split [n], in
 * leaf order as the reference's list walk updates it. */
int  orc_criterionsmth(const orc_mesh *m, int LEVEL, int *split);
/* refineh (refineh.c:3-61; merge.c:20-55; split.c:425-434): merge the families
 * whose eight leaf kids all coarsen, then split the leaves with split > level;
 * the new leaves in DFS order with their state [NEQ], SI [2] and split.  The
 * outputs hold up to 8 n leaves; *nout receives the count. */
int  orc_refineh(const orc_mesh *m, const orc_params *p, const int *split, const double *S, int *nout,
                 int *root_o, int *level_o, int *ci_o, double *S_o, int *SI_o, int *split_o);
/* The boundary ghost of a face state (m4_bc.c:3-34 with bc_symmetry.c,
 * bc_reflective.c, bc_wall.c): the ghost velocity u,v,w [3] and the ghost's
 * conservative vector [NEQ] (its vec, which the published code leaves
 * unmirrored) for BC code `code` on a face with frame n, t1, t2. */
void orc_bc_ghost(const orc_params *p, const double *sf, const double *n, const double *t1, const double *t2,
                  int code, double *uvw, double *vec);

/* Largest stable time step dt = cfl * min_cells h / max_faces(|u.n| + c). */
double orc_cfl_dt(const orc_mesh *m, const orc_cell *cells, const orc_params *p,
                  const double *S, double cfl);

#ifdef __cplusplus
}
#endif
#endif
