/*
 * forestfv.h -- host side of the MI355X ForestFV explicit update.
 *
 * C structures and functions that own the forest of octrees on the host
 * (leaf list, connectivity, AMR) and pack it for the device.  The device
 * entry points that replace the reference's stage functions are declared in
 * forestfv_dev.h.
 *
 * Conventions kept from the reference (src/c/strdata.h:219-361):
 *  - root trees are axis-aligned cubic hexahedra; a root has level 1;
 *  - a hex splits into 8 kids ordered ikid = i + 2j + 4k (brchconn.c:23-26);
 *  - leaves are listed depth-first by kid index, root trees in order
 *    (rebuildlist.c:17-77);
 *  - faces are numbered f0 y-, f1 x+, f2 y+, f3 x-, f4 z-, f5 z+
 *    (operatetp.c:60-69), opposite pairs 0-2, 1-3, 4-5 (operatetp.c:268-279);
 *  - the tag of a leaf is sum_{l=2..level} 10^(l-2) (1 + ikid_l)
 *    (tagaddress.c:3-11), so at most FV_MAXLEVEL levels fit in an int32;
 *  - lfc_neigs[f] is 0 on a domain boundary, 1 for a same-level or coarser
 *    neighbour, 4 for four finer neighbours listed in fcrefnd2elnd order
 *    (elconnfin.c:14-22).
 */
#ifndef FORESTFV_H
#define FORESTFV_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FV_MAXMAT   4
#define FV_MAXNEQ   16
#define FV_MAXLEVEL 10

/* status codes returned by every entry point (the reference calls exit(0)) */
#define FV_OK          0
#define FV_ERR_NAN    -1   /* NaN in a residual or state (m4_residual.c:182-190) */
#define FV_ERR_ARG    -2   /* bad argument or unsupported configuration */
#define FV_ERR_HIP    -3   /* HIP runtime error */
#define FV_ERR_MESH   -4   /* octree not 2:1 face balanced, or too deep */
#define FV_ERR_NOMEM  -5

/* face kinds carried per cell in the device info word */
#define FV_FK_BOUNDARY 0
#define FV_FK_SAME     1   /* same-level neighbour; this cell computes the flux */
#define FV_FK_MIRROR   2   /* same-level neighbour earlier in leaf order computes it (m4_residual.c:62-85) */
#define FV_FK_COARSE   3   /* neighbour one level coarser (2:1) */
#define FV_FK_FINE     4   /* four neighbours one level finer (1:4) */

/* RK stage tables (RK.c:17-45) */
#define FV_RK_REF2  2      /* RK_STEPS=2 exactly as the reference runs it: cases 1,2 */
#define FV_RK_SSP3  3      /* RK_STEPS=3: SSP-RK3, cases 1,2,3 */
#define FV_RK_SSP2 22      /* SSP-RK2 (Heun): cases 1,4 -- not a reference table */

/* Model-4 parameters (globals of strdata.h:47-171 and m4_input.c). */
typedef struct {
  int    nmat;                /* NMATERIALS */
  double gamma[FV_MAXMAT];    /* MATERGAMA  */
  double pinf[FV_MAXMAT];     /* MATERPINF  */
  double amin;                /* AMIN       */
  double avf_lim;             /* AVF_LIM    */
  int    prelax;              /* PRELAX: 0 none, 1 closed-form two-phase, 2 damped Newton on F_pe, 3 Newton */
  int    rk;                  /* FV_RK_* */
  int    limiter;             /* 0 minmod (m4_reconctruct.c:114-119), 1 MC (new), 2 per leaf (new) */
  double dt;                  /* DT */
  int    bc[6];               /* BC code of the -x,+x,-y,+y,-z,+z sides (BCOVERWRT) */
  int    fix;                 /* FV_FIX_* corrections of reference defects; 0 = literal reference */
  int    primtv;              /* PRIMTV: 0 conservative, 1 primitive reconstruction (m4_reconctruct.c:126-181) */
  int    bc_fixed_p;          /* 1: the ghost state of an outlet (code 2) / inlet (code 3) face holds every
                                 material at pressure p_outlet / p_inlet, as m4_bc_outlet / m4_bc_inlet do for
                                 CASE 3400 and 3410 (m4_bc.c:44-113, 121-186; 1e6 and 150e6 Pa there);
                                 0: those faces copy the own face state (every other CASE) */
  double p_inlet, p_outlet;
} fv_params;

/* Corrections of the published model-4 residual (0 reproduces it literally):
 *  FV_FIX_NONCONS: subtract the non-conservative alpha / alpha-rho-e terms
 *    (m4_residual.c:158-176 adds them against RK.c:25's S = SN - dt RHS / V)
 *    and reuse the star velocity unnegated on the mirrored side
 *    (m4_residual.c:72-74 negates a global-frame vector);
 *  FV_FIX_ENERGY: star-state energy with -s11K uK (m4_star_region.c:46 has +). */
#define FV_FIX_NONCONS 1
#define FV_FIX_ENERGY  2
#define FV_FIX_ALL     3

/* Target-level criteria (criterionsplt.c, m4_input.c:32-56). */
typedef struct {
  int    uniform_on, uniform_lvl;
  int    prs_on, prs_lvl;   double prs_thr;
  int    rho_on, rho_lvl;   double rho_thr;
  int    vel_on, vel_lvl;   double vel_thr;
  int    inter0_on, inter0_lvl, inter0_lrs;
  int    inter1_on, inter1_lvl, inter1_lrs;
} fv_crit;

/* Forest of octrees over a brick of nroot[0] x nroot[1] x nroot[2] cubic
 * roots (root r at brick position r % nx, (r / nx) % ny, r / (nx ny)).
 * Every per-leaf array is in DFS leaf order. */
typedef struct fv_mesh {
  int     n;                  /* leaves */
  int     nroot[3];
  double  origin[3];
  double  rootsize;
  int     maxlevel;           /* deepest leaf level */
  int    *root;               /* [n] */
  int    *level;              /* [n] */
  int    *ci;                 /* [n][3] integer position at own level inside its root */
  int    *tag;                /* [n] decimal tag (tagaddress.c) */
  int    *lfc_neigs;          /* [n][6] */
  int    *neigtr;             /* [n][6][4] leaf indices, -1 unused */
  int    *bc;                 /* [n][6] BC code, 0 interior */
  int    *SI;                 /* [n][2] interface layers */
  int     bcside[6];
  /* lookup (root, level, i, j, k) -> leaf */
  uint64_t *hkey;
  int      *hval;
  size_t    hcap;
} fv_mesh;

/* Per-level geometry of axis-aligned cubic cells and the face frames,
 * computed with the reference's formulas (normalvector.c, tagvector.c,
 * meshproperties.c, distance.c) on a cube of the level's edge length. */
typedef struct {
  double frame[6][9];                 /* n, t1, t2 of each face index */
  double h[FV_MAXLEVEL + 2];          /* edge length */
  double area[FV_MAXLEVEL + 2];       /* face area */
  double vol[FV_MAXLEVEL + 2];        /* volume */
  double dcf[3][FV_MAXLEVEL + 2];     /* dist_cf: [0] same level, [1] finer nbr, [2] coarser nbr */
  double dcc[3][FV_MAXLEVEL + 2];     /* dist_cc */
} fv_geo;

/* Device-bound image of a mesh: plain arrays, owned by the caller. */
typedef struct {
  int            n;
  int           *nbr;      /* [6][n] same/coarse: leaf index; fine: row of fine4; boundary: -bc */
  uint32_t      *info;     /* [n] bits 0-3 level, bit 4 SI[0]!=0, bits 8+3f face kind of face f */
  int            nfine;
  int           *fine4;    /* [nfine][4] the four finer neighbours in neigtr order */
  fv_geo         geo;
} fv_pack;

/* ---- mesh ---- */
fv_mesh *fv_mesh_uniform(const int nroot[3], const double origin[3], double rootsize,
                         int level, const int bcside[6]);
/* Leaves in any order; sorted to DFS order.  Returns NULL and sets *status
 * (FV_ERR_MESH) when the forest is not 2:1 face balanced. */
fv_mesh *fv_mesh_from_leaves(const int nroot[3], const double origin[3], double rootsize,
                             const int bcside[6], int n, const int *root, const int *level,
                             const int *ci, int *status);
void     fv_mesh_free(fv_mesh *m);
int      fv_mesh_find(const fv_mesh *m, int root, int level, int i, int j, int k);
void     fv_mesh_centre(const fv_mesh *m, int leaf, double c[3]);
int      fv_tag_from_address(int level, const int *ci);
int      fv_address_from_tag(int tag, int *level, int *ci);

/* ---- device image ---- */
int      fv_mesh_pack(const fv_mesh *m, fv_pack *pk);
void     fv_pack_free(fv_pack *pk);
void     fv_geo_build(double rootsize, fv_geo *g);

/* ---- AMR (criterionsmth.c, refineh.c, split.c, merge.c) ---- */
/* LEVEL+1 passes of sibling consensus and 2:1 smoothing, in leaf order. */
int      fv_criterionsmth(const fv_mesh *m, int LEVEL, int *split);
/* Merge then split by target levels; S ([n][NEQ], leaf order) is remapped.
 * On success *m and *S are replaced by the new mesh and state. */
int      fv_mesh_adapt(fv_mesh **m, const fv_params *p, const int *split, double **S);

/* ---- physics on the host (used at regrid time and for initial data) ---- */
int      fv_neq(int nmat);
void     fv_relax_host(const fv_params *p, double *vec);

/* ---- initial conditions (m4_init_domain.c / init_S.c:70-107) ---- */
/* cell state of pure material mat (init_S.c:70-107) with pressure p and velocity u */
void     fv_init_pure(const fv_params *p, const double *rho_init, int mat, double pres,
                      const double u[3], double *S);
/* two-state split at x = xs (Sod, CASE 0 of m4_init_domain.c:22-28) */
int      fv_init_split_x(const fv_mesh *m, const fv_params *p, const double *rho_init,
                         const double *p_init, double xs, double *S);
/* spherical bubble of material 1 in material 0, volume fraction smeared with
 * tanh over `smear` cells of the finest level (new case) */
int      fv_init_bubble(const fv_mesh *m, const fv_params *p, const double *rho_init,
                        const double *p_init, const double centre[3], double radius,
                        double smear, unsigned seed, double perturb, double *S);

/* ---- restart and field files (restart_save.c:12-45, restart_load.c:14-45,
 * export_ultra_fast in exportfield.c:45-97) ----
 * Text (binary = 0): header "istep time tleaves ntrees NEQ", then per leaf in
 * DFS order "root tag S[0..NEQ-1]" at %15.10e, as the reference writes and
 * reads it.  Binary (binary = 1, new): the same content without rounding,
 * magic "FVDUMP01".  S is [n][NEQ] in leaf order. */
int      fv_restart_write(const char *path, const fv_mesh *m, const double *S, int neq, int istep,
                          double time, int binary);
/* Either format (told apart by the magic): a new forest over the given brick of
 * roots, in DFS order, and its state; *S is malloc'd, the caller frees it. */
int      fv_restart_read(const char *path, const int nroot[3], const double origin[3], double rootsize,
                         const int bcside[6], fv_mesh **m, double **S, int *neq, int *istep, double *time);

#ifdef __cplusplus
}
#endif
#endif
