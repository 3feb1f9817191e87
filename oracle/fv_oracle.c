/*
 * fv_oracle.c -- CPU restatement of ForestFV's model-4 explicit update.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP path is tested against (see fv_oracle.h).
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 *
 * Every function names the reference lines it restates.  pow(x,2.0) in the
 * reference is written x*x here: IEEE multiplication is correctly rounded
 * and glibc's pow is within 0.52 ulp, so the two agree except on squares
 * that fall exactly on a rounding midpoint.
 */
#include "fv_oracle.h"
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

#define SQ(x) ((x) * (x))
#define MAXF(a, b) (((a) > (b)) ? (a) : (b))   /* strdata.h:27 */
#define MINF(a, b) (((a) < (b)) ? (a) : (b))   /* strdata.h:31 */

/* ------------------------------------------------------------------ */
/* hex tables (operatetp.c)                                            */
/* ------------------------------------------------------------------ */
/* fcnd2elnd, typ 0 (operatetp.c:60-69): face f, face node q -> element node */
static const int FCND[6][4] = {
  {0, 1, 5, 4}, {3, 7, 5, 1}, {2, 6, 7, 3}, {2, 0, 4, 6}, {2, 3, 1, 0}, {6, 4, 5, 7}};
/* fcopp, typ 0 (operatetp.c:268-279) */
static const int FCOPP[6] = {2, 3, 0, 1, 5, 4};
/* face -> axis, sign (face order y-, x+, y+, x-, z-, z+) */
static const int FAX[6] = {1, 0, 1, 0, 2, 2};
static const int FSG[6] = {-1, 1, 1, -1, -1, 1};
/* elnd2fcnd typ 0 (operatetp.c:22-29): -1 when node not on face */
static const int ELND2FCND[6][8] = {
  {0, 1, -1, -1, 3, 2, -1, -1}, {-1, 3, -1, 0, -1, 2, -1, 1},
  {-1, -1, 0, 3, -1, -1, 1, 2}, {1, -1, 0, -1, 2, -1, 3, -1},
  {3, 2, 0, 1, -1, -1, -1, -1}, {-1, -1, -1, -1, 1, 2, 0, 3}};

int orc_neq(int nmat) { return nmat + 3 + 1 + (nmat - 1) + nmat; }   /* m4_input.c:121-131 */

typedef struct {
  int nmat, neq, ncons;
  int i0, i1, i2, i3, i4;     /* eqtypi */
  int n0, n3, n4;             /* eqtypn[0], eqtypn[3], eqtypn[4] */
} layout_t;

static layout_t mk_layout(int nmat) {   /* m4_input.c:121-159 */
  layout_t L;
  L.nmat = nmat;
  L.neq = orc_neq(nmat);
  L.ncons = nmat + 3 + 1;
  L.n0 = nmat; L.n3 = nmat - 1; L.n4 = nmat;
  L.i0 = 0; L.i1 = L.i0 + nmat; L.i2 = L.i1 + 3; L.i3 = L.i2 + 1; L.i4 = L.i3 + L.n3;
  return L;
}

/* ------------------------------------------------------------------ */
/* geometry                                                            */
/* ------------------------------------------------------------------ */
typedef struct { double x[8], y[8], z[8]; } nodes_t;

static void root_nodes(const orc_mesh *m, int r, nodes_t *nd) {
  int rx = r % m->nroot[0], ry = (r / m->nroot[0]) % m->nroot[1], rz = r / (m->nroot[0] * m->nroot[1]);
  for (int q = 0; q < 8; q++) {
    nd->x[q] = m->origin[0] + (double)(rx + (q & 1)) * m->rootsize;
    nd->y[q] = m->origin[1] + (double)(ry + ((q >> 1) & 1)) * m->rootsize;
    nd->z[q] = m->origin[2] + (double)(rz + ((q >> 2) & 1)) * m->rootsize;
  }
}

/* split.c:68-97: node coordinates of kid ikid of a hex split into 8 */
static void kid_nodes(const nodes_t *p, int ikid, nodes_t *k) {
  for (int ind = 0; ind < 8; ind++) {
    int onfc = 0, onifc = 0;
    for (int f = 0; f < 6; f++)
      if (ELND2FCND[f][ind] != -1 && ELND2FCND[f][ikid] != -1) { onfc++; onifc = f; }
    if (onfc == 3) {
      k->x[ind] = p->x[ind]; k->y[ind] = p->y[ind]; k->z[ind] = p->z[ind];
    } else if (onfc == 2) {
      k->x[ind] = 0.5 * (p->x[ind] + p->x[ikid]);
      k->y[ind] = 0.5 * (p->y[ind] + p->y[ikid]);
      k->z[ind] = 0.5 * (p->z[ind] + p->z[ikid]);
    } else if (onfc == 1) {
      const int *q = FCND[onifc];
      k->x[ind] = 0.25 * (p->x[q[0]] + p->x[q[1]] + p->x[q[2]] + p->x[q[3]]);
      k->y[ind] = 0.25 * (p->y[q[0]] + p->y[q[1]] + p->y[q[2]] + p->y[q[3]]);
      k->z[ind] = 0.25 * (p->z[q[0]] + p->z[q[1]] + p->z[q[2]] + p->z[q[3]]);
    } else {
      k->x[ind] = 0.125 * (p->x[0] + p->x[1] + p->x[2] + p->x[3] + p->x[4] + p->x[5] + p->x[6] + p->x[7]);
      k->y[ind] = 0.125 * (p->y[0] + p->y[1] + p->y[2] + p->y[3] + p->y[4] + p->y[5] + p->y[6] + p->y[7]);
      k->z[ind] = 0.125 * (p->z[0] + p->z[1] + p->z[2] + p->z[3] + p->z[4] + p->z[5] + p->z[6] + p->z[7]);
    }
  }
}

/* nodes of the octant (root r, level lvl, integer position c) by descending
 * from the root with kid index ikid = i + 2j + 4k per level (brchconn.c:23-26) */
static void octant_nodes(const orc_mesh *m, int r, int lvl, const int *c, nodes_t *out) {
  nodes_t a, b;
  root_nodes(m, r, &a);
  for (int l = 2; l <= lvl; l++) {
    int sh = lvl - l;
    int ikid = ((c[0] >> sh) & 1) + 2 * ((c[1] >> sh) & 1) + 4 * ((c[2] >> sh) & 1);
    kid_nodes(&a, ikid, &b);
    a = b;
  }
  *out = a;
}

static void centre(const nodes_t *nd, double *xc, double *yc, double *zc) {  /* distance.c:101-118 */
  *xc = 0.0; *yc = 0.0; *zc = 0.0;
  for (int q = 0; q < 8; q++) { *xc += nd->x[q]; *yc += nd->y[q]; *zc += nd->z[q]; }
  *xc /= 8.0; *yc /= 8.0; *zc /= 8.0;
}

static double face_area(const nodes_t *nd, int f) {   /* meshproperties.c:42-73 */
  double A = 0.0;
  const int *q = FCND[f];
  for (int it = 0; it < 2; it++) {
    double dx1 = nd->x[q[1 + it]] - nd->x[q[0]], dx2 = nd->x[q[2 + it]] - nd->x[q[0]];
    double dy1 = nd->y[q[1 + it]] - nd->y[q[0]], dy2 = nd->y[q[2 + it]] - nd->y[q[0]];
    double dz1 = nd->z[q[1 + it]] - nd->z[q[0]], dz2 = nd->z[q[2 + it]] - nd->z[q[0]];
    A += fabs(0.5 * sqrt(SQ(dy1 * dz2 - dz1 * dy2) + SQ(dz1 * dx2 - dx1 * dz2) + SQ(dx1 * dy2 - dy1 * dx2)));
  }
  return A;
}

static double volume(const nodes_t *nd) {   /* meshproperties.c:76-125 */
  static const int vn[6][3] = {{1, 3, 5}, {5, 3, 7}, {5, 7, 4}, {3, 2, 7}, {7, 2, 6}, {7, 6, 4}};
  double V = 0.0;
  for (int it = 0; it < 6; it++) {
    double dx[3], dy[3], dz[3];
    for (int in = 0; in < 3; in++) {
      dx[in] = nd->x[vn[it][in]] - nd->x[0];
      dy[in] = nd->y[vn[it][in]] - nd->y[0];
      dz[in] = nd->z[vn[it][in]] - nd->z[0];
    }
    double c0 = dy[0] * dz[1] - dz[0] * dy[1];
    double c1 = dz[0] * dx[1] - dx[0] * dz[1];
    double c2 = dx[0] * dy[1] - dy[0] * dx[1];
    V += (c0 * dx[2] + c1 * dy[2] + c2 * dz[2]) / 6.0;
  }
  return V;
}

static void normals(const nodes_t *nd, orc_cell *cl) {   /* normalvector.c:15-30, tagvector.c:15-39 */
  for (int f = 0; f < 6; f++) {
    const int *q = FCND[f];
    double ax = nd->x[q[0]], ay = nd->y[q[0]], az = nd->z[q[0]];
    double bx = nd->x[q[1]], by = nd->y[q[1]], bz = nd->z[q[1]];
    double cx = nd->x[q[2]], cy = nd->y[q[2]], cz = nd->z[q[2]];
    double n0 = (by - ay) * (cz - az) - (cy - ay) * (bz - az);
    double n1 = (bx - ax) * (cz - az) - (cx - ax) * (bz - az);
    double n2 = (bx - ax) * (cy - ay) - (cx - ax) * (by - ay);
    double norm = sqrt(SQ(n0) + SQ(n1) + SQ(n2));
    cl->nx[f] = n0 / norm;
    cl->ny[f] = -(n1) / norm;
    cl->nz[f] = n2 / norm;
    double tn = sqrt(SQ(by - ay) + SQ(bx - ax) + SQ(bz - az));
    cl->nxt1[f] = (bx - ax) / tn;
    cl->nyt1[f] = (by - ay) / tn;
    cl->nzt1[f] = (bz - az) / tn;
    double nn = sqrt(SQ((cl->ny[f] * cl->nzt1[f]) - (cl->nyt1[f] * cl->nz[f])) +
                     SQ((cl->nx[f] * cl->nzt1[f]) - (cl->nxt1[f] * cl->nz[f])) +
                     SQ((cl->nx[f] * cl->nyt1[f]) - (cl->nxt1[f] * cl->ny[f])));
    cl->nxt2[f] = (((cl->ny[f]) * (cl->nzt1[f])) - ((cl->nyt1[f]) * (cl->nz[f]))) / nn;
    cl->nyt2[f] = -(((cl->nx[f]) * (cl->nzt1[f])) - ((cl->nxt1[f]) * (cl->nz[f]))) / nn;
    cl->nzt2[f] = (((cl->nx[f]) * (cl->nyt1[f])) - ((cl->nxt1[f]) * (cl->ny[f]))) / nn;
  }
}

int orc_geometry(const orc_mesh *m, orc_cell *cells) {
  for (int i = 0; i < m->n; i++) {
    orc_cell *cl = &cells[i];
    nodes_t nd;
    const int *ci = &m->ci[3 * i];
    octant_nodes(m, m->root[i], m->level[i], ci, &nd);
    normals(&nd, cl);
    cl->Vol = volume(&nd);
    centre(&nd, &cl->xc, &cl->yc, &cl->zc);
    for (int f = 0; f < 6; f++) {
      int lfc = m->lfc_neigs[6 * i + f];
      int nsfc = lfc > 1 ? lfc : 1;                       /* elconnfin.c:103 */
      for (int k = 0; k < 4; k++) cl->Area[f][k] = 0.0;
      if (nsfc == 1) {
        cl->Area[f][0] = face_area(&nd, f);
      } else {
        for (int k = 0; k < nsfc; k++) {
          int j = m->neigtr[24 * i + 4 * f + k];
          nodes_t nn;
          octant_nodes(m, m->root[j], m->level[j], &m->ci[3 * j], &nn);
          cl->Area[f][k] = face_area(&nn, FCOPP[f]);
        }
      }
      /* distance.c:120-167 */
      cl->dist_cf[f] = 0.0; cl->dist_cc[f] = 0.0;
      if (lfc == 0) continue;   /* boundary: the sentinel neighbour matches no case */
      int j = m->neigtr[24 * i + 4 * f];
      int ldiff = m->level[j] - m->level[i];
      double xc, yc, zc, xn, yn, zn, d;
      if (ldiff == 0) {
        nodes_t nn;
        octant_nodes(m, m->root[j], m->level[j], &m->ci[3 * j], &nn);
        centre(&nd, &xc, &yc, &zc); centre(&nn, &xn, &yn, &zn);
        d = sqrt(SQ(xc - xn) + SQ(yc - yn) + SQ(zc - zn));
        cl->dist_cf[f] = 0.5 * d; cl->dist_cc[f] = d;
      } else if (ldiff == 1) {
        int pc[3] = {m->ci[3 * j] >> 1, m->ci[3 * j + 1] >> 1, m->ci[3 * j + 2] >> 1};
        nodes_t nn;
        octant_nodes(m, m->root[j], m->level[j] - 1, pc, &nn);
        centre(&nd, &xc, &yc, &zc); centre(&nn, &xn, &yn, &zn);
        d = sqrt(SQ(xc - xn) + SQ(yc - yn) + SQ(zc - zn));
        cl->dist_cf[f] = 0.5 * d; cl->dist_cc[f] = 0.5 * d + 0.25 * d;
      } else if (ldiff == -1) {
        int pc[3] = {ci[0] >> 1, ci[1] >> 1, ci[2] >> 1};
        nodes_t np, nn;
        octant_nodes(m, m->root[i], m->level[i] - 1, pc, &np);
        octant_nodes(m, m->root[j], m->level[j], &m->ci[3 * j], &nn);
        centre(&np, &xc, &yc, &zc); centre(&nn, &xn, &yn, &zn);
        d = sqrt(SQ(xc - xn) + SQ(yc - yn) + SQ(zc - zn));
        cl->dist_cf[f] = 0.25 * d; cl->dist_cc[f] = 0.25 * d + 0.5 * d;
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* SOL (strdata.h SOL) and state conversions                           */
/* ------------------------------------------------------------------ */
typedef struct {
  double r, ra[ORC_MAXMAT], vec[ORC_MAXNEQ], Y[ORC_MAXMAT];
  double u, v, w, e;
  double avf[ORC_MAXMAT + 1], phi[ORC_MAXMAT], are[ORC_MAXMAT];
  double st[3][3], c2Y[ORC_MAXMAT], p_hydro[ORC_MAXMAT];
} sol_t;

/* m4_facevalues.c:20-106 and m4_facercnstr.c:20-106 share this body; the
 * first reads the cell state S, the second a face state SF, and both take the
 * cell-centre volume fraction from S for phi. */
static void fill_sol(const layout_t *L, const orc_params *p, const double *X, const double *Sc, sol_t *s) {
  double amin_100 = p->amin / 100.0;
  double at[ORC_MAXMAT + 1], ac[ORC_MAXMAT + 1], avfsum, press;
  s->r = 0.0;
  for (int v = 0; v < L->n0; ++v) { s->r += X[v]; s->ra[v] = X[v]; s->vec[v] = X[v]; }
  for (int v = 0; v < L->n0; ++v) s->Y[v] = s->ra[v] / s->r;
  int v1 = L->i1;
  s->u = X[v1 + 0] / s->r; s->v = X[v1 + 1] / s->r; s->w = X[v1 + 2] / s->r;
  s->vec[v1 + 0] = X[v1 + 0]; s->vec[v1 + 1] = X[v1 + 1]; s->vec[v1 + 2] = X[v1 + 2];
  s->e = X[L->i2] / s->r;
  s->vec[L->i2] = X[L->i2];
  for (int iv = 0; iv < L->n3; ++iv) { at[iv] = X[iv + L->i3]; ac[iv] = Sc[iv + L->i3]; }
  avfsum = 0.0;
  for (int v = 0; v < L->n3; ++v) {
    if (at[v] <= amin_100) at[v] = amin_100;
    else if (at[v] >= (1.0 - (amin_100))) at[v] = 1.0 - (amin_100);
    if (ac[v] <= amin_100) ac[v] = amin_100;
    else if (ac[v] >= (1.0 - (amin_100))) ac[v] = 1.0 - (amin_100);
    avfsum += at[v];
  }
  if (avfsum > (1.0 - (amin_100))) {
    double avf_dif = avfsum - (1.0 - (amin_100));
    for (int j = 0; j < L->n3; ++j) { double per = at[j] / avfsum; at[j] = at[j] - per * avf_dif; }
    avfsum = 1.0 - (amin_100);
  }
  at[L->n3] = 1.0 - avfsum;
  avfsum = 0.0;
  for (int v = 0; v < L->n3; ++v) { s->avf[v] = at[v]; avfsum += at[v]; s->vec[L->i3 + v] = at[v]; }
  s->avf[L->n3] = 1.0 - avfsum;
  for (int v = 0; v < L->n3; ++v) s->phi[v] = ac[v] - p->avf_lim;
  for (int iv = 0; iv < L->n4; ++iv) { s->are[iv] = X[iv + L->i4]; s->vec[L->i4 + iv] = X[iv + L->i4]; }
  for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) s->st[a][b] = 0.0;
  for (int v = 0; v < L->n4; ++v) {
    press = ((p->gamma[v] - 1.0) * (s->are[v]) / at[v]) - p->gamma[v] * p->pinf[v];
    s->c2Y[v] = (at[v] / s->r) * p->gamma[v] * (press + p->pinf[v]);
    s->st[0][0] += -at[v] * press;
    s->st[1][1] += -at[v] * press;
    s->st[2][2] += -at[v] * press;
    s->p_hydro[v] = press;
  }
}

/* face frame of one cell face */
typedef struct { double nx, ny, nz, mx, my, mz, lx, ly, lz; } frame_t;

static frame_t mkframe(const orc_cell *c, int f) {
  frame_t F;
  F.nx = c->nx[f]; F.ny = c->ny[f]; F.nz = c->nz[f];
  F.mx = c->nxt1[f]; F.my = c->nyt1[f]; F.mz = c->nzt1[f];
  F.lx = c->nxt2[f]; F.ly = c->nyt2[f]; F.lz = c->nzt2[f];
  return F;
}

/* m4_bc_inlet / m4_bc_outlet with CASE 3400 or 3410 (m4_bc.c:44-113, 121-186):
 * every material of the ghost at pressure pb, with the ghost's own fractions,
 * densities and velocity; are, c2Y, e and rho E follow.  The reference writes
 * materials 0 and 1 (its injector cases have two); this covers all of them. */
static void fixed_pressure_ghost(const layout_t *L, const orc_params *p, double pb, sol_t *s) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) s->st[a][b] = (a == b) ? -pb : 0.0;
  double re = 0.0;
  for (int k = 0; k < L->n4; k++) {
    s->p_hydro[k] = pb;
    s->are[k] = s->avf[k] * (pb + p->gamma[k] * p->pinf[k]) / (p->gamma[k] - 1.0);
    s->vec[L->i4 + k] = s->are[k];
    s->c2Y[k] = (s->avf[k] / s->r) * p->gamma[k] * (pb + p->pinf[k]);
    re += s->are[k];
  }
  re = re + 0.5 * s->r * (SQ(s->u) + SQ(s->v) + SQ(s->w));
  s->e = re / s->r;
  s->vec[L->i2] = re;
}

/* boundary ghost velocity (bc_symmetry.c, bc_reflective.c:12-28, bc_wall.c:12-26,
 * m4_bc.c:3-34).  Only sol_R->u,v,w change; vec keeps the unmirrored momentum. */
static void apply_bc(const layout_t *L, const orc_params *p, int code, const frame_t *F, const sol_t *sL, sol_t *sR) {
  if (code == 1 || code == 4 || code == 5) {
    double u = sL->u, v = sL->v, w = sL->w;
    double ul = u * F->nx + v * F->ny + w * F->nz;
    double vl = u * F->mx + v * F->my + w * F->mz;
    double wl = u * F->lx + v * F->ly + w * F->lz;
    double ur = -ul, vr = vl, wr = wl;
    if (code == 5) { vr = -vl; wr = -wl; }
    sR->u = ur * F->nx + vr * F->mx + wr * F->lx;
    sR->v = ur * F->ny + vr * F->my + wr * F->ly;
    sR->w = ur * F->nz + vr * F->mz + wr * F->lz;
  }
  /* codes 2 and 3 act only for the reference's CASE 3400/3410: a plain copy otherwise */
  if (p->bc_fixed_p && (code == 2 || code == 3)) fixed_pressure_ghost(L, p, code == 3 ? p->p_inlet : p->p_outlet, sR);
}

/* rotate_vectors.c:128-188 */
static void rotate_tensor(double t[9], double st[3][3], const frame_t *F) {
  double vec[3][3] = {{F->nx, F->ny, F->nz}, {F->mx, F->my, F->mz}, {F->lx, F->ly, F->lz}};
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
  double flux_adv[ORC_MAXNEQ][3];
  double flux_hllc[ORC_MAXNEQ];
  double Us[ORC_MAXNEQ];
  double u_nc, v_nc, w_nc;
  double avf_S[ORC_MAXMAT], are_S[ORC_MAXMAT];
} hllc_ws;

/* m4_flux.c:3-73 */
static void m4_flux(const layout_t *L, const sol_t *s, int flag_ifs, hllc_ws *W) {
  for (int v = 0; v < L->n0; v++) {
    int iv = v + L->i0;
    W->flux_adv[iv][0] = s->u * s->ra[v];
    W->flux_adv[iv][1] = s->v * s->ra[v];
    W->flux_adv[iv][2] = s->w * s->ra[v];
  }
  int iv = L->i1;
  W->flux_adv[iv][0] = s->u * s->u * s->r - s->st[0][0];
  W->flux_adv[iv][1] = s->v * s->u * s->r - s->st[0][1];
  W->flux_adv[iv][2] = s->w * s->u * s->r - s->st[0][2];
  iv = L->i1 + 1;
  W->flux_adv[iv][0] = s->u * s->v * s->r - s->st[1][0];
  W->flux_adv[iv][1] = s->v * s->v * s->r - s->st[1][1];
  W->flux_adv[iv][2] = s->w * s->v * s->r - s->st[1][2];
  iv = L->i1 + 2;
  W->flux_adv[iv][0] = s->u * s->w * s->r - s->st[2][0];
  W->flux_adv[iv][1] = s->v * s->w * s->r - s->st[2][1];
  W->flux_adv[iv][2] = s->w * s->w * s->r - s->st[2][2];
  iv = L->i2;
  W->flux_adv[iv][0] = s->r * s->e * s->u - s->st[0][0] * s->u - s->st[0][1] * s->v - s->st[0][2] * s->w;
  W->flux_adv[iv][1] = s->r * s->e * s->v - s->st[1][0] * s->u - s->st[1][1] * s->v - s->st[1][2] * s->w;
  W->flux_adv[iv][2] = s->r * s->e * s->w - s->st[2][0] * s->u - s->st[2][1] * s->v - s->st[2][2] * s->w;
  if (flag_ifs == 0 || flag_ifs == 3) {
    W->u_nc = s->u; W->v_nc = s->v; W->w_nc = s->w;
    for (int v = 0; v < L->n3; ++v) W->avf_S[v] = s->avf[v];
    for (int v = 0; v < L->n4; ++v) W->are_S[v] = s->are[v];
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
}

/* m4_star_region.c:3-87 */
static void m4_star_region(const layout_t *L, const orc_params *p, const frame_t *F,
                           const sol_t *sL, const sol_t *sR, const sol_t *s, int flag_ifs,
                           double u_S, double SL, double SR, double uL, double vL, double wL,
                           double uR, double vR, double wR, double s11L, double s12L, double s13L,
                           double s11R, double s12R, double s13R, hllc_ws *W) {
  double SK, uK, vK, wK, s11K, s12K, s13K;
  if (flag_ifs == 1) { SK = SL; uK = uL; vK = vL; wK = wL; s11K = s11L; s12K = s12L; s13K = s13L; }
  else { SK = SR; uK = uR; vK = vR; wK = wR; s11K = s11R; s12K = s12R; s13K = s13R; }
  double ar_S[ORC_MAXMAT], r_S = 0.0;
  for (int v = 0; v < L->n0; ++v) { ar_S[v] = s->ra[v] * (SK - uK) / (SK - u_S); r_S += ar_S[v]; }
  double s11_S = ((uR - SR) * sR->r * (s11L) - (uL - SL) * sL->r * (s11R) + (uL - SL) * sL->r * (uR - SR) * sR->r * (uR - uL)) /
                 ((uR - SR) * sR->r - (uL - SL) * sL->r);
  double s12_S = 0.0, s13_S = 0.0;
  double v_S = vK + (s12_S - (s12K)) / ((uK - SK) * s->r);
  double w_S = wK + (s13_S - (s13K)) / ((uK - SK) * s->r);
  double vel[3];
  vel[0] = u_S * F->nx + v_S * F->mx + w_S * F->lx;     /* rotate_vector_ltg_v2 */
  vel[1] = u_S * F->ny + v_S * F->my + w_S * F->ly;
  vel[2] = u_S * F->nz + v_S * F->mz + w_S * F->lz;
  double re = s->r * s->e;
  double Em_S;
  if (p->fix & ORC_FIX_ENERGY)
    Em_S = (re * (uK - SK) - s11K * uK - s12K * vK - s13K * wK + s11_S * u_S + s12_S * v_S + s13_S * w_S) / (r_S * (u_S - SK));
  else
    Em_S = (re * (uK - SK) + s11K * uK - s12K * vK - s13K * wK + s11_S * u_S + s12_S * v_S + s13_S * w_S) / (r_S * (u_S - SK));
  double e_S[ORC_MAXMAT];
  for (int v = 0; v < L->n4; ++v) {
    double rm = (s->ra[v] / s->avf[v]);
    double rm_S = (s->ra[v] / s->avf[v]) * (uK - SK) / (u_S - SK);
    double p_k = s->p_hydro[v];
    double g = p->gamma[v], pi = p->pinf[v];
    double p_S = (p_k + pi) * (((g - 1.0) * rm + (g + 1.0) * rm_S) / ((g - 1.0) * rm_S + (g + 1.0) * rm)) - pi;
    e_S[v] = (p_S + g * pi) / (rm_S * (g - 1.0));
  }
  for (int v = 0; v < L->n0; ++v) W->Us[L->i0 + v] = ar_S[v];
  for (int v = 0; v < 3; ++v) W->Us[L->i1 + v] = r_S * vel[v];
  W->Us[L->i2] = r_S * Em_S;
  W->u_nc = vel[0]; W->v_nc = vel[1]; W->w_nc = vel[2];
  for (int v = 0; v < L->n3; ++v) W->avf_S[v] = s->avf[v];
  for (int v = 0; v < L->n4; ++v) W->are_S[v] = ar_S[v] * e_S[v];
}

/* m4_hllc.c:3-61 with m4_wave_speeds.c:3-76 */
static void m4_hllc(const layout_t *L, const orc_params *p, const frame_t *F, int bc,
                    const sol_t *sL, const sol_t *sR, hllc_ws *W) {
  double u, v, w, uL, vL, wL, uR, vR, wR;
  u = sL->u; v = sL->v; w = sL->w;
  uL = u * F->nx + v * F->ny + w * F->nz;
  vL = u * F->mx + v * F->my + w * F->mz;
  wL = u * F->lx + v * F->ly + w * F->lz;
  u = sR->u; v = sR->v; w = sR->w;
  uR = u * F->nx + v * F->ny + w * F->nz;
  vR = u * F->mx + v * F->my + w * F->mz;
  wR = u * F->lx + v * F->ly + w * F->lz;
  double ruL = sL->r * (uL), ruR = sR->r * (uR);
  double c2YL = 0.0, c2YR = 0.0;
  for (int iv = 0; iv < L->nmat; ++iv) { c2YL += sL->c2Y[iv]; c2YR += sR->c2Y[iv]; }
  double cL = sqrt(c2YL), cR = sqrt(c2YR);
  double SR = MAXF((uL + cL), (uR + cR));
  double SL = MINF((uL - cL), (uR - cR));
  double t[9];
  double stL[3][3], stR[3][3];
  memcpy(stL, sL->st, sizeof stL); memcpy(stR, sR->st, sizeof stR);
  rotate_tensor(t, stL, F);
  double s11L = t[0], s12L = t[1], s13L = t[2];
  rotate_tensor(t, stR, F);
  double s11R = t[0], s12R = t[1], s13R = t[2];
  double u_S = ((sL->r * SQ((uL)) - (s11L)) - (sR->r * SQ((uR)) - (s11R)) - ((SL)*ruL) + ((SR)*ruR)) /
               (ruL - ruR - ((SL)*sL->r) + ((SR)*sR->r));
  if (bc == 4) u_S = 0.0;
  for (int iv = 0; iv < L->neq; ++iv) W->flux_hllc[iv] = 0.0;
  int end = 0;
  if ((0.0 <= SL) && !end) { end = 1; m4_flux(L, sL, 0, W); }
  if (((SL <= 0.0) && (0.0 <= u_S)) && !end) {
    end = 1;
    m4_star_region(L, p, F, sL, sR, sL, 1, u_S, SL, SR, uL, vL, wL, uR, vR, wR, s11L, s12L, s13L, s11R, s12R, s13R, W);
    m4_flux(L, sL, 1, W);
    for (int iv = 0; iv < L->ncons; ++iv) W->flux_hllc[iv] = SL * (W->Us[iv] - sL->vec[iv]);
  }
  if (((u_S <= 0.0) && (0.0 <= SR)) && !end) {
    end = 1;
    m4_star_region(L, p, F, sL, sR, sR, 2, u_S, SL, SR, uL, vL, wL, uR, vR, wR, s11L, s12L, s13L, s11R, s12R, s13R, W);
    m4_flux(L, sR, 2, W);
    for (int iv = 0; iv < L->ncons; ++iv) W->flux_hllc[iv] = SR * (W->Us[iv] - sR->vec[iv]);
  }
  if ((SR <= 0.0) && !end) { end = 1; m4_flux(L, sR, 3, W); }
}

/* ------------------------------------------------------------------ */
/* relaxation (m4_relaxationvec.c)                                      */
/* ------------------------------------------------------------------ */
static void vf_limit_relax(const layout_t *L, const orc_params *p, double *avf) {
  double a = p->amin / 100.0, avfsum = 0.0;
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

/* F_pe.c:18-27 and :64-73 (i_paper = 1, Favrie): sum_k ar_k tau_k(pe) - 1 and its derivative */
static double fpe(int nm, const double *G, const double *P, const double *ar, const double *t0, const double *p0,
                  double pe) {
  double F = 0.0;
  for (int iv = 0; iv < nm; ++iv)
    F += ar[iv] * t0[iv] * ((p0[iv] + G[iv] * P[iv] + (G[iv] - 1.0) * pe) / (pe + G[iv] * P[iv] + (G[iv] - 1.0) * pe));
  return F + -1.0;
}
static double dfpe(int nm, const double *G, const double *P, const double *ar, const double *t0, const double *p0,
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

static void relax_vec(const layout_t *L, const orc_params *p, double *vec) {
  double ar[ORC_MAXMAT], are[ORC_MAXMAT], avf[ORC_MAXMAT + 1];
  double rho = 0.0;
  const double *G = p->gamma, *P = p->pinf;
  for (int iv = 0; iv < L->n0; ++iv) { ar[iv] = vec[iv + L->i0]; are[iv] = vec[iv + L->i4]; rho += ar[iv]; }
  int iv1 = L->i1;
  double u = vec[iv1 + 0] / rho, v = vec[iv1 + 1] / rho, w = vec[iv1 + 2] / rho;
  double rE = vec[L->i2];
  for (int iv = 0; iv < L->n3; ++iv) avf[iv] = vec[iv + L->i3];
  vf_limit_relax(L, p, avf);
  if (p->prelax == 1) {   /* m4_relaxationvec.c:63-127 */
    double t0[ORC_MAXMAT], p0[ORC_MAXMAT];
    for (int iv = 0; iv < L->n0; ++iv) {
      t0[iv] = 1.0 / (ar[iv] / avf[iv]);
      p0[iv] = (are[iv] * (G[iv] - 1.0) - avf[iv] * G[iv] * P[iv]) / avf[iv];
    }
    double A1 = (avf[0] / G[0]) * (p0[0] + P[0]) / ((avf[0] / G[0]) + (avf[1] / G[1]));
    double A2 = (avf[1] / G[1]) * (p0[1] + P[1]) / ((avf[0] / G[0]) + (avf[1] / G[1]));
    double pe = 0.5 * (A1 + A2 - (P[0] + P[1])) + sqrt(0.25 * (SQ((A2 - A1 - (P[1] - P[0])))) + A1 * A2);
    for (int iv = 0; iv < L->n0; ++iv) {
      double t1 = t0[iv] * (p0[iv] + G[iv] * P[iv] + (G[iv] - 1.0) * pe) / (pe + G[iv] * P[iv] + (G[iv] - 1.0) * pe);
      double rm = 1.0 / t1;
      avf[iv] = ar[iv] / rm;
    }
    vf_limit_relax(L, p, avf);
    double e_hyd = rE / rho - 0.5 * (SQ((u)) + SQ((v)) + SQ((w)));
    double num = 0.0, den = 0.0;
    for (int iv = 0; iv < L->n0; ++iv) {
      num += avf[iv] * G[iv] * P[iv] / (G[iv] - 1.0);
      den += avf[iv] / (G[iv] - 1.0);
    }
    double pr = (rho * e_hyd - num) / den;
    for (int iv = 0; iv < L->n0; ++iv) are[iv] = (avf[iv] * (pr + G[iv] * P[iv]) / (G[iv] - 1.0));
  } else if (p->prelax == 2) {   /* m4_relaxationvec.c:128-297, F_pe.c:3-123 (i_paper = 1) */
    /* The published branch fills local t0_m / p0_m / ar_R while F_pe and DF_pe
     * read the RUN fields of the same names, which are allocated
     * (createutility.c:121-127) and never written; the Newton iteration below
     * uses the local values, which is what the branch computes them for. */
    const double eps_big = 1.0e-8;
    const int nm = L->nmat;
    double t0[ORC_MAXMAT], p0[ORC_MAXMAT];
    double max_avf = -1.0;
    int principal = 0;
    for (int iv = 0; iv < nm; ++iv) {
      t0[iv] = 1.0 / (ar[iv] / avf[iv]);
      p0[iv] = (are[iv] * (G[iv] - 1.0) - avf[iv] * G[iv] * P[iv]) / avf[iv];
      if (p0[iv] < -(1.0 - eps_big) * P[iv] + eps_big) p0[iv] = -(1.0 - eps_big) * P[iv] + 1.0;
      if (avf[iv] > max_avf) { max_avf = avf[iv]; principal = iv; }
    }
    double pe_old = 0.001, pe_new = pe_old;
    int iter = 0;
    double cond_1 = fabs(pe_new - pe_old);
    double cond_2 = fabs(fpe(nm, G, P, ar, t0, p0, pe_new));
    while (((cond_2 >= 1.0e-12) || (cond_1 >= 1.0e-4)) && (iter <= 101)) {
      pe_new = pe_old - 0.8 * (fpe(nm, G, P, ar, t0, p0, pe_old) / dfpe(nm, G, P, ar, t0, p0, pe_old));
      cond_1 = fabs(pe_new - pe_old);
      cond_2 = fabs(fpe(nm, G, P, ar, t0, p0, pe_new));
      if (iter >= 100 && ((fabs(pe_old - pe_new) < 1.0e3) || isnan(pe_new))) { pe_new = NAN; break; }
      pe_old = pe_new;
      iter++;
    }
    if (pe_new < (-(1.0 - eps_big) * P[principal] + eps_big)) pe_new = -(1.0 - eps_big) * P[principal] + 1.0;
    const double pe = pe_new;
    for (int iv = 0; iv < nm; ++iv) {
      double t1 = t0[iv] * (p0[iv] + G[iv] * P[iv] + (G[iv] - 1.0) * pe) / (pe + G[iv] * P[iv] + (G[iv] - 1.0) * pe);
      double rm = 1.0 / t1;
      avf[iv] = ar[iv] / rm;
    }
    vf_limit_relax(L, p, avf);
    double e_hyd = rE / rho - 0.5 * (SQ((u)) + SQ((v)) + SQ((w)));
    double num = 0.0, den = 0.0;
    for (int iv = 0; iv < nm; ++iv) {
      num += avf[iv] * G[iv] * P[iv] / (G[iv] - 1.0);
      den += avf[iv] / (G[iv] - 1.0);
    }
    double pr = (rho * e_hyd - num) / den;
    for (int iv = 0; iv < nm; ++iv) are[iv] = (avf[iv] * (pr + G[iv] * P[iv]) / (G[iv] - 1.0));
  } else if (p->prelax == 3) {   /* m4_relaxationvec.c:298-479 */
    const double eps_sml = 1.0e-12, eps_big = 1.0e-8;
    double rhoKs[ORC_MAXMAT], pk[ORC_MAXMAT];
    int nm = L->nmat;
    for (int iv = 0; iv < nm; ++iv) {
      if (avf[iv] < 0.0 || ar[iv] < 0.0) { avf[iv] = eps_sml; ar[iv] = eps_sml; are[iv] = eps_sml; }
      if (avf[iv] >= 1.0) avf[iv] = 1.0 - eps_sml;
    }
    int irelax = 1;
    for (int iv = 0; iv < nm; ++iv) if (avf[iv] > 1.0 - eps_sml) irelax = 0;
    if (irelax) {
      double pr = 0.0;
      for (int iv = 0; iv < nm; ++iv) {
        if (avf[iv] > eps_sml) {
          pk[iv] = (are[iv]) / avf[iv] * (G[iv] - 1.0) - G[iv] * P[iv];
          if (pk[iv] <= -(1.0 - eps_big) * P[iv] + eps_big) pk[iv] = -(1.0 - eps_big) * P[iv] + eps_big;
        } else pk[iv] = 0.0;
        pr += avf[iv] * pk[iv];
      }
      int iter = 0;
      double F = 1.0e-9, dF = 1.0e9;
      for (int iv = 0; iv < nm; ++iv) rhoKs[iv] = 0.0;
      while ((fabs(F) > 1.0e-12) && (iter < 500)) {
        pr = pr - F / dF;
        iter = iter + 1;
        for (int iv = 0; iv < nm; ++iv)
          if (pr < -(1.0 - eps_big) * P[iv] + eps_big) pr = -(1.0 - eps_big) * P[iv] + 1.0;
        F = -1.0; dF = 0.0;
        for (int iv = 0; iv < nm; ++iv) {
          if (avf[iv] > eps_sml) {
            rhoKs[iv] = ar[iv] / MAXF(avf[iv], eps_sml) * pow((pr + P[iv]) / (pk[iv] + P[iv]), 1.0 / G[iv]);
            F = F + ar[iv] / rhoKs[iv];
            dF = dF - ar[iv] / (G[iv] * rhoKs[iv] * (pr + P[iv]));
          }
        }
      }
      for (int iv = 0; iv < nm; ++iv) if (avf[iv] > eps_sml) avf[iv] = ar[iv] / rhoKs[iv];
      vf_limit_relax(L, p, avf);
    }
    double e_hyd = rE / rho - 0.5 * (SQ((u)) + SQ((v)) + SQ((w)));
    double num = 0.0, den = 0.0;
    for (int iv = 0; iv < nm; ++iv) {
      num += avf[iv] * G[iv] * P[iv] / (G[iv] - 1.0);
      den += avf[iv] / (G[iv] - 1.0);
    }
    double pr = (rho * e_hyd - num) / den;
    for (int iv = 0; iv < nm; ++iv) are[iv] = (avf[iv] * (pr + G[iv] * P[iv]) / (G[iv] - 1.0));
  }
  for (int iv = 0; iv < L->n3; ++iv) vec[iv + L->i3] = avf[iv];
  for (int iv = 0; iv < L->n4; ++iv) vec[iv + L->i4] = are[iv];
}

void orc_relax(const orc_params *p, double *vec) {
  layout_t L = mk_layout(p->nmat);
  relax_vec(&L, p, vec);
}

/* ------------------------------------------------------------------ */
/* face interpolation (m4_faceinterp.c:59-88)                           */
/* ------------------------------------------------------------------ */
/* avf_limit.c:3-21 on a copy of the volume fractions */
static void avf_limit(const layout_t *L, const orc_params *p, double *X) {
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

/* m4_facevalues_MUSCL.c:3-44, iside=1 */
static void muscl_neighbour(const layout_t *L, const orc_params *p, const orc_mesh *m,
                            const double *S, int i, int f, double *X) {
  int bc = m->bc[6 * i + f];
  if (bc == 0) {
    int lfc = m->lfc_neigs[6 * i + f];
    if (lfc == 1) {
      int j = m->neigtr[24 * i + 4 * f];
      for (int iv = 0; iv < L->neq; iv++) X[iv] = S[(size_t)j * L->neq + iv];
    } else {
      for (int iv = 0; iv < L->neq; iv++) X[iv] = 0.0;
      for (int k = 0; k < lfc; k++) {
        int j = m->neigtr[24 * i + 4 * f + k];
        for (int iv = 0; iv < L->neq; iv++) X[iv] += S[(size_t)j * L->neq + iv] / ((double)lfc);
      }
    }
  } else {
    /* m4_facevalues(imode=1, bc): the ghost vector is the own state with the
     * clipped volume fraction (m4_facevalues.c:20-88); m4_bc only moves sol->u,v,w */
    sol_t s;
    const double *Si = &S[(size_t)i * L->neq];
    fill_sol(L, p, Si, Si, &s);
    for (int iv = 0; iv < L->neq; iv++) X[iv] = s.vec[iv];
  }
  avf_limit(L, p, X);
}

/* m4_cons2primtv.c:3-40 */
static void cons2prim(const layout_t *L, const orc_params *p, const double *cons, double *prim) {
  double avf[ORC_MAXMAT + 1], avfsum = 0.0, rho = 0.0;
  for (int i = 0; i < L->n3; ++i) { prim[i] = cons[L->i3 + i]; avf[i] = cons[L->i3 + i]; avfsum += avf[i]; }
  avf[L->n3] = 1.0 - avfsum;
  for (int i = 0; i < L->n0; ++i) { prim[L->n3 + i] = cons[i]; rho += cons[i]; }
  int ip = L->n3 + L->n0;
  prim[ip + 0] = cons[L->i1 + 0] / rho;
  prim[ip + 1] = cons[L->i1 + 1] / rho;
  prim[ip + 2] = cons[L->i1 + 2] / rho;
  ip = L->n3 + L->n0 + 3;
  for (int i = 0; i < L->n4; ++i)
    prim[ip + i] = ((p->gamma[i] - 1.0) * (cons[L->i4 + i]) / avf[i]) - p->gamma[i] * p->pinf[i];
}

/* m4_prim2conc.c:3-74 */
static void prim2cons(const layout_t *L, const orc_params *p, const double *prim, double *cons) {
  double avf[ORC_MAXMAT + 1], avfsum = 0.0, rho = 0.0, sumare = 0.0, a10 = p->amin / 10.0, a100 = p->amin / 100.0;
  for (int i = 0; i < L->n3; ++i) {
    avf[i] = prim[i];
    if (avf[i] <= a10) avf[i] = a10;
    if (avf[i] >= (1.0 - (2.0 * a10))) avf[i] = 1.0 - (2.0 * a10);
    avfsum += avf[i];
  }
  if (avfsum >= (1.0 - a100)) {
    double dif = avfsum - (1.0 - a100);
    for (int j = 0; j < L->n3; ++j) { double per = avf[j] / avfsum; avf[j] = avf[j] - per * dif; }
    avfsum = 1.0 - a100;
  }
  avf[L->n3] = 1.0 - avfsum;
  for (int i = 0; i < L->n0; ++i) { cons[i] = prim[L->n3 + i]; rho += prim[L->n3 + i]; }
  int ip = L->n3 + L->n0;
  for (int i = 0; i < 3; ++i) cons[L->i1 + i] = rho * prim[ip + i];
  for (int i = 0; i < L->n3; ++i) cons[L->i3 + i] = avf[i];
  ip = L->n3 + L->n0 + 3;
  for (int i = 0; i < L->n4; ++i) {
    cons[L->i4 + i] = (avf[i] * (prim[ip + i] + p->gamma[i] * p->pinf[i]) / (p->gamma[i] - 1.0));
    sumare += cons[L->i4 + i];
  }
  ip = L->n3 + L->n0;
  cons[L->i2] = sumare + 0.5 * rho * (SQ(prim[ip + 0]) + SQ(prim[ip + 1]) + SQ(prim[ip + 2]));
}

void orc_cons2prim(const orc_params *p, const double *cons, double *prim) {
  layout_t L = mk_layout(p->nmat);
  cons2prim(&L, p, cons, prim);
}
void orc_prim2cons(const orc_params *p, const double *prim, double *cons) {
  layout_t L = mk_layout(p->nmat);
  prim2cons(&L, p, prim, cons);
}

/* minmod / MC slope of one variable (m4_reconctruct.c:105-128) */
static double muscl_slope(int limiter, double df, double db, double dp, double dm) {
  double g = 0.0;
  if (limiter == 0) {
    double product = (-db / dm) * (df / dp);
    if (product > 0.0) {
      if (fabs(db / dm) > fabs(df / dp)) g = 1.0 * df / dp;
      else g = -1.0 * db / dm;
    } else g = 0.0;
  } else {
    /* monotonised central: minmod(2a, 2b, (a+b)/2) with a backward, b forward slope */
    double a = -db / dm, b = df / dp;
    if (a * b > 0.0) {
      double c = 0.5 * (a + b);
      double m2 = MINF(2.0 * fabs(a), 2.0 * fabs(b));
      double mag = MINF(m2, fabs(c));
      g = (c > 0.0) ? mag : -mag;
    } else g = 0.0;
  }
  return g;
}

static const unsigned char *g_cell_limiter = NULL;
void orc_set_cell_limiter(const unsigned char *lim) { g_cell_limiter = lim; }

static void faceinterp_leaf(const layout_t *L, const orc_params *p, const orc_mesh *m,
                            const orc_cell *cl, const double *S, int i, double *SFi) {
  const int limiter = (p->limiter == 2 && g_cell_limiter) ? (int)g_cell_limiter[i] : p->limiter;
  double SLc[ORC_MAXNEQ], Xf[ORC_MAXNEQ], Xb[ORC_MAXNEQ], df[ORC_MAXNEQ], db[ORC_MAXNEQ];
  for (int iv = 0; iv < L->neq; iv++) SLc[iv] = S[(size_t)i * L->neq + iv];
  avf_limit(L, p, SLc);
  for (int f = 0; f < 6; ++f) {
    int fo = FCOPP[f];
    /* m4_deltavalues.c:23-57 */
    muscl_neighbour(L, p, m, S, i, f, Xf);
    for (int iv = 0; iv < L->neq; ++iv) df[iv] = Xf[iv] - SLc[iv];
    muscl_neighbour(L, p, m, S, i, fo, Xb);
    for (int iv = 0; iv < L->neq; ++iv) db[iv] = Xb[iv] - SLc[iv];
    /* m4_reconctruct.c:93-183 */
    double *SF = &SFi[f * L->neq];
    int bcf = m->bc[6 * i + f];
    const double a = p->amin / 100.0;
    if (p->primtv) {
      /* deltas, slopes and the face value in primitive variables; the volume
       * fraction is clipped, m4_prim2conc returns to conservative variables */
      const int np = 3 * L->nmat + 2;
      double pc[ORC_MAXNEQ], pf[ORC_MAXNEQ], pb[ORC_MAXNEQ], pr[ORC_MAXNEQ];
      cons2prim(L, p, SLc, pc);
      cons2prim(L, p, Xf, pf);
      cons2prim(L, p, Xb, pb);
      for (int iv = 0; iv < np; ++iv) {
        double dfp = pf[iv] - pc[iv], dbp = pb[iv] - pc[iv], g = 0.0;
        if (bcf == 0 && fabs(dfp) > 1.0e-10) g = muscl_slope(limiter, dfp, dbp, cl->dist_cc[f], cl->dist_cc[fo]);
        pr[iv] = pc[iv] + cl->dist_cf[f] * g;
      }
      double avf[ORC_MAXMAT + 1], avfsum = 0.0;
      for (int iv = 0; iv < L->n3; ++iv) {
        avf[iv] = pr[iv];
        if (avf[iv] <= a) avf[iv] = a;
        else if (avf[iv] >= (1.0 - (2.0 * a))) avf[iv] = 1.0 - (2.0 * a);
        avfsum += avf[iv];
      }
      if (avfsum >= (1.0 - a)) {
        double dif = avfsum - (1.0 - a);
        for (int j = 0; j < L->n3; ++j) { double per = avf[j] / avfsum; avf[j] = avf[j] - per * dif; }
        avfsum = 1.0 - a;
      }
      for (int iv = 0; iv < L->n3; ++iv) pr[iv] = avf[iv];
      prim2cons(L, p, pr, SF);
      relax_vec(L, p, SF);
      continue;
    }
    for (int iv = 0; iv < L->neq; ++iv) {
      double g = 0.0;
      if (bcf == 0 && fabs(df[iv]) > 1.0e-10) g = muscl_slope(limiter, df[iv], db[iv], cl->dist_cc[f], cl->dist_cc[fo]);
      SF[iv] = SLc[iv] + cl->dist_cf[f] * g;
    }
    /* VF limit (m4_reconctruct.c:147-171) */
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
      avfsum = 1.0 - a;
    }
    for (int iv = 0; iv < L->n3; ++iv) SF[iv + L->i3] = avf[iv];
    relax_vec(L, p, SF);
  }
}

int orc_faceinterp(const orc_mesh *m, const orc_cell *cells, const orc_params *p,
                   const double *S, double *SF) {
  layout_t L = mk_layout(p->nmat);
  for (int i = 0; i < m->n; i++) faceinterp_leaf(&L, p, m, &cells[i], S, i, &SF[(size_t)i * 6 * L.neq]);
  return 0;
}

/* ------------------------------------------------------------------ */
/* residual (m4_residual.c:7-197)                                       */
/* ------------------------------------------------------------------ */
typedef struct {
  double *flux_face;   /* [n][6][NEQ] */
  double *u_face;      /* [n][6][3]  */
  int    *flux_flag;   /* [n][6]     */
} resid_ws;

/* the computed branch of the face loop (m4_residual.c:112-149): flux_face[f] and u_face[f] of leaf i */
static void face_flux(const layout_t *L, const orc_params *p, const orc_mesh *m, const orc_cell *cells,
                      const double *S, const double *SF, int i, int f, int Nneig, int bc, double *ff, double *uf) {
  const orc_cell *cl = &cells[i];
  const int neq = L->neq;
  const double *Si = &S[(size_t)i * neq];
  uf[0] = 0.0; uf[1] = 0.0; uf[2] = 0.0;
  for (int iv = 0; iv < neq; ++iv) ff[iv] = 0.0;
  frame_t F = mkframe(cl, f);
  sol_t sL, sR;
  fill_sol(L, p, &SF[((size_t)i * 6 + f) * neq], Si, &sL);          /* m4_facercnstr imode 0 */
  for (int k = 0; k < Nneig; ++k) {
    if (bc == 0) {
      int j = m->neigtr[24 * i + 4 * f + k];
      fill_sol(L, p, &SF[((size_t)j * 6 + FCOPP[f]) * neq], &S[(size_t)j * neq], &sR);
    } else {
      fill_sol(L, p, &SF[((size_t)i * 6 + f) * neq], Si, &sR);
      apply_bc(L, p, bc, &F, &sL, &sR);
    }
    hllc_ws W;
    m4_hllc(L, p, &F, bc, &sL, &sR, &W);
    for (int iv = 0; iv < neq; ++iv) {
      ff[iv] += (cl->Area[f][k]) *
                (((W.flux_adv[iv][0] - 0.0) * cl->nx[f] + (W.flux_adv[iv][1] - 0.0) * cl->ny[f] +
                  (W.flux_adv[iv][2] - 0.0) * cl->nz[f]) + W.flux_hllc[iv]);
    }
    uf[0] += W.u_nc / ((double)Nneig);
    uf[1] += W.v_nc / ((double)Nneig);
    uf[2] += W.w_nc / ((double)Nneig);
  }
}

/* par = 0: the sequential flag protocol of m4_residual.c:62-85.  par = 1: the
 * same faces, with a mirrored face recomputed exactly as the earlier leaf
 * computes it instead of read from it, so leaves are independent (OpenMP) and
 * the result is bit-identical. */
static int residual_leaf(const layout_t *L, const orc_params *p, const orc_mesh *m,
                         const orc_cell *cells, const double *S, const double *SF, int i,
                         resid_ws *R, double *RHS, int par) {
  const orc_cell *cl = &cells[i];
  const int neq = L->neq;
  double *rhs = &RHS[(size_t)i * neq];
  for (int iv = 0; iv < neq; ++iv) rhs[iv] = 0.0;
  sol_t sc;
  const double *Si = &S[(size_t)i * neq];
  fill_sol(L, p, Si, Si, &sc);
  double source_avf[ORC_MAXMAT], source_st[ORC_MAXMAT][3][3];
  for (int iv = 0; iv < L->n3; ++iv) source_avf[iv] = sc.avf[iv];
  for (int iv = 0; iv < L->n4; ++iv) {
    for (int a = 0; a < 3; ++a) {
      for (int b = 0; b < 3; ++b) source_st[iv][a][b] = 0.0;
      source_st[iv][a][a] = -sc.avf[iv] * sc.p_hydro[iv];
    }
  }
  int SI0 = m->SI[2 * i];
  for (int f = 0; f < 6; f++) {
    double *ff = &R->flux_face[((size_t)i * 6 + f) * neq];
    double *uf = &R->u_face[((size_t)i * 6 + f) * 3];
    uf[0] = 0.0; uf[1] = 0.0; uf[2] = 0.0;
    int bc = m->bc[6 * i + f];
    int lfc = m->lfc_neigs[6 * i + f];
    int Nneig = 1, compute = 1;
    if (bc == 0) {
      if (lfc == 1) {
        int j = m->neigtr[24 * i + 4 * f];
        int fo = FCOPP[f];
        if (m->lfc_neigs[6 * j + fo] == 1) {       /* 1:1 */
          if (par) {
            if (SI0 == 0 && m->SI[2 * j] == 0 && j < i) {
              double ffj[ORC_MAXNEQ], ufj[3];
              face_flux(L, p, m, cells, S, SF, j, fo, 1, 0, ffj, ufj);
              if (p->fix & ORC_FIX_NONCONS) { uf[0] = ufj[0]; uf[1] = ufj[1]; uf[2] = ufj[2]; }
              else { uf[0] = -ufj[0]; uf[1] = -ufj[1]; uf[2] = -ufj[2]; }
              for (int iv = 0; iv < neq; ++iv) ff[iv] = -ffj[iv];
              compute = 0;
            }
          } else if (SI0 == 0) {                   /* one rank: parts always equal */
            if (R->flux_flag[6 * i + f] == 1) {
              const double *uo = &R->u_face[((size_t)j * 6 + fo) * 3];
              const double *fo_ = &R->flux_face[((size_t)j * 6 + fo) * neq];
              if (p->fix & ORC_FIX_NONCONS) { uf[0] = uo[0]; uf[1] = uo[1]; uf[2] = uo[2]; }
              else { uf[0] = -uo[0]; uf[1] = -uo[1]; uf[2] = -uo[2]; }
              for (int iv = 0; iv < neq; ++iv) ff[iv] = -fo_[iv];
              compute = 0;
            } else {
              R->flux_flag[6 * j + fo] = 1;
              compute = 1;
            }
          }
        }
      } else {
        Nneig = lfc;
      }
    }
    double Area_f = 0.0;
    for (int k = 0; k < Nneig; ++k) Area_f += cl->Area[f][k];
    if (compute) face_flux(L, p, m, cells, S, SF, i, f, Nneig, bc, ff, uf);
    for (int iv = 0; iv < neq; ++iv) rhs[iv] += ff[iv];
    double u_nc = uf[0], nx = cl->nx[f];
    double v_nc = uf[1], ny = cl->ny[f];
    double w_nc = uf[2], nz = cl->nz[f];
    const int fixnc = (p->fix & ORC_FIX_NONCONS) != 0;
    for (int iv = 0; iv < L->n3; ++iv) {
      double t = Area_f * (source_avf[iv] * u_nc * nx + source_avf[iv] * v_nc * ny + source_avf[iv] * w_nc * nz);
      if (fixnc) rhs[L->i3 + iv] -= t; else rhs[L->i3 + iv] += t;
    }
    for (int iv = 0; iv < L->n4; ++iv) {
      double (*s)[3] = source_st[iv];
      double t = Area_f * (s[0][0] * u_nc * nx + s[0][1] * v_nc * nx + s[0][2] * w_nc * nx +
                           s[1][0] * u_nc * ny + s[1][1] * v_nc * ny + s[1][2] * w_nc * ny +
                           s[2][0] * u_nc * nz + s[2][1] * v_nc * nz + s[2][2] * w_nc * nz);
      if (fixnc) rhs[L->i4 + iv] -= t; else rhs[L->i4 + iv] += t;
    }
  }
  for (int v = 0; v < neq; v++) if (isnan(rhs[v])) return -1;
  return 0;
}

static int rhs_all(const layout_t *L, const orc_params *p, const orc_mesh *m, const orc_cell *cells,
                   const double *S, const double *SF, double *RHS, resid_ws *R, int par) {
  if (par) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int i = 0; i < m->n; i++) bad |= residual_leaf(L, p, m, cells, S, SF, i, R, RHS, 1) != 0;
    return bad ? -1 : 0;
  }
  memset(R->flux_flag, 0, sizeof(int) * 6 * (size_t)m->n);   /* RK.c:11-13 resets flags each stage */
  for (int i = 0; i < m->n; i++) {
    int rc = residual_leaf(L, p, m, cells, S, SF, i, R, RHS, 0);
    if (rc) return rc;
  }
  return 0;
}

/* RK.c:17-45 */
static void rk_update(const layout_t *L, int cs, double dt, double Vol, const double *SN, double *S, const double *RHS) {
  for (int e = 0; e < L->neq; e++) {
    switch (cs) {
      case 1: S[e] = SN[e] - dt * (1.0 / Vol) * RHS[e]; break;
      case 2: S[e] = (3.0 / 4.0) * SN[e] + (1.0 / 4.0) * S[e] - (1.0 / 4.0) * dt * RHS[e] / Vol; break;
      case 3: S[e] = (1.0 / 3.0) * SN[e] + (2.0 / 3.0) * S[e] - (2.0 / 3.0) * dt * RHS[e] / Vol; break;
      case 4: S[e] = (1.0 / 2.0) * SN[e] + (1.0 / 2.0) * S[e] - (1.0 / 2.0) * dt * RHS[e] / Vol; break;
    }
  }
}

static int nstages(int rk) { return rk == 3 ? 3 : 2; }
static int stage_case(int rk, int irk) {
  if (rk == 22) return irk == 1 ? 1 : 4;
  return irk;
}

static int stage_impl(const layout_t *L, const orc_mesh *m, const orc_cell *cells, const orc_params *p,
                      int irk, const double *SN, double *S, double *SF, double *RHS, resid_ws *R, int par) {
  size_t neq = L->neq;
  const int n = m->n;
#pragma omp parallel for schedule(static) if (par)
  for (int i = 0; i < n; i++) faceinterp_leaf(L, p, m, &cells[i], S, i, &SF[(size_t)i * 6 * neq]);
  int rc = rhs_all(L, p, m, cells, S, SF, RHS, R, par);
  if (rc) return rc;
  int cs = stage_case(p->rk, irk);
#pragma omp parallel for schedule(static) if (par)
  for (int i = 0; i < n; i++) {
    rk_update(L, cs, p->dt, cells[i].Vol, &SN[i * neq], &S[i * neq], &RHS[i * neq]);
    relax_vec(L, p, &S[i * neq]);   /* m4_relaxation_fluid.c:10-18 */
  }
  return 0;
}

static int alloc_ws(const orc_mesh *m, size_t neq, double **SF, double **RHS, resid_ws *R) {
  size_t n = m->n;
  *SF = malloc(sizeof(double) * n * 6 * neq);
  *RHS = malloc(sizeof(double) * n * neq);
  R->flux_face = malloc(sizeof(double) * n * 6 * neq);
  R->u_face = malloc(sizeof(double) * n * 6 * 3);
  R->flux_flag = malloc(sizeof(int) * n * 6);
  return (*SF && *RHS && R->flux_face && R->u_face && R->flux_flag) ? 0 : -2;
}

static void free_ws(double *SF, double *RHS, resid_ws *R) {
  free(SF); free(RHS); free(R->flux_face); free(R->u_face); free(R->flux_flag);
}

int orc_stage(const orc_mesh *m, const orc_cell *cells, const orc_params *p,
              int irk, const double *SN, double *S, double *SF_out, double *RHS_out) {
  layout_t L = mk_layout(p->nmat);
  double *SF, *RHS; resid_ws R;
  if (alloc_ws(m, L.neq, &SF, &RHS, &R)) return -2;
  int rc = stage_impl(&L, m, cells, p, irk, SN, S, SF, RHS, &R, 0);
  if (SF_out) memcpy(SF_out, SF, sizeof(double) * m->n * 6 * L.neq);
  if (RHS_out) memcpy(RHS_out, RHS, sizeof(double) * m->n * L.neq);
  free_ws(SF, RHS, &R);
  return rc;
}

static int advance_impl(const orc_mesh *m, const orc_cell *cells, const orc_params *p, double *S, int nsteps,
                        int par);

int orc_advance(const orc_mesh *m, const orc_cell *cells, const orc_params *p, double *S, int nsteps) {
  return advance_impl(m, cells, p, S, nsteps, 0);
}

int orc_advance_par(const orc_mesh *m, const orc_cell *cells, const orc_params *p, double *S, int nsteps,
                    int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  return advance_impl(m, cells, p, S, nsteps, 1);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static int advance_impl(const orc_mesh *m, const orc_cell *cells, const orc_params *p, double *S, int nsteps,
                        int par) {
  layout_t L = mk_layout(p->nmat);
  size_t nv = (size_t)m->n * L.neq;
  double *SF, *RHS; resid_ws R;
  double *SN = malloc(sizeof(double) * nv);
  if (!SN || alloc_ws(m, L.neq, &SF, &RHS, &R)) return -2;
  int rc = 0;
  for (int step = 0; step < nsteps && !rc; step++) {
    memcpy(SN, S, sizeof(double) * nv);     /* RK(0) */
    for (int irk = 1; irk <= nstages(p->rk) && !rc; irk++)
      rc = stage_impl(&L, m, cells, p, irk, SN, S, SF, RHS, &R, par);
  }
  free(SN);
  free_ws(SF, RHS, &R);
  return rc;
}

/* ------------------------------------------------------------------ */
/* pointwise entry points for closed-form tests                          */
/* ------------------------------------------------------------------ */
void orc_hllc_state(const orc_params *p, const double *sfL, const double *sfR,
                    const double *n, const double *t1, const double *t2, int bc,
                    double *flux, double *ustar) {
  layout_t L = mk_layout(p->nmat);
  frame_t F = {n[0], n[1], n[2], t1[0], t1[1], t1[2], t2[0], t2[1], t2[2]};
  sol_t sL, sR;
  fill_sol(&L, p, sfL, sfL, &sL);
  fill_sol(&L, p, sfR, sfR, &sR);
  if (bc) apply_bc(&L, p, bc, &F, &sL, &sR);
  hllc_ws W;
  m4_hllc(&L, p, &F, bc, &sL, &sR, &W);
  for (int iv = 0; iv < L.neq; ++iv)
    flux[iv] = ((W.flux_adv[iv][0] - 0.0) * n[0] + (W.flux_adv[iv][1] - 0.0) * n[1] + (W.flux_adv[iv][2] - 0.0) * n[2]) + W.flux_hllc[iv];
  ustar[0] = W.u_nc; ustar[1] = W.v_nc; ustar[2] = W.w_nc;
}

void orc_bc_ghost(const orc_params *p, const double *sf, const double *n, const double *t1, const double *t2,
                  int code, double *uvw, double *vec) {
  layout_t L = mk_layout(p->nmat);
  frame_t F = {n[0], n[1], n[2], t1[0], t1[1], t1[2], t2[0], t2[1], t2[2]};
  sol_t sL, sR;
  fill_sol(&L, p, sf, sf, &sL);
  sR = sL;
  apply_bc(&L, p, code, &F, &sL, &sR);
  uvw[0] = sR.u; uvw[1] = sR.v; uvw[2] = sR.w;
  for (int iv = 0; iv < L.neq; ++iv) vec[iv] = sR.vec[iv];
}

void orc_physical_flux(const orc_params *p, const double *vec, const double *n, double *flux) {
  layout_t L = mk_layout(p->nmat);
  sol_t s;
  fill_sol(&L, p, vec, vec, &s);
  hllc_ws W;
  m4_flux(&L, &s, 0, &W);
  for (int iv = 0; iv < L.neq; ++iv)
    flux[iv] = W.flux_adv[iv][0] * n[0] + W.flux_adv[iv][1] * n[1] + W.flux_adv[iv][2] * n[2];
}

/* ------------------------------------------------------------------ */
/* indicator (interface_location.c, criterionsplt.c)                     */
/* ------------------------------------------------------------------ */
static int inter_max(const orc_crit *c) {   /* m4_input.c:52-54 */
  if (c->inter0_on == 1 && c->inter1_on == 1) return c->inter0_lrs + c->inter1_lrs;
  if (c->inter0_on == 1 && c->inter1_on == 0) return c->inter0_lrs;
  return 0;
}

int orc_interface_location(const orc_mesh *m, const orc_params *p, const orc_crit *c,
                           const double *S, int *SI) {
  layout_t L = mk_layout(p->nmat);
  size_t neq = L.neq;
  for (int i = 0; i < m->n; i++) {
    SI[2 * i] = 0; SI[2 * i + 1] = 0;
    int flag = 0;
    for (int f = 0; f < 6 && !flag; f++) {
      if (m->bc[6 * i + f] != 0) continue;
      int lfc = m->lfc_neigs[6 * i + f];
      int nsfc = lfc > 1 ? lfc : 1;
      for (int k = 0; k < nsfc; k++) {
        int j = m->neigtr[24 * i + 4 * f + k];
        sol_t sL, sR;
        fill_sol(&L, p, &S[i * neq], &S[i * neq], &sL);
        fill_sol(&L, p, &S[j * neq], &S[j * neq], &sR);
        for (int v = 0; v < L.n3; ++v) {
          if (sL.phi[v] * sR.phi[v] < 0.0) {
            if (v == 0) SI[2 * i] = 1; else SI[2 * i + 1] = 1;
            flag = 1;
          }
        }
      }
    }
  }
  for (int lvl = 2; lvl < inter_max(c); lvl++) {
    for (int i = 0; i < m->n; i++) {
      int flag = 0;
      for (int f = 0; f < 6 && !flag; f++) {
        if (m->bc[6 * i + f] != 0) continue;
        int lfc = m->lfc_neigs[6 * i + f];
        int nsfc = lfc > 1 ? lfc : 1;
        for (int k = 0; k < nsfc; k++) {
          int j = m->neigtr[24 * i + 4 * f + k];
          for (int v = 0; v < 2; ++v) {
            if (SI[2 * j + v] == (lvl - 1) && SI[2 * i + v] == 0) { SI[2 * i + v] = lvl; flag = 1; }
          }
        }
      }
    }
  }
  return 0;
}

static double rel_jump(double x_cur, double x_nb) {
  const double eps = 1.0e-8;
  if ((fabs(x_nb) > eps) || (fabs(x_cur) > eps)) return fabs(x_nb - x_cur) / MAXF(fabs(x_nb), fabs(x_cur));
  return 0.0;
}

int orc_criterionsplt(const orc_mesh *m, const orc_params *p, const orc_crit *c,
                      const double *S, const int *SI, int *split) {
  layout_t L = mk_layout(p->nmat);
  size_t neq = L.neq;
  for (int i = 0; i < m->n; i++) {
    int lt = 1;
    if (c->uniform_on == 1) lt = MAXF(lt, c->uniform_lvl);
    if (c->prs_on == 1 || c->rho_on == 1 || c->vel_on == 1) {   /* criterion_grad, criterionsplt.c:40-140 */
      double gp = 0.0, gr = 0.0, gv = 0.0, grho = 0.0;
      double lg = 1;
      for (int f = 0; f < 6; f++) {
        if (m->bc[6 * i + f] != 0) continue;
        int lfc = m->lfc_neigs[6 * i + f];
        int nsfc = lfc > 1 ? lfc : 1;
        for (int k = 0; k < nsfc; k++) {
          int j = m->neigtr[24 * i + 4 * f + k];
          sol_t sL, sR;
          fill_sol(&L, p, &S[i * neq], &S[i * neq], &sL);
          fill_sol(&L, p, &S[j * neq], &S[j * neq], &sR);
          if (c->prs_on == 1) {
            double pl = 0.0, pr = 0.0;
            for (int iv = 0; iv < L.nmat; ++iv) { pl += sL.avf[iv] * sL.p_hydro[iv]; pr += sR.avf[iv] * sR.p_hydro[iv]; }
            gp = MAXF(rel_jump(pl, pr), gp);
          }
          if (c->rho_on == 1) {
            grho = rel_jump(sL.r, sR.r);
            gr = MAXF(grho, gr);
          }
          if (c->vel_on == 1) {
            /* criterionsplt.c:113: the maximum is taken over the density jump */
            gv = MAXF(grho, gv);
          }
        }
      }
      if (c->prs_on == 1 && fabs(gp) >= c->prs_thr) lg = c->prs_lvl;
      if (c->rho_on == 1 && fabs(gr) >= c->rho_thr) lg = c->rho_lvl;
      if (c->vel_on == 1 && fabs(gv) >= c->vel_thr) lg = c->vel_lvl;
      lt = (int)MAXF((double)lt, lg);
    }
    if (c->inter0_on != 0) {
      for (int v = 0; v < 2; ++v) {
        int s = SI[2 * i + v];
        if (s >= 1 && s < c->inter0_lrs) lt = MAXF(lt, c->inter0_lvl);
        if (s >= c->inter0_lrs && s < c->inter1_lrs) lt = MAXF(lt, c->inter1_lvl);
      }
    }
    split[i] = lt;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* CFL time step (new: the reference runs a fixed DT)                    */
/* ------------------------------------------------------------------ */
double orc_cfl_dt(const orc_mesh *m, const orc_cell *cells, const orc_params *p, const double *S, double cfl) {
  (void)cells;
  layout_t L = mk_layout(p->nmat);
  double dt = INFINITY;
  for (int i = 0; i < m->n; i++) {
    sol_t s;
    fill_sol(&L, p, &S[(size_t)i * L.neq], &S[(size_t)i * L.neq], &s);
    double c2 = 0.0;
    for (int iv = 0; iv < L.nmat; ++iv) c2 += s.c2Y[iv];
    double c = sqrt(c2);
    double lam = MAXF(MAXF(fabs(s.u), fabs(s.v)), fabs(s.w)) + c;
    double h = m->rootsize / (double)(1 << (m->level[i] - 1));
    double d = h / lam;
    if (d < dt) dt = d;
  }
  return cfl * dt;
}
