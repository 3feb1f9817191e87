/*
 * fv_mesh.c -- forest of octrees on the host: leaf list, connectivity,
 * decimal tags, AMR (merge/split with state transfer) and the packed image
 * the device kernels read.
 *
 * The octree is kept as a linear list of leaves in depth-first order
 * (rebuildlist.c:17-77) with a hash from (root, level, i, j, k) to the leaf,
 * instead of the reference's pointer-linked BRANCH tree; the connectivity it
 * produces is the reference's (brchconn.c, elconn.c, elconnfin.c):
 * same-level neighbour, else the coarser leaf covering the face (2:1), else
 * the four finer leaves in fcrefnd2elnd order (1:4).
 */
#include "forestfv.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>

/* face -> axis, direction (operatetp.c:60-69 node lists) */
static const int FAX[6] = {1, 0, 1, 0, 2, 2};
static const int FSG[6] = {-1, 1, 1, -1, -1, 1};
static const int FCOPP[6] = {2, 3, 0, 1, 5, 4};                 /* operatetp.c:268-279 */
static const int BCSIDE[6] = {2, 1, 3, 0, 4, 5};                /* face -> -x,+x,-y,+y,-z,+z */
/* fcrefnd2elnd, typ 0 (operatetp.c:304-313): reflected face node -> element node */
static const int FCREF[6][4] = {
  {0, 4, 5, 1}, {3, 1, 5, 7}, {2, 3, 7, 6}, {2, 6, 4, 0}, {2, 0, 1, 3}, {6, 7, 5, 4}};

/* ------------------------------------------------------------------ */
/* hash (root, level, i, j, k) -> leaf                                  */
/* ------------------------------------------------------------------ */
static uint64_t mkkey(int r, int l, int i, int j, int k) {
  return (((uint64_t)r << 40) | ((uint64_t)l << 36) | ((uint64_t)i << 24) | ((uint64_t)j << 12) | (uint64_t)k) + 1u;
}

static size_t hslot(uint64_t key, size_t cap) {
  key ^= key >> 33; key *= 0xff51afd7ed558ccdULL; key ^= key >> 33; key *= 0xc4ceb9fe1a85ec53ULL; key ^= key >> 33;
  return (size_t)key & (cap - 1);
}

static int hash_build(fv_mesh *m) {
  size_t cap = 16;
  while (cap < (size_t)m->n * 2 + 2) cap <<= 1;
  m->hcap = cap;
  m->hkey = calloc(cap, sizeof(uint64_t));
  m->hval = malloc(cap * sizeof(int));
  if (!m->hkey || !m->hval) return FV_ERR_NOMEM;
  for (int i = 0; i < m->n; i++) {
    uint64_t key = mkkey(m->root[i], m->level[i], m->ci[3 * i], m->ci[3 * i + 1], m->ci[3 * i + 2]);
    size_t s = hslot(key, cap);
    while (m->hkey[s] != 0) {
      if (m->hkey[s] == key) return FV_ERR_MESH;    /* duplicate leaf */
      s = (s + 1) & (cap - 1);
    }
    m->hkey[s] = key; m->hval[s] = i;
  }
  return FV_OK;
}

int fv_mesh_find(const fv_mesh *m, int r, int l, int i, int j, int k) {
  if (l < 1 || l > FV_MAXLEVEL) return -1;
  int side = 1 << (l - 1);
  if (i < 0 || j < 0 || k < 0 || i >= side || j >= side || k >= side) return -1;
  uint64_t key = mkkey(r, l, i, j, k);
  size_t s = hslot(key, m->hcap);
  while (m->hkey[s] != 0) {
    if (m->hkey[s] == key) return m->hval[s];
    s = (s + 1) & (m->hcap - 1);
  }
  return -1;
}

/* ------------------------------------------------------------------ */
/* tags (tagaddress.c:3-11, communicate.c:252-271)                      */
/* ------------------------------------------------------------------ */
int fv_tag_from_address(int level, const int *ci) {
  int tag = 0, p10 = 1;
  for (int l = 2; l <= level; l++) {
    int sh = level - l;
    int ikid = ((ci[0] >> sh) & 1) + 2 * ((ci[1] >> sh) & 1) + 4 * ((ci[2] >> sh) & 1);
    tag += p10 * (1 + ikid);
    p10 *= 10;
  }
  return tag;
}

int fv_address_from_tag(int tag, int *level, int *ci) {
  int l = 1;
  ci[0] = ci[1] = ci[2] = 0;
  if (tag < 0) return FV_ERR_ARG;
  while (tag > 0) {
    int d = tag % 10;
    if (d < 1 || d > 8) return FV_ERR_ARG;
    int ikid = d - 1;
    ci[0] = 2 * ci[0] + (ikid & 1);
    ci[1] = 2 * ci[1] + ((ikid >> 1) & 1);
    ci[2] = 2 * ci[2] + ((ikid >> 2) & 1);
    tag /= 10;
    l++;
  }
  *level = l;
  return FV_OK;
}

/* ------------------------------------------------------------------ */
/* construction                                                        */
/* ------------------------------------------------------------------ */
static fv_mesh *mesh_alloc(const int nroot[3], const double origin[3], double rootsize,
                           const int bcside[6], int n) {
  fv_mesh *m = calloc(1, sizeof(fv_mesh));
  if (!m) return NULL;
  m->n = n;
  for (int a = 0; a < 3; a++) { m->nroot[a] = nroot[a]; m->origin[a] = origin[a]; }
  m->rootsize = rootsize;
  for (int f = 0; f < 6; f++) m->bcside[f] = bcside[f];
  size_t N = n > 0 ? (size_t)n : 1;
  m->root = malloc(N * sizeof(int));
  m->level = malloc(N * sizeof(int));
  m->ci = malloc(3 * N * sizeof(int));
  m->tag = malloc(N * sizeof(int));
  m->lfc_neigs = malloc(6 * N * sizeof(int));
  m->neigtr = malloc(24 * N * sizeof(int));
  m->bc = malloc(6 * N * sizeof(int));
  m->SI = calloc(2 * N, sizeof(int));
  if (!m->root || !m->level || !m->ci || !m->tag || !m->lfc_neigs || !m->neigtr || !m->bc || !m->SI) {
    fv_mesh_free(m);
    return NULL;
  }
  return m;
}

void fv_mesh_free(fv_mesh *m) {
  if (!m) return;
  free(m->root); free(m->level); free(m->ci); free(m->tag);
  free(m->lfc_neigs); free(m->neigtr); free(m->bc); free(m->SI);
  free(m->hkey); free(m->hval);
  free(m);
}

/* neighbour root across face f of root r, or -1 at the domain boundary */
static int root_across(const fv_mesh *m, int r, int f) {
  int p[3] = {r % m->nroot[0], (r / m->nroot[0]) % m->nroot[1], r / (m->nroot[0] * m->nroot[1])};
  p[FAX[f]] += FSG[f];
  if (p[FAX[f]] < 0 || p[FAX[f]] >= m->nroot[FAX[f]]) return -1;
  return p[0] + m->nroot[0] * (p[1] + m->nroot[1] * p[2]);
}

static int connect(fv_mesh *m) {
  m->maxlevel = 1;
  for (int i = 0; i < m->n; i++) {
    int r = m->root[i], l = m->level[i];
    if (l > m->maxlevel) m->maxlevel = l;
    m->tag[i] = fv_tag_from_address(l, &m->ci[3 * i]);
    int side = 1 << (l - 1);
    for (int f = 0; f < 6; f++) {
      int *nt = &m->neigtr[24 * i + 4 * f];
      nt[0] = nt[1] = nt[2] = nt[3] = -1;
      int c[3] = {m->ci[3 * i], m->ci[3 * i + 1], m->ci[3 * i + 2]};
      int a = FAX[f];
      int rr = r;
      c[a] += FSG[f];
      if (c[a] < 0 || c[a] >= side) {
        rr = root_across(m, r, f);
        c[a] = (c[a] < 0) ? side - 1 : 0;
      }
      if (rr < 0) {                           /* domain boundary */
        m->lfc_neigs[6 * i + f] = 0;
        m->bc[6 * i + f] = m->bcside[BCSIDE[f]];
        if (m->bc[6 * i + f] == 0) return FV_ERR_ARG;   /* a boundary face needs a non-zero code */
        continue;
      }
      m->bc[6 * i + f] = 0;
      int j = fv_mesh_find(m, rr, l, c[0], c[1], c[2]);
      if (j >= 0) { m->lfc_neigs[6 * i + f] = 1; nt[0] = j; continue; }
      if (l > 1) {
        j = fv_mesh_find(m, rr, l - 1, c[0] >> 1, c[1] >> 1, c[2] >> 1);
        if (j >= 0) { m->lfc_neigs[6 * i + f] = 1; nt[0] = j; continue; }
      }
      /* four finer leaves, elconnfin.c:14-22 with neigag = 0 for aligned roots */
      int fo = FCOPP[f];
      for (int q = 0; q < 4; q++) {
        int nd = FCREF[fo][q];
        int k = fv_mesh_find(m, rr, l + 1, 2 * c[0] + (nd & 1), 2 * c[1] + ((nd >> 1) & 1), 2 * c[2] + ((nd >> 2) & 1));
        if (k < 0) return FV_ERR_MESH;
        nt[q] = k;
      }
      m->lfc_neigs[6 * i + f] = 4;
    }
  }
  return FV_OK;
}

/* interleave 3 x 10-bit coordinates, x lowest (ikid = i + 2j + 4k) */
static uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  uint64_t key = 0;
  for (int b = 0; b < 10; b++)
    key |= ((uint64_t)((x >> b) & 1) << (3 * b)) | ((uint64_t)((y >> b) & 1) << (3 * b + 1)) |
           ((uint64_t)((z >> b) & 1) << (3 * b + 2));
  return key;
}

static uint64_t dfs_key(int r, int l, const int *c) {
  int sh = FV_MAXLEVEL - l;
  return ((uint64_t)r << 32) | morton3((uint32_t)c[0] << sh, (uint32_t)c[1] << sh, (uint32_t)c[2] << sh);
}

typedef struct { uint64_t key; int idx; } kv_t;
static int kv_cmp(const void *a, const void *b) {
  uint64_t x = ((const kv_t *)a)->key, y = ((const kv_t *)b)->key;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static fv_mesh *finish(fv_mesh *m, int *status) {
  int rc = hash_build(m);
  if (rc == FV_OK) rc = connect(m);
  if (status) *status = rc;
  if (rc != FV_OK) { fv_mesh_free(m); return NULL; }
  return m;
}

fv_mesh *fv_mesh_from_leaves(const int nroot[3], const double origin[3], double rootsize,
                             const int bcside[6], int n, const int *root, const int *level,
                             const int *ci, int *status) {
  fv_mesh *m = mesh_alloc(nroot, origin, rootsize, bcside, n);
  if (!m) { if (status) *status = FV_ERR_NOMEM; return NULL; }
  kv_t *kv = malloc((size_t)(n > 0 ? n : 1) * sizeof(kv_t));
  if (!kv) { fv_mesh_free(m); if (status) *status = FV_ERR_NOMEM; return NULL; }
  int nr = nroot[0] * nroot[1] * nroot[2];
  for (int i = 0; i < n; i++) {
    if (level[i] < 1 || level[i] > FV_MAXLEVEL || root[i] < 0 || root[i] >= nr) {
      free(kv); fv_mesh_free(m); if (status) *status = FV_ERR_ARG; return NULL;
    }
    kv[i].key = dfs_key(root[i], level[i], &ci[3 * i]);
    kv[i].idx = i;
  }
  qsort(kv, n, sizeof(kv_t), kv_cmp);
  for (int i = 0; i < n; i++) {
    int s = kv[i].idx;
    m->root[i] = root[s]; m->level[i] = level[s];
    m->ci[3 * i] = ci[3 * s]; m->ci[3 * i + 1] = ci[3 * s + 1]; m->ci[3 * i + 2] = ci[3 * s + 2];
  }
  free(kv);
  return finish(m, status);
}

fv_mesh *fv_mesh_uniform(const int nroot[3], const double origin[3], double rootsize,
                         int level, const int bcside[6]) {
  if (level < 1 || level > FV_MAXLEVEL) return NULL;
  int side = 1 << (level - 1);
  long nr = (long)nroot[0] * nroot[1] * nroot[2];
  long per = (long)side * side * side;
  if (nr * per > 0x7fffffffL) return NULL;
  fv_mesh *m = mesh_alloc(nroot, origin, rootsize, bcside, (int)(nr * per));
  if (!m) return NULL;
  long i = 0;
  for (int r = 0; r < nr; r++) {
    for (long q = 0; q < per; q++, i++) {   /* DFS order within a uniform root is Morton order */
      int c[3] = {0, 0, 0};
      for (int b = 0; b < level - 1; b++) {
        c[0] |= (int)((q >> (3 * b)) & 1) << b;
        c[1] |= (int)((q >> (3 * b + 1)) & 1) << b;
        c[2] |= (int)((q >> (3 * b + 2)) & 1) << b;
      }
      m->root[i] = r; m->level[i] = level;
      m->ci[3 * i] = c[0]; m->ci[3 * i + 1] = c[1]; m->ci[3 * i + 2] = c[2];
    }
  }
  return finish(m, NULL);
}

void fv_mesh_centre(const fv_mesh *m, int i, double c[3]) {
  int r = m->root[i];
  int p[3] = {r % m->nroot[0], (r / m->nroot[0]) % m->nroot[1], r / (m->nroot[0] * m->nroot[1])};
  double h = m->rootsize / (double)(1 << (m->level[i] - 1));
  for (int a = 0; a < 3; a++) c[a] = m->origin[a] + (double)p[a] * m->rootsize + ((double)m->ci[3 * i + a] + 0.5) * h;
}

/* ------------------------------------------------------------------ */
/* geometry tables (normalvector.c, tagvector.c, meshproperties.c,      */
/* distance.c) on a cube of each level's edge length                    */
/* ------------------------------------------------------------------ */
static const int FCND[6][4] = {
  {0, 1, 5, 4}, {3, 7, 5, 1}, {2, 6, 7, 3}, {2, 0, 4, 6}, {2, 3, 1, 0}, {6, 4, 5, 7}};
#define SQ(x) ((x) * (x))

void fv_geo_build(double rootsize, fv_geo *g) {
  memset(g, 0, sizeof *g);
  double X[8], Y[8], Z[8];
  for (int q = 0; q < 8; q++) { X[q] = (double)(q & 1); Y[q] = (double)((q >> 1) & 1); Z[q] = (double)((q >> 2) & 1); }
  for (int f = 0; f < 6; f++) {
    const int *q = FCND[f];
    double ax = X[q[0]], ay = Y[q[0]], az = Z[q[0]], bx = X[q[1]], by = Y[q[1]], bz = Z[q[1]];
    double cx = X[q[2]], cy = Y[q[2]], cz = Z[q[2]];
    double n0 = (by - ay) * (cz - az) - (cy - ay) * (bz - az);
    double n1 = (bx - ax) * (cz - az) - (cx - ax) * (bz - az);
    double n2 = (bx - ax) * (cy - ay) - (cx - ax) * (by - ay);
    double nrm = sqrt(SQ(n0) + SQ(n1) + SQ(n2));
    double nx = n0 / nrm, ny = -(n1) / nrm, nz = n2 / nrm;
    double tn = sqrt(SQ(by - ay) + SQ(bx - ax) + SQ(bz - az));
    double tx = (bx - ax) / tn, ty = (by - ay) / tn, tz = (bz - az) / tn;
    double nn = sqrt(SQ(ny * tz - ty * nz) + SQ(nx * tz - tx * nz) + SQ(nx * ty - tx * ny));
    double *F = g->frame[f];
    F[0] = nx; F[1] = ny; F[2] = nz; F[3] = tx; F[4] = ty; F[5] = tz;
    F[6] = ((ny * tz) - (ty * nz)) / nn;
    F[7] = -((nx * tz) - (tx * nz)) / nn;
    F[8] = ((nx * ty) - (tx * ny)) / nn;
  }
  static const int vn[6][3] = {{1, 3, 5}, {5, 3, 7}, {5, 7, 4}, {3, 2, 7}, {7, 2, 6}, {7, 6, 4}};
  for (int l = 1; l <= FV_MAXLEVEL + 1; l++) {
    double h = rootsize / (double)(1 << (l - 1));
    g->h[l] = h;
    double A = 0.0;                       /* face f0 of a cube of edge h, two triangles */
    double x[8], y[8], z[8];
    for (int q = 0; q < 8; q++) { x[q] = X[q] * h; y[q] = Y[q] * h; z[q] = Z[q] * h; }
    for (int it = 0; it < 2; it++) {
      const int *q = FCND[0];
      double dx1 = x[q[1 + it]] - x[q[0]], dx2 = x[q[2 + it]] - x[q[0]];
      double dy1 = y[q[1 + it]] - y[q[0]], dy2 = y[q[2 + it]] - y[q[0]];
      double dz1 = z[q[1 + it]] - z[q[0]], dz2 = z[q[2 + it]] - z[q[0]];
      A += fabs(0.5 * sqrt(SQ(dy1 * dz2 - dz1 * dy2) + SQ(dz1 * dx2 - dx1 * dz2) + SQ(dx1 * dy2 - dy1 * dx2)));
    }
    g->area[l] = A;
    double V = 0.0;
    for (int it = 0; it < 6; it++) {
      double dx[3], dy[3], dz[3];
      for (int in = 0; in < 3; in++) { dx[in] = x[vn[it][in]] - x[0]; dy[in] = y[vn[it][in]] - y[0]; dz[in] = z[vn[it][in]] - z[0]; }
      double c0 = dy[0] * dz[1] - dz[0] * dy[1], c1 = dz[0] * dx[1] - dx[0] * dz[1], c2 = dx[0] * dy[1] - dy[0] * dx[1];
      V += (c0 * dx[2] + c1 * dy[2] + c2 * dz[2]) / 6.0;
    }
    g->vol[l] = V;
    double d = sqrt(SQ(h));               /* centre distance to a same-level neighbour */
    g->dcf[0][l] = 0.5 * d;  g->dcc[0][l] = d;
    g->dcf[1][l] = 0.5 * d;  g->dcc[1][l] = 0.5 * d + 0.25 * d;            /* finer neighbour */
    double D = sqrt(SQ(2.0 * h));         /* own parent centre to the coarser neighbour */
    g->dcf[2][l] = 0.25 * D; g->dcc[2][l] = 0.25 * D + 0.5 * D;
  }
}

/* ------------------------------------------------------------------ */
/* device image                                                         */
/* ------------------------------------------------------------------ */
int fv_mesh_pack(const fv_mesh *m, fv_pack *pk) {
  int n = m->n;
  memset(pk, 0, sizeof *pk);
  pk->n = n;
  pk->nbr = malloc(6 * (size_t)n * sizeof(int));
  pk->info = malloc((size_t)n * sizeof(uint32_t));
  int nfine = 0;
  for (int i = 0; i < 6 * n; i++) if (m->lfc_neigs[i] == 4) nfine++;
  pk->nfine = nfine;
  pk->fine4 = malloc(4 * (size_t)(nfine > 0 ? nfine : 1) * sizeof(int));
  if (!pk->nbr || !pk->info || !pk->fine4) { fv_pack_free(pk); return FV_ERR_NOMEM; }
  int kf = 0;
  for (int i = 0; i < n; i++) {
    uint32_t info = (uint32_t)m->level[i] | (m->SI[2 * i] != 0 ? 16u : 0u);
    for (int f = 0; f < 6; f++) {
      int lfc = m->lfc_neigs[6 * i + f];
      int kind, nb;
      if (lfc == 0) { kind = FV_FK_BOUNDARY; nb = -m->bc[6 * i + f]; }
      else if (lfc == 4) {
        kind = FV_FK_FINE; nb = kf;
        for (int q = 0; q < 4; q++) pk->fine4[4 * kf + q] = m->neigtr[24 * i + 4 * f + q];
        kf++;
      } else {
        int j = m->neigtr[24 * i + 4 * f];
        nb = j;
        if (m->level[j] < m->level[i]) kind = FV_FK_COARSE;
        else if (m->SI[2 * i] == 0 && m->SI[2 * j] == 0 && j < i) kind = FV_FK_MIRROR;
        else kind = FV_FK_SAME;
      }
      pk->nbr[(size_t)f * n + i] = nb;
      info |= (uint32_t)kind << (8 + 3 * f);
    }
    pk->info[i] = info;
  }
  fv_geo_build(m->rootsize, &pk->geo);
  return FV_OK;
}

void fv_pack_free(fv_pack *pk) {
  free(pk->nbr); free(pk->info); free(pk->fine4);
  pk->nbr = NULL; pk->info = NULL; pk->fine4 = NULL;
}

/* ------------------------------------------------------------------ */
/* AMR                                                                  */
/* ------------------------------------------------------------------ */
static int kid_index(const int *c) { return (c[0] & 1) + 2 * (c[1] & 1) + 4 * (c[2] & 1); }

/* criterionsmth.c:3-58, sequential in leaf order as the reference runs it */
int fv_criterionsmth(const fv_mesh *m, int LEVEL, int *split) {
  int n = m->n;
  for (int pass = 0; pass < LEVEL + 1; pass++) {
    for (int i = 0; i < n; i++) {
      int l = m->level[i];
      if (l > 1) {
        /* siblings: the family is complete and all leaves iff 8 consecutive leaves share the parent */
        int first = i - kid_index(&m->ci[3 * i]);
        int fam = 0;
        if (first >= 0 && first + 7 < n) {
          fam = 1;
          for (int k = 0; k < 8 && fam; k++) {
            int s = first + k;
            if (m->level[s] != l || m->root[s] != m->root[i] || kid_index(&m->ci[3 * s]) != k ||
                (m->ci[3 * s] >> 1) != (m->ci[3 * i] >> 1) || (m->ci[3 * s + 1] >> 1) != (m->ci[3 * i + 1] >> 1) ||
                (m->ci[3 * s + 2] >> 1) != (m->ci[3 * i + 2] >> 1))
              fam = 0;
          }
        }
        int imerge = 0;
        if (fam) for (int k = 0; k < 8; k++) if (split[first + k] < l) imerge++;
        if (fam && imerge == 8) split[i] = l - 1;
        else if (l > split[i]) split[i] = l;
      }
      if (l < split[i]) split[i] = l + 1;
    }
    for (int i = 0; i < n; i++) {
      int maxl = 1;
      for (int f = 0; f < 6; f++) {
        if (m->bc[6 * i + f] != 0) continue;
        int lfc = m->lfc_neigs[6 * i + f];
        for (int q = 0; q < lfc; q++) {
          int j = m->neigtr[24 * i + 4 * f + q];
          if (split[j] > maxl) maxl = split[j];
        }
      }
      if (maxl - split[i] > 1) split[i] = maxl - 1;
    }
  }
  return FV_OK;
}

int fv_mesh_adapt(fv_mesh **pm, const fv_params *p, const int *split, double **pS) {
  fv_mesh *m = *pm;
  int n = m->n, neq = fv_neq(p->nmat);
  const double *S = *pS;
  /* count */
  long nn = 0;
  for (int i = 0; i < n;) {
    int l = m->level[i];
    int merged = 0;
    if (l > 1 && kid_index(&m->ci[3 * i]) == 0 && i + 7 < n) {
      int ok = 1;
      for (int k = 0; k < 8 && ok; k++) {
        int s = i + k;
        if (m->level[s] != l || m->root[s] != m->root[i] || kid_index(&m->ci[3 * s]) != k ||
            (m->ci[3 * s] >> 1) != (m->ci[3 * i] >> 1) || (m->ci[3 * s + 1] >> 1) != (m->ci[3 * i + 1] >> 1) ||
            (m->ci[3 * s + 2] >> 1) != (m->ci[3 * i + 2] >> 1) || !(split[s] < l))
          ok = 0;
      }
      merged = ok;
    }
    if (merged) { nn += 1; i += 8; }
    else if (split[i] > l) { if (l + 1 > FV_MAXLEVEL) return FV_ERR_MESH; nn += 8; i += 1; }
    else { nn += 1; i += 1; }
  }
  if (nn > 0x7fffffffL) return FV_ERR_MESH;
  fv_mesh *q = mesh_alloc(m->nroot, m->origin, m->rootsize, m->bcside, (int)nn);
  double *S2 = malloc(sizeof(double) * (size_t)nn * neq);
  if (!q || !S2) { fv_mesh_free(q); free(S2); return FV_ERR_NOMEM; }
  long o = 0;
  for (int i = 0; i < n;) {
    int l = m->level[i];
    int merged = 0;
    if (l > 1 && kid_index(&m->ci[3 * i]) == 0 && i + 7 < n) {
      int ok = 1;
      for (int k = 0; k < 8 && ok; k++) {
        int s = i + k;
        if (m->level[s] != l || m->root[s] != m->root[i] || kid_index(&m->ci[3 * s]) != k ||
            (m->ci[3 * s] >> 1) != (m->ci[3 * i] >> 1) || (m->ci[3 * s + 1] >> 1) != (m->ci[3 * i + 1] >> 1) ||
            (m->ci[3 * s + 2] >> 1) != (m->ci[3 * i + 2] >> 1) || !(split[s] < l))
          ok = 0;
      }
      merged = ok;
    }
    if (merged) {                          /* merge.c:20-55: average of the kids, then relaxation */
      q->root[o] = m->root[i]; q->level[o] = l - 1;
      for (int a = 0; a < 3; a++) q->ci[3 * o + a] = m->ci[3 * i + a] >> 1;
      double *d = &S2[o * neq];
      for (int iv = 0; iv < neq; iv++) {
        d[iv] = 0.0;
        for (int k = 0; k < 8; k++) d[iv] += S[(size_t)(i + k) * neq + iv] / ((double)8);
      }
      fv_relax_host(p, d);
      q->SI[2 * o] = 0; q->SI[2 * o + 1] = 0;
      o += 1; i += 8;
    } else if (split[i] > l) {             /* split.c:425-434: kids copy the parent state */
      for (int k = 0; k < 8; k++) {
        q->root[o] = m->root[i]; q->level[o] = l + 1;
        q->ci[3 * o] = 2 * m->ci[3 * i] + (k & 1);
        q->ci[3 * o + 1] = 2 * m->ci[3 * i + 1] + ((k >> 1) & 1);
        q->ci[3 * o + 2] = 2 * m->ci[3 * i + 2] + ((k >> 2) & 1);
        memcpy(&S2[o * neq], &S[(size_t)i * neq], sizeof(double) * neq);
        q->SI[2 * o] = m->SI[2 * i]; q->SI[2 * o + 1] = 0;
        o++;
      }
      i += 1;
    } else {
      q->root[o] = m->root[i]; q->level[o] = l;
      for (int a = 0; a < 3; a++) q->ci[3 * o + a] = m->ci[3 * i + a];
      memcpy(&S2[o * neq], &S[(size_t)i * neq], sizeof(double) * neq);
      q->SI[2 * o] = m->SI[2 * i]; q->SI[2 * o + 1] = m->SI[2 * i + 1];
      o += 1; i += 1;
    }
  }
  int st;
  q = finish(q, &st);
  if (!q) { free(S2); return st; }
  fv_mesh_free(m);
  free(*pS);
  *pm = q;
  *pS = S2;
  return FV_OK;
}
