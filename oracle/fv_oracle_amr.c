/*
 * fv_oracle_amr.c -- CPU restatement of ForestFV's mesh-adaptation flags and
 * merge/split (criterionsmth.c, refineh.c, merge.c, split.c).
 *
 * TEST INFRASTRUCTURE ONLY (see fv_oracle.h): the oracle of the device
 * regrid and of the host AMR code, never linked into the product.
 *
 * The reference walks its leaf list (run->topo->locl -> lnxt, DFS order) and
 * updates BRANCH.split in place, so a leaf reads the new values of the leaves
 * before it and the old values of the leaves after it.  These loops do the
 * same, over the leaf arrays of orc_mesh in the same order.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fv_oracle.h"
#include "fv_oracle_m5.h"

#define MAXI(a, b) ((a) > (b) ? (a) : (b))

/* key of a leaf's parent: (root, level - 1, ci >> 1) */
static uint64_t parent_key(const orc_mesh *m, int i) {
  const int *c = &m->ci[3 * i];
  return ((uint64_t)m->root[i] << 40) | ((uint64_t)(m->level[i] - 1) << 32) | ((uint64_t)(c[0] >> 1) << 20) |
         ((uint64_t)(c[1] >> 1) << 10) | (uint64_t)(c[2] >> 1);
}

typedef struct { uint64_t key; int leaf; } keyed_t;

static int cmp_keyed(const void *a, const void *b) {
  const keyed_t *x = (const keyed_t *)a, *y = (const keyed_t *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->leaf - y->leaf;
}

/* For every leaf of level > 1: the leaves that share its parent (at most 8),
 * as rows of fam [n][8] (-1 unused), and their count famn [n].  A parent whose
 * eight kids are all leaves has famn == 8; a kid that is itself refined is no
 * leaf and is missing from the row (prnt->kids[ikid]->nkids != 0). */
static int families(const orc_mesh *m, int *fam, int *famn) {
  const int n = m->n;
  keyed_t *k = (keyed_t *)malloc(sizeof(keyed_t) * (size_t)(n > 0 ? n : 1));
  if (!k) return -2;
  int nk = 0;
  for (int i = 0; i < n; i++) {
    famn[i] = 0;
    for (int j = 0; j < 8; j++) fam[8 * i + j] = -1;
    if (m->level[i] > 1) { k[nk].key = parent_key(m, i); k[nk].leaf = i; nk++; }
  }
  qsort(k, (size_t)nk, sizeof(keyed_t), cmp_keyed);
  for (int a = 0; a < nk;) {
    int b = a;
    while (b < nk && k[b].key == k[a].key) b++;
    if (b - a > 8) { free(k); return -2; }
    for (int q = a; q < b; q++) {
      const int i = k[q].leaf;
      famn[i] = b - a;
      for (int r = a; r < b; r++) fam[8 * i + (r - a)] = k[r].leaf;
    }
    a = b;
  }
  free(k);
  return 0;
}

/* criterionsmth (criterionsmth.c:13-56), LEVEL + 1 passes; split [n] in place. */
int orc_criterionsmth(const orc_mesh *m, int LEVEL, int *split) {
  const int n = m->n;
  int *fam = (int *)malloc(sizeof(int) * 8 * (size_t)(n > 0 ? n : 1));
  int *famn = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!fam || !famn) { free(fam); free(famn); return -2; }
  int rc = families(m, fam, famn);
  if (rc != 0) { free(fam); free(famn); return rc; }
  for (int ipass = 0; ipass < LEVEL + 1; ipass++) {
    /* criterionsmth.c:15-34: sibling consensus, then at most one level finer */
    for (int i = 0; i < n; i++) {
      const int lvl = m->level[i];
      if (lvl > 1) {
        int imerge = 0;
        for (int q = 0; q < famn[i]; q++) {
          const int j = fam[8 * i + q];
          if (split[j] < m->level[j]) imerge++;
        }
        if (imerge == 8) split[i] = lvl - 1;   /* prnt->nkids == 8 for a hex */
        else if (lvl > split[i]) split[i] = lvl;
      }
      if (lvl < split[i]) split[i] = lvl + 1;
    }
    /* criterionsmth.c:38-53: no face neighbour more than one level finer */
    for (int i = 0; i < n; i++) {
      int max_lvl = 1;
      for (int f = 0; f < 6; f++) {
        const int nn = m->lfc_neigs[6 * i + f];
        for (int ing = 0; ing < nn; ing++)
          if (m->bc[6 * i + f] == 0) max_lvl = MAXI(split[m->neigtr[(6 * i + f) * 4 + ing]], max_lvl);
      }
      if (max_lvl - split[i] > 1) split[i] = max_lvl - 1;
    }
  }
  free(fam);
  free(famn);
  return 0;
}

/* refineh (refineh.c:3-61) with the state rules of merge.c:20-55 and
 * split.c:425-434.  Merge pass: a family of eight leaf kids that all have
 * split < level becomes its parent, whose state is the kids' sum of S/8 in
 * kid order followed by m4_relaxationvec, whose split is the kids' largest
 * and whose SI is 0.  Split pass: a leaf with split > level becomes its
 * eight kids in ikid = i + 2j + 4k order, each a copy of the parent's state,
 * SI[0] and split, with SI[1] = 0 (leafallocation.c:19).  Both passes keep
 * the DFS order, as the list surgery of merge() and split() does.
 * Outputs are sized by the caller for up to 8 n leaves; *nout the count. */
static int refineh_impl(const orc_mesh *m, int neq, int (*relax)(const void *, double *), const void *rp,
                        const int *split, const double *S, int *nout, int *root_o, int *level_o, int *ci_o,
                        double *S_o, int *SI_o, int *split_o) {
  const int n = m->n;
  int *fam = (int *)malloc(sizeof(int) * 8 * (size_t)(n > 0 ? n : 1));
  int *famn = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  /* the list after the merge pass */
  int *r1 = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *l1 = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *c1 = (int *)malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
  int *s1 = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *si1 = (int *)malloc(sizeof(int) * 2 * (size_t)(n > 0 ? n : 1));
  double *S1 = (double *)malloc(sizeof(double) * (size_t)neq * (size_t)(n > 0 ? n : 1));
  int rc = (fam && famn && r1 && l1 && c1 && s1 && si1 && S1) ? families(m, fam, famn) : -2;
  int n1 = 0;
  for (int i = 0; rc == 0 && i < n;) {
    int merged = 0;
    if (m->level[i] > 1 && famn[i] == 8 && fam[8 * i] == i) {
      int nmerge = 0;
      for (int q = 0; q < 8; q++) {
        const int j = fam[8 * i + q];
        if (split[j] < m->level[j]) nmerge++;
      }
      if (nmerge == 8) {
        /* the family is contiguous in DFS order: kids i .. i + 7 */
        for (int q = 0; q < 8; q++)
          if (fam[8 * i + q] != i + q) rc = -2;
        if (rc != 0) break;
        r1[n1] = m->root[i];
        l1[n1] = m->level[i] - 1;
        for (int d = 0; d < 3; d++) c1[3 * n1 + d] = m->ci[3 * i + d] >> 1;
        int sp = 0;
        for (int q = 0; q < 8; q++) sp = MAXI(sp, split[i + q]);
        s1[n1] = sp;
        si1[2 * n1] = 0; si1[2 * n1 + 1] = 0;
        double *d = &S1[(size_t)n1 * neq];
        for (int iv = 0; iv < neq; iv++) {
          d[iv] = 0.0;
          for (int q = 0; q < 8; q++) d[iv] += S[(size_t)(i + q) * neq + iv] / ((double)8);
        }
        if (relax(rp, d) != 0) rc = -1;   /* m4/m5_relaxationvec, whatever PRELAX is (merge.c:54-59) */
        n1++;
        i += 8;
        merged = 1;
      }
    }
    if (!merged) {
      r1[n1] = m->root[i];
      l1[n1] = m->level[i];
      for (int d = 0; d < 3; d++) c1[3 * n1 + d] = m->ci[3 * i + d];
      s1[n1] = split[i];
      si1[2 * n1] = m->SI[2 * i]; si1[2 * n1 + 1] = m->SI[2 * i + 1];
      memcpy(&S1[(size_t)n1 * neq], &S[(size_t)i * neq], sizeof(double) * (size_t)neq);
      n1++;
      i++;
    }
  }
  int n2 = 0;
  for (int i = 0; rc == 0 && i < n1; i++) {
    if (s1[i] > l1[i]) {
      for (int ikid = 0; ikid < 8; ikid++) {
        root_o[n2] = r1[i];
        level_o[n2] = l1[i] + 1;
        ci_o[3 * n2 + 0] = 2 * c1[3 * i + 0] + (ikid & 1);
        ci_o[3 * n2 + 1] = 2 * c1[3 * i + 1] + ((ikid >> 1) & 1);
        ci_o[3 * n2 + 2] = 2 * c1[3 * i + 2] + ((ikid >> 2) & 1);
        memcpy(&S_o[(size_t)n2 * neq], &S1[(size_t)i * neq], sizeof(double) * (size_t)neq);
        SI_o[2 * n2] = si1[2 * i]; SI_o[2 * n2 + 1] = 0;
        split_o[n2] = s1[i];
        n2++;
      }
    } else {
      root_o[n2] = r1[i];
      level_o[n2] = l1[i];
      for (int d = 0; d < 3; d++) ci_o[3 * n2 + d] = c1[3 * i + d];
      memcpy(&S_o[(size_t)n2 * neq], &S1[(size_t)i * neq], sizeof(double) * (size_t)neq);
      SI_o[2 * n2] = si1[2 * i]; SI_o[2 * n2 + 1] = si1[2 * i + 1];
      split_o[n2] = s1[i];
      n2++;
    }
  }
  *nout = (rc == 0) ? n2 : 0;
  free(fam); free(famn); free(r1); free(l1); free(c1); free(s1); free(si1); free(S1);
  return rc;
}

static int relax4(const void *p, double *d) {
  orc_relax((const orc_params *)p, d);
  return 0;
}

int orc_refineh(const orc_mesh *m, const orc_params *p, const int *split, const double *S, int *nout,
                int *root_o, int *level_o, int *ci_o, double *S_o, int *SI_o, int *split_o) {
  return refineh_impl(m, orc_neq(p->nmat), relax4, p, split, S, nout, root_o, level_o, ci_o, S_o, SI_o, split_o);
}

static int relax5(const void *p, double *d) { return orc_m5_relax((const orc_m5_params *)p, d); }

int orc_m5_refineh(const orc_mesh *m, const orc_m5_params *p, const int *split, const double *S, int *nout,
                   int *root_o, int *level_o, int *ci_o, double *S_o, int *SI_o, int *split_o) {
  return refineh_impl(m, orc_m5_neq(p->nmat), relax5, p, split, S, nout, root_o, level_o, ci_o, S_o, SI_o, split_o);
}
