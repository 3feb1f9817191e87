/* fv_io.c -- restart and field files of a forest and its state.
 *
 * Text: the layout of restart_save.c:12-45 and export_ultra_fast
 * (exportfield.c:45-97): a header "istep time tleaves ntrees NEQ", then one
 * line per leaf in DFS order, "root tag" and the NEQ state values at %15.10e.
 * restart_load.c:14-45 reads it back.  Text keeps 11 significant digits.
 *
 * Binary (new, lossless): magic "FVDUMP01", int32 istep, float64 time, int32
 * n, int32 ntrees, int32 NEQ, int32 root[n], int32 tag[n], float64 S[n][NEQ],
 * in the host's byte order.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "forestfv.h"
#include "forestfv_dev.h"

static const char MAGIC[8] = {'F', 'V', 'D', 'U', 'M', 'P', '0', '1'};

static int write_leaves(const char *path, int n, int ntrees, const int *root, const int *tag, const double *S, int neq,
                        int istep, double time, int binary) {
  FILE *f = fopen(path, binary ? "wb" : "w");
  if (!f) return FV_ERR_ARG;
  int ok = 1;
  if (binary) {
    int32_t hdr[4] = {n, ntrees, neq, 0};
    int32_t st = istep;
    ok = fwrite(MAGIC, 1, 8, f) == 8 && fwrite(&st, sizeof st, 1, f) == 1 && fwrite(&time, sizeof time, 1, f) == 1 &&
         fwrite(hdr, sizeof(int32_t), 3, f) == 3;
    ok = ok && fwrite(root, sizeof(int32_t), (size_t)n, f) == (size_t)n;
    ok = ok && fwrite(tag, sizeof(int32_t), (size_t)n, f) == (size_t)n;
    ok = ok && fwrite(S, sizeof(double), (size_t)n * neq, f) == (size_t)n * neq;
  } else {
    ok = fprintf(f, "%d %15.10e %d %d %d \n", istep, time, n, ntrees, neq) > 0;
    for (int i = 0; i < n && ok; i++) {
      ok = fprintf(f, "%d %d ", root[i], tag[i]) > 0;
      for (int v = 0; v < neq && ok; v++) ok = fprintf(f, " %15.10e ", S[(size_t)i * neq + v]) > 0;
      ok = ok && fputc('\n', f) != EOF;
    }
  }
  if (fclose(f) != 0) ok = 0;
  return ok ? FV_OK : FV_ERR_ARG;
}

int fv_restart_write(const char *path, const fv_mesh *m, const double *S, int neq, int istep, double time,
                     int binary) {
  if (!path || !m || !S || neq <= 0) return FV_ERR_ARG;
  return write_leaves(path, m->n, m->nroot[0] * m->nroot[1] * m->nroot[2], m->root, m->tag, S, neq, istep, time,
                      binary);
}

int fv_dev_write_restart(fv_dev *ctx, const char *path, int istep, double time, int binary) {
  if (!ctx || !path) return FV_ERR_ARG;
  int n = 0, neq = 0, ntrees = 0, rc;
  if ((rc = fv_dev_size(ctx, &n, &neq)) != FV_OK) return rc;
  if ((rc = fv_dev_get_ntrees(ctx, &ntrees)) != FV_OK) return rc;
  int *root = malloc(sizeof(int) * (size_t)n), *level = malloc(sizeof(int) * (size_t)n);
  int *ci = malloc(sizeof(int) * 3 * (size_t)n), *tag = malloc(sizeof(int) * (size_t)n);
  double *S = malloc(sizeof(double) * (size_t)n * neq);
  if (!root || !level || !ci || !tag || !S) rc = FV_ERR_NOMEM;
  if (rc == FV_OK) rc = fv_dev_get_forest(ctx, root, level, ci, NULL);
  if (rc == FV_OK) rc = fv_dev_get_state(ctx, S);
  if (rc == FV_OK) {
    for (int i = 0; i < n; i++) tag[i] = fv_tag_from_address(level[i], &ci[3 * i]);
    rc = write_leaves(path, n, ntrees, root, tag, S, neq, istep, time, binary);
  }
  free(root); free(level); free(ci); free(tag); free(S);
  return rc;
}

int fv_restart_read(const char *path, const int nroot[3], const double origin[3], double rootsize,
                    const int bcside[6], fv_mesh **pm, double **pS, int *pneq, int *pistep, double *ptime) {
  if (!path || !pm || !pS) return FV_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) return FV_ERR_ARG;
  char mg[8];
  const int binary = fread(mg, 1, 8, f) == 8 && memcmp(mg, MAGIC, 8) == 0;
  int istep = 0, n = 0, ntrees = 0, neq = 0, rc = FV_OK;
  double time = 0.0;
  int *root = NULL, *tag = NULL, *level = NULL, *ci = NULL;
  double *Sf = NULL, *S = NULL;
  fv_mesh *m = NULL;
  if (binary) {
    int32_t st, hdr[3];
    if (fread(&st, sizeof st, 1, f) != 1 || fread(&time, sizeof time, 1, f) != 1 || fread(hdr, sizeof(int32_t), 3, f) != 3) {
      rc = FV_ERR_ARG;
      goto done;
    }
    istep = st; n = hdr[0]; ntrees = hdr[1]; neq = hdr[2];
  } else {
    rewind(f);
    if (fscanf(f, "%d %lf %d %d %d", &istep, &time, &n, &ntrees, &neq) != 5) { rc = FV_ERR_ARG; goto done; }
  }
  if (n <= 0 || neq <= 0 || ntrees != nroot[0] * nroot[1] * nroot[2]) { rc = FV_ERR_ARG; goto done; }
  root = malloc(sizeof(int) * (size_t)n);
  tag = malloc(sizeof(int) * (size_t)n);
  level = malloc(sizeof(int) * (size_t)n);
  ci = malloc(sizeof(int) * 3 * (size_t)n);
  Sf = malloc(sizeof(double) * (size_t)n * neq);
  S = malloc(sizeof(double) * (size_t)n * neq);
  if (!root || !tag || !level || !ci || !Sf || !S) { rc = FV_ERR_NOMEM; goto done; }
  if (binary) {
    if (fread(root, sizeof(int32_t), (size_t)n, f) != (size_t)n || fread(tag, sizeof(int32_t), (size_t)n, f) != (size_t)n ||
        fread(Sf, sizeof(double), (size_t)n * neq, f) != (size_t)n * neq) {
      rc = FV_ERR_ARG;
      goto done;
    }
  } else {
    for (int i = 0; i < n; i++) {
      if (fscanf(f, "%d %d", &root[i], &tag[i]) != 2) { rc = FV_ERR_ARG; goto done; }
      for (int v = 0; v < neq; v++)
        if (fscanf(f, "%lf", &Sf[(size_t)i * neq + v]) != 1) { rc = FV_ERR_ARG; goto done; }
    }
  }
  for (int i = 0; i < n; i++) {
    if (root[i] < 0 || root[i] >= ntrees || fv_address_from_tag(tag[i], &level[i], &ci[3 * i]) != FV_OK) {
      rc = FV_ERR_ARG;
      goto done;
    }
  }
  m = fv_mesh_from_leaves(nroot, origin, rootsize, bcside, n, root, level, ci, &rc);
  if (!m) goto done;
  /* the forest is in DFS order; place each file row at its leaf */
  for (int i = 0; i < n; i++) {
    int j = fv_mesh_find(m, root[i], level[i], ci[3 * i], ci[3 * i + 1], ci[3 * i + 2]);
    if (j < 0) { rc = FV_ERR_MESH; goto done; }
    memcpy(&S[(size_t)j * neq], &Sf[(size_t)i * neq], sizeof(double) * neq);
  }
  *pm = m; m = NULL;
  *pS = S; S = NULL;
  if (pneq) *pneq = neq;
  if (pistep) *pistep = istep;
  if (ptime) *ptime = time;
done:
  fclose(f);
  free(root); free(tag); free(level); free(ci); free(Sf); free(S);
  if (m) fv_mesh_free(m);
  return rc;
}
