/* fv_run.c -- C driver of the device path: the reference's time loop
 * (loop.c:12-30: advancetimeexplicit every step, meshadaptation every
 * ADAPTHSTEP steps, geometric pre-adaptation as in main.c:107-134) written
 * against the C ABI of forestfv.h / forestfv_dev.h, with no Python.
 *
 *   fv_run bubble <base> <max> <layers> <steps> <every> [prec] [bc] [regrid]
 *       air bubble (R = 0.4 mm, 1 bar) in water (50 bar) in a 4 mm cube,
 *       octree levels base..max on the pressure-jump and interface-layer
 *       indicator, regrid before every <every>-th step, interface layers
 *       after every step, CFL 0.4 (dt fixed between regrids), SSP-RK3, PRELAX 1;
 *       prec 0 plain FP64 (default), 1 mixed, 2 bit-identical; bc the code of
 *       all six sides (default 2, outlet; 4 closes the box); regrid 1 (default)
 *       adapts on the device (fv_dev_regrid), 0 through the host mesh
 *   fv_run collapse <base> <max> <layers> <every> <prec> <maxsteps> [fixed_dt] [out]
 *       the bubble to its first collapse (see run_collapse)
 *   fv_run sod <level> <steps> [prec] [out]
 *       Sod shock tube (config C0) on a uniform octree, SSP-RK3, PRELAX 0,
 *       CFL 0.5, closed box; with <out> the final state is written there as
 *       int32 n, int32 NEQ and n x NEQ doubles in leaf order
 *
 * The last line reports the relative drift of the partial masses and of the
 * total energy (volume-weighted sums over the leaves).
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "forestfv.h"
#include "forestfv_dev.h"

#define CHECK(x)                                                                 \
  do {                                                                           \
    int rc_ = (x);                                                               \
    if (rc_ != FV_OK) {                                                          \
      fprintf(stderr, "fv_run: %s -> %d (%s)\n", #x, rc_, fv_dev_last_error()); \
      exit(1);                                                                   \
    }                                                                            \
  } while (0)

static double now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* volume-weighted sum of field v over the leaves */
static long double total(const fv_mesh *m, const fv_geo *g, const double *S, int neq, int v) {
  long double s = 0.0L;
  for (int i = 0; i < m->n; i++) s += (long double)S[(size_t)i * neq + v] * (long double)g->vol[m->level[i]];
  return s;
}

static void report_levels(const char *what, const fv_mesh *m) {
  int cnt[FV_MAXLEVEL + 2] = {0};
  for (int i = 0; i < m->n; i++) cnt[m->level[i]]++;
  printf("%s: %d leaves, per level:", what, m->n);
  for (int l = 1; l <= FV_MAXLEVEL; l++)
    if (cnt[l]) printf(" L%d=%d", l, cnt[l]);
  printf("\n");
  fflush(stdout);
}

/* indicator on the device, 2:1 smoothing and merge/split on the host, new image */
static void regrid(fv_dev *dev, fv_mesh **m, const fv_params *p, const fv_crit *cr, int maxlevel, double **S) {
  const int n = (*m)->n;
  int *SI = malloc(sizeof(int) * 2 * (size_t)n), *split = malloc(sizeof(int) * (size_t)n);
  if (!SI || !split) { fprintf(stderr, "fv_run: out of memory\n"); exit(1); }
  CHECK(fv_dev_indicator(dev, cr, SI, split));
  memcpy((*m)->SI, SI, sizeof(int) * 2 * (size_t)n);
  CHECK(fv_criterionsmth(*m, maxlevel, split));
  CHECK(fv_mesh_adapt(m, p, split, S));
  free(SI);
  free(split);
  fv_pack pk;
  CHECK(fv_mesh_pack(*m, &pk));
  CHECK(fv_dev_reattach(dev, &pk));
  fv_pack_free(&pk);
  CHECK(fv_dev_set_state(dev, *S));
}

/* the host forest of the device mesh (after device regrids) */
static fv_mesh *forest_of(fv_dev *dev, const fv_mesh *like) {
  int n = 0, neq = 0;
  CHECK(fv_dev_size(dev, &n, &neq));
  int *root = malloc(sizeof(int) * (size_t)n), *level = malloc(sizeof(int) * (size_t)n);
  int *ci = malloc(sizeof(int) * 3 * (size_t)n), *SI = malloc(sizeof(int) * 2 * (size_t)n);
  if (!root || !level || !ci || !SI) { fprintf(stderr, "fv_run: out of memory\n"); exit(1); }
  CHECK(fv_dev_get_forest(dev, root, level, ci, SI));
  int st = 0;
  fv_mesh *m = fv_mesh_from_leaves(like->nroot, like->origin, like->rootsize, like->bcside, n, root, level, ci, &st);
  if (!m) { fprintf(stderr, "fv_run: fv_mesh_from_leaves -> %d\n", st); exit(1); }
  memcpy(m->SI, SI, sizeof(int) * 2 * (size_t)n);
  free(root); free(level); free(ci); free(SI);
  return m;
}

typedef struct {
  fv_params p;
  fv_crit cr;
  fv_mesh *m;
  double *S;
  fv_dev *dev;
  double L, R;
} bubble_t;

/* the water/air bubble of the configs on a forest pre-adapted to levels
 * base..maxlevel (main.c:107-134: initial data, indicator, smoothing, refine) */
static void setup_bubble(bubble_t *B, int base, int maxlevel, int layers, int prec, int bc) {
  const double L = 4e-3, R = 0.4e-3;
  fv_params p;
  memset(&p, 0, sizeof p);
  p.nmat = 2;
  p.gamma[0] = 6.12; p.pinf[0] = 3.43e8;   /* water */
  p.gamma[1] = 1.4;  p.pinf[1] = 0.0;      /* air */
  p.amin = 1e-6; p.avf_lim = 0.5; p.prelax = 1; p.rk = FV_RK_SSP3; p.limiter = 0; p.dt = 1e-9;
  p.fix = FV_FIX_ALL; p.primtv = 0;
  for (int f = 0; f < 6; f++) p.bc[f] = bc;
  fv_crit cr;
  memset(&cr, 0, sizeof cr);
  cr.uniform_on = 1; cr.uniform_lvl = base;
  cr.prs_on = 1; cr.prs_lvl = maxlevel; cr.prs_thr = 0.2;
  cr.inter0_on = 1; cr.inter0_lvl = maxlevel; cr.inter0_lrs = layers;
  const double rho0[2] = {1000.0, 1.2}, p0[2] = {5e6, 1e5}, c[3] = {0.5 * L, 0.5 * L, 0.5 * L};
  const int nroot[3] = {1, 1, 1};
  const double origin[3] = {0.0, 0.0, 0.0};
  const int neq = fv_neq(2);
  fv_mesh *m = fv_mesh_uniform(nroot, origin, L, base, p.bc);
  if (!m) { fprintf(stderr, "fv_run: mesh\n"); exit(1); }
  double *S = malloc(sizeof(double) * (size_t)m->n * neq);
  fv_dev *dev = NULL;
  for (int pass = 0; pass <= maxlevel - base; pass++) {
    CHECK(fv_init_bubble(m, &p, rho0, p0, c, R, 2.0, 1234u, 0.01, S));
    fv_pack pk;
    CHECK(fv_mesh_pack(m, &pk));
    if (!dev) CHECK(fv_dev_attach(&dev, &p, &pk));
    else CHECK(fv_dev_reattach(dev, &pk));
    fv_pack_free(&pk);
    CHECK(fv_dev_set_precision(dev, prec));
    CHECK(fv_dev_set_state(dev, S));
    regrid(dev, &m, &p, &cr, maxlevel, &S);
  }
  CHECK(fv_init_bubble(m, &p, rho0, p0, c, R, 2.0, 1234u, 0.01, S));
  CHECK(fv_dev_set_state(dev, S));
  CHECK(fv_dev_interface_location(dev, &cr, NULL));
  B->p = p; B->cr = cr; B->m = m; B->S = S; B->dev = dev; B->L = L; B->R = R;
}

static int run_bubble(int base, int maxlevel, int layers, int steps, int every, int prec, int bc, int dev_regrid) {
  const int neq = fv_neq(2);
  double t0 = now();
  bubble_t B;
  setup_bubble(&B, base, maxlevel, layers, prec, bc);
  fv_params p = B.p;
  fv_crit cr = B.cr;
  fv_mesh *m = B.m;
  double *S = B.S;
  fv_dev *dev = B.dev;
  const double L = B.L;
  if (dev_regrid) CHECK(fv_dev_set_forest(dev, m));
  report_levels("initial mesh", m);
  printf("setup %.2f s\n", now() - t0);

  fv_geo g;
  fv_geo_build(L, &g);
  long double m0 = total(m, &g, S, neq, 0), m1 = total(m, &g, S, neq, 1), e0 = total(m, &g, S, neq, 5);
  double t_step = 0.0, t_regrid = 0.0, simt = 0.0;
  long long cell_stages = 0;
  int nleaves = m->n, nregrid = 0;
  for (int step = 0; step < steps; step++) {
    /* loop.c:19: adaptation before the step when (istep+1) % ADAPTHSTEP == 0 */
    const int adapt = (step + 1) % every == 0;
    if (adapt) {
      double ta = now();
      if (dev_regrid) {
        int nn = 0;
        CHECK(fv_dev_regrid(dev, &cr, maxlevel, &nn));
        nleaves = nn;
      } else {
        CHECK(fv_dev_get_state(dev, S));
        regrid(dev, &m, &p, &cr, maxlevel, &S);
        nleaves = m->n;
      }
      t_regrid += now() - ta;
      nregrid++;
    }
    double ta = now(), dt;
    if (step == 0 || adapt) {
      CHECK(fv_dev_cfl_dt(dev, 0.4, &dt));
      CHECK(fv_dev_set_dt(dev, dt));
      p.dt = dt;
    }
    CHECK(fv_dev_advance(dev, 1));
    /* loop.c:23: interface_location after every step */
    CHECK(fv_dev_interface_location(dev, &cr, NULL));
    t_step += now() - ta;
    simt += p.dt;
    cell_stages += 3LL * nleaves;
  }
  if (dev_regrid) {
    fv_mesh *q = forest_of(dev, m);
    fv_mesh_free(m);
    m = q;
    S = realloc(S, sizeof(double) * (size_t)m->n * neq);
    if (!S) { fprintf(stderr, "fv_run: out of memory\n"); return 1; }
  }
  CHECK(fv_dev_get_state(dev, S));
  report_levels("final mesh", m);
  long double m0f = total(m, &g, S, neq, 0), m1f = total(m, &g, S, neq, 1), e0f = total(m, &g, S, neq, 5);
  printf("steps %d, simulated %.3e s, stepping %.3f s (%.1f Mcell-updates/s per RK stage incl. host calls), "
         "%d %s regrids %.3f s (%.2f ms each), loop %.1f Mcell-updates/s\n",
         steps, simt, t_step, (double)cell_stages / t_step / 1e6, nregrid, dev_regrid ? "device" : "host", t_regrid,
         nregrid ? 1e3 * t_regrid / nregrid : 0.0, (double)cell_stages / (t_step + t_regrid) / 1e6);
  printf("fv_run bubble: drift water %.3Le air %.3Le energy %.3Le\n", fabsl(m0f - m0) / m0, fabsl(m1f - m1) / m1,
         fabsl(e0f - e0) / e0);
  fv_dev_detach(dev);
  fv_mesh_free(m);
  free(S);
  return 0;
}

/* Run the bubble to its first collapse: the loop of loop.c with a device regrid
 * every `every` steps (0: fixed mesh, no interface layers), the CFL step
 * recomputed every 5 steps (or dt fixed when fixed_dt > 0), and the bubble
 * volume V_domain - integral of alpha_water from fv_dev_totals every 2 steps.
 * The first collapse is the minimum of the volume, taken once it has fallen
 * below half its start and rebounded by 2 %.  One JSON line reports it against
 * the Rayleigh time 0.915 R0 sqrt(rho_l / (p_l - p_g)); with `out` the final
 * state is written as in the sod mode. */
static int run_collapse(int base, int maxlevel, int layers, int every, int prec, int maxsteps, double fixed_dt,
                        const char *out) {
  const int neq = fv_neq(2);
  double t0 = now();
  bubble_t B;
  setup_bubble(&B, base, maxlevel, layers, prec, 2);
  const double t_setup = now() - t0;
  fv_dev *dev = B.dev;
  if (every > 0) CHECK(fv_dev_set_forest(dev, B.m));
  const double Vdom = B.L * B.L * B.L;
  double tot[16];
  CHECK(fv_dev_totals(dev, tot));
  const double Vb0 = Vdom - tot[6], mw0 = tot[0], ma0 = tot[1];
  double Vmin = Vb0, t_min = 0.0, simt = 0.0, dt = fixed_dt;
  double t_step = 0.0, t_regrid = 0.0;
  long long cell_stages = 0;
  int n = B.m->n, nregrid = 0, step = 0, collapsed = 0, nmax = n, nmin = n;
  for (step = 0; step < maxsteps && !collapsed; step++) {
    const int adapt = every > 0 && (step + 1) % every == 0;
    if (adapt) {
      double ta = now();
      CHECK(fv_dev_regrid(dev, &B.cr, maxlevel, &n));
      t_regrid += now() - ta;
      nregrid++;
      if (n > nmax) nmax = n;
      if (n < nmin) nmin = n;
    }
    double ta = now();
    if (fixed_dt <= 0.0 && (step % 5 == 0 || adapt)) CHECK(fv_dev_cfl_dt(dev, 0.4, &dt));
    CHECK(fv_dev_set_dt(dev, dt));
    CHECK(fv_dev_advance(dev, 1));
    if (every > 0) CHECK(fv_dev_interface_location(dev, &B.cr, NULL));
    t_step += now() - ta;
    simt += dt;
    cell_stages += 3LL * n;
    if (step % 2 == 1) {
      CHECK(fv_dev_totals(dev, tot));
      const double Vb = Vdom - tot[6];
      if (Vb < Vmin) { Vmin = Vb; t_min = simt; }
      if (Vmin < 0.5 * Vb0 && Vb > 1.02 * Vmin) collapsed = 1;
    }
  }
  CHECK(fv_dev_totals(dev, tot));
  const double t_ray = 0.915 * B.R * sqrt(1000.0 / (5e6 - 1e5));
  printf("{\"mode\": \"collapse\", \"collapsed\": %s, \"steps\": %d, \"t_end_s\": %.6e, \"t_collapse_s\": %.6e, "
         "\"rayleigh_time_s\": %.6e, \"t_collapse_over_rayleigh\": %.4f, \"R_min_over_R0\": %.4f, "
         "\"leaves_start\": %d, \"leaves_min\": %d, \"leaves_max\": %d, \"leaves_end\": %d, \"regrids\": %d, "
         "\"regrid_ms\": %.3f, \"stepping_s\": %.3f, \"setup_s\": %.2f, "
         "\"stage_rate_Mcell_s\": %.1f, \"loop_rate_Mcell_s\": %.1f, \"water_mass_change\": %.3e, "
         "\"air_mass_change\": %.3e, \"precision\": %d}\n",
         collapsed ? "true" : "false", step, simt, t_min, t_ray, t_min / t_ray, cbrt(Vmin / Vb0), B.m->n, nmin, nmax, n,
         nregrid, nregrid ? 1e3 * t_regrid / nregrid : 0.0, t_step, t_setup, (double)cell_stages / t_step / 1e6,
         (double)cell_stages / (t_step + t_regrid) / 1e6, (tot[0] - mw0) / mw0, (tot[1] - ma0) / ma0, prec);
  fflush(stdout);
  if (out) {
    double *S = malloc(sizeof(double) * (size_t)n * neq);
    CHECK(fv_dev_get_state(dev, S));
    FILE *f = fopen(out, "wb");
    int hdr[2] = {n, neq};
    if (!S || !f || fwrite(hdr, sizeof(int), 2, f) != 2 || fwrite(S, sizeof(double), (size_t)n * neq, f) != (size_t)n * neq ||
        fclose(f) != 0) {
      fprintf(stderr, "fv_run: cannot write %s\n", out);
      return 1;
    }
    free(S);
  }
  fv_dev_detach(dev);
  fv_mesh_free(B.m);
  free(B.S);
  return 0;
}

static int run_sod(int level, int steps, int prec, const char *out) {
  fv_params p;
  memset(&p, 0, sizeof p);
  p.nmat = 2;
  p.gamma[0] = 1.4; p.gamma[1] = 1.4;
  p.amin = 1e-8; p.avf_lim = 0.5; p.prelax = 0; p.rk = FV_RK_SSP3; p.limiter = 0;
  p.dt = 0.5 * (1.0 / (1 << (level - 1))) / 2.0;
  p.fix = FV_FIX_ALL;
  for (int f = 0; f < 6; f++) p.bc[f] = 4;
  const int nroot[3] = {1, 1, 1};
  const double origin[3] = {0.0, 0.0, 0.0};
  const double rho[2] = {1.0, 0.125}, pr[2] = {1.0, 0.1};
  const int neq = fv_neq(2);
  fv_mesh *m = fv_mesh_uniform(nroot, origin, 1.0, level, p.bc);
  double *S = malloc(sizeof(double) * (size_t)m->n * neq);
  CHECK(fv_init_split_x(m, &p, rho, pr, 0.5, S));
  fv_pack pk;
  CHECK(fv_mesh_pack(m, &pk));
  fv_dev *dev = NULL;
  CHECK(fv_dev_attach(&dev, &p, &pk));
  CHECK(fv_dev_set_precision(dev, prec));
  fv_geo g = pk.geo;
  fv_pack_free(&pk);
  CHECK(fv_dev_set_state(dev, S));
  long double m0 = total(m, &g, S, neq, 0), m1 = total(m, &g, S, neq, 1), e0 = total(m, &g, S, neq, 5);
  double t0 = now();
  CHECK(fv_dev_advance(dev, steps));
  double t = now() - t0;
  CHECK(fv_dev_get_state(dev, S));
  long double m0f = total(m, &g, S, neq, 0), m1f = total(m, &g, S, neq, 1), e0f = total(m, &g, S, neq, 5);
  printf("sod: %d leaves, %d steps in %.3f s\n", m->n, steps, t);
  if (out) {
    FILE *f = fopen(out, "wb");
    int hdr[2] = {m->n, neq};
    if (!f || fwrite(hdr, sizeof(int), 2, f) != 2 ||
        fwrite(S, sizeof(double), (size_t)m->n * neq, f) != (size_t)m->n * neq || fclose(f) != 0) {
      fprintf(stderr, "fv_run: cannot write %s\n", out);
      return 1;
    }
  }
  printf("fv_run sod: drift left %.3Le right %.3Le energy %.3Le\n", fabsl(m0f - m0) / m0, fabsl(m1f - m1) / m1,
         fabsl(e0f - e0) / e0);
  fv_dev_detach(dev);
  fv_mesh_free(m);
  free(S);
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 7 && strcmp(argv[1], "bubble") == 0)
    return run_bubble(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]),
                      argc > 7 ? atoi(argv[7]) : FV_PREC_FP64, argc > 8 ? atoi(argv[8]) : 2,
                      argc > 9 ? atoi(argv[9]) : 1);
  if (argc >= 8 && strcmp(argv[1], "collapse") == 0)
    return run_collapse(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), atoi(argv[7]),
                        argc > 8 ? atof(argv[8]) : 0.0, argc > 9 ? argv[9] : NULL);
  if (argc >= 4 && strcmp(argv[1], "sod") == 0)
    return run_sod(atoi(argv[2]), atoi(argv[3]), argc > 4 ? atoi(argv[4]) : FV_PREC_FP64, argc > 5 ? argv[5] : NULL);
  fprintf(stderr,
          "usage: fv_run bubble <base> <max> <layers> <steps> <every> [prec] [bc] [regrid]\n"
          "       fv_run collapse <base> <max> <layers> <every> <prec> <maxsteps> [fixed_dt] [out]\n"
          "       fv_run sod <level> <steps> [prec] [out]\n");
  return 2;
}
