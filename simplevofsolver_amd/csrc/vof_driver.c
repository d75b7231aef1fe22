/*
 * vof_driver.c -- step-count driver over the library's C ABI: the
 * reference's src/main.c:41-151 loop with a step count instead of the
 * timemax / wtimemax stop, and its save layout (src/save.c:72-101,
 * main.c:17-32): <out>/step%010d/ holding step, time, domain, fluid and
 * interface NPY files, each a valid initial-condition directory.
 *
 * usage: vofdrive <ic_dir> <out_dir> <nsteps> [save_every]
 * Parameters come from the environment as in the reference (Ra Pr We
 * coef_dt_adv coef_dt_dif).  With log_rate set, the logging module
 * (logging/main.c:82-104) writes <out>/log/{progress,divergence,momentum,
 * energy,nusselt,vof}.dat in the reference's formats from the device
 * reductions; with stat_rate (and stat_after) set, statistics are collected
 * (statistics.c) and written to <out>/stat/step%010zu at the end.
 * Built per dimensionality:
 *   gcc -std=c99 -DNDIMS=3 -Iinclude vof_driver.c -Lsimplevofsolver_amd -lvofhip3d
 */
#include <math.h>
#include <float.h>
#include <time.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <sys/stat.h>
#include <sys/types.h>
#include "vof_solver.h"

static int save_entrypoint(const char * out, size_t step, double time, const domain_t * domain,
                           const fluid_t * fluid, const interface_t * interface) {
  char dir[4096];
  snprintf(dir, sizeof dir, "%s/step%010zu", out, step);
  if (mkdir(dir, 0777) != 0 && errno != EEXIST) {
    perror(dir);
    return 1;
  }
  if (vof_write_step_time(dir, step, time)) return 1;
  if (domain_save(dir, domain)) return 1;
  if (fluid_save(dir, domain, fluid)) return 1;
  if (interface_save(dir, domain, interface)) return 1;
  return 0;
}

static double wall_time(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static FILE * open_log(const char * out, const char * name) {
  char path[4096];
  snprintf(path, sizeof path, "%s/log/%s", out, name);
  return fopen(path, "a");
}

/* logging/main.c:82-104 with the values of logging/{divergence,momentum,
 * energy,interface}.c and logging/nusselt/main.c from device reductions */
static int check_and_output(const char * out, const domain_t * domain, size_t step, double time, double dt,
                            double wtime, const fluid_t * fluid, const interface_t * interface) {
  double d[16];
  double nu[4];
  if (vof_diagnostics(domain, fluid, interface, d)) return 1;
  if (vof_nusselt(domain, fluid, nu)) return 1;
  /* d = {max|div|, vof min, vof max, volume, ke[NDIMS], thermal, mom[NDIMS], area} */
  const double * ke = d + 4;
  const double * mom = d + 4 + NDIMS + 1;
  FILE * fp;
  if ((fp = open_log(out, "progress.dat"))) {
    fprintf(fp, "step %zu, time %.1f, dt %.2e, elapsed %.1f [sec]\n", step, time, dt, wtime);
    fclose(fp);
  }
  printf("step %zu, time %.1f, dt %.2e, elapsed %.1f [sec]\n", step, time, dt, wtime);
  if ((fp = open_log(out, "divergence.dat"))) {
    fprintf(fp, "%8.2f % .1e\n", time, d[0]);
    fclose(fp);
  }
  if ((fp = open_log(out, "momentum.dat"))) {
    fprintf(fp, "%8.2f ", time);
    for (int n = 0; n < NDIMS; n++) fprintf(fp, "% 18.15e%c", mom[n], NDIMS - 1 == n ? '\n' : ' ');
    fclose(fp);
  }
  if ((fp = open_log(out, "energy.dat"))) {
    fprintf(fp, "%8.2f ", time);
    for (int n = 0; n <= NDIMS; n++) fprintf(fp, "% 18.15e%c", ke[n], NDIMS == n ? '\n' : ' ');
    fclose(fp);
  }
  if ((fp = open_log(out, "nusselt.dat"))) {
    fprintf(fp, "%8.2f ", time);
    for (int n = 0; n < 4; n++) fprintf(fp, "% 18.15e%c", nu[n], 3 == n ? '\n' : ' ');
    fclose(fp);
  }
  if ((fp = open_log(out, "vof.dat"))) {
    fprintf(fp, "%8.2f ", time);
    fprintf(fp, "% 18.15e ", fabs(d[1]) < 1.e-100 ? 1.e-99 : d[1]);
    fprintf(fp, "% 18.15e ", d[2]);
    fprintf(fp, "% 18.15e\n", d[3]);
    fclose(fp);
  }
  return 0;
}

int main(int argc, char * argv[]) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s <ic_dir> <out_dir> <nsteps> [save_every]\n", argv[0]);
    return 2;
  }
  const char * ic = argv[1];
  const char * out = argv[2];
  const long nsteps = strtol(argv[3], NULL, 10);
  const long save_every = argc > 4 ? strtol(argv[4], NULL, 10) : 0;
  size_t step = 0;
  double time = 0.;
  domain_t domain;
  fluid_t fluid;
  interface_t interface;
  memset(&domain, 0, sizeof domain);
  memset(&fluid, 0, sizeof fluid);
  memset(&interface, 0, sizeof interface);
  if (vof_read_step_time(ic, &step, &time) || domain_init(ic, &domain) ||
      fluid_init(ic, &domain, &fluid) || interface_init(ic, &domain, &interface)) {
    fprintf(stderr, "init failed: %s\n", vof_last_error());
    return 1;
  }
  if (mkdir(out, 0777) != 0 && errno != EEXIST) {
    perror(out);
    return 1;
  }
  printf("step: %zu, time: % .7e\n", step, time);
  /* logging.init (logging/main.c:20-38) */
  const char * log_rate_s = getenv("log_rate");
  const double log_rate = log_rate_s ? strtod(log_rate_s, NULL) : 0.;
  double log_next = log_rate > 0. ? log_rate * ceil(fmax(DBL_EPSILON, time) / log_rate) : 0.;
  if (log_rate > 0.) {
    char dir[4096];
    snprintf(dir, sizeof dir, "%s/log", out);
    if (mkdir(dir, 0777) != 0 && errno != EEXIST) {
      perror(dir);
      return 1;
    }
  }
  const int stats_on = getenv("stat_rate") != NULL;
  if (stats_on && statistics.init(&domain, time)) {
    fprintf(stderr, "statistics.init failed: %s\n", vof_last_error());
    return 1;
  }
  const double tic = wall_time();
  for (long n = 0; n < nsteps; n++) {
    double dt = 0.;
    if (integrate(&domain, &fluid, &interface, &dt)) {
      fprintf(stderr, "integrate failed: %s\n", vof_last_error());
      return 1;
    }
    step += 1;
    time += dt;
    if (log_rate > 0. && log_next < time) {
      if (check_and_output(out, &domain, step, time, dt, wall_time() - tic, &fluid, &interface)) {
        fprintf(stderr, "logging failed: %s\n", vof_last_error());
        return 1;
      }
      log_next += log_rate;
    }
    if (stats_on && statistics.get_next_time() < time) {
      if (statistics.collect(&domain, &fluid)) {
        fprintf(stderr, "statistics.collect failed: %s\n", vof_last_error());
        return 1;
      }
    }
    if (save_every > 0 && (n + 1) % save_every == 0 && n + 1 < nsteps) {
      if (save_entrypoint(out, step, time, &domain, &fluid, &interface)) return 1;
    }
  }
  if (save_entrypoint(out, step, time, &domain, &fluid, &interface)) {
    fprintf(stderr, "save failed: %s\n", vof_last_error());
    return 1;
  }
  if (stats_on) {
    char dir[4096];
    snprintf(dir, sizeof dir, "%s/stat/step%010zu", out, step);
    if (vof_statistics_output_to(&domain, dir)) {
      fprintf(stderr, "statistics output failed: %s\n", vof_last_error());
      return 1;
    }
  }
  printf("step: %zu, time: % .7e\n", step, time);
  interface_finalise(&domain, &interface);
  fluid_finalise(&domain, &fluid);
  domain_finalise(&domain);
  return 0;
}
