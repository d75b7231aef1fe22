/*
 * vof_driver.c -- step-count driver over the library's C ABI: the
 * reference's src/main.c:41-151 loop with a step count instead of the
 * timemax / wtimemax stop, and its save layout (src/save.c:72-101,
 * main.c:17-32): <out>/step%010d/ holding step, time, domain, fluid and
 * interface NPY files, each a valid initial-condition directory.
 *
 * usage: vofdrive <ic_dir> <out_dir> <nsteps> [save_every]
 * Parameters come from the environment as in the reference (Ra Pr We
 * coef_dt_adv coef_dt_dif).  Built per dimensionality:
 *   gcc -std=c99 -DNDIMS=3 -Iinclude vof_driver.c -Lsimplevofsolver_amd -lvofhip3d
 */
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
  for (long n = 0; n < nsteps; n++) {
    double dt = 0.;
    if (integrate(&domain, &fluid, &interface, &dt)) {
      fprintf(stderr, "integrate failed: %s\n", vof_last_error());
      return 1;
    }
    step += 1;
    time += dt;
    if (save_every > 0 && (n + 1) % save_every == 0 && n + 1 < nsteps) {
      if (save_entrypoint(out, step, time, &domain, &fluid, &interface)) return 1;
    }
  }
  if (save_entrypoint(out, step, time, &domain, &fluid, &interface)) {
    fprintf(stderr, "save failed: %s\n", vof_last_error());
    return 1;
  }
  printf("step: %zu, time: % .7e\n", step, time);
  interface_finalise(&domain, &interface);
  fluid_finalise(&domain, &fluid);
  domain_finalise(&domain);
  return 0;
}
