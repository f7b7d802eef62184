/*
 * tc_main.c -- the solver executable, with the semantics of src/main.c:21-128 of the
 * reference: read step and time from the initial-condition directory, set up domain and
 * fluid, then integrate until timemax or wtimemax, logging, saving and collecting
 * statistics on their schedules, and save the final field.
 *
 * The fields stay on the device between host consumers (TC_COHERENCE=lazy unless the
 * environment says otherwise); log lines come from device reductions and checkpoints copy
 * the fields back only when they are written.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include "domain.h"
#include "fluid.h"
#include "fluid_solver.h"
#include "integrate.h"
#include "statistics.h"
#include "save.h"
#include "logging.h"
#include "config.h"
#include "fileio.h"

static double wall_time(void){
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char * argv[]){
  const double tic = wall_time();
  if(2 != argc){
    printf("directory name should be given as input\n");
    return 1;
  }
  const char * dirname_ic = argv[1];
  if(0 != fileio.init()) return 1;
  size_t step = 0;
  if(0 != fileio.r_serial(dirname_ic, "step", 0, NULL, fileio.npy_size_t, sizeof(size_t), &step)) return 1;
  double time = 0.;
  if(0 != fileio.r_serial(dirname_ic, "time", 0, NULL, fileio.npy_double, sizeof(double), &time)) return 1;
  setenv("TC_COHERENCE", "lazy", 0);
  domain_t domain = {0};
  if(0 != domain_init(dirname_ic, &domain)) return 1;
  fluid_t fluid = {0};
  if(0 != fluid_init(dirname_ic, &domain, &fluid)) return 1;
  if(0 != logging.init(&domain, time)) return 1;
  if(0 != save.init(&domain, time)) return 1;
  if(0 != statistics.init(&domain, time)) return 1;
  double timemax = 0.;
  if(0 != config.get_double("timemax", &timemax)) return 1;
  double wtimemax = 0.;
  if(0 != config.get_double("wtimemax", &wtimemax)) return 1;
  printf("step: %zu, time: % .7e\n", step, time);
  printf("timemax: % .7e, wtimemax: % .7e\n", timemax, wtimemax);
  fflush(stdout);
  int rc = 0;
  for(;;){
    double dt = 0.;
    if(0 != integrate(&domain, &fluid, &dt)){
      rc = 1;
      break;
    }
    step += 1;
    time += dt;
    const double toc = wall_time();
    if(timemax < time) break;
    if(wtimemax < toc - tic) break;
    if(logging.get_next_time() < time){
      logging.check_and_output(&domain, step, time, dt, toc - tic, &fluid);
    }
    if(save.get_next_time() < time){
      save.output(&domain, step, time, &fluid);
    }
    if(statistics.get_next_time() < time){
      statistics.collect(&domain, &fluid);
    }
  }
  if(0 == rc){
    save.output(&domain, step, time, &fluid);
    statistics.output(&domain, step);
  }
  tc_fluid_finalise(&fluid);
  return rc;
}
