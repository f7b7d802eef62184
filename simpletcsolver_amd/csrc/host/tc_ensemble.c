/*
 * tc_ensemble.c -- a batch of independent flows of uneven sizes advanced together.
 *
 * The reference runs one flow per process.  Here n solvers are spread over a few HIP streams
 * (lanes); after one eager step, each RK substep of the whole batch (dt and factor tables in
 * the first, then predict, Poisson solve and correction of every member) is recorded once
 * into a hipGraph whose lanes are independent branches, so a step of the batch is three graph
 * launches whatever n is, and members of different lanes may run concurrently.  The members' double buffers
 * alternate from step to step, so the substeps of two consecutive steps are recorded (six
 * graphs) and replayed by step parity; the members' host-side bookkeeping advances with each
 * replayed step exactly as an eager step would advance it.
 *
 * The graphs hold the members' buffers, launch shapes, closure network and parameters as they
 * were when recorded.  Every member carries a generation counter that any change the graphs
 * cannot see bumps (tc_internal.h); a step finds a member whose counter moved since the
 * recording, drops the graphs and records again, so replaced fields, a closure coupled,
 * uncoupled or destroyed, new shapes or a member stepped on its own are always taken up.
 *
 * Launch shapes: tc_ensemble_autotune times the predictor of every member over the candidate
 * tiles and keeps the fastest, per member (members of the same shape reuse the first result).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime_api.h>
#include "tcstep.h"
#include "tc_internal.h"

#define TC_ENS_LANES 4   /* streams the members are spread over (hardware queues per process) */

struct tc_ensemble_s {
  int n;
  tc_solver_t ** s;
  hipStream_t stream;         /* root of the recorded graphs and their launches */
  int nlanes;
  hipStream_t lane[TC_ENS_LANES];   /* member m works on lane m % nlanes */
  hipEvent_t fork, join[TC_ENS_LANES];
  int device;
  int recorded;
  long long steps;            /* steps taken since the recording (parity of the replay) */
  unsigned long long * gen;   /* members' generation counters when the graphs were recorded */
  hipGraph_t graph[2][3];
  hipGraphExec_t exec[2][3];
};

int tc_ensemble_create(int n, const tc_params_t * params, const double * const * xf,
    const double * const * xc, tc_ensemble_t ** out){
  *out = NULL;
  if(n < 1) return tc_fail("tc_ensemble_create: n = %d", n);
  tc_ensemble_t * e = calloc(1, sizeof(*e));
  if(!e) return tc_fail("host allocation failed");
  e->s = calloc((size_t)n, sizeof(tc_solver_t *));
  e->gen = calloc((size_t)n, sizeof(unsigned long long));
  if(!e->s || !e->gen){ free(e->s); free(e->gen); free(e); return tc_fail("host allocation failed"); }
  e->device = params[0].device;
  if(hipSetDevice(e->device) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess){
    free(e->s);
    free(e->gen);
    free(e);
    return tc_fail("tc_ensemble_create: stream creation failed");
  }
  e->nlanes = n < TC_ENS_LANES ? n : TC_ENS_LANES;
  int bad = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess;
  for(int l = 0; l < e->nlanes && !bad; l++){
    bad = hipStreamCreateWithFlags(&e->lane[l], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&e->join[l], hipEventDisableTiming) != hipSuccess;
  }
  if(bad){
    tc_ensemble_destroy(e);
    return tc_fail("tc_ensemble_create: stream or event creation failed");
  }
  for(int m = 0; m < n; m++){
    if(params[m].device != e->device){
      tc_ensemble_destroy(e);
      return tc_fail("tc_ensemble_create: all members must use device %d", e->device);
    }
    if(tc_create(&params[m], xf[m], xc[m], &e->s[m]) || tc_share_stream(e->s[m], e->lane[m % e->nlanes])){
      tc_ensemble_destroy(e);
      return 1;
    }
    e->n = m + 1;
  }
  *out = e;
  return 0;
}

tc_solver_t * tc_ensemble_member(tc_ensemble_t * e, int m){
  if(!e || m < 0 || m >= e->n) return NULL;
  return e->s[m];
}

static void drop_graphs(tc_ensemble_t * e){
  for(int p = 0; p < 2; p++){
    for(int r = 0; r < 3; r++){
      if(e->exec[p][r]) hipGraphExecDestroy(e->exec[p][r]);
      if(e->graph[p][r]) hipGraphDestroy(e->graph[p][r]);
      e->exec[p][r] = NULL;
      e->graph[p][r] = NULL;
    }
  }
  e->recorded = 0;
}

static int sync_all(tc_ensemble_t * e){
  if(e->stream) HIPOK(hipStreamSynchronize(e->stream));
  for(int l = 0; l < e->nlanes; l++) if(e->lane[l]) HIPOK(hipStreamSynchronize(e->lane[l]));
  return 0;
}

int tc_ensemble_destroy(tc_ensemble_t * e){
  if(!e) return 0;
  hipSetDevice(e->device);
  sync_all(e);
  drop_graphs(e);
  for(int m = 0; m < e->n; m++) tc_destroy(e->s[m]);
  for(int l = 0; l < e->nlanes; l++){
    if(e->lane[l]) hipStreamDestroy(e->lane[l]);
    if(e->join[l]) hipEventDestroy(e->join[l]);
  }
  if(e->fork) hipEventDestroy(e->fork);
  if(e->stream) hipStreamDestroy(e->stream);
  free(e->s);
  free(e->gen);
  free(e);
  return 0;
}

int tc_ensemble_autotune(tc_ensemble_t * e, int reps, int * kb, int * ty){
  HIPOK(hipSetDevice(e->device));
  for(int m = 0; m < e->n; m++){
    const tc_params_t * p = &e->s[m]->p;
    int done = 0;
    for(int q = 0; q < m && !done; q++){
      const tc_params_t * pq = &e->s[q]->p;
      if(pq->nx == p->nx && pq->ny == p->ny && pq->nz == p->nz && pq->precision == p->precision &&
         pq->scalar == p->scalar && pq->implicit[0] == p->implicit[0] && pq->implicit[1] == p->implicit[1] &&
         pq->implicit[2] == p->implicit[2]){
        memcpy(e->s[m]->tune, e->s[q]->tune, sizeof(e->s[m]->tune));
        done = 1;
      }
    }
    if(!done && tc_autotune(e->s[m], reps, NULL, NULL, NULL)) return 1;
    if(kb) kb[m] = e->s[m]->tune[TC_TUNE_PREDICT][0];
    if(ty) ty[m] = e->s[m]->tune[TC_TUNE_PREDICT][1];
  }
  /* recorded graphs hold the old shapes */
  if(e->recorded){
    if(sync_all(e)) return 1;
    drop_graphs(e);
  }
  return 0;
}

/* host-side state of a member that enqueueing a substep changes (buffer parity, CFL state):
 * saved before a recording so that a failed capture, which ran nothing, leaves the member's
 * bookkeeping matching its buffers */
typedef struct { int cur[TC_NFIELD]; int dt_valid; } member_state_t;

/* one eager step (so that every member's CFL state is the steady one), then the substeps of
 * two steps recorded; nothing runs while recording, and the members' host state returns to
 * where it was after the two recorded steps (even parity) */
static int record(tc_ensemble_t * e){
  if(sync_all(e)) return 1;
  drop_graphs(e);
  for(int m = 0; m < e->n; m++){
    if(tc_enqueue_step(e->s[m])) return 1;
  }
  if(sync_all(e)) return 1;
  member_state_t * saved = malloc(sizeof(member_state_t) * (size_t)e->n);
  if(!saved) return tc_fail("host allocation failed");
  for(int m = 0; m < e->n; m++){
    memcpy(saved[m].cur, e->s[m]->cur, sizeof(saved[m].cur));
    saved[m].dt_valid = e->s[m]->dt_valid;
  }
  int rc = 0;
  for(int p = 0; p < 2 && !rc; p++){
    for(int r = 0; r < 3 && !rc; r++){
      /* the lanes fork from the root stream and join it again: the members' kernels form
       * independent branches of the graph */
      hipError_t ce = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
      if(ce != hipSuccess){
        rc = tc_fail("stream capture could not begin: %s", hipGetErrorString(ce));
        break;
      }
      int bad = hipEventRecord(e->fork, e->stream) != hipSuccess;
      for(int l = 0; l < e->nlanes && !bad; l++) bad = hipStreamWaitEvent(e->lane[l], e->fork, 0) != hipSuccess;
      for(int m = 0; m < e->n && !bad; m++) bad = tc_enqueue_substep(e->s[m], r);
      for(int l = 0; l < e->nlanes && !bad; l++){
        bad = hipEventRecord(e->join[l], e->lane[l]) != hipSuccess ||
              hipStreamWaitEvent(e->stream, e->join[l], 0) != hipSuccess;
      }
      hipGraph_t g = NULL;
      ce = hipStreamEndCapture(e->stream, &g);
      if(bad || ce != hipSuccess){
        if(g) hipGraphDestroy(g);
        rc = bad ? 1 : tc_fail("stream capture failed: %s", hipGetErrorString(ce));
        break;
      }
      e->graph[p][r] = g;
      ce = hipGraphInstantiate(&e->exec[p][r], g, NULL, NULL, 0);
      if(ce != hipSuccess){
        e->exec[p][r] = NULL;
        rc = tc_fail("graph instantiation failed: %s", hipGetErrorString(ce));
      }
    }
  }
  if(rc){
    /* nothing recorded ran: the members' bookkeeping goes back to the eager step's end */
    for(int m = 0; m < e->n; m++){
      memcpy(e->s[m]->cur, saved[m].cur, sizeof(saved[m].cur));
      e->s[m]->dt_valid = saved[m].dt_valid;
    }
    free(saved);
    drop_graphs(e);
    return 1;
  }
  free(saved);
  for(int m = 0; m < e->n; m++) e->gen[m] = e->s[m]->gen;
  e->recorded = 1;
  e->steps = 0;
  return 0;
}

/* a member changed since the recording in a way the graphs cannot see */
static int stale(const tc_ensemble_t * e){
  for(int m = 0; m < e->n; m++) if(e->s[m]->gen != e->gen[m]) return 1;
  return 0;
}

int tc_ensemble_step(tc_ensemble_t * e, int nsteps){
  HIPOK(hipSetDevice(e->device));
  if(nsteps < 1) return 0;
  int left = nsteps;
  if(e->recorded && stale(e)){
    if(sync_all(e)) return 1;
    drop_graphs(e);
  }
  if(!e->recorded){
    for(int m = 0; m < e->n; m++){
      if(e->s[m]->prof_on || e->s[m]->diag_armed || e->s[m]->stat_armed)
        return tc_fail("tc_ensemble_step: members may not profile or record diagnostics");
    }
    if(record(e)) return 1;   /* includes one eager step */
    left -= 1;
  }
  if(left < 1) return 0;
  /* the graphs run after whatever the members' lanes hold, and the lanes continue after the
   * graphs (reads of a member's fields are then ordered after the replayed steps) */
  for(int l = 0; l < e->nlanes; l++){
    HIPOK(hipEventRecord(e->join[l], e->lane[l]));
    HIPOK(hipStreamWaitEvent(e->stream, e->join[l], 0));
  }
  for(int q = 0; q < left; q++){
    const int p = (int)(e->steps & 1);
    for(int r = 0; r < 3; r++) HIPOK(hipGraphLaunch(e->exec[p][r], e->stream));
    for(int m = 0; m < e->n; m++) tc_advance_host_state(e->s[m]);
    e->steps += 1;
  }
  HIPOK(hipEventRecord(e->fork, e->stream));
  for(int l = 0; l < e->nlanes; l++) HIPOK(hipStreamWaitEvent(e->lane[l], e->fork, 0));
  return 0;
}

int tc_ensemble_synchronize(tc_ensemble_t * e){
  HIPOK(hipSetDevice(e->device));
  return sync_all(e);
}

/* drop the recorded graphs; the next step records again.  Changes made through the library
 * (fields, closure, shapes, solo steps) are detected without it */
int tc_ensemble_invalidate(tc_ensemble_t * e){
  HIPOK(hipSetDevice(e->device));
  if(sync_all(e)) return 1;
  drop_graphs(e);
  return 0;
}
