/*
 * ecogen_pool.h -- batched and served solves of the explicit finite-volume
 * step on one MI355X.
 *
 * The reference runs one solve per process (Run::solver, src/Run.cpp:204-277,
 * driven by main, src/main.cpp:60-118); it has no batching or serving layer,
 * so these entry points have no reference counterpart.  Each case admitted to
 * a pool is one independent reference solve (same configuration semantics as
 * eco_config, same time loop as eco_run), and its result is bit-identical to
 * the same case run alone through eco_create/eco_set_state/eco_run with the
 * same precision.
 *
 * A pool holds many cases of uneven size resident in HBM.  Cases that share a
 * kernel class (model, dimension, order, precision, limiter form) advance
 * together: one persistent launch per stage and class covers every tile of
 * every active case through per-case offset tables, and chunks of steps are
 * replayed from a captured hipGraph.  Cases finish at their own targets; a
 * finished case's tiles are dropped from the tables at the next poll.
 *
 * Device memory comes from one arena allocated at creation (hbm_budget
 * bytes); a case is admitted only when its buffers fit (ECO_ERR_BUDGET
 * otherwise), and its memory returns to the arena when it is released.
 *
 * Threading: a pool is driven from one host thread.  Pools are independent.
 */
#ifndef ECOGEN_POOL_H
#define ECOGEN_POOL_H

#include "ecogen.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ECO_ERR_BUDGET (-6) /* the case does not fit the pool's free HBM budget now */
#define ECO_ERR_FULL (-7)   /* the pool's case table is full */

typedef struct eco_pool eco_pool;

typedef struct {
  unsigned long long hbm_budget; /* bytes of device memory for case state (arena size) */
  int max_cases;                 /* capacity of the case table (0 = 8192) */
  int steps_per_chunk;           /* steps per graph replay, rounded up to a multiple of 6 (0 = 24) */
  int use_graph;                 /* 1: replay captured hipGraphs, 0: direct launches */
  int grid_blocks;               /* persistent workgroups per launch (0 = occupancy x CUs) */
} eco_pool_config;

typedef struct {
  int active;                    /* cases admitted and not released */
  int running;                   /* active cases that have not reached their target */
  unsigned long long bytes_used; /* arena bytes held by cases */
  unsigned long long bytes_free; /* largest admissible allocation right now */
  long long steps_launched;      /* pool steps enqueued since creation */
  long long tiles_launched;      /* case tiles processed by the stage kernels (stage-tile count) */
  int graph_captures;            /* hipGraph (re)captures */
} eco_pool_stats;

/* Creates a pool and its arena. */
int eco_pool_create(const eco_pool_config *pc, eco_pool **out);

/* Device bytes one case of this configuration holds in the arena. */
unsigned long long eco_pool_case_bytes(const eco_config *cfg);

/* Admits one solve: uploads its initial primitive state (host layout of
   ecogen.h, E ignored) and sets its target (nsteps > 0: exactly nsteps steps;
   nsteps <= 0: until t >= t_final compared as float, src/Run.cpp:239-244).
   The upload overlaps the pool's running work; the case joins at the next
   chunk.  case_id receives the case handle.  A configuration with transports
   (ntransport > 0) is refused with ECO_ERR_UNSUPPORTED: it runs in a context. */
int eco_pool_admit(eco_pool *p, const eco_config *cfg, const double *host_state, int nsteps, double t_final,
                   int *case_id);

/* Enqueues steps_per_chunk steps of every running case (no host wait). */
int eco_pool_advance(eco_pool *p);

/* Waits until at most `keep` chunks are still in flight and reports the
   cases that reached their target in the completed chunks (each case once).
   done_ids receives up to max_ids handles; ndone the count written. */
int eco_pool_poll(eco_pool *p, int keep, int *done_ids, int max_ids, int *ndone);

/* Downloads a case's current state (host layout, E included) and status.
   Valid for running and finished cases; synchronises the pool. */
int eco_pool_result(eco_pool *p, int case_id, double *host_state, eco_status *st);

/* Returns a finished case's memory to the arena (deferred until no enqueued
   work refers to it) and frees its handle. */
int eco_pool_release(eco_pool *p, int case_id);

int eco_pool_stats_get(eco_pool *p, eco_pool_stats *s);

/* Waits for all enqueued work and frees the pool. */
void eco_pool_destroy(eco_pool *p);

/* ---- serving loop -------------------------------------------------------
   Runs a request stream in real time on a pool: request i becomes visible
   at arrival_s[i] seconds after the call starts; visible requests are
   admitted in arrival order as the HBM budget allows (continuous batching:
   new cases join the running chunks), every case runs to its own target, and
   each finished case's state is written to out_states[i] (host layout, may be
   NULL) before its memory is released.  Times are seconds from the start of
   the call: admit_s[i] (admission enqueued), finish_s[i] (result on the
   host).  Returns ECO_OK when every request finished. */
typedef struct {
  const eco_config *cfg;     /* configuration of the request */
  const double *state;       /* initial primitive state (host layout) */
  int nsteps;                /* target as in eco_pool_admit */
  double t_final;
  double arrival_s;          /* arrival time, seconds from the start */
  double *out_state;         /* result buffer (host layout) or NULL */
} eco_request;

typedef struct {
  double admit_s, finish_s;  /* seconds from the start of eco_serve */
  int steps;                 /* steps the case took */
  double t;                  /* physical time reached */
  uint32_t flags;            /* ECO_WARN_* of the case */
} eco_request_result;

int eco_serve(eco_pool *p, const eco_request *req, int nreq, eco_request_result *res);

#ifdef __cplusplus
}
#endif
#endif
