/*
 * Request serving over the C ABI: a stream of independent 3-D elastic solves
 * (each a periodic cube brick with its own order, layered material, Ricker
 * source and step count) is admitted under an HBM budget as the requests
 * arrive, kept resident on the device, and advanced together: every launch
 * of the batch advances all admitted requests by one LSRK stage (continuous
 * batching).  The batch is rebuilt whenever a request joins or leaves, at
 * chunk boundaries of a few steps.
 *
 * The reference has no serving layer; this sits on top of the same entry
 * points an application uses (bfamx_subdomain_upload, bfamx_batch_*), the
 * analogue of bfamo_device_malloc's budget check (bfam.c:8561-8589) applied
 * to whole requests.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bfamx_internal.h"

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void sleep_s(double s)
{
  if (s <= 0)
    return;
  struct timespec ts;
  ts.tv_sec = (time_t)s;
  ts.tv_nsec = (long)((s - (double)ts.tv_sec) * 1e9);
  nanosleep(&ts, NULL);
}

int bfamx_serve_generate(uint64_t seed, int nreq, double rate_per_s,
                         int n_min, int n_max, int N_min, int N_max,
                         int steps_min, int steps_max, bfamx_request_t *reqs)
{
  if (nreq < 0 || !reqs || rate_per_s <= 0 || n_min < 1 || n_max < n_min ||
      N_min < 3 || N_max > 7 || N_max < N_min || steps_min < 1 ||
      steps_max < steps_min)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_serve_generate: bad arguments");
  bfamx_pcg32_t g;
  bfamx_pcg32_seed(&g, seed, 91);
  double t = 0;
  for (int i = 0; i < nreq; ++i)
  {
    /* exponential inter-arrival times: a Poisson stream of the given rate */
    const double u = bfamx_pcg32_uniform(&g);
    t += -log(1.0 - u) / rate_per_s;
    reqs[i].arrival = t;
    reqs[i].n = n_min + (int)bfamx_pcg32_bounded(&g, (uint32_t)(n_max - n_min + 1));
    reqs[i].N = N_min + (int)bfamx_pcg32_bounded(&g, (uint32_t)(N_max - N_min + 1));
    reqs[i].nsteps =
        steps_min + (int)bfamx_pcg32_bounded(&g, (uint32_t)(steps_max - steps_min + 1));
    reqs[i].seed = ((uint64_t)bfamx_pcg32_next(&g) << 32) | bfamx_pcg32_next(&g);
  }
  return BFAMX_OK;
}

size_t bfamx_request_bytes(const bfamx_request_t *rq, bfamx_precision_t prec)
{
  if (!rq || rq->n < 1 || rq->N < 1)
    return 0;
  const int32_t K = rq->n * rq->n * rq->n;
  return bfamx_subdomain_bytes(3, rq->N, K, BFAMX_ELASTIC3D, prec);
}

/*
 * Host setup of one request: unit periodic cube of n^3 elements (Morton
 * order), LGL order N, material in horizontal layers (2..5 layers with
 * rho ~ U[1.5,3], cp ~ U[3,6], cs = cp / U[1.6,2]), explosive Ricker source
 * at a random point of [0.25,0.75]^3 peaking at half the run,
 * dt = 0.5 h / (cp_max (N+1)^2).  Everything here is host work (no device
 * call), so a worker thread can prepare requests while batches step.
 */
typedef struct
{
  int n, N, Np;
  int32_t K;
  double dt, f0, t0;
  int32_t src_elem;
  int32_t *vM, *vP;
  double *x0, *h, *mat, *wts; /* mat: rho[K], cp[K], cs[K] */
  double r[8], w[8], wi[8], Dr[64];
} request_setup_t;

static void request_setup_free(request_setup_t *s)
{
  free(s->vM);
  free(s->vP);
  free(s->x0);
  free(s->h);
  free(s->mat);
  free(s->wts);
  memset(s, 0, sizeof(*s));
}

static int request_prepare(const bfamx_request_t *rq, request_setup_t *s)
{
  memset(s, 0, sizeof(*s));
  if (!rq || rq->n < 1 || rq->N < 3 || rq->N > 7 || rq->nsteps < 1)
    return bfamx_fail(BFAMX_EINVAL, "request: bad request");
  const int n = rq->n, N = rq->N, n1 = N + 1;
  const int Np = n1 * n1 * n1, nfp = n1 * n1;
  const int32_t K = n * n * n;
  const int nn[3] = {n, n, n}, per[3] = {1, 1, 1};
  const double L[3] = {1, 1, 1};
  s->n = n;
  s->N = N;
  s->Np = Np;
  s->K = K;
  int rc = BFAMX_OK;
  int32_t *EToE = malloc(sizeof(int32_t) * 6 * K);
  int8_t *EToF = malloc(6 * (size_t)K);
  int32_t *lex = malloc(sizeof(int32_t) * K);
  s->x0 = malloc(sizeof(double) * 3 * K);
  s->h = malloc(sizeof(double) * 3 * K);
  s->vM = malloc(sizeof(int32_t) * 6 * (size_t)K * nfp);
  s->vP = malloc(sizeof(int32_t) * 6 * (size_t)K * nfp);
  s->mat = malloc(sizeof(double) * 3 * K);
  s->wts = malloc(sizeof(double) * Np);
  if (!EToE || !EToF || !lex || !s->x0 || !s->h || !s->vM || !s->vP || !s->mat || !s->wts)
  {
    rc = bfamx_fail(BFAMX_EINVAL, "request: host allocation failed");
    goto done;
  }
  if ((rc = bfamx_brick(3, nn, per, L, 1, EToE, EToF, lex, s->x0, s->h)) ||
      (rc = bfamx_lgl(N, s->r, s->w)) || (rc = bfamx_diff_matrix(N, s->r, s->Dr)) ||
      (rc = bfamx_build_vmaps(3, N, K, EToE, EToF, s->vM, s->vP)))
    goto done;
  for (int i = 0; i < n1; ++i)
    s->wi[i] = 1.0 / s->w[i];

  /* layered material */
  bfamx_pcg32_t g;
  bfamx_pcg32_seed(&g, rq->seed, 17);
  const int nl = 2 + (int)bfamx_pcg32_bounded(&g, 4);
  double zb[6], lay[6][3];
  for (int l = 0; l < nl; ++l)
  {
    const double rho = 1.5 + 1.5 * bfamx_pcg32_uniform(&g);
    const double cp = 3.0 + 3.0 * bfamx_pcg32_uniform(&g);
    lay[l][0] = rho;
    lay[l][1] = cp;
    lay[l][2] = cp / (1.6 + 0.4 * bfamx_pcg32_uniform(&g));
  }
  for (int l = 0; l < nl - 1; ++l)
    zb[l] = (double)(l + 1) / nl + (bfamx_pcg32_uniform(&g) - 0.5) / (2.0 * nl);
  double cmax = 0;
  for (int32_t e = 0; e < K; ++e)
  {
    const double zc = s->x0[3 * e + 2] + 0.5 * s->h[3 * e + 2];
    int l = 0;
    while (l < nl - 1 && zc > zb[l])
      ++l;
    for (int c = 0; c < 3; ++c)
      s->mat[c * K + e] = lay[l][c];
    if (lay[l][1] > cmax)
      cmax = lay[l][1];
  }
  s->dt = 0.5 * (1.0 / n) / (cmax * n1 * n1);

  /* Ricker point source */
  double xs[3], rs[3], lx[3][8];
  int ijk[3];
  for (int d = 0; d < 3; ++d)
  {
    xs[d] = 0.25 + 0.5 * bfamx_pcg32_uniform(&g);
    ijk[d] = (int)(xs[d] * n);
    if (ijk[d] > n - 1)
      ijk[d] = n - 1;
  }
  const int32_t lexs = ijk[0] + n * (ijk[1] + n * ijk[2]);
  int32_t es = 0;
  for (int32_t e = 0; e < K; ++e)
    if (lex[e] == lexs)
      es = e;
  double J = 1;
  for (int d = 0; d < 3; ++d)
  {
    rs[d] = 2.0 * (xs[d] - s->x0[3 * es + d]) / s->h[3 * es + d] - 1.0;
    bfamx_lagrange_eval(N, s->r, rs[d], lx[d]);
    J *= 0.5 * s->h[3 * es + d];
  }
  for (int q = 0; q < Np; ++q)
  {
    const int i = q % n1, j = (q / n1) % n1, k = q / (n1 * n1);
    s->wts[q] = lx[0][i] * lx[1][j] * lx[2][k] / (J * s->w[i] * s->w[j] * s->w[k]);
  }
  const double T = rq->nsteps * s->dt;
  s->src_elem = es;
  s->f0 = 2.4 / T;
  s->t0 = 0.5 * T;
done:
  free(EToE);
  free(EToF);
  free(lex);
  if (rc)
    request_setup_free(s);
  return rc;
}

/* the prepared request on the device, source attached, at rest at t = 0 */
static int request_upload(bfamx_device_t *dev, const request_setup_t *s,
                          bfamx_precision_t prec, bfamx_subdomain_t **out)
{
  bfamx_dgx_t dg = {3, s->N, s->Np, s->K, s->r, s->w, s->wi, s->Dr, s->vM, s->vP};
  bfamx_physics_desc_t ph = {BFAMX_ELASTIC3D, prec, s->h, s->x0, s->mat, s->mat + s->K,
                             s->mat + 2 * s->K};
  bfamx_lsrk_t ts;
  bfamx_lsrk_init(&ts, BFAMX_LSRK_KC54);
  bfamx_subdomain_t *sub = NULL;
  int rc = bfamx_subdomain_upload(dev, &dg, &ph, &ts, &sub);
  if (rc)
    return rc;
  static const double amp[9] = {0, 0, 0, 1, 1, 1, 0, 0, 0};
  if ((rc = bfamx_set_source(sub, s->src_elem, s->wts, amp, s->f0, s->t0)) ||
      (rc = bfamx_set_fields(sub, NULL, 0.0, s->dt)))
  {
    bfamx_subdomain_free(sub);
    return rc;
  }
  *out = sub;
  return BFAMX_OK;
}

int bfamx_request_build(bfamx_device_t *dev, const bfamx_request_t *rq,
                        bfamx_precision_t prec, bfamx_subdomain_t **out,
                        double *dt_out)
{
  if (!dev || !rq || !out)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_request_build: bad request");
  request_setup_t s;
  int rc = request_prepare(rq, &s);
  if (rc)
    return rc;
  rc = request_upload(dev, &s, prec, out);
  if (!rc && dt_out)
    *dt_out = s.dt;
  request_setup_free(&s);
  return rc;
}

int bfamx_request_inputs(const bfamx_request_t *rq, double *rho, double *cp, double *cs,
                         int32_t *src_elem, double *src_weights, double *f0, double *t0,
                         double *dt)
{
  if (!rq)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_request_inputs: NULL request");
  request_setup_t s;
  int rc = request_prepare(rq, &s);
  if (rc)
    return rc;
  const size_t K = (size_t)s.K;
  if (rho)
    memcpy(rho, s.mat, sizeof(double) * K);
  if (cp)
    memcpy(cp, s.mat + K, sizeof(double) * K);
  if (cs)
    memcpy(cs, s.mat + 2 * K, sizeof(double) * K);
  if (src_elem)
    *src_elem = s.src_elem;
  if (src_weights)
    memcpy(src_weights, s.wts, sizeof(double) * (size_t)s.Np);
  if (f0)
    *f0 = s.f0;
  if (t0)
    *t0 = s.t0;
  if (dt)
    *dt = s.dt;
  request_setup_free(&s);
  return BFAMX_OK;
}

/*
 * Preparation ahead of admission: a host worker builds each request's setup
 * once it has arrived, at most PREP_AHEAD requests ahead of the admission
 * point, so that admitting a request costs only its upload.
 */
enum
{
  PREP_AHEAD = 8
};

typedef struct
{
  const bfamx_request_t *reqs;
  int nreq;
  double t0;
  request_setup_t *setup;
  int *state; /* 0 not yet, 1 ready, 2 failed */
  int admitted; /* requests the serve loop has taken */
  int stop;
  pthread_mutex_t mu;
  pthread_cond_t cv;
} prep_t;

/* wait until `arrival` seconds after the start, or until stopped (1) */
static int wait_arrival(prep_t *p, double arrival)
{
  pthread_mutex_lock(&p->mu);
  for (;;)
  {
    const double left = arrival - (now_s() - p->t0);
    if (p->stop || left <= 0)
      break;
    struct timespec ts;
    clock_gettime(CLOCK_REALTIME, &ts);
    const double step = left < 0.05 ? left : 0.05;
    ts.tv_nsec += (long)(step * 1e9);
    ts.tv_sec += ts.tv_nsec / 1000000000L;
    ts.tv_nsec %= 1000000000L;
    pthread_cond_timedwait(&p->cv, &p->mu, &ts);
  }
  const int stop = p->stop;
  pthread_mutex_unlock(&p->mu);
  return stop;
}

static void *prep_main(void *arg)
{
  prep_t *p = arg;
  for (int i = 0; i < p->nreq; ++i)
  {
    pthread_mutex_lock(&p->mu);
    while (!p->stop && i >= p->admitted + PREP_AHEAD)
      pthread_cond_wait(&p->cv, &p->mu);
    const int stop = p->stop;
    pthread_mutex_unlock(&p->mu);
    if (stop)
      break;
    /* a request is prepared only once it has arrived; a stop ends the wait */
    if (wait_arrival(p, p->reqs[i].arrival))
      break;
    const int rc = request_prepare(&p->reqs[i], &p->setup[i]);
    pthread_mutex_lock(&p->mu);
    p->state[i] = rc ? 2 : 1;
    pthread_cond_broadcast(&p->cv);
    pthread_mutex_unlock(&p->mu);
  }
  return NULL;
}

typedef struct
{
  int idx;
  bfamx_subdomain_t *sub;
  int left; /* steps still to run */
  size_t bytes;
} active_t;

static int cmp_double(const void *a, const void *b)
{
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

static double quantile(const double *sorted, int n, double p)
{
  if (n == 0)
    return 0;
  /* nearest-rank */
  int k = (int)ceil(p * n) - 1;
  if (k < 0)
    k = 0;
  if (k > n - 1)
    k = n - 1;
  return sorted[k];
}

int bfamx_serve(bfamx_device_t *dev, const bfamx_request_t *reqs, int nreq,
                bfamx_precision_t prec, size_t budget_bytes, int max_chunk,
                bfamx_request_result_t *res, bfamx_serve_stats_t *stats)
{
  return bfamx_serve_cb(dev, reqs, nreq, prec, budget_bytes, max_chunk, NULL, NULL, res,
                        stats);
}

int bfamx_serve_cb(bfamx_device_t *dev, const bfamx_request_t *reqs, int nreq,
                   bfamx_precision_t prec, size_t budget_bytes, int max_chunk,
                   bfamx_serve_done_t done, void *user, bfamx_request_result_t *res,
                   bfamx_serve_stats_t *stats)
{
  if (!dev || nreq < 0 || (nreq && (!reqs || !res)) || !stats || max_chunk < 1 ||
      prec == BFAMX_FP64_PLAIN)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_serve: bad arguments");
  for (int i = 0; i < nreq; ++i)
  {
    if (i && reqs[i].arrival < reqs[i - 1].arrival)
      return bfamx_fail(BFAMX_EINVAL, "bfamx_serve: arrivals must be sorted");
    if (reqs[i].n < 1 || reqs[i].N < 3 || reqs[i].N > 7 || reqs[i].nsteps < 1)
      return bfamx_fail(BFAMX_EINVAL, "bfamx_serve: request %d is malformed", i);
    if (bfamx_request_bytes(&reqs[i], prec) > budget_bytes)
      return bfamx_fail(BFAMX_EBUDGET,
                        "bfamx_serve: request %d needs %zu bytes, budget %zu", i,
                        bfamx_request_bytes(&reqs[i], prec), budget_bytes);
  }
  memset(stats, 0, sizeof(*stats));
  memset(res, 0, sizeof(*res) * (size_t)nreq);
  active_t *act = calloc((size_t)nreq + 1, sizeof(active_t));
  bfamx_subdomain_t **subs = calloc((size_t)nreq + 1, sizeof(bfamx_subdomain_t *));
  prep_t prep;
  memset(&prep, 0, sizeof(prep));
  prep.setup = calloc((size_t)nreq + 1, sizeof(request_setup_t));
  prep.state = calloc((size_t)nreq + 1, sizeof(int));
  if (!act || !subs || !prep.setup || !prep.state)
  {
    free(act);
    free(subs);
    free(prep.setup);
    free(prep.state);
    return bfamx_fail(BFAMX_EINVAL, "bfamx_serve: host allocation failed");
  }
  int rc = BFAMX_OK, nact = 0, next = 0, ndone = 0, dirty = 0;
  size_t resident = 0;
  bfamx_batch_t *batch = NULL;
  double dof_stages = 0;
  const double t0 = now_s();
  prep.reqs = reqs;
  prep.nreq = nreq;
  prep.t0 = t0;
  pthread_mutex_init(&prep.mu, NULL);
  pthread_cond_init(&prep.cv, NULL);
  pthread_t worker;
  const int have_worker = nreq > 0 && pthread_create(&worker, NULL, prep_main, &prep) == 0;
  if (nreq > 0 && !have_worker)
  {
    rc = bfamx_fail(BFAMX_EINVAL, "bfamx_serve: cannot start the preparing thread");
    goto out;
  }

  while (ndone < nreq)
  {
    double t = now_s() - t0;
    /* admission: arrival order, while the budget holds; the worker has
     * prepared the request's host setup (or is finishing it) */
    while (next < nreq && reqs[next].arrival <= t)
    {
      const size_t need = bfamx_request_bytes(&reqs[next], prec);
      if (resident + need > budget_bytes)
        break;
      pthread_mutex_lock(&prep.mu);
      while (prep.state[next] == 0)
        pthread_cond_wait(&prep.cv, &prep.mu);
      const int state = prep.state[next];
      pthread_mutex_unlock(&prep.mu);
      if (state != 1)
      {
        rc = bfamx_fail(BFAMX_EINVAL, "bfamx_serve: request %d could not be prepared", next);
        goto out;
      }
      active_t *a = &act[nact];
      a->idx = next;
      a->left = reqs[next].nsteps;
      a->bytes = need;
      res[next].arrival = reqs[next].arrival;
      res[next].admitted = t;
      res[next].dof = 9LL * reqs[next].n * reqs[next].n * reqs[next].n *
                      (int64_t)((reqs[next].N + 1) * (reqs[next].N + 1) * (reqs[next].N + 1));
      rc = request_upload(dev, &prep.setup[next], prec, &a->sub);
      request_setup_free(&prep.setup[next]);
      pthread_mutex_lock(&prep.mu);
      prep.admitted = next + 1;
      pthread_cond_broadcast(&prep.cv);
      pthread_mutex_unlock(&prep.mu);
      if (rc)
        goto out;
      resident += need;
      if (resident > stats->peak_bytes)
        stats->peak_bytes = resident;
      ++nact;
      ++next;
      dirty = 1;
      t = now_s() - t0;
    }
    if (nact > stats->max_active)
      stats->max_active = nact;
    if (nact == 0)
    {
      /* idle until the next arrival */
      sleep_s(reqs[next].arrival - (now_s() - t0));
      continue;
    }
    if (dirty)
    {
      if (batch)
        bfamx_batch_free(batch);
      batch = NULL;
      for (int i = 0; i < nact; ++i)
        subs[i] = act[i].sub;
      if ((rc = bfamx_batch_create(dev, subs, nact, &batch)))
        goto out;
      /* mixed orders: one launch or one per order, timed once per order
       * set and cached on the device; a failure keeps the default */
      if (bfamx_batch_autotune(batch, 2, 0, NULL))
        bfamx_clear_error();
      dirty = 0;
      ++stats->nbatches;
    }
    int chunk = max_chunk;
    for (int i = 0; i < nact; ++i)
      if (act[i].left < chunk)
        chunk = act[i].left;
    if ((rc = bfamx_batch_step(batch, chunk)) ||
        (rc = bfamx_device_synchronize(dev)))
      goto out;
    stats->nlaunch += 5 * chunk;
    t = now_s() - t0;
    for (int i = 0; i < nact;)
    {
      act[i].left -= chunk;
      dof_stages += 5.0 * chunk * (double)res[act[i].idx].dof;
      if (act[i].left > 0)
      {
        ++i;
        continue;
      }
      bfamx_request_result_t *r = &res[act[i].idx];
      r->finished = t;
      r->latency = t - r->arrival;
      if ((rc = bfamx_reduce_energy(act[i].sub, &r->energy)))
        goto out;
      if (done)
        done(act[i].idx, act[i].sub, user);
      bfamx_subdomain_free(act[i].sub);
      resident -= act[i].bytes;
      act[i] = act[nact - 1];
      --nact;
      ++ndone;
      dirty = 1;
    }
  }
  {
    const double wall = now_s() - t0;
    stats->wall_s = wall;
    stats->dof_stage_per_s = wall > 0 ? dof_stages / wall : 0;
    double *lat = malloc(sizeof(double) * (size_t)(nreq ? nreq : 1));
    for (int i = 0; i < nreq; ++i)
      lat[i] = res[i].latency;
    qsort(lat, (size_t)nreq, sizeof(double), cmp_double);
    stats->p50_s = quantile(lat, nreq, 0.50);
    stats->p99_s = quantile(lat, nreq, 0.99);
    stats->max_s = nreq ? lat[nreq - 1] : 0;
    free(lat);
  }
out:
  if (have_worker)
  {
    pthread_mutex_lock(&prep.mu);
    prep.stop = 1;
    pthread_cond_broadcast(&prep.cv);
    pthread_mutex_unlock(&prep.mu);
    pthread_join(worker, NULL);
  }
  for (int i = 0; i < nreq; ++i)
    if (prep.state[i] == 1)
      request_setup_free(&prep.setup[i]);
  pthread_mutex_destroy(&prep.mu);
  pthread_cond_destroy(&prep.cv);
  if (batch)
    bfamx_batch_free(batch);
  for (int i = 0; i < nact; ++i)
    bfamx_subdomain_free(act[i].sub);
  free(act);
  free(subs);
  free(prep.setup);
  free(prep.state);
  return rc;
}
